"""The measured workloads at their full size, each against an independent FP64
reference on the device.

- 15,000 x 15,000 pair, growth_iters=3, 6-term split-BF16 cost: the whole
  growth loop against the plain-GPU FP64 restatement of the reference
  (wot_amd.plain_gpu, pinned to the reference's golden vectors on the CPU by
  tests/test_plain_baseline.py) on the FP64 cost.
- 38-pair time course x 9 (eps, lambda1): sampled (pair, eps) plans against
  plain-GPU FP64 solves; the column-marginal identity on all 342 solves.
- 8,000 x 8,000 gradient: the adjoint kernel against FP64 torch autograd of
  the primal at the same converged plan.
- 60,000 x 60,000 never-stored solve: the column-marginal identity with the
  FP64 direct cost, the median checked by counting FP64 distances below and
  above it (independent of the radix select), and a 5-map 12,000-cell chain
  against an FP64 GEMM chain.

Gates (the target spec): plan relative Frobenius error <= 1e-5, marginals
<= 1e-6 relative.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SOLVER = dict(epsilon=0.05, lambda1=1.0, lambda2=50.0, epsilon0=1.0, tau=1e4, scaling_iter=3000,
              inner_iter_max=50, extra_iter=1000)
PLAN_TOL = 1e-5
MARG_TOL = 1e-6


@pytest.fixture(scope='module')
def mods():
    import torch
    assert torch.cuda.is_available(), 'GPU tests need a ROCm device'
    import wot_amd.ot as ot
    from wot_amd import plain_gpu
    from wot_amd.ot import device as dev
    return torch, ot, dev, plain_gpu


def _fp64_cost(torch, dev, x, y):
    """sqeuclid / median in FP64 on the device (numpy median semantics)."""
    X = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    Y = torch.from_numpy(np.ascontiguousarray(y)).cuda()
    D = dev.sqeuclid_f64(X, Y)
    v = torch.sort(D.view(-1)).values
    n = v.numel()
    med = (v[n // 2 - 1] + v[n // 2]) / 2 if n % 2 == 0 else v[n // 2]
    del v
    return D / med


def _rel(torch, a, b):
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


def _marg(torch, a, b):
    return max(float((a.sum(k) - b.sum(k)).abs().max() / b.sum(k).abs().max()) for k in (0, 1))


@pytest.mark.timeout(600)
def test_pair_15k_growth_loop_vs_plain_fp64(mods):
    torch, ot, dev, plain = mods
    from wot_amd.synthetic import day_pair
    dp = day_pair(15000, 15000, dim=30, n_components=10, seed=1)
    cost = ot.DeviceCost.from_coordinates(dp.x, dp.y, mode='bf16x6')
    R, learned = ot.compute_transport_matrix(solver=ot.transport_stablev2, C=cost, G=dp.growth, growth_iters=3,
                                             device_out=True, **SOLVER)
    del cost
    C = _fp64_cost(torch, dev, dp.x, dp.y)
    G = torch.from_numpy(dp.growth).cuda()
    ref_learned = []
    Rp = None
    for it in range(3):
        if it > 0:
            G = Rp.sum(1)
        ref_learned.append(G)
        del Rp
        Rp = plain.transport_stablev2(C, G=G, **SOLVER)
    del C
    e, m = _rel(torch, R, Rp), _marg(torch, R, Rp)
    print('15k pair, growth_iters=3: rel frob %.3e, marginals %.3e' % (e, m))
    assert e <= PLAN_TOL and m <= MARG_TOL
    for a, b in zip(learned, ref_learned):
        a = torch.as_tensor(a).cuda() if not isinstance(a, torch.Tensor) else a
        assert float((a - b).abs().max() / b.abs().max()) <= MARG_TOL


@pytest.mark.timeout(900)
def test_time_course_38_pairs_9_combos(mods):
    torch, ot, dev, plain = mods
    from wot_amd.ot import batch
    from wot_amd.synthetic import time_course
    coords, growth, counts = time_course(n_days=39, min_cells=1000, max_cells=20000, dim=30, n_components=20, seed=2)
    sched = {k: v for k, v in SOLVER.items() if k not in ('epsilon', 'lambda1')}
    res = batch.solve_time_course(coords, growth, epsilons=(0.01, 0.05, 0.1), lambda1s=(0.5, 1.0, 5.0), **sched)
    assert len(res.combos) == 9
    pair_ij = [int(counts[t]) * int(counts[t + 1]) for t in range(len(counts) - 1)]
    # column-marginal identity of every solve: colsum(R J) = q I exp(-g / lambda2)
    buf = torch.empty(max(pair_ij), dtype=torch.float64, device='cuda')
    worst = 0.0
    for p in range(len(pair_ij)):
        n_i, n_j = res.shapes[p]
        q = float(np.mean(growth[p]))
        for e in range(9):
            R = res.plan(p, e, out=buf[:n_i * n_j].view(n_i, n_j))
            want = q * n_i * torch.exp(-res.g[p][e] / sched['lambda2'])
            got = R.sum(0) * n_j
            assert torch.isfinite(got).all()
            worst = max(worst, float((got - want).abs().max() / want.abs().max()))
    print('time course: marginal identity over 342 solves %.3e' % worst)
    assert worst <= MARG_TOL
    del buf
    # sampled plans against plain-GPU FP64 solves: smallest, median, largest pair x eps
    order = np.argsort(pair_ij)
    for p in (int(order[0]), int(order[len(order) // 2]), int(order[-1])):
        C = _fp64_cost(torch, dev, coords[p], coords[p + 1])
        for eps in (0.01, 0.05, 0.1):
            e = res.combos.index((eps, 1.0))
            Rp = plain.transport_stablev2(C, G=growth[p], epsilon=eps, lambda1=1.0, **sched)
            R = res.plan(p, e, device_out=True)
            err, m = _rel(torch, R, Rp), _marg(torch, R, Rp)
            print('pair %d (%d x %d) eps %.2f: rel frob %.3e marginals %.3e' % (p, *R.shape, eps, err, m))
            assert err <= PLAN_TOL and m <= MARG_TOL
            del Rp, R
        del C
        torch.cuda.empty_cache()


@pytest.mark.timeout(600)
def test_gradient_8k_vs_fp64_autograd(mods):
    torch, ot, dev, plain = mods
    from wot_amd.ot import adjoint as adj
    from wot_amd.synthetic import day_pair
    dp = day_pair(8000, 8000, dim=30, n_components=10, seed=5)
    x = torch.from_numpy(dp.x).float().cuda().requires_grad_(True)
    lg = torch.from_numpy(np.log(dp.growth)).float().cuda().requires_grad_(True)
    y = torch.from_numpy(dp.y).float().cuda()
    val = ot.ot_objective(x, lg, y)
    val.backward()
    assert x.grad.dtype == torch.float32 and torch.isfinite(x.grad).all() and torch.isfinite(lg.grad).all()
    # FP64 autograd of the primal at the same converged plan (envelope theorem)
    X = x.detach().double()
    Y = y.double()
    G = torch.exp(lg.detach().double())
    cost = ot.DeviceCost.from_coordinates(X, Y)
    prm = dict(adj.DEFAULTS)
    r = dev.solve(cost, G, 'duality_gap', prm, want_plan=False, want_potentials=True)
    eps = r['stats']['epsilon_final']
    gx, glg = adj.adjoint(cost, X, Y, G, r['f'], r['g'], prm['lambda1'], prm['lambda2'], eps)
    xx = X.clone().requires_grad_(True)
    ll = lg.detach().double().clone().requires_grad_(True)
    D = (xx * xx).sum(1)[:, None] + (Y * Y).sum(1)[None, :] - 2.0 * xx @ Y.T
    D = D + (dev.sqeuclid_f64(X, Y) - D).detach()
    mid = cost.median_index.cpu().numpy()
    flat = D.reshape(-1)
    C = D / (0.5 * (flat[int(mid[0])] + flat[int(mid[1])]))
    n_i, n_j = C.shape
    p = torch.exp(ll)
    q = p.mean()
    with torch.no_grad():
        R = torch.exp((r['f'][:, None] + r['g'][None, :] - C.detach()) / eps)
    K = torch.exp(-C / eps)
    xr, yc = R.sum(1) / n_j, R.sum(0) / n_i
    P = (prm['lambda1'] * (1.0 / n_i * (xr * torch.log(xr / p) - xr + p)).sum()
         + prm['lambda2'] * (1.0 / n_j * (yc * torch.log(yc / q) - yc + q)).sum()
         + (eps * (R * torch.log(R) - R + K).sum() + (R * C).sum()) / (n_i * n_j))
    P.backward()
    e_x = float(torch.linalg.norm(gx - xx.grad) / torch.linalg.norm(xx.grad))
    e_g = float(torch.linalg.norm(glg - ll.grad) / torch.linalg.norm(ll.grad))
    # the FP32 gradients of the autograd function are the FP64 adjoint rounded
    e_x32 = float(torch.linalg.norm(x.grad.double() - xx.grad) / torch.linalg.norm(xx.grad))
    print('8k gradient vs FP64 autograd: x %.3e, log g %.3e (FP32 output %.3e)' % (e_x, e_g, e_x32))
    assert float(val) == pytest.approx(float(P), rel=1e-6)
    assert e_x <= 1e-5 and e_g <= 1e-6 and e_x32 <= 1e-5


@pytest.mark.timeout(900)
def test_never_stored_60k_and_chain(mods):
    torch, ot, dev, plain = mods
    from wot_amd.ot import large
    import wot_amd.tmap as tm
    from wot_amd.synthetic import day_pair, time_course
    n = 60000
    dp = day_pair(n, n, dim=50, n_components=30, seed=6)
    X = torch.from_numpy(dp.x).cuda()
    Y = torch.from_numpy(dp.y).cuda()
    res = large.transport_stablev2_coords(X, Y, torch.from_numpy(dp.growth).cuda(), terms=6, **SOLVER)
    assert torch.isfinite(res.f).all() and torch.isfinite(res.g).all()
    med = float(res.median.cpu()[0])
    # median: at most half the distances below it and at most half above it,
    # and it is the mean of the largest one below and the smallest one above
    below = above = 0
    lo, hi = -np.inf, np.inf
    colsum = torch.zeros(n, dtype=torch.float64, device='cuda')
    step = 80
    for r0 in range(0, n, step):
        # FP64 direct differences in torch (independent of the library's kernels)
        D = ((X[r0:r0 + step, None, :] - Y[None, :, :]) ** 2).sum(-1)
        below += int((D < med).sum())
        above += int((D > med).sum())
        lo = max(lo, float(D[D < med].max()))
        hi = min(hi, float(D[D > med].min()))
        C = D / med
        colsum += torch.exp((res.f[r0:r0 + step, None] + res.g[None, :] - C) / res.epsilon).sum(0)
        del D, C
    assert below <= n * n // 2 and above <= n * n // 2, (below, above)
    assert med == pytest.approx(0.5 * (lo + hi), rel=1e-15, abs=0.0)
    # column-marginal identity of the returned plan (colsum of R = q I exp(-g/lambda2)),
    # evaluated with the FP64 direct cost: the 6-term cost error enters here
    want = float(np.mean(dp.growth)) * n * torch.exp(-res.g / SOLVER['lambda2'])
    ident = float((colsum - want).abs().max() / want.abs().max())
    print('60k never stored: median %.12g (below %d, above %d), marginal identity %.3e' % (med, below, above, ident))
    assert ident <= PLAN_TOL
    del colsum, X, Y, res
    torch.cuda.empty_cache()
    # 5-map chain of 12,000-cell days: FP32 MFMA GEMMs against FP64 products
    coords, growth, _ = time_course(dim=50, n_components=30, seed=7, counts=[12000] * 6)
    maps = []
    for t in range(5):
        cst = ot.DeviceCost.from_coordinates(coords[t], coords[t + 1])
        maps.append(ot.transport_stablev2(cst, G=growth[t], device_out=True, **SOLVER))
        del cst
    chain32 = tm.chain_product([m.float() for m in maps]).double()
    ref = maps[0]
    for m in maps[1:]:
        ref = ref @ m
    rel = _rel(torch, chain32, ref)
    rn = chain32 / chain32.sum(1, keepdim=True)
    rr = ref / ref.sum(1, keepdim=True)
    row = float((rn - rr).abs().max() / rr.abs().max())
    print('5-map 12k chain: rel frob %.3e, row-normalised max %.3e' % (rel, row))
    assert rel <= PLAN_TOL and row <= PLAN_TOL
