"""GPU adjoint (wotk_ot_adjoint behind ot_objective) against the FP64 oracle
gradient and an FP64 torch autograd reference of the primal at the same
converged potentials (envelope theorem: the plan is held fixed)."""
import numpy as np
import pytest

from oracle import ot_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def mods():
    import torch
    assert torch.cuda.is_available()
    import wot_amd.ot as ot
    from wot_amd.ot import adjoint, device
    return torch, ot, adjoint, device


def _autograd_reference(torch, X, Y, lg, f, g, eps, lam1, lam2, mid_idx, dt=1.0, c_values=None):
    """FP64 torch autograd of primal() with R = exp((f + g - C)/eps) held fixed.
    ``c_values``: take the cost's values from this matrix (the device's FP32
    cost, widened), also where the median's derivative multiplies them, while
    the gradient still flows through D and the median."""
    x = X.clone().requires_grad_(True)
    l = lg.clone().requires_grad_(True)
    D = (x * x).sum(1)[:, None] + (Y * Y).sum(1)[None, :] - 2.0 * x @ Y.T
    Dd = ((X[:, None, :] - Y[None, :, :]) ** 2).sum(-1) if X.shape[0] * Y.shape[0] <= 4_000_000 else None
    if Dd is not None:  # exact direct-difference values, differentiable through the expansion
        D = D + (Dd - D).detach()
    flat = D.reshape(-1)
    n = flat.numel()
    if n % 2:
        m = flat[mid_idx[0]]
    else:
        m = 0.5 * (flat[mid_idx[0]] + flat[mid_idx[1]])
    if c_values is None:
        C = D / m
    else:  # value c_values; dC/dD = 1/m and dC/dm = -c_values/m, as the kernel takes them
        md = m.detach()
        C = c_values + (D - D.detach()) / md - c_values * (m - md) / md
    n_i, n_j = C.shape
    p = torch.exp(dt * l)
    q = p.mean()
    with torch.no_grad():
        R = torch.exp((f[:, None] + g[None, :] - C.detach()) / eps)
    K = torch.exp(-C / eps)
    xr = R.sum(1) / n_j
    yc = R.sum(0) / n_i
    F1 = lam1 * (1.0 / n_i * (xr * torch.log(xr / p) - xr + p)).sum()
    F2 = lam2 * (1.0 / n_j * (yc * torch.log(yc / q) - yc + q)).sum()
    logR = torch.log(R)
    P = F1 + F2 + (eps * (R * logR - R + K).sum() + (R * C).sum()) / (n_i * n_j)
    P.backward()
    return P.detach(), x.grad, l.grad


@pytest.mark.parametrize('n_i,n_j,d,solver', [(300, 400, 30, 'fixed_iters'), (1000, 1200, 30, 'duality_gap')])
def test_adjoint_kernel_vs_fp64_autograd(mods, n_i, n_j, d, solver):
    torch, ot, adjoint, device = mods
    from wot_amd.synthetic import day_pair
    dp = day_pair(n_i, n_j, dim=d, seed=31)
    X = torch.from_numpy(dp.x).cuda()
    Y = torch.from_numpy(dp.y).cuda()
    lg = torch.from_numpy(np.log(dp.growth)).cuda()
    cost = ot.DeviceCost.from_coordinates(X, Y)
    G = torch.exp(lg)
    prm = dict(adjoint.DEFAULTS)
    res = device.solve(cost, G, solver, prm, want_plan=False, want_potentials=True)
    eps = res['stats']['epsilon_final']
    gx, glg = adjoint.adjoint(cost, X, Y, G, res['f'], res['g'], 1.0, 50.0, eps)
    mid = cost.median_index.cpu().numpy()
    P, rgx, rglg = _autograd_reference(torch, X, Y, lg, res['f'], res['g'], eps, 1.0, 50.0, mid)
    ex = float(torch.linalg.norm(gx - rgx) / torch.linalg.norm(rgx))
    eg = float(torch.linalg.norm(glg - rglg) / torch.linalg.norm(rglg))
    print('adjoint vs autograd: grad_x %.3e grad_log_g %.3e' % (ex, eg))
    # the FP32 cost entries enter both sides identically except through C in
    # the reference (recomputed in FP64 here): bound by the FP32 rounding of C
    assert ex < 1e-5 and eg < 1e-6


@pytest.mark.parametrize('n_i,n_j,d,eps,odd_ld', [
    (300, 413, 40, 0.05, True),    # three 16-wide coordinate tiles, ldc = 414 (scalar cost loads)
    (257, 300, 64, 0.05, False),   # four tiles, a ragged last row block
    (100, 5000, 30, 0.05, False),  # 20 column splits, the last one ragged
    (200, 250, 70, 0.05, False),   # d > 64: the LDS kernel
    (300, 400, 30, 0.002, False),  # exp(-C/eps) below 2^-1000 for ~9% of the entries: R formed directly
    (1, 3, 5, 0.05, False),        # one row, three columns: a single masked step
    (5, 1, 30, 0.05, False),       # one column
])
def test_adjoint_shapes_against_autograd(mods, n_i, n_j, d, eps, odd_ld):
    """Adjoint at arbitrary potentials (the partial derivative at a fixed plan
    needs no optimum) against FP64 autograd on the same FP32 cost values."""
    torch, ot, adjoint, device = mods
    rng = np.random.default_rng(n_i + n_j + d)
    X = torch.from_numpy(rng.normal(size=(n_i, d))).cuda()
    Y = torch.from_numpy(rng.normal(size=(n_j, d)) + 0.3).cuda()
    lg = torch.from_numpy(rng.normal(scale=0.3, size=n_i)).cuda()
    cost = ot.DeviceCost.from_coordinates(X, Y)
    if odd_ld:
        C2 = torch.zeros((n_i, n_j + 1), dtype=torch.float32, device=X.device)
        C2[:, :n_j] = cost.C[:, :n_j]
        cost = ot.DeviceCost(C2, cost.CT, n_i, n_j, median=cost.median, median_index=cost.median_index)
        assert cost.ldc % 4 != 0
    Cv = cost.C[:, :n_j].double()
    # potentials that keep every row's and column's plan mass positive
    f = 0.5 * Cv.min(1).values
    g = 0.5 * Cv.min(0).values
    G = torch.exp(lg)
    gx, glg = adjoint.adjoint(cost, X, Y, G, f, g, 1.0, 50.0, eps)
    mid = cost.median_index.cpu().numpy()
    _, rgx, rglg = _autograd_reference(torch, X, Y, lg, f, g, eps, 1.0, 50.0, mid, c_values=Cv)
    assert torch.isfinite(rgx).all() and torch.isfinite(rglg).all()
    ex = float(torch.linalg.norm(gx - rgx) / torch.linalg.norm(rgx))
    eg = float(torch.linalg.norm(glg - rglg) / torch.linalg.norm(rglg))
    print('adjoint %dx%d d=%d eps=%g vs autograd: grad_x %.3e grad_log_g %.3e' % (n_i, n_j, d, eps, ex, eg))
    # same cost values on both sides: FP64 rounding only (an entry left out or
    # misplaced would show at ~1e-3)
    assert ex < 1e-12 and eg < 1e-12


def test_adjoint_vs_oracle_gradient(mods):
    """End to end: GPU solve + adjoint vs the FP64 oracle's converged gradient."""
    torch, ot, adjoint, device = mods
    from wot_amd.synthetic import day_pair
    dp = day_pair(120, 150, dim=5, seed=3)
    lg = np.log(dp.growth)
    prm = dict(lambda1=1.0, lambda2=50.0, epsilon=0.05, scaling_iter=3000, tau=1e4, epsilon0=1.0,
               extra_iter=1000, inner_iter_max=50)
    val, gx, glg, _ = O.objective_and_grad(dp.x, lg, dp.y, **prm)
    x = torch.from_numpy(dp.x).cuda().requires_grad_(True)
    l = torch.from_numpy(lg).cuda().requires_grad_(True)
    y = torch.from_numpy(dp.y).cuda()
    out = ot.ot_objective(x, l, y, solver='fixed_iters', **prm)
    out.backward()
    assert float(out.detach()) == pytest.approx(val, rel=1e-6)
    ex = np.linalg.norm(x.grad.cpu().numpy() - gx) / np.linalg.norm(gx)
    eg = np.linalg.norm(l.grad.cpu().numpy() - glg) / np.linalg.norm(glg)
    print('gpu vs oracle gradient: %.3e %.3e' % (ex, eg))
    assert ex < 1e-4 and eg < 1e-5


def test_ot_objective_fp32_inputs(mods):
    """FP32 embeddings and log growth (the differentiation workload's dtypes)."""
    torch, ot, adjoint, device = mods
    from wot_amd.synthetic import day_pair
    dp = day_pair(200, 220, dim=30, seed=8)
    x = torch.from_numpy(dp.x).float().cuda().requires_grad_(True)
    l = torch.from_numpy(np.log(dp.growth)).float().cuda().requires_grad_(True)
    y = torch.from_numpy(dp.y).float().cuda()
    out = ot.ot_objective(x, l, y)
    out.backward()
    assert out.dtype == torch.float32 and x.grad.dtype == torch.float32 and l.grad.dtype == torch.float32
    assert torch.isfinite(x.grad).all() and torch.isfinite(l.grad).all()
