"""Batched time course (all day pairs x an (epsilon, lambda1) grid in one
persistent sweep per half-iteration) against the FP64 oracle, solve by solve."""
import numpy as np
import pytest

from oracle import ot_oracle as O

pytestmark = pytest.mark.gpu

SCHED = dict(scaling_iter=300, extra_iter=100, inner_iter_max=50, tau=1e4)


@pytest.fixture(scope='module')
def batch():
    import torch
    assert torch.cuda.is_available()
    from wot_amd.ot import batch
    return batch


def _course(seed, counts, dim=10):
    from wot_amd.synthetic import time_course
    coords, growth, _ = time_course(dim=dim, n_components=4, seed=seed, counts=counts)
    return coords, growth


def _check(res, coords, growth, epsilon0, lambda2=50.0, tol=1e-5):
    worst = 0.0
    for p in range(len(coords) - 1):
        C = O.default_cost_matrix(coords[p], coords[p + 1])
        for e, (eps, lam1) in enumerate(res.combos):
            R = res.plan(p, e)
            Rref = O.transport_stablev2(C, lam1, lambda2, eps, G=growth[p], epsilon0=epsilon0, **SCHED)
            err = np.linalg.norm(R - Rref) / np.linalg.norm(Rref)
            marg = max(np.abs(R.sum(a) - Rref.sum(a)).max() / np.abs(Rref.sum(a)).max() for a in (0, 1))
            worst = max(worst, err)
            assert err <= tol and marg <= 1e-6, (p, eps, lam1, err, marg)
    return worst


@pytest.mark.parametrize('share_exp', [True, False])
def test_batch_grid_matches_oracle(batch, share_exp):
    coords, growth = _course(1, [150, 320, 200, 410])
    res = batch.solve_time_course(coords, growth, epsilons=(0.01, 0.05, 0.1), lambda1s=(0.5, 1.0, 5.0),
                                  graph_iters=50, share_exp=share_exp, **SCHED)
    assert len(res.combos) == 9
    assert res.stats['shared_group'] == (3 if share_exp else 0)
    worst = _check(res, coords, growth, 1.0)
    print('batch worst rel frob %.3e (share_exp=%s, exact rows %d)' % (worst, share_exp, res.stats['exact_rows']))


@pytest.mark.parametrize('inner_iter_max', [50, 5])
def test_batch_tiny_days(batch, inner_iter_max):
    """Days of 1, 2, 3 and 5 cells beside a 130-cell day: single-row and
    single-column pairs, widths below one 4-column group and far below a tile,
    every (eps, lambda1) combination against the oracle.  With eps changing
    every 5 iterations the schedule converges inside the run, so the shared
    modes (1, 1), (r, r), (2, r) and (2, 10) all run."""
    sched = dict(SCHED, inner_iter_max=inner_iter_max)
    coords, growth = _course(21, [1, 3, 2, 5, 130, 1])
    res = batch.solve_time_course(coords, growth, epsilons=(0.01, 0.05, 0.1), lambda1s=(0.5, 1.0, 5.0),
                                  graph_iters=50, **sched)
    worst = 0.0
    for p in range(len(coords) - 1):
        C = O.default_cost_matrix(coords[p], coords[p + 1])
        for e, (eps, lam1) in enumerate(res.combos):
            R = res.plan(p, e)
            Rref = O.transport_stablev2(C, lam1, 50.0, eps, G=growth[p], epsilon0=1.0, **sched)
            err = np.linalg.norm(R - Rref) / np.linalg.norm(Rref)
            marg = max(np.abs(R.sum(a) - Rref.sum(a)).max() / np.abs(Rref.sum(a)).max() for a in (0, 1))
            worst = max(worst, err)
            assert err <= 1e-5 and marg <= 1e-6, (p, C.shape, eps, lam1, err, marg)
    print('tiny days (inner_iter_max %d): worst rel frob %.3e' % (inner_iter_max, worst))


@pytest.mark.parametrize('epsilons,lambda1s,group', [
    ((0.05, 0.1), (1.0, 5.0), 2),          # 2 groups of 2
    ((0.05, 0.1), (0.5, 1.0, 5.0), 3),     # 2 groups of 3
    ((0.05,), (0.5, 5.0), 2),              # 1 group of 2
    ((0.01, 0.05, 0.1), (1.0,), 0),        # no equal epsilons: per-combination kernel
])
def test_shared_exponential_group_shapes(batch, epsilons, lambda1s, group):
    coords, growth = _course(7, [97, 210, 160])
    res = batch.solve_time_course(coords, growth, epsilons=epsilons, lambda1s=lambda1s, graph_iters=50, **SCHED)
    assert res.stats['shared_group'] == group
    _check(res, coords, growth, 1.0)


def test_shared_exponential_far_potentials_fall_back(batch):
    """eps = eps0 = 0.005 with lambda1 from 0.001 to 1000: the combinations'
    potentials and column factors drift so far apart (up to 86 / k, measured
    with an FP64 log-domain iteration on the CPU) that many factored row sums
    could lose mass to FP32 underflow; those rows are recomputed exactly and
    every plan still meets the gate (C max / eps = 604, so the oracle's kernel
    stays representable in FP64)."""
    import wot_amd.ot as ot
    coords, growth = _course(8, [130, 170])
    lams = (0.001, 1.0, 1000.0)
    res = batch.solve_time_course(coords, growth, epsilons=(0.005,), lambda1s=lams, epsilon0=0.005, graph_iters=0,
                                  **SCHED)
    assert res.stats['shared_group'] == 3
    print('exact rows %d' % res.stats['exact_rows'])
    assert res.stats['exact_rows'] > 0
    # against the single-pair solver on the same FP32 cost: the batch path adds no error
    cost = ot.DeviceCost.from_coordinates(coords[0], coords[1])
    C = O.default_cost_matrix(coords[0], coords[1])
    for e, lam1 in enumerate(lams):
        R1 = ot.transport_stablev2(cost, G=growth[0], lambda1=lam1, lambda2=50.0, epsilon=0.005, epsilon0=0.005,
                                   **SCHED)
        R2 = res.plan(0, e)
        assert np.linalg.norm(R2 - R1) / np.linalg.norm(R1) <= 1e-6, lam1
        # against the FP64 oracle: the plan gate holds; at C / eps ~ 600 the FP32
        # rounding of the stored cost (x 600 in the exponent) puts the marginals
        # near 1e-6 for either solver
        Rref = O.transport_stablev2(C, lam1, 50.0, 0.005, G=growth[0], epsilon0=0.005, **SCHED)
        err = np.linalg.norm(R2 - Rref) / np.linalg.norm(Rref)
        marg = [max(np.abs(R.sum(a) - Rref.sum(a)).max() / np.abs(Rref.sum(a)).max() for a in (0, 1)) for R in (R1, R2)]
        print('lambda1 %g: plan %.2e, marginals single %.2e batch %.2e' % (lam1, err, marg[0], marg[1]))
        assert err <= 1e-5 and marg[1] <= 2e-6


def test_batch_graph_replay_is_deterministic(batch):
    coords, growth = _course(2, [130, 257, 96])
    kw = dict(epsilons=(0.05, 0.1), lambda1s=(1.0, 5.0), **SCHED)
    a = batch.solve_time_course(coords, growth, graph_iters=50, **kw)
    b = batch.solve_time_course(coords, growth, graph_iters=0, **kw)
    c = batch.solve_time_course(coords, growth, graph_iters=7, **kw)  # 400 = 57*7 + 1: replays + a remainder
    for p in range(2):
        assert np.array_equal(a.f[p].cpu().numpy(), b.f[p].cpu().numpy())
        assert np.array_equal(a.g[p].cpu().numpy(), c.g[p].cpu().numpy())


def test_batch_fallback_rows(batch):
    """epsilon0 = eps = 0.01 with an outlier source cell: its first-iteration
    terms all underflow FP32 under the fixed shift, so the online-maximum pass
    must take over (FP64 stays finite: min C / eps < 700)."""
    coords, growth = _course(3, [120, 140])
    coords[0] = coords[0].copy()
    coords[0][0] += 8.0  # far from every target cell
    res = batch.solve_time_course(coords, growth, epsilons=(0.01,), lambda1s=(1.0,), epsilon0=0.01,
                                  graph_iters=0, **SCHED)
    C = O.default_cost_matrix(coords[0], coords[1])
    assert C[0].min() / 0.01 * 1.4427 > 149 and C[0].min() / 0.01 < 700
    assert res.stats['exact_rows'] > 0
    _check(res, coords, growth, 0.01)
    # the same outlier with the shared-exponential kernel (3 combinations of eps 0.01)
    res = batch.solve_time_course(coords, growth, epsilons=(0.01,), lambda1s=(0.5, 1.0, 5.0), epsilon0=0.01,
                                  graph_iters=0, **SCHED)
    assert res.stats['shared_group'] == 3 and res.stats['exact_rows'] > 0
    _check(res, coords, growth, 0.01)


def test_batch_single_combo_equals_single_solver(batch):
    """E = 1 batch == the single-pair solver on the same cost."""
    import wot_amd.ot as ot
    coords, growth = _course(4, [200, 230])
    res = batch.solve_time_course(coords, growth, epsilons=(0.05,), lambda1s=(1.0,), graph_iters=0, **SCHED)
    cost = ot.DeviceCost.from_coordinates(coords[0], coords[1])
    R1 = ot.transport_stablev2(cost, G=growth[0], lambda1=1.0, lambda2=50.0, epsilon=0.05, epsilon0=1.0, **SCHED)
    R2 = res.plan(0, 0)
    assert np.linalg.norm(R1 - R2) / np.linalg.norm(R1) < 1e-6


REAL = 255  # a group whose k ratio to the base group's is not an integer
COMPILED = (2 | (10 << 8), 2 | (4 << 8), 2 | (REAL << 8), 1 | (1 << 8), REAL | (REAL << 8))


def _expected_mode(ratios):
    """The library's choice for ratios k_g / k_base of groups 0 and 1: integer
    powers within 1e-10, else the real ratio; the first compiled of (p1, p0),
    (p1, real), (real, real)."""
    pw = [int(round(r)) if abs(r - round(r)) <= 1e-10 * round(r) and round(r) >= 1 else REAL for r in ratios]
    for want in (pw[1] | (pw[0] << 8), pw[1] | (REAL << 8), REAL | (REAL << 8)):
        if want in COMPILED:
            return want
    raise AssertionError('no compiled mode')


# eps grids whose k = log2(e)/eps stand in integer ratios to the largest eps's:
# (groups 1, 0) take its exponential squared and to the 10th / 4th power, or
# group 0 takes the real ratio 10/3 of its exponent
@pytest.mark.parametrize('epsilons,mode', [
    ((0.01, 0.05, 0.1), 2 | (10 << 8)),
    ((0.025, 0.05, 0.1), 2 | (4 << 8)),
    ((0.03, 0.05, 0.1), 2 | (REAL << 8)),
])
def test_shared_eps_powers_match_oracle(batch, epsilons, mode):
    """With WOTK_BATCH_SHARE_EPS, once the eps schedule has converged (here
    from iteration 130 of 400: eps changes every 5 iterations, and
    |k_g / k_base - m| <= 1e-10 m from e^-26..28 on) the smaller-eps groups
    take powers of the largest eps's exponential, before that its exponent
    times the real ratio; every plan still meets the oracle gate, and differs
    from the run with each group's own exponent by far less than the gate."""
    sched = dict(scaling_iter=300, extra_iter=100, inner_iter_max=5, tau=1e4)
    coords, growth = _course(11, [150, 320, 200])
    lams = (0.5, 1.0, 5.0)
    combos, eps, l1, l2 = batch.grid_arrays(2, epsilons, lams, 50.0)
    tc = batch.prepare_time_course(coords, growth)
    with batch.TimeCourseRun(tc, eps, l1, l2, epsilon0=1.0, graph_iters=50, share_eps=True, combos=combos,
                             **{k: v for k, v in sched.items() if k != 'tau'}) as run:
        run.advance(run.n_iter)
        modes = run.modes()
        res = run.result()
    # the host's choice per iteration
    e0 = np.array(epsilons)
    want = np.zeros(run.n_iter, dtype=np.int32)
    for n in range(run.n_iter):
        cur = (1.0 - e0) * np.exp(-float(min(n, 300) // 5)) + e0
        want[n] = _expected_mode(cur[2] / cur[:2])
    assert np.array_equal(modes, want), (np.unique(modes), np.unique(want))
    # eps0 for every group in the schedule's first step: one exponential for all
    # nine; then real ratios until they converge to the grid's
    assert (want[:5] == 1 | (1 << 8)).all() and (want[5:100] == REAL | (REAL << 8)).all()
    first = 5 + int(np.argmax(want[5:] != REAL | (REAL << 8)))
    assert 100 < first < 200 and (want[200:] == mode).all()
    worst = 0.0
    for p in range(2):
        C = O.default_cost_matrix(coords[p], coords[p + 1])
        for e, (ep, lam1) in enumerate(combos):
            R = res.plan(p, e)
            Rref = O.transport_stablev2(C, lam1, 50.0, ep, G=growth[p], epsilon0=1.0, **sched)
            err = np.linalg.norm(R - Rref) / np.linalg.norm(Rref)
            marg = max(np.abs(R.sum(a) - Rref.sum(a)).max() / np.abs(Rref.sum(a)).max() for a in (0, 1))
            worst = max(worst, err)
            assert err <= 1e-5 and marg <= 1e-6, (p, ep, lam1, err, marg)
    own = batch.solve_prepared(tc, epsilons, lams, epsilon0=1.0, graph_iters=50, share_eps=False, **sched)
    diff = max(np.linalg.norm(res.plan(p, e) - own.plan(p, e)) / np.linalg.norm(own.plan(p, e))
               for p in range(2) for e in range(9))
    print('mode %d from iteration %d: worst vs oracle %.2e, vs own exponentials %.2e, exact rows %d'
          % (mode, first, worst, diff, res.stats['exact_rows']))
    assert diff <= 2e-6


def test_shared_eps_modes_follow_the_schedule(batch):
    """Without the flag every group keeps its own exponent; with it every
    iteration shares the base group's: one exponential while all groups are at
    eps0 (the schedule's first step), the real ratio while the ratios are not
    integers."""
    coords, growth = _course(12, [90, 130])
    combos, eps, l1, l2 = batch.grid_arrays(1, (0.01, 0.05, 0.1), (0.5, 1.0, 5.0), 50.0)
    tc = batch.prepare_time_course(coords, growth)
    for share_eps, sched in ((False, dict(scaling_iter=300, extra_iter=100, inner_iter_max=5)),
                             (True, dict(scaling_iter=300, extra_iter=100, inner_iter_max=50))):
        with batch.TimeCourseRun(tc, eps, l1, l2, epsilon0=1.0, graph_iters=50, share_eps=share_eps,
                                 combos=combos, **sched) as run:
            m = run.modes()
            if share_eps:
                assert (m[:50] == 1 | (1 << 8)).all() and (m[50:] == REAL | (REAL << 8)).all()
            else:
                assert (m == 0).all()


def test_shared_eps_per_pair_grids(batch):
    """Per-pair eps (a day-pair configuration): pair 0 on {0.01, 0.05, 0.1},
    pair 1 on {0.02, 0.05, 0.1}.  Group 1 has the power 2 in both, group 0
    has 10 in one and 5 in the other, so once the schedule has converged the
    common mode is (2, real); every plan meets the oracle gate."""
    sched = dict(scaling_iter=300, extra_iter=100, inner_iter_max=5)
    coords, growth = _course(13, [140, 230, 190])
    lams = (0.5, 1.0, 5.0)
    grids = [(0.01, 0.05, 0.1), (0.02, 0.05, 0.1)]
    eps = np.array([[e for e in g for _ in lams] for g in grids])
    l1 = np.array([[l for _ in g for l in lams] for g in grids])
    l2 = np.full_like(eps, 50.0)
    tc = batch.prepare_time_course(coords, growth)
    with batch.TimeCourseRun(tc, eps, l1, l2, epsilon0=1.0, graph_iters=50, share_eps=True, **sched) as run:
        run.advance(run.n_iter)
        modes = run.modes()
        res = run.result()
    assert (modes[:5] == 1 | (1 << 8)).all()
    assert (modes[250:] == 2 | (REAL << 8)).all(), np.unique(modes[250:])
    for p in range(2):
        C = O.default_cost_matrix(coords[p], coords[p + 1])
        for e in range(9):
            ep, lam1 = eps[p, e], l1[p, e]
            R = res.plan(p, e)
            Rref = O.transport_stablev2(C, lam1, 50.0, ep, G=growth[p], epsilon0=1.0, tau=1e4, **sched)
            err = np.linalg.norm(R - Rref) / np.linalg.norm(Rref)
            marg = max(np.abs(R.sum(a) - Rref.sum(a)).max() / np.abs(Rref.sum(a)).max() for a in (0, 1))
            assert err <= 1e-5 and marg <= 1e-6, (p, ep, lam1, err, marg)
