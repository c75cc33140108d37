"""Never-materialised solver (cost regenerated per sweep on MFMA) against the
stored-cost solver (FP32 cost, FP64 exponents; itself checked against the
FP64 oracle) on the same inputs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCHED = dict(lambda1=1.0, lambda2=50.0, epsilon=0.05, scaling_iter=300, tau=1e4, epsilon0=1.0, extra_iter=100,
             inner_iter_max=50)


@pytest.fixture(scope='module')
def mods():
    import torch
    assert torch.cuda.is_available()
    import wot_amd.ot as ot
    from wot_amd.ot import device as dev
    from wot_amd.ot import large
    return torch, ot, dev, large


def _compare(torch, ot, dev, large, x, y, G, terms, sched):
    res = large.transport_stablev2_coords(x, y, G, terms=terms, **sched)
    cost = ot.DeviceCost.from_coordinates(x, y)
    ref = dev.solve(cost, G, 'fixed_iters', sched, want_plan=True, want_potentials=True)
    assert float(res.median.cpu()[0]) == float(cost.median.cpu()[0])
    rows = torch.arange(x.shape[0], device='cuda')
    R = res.plan_rows(x, y, rows)
    Rref = ref['plan']
    err = float(torch.linalg.norm(R - Rref) / torch.linalg.norm(Rref))
    return err, res


@pytest.mark.parametrize('terms,tol', [(6, 1e-5), (3, 1e-3)])
def test_fused_matches_stored_cost(mods, terms, tol):
    torch, ot, dev, large = mods
    from wot_amd.synthetic import day_pair
    dp = day_pair(700, 900, dim=50, n_components=8, seed=41)
    err, res = _compare(torch, ot, dev, large, dp.x, dp.y, dp.growth, terms, SCHED)
    print('fused terms=%d rel frob vs stored-cost solver: %.3e' % (terms, err))
    assert err < tol


def test_fused_repair_rows(mods):
    """eps0 = eps = 0.01 with an outlier source cell: its fixed-shift row sum
    underflows in the first sweep and the repair kernel must recompute it."""
    torch, ot, dev, large = mods
    from wot_amd.synthetic import time_course
    coords, growth, _ = time_course(dim=10, n_components=4, seed=3, counts=[120, 140])
    x = coords[0].copy()
    x[0] += 8.0
    sched = dict(SCHED, epsilon=0.01, epsilon0=0.01)
    err, res = _compare(torch, ot, dev, large, x, coords[1], growth[0], 6, sched)
    assert np.isfinite(res.f.cpu().numpy()).all()
    assert err < 1e-5, err


def test_fused_odd_sizes_and_dims(mods):
    torch, ot, dev, large = mods
    from wot_amd.synthetic import day_pair
    dp = day_pair(97, 131, dim=7, n_components=3, seed=5)
    err, _ = _compare(torch, ot, dev, large, dp.x, dp.y, dp.growth, 6, SCHED)
    assert err < 1e-5, err


def test_fused_row_sums_and_growth_loop(mods):
    """Row sums of the returned plan (the next growth iteration's G) and the
    growth loop against the stored-cost solver's loop."""
    torch, ot, dev, large = mods
    from wot_amd.synthetic import day_pair
    dp = day_pair(300, 420, dim=30, seed=9)
    res, learned = large.transport_map_coords(dp.x, dp.y, dp.growth, growth_iters=3, terms=6, **SCHED)
    cost = ot.DeviceCost.from_coordinates(dp.x, dp.y)
    R, lg = ot.compute_transport_matrix(solver=ot.transport_stablev2, C=cost, G=dp.growth, growth_iters=3, **SCHED)
    want = R.sum(1)
    got = res.row_sums.cpu().numpy()
    assert np.abs(got - want).max() / np.abs(want).max() < 1e-6
    for a, b in zip(learned[1:], lg[1:]):
        assert np.abs(a.cpu().numpy() - b).max() / np.abs(b).max() < 1e-6


def _vs_oracle(torch, large, x, y, G, terms, sched):
    """Plan of the never-stored solve (rows regenerated with the FP64 cost)
    against the FP64 oracle's transport_stablev2 on the reference cost."""
    from oracle import ot_oracle as O
    res = large.transport_stablev2_coords(x, y, G, terms=terms, **sched)
    R = res.plan_rows(x, y, torch.arange(x.shape[0], device='cuda')).cpu().numpy()
    Ro = O.transport_stablev2(O.default_cost_matrix(x, y), G=G, **sched)
    err = np.linalg.norm(R - Ro) / np.linalg.norm(Ro)
    marg = max(np.abs(R.sum(a) - Ro.sum(a)).max() / np.abs(Ro.sum(a)).max() for a in (0, 1))
    return err, marg


# The cross term x'.y' comes from split-BF16 MFMA products accumulated in two
# FP32 chains (hi.hi, and the smaller products) that are added in FP64: about
# 1e-8 RMS (1.5e-7 max) absolute error in the normalised cost, which enters
# each plan entry as |dC| / eps.  At eps = 0.05 the plan and marginals meet
# the 1e-5 / 1e-6 gate down to d = 7; at eps = 0.01 the amplification is 5x
# larger and that case is held to 1e-5 on the marginals (the stored-cost
# solver keeps the FP64 cost and meets 1e-6 there: tests/test_gpu_parity.py).
@pytest.mark.parametrize('n_i,n_j,dim,seed,marg_tol', [(700, 900, 50, 41, 1e-6), (97, 131, 30, 5, 1e-6),
                                                       (97, 131, 7, 5, 1e-6)])
def test_fused_vs_oracle(mods, n_i, n_j, dim, seed, marg_tol):
    """The never-stored solve (6-term split-BF16 cross term) against the FP64
    oracle restating wot/ot/optimal_transport.py:167-236 on the same inputs."""
    torch, ot, dev, large = mods
    from wot_amd.synthetic import day_pair
    dp = day_pair(n_i, n_j, dim=dim, n_components=5, seed=seed)
    err, marg = _vs_oracle(torch, large, dp.x, dp.y, dp.growth, 6, SCHED)
    print('fused vs oracle %dx%d d=%d: rel frob %.3e marginals %.3e' % (n_i, n_j, dim, err, marg))
    assert err <= 1e-5 and marg <= marg_tol


@pytest.mark.parametrize('n_i,n_j,dim', [(1, 5, 3), (5, 1, 3), (3, 2, 7), (33, 17, 30), (40, 200, 50)])
def test_fused_small_shapes_vs_oracle(mods, n_i, n_j, dim):
    """Pairs below one 32-row MFMA tile and one 32-column tile (single rows and
    columns, clamped indices and masked entries) against the FP64 oracle."""
    torch, ot, dev, large = mods
    from wot_amd.synthetic import day_pair
    dp = day_pair(n_i, n_j, dim=dim, n_components=3, seed=17)
    err, marg = _vs_oracle(torch, large, dp.x, dp.y, dp.growth, 6, SCHED)
    print('fused vs oracle %dx%d d=%d: rel frob %.3e marginals %.3e' % (n_i, n_j, dim, err, marg))
    assert err <= 1e-5 and marg <= 1e-6, (err, marg)


def test_fused_outlier_vs_oracle(mods):
    """eps0 = eps = 0.01 with an outlier source cell (its row is recomputed by
    the repair kernel), against the FP64 oracle."""
    torch, ot, dev, large = mods
    from wot_amd.synthetic import time_course
    coords, growth, _ = time_course(dim=10, n_components=4, seed=3, counts=[120, 140])
    x = coords[0].copy()
    x[0] += 8.0
    sched = dict(SCHED, epsilon=0.01, epsilon0=0.01)
    err, marg = _vs_oracle(torch, large, x, coords[1], growth[0], 6, sched)
    print('fused vs oracle, eps = 0.01 with an outlier: rel frob %.3e marginals %.3e' % (err, marg))
    assert err <= 1e-5 and marg <= 1e-5, (err, marg)  # eps = 0.01: see the note above test_fused_vs_oracle


def test_fused_nan_growth_propagates(mods):
    """A NaN growth entry gives NaN potentials for its row (the reference's
    arithmetic propagates it), never stale values."""
    torch, ot, dev, large = mods
    from wot_amd.synthetic import day_pair
    dp = day_pair(64, 80, dim=8, n_components=3, seed=12)
    G = dp.growth.copy()
    G[5] = np.nan
    res = large.transport_stablev2_coords(dp.x, dp.y, G, terms=6, want_row_sums=True,
                                          **dict(SCHED, scaling_iter=60, extra_iter=10))
    f = res.f.cpu().numpy()
    rs = res.row_sums.cpu().numpy()
    assert np.isnan(f[5]) and np.isnan(rs[5])


GAP = dict(lambda1=1.0, lambda2=50.0, epsilon=0.05, batch_size=5, tolerance=1e-8, tau=1e4, epsilon0=1.0,
           max_iter=1e7)


@pytest.mark.parametrize('n_i,n_j,dim,seed', [(300, 360, 30, 14), (120, 150, 10, 15)])
def test_fused_duality_gap_vs_oracle(mods, n_i, n_j, dim, seed):
    """The duality-gap schedule on the never-stored cost (wot/ot/optimal_transport.py:67-164)
    against the FP64 oracle on the reference cost: plan, marginals, converged
    gap, and the iteration count within one batch."""
    torch, ot, dev, large = mods
    from oracle import ot_oracle as O
    from wot_amd.synthetic import day_pair
    dp = day_pair(n_i, n_j, dim=dim, n_components=5, seed=seed)
    st = {}
    res = ot.optimal_transport_duality_gap(None, dp.growth, coords=(dp.x, dp.y), stats=st, **GAP)
    R = res.dense().cpu().numpy()
    so = {}
    Ro = O.optimal_transport_duality_gap(O.default_cost_matrix(dp.x, dp.y), dp.growth, stats=so, **GAP)
    err = np.linalg.norm(R - Ro) / np.linalg.norm(Ro)
    marg = max(np.abs(R.sum(a) - Ro.sum(a)).max() / np.abs(Ro.sum(a)).max() for a in (0, 1))
    print('fused duality gap vs oracle %dx%d: rel frob %.3e marginals %.3e, iterations %d vs %d, gap %.3e vs %.3e'
          % (n_i, n_j, err, marg, st['iterations'], so['iterations'], st['gap'], so['gap']))
    assert err <= 1e-5 and marg <= 1e-6
    assert 0.0 <= st['gap'] <= GAP['tolerance']
    assert abs(st['iterations'] - so['iterations']) <= GAP['batch_size']


def test_fused_duality_gap_max_iter_early_return(mods, caplog):
    """max_iter inside the dual-evolution scalings: the reference's warning and
    its un-normalised plan (optimal_transport.py:143-145), against the oracle
    stopped at the same iteration."""
    torch, ot, dev, large = mods
    import logging
    from oracle import ot_oracle as O
    from wot_amd.synthetic import day_pair
    dp = day_pair(90, 110, dim=12, n_components=4, seed=16)
    prm = dict(GAP, max_iter=37)
    with caplog.at_level(logging.WARNING, logger='wot'):
        res = ot.optimal_transport_duality_gap(None, dp.growth, coords=(dp.x, dp.y), **prm)
    assert 'Reached max_iter' in caplog.text
    assert res.stats['max_iter_reached'] and res.stats['iterations'] == 37
    R = res.dense().cpu().numpy()
    Ro = O.optimal_transport_duality_gap(O.default_cost_matrix(dp.x, dp.y), dp.growth, **prm)
    assert np.linalg.norm(R - Ro) / np.linalg.norm(Ro) <= 1e-5
