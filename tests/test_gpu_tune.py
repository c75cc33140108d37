"""Every compiled tile variant of the two sweep kernels gives the oracle's
plans (they differ only in the order of FP32 partial sums inside a row), and
the autotuner times all variants, installs the fastest and persists it."""
import json

import numpy as np
import pytest

from oracle import ot_oracle as O

pytestmark = pytest.mark.gpu

SCHED = dict(scaling_iter=300, extra_iter=100, inner_iter_max=50, tau=1e4)


@pytest.fixture(scope='module')
def tune():
    import torch
    assert torch.cuda.is_available()
    from wot_amd import tune as t
    yield t
    t.reset()


@pytest.fixture(autouse=True)
def private_cache(tmp_path, monkeypatch):
    monkeypatch.setenv('WOTK_TILE_CACHE', str(tmp_path / 'tiles.json'))


def _course(seed, counts, dim=10):
    from wot_amd.synthetic import time_course
    coords, growth, _ = time_course(dim=dim, n_components=4, seed=seed, counts=counts)
    return coords, growth


def test_every_row_sweep_variant_matches_oracle(tune):
    import wot_amd.ot as ot
    from wot_amd.synthetic import day_pair
    dp = day_pair(333, 517, dim=12, n_components=4, seed=21)  # ragged: neither side a multiple of any tile
    C = O.default_cost_matrix(dp.x, dp.y)
    Rref = O.transport_stablev2(C, 1.0, 50.0, 0.05, G=dp.growth, epsilon0=1.0, **SCHED)
    cost = ot.DeviceCost.from_coordinates(dp.x, dp.y)
    plans = {}
    for tile in tune.variants(tune.ROW_SWEEP):
        tune.set_tile(tune.ROW_SWEEP, -1, tile)
        R = ot.transport_stablev2(cost, G=dp.growth, epsilon=0.05, lambda1=1.0, lambda2=50.0, epsilon0=1.0,
                                  **SCHED)
        err = np.linalg.norm(R - Rref) / np.linalg.norm(Rref)
        marg = max(np.abs(R.sum(a) - Rref.sum(a)).max() / np.abs(Rref.sum(a)).max() for a in (0, 1))
        assert err <= 1e-5 and marg <= 1e-6, (tile, err, marg)
        plans[tile] = R
    tune.reset()
    base = plans[tune.variants(tune.ROW_SWEEP)[0]]
    for tile, R in plans.items():
        assert np.linalg.norm(R - base) / np.linalg.norm(base) <= 1e-6, tile


def test_every_batch_variant_matches_oracle(tune):
    from wot_amd.ot import batch
    coords, growth = _course(5, [141, 263, 90])
    grid = dict(epsilons=(0.01, 0.1), lambda1s=(1.0, 5.0))
    refs = {}
    for p in range(2):
        C = O.default_cost_matrix(coords[p], coords[p + 1])
        for e, (eps, lam1) in enumerate([(a, b) for a in grid['epsilons'] for b in grid['lambda1s']]):
            refs[p, e] = O.transport_stablev2(C, lam1, 50.0, eps, G=growth[p], epsilon0=1.0, **SCHED)
    for kernel, share in ((tune.BATCH_SWEEP, False), (tune.BATCH_SHARED, True)):
        for tile in tune.variants(kernel):
            tune.set_tile(kernel, -1, tile)
            res = batch.solve_time_course(coords, growth, graph_iters=50, share_exp=share, **grid, **SCHED)
            assert res.stats['shared_group'] == (2 if share else 0)
            for (p, e), Rref in refs.items():
                R = res.plan(p, e)
                err = np.linalg.norm(R - Rref) / np.linalg.norm(Rref)
                assert err <= 1e-5, (kernel, tile, p, e, err)
    tune.reset()


def test_autotune_row_sweep_installs_and_persists(tune, tmp_path):
    import wot_amd.ot as ot
    from wot_amd.synthetic import day_pair
    dp = day_pair(700, 900, dim=10, n_components=4, seed=3)
    cost = ot.DeviceCost.from_coordinates(dp.x, dp.y)
    r = tune.autotune_row_sweep(cost, iters=5, repeats=1)
    assert set(r.timings) == set(tune.variants(tune.ROW_SWEEP))
    assert all(v > 0 for v in r.timings.values())
    assert r.best == min(r.timings, key=r.timings.get)
    for b in r.buckets:
        assert tune.get_tile(tune.ROW_SWEEP, b) == r.best
    data = json.load(open(tmp_path / 'tiles.json'))
    keys = [k for k in data['entries'] if k.split('|')[1] == 'row_sweep']
    assert len(keys) == len(r.buckets)
    # a fresh load restores the tuned tile
    tune.reset()
    assert tune.apply_persisted(force=True) == len(r.buckets)
    for b in r.buckets:
        assert tune.get_tile(tune.ROW_SWEEP, b) == r.best
    tune.reset()


def test_autotune_time_course(tune):
    from wot_amd.ot import batch
    coords, growth = _course(6, [300, 420, 380])
    tc = batch.prepare_time_course(coords, growth)
    r = tune.autotune_time_course(tc, (0.05, 0.1), (1.0,), iters=4, repeats=1)  # no equal epsilons
    assert r.kernel == tune.BATCH_SWEEP
    assert set(r.timings) == set(tune.variants(tune.BATCH_SWEEP))
    assert tune.get_tile(tune.BATCH_SWEEP, r.buckets[0]) == r.best
    assert r.buckets[0] == tune.bucket(tune.BATCH_SWEEP, 0, tc.mean_row_length(), 2)
    r = tune.autotune_time_course(tc, (0.05,), (1.0, 5.0), iters=4, repeats=1)  # one group of 2
    assert r.kernel == tune.BATCH_SHARED
    assert set(r.timings) == set(tune.variants(tune.BATCH_SHARED))
    assert tune.get_tile(tune.BATCH_SHARED, r.buckets[0]) == r.best
    tune.reset()


def test_every_shared_variant_repairs_fallback_tiles(tune):
    """An outlier source cell with eps = eps0 = 0.01 makes its tile's fixed-shift
    sums underflow: with every tile shape of the shared-exponential kernel
    (register and LDS-staged, 1 to 16 rows per tile) the repair kernel must
    recompute those tiles and the plans must match the oracle."""
    from wot_amd.ot import batch
    coords, growth = _course(3, [120, 140])
    coords[0] = coords[0].copy()
    coords[0][0] += 8.0
    C = O.default_cost_matrix(coords[0], coords[1])
    lams = (0.5, 1.0, 5.0)
    refs = [O.transport_stablev2(C, l, 50.0, 0.01, G=growth[0], epsilon0=0.01, **SCHED) for l in lams]
    for tile in tune.variants(tune.BATCH_SHARED):
        tune.set_tile(tune.BATCH_SHARED, -1, tile)
        res = batch.solve_time_course(coords, growth, epsilons=(0.01,), lambda1s=lams, epsilon0=0.01,
                                      graph_iters=0, **SCHED)
        assert res.stats['shared_group'] == 3 and res.stats['exact_rows'] > 0, tile
        for e, Rref in enumerate(refs):
            R = res.plan(0, e)
            err = np.linalg.norm(R - Rref) / np.linalg.norm(Rref)
            assert err <= 1e-5, (tile, e, err)
    tune.reset()
