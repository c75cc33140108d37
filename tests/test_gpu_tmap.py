"""Populations through device-resident transport maps (libwotk
wotk_tmap_propagate via wot_amd.tmap) against the FP64 oracle."""
import os

import numpy as np
import pandas as pd
import pytest

from oracle import tmap_oracle as O

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(__file__), 'golden', 'transport_maps')


@pytest.fixture(scope='module')
def tm():
    import torch
    assert torch.cuda.is_available()
    import wot_amd.tmap as tm
    return tm


def _dev(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _walk(tm, maps_np, P, direction, normalize, dtype='float64', outs_all=False, transposed=False):
    """P: k x cells (reference layout).  Returns k x cells after the walk (and every step if outs_all)."""
    import torch
    dms = [tm.DeviceMap.upload(m, dtype) for m in maps_np]
    if transposed:
        for d in dms:
            d.make_transposed()
    Pd = _dev(P.T.copy())
    k = P.shape[0]
    if outs_all:
        outs = [torch.empty(((m.shape[1] if direction == tm.PUSH_FORWARD else m.shape[0]), k), dtype=torch.float64,
                            device='cuda') for m in maps_np]
        tm.propagate(dms, Pd, direction, normalize=normalize, outs=outs)
        return [o.cpu().numpy().T for o in outs]
    return tm.propagate(dms, Pd, direction, normalize=normalize).cpu().numpy().T


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def fixture_maps():
    return [pd.read_csv(os.path.join(FIX, 't%d_t%d.csv' % (d, d + 1)), index_col='id').values.astype(np.float64)
            for d in range(4)]


def test_fixture_maps_exact(tm):
    """The reference's integer maps: unnormalised walks are exact integers;
    every single-cell population at every day, pushed to the end and pulled
    to the start."""
    maps = fixture_maps()
    sizes = [maps[0].shape[0]] + [m.shape[1] for m in maps]
    for d in range(5):
        P = np.eye(sizes[d])
        if d < 4:
            got = _walk(tm, maps[d:], P, tm.PUSH_FORWARD, False)
            assert np.array_equal(got, O.push_forward(maps, P, d, 4, normalize=False)), d
            got = _walk(tm, maps[d:], P, tm.PUSH_FORWARD, True)
            np.testing.assert_allclose(got, O.push_forward(maps, P, d, 4), rtol=1e-15, atol=0)
        if d > 0:
            back = maps[:d][::-1]
            got = _walk(tm, back, P, tm.PULL_BACK, False)
            assert np.array_equal(got, O.pull_back(maps, P, d, 0, normalize=False)), d
            got = _walk(tm, back, P, tm.PULL_BACK, True)
            np.testing.assert_allclose(got, O.pull_back(maps, P, d, 0), rtol=1e-15, atol=0)


SHAPES = [(1, 1), (5, 3), (63, 64), (64, 65), (65, 63), (130, 257), (1000, 700), (2049, 1531)]


@pytest.mark.parametrize('dtype', ['float64', 'float32'])
@pytest.mark.parametrize('k', [1, 3, 16, 17, 40])
def test_single_map_random(tm, dtype, k):
    """One map, every direction and normalisation; FP32 maps are compared
    with the oracle on the same FP32-rounded entries (the kernel widens them
    exactly, so only the summation order differs)."""
    rng = np.random.default_rng(100 + k)
    for (a, b) in SHAPES:
        M = rng.random((a, b)) ** 3
        Mref = M.astype(np.float32).astype(np.float64) if dtype == 'float32' else M
        P = rng.random((k, a))
        Q = rng.random((k, b))
        for normalize in (False, True):
            got = _walk(tm, [M], P, tm.PUSH_FORWARD, normalize, dtype)
            e = rel(got, O.push_forward([Mref], P, 0, 1, normalize))
            assert e < 1e-13, ('push', a, b, k, normalize, e)
            want = O.pull_back([Mref], Q, 1, 0, normalize)
            got = _walk(tm, [M], Q, tm.PULL_BACK, normalize, dtype)
            e = rel(got, want)
            assert e < 1e-13, ('pull', a, b, k, normalize, e)
            got = _walk(tm, [M], Q, tm.PULL_BACK, normalize, dtype, transposed=True)
            e = rel(got, want)
            assert e < 1e-13, ('pull via transposed copy', a, b, k, normalize, e)


def test_chain_ragged_days_all_steps(tm):
    rng = np.random.default_rng(11)
    sizes = [300, 1, 77, 1024, 513, 64, 900]
    maps = [rng.random((sizes[t], sizes[t + 1])) for t in range(len(sizes) - 1)]
    P = rng.random((5, sizes[0]))
    steps = _walk(tm, maps, P, tm.PUSH_FORWARD, True, outs_all=True)
    for s, got in enumerate(steps):
        assert rel(got, O.push_forward(maps, P, 0, s + 1)) < 1e-13, s
    Q = rng.random((5, sizes[-1]))
    last = len(sizes) - 1
    for transposed in (False, True):
        steps = _walk(tm, maps[::-1], Q, tm.PULL_BACK, True, outs_all=True, transposed=transposed)
        for s, got in enumerate(steps):
            assert rel(got, O.pull_back(maps, Q, last, last - s - 1)) < 1e-13, (s, transposed)


def test_transpose_is_exact(tm):
    import torch
    rng = np.random.default_rng(14)
    for (a, b) in [(1, 1), (31, 33), (64, 64), (1000, 77), (129, 2050)]:
        for dt in ('float64', 'float32'):
            d = tm.DeviceMap.upload(rng.random((a, b)), dt).make_transposed()
            assert d.XT.shape == (b, a)
            assert torch.equal(d.XT, d.X.T)


def test_strided_maps_take_the_scalar_path(tm):
    """Maps whose rows are not 16-byte aligned (odd leading dimension) read
    element by element; results agree with the oracle."""
    import torch
    rng = np.random.default_rng(12)
    for (a, b, ld) in [(100, 130, 131), (257, 64, 67)]:
        big = rng.random((a, ld))
        M = big[:, :b]
        Pd = _dev(rng.random((a, 7)))
        for dt, tdt in (('float64', torch.float64), ('float32', torch.float32)):
            X = _dev(big).to(tdt)[:, :b]
            dm = tm.DeviceMap(X)
            assert dm.X.stride(0) == ld
            Mref = M.astype(np.float32).astype(np.float64) if dt == 'float32' else M
            got = tm.propagate([dm], Pd, tm.PUSH_FORWARD, normalize=False).cpu().numpy()
            want = O.push_forward([Mref], Pd.cpu().numpy().T, 0, 1, normalize=False).T
            assert rel(got, want) < 1e-13
            Qd = _dev(rng.random((b, 3)))
            got = tm.propagate([dm], Qd, tm.PULL_BACK, normalize=True).cpu().numpy()
            want = O.pull_back([Mref], Qd.cpu().numpy().T, 1, 0).T
            assert rel(got, want) < 1e-13


def test_deterministic(tm):
    rng = np.random.default_rng(13)
    M = rng.random((3000, 2500))
    P = rng.random((9, 3000))
    a = _walk(tm, [M], P, tm.PUSH_FORWARD, True)
    b = _walk(tm, [M], P, tm.PUSH_FORWARD, True)
    assert np.array_equal(a, b)
    Q = rng.random((9, 2500))
    assert np.array_equal(_walk(tm, [M], Q, tm.PULL_BACK, True), _walk(tm, [M], Q, tm.PULL_BACK, True))


def test_zero_population_is_nan_like_numpy(tm):
    M = np.ones((4, 3))
    P = np.array([[1.0, 0, 0, 0], [0, 0, 0, 0]])
    got = _walk(tm, [M], P, tm.PUSH_FORWARD, True)
    np.testing.assert_allclose(got[0], [1 / 3] * 3)
    assert np.isnan(got[1]).all()


def test_large_map_adjoint_and_mass(tm):
    """15000 x 15000 FP64 map (1.8 GB): <pull(q), p> = <q, push(p)> and the
    pushed mass equals sum_i p_i * rowsum_i (FP64 device reductions)."""
    import torch
    g = torch.Generator(device='cuda').manual_seed(5)
    n = 15000
    X = torch.rand((n, n), dtype=torch.float64, device='cuda', generator=g)
    dm = tm.DeviceMap(X)
    P = torch.rand((n, 10), dtype=torch.float64, device='cuda', generator=g)
    Q = torch.rand((n, 10), dtype=torch.float64, device='cuda', generator=g)
    push = tm.propagate([dm], P, tm.PUSH_FORWARD, normalize=False)
    pull = tm.propagate([dm], Q, tm.PULL_BACK, normalize=False)
    lhs = (pull * P).sum(0)
    rhs = (Q * push).sum(0)
    assert torch.allclose(lhs, rhs, rtol=1e-12, atol=0)
    dm.make_transposed()
    pull_t = tm.propagate([dm], Q, tm.PULL_BACK, normalize=False)
    assert torch.allclose(pull_t, pull, rtol=1e-12, atol=0)
    dm.XT = None
    mass = (P * X.sum(1, keepdim=True)).sum(0)
    assert torch.allclose(push.sum(0), mass, rtol=1e-12, atol=0)
    # a sampled block against a library FP64 product
    ref = (P[:, :3].T @ X[:, :256]).T
    assert torch.allclose(push[:256, :3], ref, rtol=1e-12, atol=0)
    del X


# ---------------------------------------------------------------------------
# TransportMapModel over the device maps
# ---------------------------------------------------------------------------
def _course(rng, sizes):
    maps = [rng.random((sizes[t], sizes[t + 1])) ** 2 for t in range(len(sizes) - 1)]
    ids = [['d%d_c%d' % (t, i) for i in range(n)] for t, n in enumerate(sizes)]
    meta = pd.DataFrame({'day': np.concatenate([[float(t)] * n for t, n in enumerate(sizes)])},
                        index=sum(ids, []))
    from wot_amd import _anndata
    tmaps = {(float(t), float(t + 1)): _anndata.AnnData(maps[t], pd.DataFrame(index=ids[t]),
                                                        pd.DataFrame(index=ids[t + 1]))
             for t in range(len(maps))}
    return maps, meta, tmaps, ids


@pytest.mark.parametrize('dtype,transposed', [('float64', True), ('float32', True), ('float64', False)])
def test_model_walks_match_oracle(tm, dtype, transposed):
    rng = np.random.default_rng(21)
    sizes = [120, 95, 200, 64, 150]
    maps, meta, tmaps, ids = _course(rng, sizes)
    mref = [m.astype(np.float32).astype(np.float64) for m in maps] if dtype == 'float32' else maps
    model = tm.TransportMapModel(tmaps, meta, dtype=dtype, transposed_maps=transposed)
    pops = [tm.Population(1.0, rng.random(95), 'a'), tm.Population(1.0, (rng.random(95) > 0.7) * 1.0, 'b')]
    P = np.vstack([p.p for p in pops])
    out = model.push_forward(*pops, to_time=4.0)
    assert isinstance(out, list) and len(out) == 2 and out[0].time == 4.0
    assert rel(np.vstack([p.p for p in out]), O.push_forward(mref, P, 1, 4)) < 1e-13
    one = model.push_forward(pops[0])
    assert isinstance(one, tm.Population) and one.time == 2.0
    assert rel(one.p, O.push_forward(mref, P[:1], 1, 2)[0]) < 1e-13
    back = model.pull_back(*pops, normalize=False, as_list=True)
    assert rel(np.vstack([p.p for p in back]), O.pull_back(mref, P, 1, 0, normalize=False)) < 1e-13
    same = model.push_forward(pops[0], to_time=1.0)  # no map: the input, unnormalised
    np.testing.assert_array_equal(same.p, pops[0].p)
    anc = model.ancestors(*pops, at_time=0.0)
    desc = model.descendants(*pops, at_time=3.0)
    assert rel(np.vstack([p.p for p in anc]), O.pull_back(mref, P, 1, 0)) < 1e-13
    assert rel(np.vstack([p.p for p in desc]), O.push_forward(mref, P, 1, 3)) < 1e-13

    day3 = [tm.Population(3.0, (rng.random(64) > 0.5) * 1.0, 's%d' % i) for i in range(3)]
    P3 = np.vstack([p.p for p in day3])
    f = model.fates(day3)
    want = O.fates(mref, P3, 3)
    assert f.X.shape == want.shape and rel(f.X, want) < 1e-13
    assert list(f.var.index)[:3] == ['s0', 's1', 's2']
    assert len(f.obs) == sum(sizes[:4])

    traj = model.trajectories(day3)
    want = O.trajectories(mref, P3, 3)
    assert traj.X.shape == want.shape and rel(traj.X, want) < 1e-13
    assert list(traj.var.index) == ['s0', 's1', 's2']

    start = [tm.Population(1.0, (rng.random(95) > 0.6) * 1.0, 'x%d' % i) for i in range(2)]
    tt = model.transition_table(start, day3)
    want = O.transition_table(mref, np.vstack([p.p for p in start]), 1, P3, 3)
    assert tt.X.shape == want.shape and rel(tt.X, want) < 1e-13
    assert list(tt.obs.index)[:2] == ['x0', 'x1']

    from wot_amd import _anndata
    csm = _anndata.AnnData((rng.random((sum(sizes), 4)) > 0.5) * 1.0, pd.DataFrame(index=meta.index),
                           pd.DataFrame(index=['A', 'B', 'C', 'D']))
    tps, census = model.ancestor_census(csm, *pops)
    assert tps == [0.0, 1.0, 2.0, 3.0, 4.0]
    assert census.shape == (2, 5, 4)
    offs = np.cumsum([0] + sizes)
    for d in range(5):
        Pd = O.pull_back(mref, P, 1, d) if d < 1 else (O.push_forward(mref, P, 1, d) if d > 1 else P)
        want = O.census(Pd, csm.X[offs[d]:offs[d + 1]])
        assert rel(census[:, d, :], want) < 1e-12, d


def test_model_on_fixture_maps_by_hand(tm):
    maps = fixture_maps()
    sizes = [4, 2, 5, 6, 3]
    ids = [['c%d-t%d' % (i + 1, t) for i in range(n)] for t, n in enumerate(sizes)]
    meta = pd.DataFrame({'day': np.concatenate([[t] * n for t, n in enumerate(sizes)])}, index=sum(ids, []))
    from wot_amd import _anndata
    tmaps = {(t, t + 1): _anndata.AnnData(maps[t], pd.DataFrame(index=ids[t]), pd.DataFrame(index=ids[t + 1]))
             for t in range(4)}
    model = tm.TransportMapModel(tmaps, meta)
    pop = model.population_from_ids(['c1-t0'], at_time=0)[0]
    two = model.push_forward(pop, to_time=2)
    np.testing.assert_allclose(two.p, np.array([5, 6, 7, 8, 9]) / 35, rtol=0, atol=1e-15)
    q = model.population_from_ids(['c1-t1'], at_time=1)[0]
    np.testing.assert_allclose(model.pull_back(q).p, [0.1, 0.2, 0.3, 0.4], rtol=0, atol=1e-15)


def test_model_from_time_course(tm):
    """Maps materialised on the device from a batched solve feed the model
    directly; results match the oracle on the same (downloaded) maps."""
    from wot_amd.ot import batch
    rng = np.random.default_rng(31)
    sizes = [150, 180, 120, 160]
    coords = [rng.standard_normal((n, 5)) + 0.3 * t for t, n in enumerate(sizes)]
    growth = [np.exp(0.2 * rng.standard_normal(n)) for n in sizes]
    res = batch.solve_time_course(coords, growth, epsilons=(0.05,), lambda1s=(1.0,), scaling_iter=200,
                                  extra_iter=0, inner_iter_max=50)
    ids = [['t%d_%d' % (t, i) for i in range(n)] for t, n in enumerate(sizes)]
    meta = pd.DataFrame({'day': np.concatenate([[float(t)] * n for t, n in enumerate(sizes)])}, index=sum(ids, []))
    model = tm.TransportMapModel.from_time_course(res, [0.0, 1.0, 2.0, 3.0], meta)
    maps = [res.plan(p, 0) for p in range(3)]
    P = rng.random((2, 150))
    out = model.push_forward(tm.Population(0.0, P[0]), tm.Population(0.0, P[1]), to_time=3.0)
    assert rel(np.vstack([p.p for p in out]), O.push_forward(maps, P, 0, 3)) < 1e-13
    cp = model.get_coupling(1.0, 2.0)
    np.testing.assert_array_equal(cp.X, maps[1])
    assert list(cp.obs.index[:2]) == ['t1_0', 't1_1']


def test_otmodel_h5ad_course_through_from_directory(tm, tmp_path):
    """OTModel writes a 3-day course as h5ad (the default format, through the
    HDF5 C library); TransportMapModel.from_directory reads ids and days from
    the files and walks the maps on the device; results match the oracle on
    the maps read back."""
    import wot_amd.ot as ot
    from wot_amd import _anndata, io
    rng = np.random.default_rng(41)
    sizes = [60, 80, 70]
    X = np.vstack([rng.standard_normal((n, 5)) + 0.2 * t for t, n in enumerate(sizes)])
    days = np.concatenate([[float(t)] * n for t, n in enumerate(sizes)])
    obs = pd.DataFrame({'day': days}, index=['c%d' % i for i in range(len(days))])
    adata = _anndata.AnnData(X, obs, pd.DataFrame(index=['f%d' % k for k in range(5)]))
    model = ot.OTModel(adata, local_pca=0, solver='fixed_iters', scaling_iter=300, extra_iter=0)
    out = str(tmp_path / 'tmaps')
    model.compute_all_transport_maps(tmap_out=out)
    paths = [out + '_0.0_1.0.h5ad', out + '_1.0_2.0.h5ad']
    maps = [io.read_dataset(p).X for p in paths]
    assert maps[0].shape == (60, 80) and maps[1].shape == (80, 70)
    tmm = tm.TransportMapModel.from_directory(out)
    assert tmm.timepoints == [0.0, 1.0, 2.0]
    assert list(tmm.meta.index) == list(obs.index)
    pops = tmm.population_from_cell_sets({'a': ['c%d' % i for i in range(60, 100)],
                                          'b': ['c%d' % i for i in range(100, 140)]}, at_time=1)
    P = np.vstack([p.p for p in pops])
    traj = tmm.trajectories(pops)
    assert rel(traj.X, O.trajectories(maps, P, 1)) < 1e-13
    f = tmm.fates(pops)
    assert rel(f.X, O.fates(maps, P, 1)) < 1e-13


def test_model_get_coupling_chains_in_fp64(tm):
    """Non-consecutive pairs chain the device maps in FP64 like np.dot
    (chaining.py:4-40): equal to the oracle chain to rounding."""
    from oracle import ot_oracle
    rng = np.random.default_rng(51)
    sizes = [90, 70, 110, 60]
    maps, meta, tmaps, ids = _course(rng, sizes)
    model = tm.TransportMapModel(tmaps, meta)
    cp = model.get_coupling(0.0, 3.0)
    want = ot_oracle.chain(maps)
    assert cp.X.shape == (90, 60)
    assert rel(cp.X, want) < 1e-14
    assert list(cp.obs.index) == ids[0] and list(cp.var.index) == ids[3]
    cp = model.get_coupling(1.0, 3.0)
    assert rel(cp.X, ot_oracle.chain(maps[1:])) < 1e-14
