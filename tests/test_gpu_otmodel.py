"""OTModel drop-in (wot/ot/ot_model.py) end to end on the GPU: the reference's
own tests (tests/test_transport.py), the PCA path, learned growth obs, file
output and configuration errors, against the FP64 oracle."""
import os

import numpy as np
import pandas as pd
import pytest

from oracle import ot_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def amd():
    import torch
    assert torch.cuda.is_available()
    import wot_amd.ot as ot
    from wot_amd import _anndata
    return ot, _anndata


def _adata(_anndata, X, days, growth=None):
    obs = pd.DataFrame({'day': days}, index=['c%d' % i for i in range(len(days))])
    if growth is not None:
        obs['cell_growth_rate'] = growth
    var = pd.DataFrame(index=['f%d' % k for k in range(X.shape[1])])
    return _anndata.AnnData(X, obs, var)


def test_reference_costs_single_tmap(amd):
    """tests/test_transport.py:11-34 of the reference."""
    ot, _anndata = amd
    adata = _adata(_anndata, np.random.rand(6, 1000), [1, 1, 1, 2, 2, 2])
    cost = np.array([[0, 100, 100], [100, 0, 100], [100, 100, 0]])
    model = ot.OTModel(adata, epsilon=0.01, lambda1=1, lambda2=50)
    tmap = model.compute_transport_map(1, 2, cost_matrix=cost)
    assert np.allclose(tmap.X, np.eye(3), atol=0.01, rtol=0)


def test_reference_costs_multiple_tmaps(amd, tmp_path):
    """tests/test_transport.py:36-65 of the reference (npz output: no h5py here)."""
    ot, _anndata = amd
    from wot_amd.io import read_npz_tmap
    adata = _adata(_anndata, np.random.rand(9, 1000), [1, 1, 1, 2, 2, 2, 3, 3, 3])
    cost = np.array([[0, 100, 100], [100, 0, 100], [100, 100, 0]])
    model = ot.OTModel(adata, epsilon=0.01, lambda1=1, lambda2=50)
    out = os.path.join(str(tmp_path), 'tmaps_out', 'tmap')
    model.compute_all_transport_maps(tmap_out=out, cost_matrices=[cost, cost], output_file_format='npz')
    for name in ('tmap_1_2.npz', 'tmap_2_3.npz'):
        X, obs, var = read_npz_tmap(os.path.join(str(tmp_path), 'tmaps_out', name))
        assert np.allclose(X, np.eye(3), atol=0.01, rtol=0)


def test_otmodel_coordinates_growth_loop(amd, tmp_path):
    """PCA coordinates given directly (local_pca=0), fixed_iters, growth_iters=3:
    plan and obs g0..g3 against the oracle; learned growth file."""
    ot, _anndata = amd
    from wot_amd.synthetic import day_pair
    dp = day_pair(300, 350, dim=30, seed=12)
    X = np.vstack([dp.x, dp.y])
    days = [0.0] * 300 + [1.0] * 350
    growth = np.concatenate([dp.growth, np.ones(350)])
    adata = _adata(_anndata, X, days, growth)
    model = ot.OTModel(adata, local_pca=0, solver='fixed_iters', growth_iters=3, scaling_iter=600, extra_iter=100)
    tmap = model.compute_transport_map(0.0, 1.0)
    C = O.default_cost_matrix(dp.x, dp.y)
    prm = dict(model.ot_config)
    R, learned = O.compute_transport_matrix(O.transport_stablev2, **dict(prm, C=C, G=dp.growth))
    err = np.linalg.norm(tmap.X - R) / np.linalg.norm(R)
    assert err < 1e-5, err
    want = O.growth_obs(learned, R, 1.0)
    assert list(tmap.obs.columns) == ['g0', 'g1', 'g2', 'g3']
    for k, v in want.items():
        np.testing.assert_allclose(tmap.obs[k].values, v, rtol=1e-6)
    assert list(tmap.obs.index) == ['c%d' % i for i in range(300)]
    assert list(tmap.var.index) == ['c%d' % i for i in range(300, 650)]
    # compute_all_transport_maps writes the learned growth file when growth_iters > 1
    out = os.path.join(str(tmp_path), 'tm')
    model.compute_all_transport_maps(tmap_out=out, output_file_format='npz')
    g = pd.read_csv(out + '_g.txt', sep='\t', index_col='id')
    assert list(g.columns) == ['g0', 'g1', 'g2', 'g3'] and len(g) == 300


def test_otmodel_local_pca_duality_gap(amd):
    """Default path: local PCA (device) + duality-gap solver, against the
    oracle pipeline with the reference's own sklearn PCA call."""
    ot, _anndata = amd
    rng = np.random.default_rng(3)
    genes = 200
    base = rng.random((1, genes)) * 3
    X = np.vstack([base + rng.standard_normal((120, genes)), base + 0.3 + rng.standard_normal((140, genes))])
    adata = _adata(_anndata, X, [1.0] * 120 + [2.0] * 140)
    model = ot.OTModel(adata)
    tmap = model.compute_transport_map(1.0, 2.0)
    from oracle import pca_oracle
    p0, p1, pca, _ = pca_oracle.reference_pca(X[:120], X[120:], 30)
    eig = np.diag(pca.singular_values_)
    C = O.default_cost_matrix(p0, p1, eig)
    R = O.optimal_transport_duality_gap(C, np.ones(120), **model.ot_config)
    err = np.linalg.norm(tmap.X - R) / np.linalg.norm(R)
    assert err < 1e-5, err


def test_otmodel_errors(amd):
    ot, _anndata = amd
    adata = _adata(_anndata, np.random.rand(6, 5), [1, 1, 1, 2, 2, 2])
    with pytest.raises(ValueError, match='Unknown solver'):
        ot.OTModel(adata, solver='nope')
    cfg = pd.DataFrame({'t0': [1.0], 't1': [2.0], 'epsilon': [0.1]})
    model = ot.OTModel(adata, config=cfg, local_pca=0)
    with pytest.raises(ValueError, match='not present in day_pairs'):
        model.compute_transport_map(2.0, 3.0)
    assert model.day_pairs == {(1.0, 2.0): {'epsilon': 0.1}}
    assert ot.OTModel(adata, local_pca=0).compute_transport_map(1, 5) is None  # no cells at day 5


def test_compute_all_transport_maps_batched_matches_sequential(amd, tmp_path):
    """fixed_iters, per-pair epsilon / lambda1 / lambda2 from a day-pair
    configuration, growth_iters=2: the batched run (all pairs in one device
    run) writes the same maps, learned growth and _g.txt as the pair-by-pair
    solves, and each map matches the FP64 oracle's growth loop."""
    ot, _anndata = amd
    from wot_amd.io import read_npz_tmap
    from wot_amd.synthetic import time_course
    coords, growth, counts = time_course(n_days=4, min_cells=60, max_cells=160, dim=8, n_components=3, seed=21)
    X = np.vstack(coords)
    days = np.repeat(np.arange(4, dtype=float), counts)
    adata = _adata(_anndata, X, days, growth=np.concatenate(growth))
    cfg = tmp_path / 'pairs.txt'
    cfg.write_text('t0\tt1\tepsilon\tlambda1\tlambda2\n0\t1\t0.05\t1\t50\n1\t2\t0.1\t2\t40\n2\t3\t0.03\t0.5\t60\n')
    prm = dict(local_pca=0, solver='fixed_iters', growth_iters=2, scaling_iter=400, extra_iter=100,
               config=str(cfg))
    out_b = os.path.join(str(tmp_path), 'b', 'tm')
    out_s = os.path.join(str(tmp_path), 's', 'tm')
    ot.OTModel(adata, **prm).compute_all_transport_maps(tmap_out=out_b, output_file_format='npz', batch=True)
    ot.OTModel(adata, **prm).compute_all_transport_maps(tmap_out=out_s, output_file_format='npz', batch=False)
    per_pair = {(0.0, 1.0): (0.05, 1.0, 50.0), (1.0, 2.0): (0.1, 2.0, 40.0), (2.0, 3.0): (0.03, 0.5, 60.0)}
    for t in range(3):
        name = '_{}_{}.npz'.format(float(t), float(t + 1))
        Xb, obs_b, var_b = read_npz_tmap(out_b + name)
        Xs, obs_s, var_s = read_npz_tmap(out_s + name)
        assert list(obs_b.index) == list(obs_s.index) and list(var_b) == list(var_s)
        assert np.linalg.norm(Xb - Xs) / np.linalg.norm(Xs) <= 1e-6
        for g in ('g0', 'g1', 'g2'):
            assert np.abs(obs_b[g].values - obs_s[g].values).max() / np.abs(obs_s[g].values).max() <= 1e-6
        eps, l1, l2 = per_pair[(float(t), float(t + 1))]
        Ro, _ = O.compute_transport_matrix(O.transport_stablev2, C=O.default_cost_matrix(coords[t], coords[t + 1]),
                                           G=growth[t], growth_iters=2, epsilon=eps, lambda1=l1, lambda2=l2,
                                           epsilon0=1.0, tau=1e4, scaling_iter=400, extra_iter=100,
                                           inner_iter_max=50)
        assert np.linalg.norm(Xb - Ro) / np.linalg.norm(Ro) <= 1e-5
        for a in (0, 1):
            assert np.abs(Xb.sum(a) - Ro.sum(a)).max() / np.abs(Ro.sum(a)).max() <= 1e-6
    gb = pd.read_csv(out_b + '_g.txt', sep='\t', index_col=0)
    gs = pd.read_csv(out_s + '_g.txt', sep='\t', index_col=0)
    assert list(gb.index) == list(gs.index) and list(gb.columns) == ['g0', 'g1', 'g2']
    np.testing.assert_allclose(gb.values, gs.values, rtol=1e-6)


def test_compute_all_transport_maps_batch_refusals(amd, tmp_path):
    """batch=True with the duality-gap solver or given cost matrices raises;
    'auto' falls back to the pair-by-pair loop (same files)."""
    ot, _anndata = amd
    adata = _adata(_anndata, np.random.rand(9, 5), [1, 1, 1, 2, 2, 2, 3, 3, 3])
    with pytest.raises(ValueError, match="cannot batch"):
        ot.OTModel(adata, local_pca=0).compute_all_transport_maps(tmap_out=str(tmp_path / 'a' / 't'),
                                                                  output_file_format='npz', batch=True)
    ot.OTModel(adata, local_pca=0).compute_all_transport_maps(tmap_out=str(tmp_path / 'b' / 't'),
                                                              output_file_format='npz')
    assert sorted(os.listdir(str(tmp_path / 'b'))) == ['t_1_2.npz', 't_2_3.npz']
