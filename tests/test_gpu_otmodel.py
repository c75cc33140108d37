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
