"""OT validation on the device (wot_amd.ot.validation: interpolation kernels,
exact EMD on the device cost) against the numpy / scipy oracle."""
import numpy as np
import pandas as pd
import pytest

from oracle import pca_oracle as PO
from oracle import validation_oracle as V

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def val():
    import torch
    assert torch.cuda.is_available()
    from wot_amd.ot import validation
    return validation


def test_interpolation_same_stream_same_cells(val):
    rng = np.random.default_rng(1)
    p0 = rng.standard_normal((37, 6))
    p1 = rng.standard_normal((53, 6)) + 1.0
    tmap = rng.random((37, 53)) ** 4
    g = np.exp(0.3 * rng.standard_normal(37))
    for seed in (0, 7):
        np.random.seed(seed)
        a = [val.interpolate_with_ot(p0, p1, tmap, 0.4, 500), val.interpolate_randomly(p0, p1, 0.4, 300),
             val.interpolate_randomly_with_growth(p0, p1, 0.4, 200, g)]
        np.random.seed(seed)
        b = [V.interpolate_with_ot(p0, p1, tmap, 0.4, 500), V.interpolate_randomly(p0, p1, 0.4, 300),
             V.interpolate_randomly_with_growth(p0, p1, 0.4, 200, g)]
        for x, y in zip(a, b):
            assert x.shape == y.shape
            np.testing.assert_allclose(x, y, rtol=1e-15, atol=1e-15)


def test_interpolation_errors(val):
    with pytest.raises(ValueError, match='genes do not match'):
        val.interpolate_with_ot(np.ones((2, 3)), np.ones((2, 4)), np.ones((2, 2)), 0.5, 4)
    with pytest.raises(ValueError, match='Tmap size'):
        val.interpolate_with_ot(np.ones((2, 3)), np.ones((2, 3)), np.ones((3, 2)), 0.5, 4)
    with pytest.raises(ValueError, match='NaN'):
        val.interpolate_with_ot(np.ones((2, 3)), np.ones((2, 3)), np.zeros((2, 2)), 0.5, 4)


def test_earth_mover_distance(val):
    from wot_amd.ot.emd import earth_mover_distance
    rng = np.random.default_rng(2)
    c1 = rng.standard_normal((40, 5))
    c2 = rng.standard_normal((55, 5)) + 0.5
    eig = np.diag(rng.random(5) + 0.5)
    w1 = rng.random(40)
    got = earth_mover_distance(c1, c2, eig, weights1=w1)
    M = V.sqeuclid(c1.dot(eig), c2.dot(eig))
    want = np.sqrt(V.emd2_lp(w1 / w1.sum(), np.ones(55) / 55, M))
    assert got == pytest.approx(want, rel=1e-10)


def _model(n_per_day=(40, 50, 45), genes=20, covariates=1, growth=True, local_pca=3):
    import wot_amd.ot as ot
    from wot_amd import _anndata
    rng = np.random.default_rng(3)
    W = rng.standard_normal((4, genes))
    X = np.vstack([rng.standard_normal((n, 4)) @ W + 0.2 * t + 0.1 * rng.standard_normal((n, genes))
                   for t, n in enumerate(n_per_day)])
    days = np.concatenate([[float(t)] * n for t, n in enumerate(n_per_day)])
    obs = pd.DataFrame({'day': days, 'covariate': rng.integers(0, covariates, len(days))},
                       index=['c%d' % i for i in range(len(days))])
    if growth:
        obs['cell_growth_rate'] = np.exp(0.2 * rng.standard_normal(len(days)))
    adata = _anndata.AnnData(X, obs, pd.DataFrame(index=['g%d' % k for k in range(genes)]))
    return ot.OTModel(adata, local_pca=local_pca, solver='fixed_iters', scaling_iter=300, extra_iter=0), X, obs


def test_summary_single_covariate_matches_oracle(val):
    """One triplet, one covariate, no growth: rows I, R, F, L in the
    reference's order; every distance against the oracle pipeline on the same
    maps and the same numpy stream."""
    model, X, obs = _model(growth=False)
    np.random.seed(11)
    df = val.compute_validation_summary(model, interp_size=60)
    assert list(df['name']) == ['I', 'R', 'F', 'L']
    assert set(df['interval_mid']) == {1.0} and not df['full'].any()
    d0, d05, d1 = (obs['day'] == 0).values, (obs['day'] == 1).values, (obs['day'] == 2).values
    p1, p2, pca, mean = PO.reference_pca(X[d0], X[d1], 3, svd_solver='full')
    S = pca.singular_values_
    x05 = V.project_middle(X[d0], X[d1], X[d05], pca.components_, S, mean)
    eig = np.diag(S)
    tmap = model.compute_transport_map(0.0, 2.0, covariate=(0, 0)).X
    np.random.seed(11)
    i05 = V.interpolate_with_ot(p1, p2, tmap, 0.5, 60)
    r05 = V.interpolate_randomly(p1, p2, 0.5, 60)

    def emd(a, b):
        M = V.sqeuclid(a.dot(eig), b.dot(eig))
        return np.sqrt(V.emd2_lp(np.ones(len(a)) / len(a), np.ones(len(b)) / len(b), M))

    want = [emd(i05, x05), emd(r05, x05), emd(p1, x05), emd(p2, x05)]
    np.testing.assert_allclose(df['distance'].values, want, rtol=1e-6)


def test_summary_covariates_growth_and_full(val):
    """Two covariates, growth, full distances: the reference's row set."""
    model, X, obs = _model(covariates=2, growth=True, local_pca=0)
    np.random.seed(5)
    df = val.compute_validation_summary(model, interp_size=40, compute_full_distances=True)
    full = df[df['full']]
    assert list(full['name']) == ['Rg', 'R', 'I', 'I1', 'I2', 'F', 'L', 'A']
    part = df[~df['full']]
    assert set(part['name']) == {'P', 'I', 'R', 'Rg', 'F', 'L'}
    assert (df['distance'] > 0).all() and np.isfinite(df['distance']).all()
    # deterministic rows do not depend on the stream
    np.random.seed(6)
    df2 = val.compute_validation_summary(model, interp_size=40, compute_full_distances=True)
    det = df['name'].isin(['P', 'F', 'L', 'A']).values
    np.testing.assert_allclose(df['distance'].values[det], df2['distance'].values[det], rtol=1e-12)


def test_summary_emds_on_host_threads_match_in_line(val):
    """The EMDs run on a pool of host threads; with the same seed the
    summary equals the one computed with every EMD in line, row for row."""
    model, X, obs = _model(covariates=2, growth=True, local_pca=0)
    np.random.seed(7)
    a = val.compute_validation_summary(model, interp_size=60, compute_full_distances=True, emd_workers=1)
    np.random.seed(7)
    b = val.compute_validation_summary(model, interp_size=60, compute_full_distances=True, emd_workers=4)
    assert list(a['name']) == list(b['name'])
    assert b['distance'].dtype == np.float64
    np.testing.assert_array_equal(a['distance'].values, b['distance'].values)
