"""Local PCA on the device (wot_amd.ot.pca, libwotk wotk_pca_center) against
the numpy / sklearn oracle (oracle/pca_oracle.py)."""
import numpy as np
import pytest

from oracle import pca_oracle as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def pca():
    import torch
    assert torch.cuda.is_available()
    from wot_amd.ot import pca
    return pca


def _days(rng, n1, n2, g, rank=6, noise=0.05):
    W = rng.standard_normal((rank, g))
    a = rng.standard_normal((n1, rank)) @ W + noise * rng.standard_normal((n1, g))
    b = rng.standard_normal((n2, rank)) @ W + noise * rng.standard_normal((n2, g)) + 0.3
    return a, b


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize('shape', [(1, 1, 3), (7, 5, 300), (130, 70, 97), (300, 257, 1031),
                                   (300, 257, 1100), (40, 33, 2052),  # 16-byte rows: ragged chunks and row blocks
                                   (1500, 1200, 260), (1500, 1200, 97),  # 85 streaming row blocks: two-level column means
                                   (600000, 500000, 4)])  # 68,750 16-row copy blocks on a 65,535-block grid
@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_center_kernel(pca, shape, dtype):
    import torch
    rng = np.random.default_rng(sum(shape))
    n1, n2, g = shape
    x = rng.random((n1 + n2, g)).astype(dtype)
    A, mu, nu = pca.center(torch.from_numpy(x).cuda())
    x64 = x.astype(np.float64)
    want_mu = x64.mean(0)
    want_A = P.double_centered(x64[:n1], x64[n1:])
    np.testing.assert_allclose(mu.cpu().numpy(), want_mu, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(A.cpu().numpy(), want_A, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(nu.cpu().numpy(), (x64 - want_mu).mean(1), rtol=1e-10, atol=1e-13)


@pytest.mark.parametrize('shape', [(30, 40, 200), (120, 90, 50), (500, 600, 2000)])
@pytest.mark.parametrize('method', ['exact', 'subspace'])
def test_fit_matches_exact_svd(pca, shape, method):
    """Singular values and the cells' pairwise squared distances (sign-free)
    against numpy's SVD of the doubly-centered matrix."""
    rng = np.random.default_rng(7)
    a, b = _days(rng, *shape)
    k = 5
    p1, p2, res, mean = pca.compute_pca(a, b, k, method=method)
    coords, S, V = P.exact_coordinates(a, b, k)
    assert rel(res.singular_values_, S) < 1e-10
    Z = np.vstack([p1, p2]) * res.singular_values_
    assert rel(P.sqdist(Z), P.sqdist(coords)) < 1e-9
    # with the same sign convention the coordinates themselves agree
    assert rel(np.vstack([p1, p2]), V) < 1e-8
    np.testing.assert_allclose(mean, np.vstack([a, b]).mean(0), rtol=1e-12)


def test_reference_call_parity(pca):
    """The reference's own call (sklearn 'auto', random_state 58951) on a
    day pair: cost-relevant distances agree."""
    rng = np.random.default_rng(8)
    a, b = _days(rng, 150, 180, 400)
    p1, p2, res, _ = pca.compute_pca(a, b, 30)
    r1, r2, rp, _ = P.reference_pca(a, b, 30)
    assert rel(res.singular_values_, rp.singular_values_) < 1e-8
    Z = np.vstack([p1, p2]) * res.singular_values_
    R = np.vstack([r1, r2]) * rp.singular_values_
    assert rel(P.sqdist(Z), P.sqdist(R)) < 1e-7


def test_float32_input_and_errors(pca):
    rng = np.random.default_rng(9)
    a, b = _days(rng, 60, 70, 90)
    p1, p2, res, _ = pca.compute_pca(a.astype(np.float32), b.astype(np.float32), 4)
    coords, S, _ = P.exact_coordinates(a.astype(np.float32), b.astype(np.float32), 4)
    assert rel(res.singular_values_, S) < 1e-9
    with pytest.raises(ValueError, match='n_components'):
        pca.compute_pca(a[:, :3], b[:, :3], 4)
