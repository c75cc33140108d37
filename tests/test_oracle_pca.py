"""Pins the local-PCA oracle (oracle/pca_oracle.py) with sklearn and numpy."""
import numpy as np
import pytest

from oracle import pca_oracle as P


def _days(rng, n1, n2, g, rank=6):
    W = rng.standard_normal((rank, g))
    a = rng.standard_normal((n1, rank)) @ W + 0.05 * rng.standard_normal((n1, g))
    b = rng.standard_normal((n2, rank)) @ W + 0.05 * rng.standard_normal((n2, g)) + 0.3
    return a, b


@pytest.mark.parametrize('shape', [(30, 40, 200), (120, 90, 50)])
def test_sklearn_exact_solver_equals_numpy_svd_of_double_centered(shape):
    rng = np.random.default_rng(1)
    a, b = _days(rng, *shape)
    k = 5
    p1, p2, pca, mean = P.reference_pca(a, b, k, svd_solver='full')
    coords_sk = np.vstack([p1, p2]) * pca.singular_values_
    coords, S, V = P.exact_coordinates(a, b, k)
    np.testing.assert_allclose(pca.singular_values_, S, rtol=1e-10)
    np.testing.assert_allclose(np.abs(np.vstack([p1, p2])), np.abs(V), atol=1e-9)
    np.testing.assert_allclose(P.sqdist(coords_sk), P.sqdist(coords), rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(mean, np.vstack([a, b]).mean(0), rtol=1e-14)


def test_double_centering_invariants_and_spectrum():
    rng = np.random.default_rng(2)
    a, b = _days(rng, 20, 25, 60)
    A = P.double_centered(a, b)
    np.testing.assert_allclose(A.mean(0), 0, atol=1e-12)   # per gene
    np.testing.assert_allclose(A.mean(1), 0, atol=1e-12)   # per cell
    _, S, _ = P.exact_coordinates(a, b, 8)
    w = np.sort(np.linalg.eigvalsh(A @ A.T))[::-1][:8]
    np.testing.assert_allclose(S ** 2, w, rtol=1e-9)


def test_reference_call_small_input_matches_exact():
    """The reference's own call ('auto' solver): on inputs this small sklearn
    picks an exact solver, so its distances equal the exact ones."""
    rng = np.random.default_rng(3)
    a, b = _days(rng, 40, 50, 120)
    p1, p2, pca, _ = P.reference_pca(a, b, 4)
    coords, S, _ = P.exact_coordinates(a, b, 4)
    np.testing.assert_allclose(pca.singular_values_, S, rtol=1e-8)
    Z = np.vstack([p1, p2]) * pca.singular_values_
    np.testing.assert_allclose(P.sqdist(Z), P.sqdist(coords), rtol=1e-7, atol=1e-8)
