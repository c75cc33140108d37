"""Pins the validation oracle and the exact EMD (libwotk's network simplex,
host code, runs without a GPU) against scipy's exact solvers."""
import numpy as np
import pytest

from oracle import validation_oracle as V


def test_emd_network_simplex_vs_assignment():
    from wot_amd.ot.emd import emd2
    rng = np.random.default_rng(0)
    for n in (1, 2, 9, 64, 300):
        x = rng.standard_normal((n, 4))
        y = rng.standard_normal((n, 4)) + 0.4
        M = V.sqeuclid(x, y)
        got = emd2(np.ones(n) / n, np.ones(n) / n, M)
        assert got == pytest.approx(V.emd2_assignment(M), rel=1e-12, abs=1e-15), n


@pytest.mark.parametrize('shape', [(1, 5), (4, 3), (17, 29), (60, 45)])
def test_emd_network_simplex_vs_lp(shape):
    from wot_amd.ot.emd import emd2
    rng = np.random.default_rng(sum(shape))
    n, m = shape
    a = rng.random(n) + 0.01
    b = rng.random(m) + 0.01
    a /= a.sum()
    b /= b.sum()
    M = rng.random((n, m)) * 10
    got, F = emd2(a, b, M, return_plan=True)
    assert got == pytest.approx(V.emd2_lp(a, b, M), rel=1e-10)
    assert (F >= -1e-15).all()
    np.testing.assert_allclose(F.sum(1), a, atol=1e-14)
    np.testing.assert_allclose(F.sum(0), b, atol=1e-14)
    assert float((F * M).sum()) == pytest.approx(got, rel=1e-13)


def test_emd_one_dimensional_monotone_coupling():
    """Squared distance on the line: the sorted (monotone) coupling is optimal."""
    from wot_amd.ot.emd import emd2
    rng = np.random.default_rng(4)
    x = np.sort(rng.standard_normal(50))
    y = np.sort(rng.standard_normal(50) * 2 + 1)
    M = (x[:, None] - y[None, :]) ** 2
    assert emd2(np.ones(50) / 50, np.ones(50) / 50, M) == pytest.approx(np.mean((x - y) ** 2), rel=1e-12)


def test_emd_errors_and_iteration_limit(caplog):
    from wot_amd.ot.emd import emd2
    with pytest.raises(ValueError, match='sum'):
        emd2(np.array([0.5, 0.6]), np.array([1.0]), np.ones((2, 1)))
    with pytest.raises(ValueError, match='negative'):
        emd2(np.array([-0.5, 1.5]), np.array([1.0]), np.ones((2, 1)))
    rng = np.random.default_rng(5)
    M = rng.random((30, 30))
    u = np.ones(30) / 30
    capped, F = emd2(u, u, M, numItermax=3, return_plan=True)
    assert 'numItermax' in caplog.text
    assert F.sum() < 1.0  # stopped early: mass still on the start arcs
    assert capped == pytest.approx(float((F * M).sum()), rel=1e-13)


def test_interpolation_oracle_semantics():
    """Uniform weights: interpolate_with_ot with an all-ones map draws like
    interpolate_randomly_with_growth with g = 1 (same stream, same pairs)."""
    rng = np.random.default_rng(6)
    p0 = rng.standard_normal((5, 3))
    p1 = rng.standard_normal((4, 3))
    np.random.seed(3)
    a = V.interpolate_with_ot(p0, p1, np.ones((5, 4)), 0.5, 40)
    np.random.seed(3)
    b = V.interpolate_randomly_with_growth(p0, p1, 0.5, 40, np.ones(5))
    np.testing.assert_array_equal(a, b)
    # every sample lies on a segment between a p0 and a p1 cell at the fraction
    for row in a:
        d = np.array([[np.abs(row - (x * 0.5 + y * 0.5)).max() for y in p1] for x in p0])
        assert d.min() < 1e-14
