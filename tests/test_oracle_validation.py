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


def _draws(p_flat, seed, size):
    np.random.seed(seed)
    return np.random.choice(len(p_flat), p=p_flat, size=size)


@pytest.mark.parametrize('frac', [0.0, 0.3, 0.5, 1.0])
def test_interpolation_weights_from_the_formula(frac):
    """interpolate_with_ot (wot/ot/util.py:140-148) samples pair (i, j) with
    probability proportional to T_ij / colsum_j^(1 - frac).  A 2 x 3 map with
    unequal column AND row sums, the weights computed by hand here, must give
    the oracle's exact draw stream (a wrong axis or exponent changes it)."""
    T = np.array([[1.0, 2.0, 0.5], [3.0, 6.0, 4.5]])   # column sums 4, 8, 5; row sums 3.5, 13.5
    colsum = np.array([4.0, 8.0, 5.0])
    w = np.array([[T[i, j] / colsum[j] ** (1.0 - frac) for j in range(3)] for i in range(2)])
    p = (w / w.sum()).ravel()
    p0 = np.array([[0.0, 0.0], [10.0, 0.0]])
    p1 = np.array([[0.0, 1.0], [0.0, 2.0], [0.0, 4.0]])
    idx = _draws(p, 11, 500)
    np.random.seed(11)
    got = V.interpolate_with_ot(p0, p1, T, frac, 500)
    want = np.array([p0[k // 3] * (1 - frac) + p1[k % 3] * frac for k in idx])
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-15)
    # the wrong axis (row sums) draws a different stream for frac < 1
    if frac < 1.0:
        rowsum = T.sum(1)
        w_bad = T / rowsum[:, None] ** (1.0 - frac)
        assert not np.array_equal(_draws((w_bad / w_bad.sum()).ravel(), 11, 500), idx)


def test_interpolation_limits():
    """frac = 1: p is proportional to T; frac = 0: every target column gets the
    same total sampling mass 1/J, whatever its column sum."""
    rng = np.random.default_rng(8)
    T = rng.random((4, 5)) * np.array([1.0, 10.0, 0.1, 3.0, 7.0])
    for frac, p_want in [(1.0, T / T.sum()), (0.0, (T / T.sum(0)) / 5.0)]:
        w = T / np.power(T.sum(axis=0), 1.0 - frac)
        np.testing.assert_allclose(w / w.sum(), p_want, rtol=1e-14)
        if frac == 0.0:
            np.testing.assert_allclose((w / w.sum()).sum(0), np.full(5, 0.2), rtol=1e-14)
        # and the oracle draws from exactly that distribution
        p0 = rng.standard_normal((4, 2))
        p1 = rng.standard_normal((5, 2))
        idx = _draws(p_want.ravel(), 21, 300)
        np.random.seed(21)
        got = V.interpolate_with_ot(p0, p1, T, frac, 300)
        want = np.array([p0[k // 5] * (1 - frac) + p1[k % 5] * frac for k in idx])
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-14)
