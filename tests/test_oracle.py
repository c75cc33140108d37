"""Pins the FP64 oracle (oracle/ot_oracle.py) independently of its own text:

* golden vectors written from the reference solver file (tests/golden/);
* library routines: scipy cdist + numpy median for the cost;
* closed forms: the I=J=1 fixed point; the balanced limit (lambda -> inf) is
  textbook Sinkhorn, checked against an independent tiny implementation;
* weak duality primal >= dual at any dual-feasible point and strong duality at
  convergence;
* brute force: on tiny problems the converged plan minimises the primal
  objective (scipy.optimize on the plan entries).
"""
import glob
import json
import os

import numpy as np
import pytest

from oracle import ot_oracle as O

GOLD = os.path.join(os.path.dirname(__file__), 'golden')


def _load(name):
    z = np.load(os.path.join(GOLD, name + '.npz'), allow_pickle=False)
    return {k: z[k] for k in z.files}


def _params(z):
    return json.loads(str(z['params']))


@pytest.mark.parametrize('n', [0, 1, 2])
def test_objective_golden(n):
    z = _load('objective_%d' % n)
    args = (z['C'], z['K'], z['R'], np.ones(len(z['p'])) / len(z['p']), np.ones(len(z['q'])) / len(z['q']),
            z['p'], z['q'], z['a'], z['b'], float(z['eps']), float(z['lambda1']), float(z['lambda2']))
    with np.errstate(all='ignore'):
        pr, du = O.primal(*args), O.dual(*args)
    # case 0 is 1x1 with R=0: the reference yields NaN (0*log 0 in fdiv)
    np.testing.assert_allclose(pr, float(z['primal']), rtol=1e-13, equal_nan=True)
    np.testing.assert_allclose(du, float(z['dual']), rtol=1e-13, equal_nan=True)


SV2 = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLD, 'sv2_*.npz')))
GAP = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLD, 'gap_*.npz')))


@pytest.mark.parametrize('name', SV2 + ['diag3_stablev2'])
def test_stablev2_golden(name):
    z = _load(name)
    R = O.transport_stablev2(C=z['C'], G=z['G'], **_params(z))
    np.testing.assert_allclose(R, z['R'], rtol=1e-9, atol=1e-14 * np.abs(z['R']).max())


@pytest.mark.parametrize('name', GAP + ['diag3_gap'])
def test_duality_gap_golden(name):
    z = _load(name)
    R = O.optimal_transport_duality_gap(C=z['C'], G=z['G'], **_params(z))
    np.testing.assert_allclose(R, z['R'], rtol=1e-9, atol=1e-14 * np.abs(z['R']).max())


@pytest.mark.parametrize('name,solver', [('ctm_stablev2', O.transport_stablev2),
                                         ('ctm_gap', O.optimal_transport_duality_gap)])
def test_growth_loop_golden(name, solver):
    z = _load(name)
    params = dict(_params(z), C=z['C'], G=z['G'])
    R, learned = O.compute_transport_matrix(solver=solver, **params)
    np.testing.assert_allclose(R, z['R'], rtol=1e-9, atol=1e-14)
    np.testing.assert_allclose(np.stack(learned), z['learned'], rtol=1e-9)
    # `**params` hands the function its own dict: the assignment at
    # optimal_transport.py:28 never reaches the caller's G
    np.testing.assert_array_equal(params['G'], z['G_after'])
    np.testing.assert_array_equal(z['G_after'], z['G'])


def test_reference_identity_case():
    """tests/test_transport.py:11-34 of the reference: diagonal cost -> identity."""
    C = np.array([[0, 100, 100], [100, 0, 100], [100, 100, 0]])
    R = O.optimal_transport_duality_gap(C, np.ones(3), 1, 50, 0.01, 5, 1e-8, 1e4, 1, 1e7)
    np.testing.assert_allclose(R, np.eye(3), atol=0.01, rtol=0)


def test_cost_matches_library():
    rng = np.random.default_rng(5)
    x, y = rng.standard_normal((30, 7)), rng.standard_normal((19, 7))  # 570 values: even count
    c = ((x[:, None, :] - y[None, :, :]) ** 2).sum(-1)  # brute force
    ref = c / np.median(c)
    np.testing.assert_allclose(O.default_cost_matrix(x, y), ref, rtol=1e-13)
    # even count -> mean of the two middle order statistics
    v = np.sort(c.ravel())
    assert np.median(c) == 0.5 * (v[len(v) // 2 - 1] + v[len(v) // 2])


def test_one_by_one_closed_form():
    """I=J=1: fixed point a = (p/(K b))^al1, b = (q/(K a))^al2 with q=p.
    In log form f = al1 (eps log p - (g - c)) ... solved in closed form:
    f + g - c = eps * (al1 al2 ... ) -> plan R = exp((f+g-c)/eps)."""
    c, p, eps, l1, l2 = 0.7, 1.3, 0.05, 1.0, 50.0
    a1, a2 = l1 / (l1 + eps), l2 / (l2 + eps)
    # f = a1 (eps log p - g + c), g = a2 (eps log p - f + c)  (dx=dy=1)
    A = np.array([[1.0, a1], [a2, 1.0]])
    rhs = np.array([a1 * (eps * np.log(p) + c), a2 * (eps * np.log(p) + c)])
    f, g = np.linalg.solve(A, rhs)
    R_fp = np.exp((f + g - c) / eps)
    st = {}
    R = O.transport_stablev2(np.array([[c]]), l1, l2, eps, 3000, np.array([p]), 1e4, 1.0, 1000, 50, stats=st)
    assert R[0, 0] == pytest.approx(R_fp, rel=1e-10)
    R2 = O.optimal_transport_duality_gap(np.array([[c]]), np.array([p]), l1, l2, eps, 5, 1e-12, 1e4, 1.0, 1e7)
    # the gap solver's final epsilon is eps0*eps = eps here; it stops on a
    # relative duality gap <= 1e-12, which bounds the plan error only to
    # first order in sqrt(gap)
    assert R2[0, 0] == pytest.approx(R_fp, rel=1e-6)


def _balanced_sinkhorn(C, p, q, eps, n):
    K = np.exp(-C / eps)
    b = np.ones(len(q))
    for _ in range(n):
        a = p / K.dot(b)
        b = q / K.T.dot(a)
    return a[:, None] * K * b[None, :]


def test_balanced_limit():
    """lambda1, lambda2 -> inf: alpha -> 1 and the update is textbook Sinkhorn
    on marginals p*I and q*J (dx = 1/I, dy = 1/J absorbed)."""
    rng = np.random.default_rng(3)
    n_i, n_j = 9, 11
    C = rng.random((n_i, n_j))
    G = np.ones(n_i)
    eps = 0.2
    R = O.transport_stablev2(C, 1e12, 1e12, eps, 2000, G, 1e4, eps, 0, 50)  # eps0 = eps: no schedule
    # with alpha = 1 the unnormalised plan diag(a) K diag(b) has row sums p*J
    # and column sums q*I; the solver returns it divided by J
    ref = _balanced_sinkhorn(C, G * n_j, np.ones(n_j) * G.mean() * n_i, eps, 2000) / n_j
    np.testing.assert_allclose(R, ref, rtol=1e-7, atol=1e-14)


def test_weak_and_strong_duality():
    rng = np.random.default_rng(11)
    n_i, n_j, eps, l1, l2 = 6, 7, 0.3, 1.0, 5.0
    C = rng.random((n_i, n_j))
    p = rng.random(n_i) + 0.5
    q = np.ones(n_j) * p.mean()
    dx, dy = np.ones(n_i) / n_i, np.ones(n_j) / n_j
    Kt = np.exp(-C / eps)
    # strong duality at the converged plan; a, b are the true scalings
    # exp(f/eps), exp(g/eps) of the converged potentials
    st = {}
    R = O.transport_stablev2(C, l1, l2, eps, 4000, p, 1e4, eps, 0, 50, stats=st) * n_j
    a, b = np.exp(st['u'] / eps), np.exp(st['v'] / eps)
    assert np.allclose(a[:, None] * Kt * b[None, :], R, rtol=1e-10)
    pri = O.primal(C, Kt, R, dx, dy, p, q, a, b, eps, l1, l2)
    dua = O.dual(C, Kt, R, dx, dy, p, q, a, b, eps, l1, l2)
    assert abs(pri - dua) / abs(pri) < 1e-9
    # weak duality at a perturbed dual point: dual(a', b') <= primal(R*) ... the
    # dual objective is evaluated with R built from the same (a', b')
    for s in range(5):
        a2 = a * np.exp(0.3 * rng.standard_normal(n_i))
        b2 = b * np.exp(0.3 * rng.standard_normal(n_j))
        R2 = a2[:, None] * Kt * b2[None, :]
        assert O.dual(C, Kt, R2, dx, dy, p, q, a2, b2, eps, l1, l2) <= pri + 1e-12


def test_brute_force_minimiser():
    """Tiny problem: the converged plan minimises primal() over all positive
    plans (scipy L-BFGS on log R)."""
    from scipy.optimize import minimize
    rng = np.random.default_rng(17)
    n_i, n_j, eps, l1, l2 = 3, 4, 0.5, 1.0, 3.0
    C = rng.random((n_i, n_j))
    p = rng.random(n_i) + 0.5
    q = np.ones(n_j) * p.mean()
    dx, dy = np.ones(n_i) / n_i, np.ones(n_j) / n_j
    Kt = np.exp(-C / eps)

    def obj(z):
        R = np.exp(z.reshape(n_i, n_j))
        return O.primal(C, Kt, R, dx, dy, p, q, None, None, eps, l1, l2)

    res = minimize(obj, np.zeros(n_i * n_j), method='L-BFGS-B', options=dict(ftol=1e-15, gtol=1e-12,
                                                                              maxiter=20000))
    R_bf = np.exp(res.x.reshape(n_i, n_j))
    R = O.transport_stablev2(C, l1, l2, eps, 5000, p, 1e4, eps, 0, 50) * n_j
    np.testing.assert_allclose(R, R_bf, rtol=2e-4)
    assert obj(np.log(R).ravel()) <= res.fun + 1e-12


def test_chain_integer_fixtures():
    """Chain of the integer maps tests/inputs/transport_maps/t*_t*.csv shapes
    (4x2, 2x5, 5x6, 6x3): the chained product equals the brute-force
    triple-loop product exactly (integers)."""
    rng = np.random.default_rng(2)
    shapes = [(4, 2), (2, 5), (5, 6), (6, 3)]
    maps = [rng.integers(0, 9, s).astype(np.float64) for s in shapes]
    out = O.chain(maps)
    bf = maps[0]
    for m in maps[1:]:
        nxt = np.zeros((bf.shape[0], m.shape[1]))
        for i in range(bf.shape[0]):
            for j in range(m.shape[1]):
                for k in range(m.shape[0]):
                    nxt[i, j] += bf[i, k] * m[k, j]
        bf = nxt
    assert np.array_equal(out, bf)


def test_adjoint_gradient_matches_finite_differences():
    """Envelope-theorem gradient of the converged objective (oracle) against
    central finite differences of the converged FP64 objective, h = 1e-4."""
    rng = np.random.default_rng(21)
    n_i, n_j, d = 9, 11, 3
    x = rng.standard_normal((n_i, d))
    y = rng.standard_normal((n_j, d)) + 0.3
    lg = 0.3 * rng.standard_normal(n_i)
    prm = dict(lambda1=1.0, lambda2=5.0, epsilon=0.2, scaling_iter=3000, epsilon0=0.2, extra_iter=0,
               inner_iter_max=50)

    def obj(xx, ll):
        return O.objective_and_grad(xx, ll, y, **prm)[0]

    val, gx, glg, _ = O.objective_and_grad(x, lg, y, **prm)
    h = 1e-4
    for (i, k) in [(0, 0), (3, 1), (5, 2), (8, 0), (2, 2)]:
        xp, xm = x.copy(), x.copy()
        xp[i, k] += h
        xm[i, k] -= h
        fd = (obj(xp, lg) - obj(xm, lg)) / (2 * h)
        assert fd == pytest.approx(gx[i, k], rel=1e-5, abs=1e-9), (i, k, fd, gx[i, k])
    for i in [0, 4, 7]:
        lp, lm = lg.copy(), lg.copy()
        lp[i] += h
        lm[i] -= h
        fd = (obj(x, lp) - obj(x, lm)) / (2 * h)
        assert fd == pytest.approx(glg[i], rel=1e-5, abs=1e-9), (i, fd, glg[i])


def test_adjoint_median_subgradient_matters():
    """Perturbing a middle element moves the median: the FD derivative of the
    objective along that coordinate includes the median term."""
    rng = np.random.default_rng(5)
    n_i, n_j, d = 6, 8, 2
    x = rng.standard_normal((n_i, d))
    y = rng.standard_normal((n_j, d))
    lg = np.zeros(n_i)
    prm = dict(lambda1=1.0, lambda2=5.0, epsilon=0.3, scaling_iter=2000, epsilon0=0.3, extra_iter=0,
               inner_iter_max=50)
    D = O.sqeuclidean(x, y)
    lo, hi = O.middle_indices(D)
    k, _ = divmod(lo, n_j)
    _, gx, _, _ = O.objective_and_grad(x, lg, y, **prm)
    h = 1e-5
    xp, xm = x.copy(), x.copy()
    xp[k, 0] += h
    xm[k, 0] -= h
    fd = (O.objective_and_grad(xp, lg, y, **prm)[0] - O.objective_and_grad(xm, lg, y, **prm)[0]) / (2 * h)
    assert fd == pytest.approx(gx[k, 0], rel=1e-5, abs=1e-10)
