"""CPU oracle for OT validation (wot/ot/util.py:109-237 and
wot/ot/optimal_transport_validation.py:15-196).

TEST INFRASTRUCTURE ONLY (never imported by the product).

* ``interpolate_with_ot``, ``interpolate_randomly``,
  ``interpolate_randomly_with_growth``: numpy restatements drawing from
  numpy's global generator as the reference does (np.random.choice)
* ``emd2_lp``: the exact transportation LP solved by scipy's HiGHS (the
  library the tests pin the network simplex against; POT is not installed)
* ``emd2_assignment``: equal sizes, uniform weights: the optimal plan is a
  permutation (Birkhoff), solved by scipy's linear_sum_assignment
* ``project_middle``: the middle-day PCA projection of :79-84
"""
import numpy as np


def interpolate_with_ot(p0, p1, tmap, interp_frac, size):
    p0 = np.asarray(p0, dtype=np.float64)
    p1 = np.asarray(p1, dtype=np.float64)
    tmap = np.asarray(tmap, dtype=np.float64)
    I, J = len(p0), len(p1)
    p = tmap / np.power(tmap.sum(axis=0), 1. - interp_frac)
    p = p.flatten(order='C')
    p = p / p.sum()
    choices = np.random.choice(I * J, p=p, size=size)
    return np.asarray([p0[i // J] * (1 - interp_frac) + p1[i % J] * interp_frac for i in choices])


def interpolate_randomly(p0, p1, t, size):
    p0 = np.asarray(p0, dtype=np.float64)
    p1 = np.asarray(p1, dtype=np.float64)
    I, J = len(p0), len(p1)
    choices = np.random.choice(I * J, size=size)
    return np.asarray([p0[i // J] * (1 - t) + p1[i % J] * t for i in choices])


def interpolate_randomly_with_growth(p0, p1, t, size, g):
    p0 = np.asarray(p0, dtype=np.float64)
    p1 = np.asarray(p1, dtype=np.float64)
    q = np.ones(p1.shape[0]) * np.average(g)
    p = np.outer(g, q).flatten(order='C')
    p = p / p.sum()
    I, J = len(p0), len(p1)
    choices = np.random.choice(I * J, p=p, size=size)
    return np.asarray([p0[i // J] * (1 - t) + p1[i % J] * t for i in choices])


def sqeuclid(a, b):
    from scipy.spatial.distance import cdist
    return cdist(np.asarray(a, np.float64), np.asarray(b, np.float64), 'sqeuclidean')


def emd2_lp(a, b, M):
    from scipy.optimize import linprog
    from scipy.sparse import lil_matrix
    n, m = M.shape
    A = lil_matrix((n + m, n * m))
    for i in range(n):
        A[i, i * m:(i + 1) * m] = 1.0
    for j in range(m):
        A[n + j, j::m] = 1.0
    res = linprog(M.ravel(), A_eq=A.tocsr(), b_eq=np.concatenate([a, b]), bounds=(0, None), method='highs')
    return res.fun


def emd2_assignment(M):
    from scipy.optimize import linear_sum_assignment
    r, c = linear_sum_assignment(M)
    return M[r, c].mean()


def project_middle(x0, x1, x05, components, singular_values, mean_shift):
    U = np.vstack([x0, x1]).T.dot(components.T).dot(np.diag(1 / singular_values))
    y = x05 - mean_shift
    return np.diag(1 / singular_values).dot(U.T.dot(y.T)).T
