"""Exact earth mover's distance (wot/ot/util.py:202-237 earth_mover_distance,
which calls POT's emd2 with numItermax=1e7): the squared-Euclidean cost
between the two clouds is built on the device (libwotk ``wotk_sqeuclid_f64``),
the transportation problem is solved exactly by libwotk's network simplex
(``wotk_emd``, host: exact pivoting is sequential)."""
import ctypes
import logging

import numpy as np
import scipy.sparse

from .. import _native
from . import device as _dev

logger = logging.getLogger('wot')


def emd2(a, b, M, numItermax=100000, return_plan=False):
    """POT's ``ot.emd2(a, b, M, numItermax)``: the optimal transport cost
    sum F_ij M_ij (host arrays); with ``return_plan`` also the plan."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())
    b = np.ascontiguousarray(np.asarray(b, dtype=np.float64).ravel())
    M = np.ascontiguousarray(np.asarray(M, dtype=np.float64))
    if M.shape != (a.size, b.size):
        raise ValueError('cost matrix %s does not match weights (%d, %d)' % (M.shape, a.size, b.size))
    cost = ctypes.c_double()
    it = ctypes.c_int64()
    status = ctypes.c_int32()
    plan = np.empty_like(M) if return_plan else None
    fp = plan.ctypes.data_as(ctypes.c_void_p) if plan is not None else None
    _native.check(_native.lib().wotk_emd(a.ctypes.data_as(ctypes.c_void_p), a.size, b.ctypes.data_as(ctypes.c_void_p),
                                         b.size, M.ctypes.data_as(ctypes.c_void_p), M.shape[1], int(numItermax),
                                         ctypes.byref(cost), fp, ctypes.byref(it), ctypes.byref(status)))
    if status.value == 1:
        logger.warning('numItermax reached before optimality. Try to increase numItermax.')
    return (cost.value, plan) if return_plan else cost.value


def _dense(x):
    return x.toarray() if scipy.sparse.issparse(x) else np.asarray(x)


def emd_inputs(cloud1, cloud2, eigenvals=None, weights1=None, weights2=None):
    """(p, q, M) of earth_mover_distance: the clouds' weights (uniform unless
    given) and their squared-Euclidean cost, built on the device, as host
    arrays (util.py:202-237)."""
    c1, c2 = _dense(cloud1), _dense(cloud2)
    if eigenvals is not None:
        c1 = c1.dot(eigenvals)
        c2 = c2.dot(eigenvals)
    p = np.ones(len(c1)) / len(c1) if weights1 is None else np.asarray(weights1, np.float64) / np.sum(weights1)
    q = np.ones(len(c2)) / len(c2) if weights2 is None else np.asarray(weights2, np.float64) / np.sum(weights2)
    M = _dev.sqeuclid_f64(c1, c2).cpu().numpy()
    return p, q, M


def earth_mover_distance(cloud1, cloud2, eigenvals=None, weights1=None, weights2=None):
    """sqrt(EMD) between two point clouds under the squared-Euclidean cost
    (util.py:202-237); ``eigenvals`` scales both clouds first."""
    p, q, M = emd_inputs(cloud1, cloud2, eigenvals, weights1, weights2)
    return np.sqrt(emd2(p, q, M, numItermax=int(1e7)))
