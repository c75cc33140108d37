"""Transport maps on the device (wot.tmap).

* ``TransportMapModel`` (wot/tmap/transport_map_model.py): push_forward /
  pull_back / fates / trajectories / transition_table / census over maps kept
  in HBM, FP64 products on the matrix cores (``wotk_tmap_propagate``).
* ``Population``, ``unique_timepoint``, ``find_path`` (wot/population.py,
  wot/tmap/util.py:52-72, wot/tmap/chaining.py:43-110).
* Long-range couplings (glue_transport_maps / chain_transport_maps,
  wot/tmap/util.py:74-94 and wot/tmap/chaining.py:4-40): maps multiplied on
  the device in FP64 like the reference's np.dot (rocBLAS DGEMM), or in FP32 on
  the matrix cores (``wotk_gemm_f32``, ``precision='float32'``; ~1e-7
  relative); as in the reference there is no normalisation.
"""
import ctypes

import numpy as np

from .. import _native, _anndata
from ..ot import device as _dev
from .population import Population, find_path, unique_timepoint  # noqa: F401
from .propagate import PULL_BACK, PUSH_FORWARD, DeviceMap, normalize_rows, propagate  # noqa: F401
from .model import TransportMapModel  # noqa: F401


def gemm_f32(A, B, out=None):
    """A (M x K) @ B (K x N) in FP32 on the device (torch tensors or arrays)."""
    torch = _dev.torch_mod()
    A = (A if isinstance(A, torch.Tensor) else torch.from_numpy(np.asarray(A))).to('cuda', torch.float32)
    B = (B if isinstance(B, torch.Tensor) else torch.from_numpy(np.asarray(B))).to('cuda', torch.float32)
    A = A.contiguous()
    B = B.contiguous()
    if A.ndim != 2 or B.ndim != 2 or A.shape[1] != B.shape[0]:
        raise ValueError('shape mismatch %s @ %s' % (tuple(A.shape), tuple(B.shape)))
    M, K = A.shape
    N = B.shape[1]
    C = out if out is not None else torch.empty((M, N), dtype=torch.float32, device=A.device)
    p = _dev._ptr
    _native.check(_native.lib().wotk_gemm_f32(p(A), A.stride(0), p(B), B.stride(0), p(C), C.stride(0), M, N, K,
                                              ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return C


def chain_product(maps):
    """T_0 T_1 ... T_n (FP32, device); returns a torch tensor."""
    out = None
    for m in maps:
        out = m if out is None else gemm_f32(out, m)
    torch = _dev.torch_mod()
    return out if isinstance(out, torch.Tensor) else torch.from_numpy(np.asarray(out, np.float32)).cuda()


def glue_transport_maps(tmap_0, tmap_1, precision='float64'):
    """Transport map t0 -> t2 from t0 -> t1 and t1 -> t2 (util.py:74-94): rows
    of tmap_1 re-ordered to tmap_0's columns, then multiplied (no column
    normalisation, as in the reference).  ``precision``: 'float64' (the
    reference's arithmetic, rocBLAS DGEMM on the device) or 'float32'
    (``wotk_gemm_f32`` on the matrix cores)."""
    if precision not in ('float64', 'float32'):
        raise ValueError("precision must be 'float64' or 'float32'")
    idx = tmap_1.obs.index.get_indexer_for(tmap_0.var.index)
    x1 = np.asarray(tmap_1.X)[idx, :]
    if precision == 'float32':
        res = gemm_f32(np.asarray(tmap_0.X, np.float32), x1.astype(np.float32)).cpu().numpy().astype(np.float64)
    else:
        torch = _dev.torch_mod()
        a = torch.from_numpy(np.ascontiguousarray(tmap_0.X, dtype=np.float64)).cuda()
        b = torch.from_numpy(np.ascontiguousarray(x1, dtype=np.float64)).cuda()
        res = (a @ b).cpu().numpy()
    return _anndata.AnnData(res, tmap_0.obs.copy(), tmap_1.var.copy())


def chain_transport_maps(tmap_model, pairs_list, precision='float64'):
    """Chain the maps of consecutive pairs (chaining.py:4-40); ``tmap_model``
    needs ``get_coupling(t0, t1)`` returning an AnnData-like map;
    ``precision`` as for glue_transport_maps."""
    pairs_list = list(pairs_list)
    for i in range(len(pairs_list) - 1):
        if pairs_list[i][1] != pairs_list[i + 1][0]:
            raise ValueError("Transport maps cannot be glued together : invalid path")
    for a, b in pairs_list:
        if a >= b:
            raise ValueError("({}, {}) is not a valid transport map : it goes backwards in time".format(a, b))
    tmap_0 = tmap_model.get_coupling(*pairs_list.pop(0))
    while len(pairs_list) > 0:
        tmap_1 = tmap_model.get_coupling(*pairs_list.pop(0))
        tmap_0 = glue_transport_maps(tmap_0, tmap_1, precision=precision)
    return tmap_0
