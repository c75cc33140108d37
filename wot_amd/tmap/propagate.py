"""Populations through device-resident transport maps (libwotk
``wotk_tmap_propagate``): the products of TransportMapModel.push_forward /
pull_back (wot/tmap/transport_map_model.py:284-293, 350-359) on the matrix
cores, FP64 accumulation, with the maps kept in HBM.

A ``DeviceMap`` is one map (rows = source-day cells, columns = target-day
cells) in FP64 or FP32.  ``propagate`` runs a whole chain of maps in one
library call; populations are CELLS x POPULATIONS matrices on the device (the
transpose of the reference's ``np.vstack([pop.p ...])``) and stay there
between maps.
"""
import ctypes

import numpy as np

from .. import _native
from ..ot import device as _dev

PUSH_FORWARD = _native.PUSH_FORWARD
PULL_BACK = _native.PULL_BACK
_DTYPES = {'float64': _native.DTYPE_F64, 'float32': _native.DTYPE_F32}


class DeviceMap:
    """One transport map resident on the device, with the cell ids of its
    rows (source day) and columns (target day) when known."""

    def __init__(self, X, obs_names=None, var_names=None):
        torch = _dev.torch_mod()
        if not isinstance(X, torch.Tensor) or X.device.type != 'cuda':
            raise TypeError('DeviceMap needs a CUDA tensor; use DeviceMap.upload for host data')
        if X.ndim != 2 or X.dtype not in (torch.float64, torch.float32) or X.stride(1) != 1:
            raise ValueError('a device map is a row-major 2-D float64 / float32 tensor')
        self.X = X
        self.obs_names = obs_names
        self.var_names = var_names
        self.XT = None  # optional transposed copy (cols x rows): pull backs stream it coalesced

    def make_transposed(self):
        """Keep a transposed copy in HBM (doubles the map's memory): every pull
        back through this map then runs the push kernel over the copy."""
        if self.XT is None:
            torch = _dev.torch_mod()
            dt = 'float64' if self.X.dtype == torch.float64 else 'float32'
            t = DeviceMap.empty(self.X.shape[1], self.X.shape[0], dt, self.X.device).X
            code = _native.DTYPE_F64 if dt == 'float64' else _native.DTYPE_F32
            _native.check(_native.lib().wotk_tmap_transpose(_dev._ptr(self.X), code, self.X.stride(0),
                                                            self.X.shape[0], self.X.shape[1], _dev._ptr(t),
                                                            t.stride(0), _dev._stream(torch)))
            self.XT = t
        return self

    @classmethod
    def empty(cls, rows, cols, dtype='float64', device='cuda', obs_names=None, var_names=None):
        """Uninitialised map whose rows start on 256-byte boundaries (the
        product kernels then read every row with 16-byte vector loads)."""
        torch = _dev.torch_mod()
        if dtype not in _DTYPES:
            raise ValueError('dtype must be float64 or float32, not %r' % (dtype,))
        tdt = torch.float64 if dtype == 'float64' else torch.float32
        align = 32 if dtype == 'float64' else 64
        ld = (int(cols) + align - 1) // align * align
        buf = torch.empty((int(rows), ld), dtype=tdt, device=device)
        return cls(buf[:, :int(cols)], obs_names, var_names)

    @classmethod
    def upload(cls, tmap, dtype='float64', device='cuda'):
        """From an AnnData-like map (X, obs, var), a numpy array or a tensor."""
        torch = _dev.torch_mod()
        if dtype not in _DTYPES:
            raise ValueError('dtype must be float64 or float32, not %r' % (dtype,))
        tdt = torch.float64 if dtype == 'float64' else torch.float32
        obs = var = None
        X = tmap
        if hasattr(tmap, 'X') and hasattr(tmap, 'obs'):
            X = tmap.X
            obs, var = tmap.obs.index, tmap.var.index
        if not isinstance(X, torch.Tensor):
            if hasattr(X, 'toarray'):  # scipy sparse
                X = X.toarray()
            X = torch.from_numpy(np.ascontiguousarray(np.asarray(X, dtype=np.float64)))
        if X.ndim != 2:
            raise ValueError('a transport map is 2-D, got shape %s' % (tuple(X.shape),))
        out = cls.empty(X.shape[0], X.shape[1], dtype, device, obs, var)
        out.X.copy_(X)
        return out

    @property
    def shape(self):
        return tuple(self.X.shape)

    @property
    def nbytes(self):
        return self.X.numel() * self.X.element_size()

    def struct(self):
        torch = _dev.torch_mod()
        dt = _native.DTYPE_F64 if self.X.dtype == torch.float64 else _native.DTYPE_F32
        t = self.XT
        return _native.TMap(self.X.data_ptr(), dt, 0, self.X.stride(0), self.X.shape[0], self.X.shape[1],
                            t.data_ptr() if t is not None else None, t.stride(0) if t is not None else 0)


def _cells_in(m, direction):
    return m.shape[0] if direction == PUSH_FORWARD else m.shape[1]


def _cells_out(m, direction):
    return m.shape[1] if direction == PUSH_FORWARD else m.shape[0]


def propagate(maps, pops, direction, normalize=True, outs=None, stats=None):
    """Push (``direction=PUSH_FORWARD``: p <- p T) or pull (``PULL_BACK``:
    p <- (T p^T)^T) the populations ``pops`` (device tensor, cells x k,
    float64) through ``maps`` (DeviceMaps, in the order they are applied).

    ``outs``: None (return the final populations, a new tensor), or a list
    with one entry per map: a cells x k float64 device tensor (row stride >= k,
    unit column stride) receiving the populations after that map, or None
    (not kept; the last entry must be a tensor).  Returns the final tensor.
    ``stats``: a dict that receives the library's kernel timer (maps applied,
    summed product-kernel ms; the call then synchronises).
    """
    torch = _dev.torch_mod()
    maps = list(maps)
    if not maps:
        raise ValueError('propagate needs at least one map')
    if pops.ndim != 2 or pops.dtype != torch.float64 or pops.stride(1) != 1:
        pops = pops.to(torch.float64).contiguous()
    n0, k = pops.shape
    if n0 != _cells_in(maps[0], direction):
        raise ValueError('populations have %d cells, the first map takes %d' % (n0, _cells_in(maps[0], direction)))
    if outs is None:
        outs = [None] * (len(maps) - 1) + [torch.empty((_cells_out(maps[-1], direction), k), dtype=torch.float64,
                                                        device=pops.device)]
    if len(outs) != len(maps) or outs[-1] is None:
        raise ValueError('outs needs one entry per map and a tensor for the last one')
    ld = None
    for s, o in enumerate(outs):
        if o is None:
            continue
        if o.dtype != torch.float64 or o.ndim != 2 or o.stride(1) != 1 or o.shape != (_cells_out(maps[s], direction), k):
            raise ValueError('outs[%d] must be a float64 %d x %d tensor with unit column stride'
                             % (s, _cells_out(maps[s], direction), k))
        if ld is None:
            ld = o.stride(0)
        elif o.stride(0) != ld:
            raise ValueError('every output needs the same row stride')
    arr = (_native.TMap * len(maps))(*[m.struct() for m in maps])
    L = _native.lib()
    nbytes = L.wotk_tmap_workspace_size(arr, len(maps), direction, k)
    ws = _dev.workspace(torch, nbytes, pops.device)
    ptrs = (ctypes.c_void_p * len(maps))(*[o.data_ptr() if o is not None else None for o in outs])
    st = _native.Stats() if stats is not None else None
    _native.check(L.wotk_tmap_propagate(arr, len(maps), direction, _dev._ptr(pops), pops.stride(0), k,
                                        1 if normalize else 0, ptrs, ld, _dev._ptr(ws), ws.numel(),
                                        ctypes.byref(st) if st is not None else None, _dev._stream(torch)))
    if st is not None:
        stats.update(maps=int(st.sweeps), kernel_ms=float(st.sweep_ms))
    return outs[-1]


def normalize_rows(X):
    """X[i, :] /= X[i, :].sum() in place on the device (float64, unit column stride)."""
    torch = _dev.torch_mod()
    if X.dtype != torch.float64 or X.ndim != 2 or X.stride(1) != 1:
        raise ValueError('normalize_rows needs a float64 2-D tensor with unit column stride')
    _native.check(_native.lib().wotk_rows_normalize_f64(_dev._ptr(X), X.shape[0], X.shape[1], X.stride(0),
                                                        _dev._stream(torch)))
    return X
