"""Device-resident cost matrices and thin wrappers over the libwotk C ABI.

Torch-ROCm supplies device memory and the stream; every numeric step runs in
libwotk's HIP kernels.  A cost lives on the device as the FP32 pair (C, C^T)
with leading dimensions padded to a multiple of 64 floats (256 B rows): the
row sweeps stream C for the source potential and C^T for the target potential.
"""
import ctypes

import numpy as np

from .. import _native

_LD_ALIGN = 64


def torch_mod():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError('wot_amd needs a ROCm GPU (torch.cuda.is_available() is False)')
    return torch


def _stream(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _ld(n):
    return int((n + _LD_ALIGN - 1) // _LD_ALIGN * _LD_ALIGN)


def workspace(torch, nbytes, device):
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


def as_device_f64(torch, x, device):
    """numpy / list / torch -> contiguous float64 device tensor."""
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=torch.float64).contiguous()
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.float64))).to(device)


class DeviceCost:
    """FP32 cost pair (C, C^T) on the device, plus its FP64 median when it was
    computed from coordinates (OTModel.compute_default_cost_matrix semantics)."""

    def __init__(self, C, CT, n_i, n_j, median=None, median_index=None):
        self.C, self.CT = C, CT
        self.n_i, self.n_j = int(n_i), int(n_j)
        self.ldc, self.ldct = C.shape[1], CT.shape[1]
        self.median = median              # device scalar tensor or None
        self.median_index = median_index  # device int64[2]: flat indices of the middle element(s)

    @property
    def shape(self):
        return (self.n_i, self.n_j)

    @property
    def device(self):
        return self.C.device

    @staticmethod
    def _alloc(torch, n_i, n_j, device):
        C = torch.empty((n_i, _ld(n_j)), dtype=torch.float32, device=device)
        CT = torch.empty((n_j, _ld(n_i)), dtype=torch.float32, device=device)
        return C, CT

    @classmethod
    def from_coordinates(cls, x, y, mode='f64', device='cuda'):
        """C = sqeuclid(x, y) / median  (wot/ot/ot_model.py:242-253)."""
        torch = torch_mod()
        X = as_device_f64(torch, x, device)
        Y = as_device_f64(torch, y, device)
        if X.ndim != 2 or Y.ndim != 2 or X.shape[1] != Y.shape[1]:
            raise ValueError('coordinates must be 2-D with the same number of columns')
        n_i, d = X.shape
        n_j = Y.shape[0]
        m = _native.COST_MODES[mode]
        L = _native.lib()
        ws = workspace(torch, L.wotk_cost_workspace_size(n_i, n_j, d, m), device)
        C, CT = cls._alloc(torch, n_i, n_j, device)
        med = torch.empty(1, dtype=torch.float64, device=device)
        midx = torch.empty(2, dtype=torch.int64, device=device)
        _native.check(L.wotk_cost_sqeuclid(_ptr(X), _ptr(Y), n_i, n_j, d, m, _ptr(C), C.shape[1], _ptr(CT),
                                           CT.shape[1], _ptr(med), _ptr(midx), _ptr(ws), ws.numel(),
                                           _stream(torch)))
        return cls(C, CT, n_i, n_j, median=med, median_index=midx)

    @classmethod
    def from_matrix(cls, cost, scale=1.0, device='cuda'):
        """A given cost matrix, used as-is (ot_model.py:307-311)."""
        torch = torch_mod()
        C64 = as_device_f64(torch, cost, device)
        if C64.ndim != 2:
            raise ValueError('cost matrix must be 2-D')
        n_i, n_j = C64.shape
        C, CT = cls._alloc(torch, n_i, n_j, device)
        L = _native.lib()
        _native.check(L.wotk_cost_import(_ptr(C64), n_i, n_j, C64.stride(0), float(scale), _ptr(C), C.shape[1],
                                         _ptr(CT), CT.shape[1], _stream(torch)))
        return cls(C, CT, n_i, n_j)

    def dense(self):
        """FP32 cost as a (I, J) torch view (no copy)."""
        return self.C[:, :self.n_j]


def ensure_cost(C, device='cuda'):
    if isinstance(C, DeviceCost):
        return C
    return DeviceCost.from_matrix(C, device=device)


def sqeuclid_f64(x, y, device='cuda'):
    """Raw FP64 squared distances on the device (I x J torch tensor)."""
    torch = torch_mod()
    X = as_device_f64(torch, x, device)
    Y = as_device_f64(torch, y, device)
    out = torch.empty((X.shape[0], Y.shape[0]), dtype=torch.float64, device=device)
    _native.check(_native.lib().wotk_sqeuclid_f64(_ptr(X), _ptr(Y), X.shape[0], Y.shape[0], X.shape[1],
                                                  _ptr(out), out.stride(0), _stream(torch)))
    return out


def median_f64(values, return_index=False):
    """Exact numpy-style median of a float64 device tensor (all values >= 0);
    with ``return_index`` also the flat indices of the two middle elements."""
    torch = torch_mod()
    v = values.contiguous().view(-1)
    L = _native.lib()
    ws = workspace(torch, L.wotk_median_workspace_size(v.numel()), v.device)
    out = torch.empty(1, dtype=torch.float64, device=v.device)
    idx = torch.empty(2, dtype=torch.int64, device=v.device)
    _native.check(L.wotk_median_f64(_ptr(v), v.numel(), _ptr(out), _ptr(idx), _ptr(ws), ws.numel(),
                                    _stream(torch)))
    return (out, idx) if return_index else out


def solve(cost, G, solver, params, want_plan=True, want_row_sums=False, want_potentials=False):
    """Run one solve on the device.  Returns a dict with torch tensors:
    plan (I x J float64, or None), row_sums, f, g and the stats dict."""
    torch = torch_mod()
    dev = cost.device
    n_i, n_j = cost.shape
    Gd = as_device_f64(torch, G, dev).view(-1)
    if Gd.numel() != n_i:
        raise ValueError('G has %d entries for %d source cells' % (Gd.numel(), n_i))
    L = _native.lib()
    from .. import tune
    tune.apply_persisted()
    ws = workspace(torch, L.wotk_solver_workspace_size(n_i, n_j), dev)
    plan = torch.empty((n_i, n_j), dtype=torch.float64, device=dev) if want_plan else None
    rows = torch.empty(n_i, dtype=torch.float64, device=dev) if want_row_sums else None
    f = torch.empty(n_i, dtype=torch.float64, device=dev) if want_potentials else None
    g = torch.empty(n_j, dtype=torch.float64, device=dev) if want_potentials else None
    st = _native.Stats()
    ldp = n_j
    if solver == 'fixed_iters':
        tau = params['tau']
        if tau is None:
            # the reference compares max(...) > None and fails the same way
            raise TypeError("'>' not supported between instances of 'float' and 'NoneType'")
        rc = L.wotk_transport_stablev2(
            _ptr(cost.C), cost.ldc, _ptr(cost.CT), cost.ldct, n_i, n_j, _ptr(Gd),
            float(params['lambda1']), float(params['lambda2']), float(params['epsilon']),
            int(params['scaling_iter']), float(tau), float(params['epsilon0']), int(params['extra_iter']),
            int(params['inner_iter_max']), _ptr(plan), ldp, _ptr(rows), _ptr(f), _ptr(g), _ptr(ws), ws.numel(),
            ctypes.byref(st), _stream(torch))
    elif solver == 'duality_gap':
        max_iter = params['max_iter']
        max_iter = int(np.ceil(max_iter)) if np.isfinite(max_iter) else np.iinfo(np.int64).max
        rc = L.wotk_transport_duality_gap(
            _ptr(cost.C), cost.ldc, _ptr(cost.CT), cost.ldct, n_i, n_j, _ptr(Gd),
            float(params['lambda1']), float(params['lambda2']), float(params['epsilon']),
            int(params['batch_size']), float(params['tolerance']), float(params['tau']),
            float(params['epsilon0']), max_iter, _ptr(plan), ldp, _ptr(rows), _ptr(f), _ptr(g), _ptr(ws),
            ws.numel(), ctypes.byref(st), _stream(torch))
    else:
        raise ValueError('Unknown solver')
    _native.check(rc)
    return dict(plan=plan, row_sums=rows, f=f, g=g, stats=st.as_dict())


def primal_dual(cost, G, f, g, lambda1, lambda2, epsilon):
    """Primal and dual objectives (optimal_transport.py:45-62) at the plan
    exp((f_i + g_j - C_ij)/eps), FP64 reductions on the device."""
    torch = torch_mod()
    dev = cost.device
    n_i, n_j = cost.shape
    L = _native.lib()
    ws = workspace(torch, L.wotk_solver_workspace_size(n_i, n_j), dev)
    out = (ctypes.c_double * 2)()
    fd, gd, Gd = (as_device_f64(torch, t, dev).view(-1) for t in (f, g, G))
    _native.check(L.wotk_primal_dual(_ptr(cost.C), cost.ldc, _ptr(cost.CT), cost.ldct, n_i, n_j, _ptr(Gd),
                                     _ptr(fd), _ptr(gd), float(lambda1), float(lambda2), float(epsilon), out,
                                     _ptr(ws), ws.numel(), _stream(torch)))
    return out[0], out[1]
