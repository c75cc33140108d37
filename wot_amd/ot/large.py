"""Large single pairs without a stored cost: ``transport_stablev2_coords``
runs the fixed_iters solve (wot/ot/optimal_transport.py:167-236) on
C = sqeuclid(x, y) / median (wot/ot/ot_model.py:242-253) regenerating C from
the coordinates in every sweep (libwotk ``wotk_transport_stablev2_coords``:
split-BF16 MFMA cross term with ``terms`` = 3 or 6 products, FP64 potentials,
exact FP64 median).  Memory is O(I + J): a 60,000 x 60,000 pair needs no cost
matrix.  The result holds the potentials; plan entries are
exp((f_i + g_j - C_ij)/eps) / J.
"""
import ctypes

from .. import _native
from . import device as _dev


class CoordsResult:
    def __init__(self, f, g, median, stats, eps_final, row_sums=None, x=None, y=None, divisor=None):
        self.f, self.g, self.median, self.stats, self.epsilon = f, g, median, stats, eps_final
        self.row_sums = row_sums  # row sums of the returned plan
        self.x, self.y = x, y     # the coordinates the plan is defined on
        # the returned plan is R / J, or R itself after a max_iter early return
        self.divisor = float(g.shape[0]) if divisor is None else float(divisor)

    @property
    def shape(self):
        return (self.f.shape[0], self.g.shape[0])

    def dense(self):
        """The whole plan R / J (FP64, device): I x J x 8 bytes."""
        torch = _dev.torch_mod()
        return self.plan_rows(self.x, self.y, torch.arange(self.f.shape[0], device=self.f.device))

    def plan_rows(self, x, y, rows):
        """Rows of the returned plan (R / divisor, FP64, exact FP64 cost) for the given source indices."""
        torch = _dev.torch_mod()
        X = _dev.as_device_f64(torch, x, self.f.device)[rows]
        Y = _dev.as_device_f64(torch, y, self.f.device)
        D = _dev.sqeuclid_f64(X, Y)
        C = D / self.median
        return torch.exp((self.f[rows][:, None] + self.g[None, :] - C) / self.epsilon) / self.divisor


def transport_stablev2_coords(x, y, G, lambda1=1.0, lambda2=50.0, epsilon=0.05, scaling_iter=3000, tau=1e4,
                              epsilon0=1.0, extra_iter=1000, inner_iter_max=50, terms=3, want_row_sums=False,
                              device='cuda', **ignored):
    torch = _dev.torch_mod()
    X = _dev.as_device_f64(torch, x, device)
    Y = _dev.as_device_f64(torch, y, device)
    Gd = _dev.as_device_f64(torch, G, device).view(-1)
    n_i, d = X.shape
    n_j = Y.shape[0]
    if Y.shape[1] != d or Gd.numel() != n_i:
        raise ValueError('shape mismatch')
    L = _native.lib()
    ws = _dev.workspace(torch, L.wotk_fused_workspace_size(n_i, n_j, d), device)
    f = torch.empty(n_i, dtype=torch.float64, device=device)
    g = torch.empty(n_j, dtype=torch.float64, device=device)
    med = torch.empty(1, dtype=torch.float64, device=device)
    rs = torch.empty(n_i, dtype=torch.float64, device=device) if want_row_sums else None
    st = _native.Stats()
    p = _dev._ptr
    _native.check(L.wotk_transport_stablev2_coords(p(X), p(Y), n_i, n_j, d, p(Gd), float(lambda1), float(lambda2),
                                                   float(epsilon), int(scaling_iter), float(tau), float(epsilon0),
                                                   int(extra_iter), int(inner_iter_max), int(terms), p(f), p(g), p(rs),
                                                   p(med),
                                                   p(ws), ws.numel(), ctypes.byref(st),
                                                   ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    stats = st.as_dict()
    return CoordsResult(f, g, med, stats, stats['epsilon_final'], rs, X, Y)


def optimal_transport_duality_gap_coords(x, y, G, lambda1=1.0, lambda2=50.0, epsilon=0.05, batch_size=5,
                                         tolerance=1e-8, tau=1e4, epsilon0=1.0, max_iter=1e7, terms=6,
                                         want_row_sums=False, device='cuda', **ignored):
    """optimal_transport_duality_gap (wot/ot/optimal_transport.py:67-164) on
    C = sqeuclid(x, y) / median without storing C (libwotk
    ``wotk_transport_duality_gap_coords``).  Returns a CoordsResult; its stats
    carry the iteration and stabilisation counts and the final gap.  A NaN gap
    raises RuntimeError as the reference does; at ``max_iter`` the reference's
    warning is logged and ``row_sums`` are those of the un-normalised plan."""
    import logging
    import numpy as np
    torch = _dev.torch_mod()
    X = _dev.as_device_f64(torch, x, device)
    Y = _dev.as_device_f64(torch, y, device)
    Gd = _dev.as_device_f64(torch, G, device).view(-1)
    n_i, d = X.shape
    n_j = Y.shape[0]
    if Y.shape[1] != d or Gd.numel() != n_i:
        raise ValueError('shape mismatch')
    L = _native.lib()
    ws = _dev.workspace(torch, L.wotk_duality_gap_coords_workspace_size(n_i, n_j, d), device)
    f = torch.empty(n_i, dtype=torch.float64, device=device)
    g = torch.empty(n_j, dtype=torch.float64, device=device)
    med = torch.empty(1, dtype=torch.float64, device=device)
    rs = torch.empty(n_i, dtype=torch.float64, device=device) if want_row_sums else None
    st = _native.Stats()
    p = _dev._ptr
    mi = int(np.ceil(max_iter)) if np.isfinite(max_iter) else np.iinfo(np.int64).max
    _native.check(L.wotk_transport_duality_gap_coords(p(X), p(Y), n_i, n_j, d, p(Gd), float(lambda1), float(lambda2),
                                                      float(epsilon), int(batch_size), float(tolerance), float(tau),
                                                      float(epsilon0), mi, int(terms), p(f), p(g), p(rs), p(med),
                                                      p(ws), ws.numel(), ctypes.byref(st),
                                                      ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    stats = st.as_dict()
    if stats.get('max_iter_reached'):
        logging.getLogger('wot').warning("Reached max_iter with duality gap still above threshold. Returning")
    return CoordsResult(f, g, med, stats, stats['epsilon_final'], rs, X, Y,
                        divisor=1.0 if stats.get('max_iter_reached') else None)


def transport_map_coords(x, y, G, growth_iters=1, **params):
    """The growth-learning loop (optimal_transport.py:10-33) around the
    never-materialised solver: iteration k > 0 solves with the previous
    plan's row sums.  Returns (last CoordsResult, learned growth list)."""
    torch = _dev.torch_mod()
    learned = []
    res = None
    G_k = G
    for it in range(growth_iters):
        if it > 0:
            G_k = res.row_sums
        learned.append(G_k)
        res = transport_stablev2_coords(x, y, G_k, want_row_sums=True, **params)
    del torch
    return res, learned
