"""Batched transport maps for a whole time course and a parameter grid.

``solve_time_course(coords, growth, epsilons, lambda1s, ...)`` computes, for
every consecutive day pair (the pairs ``OTModel.compute_all_transport_maps``
walks, wot/ot/ot_model.py:153-157) and every (epsilon, lambda1) of the
Cartesian grid (the reference's parameter sweep runs one process per
combination), the fixed_iters solve of wot/ot/optimal_transport.py:167-236.
All costs are built on the device into one buffer and the solves advance
together: one persistent sweep launch per half-iteration covers every pair
and combination (libwotk ``wotk_batch_stablev2``), optionally replayed from a
hipGraph.

The result keeps the potentials; ``plan(pair, combo)`` materialises one map
(FP64, R / J) on demand, since the plans of a full grid would not fit in HBM.
"""
import ctypes
import itertools

import numpy as np

from .. import _native
from . import device as _dev

_ALIGN = 64  # floats: every matrix starts on a 256-byte boundary


class BatchResult:
    def __init__(self, cost_buf, pairs, combos, f, g, stats, shapes, eps_final):
        self.cost_buf = cost_buf
        self.pairs = pairs          # list of BatchPair
        self.combos = combos        # list of (epsilon, lambda1)
        self.f = f                  # list per pair: tensor (E, I_p)
        self.g = g                  # list per pair: tensor (E, J_p)
        self.stats = stats
        self.shapes = shapes
        self.eps_final = eps_final  # per combination

    def plan(self, p, e, device_out=False, out=None):
        """Transport map of pair p, combination e: exp((f+g-C)/eps) / J (FP64).
        ``out``: an (I, J) float64 device view with unit column stride to write
        into (any row stride); implies device_out."""
        torch = _dev.torch_mod()
        bp = self.pairs[p]
        n_i, n_j = self.shapes[p]
        if out is None:
            out = torch.empty((n_i, n_j), dtype=torch.float64, device=self.cost_buf.device)
        elif out.dtype != torch.float64 or tuple(out.shape) != (n_i, n_j) or out.stride(1) != 1:
            raise ValueError('out must be a float64 (%d, %d) view with unit column stride' % (n_i, n_j))
        else:
            device_out = True
        C = self.cost_buf[bp.c_offset:]
        _native.check(_native.lib().wotk_plan(_dev._ptr(C), bp.ldc, n_i, n_j, _dev._ptr(self.f[p][e]),
                                              _dev._ptr(self.g[p][e]), float(self.eps_final[e]), float(n_j),
                                              _dev._ptr(out), out.stride(0), None, _dev._stream(torch)))
        return out if device_out else out.cpu().numpy()

    def row_sums(self, p, e):
        """Row sums of map (p, e): the learned growth of the next growth iteration."""
        torch = _dev.torch_mod()
        bp = self.pairs[p]
        n_i, n_j = self.shapes[p]
        out = torch.empty(n_i, dtype=torch.float64, device=self.cost_buf.device)
        C = self.cost_buf[bp.c_offset:]
        _native.check(_native.lib().wotk_plan(_dev._ptr(C), bp.ldc, n_i, n_j, _dev._ptr(self.f[p][e]),
                                              _dev._ptr(self.g[p][e]), float(self.eps_final[e]), float(n_j),
                                              None, n_j, _dev._ptr(out), _dev._stream(torch)))
        return out


def _up(n):
    return (n + _ALIGN - 1) // _ALIGN * _ALIGN


def build_costs(coords, pairs_idx, cost_mode='f64', device='cuda'):
    """All pairs' FP32 (C, C^T) in one device buffer; returns (buffer, BatchPair list, shapes)."""
    torch = _dev.torch_mod()
    L = _native.lib()
    mode = _native.COST_MODES[cost_mode]
    layout = []
    off = 0
    for (a, b) in pairs_idx:
        n_i, n_j = coords[a].shape[0], coords[b].shape[0]
        ldc, ldct = _up(n_j), _up(n_i)
        layout.append((off, off + n_i * ldc, ldc, ldct, n_i, n_j))
        off += n_i * ldc + n_j * ldct
    buf = torch.empty(off, dtype=torch.float32, device=device)
    ws_bytes = max(L.wotk_cost_workspace_size(l[4], l[5], coords[0].shape[1], mode) for l in layout)
    ws = _dev.workspace(torch, ws_bytes, device)
    med = torch.empty(1, dtype=torch.float64, device=device)
    out = []
    for (a, b), (c_off, ct_off, ldc, ldct, n_i, n_j) in zip(pairs_idx, layout):
        X = _dev.as_device_f64(torch, coords[a], device)
        Y = _dev.as_device_f64(torch, coords[b], device)
        _native.check(L.wotk_cost_sqeuclid(_dev._ptr(X), _dev._ptr(Y), n_i, n_j, X.shape[1], mode,
                                           ctypes.c_void_p(buf.data_ptr() + 4 * c_off), ldc,
                                           ctypes.c_void_p(buf.data_ptr() + 4 * ct_off), ldct, _dev._ptr(med), None,
                                           _dev._ptr(ws), ws.numel(), _dev._stream(torch)))
        out.append(_native.BatchPair(c_off, ct_off, ldc, ldct, n_i, n_j, 0))
    return buf, out, [(l[4], l[5]) for l in layout]


class TimeCourseCosts:
    """Device-resident inputs of a batched time course: every pair's FP32
    (C, C^T) in one buffer, the pair table and the concatenated growth."""

    def __init__(self, buf, bps, shapes, G, pairs):
        self.buf = buf
        self.bps = bps
        self.shapes = shapes
        self.G = G
        self.pairs = pairs
        self.arr = (_native.BatchPair * len(bps))(*bps)

    @property
    def n_pairs(self):
        return len(self.bps)

    def mean_row_length(self):
        """sum(2 I J) / sum(I + J): the row length the tile table is keyed on."""
        e = sum(2.0 * i * j for i, j in self.shapes)
        r = sum(float(i + j) for i, j in self.shapes)
        return int(e / r)


def prepare_time_course(coords, growth, pairs=None, cost_mode='f64', device='cuda'):
    """Build the costs of every consecutive day pair (or the given index pairs)."""
    torch = _dev.torch_mod()
    if pairs is None:
        pairs = [(t, t + 1) for t in range(len(coords) - 1)]
    buf, bps, shapes = build_costs(coords, pairs, cost_mode=cost_mode, device=device)
    goffs = np.cumsum([0] + [shapes[k][0] for k in range(len(pairs))])
    G = torch.cat([_dev.as_device_f64(torch, growth[a], device).view(-1) for (a, _) in pairs])
    for k, bp in enumerate(bps):
        bp.g_offset = int(goffs[k])
    return TimeCourseCosts(buf, bps, shapes, G, pairs)


def solve_time_course(coords, growth, epsilons=(0.05,), lambda1s=(1.0,), lambda2=50.0, scaling_iter=3000,
                      tau=1e4, epsilon0=1.0, extra_iter=1000, inner_iter_max=50, graph_iters=50,
                      cost_mode='f64', pairs=None, group_size=None, share_exp=True, device='cuda'):
    """Solve every consecutive day pair (or the given index pairs) for every
    (epsilon, lambda1) of the grid epsilons x lambda1s.  ``coords[t]`` are the
    PCA coordinates of day t, ``growth[t]`` its cells' growth (G = growth, dt=1).
    ``group_size`` splits the grid into batched runs of that many combinations
    (each cost entry is then read once per group instead of once overall).
    ``share_exp`` lets combinations of equal epsilon share one exponential per
    cost entry (libwotk ``WOTK_BATCH_SHARE_EXP``)."""
    tc = prepare_time_course(coords, growth, pairs=pairs, cost_mode=cost_mode, device=device)
    return solve_prepared(tc, epsilons, lambda1s, lambda2=lambda2, scaling_iter=scaling_iter, tau=tau,
                          epsilon0=epsilon0, extra_iter=extra_iter, inner_iter_max=inner_iter_max,
                          graph_iters=graph_iters, group_size=group_size, share_exp=share_exp)


def solve_prepared(tc, epsilons=(0.05,), lambda1s=(1.0,), lambda2=50.0, scaling_iter=3000, tau=1e4, epsilon0=1.0,
                   extra_iter=1000, inner_iter_max=50, graph_iters=50, group_size=None, share_exp=True):
    """``solve_time_course`` on costs already built by ``prepare_time_course``."""
    torch = _dev.torch_mod()
    from .. import tune
    tune.apply_persisted()
    device = tc.buf.device
    buf, bps, shapes, G, arr = tc.buf, tc.bps, tc.shapes, tc.G, tc.arr
    combos = list(itertools.product(epsilons, lambda1s))
    E = len(combos)
    n_iter = scaling_iter + extra_iter
    L = _native.lib()
    gsz = E if group_size is None else int(group_size)
    groups = [list(range(k, min(E, k + gsz))) for k in range(0, E, gsz)]
    ws = _dev.workspace(torch, max(L.wotk_batch_workspace_size(arr, len(bps), len(g), n_iter) for g in groups),
                        device)
    fs = [torch.empty((E, s[0]), dtype=torch.float64, device=device) for s in shapes]
    gs = [torch.empty((E, s[1]), dtype=torch.float64, device=device) for s in shapes]
    st_total = None
    for grp in groups:
        Eg = len(grp)
        f_out = torch.empty(sum(Eg * s[0] for s in shapes), dtype=torch.float64, device=device)
        g_out = torch.empty(sum(Eg * s[1] for s in shapes), dtype=torch.float64, device=device)
        eps = (ctypes.c_double * Eg)(*[combos[e][0] for e in grp])
        lam = (ctypes.c_double * Eg)(*[combos[e][1] for e in grp])
        st = _native.Stats()
        _native.check(L.wotk_batch_stablev2(_dev._ptr(buf), arr, len(bps), _dev._ptr(G), Eg, eps, lam,
                                            float(lambda2), int(scaling_iter), float(tau), float(epsilon0),
                                            int(extra_iter), int(inner_iter_max), _dev._ptr(f_out), _dev._ptr(g_out),
                                            int(graph_iters), _native.BATCH_SHARE_EXP if share_exp else 0,
                                            _dev._ptr(ws), ws.numel(), ctypes.byref(st),
                                            _dev._stream(torch)))
        a = b = 0
        for k, (n_i, n_j) in enumerate(shapes):
            fs[k][grp] = f_out[a:a + Eg * n_i].view(Eg, n_i)
            gs[k][grp] = g_out[b:b + Eg * n_j].view(Eg, n_j)
            a += Eg * n_i
            b += Eg * n_j
        d = st.as_dict()
        d['exact_rows'] = d['stabilisations']  # rows recomputed with the online maximum
        d['shared_group'] = d['gap_checks']     # combinations per shared exponential (0: none)
        if st_total is None:
            st_total = d
        else:
            st_total['sweep_ms'] += d['sweep_ms']
            st_total['sweeps'] += d['sweeps']
            st_total['stabilisations'] += d['stabilisations']
            st_total['exact_rows'] += d['exact_rows']

    def reg(e, n):
        return (epsilon0 - e) * np.exp(-n) + e
    n_end = scaling_iter // inner_iter_max
    eps_final = [reg(c[0], n_end) for c in combos]
    return BatchResult(buf, bps, combos, fs, gs, st_total, shapes, eps_final)
