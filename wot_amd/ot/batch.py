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
        self.eps_final = np.asarray(eps_final, dtype=np.float64)  # [pair][combination]

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
                                              _dev._ptr(self.g[p][e]), float(self.eps_final[p, e]), float(n_j),
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
                                              _dev._ptr(self.g[p][e]), float(self.eps_final[p, e]), float(n_j),
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
                      cost_mode='f64', pairs=None, group_size=None, share_exp=True, share_eps=True,
                      device='cuda'):
    """Solve every consecutive day pair (or the given index pairs) for every
    (epsilon, lambda1) of the grid epsilons x lambda1s.  ``coords[t]`` are the
    PCA coordinates of day t, ``growth[t]`` its cells' growth (G = growth, dt=1).
    ``group_size`` splits the grid into batched runs of that many combinations
    (each cost entry is then read once per group instead of once overall).
    ``share_exp`` lets combinations of equal epsilon share one exponential per
    cost entry (libwotk ``WOTK_BATCH_SHARE_EXP``); ``share_eps`` also lets
    three eps groups share the last group's exponent: powers of its
    exponential (exponent split exactly in FP32) where k = log2(e)/eps stands
    in an integer ratio (the converged schedule), the real multiple of its
    FP64 exponent elsewhere (``WOTK_BATCH_SHARE_EPS``)."""
    tc = prepare_time_course(coords, growth, pairs=pairs, cost_mode=cost_mode, device=device)
    return solve_prepared(tc, epsilons, lambda1s, lambda2=lambda2, scaling_iter=scaling_iter, tau=tau,
                          epsilon0=epsilon0, extra_iter=extra_iter, inner_iter_max=inner_iter_max,
                          graph_iters=graph_iters, group_size=group_size, share_exp=share_exp, share_eps=share_eps)


class TimeCourseRun:
    """One batched run advanced step by step (libwotk ``wotk_batch_begin`` /
    ``wotk_batch_advance`` / ``wotk_batch_potentials`` / ``wotk_batch_end``).

    ``epsilon``, ``lambda1``, ``lambda2``: arrays [pair][combination] (per-pair
    parameters, wot/ot/initializer.py:135-183).  ``G``: a device float64
    tensor, one growth vector per pair at ``bp.g_offset`` (``growth_per_combo``
    False) or one per (pair, combination) ([combination][I_p] from
    ``bp.g_offset``).  ``advance(n)`` runs the next n iterations of the
    schedule; ``result()`` returns the potentials as a :class:`BatchResult`;
    ``modes()`` which iterations took powers of one exponential (``share_eps``)."""

    def __init__(self, tc, epsilon, lambda1, lambda2, scaling_iter=3000, epsilon0=1.0, extra_iter=1000,
                 inner_iter_max=50, graph_iters=50, share_exp=True, G=None, growth_per_combo=False,
                 combos=None, bps=None, share_eps=True):
        torch = _dev.torch_mod()
        from .. import tune
        tune.apply_persisted()
        self.tc = tc
        self.device = tc.buf.device
        eps = np.ascontiguousarray(epsilon, dtype=np.float64)
        l1 = np.ascontiguousarray(lambda1, dtype=np.float64)
        l2 = np.ascontiguousarray(lambda2, dtype=np.float64)
        n_pairs = tc.n_pairs
        if eps.ndim != 2 or eps.shape[0] != n_pairs or l1.shape != eps.shape or l2.shape != eps.shape:
            raise ValueError('epsilon, lambda1, lambda2 must be [pair][combination] arrays of the same shape')
        self.E = eps.shape[1]
        self.n_iter = int(scaling_iter) + int(extra_iter)
        self.combos = combos if combos is not None else [tuple(r) for r in np.stack([eps[0], l1[0]], 1)]
        L = _native.lib()
        self._bps = bps if bps is not None else tc.bps
        arr = (_native.BatchPair * len(self._bps))(*self._bps)
        self._arr = arr
        self.ws = _dev.workspace(torch, L.wotk_batch_workspace_size(arr, n_pairs, self.E, self.n_iter), self.device)
        self.G = tc.G if G is None else G
        h = ctypes.c_void_p()
        P = ctypes.POINTER(ctypes.c_double)
        flags = (_native.BATCH_SHARE_EXP if share_exp else 0) | (_native.BATCH_SHARE_EPS if share_exp and share_eps else 0)
        _native.check(L.wotk_batch_begin(_dev._ptr(tc.buf), arr, n_pairs, _dev._ptr(self.G), int(bool(growth_per_combo)),
                                         self.E, eps.ctypes.data_as(P), l1.ctypes.data_as(P), l2.ctypes.data_as(P),
                                         int(scaling_iter), float(epsilon0), int(extra_iter), int(inner_iter_max),
                                         int(graph_iters), flags,
                                         _dev._ptr(self.ws), self.ws.numel(), ctypes.byref(h), _dev._stream(torch)))
        self._h = h
        self.remaining = self.n_iter

    def advance(self, n_iterations):
        """Queue the next ``n_iterations`` (at most what is left); returns the iterations left."""
        if self._h is None:
            raise RuntimeError('batch run already closed')
        torch = _dev.torch_mod()
        left = ctypes.c_int64()
        _native.check(_native.lib().wotk_batch_advance(self._h, int(n_iterations), ctypes.byref(left),
                                                       _dev._stream(torch)))
        self.remaining = int(left.value)
        return self.remaining

    def result(self):
        """Potentials (and stats) at the current iteration as a BatchResult."""
        torch = _dev.torch_mod()
        tc, E = self.tc, self.E
        shapes = tc.shapes
        f_out = torch.empty(sum(E * s[0] for s in shapes), dtype=torch.float64, device=self.device)
        g_out = torch.empty(sum(E * s[1] for s in shapes), dtype=torch.float64, device=self.device)
        epsf = (ctypes.c_double * (tc.n_pairs * E))()
        st = _native.Stats()
        _native.check(_native.lib().wotk_batch_potentials(self._h, _dev._ptr(f_out), _dev._ptr(g_out), epsf,
                                                          ctypes.byref(st), _dev._stream(torch)))
        fs, gs = [], []
        a = b = 0
        for (n_i, n_j) in shapes:
            fs.append(f_out[a:a + E * n_i].view(E, n_i))
            gs.append(g_out[b:b + E * n_j].view(E, n_j))
            a += E * n_i
            b += E * n_j
        d = st.as_dict()
        d['exact_rows'] = d['stabilisations']  # rows recomputed with the online maximum
        d['shared_group'] = d['gap_checks']     # combinations per shared exponential (0: none)
        eps_pe = np.array(epsf[:], dtype=np.float64).reshape(tc.n_pairs, E)
        return BatchResult(tc.buf, self._bps, self.combos, fs, gs, d, shapes, eps_pe)

    def modes(self):
        """Per iteration: the powers eps groups 1 (bits 0-7) and 0 (bits 8-15)
        take of the last group's exponential (``share_eps``; 255: the real
        ratio of their exponents), 0 where every group uses its own exponent."""
        if self._h is None:
            raise RuntimeError('batch run already closed')
        out = (ctypes.c_int32 * max(1, self.n_iter))()
        _native.check(_native.lib().wotk_batch_modes(self._h, out, self.n_iter))
        return np.array(out[:self.n_iter], dtype=np.int32)

    def close(self):
        if self._h is not None:
            _native.lib().wotk_batch_end(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def grid_arrays(n_pairs, epsilons, lambda1s, lambda2):
    """[pair][combination] parameter arrays of the Cartesian grid epsilons x
    lambda1s (epsilon-major), the same for every pair."""
    combos = list(itertools.product(epsilons, lambda1s))
    eps = np.tile(np.array([c[0] for c in combos], dtype=np.float64), (n_pairs, 1))
    l1 = np.tile(np.array([c[1] for c in combos], dtype=np.float64), (n_pairs, 1))
    l2 = np.full_like(eps, float(lambda2))
    return combos, eps, l1, l2


def solve_prepared(tc, epsilons=(0.05,), lambda1s=(1.0,), lambda2=50.0, scaling_iter=3000, tau=1e4, epsilon0=1.0,
                   extra_iter=1000, inner_iter_max=50, graph_iters=50, group_size=None, share_exp=True,
                   share_eps=True):
    """``solve_time_course`` on costs already built by ``prepare_time_course``."""
    torch = _dev.torch_mod()
    combos, eps, l1, l2 = grid_arrays(tc.n_pairs, epsilons, lambda1s, lambda2)
    E = len(combos)
    gsz = E if group_size is None else int(group_size)
    groups = [list(range(k, min(E, k + gsz))) for k in range(0, E, gsz)]
    device = tc.buf.device
    fs = [torch.empty((E, s[0]), dtype=torch.float64, device=device) for s in tc.shapes]
    gs = [torch.empty((E, s[1]), dtype=torch.float64, device=device) for s in tc.shapes]
    eps_final = np.zeros((tc.n_pairs, E))
    st_total = None
    for grp in groups:
        with TimeCourseRun(tc, eps[:, grp], l1[:, grp], l2[:, grp], scaling_iter=scaling_iter, epsilon0=epsilon0,
                           extra_iter=extra_iter, inner_iter_max=inner_iter_max, graph_iters=graph_iters,
                           share_exp=share_exp, share_eps=share_eps, combos=[combos[e] for e in grp]) as run:
            run.advance(run.n_iter)
            r = run.result()
        for k in range(tc.n_pairs):
            fs[k][grp] = r.f[k]
            gs[k][grp] = r.g[k]
        eps_final[:, grp] = r.eps_final
        d = r.stats
        if st_total is None:
            st_total = d
        else:
            st_total['sweep_ms'] += d['sweep_ms']
            st_total['sweeps'] += d['sweeps']
            st_total['stabilisations'] += d['stabilisations']
            st_total['exact_rows'] += d['exact_rows']
    del tau  # absorptions are exact in the log domain
    return BatchResult(tc.buf, tc.bps, combos, fs, gs, st_total, tc.shapes, eps_final)
