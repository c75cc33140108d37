"""OT validation (wot/ot/optimal_transport_validation.py:15-196 and the
interpolation helpers of wot/ot/util.py:109-200): for day triplets
(t0, t0.5, t1), populations interpolated between t0 and t1 are compared
with the real cells of t0.5 by the exact earth mover's distance.

Interpolation runs on the device (libwotk ``wotk_interp_*``: sampling
weights, inverse-CDF sampling, gather + lerp).  The uniforms are drawn from
numpy's global generator exactly as the reference's ``np.random.choice``
draws them: ``random_sample(size)`` with weights, ``randint(0, I*J, size)``
without.  A seeded run therefore picks the same cell pairs.  EMDs use
``wot_amd.ot.emd`` (device cost, exact network simplex); the summary's EMDs
are independent, so their network simplex runs on a pool of host threads
(the library call releases the GIL) while the main thread goes on with the
device work and the random draws in the reference's order.
"""
import collections
import itertools
import os
from concurrent.futures import Future, ThreadPoolExecutor

import numpy as np
import pandas as pd
import scipy.sparse

from .. import _anndata, _native
from . import device as _dev
from . import emd as _emd
from .pca import compute_pca


def _dense64(x):
    x = x.toarray() if scipy.sparse.issparse(x) else x
    return np.asarray(x, dtype=np.float64)


def _sample_rows(p0, p1, choices, frac):
    """rows p0[i // J] (1 - frac) + p1[i % J] frac for device indices ``choices``."""
    torch = _dev.torch_mod()
    P0 = _dev.as_device_f64(torch, p0, 'cuda')
    P1 = _dev.as_device_f64(torch, p1, 'cuda')
    size = int(choices.numel())
    out = torch.empty((size, P0.shape[1]), dtype=torch.float64, device='cuda')
    if size:
        _native.check(_native.lib().wotk_interp_gather(_dev._ptr(choices), size, _dev._ptr(P0), P0.stride(0),
                                                       P0.shape[0], _dev._ptr(P1), P1.stride(0), P1.shape[0],
                                                       P0.shape[1], float(frac), _dev._ptr(out), out.stride(0),
                                                       _dev._stream(torch)))
    return out.cpu().numpy()


def _choice_weighted(W):
    """np.random.choice(W.numel(), p=W / W.sum(), size) on the device: numpy's
    inverse CDF (cumsum, divide by the last entry, searchsorted right) with
    numpy's uniforms."""
    torch = _dev.torch_mod()
    flat = W.reshape(-1)
    if not bool(torch.isfinite(flat).all()):
        raise ValueError('probabilities contain NaN')
    if bool((flat < 0).any()):
        raise ValueError('probabilities are not non-negative')
    cdf = torch.cumsum(flat / flat.sum(), 0)
    cdf = cdf / cdf[-1]

    def draw(size):
        u = torch.from_numpy(np.random.random_sample(size)).cuda()
        idx = torch.empty(size, dtype=torch.int64, device='cuda')
        if size:
            _native.check(_native.lib().wotk_interp_sample(_dev._ptr(cdf), cdf.numel(), _dev._ptr(u), size,
                                                           _dev._ptr(idx), _dev._stream(torch)))
        return idx
    return draw


def interpolate_with_ot(p0, p1, tmap, interp_frac, size):
    """``size`` cells between p0 and p1 at ``interp_frac``, pairs drawn with
    probability tmap / colsum^(1 - frac) (util.py:109-148)."""
    torch = _dev.torch_mod()
    p0, p1 = _dense64(p0), _dense64(p1)
    if p0.shape[1] != p1.shape[1]:
        raise ValueError("Unable to interpolate. Number of genes do not match")
    T = tmap if isinstance(tmap, torch.Tensor) else np.asarray(tmap, dtype=np.float64)
    if p0.shape[0] != T.shape[0] or p1.shape[0] != T.shape[1]:
        raise ValueError("Unable to interpolate. Tmap size is {}, expected {}"
                         .format(tuple(T.shape), (len(p0), len(p1))))
    Td = _dev.as_device_f64(torch, T, 'cuda')
    I, J = Td.shape
    W = torch.empty((I, J), dtype=torch.float64, device='cuda')
    colsum = torch.empty(J, dtype=torch.float64, device='cuda')
    _native.check(_native.lib().wotk_interp_weights(_dev._ptr(Td), Td.stride(0), I, J, float(interp_frac),
                                                    _dev._ptr(W), _dev._ptr(colsum), _dev._stream(torch)))
    choices = _choice_weighted(W)(int(size))
    return _sample_rows(p0, p1, choices, interp_frac)


def interpolate_randomly(p0, p1, t_interpolate, size):
    """``size`` cells between uniformly drawn pairs (util.py:151-179)."""
    torch = _dev.torch_mod()
    p0, p1 = _dense64(p0), _dense64(p1)
    if p0.shape[1] != p1.shape[1]:
        raise ValueError("Unable to interpolate. Number of genes do not match")
    choices = torch.from_numpy(np.random.randint(0, len(p0) * len(p1), size=int(size)).astype(np.int64)).cuda()
    return _sample_rows(p0, p1, choices, t_interpolate)


def interpolate_randomly_with_growth(p0, p1, t, size, g):
    """Pairs drawn with probability g_i * mean(g) (util.py:182-199)."""
    torch = _dev.torch_mod()
    p0, p1 = _dense64(p0), _dense64(p1)
    g = np.asarray(g, dtype=np.float64)
    q = np.ones(p1.shape[0]) * np.average(g)
    W = torch.outer(torch.from_numpy(g).cuda(), torch.from_numpy(q).cuda())
    choices = _choice_weighted(W)(int(size))
    return _sample_rows(p0, p1, choices, t)


def _host_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        n = os.cpu_count() or 1
    return max(1, min(16, n))


class _EmdRunner:
    """Exact EMDs of one summary: costs built on the device in call order
    (calling thread), network simplex on ``workers`` host threads; at most
    2 x workers cost matrices are held at once.  ``submit`` returns the
    distance (one worker) or a Future of it."""

    def __init__(self, workers):
        self.pool = ThreadPoolExecutor(workers) if workers > 1 else None
        self.limit = 2 * workers
        self.pending = collections.deque()

    def submit(self, cloud1, cloud2, eigenvals):
        p, q, M = _emd.emd_inputs(cloud1, cloud2, eigenvals)
        if self.pool is None:
            return float(np.sqrt(_emd.emd2(p, q, M, numItermax=int(1e7))))
        fut = self.pool.submit(lambda: float(np.sqrt(_emd.emd2(p, q, M, numItermax=int(1e7)))))
        self.pending.append(fut)
        while len(self.pending) > self.limit:
            self.pending.popleft().result()
        return fut

    def close(self):
        if self.pool is not None:
            self.pool.shutdown(wait=True)


def split_anndata(ds, field):
    """{value: rows of ds with obs[field] == value} in order of appearance (wot.split_anndata)."""
    out = {}
    for v in pd.unique(ds.obs[field].values):
        out[v] = ds[(ds.obs[field] == v).values, :]
    return out


def compute_validation_summary(ot_model, day_triplets=None, interp_size=10000, compute_full_distances=False,
                               emd_workers=None):
    """Validation summary (optimal_transport_validation.py:15-196): one row
    per distance (interval_start, interval_mid, interval_end, t0, t1, cv0,
    cv1, name in P / I / F / L / R / Rg (/ I1 / I2 / A when full), distance,
    full).  ``emd_workers``: host threads for the EMDs (default: the cores
    this process may use, at most 16; 1 runs them in line)."""
    runner = _EmdRunner(_host_threads() if emd_workers is None else max(1, int(emd_workers)))
    try:
        df = _summary(ot_model, day_triplets, interp_size, compute_full_distances, runner)
    finally:
        runner.close()
    if 'distance' in df:
        df['distance'] = [d.result() if isinstance(d, Future) else d for d in df['distance']]
    return df


def _summary(ot_model, day_triplets, interp_size, compute_full_distances, runner):
    if ot_model.covariate_field not in ot_model.matrix.obs:
        ot_model.matrix.obs['covariate'] = 1
    if day_triplets is None:
        tp = np.array(ot_model.timepoints)
        day_triplets = [(tp[i], tp[i + 1], tp[i + 2]) for i in range(len(tp) - 2)]
    ot_model.day_pairs = {(tr[0], tr[2]): {} for tr in day_triplets}
    has_covariate = ot_model.covariate_field is not None
    if not has_covariate and not compute_full_distances:
        raise ValueError('No covariate specified. Please provide a covariate or compute full distances')
    rows = []
    local_pca = ot_model.ot_config['local_pca']
    day_field = ot_model.day_field
    growth_field = ot_model.cell_growth_rate_field
    obs_all = ot_model.matrix.obs
    for (t0, t05, t1) in day_triplets:
        interp_frac = (t05 - t0) / (t1 - t0)
        M = ot_model.matrix
        p0_ds = M[(M.obs[day_field] == float(t0)).values, :]
        p05_ds = M[(M.obs[day_field] == float(t05)).values, :]
        p1_ds = M[(M.obs[day_field] == float(t1)).values, :]
        eigenvals = None
        if local_pca > 0:
            x0, x1 = _dense64(p0_ds.X), _dense64(p1_ds.X)
            p0_pca, p1_pca, pca, _ = compute_pca(x0, x1, local_pca)
            k = p0_pca.shape[1]
            var = pd.DataFrame(index=pd.RangeIndex(start=0, stop=k, step=1).astype(str))
            eigenvals = np.diag(pca.singular_values_)
            # the middle day projected with the reference's formula (optimal_transport_validation.py:79-84):
            # U = [x0; x1]^T V diag(1/S), p05 = diag(1/S) U^T (x05 - mean_shift)^T, as device GEMMs
            torch = _dev.torch_mod()
            X = torch.cat([_dev.as_device_f64(torch, x0, 'cuda'), _dev.as_device_f64(torch, x1, 'cuda')], 0)
            U = (X.T @ pca.V) / pca.S[None, :]
            y = _dev.as_device_f64(torch, _dense64(p05_ds.X), 'cuda') - pca.gene_mean[None, :]
            p05_x = ((y @ U) / pca.S[None, :]).cpu().numpy()
            del X, U, y
            p0_ds = _anndata.AnnData(p0_pca, p0_ds.obs, var)
            p1_ds = _anndata.AnnData(p1_pca, p1_ds.obs, var)
            p05_ds = _anndata.AnnData(p05_x, p05_ds.obs, var)

        def record(dist, t, name, cv0, cv1, full):
            rows.append({'interval_start': t0, 'interval_mid': t05, 'interval_end': t1, 't0': t, 't1': t05,
                         'cv0': cv0, 'cv1': cv1, 'name': name, 'distance': dist, 'full': full})

        if compute_full_distances:
            tmap_full = ot_model.compute_transport_map(t0, t1)

            def full(pop, t, name, pop2=p05_ds.X):
                record(runner.submit(pop, pop2, eigenvals), t, name, '', '', True)

            if growth_field in obs_all:
                g = p0_ds.obs[growth_field].values ** interp_frac
                full(interpolate_randomly_with_growth(p0_ds.X, p1_ds.X, interp_frac, interp_size, g), t05, 'Rg')
            full(interpolate_randomly(p0_ds.X, p1_ds.X, interp_frac, interp_size), t05, 'R')
            try:
                i05 = interpolate_with_ot(p0_ds.X, p1_ds.X, tmap_full.X, interp_frac, interp_size)
                full(i05, t05, 'I')
                full(i05, t05, 'I1', p0_ds.X)
                full(i05, t05, 'I2', p1_ds.X)
            except ValueError:
                pass
            full(p0_ds.X, t0, 'F')
            full(p1_ds.X, t1, 'L')
            full(p0_ds.X, t1, 'A', p1_ds.X)

        if not has_covariate:
            continue
        cvf = ot_model.covariate_field
        p0, p05, p1 = split_anndata(p0_ds, cvf), split_anndata(p05_ds, cvf), split_anndata(p1_ds, cvf)
        for cv05 in p05.keys():
            p05_x = p05[cv05].X
            seen_first, seen_last = set(), set()

            def to_p05(pop, t, name, cv):
                record(runner.submit(pop, p05_x, eigenvals), t, name, cv, cv05, False)

            for cv05_2 in p05.keys():
                if cv05_2 != cv05:
                    to_p05(p05[cv05_2].X, t05, 'P', cv05_2)
            for cv0, cv1 in itertools.product(p0.keys(), p1.keys()):
                tmap = ot_model.compute_transport_map(t0, t1, covariate=(cv0, cv1))
                if tmap is None:
                    continue
                p0_x, p1_x = p0[cv0].X, p1[cv1].X
                i05 = interpolate_with_ot(p0_x, p1_x, tmap.X, interp_frac, interp_size)
                if growth_field in obs_all:
                    g = p0[cv0].obs[growth_field].values ** interp_frac
                    to_p05(interpolate_randomly_with_growth(p0_x, p1_x, interp_frac, interp_size, g), t05, 'Rg',
                           (cv0, cv1))
                r05 = interpolate_randomly(p0_x, p1_x, interp_frac, interp_size)
                to_p05(i05, t05, 'I', (cv0, cv1))
                to_p05(r05, t05, 'R', (cv0, cv1))
                if cv0 == cv05 and cv0 not in seen_first:
                    seen_first.add(cv0)
                    to_p05(p0_x, t0, 'F', cv0)
                if cv1 == cv05 and cv1 not in seen_last:
                    seen_last.add(cv1)
                    to_p05(p1_x, t1, 'L', cv1)
    return pd.DataFrame(rows)
