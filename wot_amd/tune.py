"""Autotuner for the tile shapes of the Sinkhorn sweep kernels.

Every sweep kernel is compiled in several tile variants (cost-matrix rows per
workgroup x threads per workgroup); the library keeps one variant per
(kernel, shape bucket) and every launch looks its tile up by the bucket of
its shape (``wotk_set_tile`` / ``wotk_get_tile`` in include/wotk.h).  The
tuner times each variant on the caller's own workload with the library's
device-side sweep timer (hipEvents around the sweep loop), stores the fastest
in the library and persists it per device and bucket in a JSON file, which
later processes load before their first solve.

    tune.autotune_row_sweep(cost)          # one stored-cost day pair
    tune.autotune_time_course(tc, eps, l1) # the batched time course

The cache file is ``$WOTK_TILE_CACHE`` (default ``~/.cache/wot_amd/tiles.json``);
``WOTK_TILE_CACHE=off`` disables loading and saving.  A variant changes only
the order of the FP32 partial sums inside a row (the results agree to
rounding); it never changes which terms are summed.
"""
import ctypes
import json
import os
import threading

from . import _native

ROW_SWEEP = _native.KERNEL_ROW_SWEEP
BATCH_SWEEP = _native.KERNEL_BATCH_SWEEP
BATCH_SHARED = _native.KERNEL_BATCH_SHARED
KERNEL_NAMES = {ROW_SWEEP: 'row_sweep', BATCH_SWEEP: 'batch_sweep', BATCH_SHARED: 'batch_shared'}

_lock = threading.Lock()
_applied = False


def variants(kernel):
    """Compiled tile variants of ``kernel`` as (rows, threads); entry 0 is the default."""
    L = _native.lib()
    n = L.wotk_tile_variants(kernel, None, 0)
    if n < 0:
        _native.check(_native.WOTK_EINVAL)
    arr = (_native.Tile * n)()
    L.wotk_tile_variants(kernel, arr, n)
    return [t.as_tuple() for t in arr]


def bucket(kernel, n_rows=0, n_cols=0, n_combos=1):
    b = _native.lib().wotk_shape_bucket(kernel, int(n_rows), int(n_cols), int(n_combos))
    if b < 0:
        raise ValueError('unknown kernel %r' % (kernel,))
    return b


def get_tile(kernel, b):
    t = _native.Tile()
    _native.check(_native.lib().wotk_get_tile(kernel, int(b), ctypes.byref(t)))
    return t.as_tuple()


def set_tile(kernel, b, tile):
    """Set the tile of bucket ``b`` (-1: every bucket); ``tile=None`` restores the default."""
    ptr = None if tile is None else ctypes.byref(_native.Tile(int(tile[0]), int(tile[1])))
    _native.check(_native.lib().wotk_set_tile(kernel, int(b), ptr))


def reset():
    for k in KERNEL_NAMES:
        set_tile(k, -1, None)


# ---------------------------------------------------------------------------
# persistence
# ---------------------------------------------------------------------------
def cache_path():
    p = os.environ.get('WOTK_TILE_CACHE', '')
    if p.lower() in ('off', '0', 'none'):
        return None
    return p or os.path.join(os.path.expanduser('~'), '.cache', 'wot_amd', 'tiles.json')


def device_key():
    import torch
    prop = torch.cuda.get_device_properties(torch.cuda.current_device())
    arch = getattr(prop, 'gcnArchName', '')
    return '%s/%s/abi%d' % (prop.name, arch, _native.lib().wotk_abi_version())


def _read(path):
    try:
        with open(path) as fh:
            data = json.load(fh)
        return data if isinstance(data, dict) and isinstance(data.get('entries'), dict) else {'entries': {}}
    except (OSError, ValueError):
        return {'entries': {}}


def _key(dev, kernel, b):
    return '%s|%s|%d' % (dev, KERNEL_NAMES[kernel], b)


def apply_persisted(force=False):
    """Load this device's tuned tiles from the cache file into the library (once per process)."""
    global _applied
    with _lock:
        if _applied and not force:
            return 0
        _applied = True
        path = cache_path()
        if path is None or not os.path.exists(path):
            return 0
        dev = device_key()
        valid = {k: set(variants(k)) for k in KERNEL_NAMES}
        names = {v: k for k, v in KERNEL_NAMES.items()}
        n = 0
        for key, ent in _read(path)['entries'].items():
            parts = key.split('|')
            if len(parts) != 3 or parts[0] != dev or parts[1] not in names:
                continue
            kernel, b = names[parts[1]], int(parts[2])
            tile = (int(ent['rows']), int(ent['threads']))
            if tile in valid[kernel]:
                set_tile(kernel, b, tile)
                n += 1
        return n


def _persist(kernel, buckets, tile, timings):
    path = cache_path()
    if path is None:
        return
    dev = device_key()
    with _lock:
        data = _read(path)
        for b in buckets:
            data['entries'][_key(dev, kernel, b)] = {
                'rows': tile[0], 'threads': tile[1],
                'us_per_sweep': {'%dx%d' % t: round(v, 3) for t, v in timings.items()}}
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        tmp = path + '.tmp%d' % os.getpid()
        with open(tmp, 'w') as fh:
            json.dump(data, fh, indent=1, sort_keys=True)
        os.replace(tmp, path)


class TuneResult:
    def __init__(self, kernel, buckets, best, timings):
        self.kernel = kernel
        self.buckets = buckets
        self.best = best            # (rows, threads)
        self.timings = timings      # (rows, threads) -> us per sweep launch

    def as_dict(self):
        return {'kernel': KERNEL_NAMES[self.kernel], 'buckets': list(self.buckets),
                'tile': {'rows': self.best[0], 'threads': self.best[1]},
                'us_per_sweep': {'%dx%d' % t: round(v, 3) for t, v in self.timings.items()}}


def _pick(kernel, buckets, measure, repeats, persist):
    apply_persisted()  # load the cache now, so a later first solve cannot overwrite the trial tiles
    timings = {}
    for tile in variants(kernel):
        for b in buckets:
            set_tile(kernel, b, tile)
        measure()  # warm-up: first launch of this variant loads its code object
        timings[tile] = min(measure() for _ in range(repeats))
    best = min(timings, key=timings.get)
    for b in buckets:
        set_tile(kernel, b, best)
    if persist:
        _persist(kernel, buckets, best, timings)
    return TuneResult(kernel, buckets, best, timings)


def autotune_row_sweep(cost, G=None, epsilon=0.05, lambda1=1.0, lambda2=50.0, iters=20, repeats=3, persist=True):
    """Tune the stored-cost sweep on ``cost`` (an ``ot.device.DeviceCost``):
    each variant runs ``iters`` fixed-epsilon Sinkhorn iterations (2 sweeps
    each) and is timed by the library's sweep timer.  Both sweep shapes of the
    pair (C and C^T) get the winning tile."""
    import torch
    from .ot import device as dev
    n_i, n_j = cost.shape
    if G is None:
        G = torch.ones(n_i, dtype=torch.float64, device=cost.device)
    params = dict(lambda1=lambda1, lambda2=lambda2, epsilon=epsilon, scaling_iter=int(iters), tau=1e4,
                  epsilon0=epsilon, extra_iter=0, inner_iter_max=int(iters))
    buckets = sorted({bucket(ROW_SWEEP, n_i, n_j), bucket(ROW_SWEEP, n_j, n_i)})

    def measure():
        out = dev.solve(cost, G, 'fixed_iters', params, want_plan=False)
        st = out['stats']
        return 1e3 * st['sweep_ms'] / max(1, st['sweeps'])

    return _pick(ROW_SWEEP, buckets, measure, repeats, persist)


def autotune_time_course(tc, epsilons, lambda1s, lambda2=50.0, iters=20, repeats=2, share_exp=True, persist=True):
    """Tune the batched sweep on a prepared time course (``ot.batch.TimeCourseCosts``)
    for the grid epsilons x lambda1s: each variant runs ``iters`` iterations
    of the whole batch; timed per half-iteration launch.  Tunes the kernel the
    grid runs on (the shared-exponential one when combinations of equal
    epsilon can share, else the per-combination one)."""
    from .ot import batch
    E = len(epsilons) * len(lambda1s)
    kw = dict(lambda2=lambda2, extra_iter=0, inner_iter_max=50, graph_iters=0, share_exp=share_exp)
    probe = batch.solve_prepared(tc, epsilons, lambda1s, scaling_iter=1, **kw)
    kernel = BATCH_SHARED if probe.stats['shared_group'] else BATCH_SWEEP
    buckets = [bucket(kernel, 0, tc.mean_row_length(), E)]

    def measure():
        r = batch.solve_prepared(tc, epsilons, lambda1s, scaling_iter=int(iters), **kw)
        return 1e3 * r.stats['sweep_ms'] / max(1, r.stats['sweeps'])

    return _pick(kernel, buckets, measure, repeats, persist)
