"""Benchmark: transport maps/sec (and Sinkhorn iterations/sec) of the fused HIP
solver on one MI355X, versus the plain-GPU torch FP64 restatement and the
FP64 CPU oracle.

Default workload (``--config timecourse``): a 38-pair time course (39 days,
cells per day ~ U[1000, 20000], 30-d PCA coordinates from a seeded 20-component
Gaussian mixture drifting in time) swept over eps {0.01, 0.05, 0.1} x lambda1
{0.5, 1, 5}: 342 fixed_iters solves (lambda2=50, eps0=1, scaling_iter=3000,
inner_iter_max=50, extra_iter=1000) advancing together in one persistent
launch per half-iteration.  One step = 200 iterations of every solve; 20
steps = one complete course, costs and medians built on the device in the
first step and every FP64 map written once in HBM after the last.  Inputs
(coordinates, growth) are resident in HBM before the timed region.
``--config pair`` times whole 15,000 x 15,000 maps (growth_iters=3) instead;
the other configs are the remaining workloads.

Prints ONE JSON line.  Usage:  python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
CFG = dict(n_source=15000, n_target=15000, dim=30, n_components=10, seed=1)
SOLVER = dict(epsilon=0.05, lambda1=1.0, lambda2=50.0, epsilon0=1.0, tau=1e4, scaling_iter=3000,
              inner_iter_max=50, extra_iter=1000, growth_iters=3)


def _traffic(name='lse_rows_kernel_15000x15000'):
    """HBM bytes per launch of a kernel from the calibrated PMC runs kept in profiles/."""
    path = os.path.join(ROOT, 'profiles', name + '.traffic.json')
    try:
        with open(path) as fh:
            return json.load(fh)['traffic_bytes']
    except (OSError, ValueError, KeyError):
        return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def one_map(ot, dev, x, y, growth, cost_mode='bf16x6'):
    """cost + median on the device, growth loop, final FP64 plan in HBM."""
    cost = dev.DeviceCost.from_coordinates(x, y, mode=cost_mode)
    params = dict(SOLVER, C=cost, G=growth, device_out=True)
    sweep_ms, sweeps = 0.0, 0
    G = params['G']
    res = None
    for it in range(SOLVER['growth_iters']):
        if it > 0:
            G = res['row_sums']
        last = it == SOLVER['growth_iters'] - 1
        res = dev.solve(cost, G, 'fixed_iters', params, want_plan=last, want_row_sums=not last,
                        want_potentials=last)
        sweep_ms += res['stats']['sweep_ms']
        sweeps += res['stats']['sweeps']
    res['G_last'] = G
    return cost, res, sweep_ms, sweeps


def cost_gemm_report(torch, x, y, G=None, solve=None):
    """Split-BF16 MFMA costs (3 and 6 terms) against the FP64 direct cost on
    one pair: error of the normalised FP32 cost (max, RMS, RMS over the nearest
    1 %% of pairs), median error, cost-build wall times, and, if ``solve``
    (fixed_iters parameters) is given, the relative Frobenius error of the plan
    computed from each split cost against the plan from the FP64 cost."""
    import wot_amd.ot as ot
    from wot_amd.ot import device as dev
    out = {'shape': [int(x.shape[0]), int(y.shape[0]), int(x.shape[1])]}
    costs = {}
    for mode in ('f64', 'bf16x3', 'bf16x6'):
        dev.DeviceCost.from_coordinates(x, y, mode=mode)  # warm-up (code objects, allocator)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        costs[mode] = dev.DeviceCost.from_coordinates(x, y, mode=mode)
        torch.cuda.synchronize()
        out['build_s_' + mode] = time.perf_counter() - t0
    cf = costs['f64'].dense()
    n_i, n_j = cf.shape
    # 1 %% quantile of the cost from a fixed sample of entries: "nearest pairs" threshold
    gen = torch.Generator(device=cf.device).manual_seed(0)
    idx = torch.randint(0, n_i * n_j, (min(n_i * n_j, 1 << 20),), device=cf.device, generator=gen)
    thr = torch.quantile(cf[idx // n_j, idx % n_j].double(), 0.01)
    step = max(1, (1 << 27) // n_j)  # row chunks of ~128 M entries
    Rf = ot.transport_stablev2(costs['f64'], G=G, device_out=True, **solve) if solve is not None else None
    for mode in ('bf16x3', 'bf16x6'):
        cb = costs[mode].dense()
        mx, ss, ss_near, n_near = 0.0, 0.0, 0.0, 0
        for r0 in range(0, n_i, step):
            a = cf[r0:r0 + step].double()
            d = cb[r0:r0 + step].double() - a
            mx = max(mx, float(d.abs().max()))
            ss += float(d.pow(2).sum())
            m = a <= thr
            ss_near += float(d[m].pow(2).sum())
            n_near += int(m.sum())
            del a, d, m
        r = {'max_abs_err': mx, 'rms_err': (ss / (n_i * n_j)) ** 0.5,
             'rms_err_nearest_1pct': (ss_near / max(n_near, 1)) ** 0.5,
             'median_rel_err': float((costs[mode].median - costs['f64'].median).abs() / costs['f64'].median)}
        if Rf is not None:
            Rb = ot.transport_stablev2(costs[mode], G=G, device_out=True, **solve)
            r['plan_rel_frob_vs_f64_cost'] = float(torch.linalg.norm(Rb - Rf) / torch.linalg.norm(Rf))
            del Rb
        out[mode] = r
    del costs, cf, Rf
    torch.cuda.empty_cache()
    return out


def _tune_pair(args, dev, x, y, g):
    """Autotune the row-sweep tile on this pair's cost (untimed)."""
    if args.no_tune:
        return None
    from wot_amd import tune
    cost = dev.DeviceCost.from_coordinates(x, y)
    r = tune.autotune_row_sweep(cost, G=g, epsilon=SOLVER['epsilon'], lambda1=SOLVER['lambda1'],
                                lambda2=SOLVER['lambda2'], persist=False)
    log('tiles: %s' % json.dumps(r.as_dict()))
    return r.as_dict()


def plain_gpu_sample(torch, x, y, growth):
    """Plain GPU (torch FP64 dense K, one pair at a time): one full fixed_iters
    solve (4000 iterations, 61 kernel rebuilds) at full size, timed after a
    short warm-up; a map is three such solves (growth_iters=3) plus the cost.
    Returns (seconds per map, detail, plan of that solve)."""
    from wot_amd import plain_gpu
    from wot_amd.ot import device as dev
    X = torch.from_numpy(x).cuda()
    Y = torch.from_numpy(y).cuda()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    c64 = dev.sqeuclid_f64(X, Y)
    v = torch.sort(c64.view(-1)).values  # numpy median: mean of the two middle values
    n = v.numel()
    med = (v[n // 2 - 1] + v[n // 2]) / 2 if n % 2 == 0 else v[n // 2]
    C = c64 / med
    del v
    torch.cuda.synchronize()
    t_cost = time.perf_counter() - t0
    plain_gpu.transport_stablev2(C, G=growth, **dict(SOLVER, scaling_iter=5, extra_iter=0))  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plan = plain_gpu.transport_stablev2(C, G=growth, **SOLVER)
    torch.cuda.synchronize()
    t_solve = time.perf_counter() - t0
    per_map = t_cost + SOLVER['growth_iters'] * t_solve
    return per_map, dict(t_cost_s=t_cost, t_solve_s=t_solve), plan


def cpu_sample(x, y, growth):
    """FP64 numpy oracle (the reference algorithm restated) timed on the host
    cores on a bounded sample of the same map: cost + median at full size,
    then one reduced solve whose two parts are timed apart -- 50 scheduled
    iterations (2 kernel builds) and 50 extra iterations (no builds) -- and
    extrapolated to 3 x (4000 iterations + 61 builds)."""
    from oracle import ot_oracle as O
    try:
        from threadpoolctl import threadpool_info
        threads = max([i.get('num_threads', 1) for i in threadpool_info()] + [1])
    except Exception:  # pragma: no cover
        threads = os.cpu_count()
    t0 = time.perf_counter()
    C = O.default_cost_matrix(x, y)
    t_cost = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.transport_stablev2(C, G=growth, **dict(SOLVER, scaling_iter=50, extra_iter=0))
    t_a = time.perf_counter() - t0   # 2 builds + 50 iterations
    t0 = time.perf_counter()
    O.transport_stablev2(C, G=growth, **dict(SOLVER, scaling_iter=50, extra_iter=50))
    t_b = time.perf_counter() - t0   # 2 builds + 100 iterations
    t_it = max((t_b - t_a) / 50.0, 0.0)
    t_build = max((t_a - 50 * t_it) / 2.0, 0.0)
    n_build = 1 + SOLVER['scaling_iter'] // SOLVER['inner_iter_max']
    n_iter = SOLVER['scaling_iter'] + SOLVER['extra_iter']
    per_map = t_cost + SOLVER['growth_iters'] * (n_build * t_build + n_iter * t_it)
    return per_map, threads, dict(t_cost_s=t_cost, t_iter_s=t_it, t_build_s=t_build, sample_wall_s=t_cost + t_a + t_b)


TC = dict(n_days=39, min_cells=1000, max_cells=20000, dim=30, n_components=20, seed=2)
GRID = dict(epsilons=(0.01, 0.05, 0.1), lambda1s=(0.5, 1.0, 5.0))


STEP_ITERS = 200  # iterations of every solve per step: 20 steps = one full 3000 + 1000 schedule


def host_info():
    """Host cores and CPU model (for the CPU baseline)."""
    model = None
    try:
        with open('/proc/cpuinfo') as fh:
            for line in fh:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:  # pragma: no cover
        pass
    try:
        from threadpoolctl import threadpool_info
        threads = max([i.get('num_threads', 1) for i in threadpool_info()] + [1])
    except Exception:  # pragma: no cover
        threads = os.cpu_count()
    return {'nproc': os.cpu_count(), 'blas_threads': threads, 'cpu_model': model}


def valu_rates(torch):
    """Issue rates of the sweep's instructions on this device (libwotk probe)."""
    import ctypes
    from wot_amd import _native
    from wot_amd.ot import device as dev
    L = _native.lib()
    sink = torch.empty(L.wotk_probe_sink_bytes() // 4, dtype=torch.float32, device='cuda')
    rates = (ctypes.c_double * 7)()
    _native.check(L.wotk_probe_valu_rates(rates, 7, dev._ptr(sink), sink.numel() * 4, dev._stream(torch)))
    names = ['exp_f32', 'fma_f64', 'fma_f32', 'add_f64_cvt_pair', 'rndne_f64_add', 'ldexp_f32', 'split_exp_entry']
    out = {n: float(v) for n, v in zip(names, rates)}
    mix = (ctypes.c_double * 4)()
    _native.check(L.wotk_probe_issue_mix(mix, 4, dev._ptr(sink), sink.numel() * 4, dev._stream(torch)))
    out['fma29_exp1_mix_by_waves_per_simd'] = {str(w): float(v) for w, v in zip((1, 2, 4, 8), mix)}
    return out


def bench_timecourse(args):
    """38 consecutive day pairs (cells per day ~ U[1000, 20000], 30-d, 20
    drifting components) x the grid eps {0.01, 0.05, 0.1} x lambda1 {0.5, 1, 5}
    = 342 fixed_iters solves advancing together (one persistent launch per
    half-iteration, hipGraph replay of 50-iteration blocks).

    A step is STEP_ITERS (200) Sinkhorn iterations of all 342 solves, so 20
    steps are one complete time course: every pair's cost and median built on
    the device from coordinates already in HBM (in the first step), the full
    3000 + 1000 schedule, the final potentials, and every FP64 map
    materialised in HBM once (one at a time into a reused buffer, after the
    last step).  value = solves per second = 342 x (steps x 200 / 4000) /
    elapsed, i.e. exactly 342 / elapsed at the default 20 steps."""
    import torch
    from wot_amd.ot import batch
    from wot_amd.synthetic import time_course
    coords, growth, counts = time_course(**TC)
    sched = dict(SOLVER)
    sched.pop('growth_iters')
    for k in ('epsilon', 'lambda1'):
        sched.pop(k)
    if args.iters is not None:
        sched['extra_iter'] = sched['extra_iter'] * args.iters // sched['scaling_iter']
        sched['scaling_iter'] = args.iters
    if args.inner_iter_max is not None:
        sched['inner_iter_max'] = args.inner_iter_max
    n_iter = sched['scaling_iter'] + sched['extra_iter']
    step_iters = min(STEP_ITERS, n_iter)
    steps_per_course = -(-n_iter // step_iters)
    pair_ij = [int(counts[t]) * int(counts[t + 1]) for t in range(len(counts) - 1)]
    sum_ij = float(sum(pair_ij))
    n_combo = len(GRID['epsilons']) * len(GRID['lambda1s'])
    n_solves = len(pair_ij) * n_combo
    # inputs resident in HBM before anything is timed
    dcoords = [torch.from_numpy(np.ascontiguousarray(c)).cuda() for c in coords]
    dgrowth = [torch.from_numpy(np.ascontiguousarray(g)).cuda() for g in growth]
    combos, eps_a, l1_a, l2_a = batch.grid_arrays(len(pair_ij), GRID['epsilons'], GRID['lambda1s'],
                                                  sched['lambda2'])
    run_kw = dict(scaling_iter=sched['scaling_iter'], epsilon0=sched['epsilon0'], extra_iter=sched['extra_iter'],
                  inner_iter_max=sched['inner_iter_max'], graph_iters=50, combos=combos,
                  share_eps=args.share_eps == 'on')
    tiles = None
    if args.tile:
        from wot_amd import tune
        tune.set_tile(tune.BATCH_SHARED, -1, _parse_tile(args.tile))
        tiles = {'kernel': 'batch_shared', 'tile': dict(zip(('rows', 'threads'), _parse_tile(args.tile)))}
    elif not args.no_tune:
        from wot_amd import tune
        tc = batch.prepare_time_course(dcoords, dgrowth)
        tiles = tune.autotune_time_course(tc, GRID['epsilons'], GRID['lambda1s'], lambda2=sched['lambda2'],
                                          persist=False).as_dict()
        del tc
        torch.cuda.empty_cache()
        log('tiles: %s' % json.dumps(tiles))
    log('time course: %d pairs, %d combinations, sum I*J = %.3e, %d iterations, %d per step' %
        (len(pair_ij), n_combo, sum_ij, n_iter, step_iters))
    if args.warmup:
        tc = batch.prepare_time_course(dcoords, dgrowth)
        with batch.TimeCourseRun(tc, eps_a, l1_a, l2_a, **run_kw) as run:
            for w in range(args.warmup):
                run.advance(step_iters)
                torch.cuda.synchronize()
                log('warmup %d' % w)
        del tc
        torch.cuda.empty_cache()
    torch.cuda.synchronize()
    plan_buf = torch.empty(max(pair_ij), dtype=torch.float64, device='cuda')
    sweep_ms, sweeps, exact_rows, shared, derived_iters, modes = 0.0, 0, 0, None, None, None
    res = tc = run = None
    courses = 0
    t0 = time.perf_counter()
    for s in range(args.steps):
        if run is None:
            tc = batch.prepare_time_course(dcoords, dgrowth)
            run = batch.TimeCourseRun(tc, eps_a, l1_a, l2_a, **run_kw)
        left = run.advance(step_iters)
        if left == 0:
            res = run.result()
            modes = run.modes()
            derived_iters = int((modes != 0).sum())
            for p in range(len(pair_ij)):
                n_i, n_j = res.shapes[p]
                for e in range(n_combo):
                    res.plan(p, e, out=plan_buf[:n_i * n_j].view(n_i, n_j))
            run.close()
            run = None
            courses += 1
            sweep_ms += res.stats['sweep_ms']
            sweeps += res.stats['sweeps']
            exact_rows += res.stats['exact_rows']
            shared = res.stats['shared_group']
        torch.cuda.synchronize()
        log('step %d done' % s)
    elapsed = time.perf_counter() - t0
    if run is not None:  # a course left unfinished by the step count
        part = run.result()
        sweep_ms += part.stats['sweep_ms']
        sweeps += part.stats['sweeps']
        exact_rows += part.stats['exact_rows']
        shared = part.stats['shared_group']
        run.close()
        run = None
    done_fraction = args.steps * step_iters / float(n_iter)
    value = n_solves * done_fraction / elapsed
    avg_sweep = (sweep_ms / 1000.0) / max(sweeps, 1)  # per half-iteration over the whole grid
    # algorithmic work of one sweep over the whole batch: 4 bytes per cost entry
    # (one read shared by all combinations) and one exp per (entry, combination)
    alg_bytes = 4.0 * sum_ij
    achieved = alg_bytes / avg_sweep / 1e9
    # exponentials per cost entry over the course's sweep modes (wotk_batch_modes): one per eps
    # group on its own or real-ratio exponent, none for an integer power of the base's
    mode_hist, exps_per_entry = None, float(len(GRID['epsilons']))
    if modes is not None and len(modes):
        def n_exp(m):
            return len(GRID['epsilons']) if m == 0 else 1 + int((m & 0xff) == 255) + int((m >> 8) == 255)
        mode_hist = {str(int(m)): int((modes == m).sum()) for m in np.unique(modes)}
        exps_per_entry = float(np.mean([n_exp(int(m)) for m in modes]))
    rates = valu_rates(torch)
    entries_per_s = sum_ij / avg_sweep
    out = {
        'metric': 'transport maps/sec and Sinkhorn iters/sec at ≤1e-5 rel err, fused vs plain GPU',
        'value': value, 'unit': 'maps/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': 1000.0 * elapsed / args.steps, 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f64/f32',
        'data': 'synthetic (seeded drifting Gaussian mixture, 39 days, 30-d PCA coordinates resident in HBM)',
        'config': {'workload': '38-pair time course x eps{0.01,0.05,0.1} x lambda1{0.5,1,5} = 342 fixed_iters '
                               'solves, one step = %d iterations of every solve, %d steps = one course' %
                               (step_iters, steps_per_course),
                   'cells_total': int(sum(counts)), 'sum_IJ': sum_ij, 'solves': n_solves, **sched,
                   'step_iterations': step_iters, 'courses_completed': courses, 'tiles': tiles},
        'timed_region': 'per course: costs + medians on the device, the full schedule, potentials, every FP64 map '
                        'written once in HBM; no host copies (the 342 FP64 maps total %.0f GB; a D2H of them is '
                        'not timed)' % (8.0 * n_combo * sum_ij / 1e9),
        'sinkhorn_iters_per_s': n_solves * args.steps * step_iters / elapsed,
        'sweep_loop_s_per_course': (sweep_ms / 1000.0) / max(courses, 1) if courses else None,
        'shared_exponential_group': shared,
        'share_eps': args.share_eps,
        'iterations_with_derived_eps_groups': derived_iters,
        'rows_recomputed_exactly': exact_rows,
        'roofline': {'bound': 'hbm', 'kernel': _tc_kernel_name({'shared_group': shared}, tiles),
                     'avg_launch_us': avg_sweep * 1e6, 'alg_bytes_per_launch': alg_bytes,
                     'achieved': achieved, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS,
                     'traffic': _traffic('batch_shared_lds_kernel_timecourse_38pairs_9combos'),
                     'traffic_source': 'profiles/batch_shared_lds_kernel_timecourse_38pairs_9combos.traffic.json '
                                       '(calibrated FETCH_SIZE + WRITE_SIZE, not measured in this run)',
                     'compute': {'note': 'the sweep is VALU-issue-bound, not HBM-bound: one FP32 cost read '
                                         'serves 9 combinations; SQ counters put the VALU busy ~100 % in every '
                                         'sweep mode (profiles/batch_shared_lds_kernel_timecourse_38pairs_9combos.md)',
                                 'entries_per_s': entries_per_s,
                                 'entry_combinations_per_s': entries_per_s * n_combo,
                                 'exp_f32_per_entry_course_average': exps_per_entry,
                                 'exp_f32_per_s': entries_per_s * exps_per_entry,
                                 'measured_exp_f32_peak_per_s': rates['exp_f32'],
                                 'exp_frac': entries_per_s * exps_per_entry / rates['exp_f32'],
                                 'sweep_modes': mode_hist,
                                 'measured_rates_per_s': rates}},
    }
    if args.no_parity:
        return out
    if not args.no_baselines:
        out.update(_timecourse_baselines(torch, coords, growth, pair_ij, n_combo, sched, n_iter, 1.0 / value * n_solves,
                                         res))
    order = np.argsort(pair_ij)
    out['cost_split_bf16'] = [cost_gemm_report(torch, coords[p], coords[p + 1])
                              for p in (int(order[0]), int(order[len(order) // 2]), int(order[-1]))]
    out['secondary_config_1'] = _secondary_pair(torch)
    return out


def _secondary_pair(torch):
    """One full map of the 15k x 15k pair (growth_iters=3, 6-term split-BF16
    cost) timed once, with its marginal identity: the single-pair path."""
    import wot_amd.ot as ot
    from wot_amd.ot import device as dev
    from wot_amd.synthetic import day_pair
    dp = day_pair(CFG['n_source'], CFG['n_target'], dim=CFG['dim'], n_components=CFG['n_components'],
                  seed=CFG['seed'])
    x = torch.from_numpy(dp.x).cuda()
    y = torch.from_numpy(dp.y).cuda()
    g = torch.from_numpy(dp.growth).cuda()
    one_map(ot, dev, x, y, g)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cost, res, sweep_ms, sweeps = one_map(ot, dev, x, y, g)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    q = float(res['G_last'].mean())
    want = q * CFG['n_source'] * torch.exp(-res['g'] / SOLVER['lambda2'])
    colsum = res['plan'].sum(0) * CFG['n_target']
    marg = float((colsum - want).abs().max() / want.abs().max())
    avg = sweep_ms / 1000.0 / sweeps
    alg = 4.0 * CFG['n_source'] * CFG['n_target']
    return {'workload': 'one day pair 15000x15000, d=30, fixed_iters, growth_iters=3, 6-term split-BF16 cost',
            'maps_per_s': 1.0 / el, 'ms_per_map': 1000.0 * el, 'marginal_identity_rel_err': marg,
            'roofline': {'bound': 'hbm', 'kernel': 'lse_rows_kernel', 'avg_launch_us': avg * 1e6,
                         'achieved': alg / avg / 1e9, 'peak': HBM_PEAK_GBS, 'frac': alg / avg / 1e9 / HBM_PEAK_GBS}}


def _parse_tile(text):
    r, t = (int(v) for v in text.lower().split('x'))
    return r, t


def _tc_kernel_name(stats, tiles):
    """Name of the batch sweep kernel the run used (shared exponentials or not, tile)."""
    if not stats.get('shared_group'):
        return 'batch_sweep_kernel'
    t = (tiles or {}).get('tile')
    if t and t['rows'] > 2:
        return 'batch_shared_lds_kernel (%d rows x %d threads)' % (t['rows'], t['threads'])
    return 'batch_shared_kernel'


def _fit_and_sum(samples, pair_ij, n_combo):
    """least-squares t = a + b * I*J over timed samples, summed over all solves"""
    x = np.array([s[0] for s in samples], dtype=np.float64)
    t = np.array([s[1] for s in samples], dtype=np.float64)
    A = np.vstack([np.ones_like(x), x]).T
    a, b = np.linalg.lstsq(A, t, rcond=None)[0]
    a, b = max(a, 0.0), max(b, 0.0)
    return n_combo * sum(a + b * ij for ij in pair_ij), a, b


def _timecourse_baselines(torch, coords, growth, pair_ij, n_combo, sched, n_iter, course_s, res):
    """Plain GPU: the torch FP64 restatement solves the smallest, median and
    largest pairs in full (eps=0.05, lambda1=1; the solve time does not depend
    on the grid point), fitted as t = a + b*I*J and summed over all 342 solves;
    its three plans also give the parity of the batched maps of the same
    (pair, combination).  CPU: the FP64 numpy oracle solves the smallest pair
    in full and times 20 iterations + one kernel build on the median and
    largest pairs; iteration and build times are fitted linearly in I*J and
    summed over every solve's 4000 iterations and 61 builds."""
    from wot_amd import plain_gpu
    from wot_amd.ot import device as dev
    order = np.argsort(pair_ij)
    picks = [int(order[0]), int(order[len(order) // 2]), int(order[-1])]
    e_ref = res.combos.index((0.05, 1.0)) if res is not None else None
    samples, parity, parity_all = [], [], []
    for p in picks:
        X = torch.from_numpy(coords[p]).cuda()
        Y = torch.from_numpy(coords[p + 1]).cuda()
        c64 = dev.sqeuclid_f64(X, Y)
        v = torch.sort(c64.view(-1)).values
        n = v.numel()
        med = (v[n // 2 - 1] + v[n // 2]) / 2 if n % 2 == 0 else v[n // 2]
        C = c64 / med
        del v, c64
        plain_gpu.transport_stablev2(C, G=growth[p], epsilon=0.05, lambda1=1.0,
                                     **dict(sched, scaling_iter=3, extra_iter=0))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        R0 = plain_gpu.transport_stablev2(C, G=growth[p], epsilon=0.05, lambda1=1.0, **sched)
        torch.cuda.synchronize()
        samples.append((pair_ij[p], time.perf_counter() - t0))
        if res is not None:
            R1 = res.plan(p, e_ref, device_out=True)
            rel = float(torch.linalg.norm(R1 - R0) / torch.linalg.norm(R0))
            marg = max(float((R1.sum(a) - R0.sum(a)).abs().max() / R0.sum(a).abs().max()) for a in (0, 1))
            parity.append({'pair': p, 'shape': list(R0.shape), 'rel_frob': rel, 'marginal_rel': marg})
            del R1
        if res is not None and p == picks[-1]:
            # every combination of the largest pair (the factored and derived-group paths too)
            del R0
            for e, (ep, lam) in enumerate(res.combos):
                if e == e_ref:
                    continue
                R0 = plain_gpu.transport_stablev2(C, G=growth[p], epsilon=ep, lambda1=lam, **sched)
                R1 = res.plan(p, e, device_out=True)
                rel = float(torch.linalg.norm(R1 - R0) / torch.linalg.norm(R0))
                marg = max(float((R1.sum(a) - R0.sum(a)).abs().max() / R0.sum(a).abs().max()) for a in (0, 1))
                parity_all.append({'combination': [ep, lam], 'rel_frob': rel, 'marginal_rel': marg})
                del R0, R1
            R0 = None
        del C, R0
        torch.cuda.empty_cache()
        log('plain GPU sample pair %d (I*J=%d): %.2f s' % (p, pair_ij[p], samples[-1][1]))
    plain_total, a, b = _fit_and_sum(samples, pair_ij, n_combo)
    from oracle import ot_oracle as O
    lam2, eps0, inner = sched['lambda2'], sched['epsilon0'], sched['inner_iter_max']
    n_build = 1 + sched['scaling_iter'] // inner
    cpu_pts = []
    t_start = time.perf_counter()
    p = picks[0]
    C = O.default_cost_matrix(coords[p], coords[p + 1])
    t0 = time.perf_counter()
    O.transport_stablev2(C, 1.0, lam2, 0.05, sched['scaling_iter'], growth[p], 1e4, eps0, sched['extra_iter'], inner)
    t_full_small = time.perf_counter() - t0
    for p in picks[1:]:
        C = O.default_cost_matrix(coords[p], coords[p + 1])
        t0 = time.perf_counter()
        O.transport_stablev2(C, 1.0, lam2, 0.05, 1, growth[p], 1e4, eps0, 0, inner)   # 1 build + 1 iteration
        ta = time.perf_counter() - t0
        t0 = time.perf_counter()
        O.transport_stablev2(C, 1.0, lam2, 0.05, 1, growth[p], 1e4, eps0, 20, inner)  # 1 build + 21 iterations
        tb = time.perf_counter() - t0
        t_it = max((tb - ta) / 20.0, 0.0)
        cpu_pts.append((pair_ij[p], t_it, max(ta - t_it, 0.0)))
        del C
    sample_wall = time.perf_counter() - t_start
    # per-solve time of pair p: the small pair measured in full; the others from
    # the linear fits of iteration and build time in I*J through the two points
    (x1, i1, b1), (x2, i2, b2) = cpu_pts
    def lin(x, y1, y2):
        return y1 + (y2 - y1) * (x - x1) / max(x2 - x1, 1)
    per_solve = []
    for k, ij in enumerate(pair_ij):
        if k == picks[0]:
            per_solve.append(t_full_small)
        else:
            per_solve.append(n_iter * max(lin(ij, i1, i2), 0.0) + n_build * max(lin(ij, b1, b2), 0.0))
    cpu_total = n_combo * sum(per_solve)
    hi = host_info()
    n_solves = len(pair_ij) * n_combo
    out = {
        'plain_gpu': {'value': n_solves / plain_total, 'unit': 'maps/s',
                      'sample': 'torch FP64 dense-K restatement: full solves of the smallest, median and largest '
                                'pairs timed, t = a + b*I*J fitted and summed over all 342 solves',
                      'fit_a_s': a, 'fit_b_s': b, 'samples': samples},
        'speedup_vs_plain_gpu': plain_total / course_s,
        'cpu_baseline': {'value': n_solves / cpu_total, 'unit': 'maps/s', 'cores': hi['blas_threads'],
                         'kind': 'port', 'nproc': hi['nproc'], 'cpu_model': hi['cpu_model'],
                         'sample': 'FP64 numpy oracle: the smallest pair solved in full (%.1f s); 20 iterations and '
                                   'one kernel build timed on the median and largest pairs, fitted linearly in I*J '
                                   'and summed over 342 solves x (4000 iterations + %d builds); sample %.1f s' %
                                   (t_full_small, n_build, sample_wall),
                         'points_ij_iter_build_s': cpu_pts},
        'speedup_vs_cpu': cpu_total / course_s,
    }
    if parity:
        out['parity_vs_plain_gpu'] = {'combination': [0.05, 1.0], 'pairs': parity,
                                      'max_rel_frob': max(q['rel_frob'] for q in parity),
                                      'max_marginal_rel': max(q['marginal_rel'] for q in parity)}
    if parity_all:
        out['parity_vs_plain_gpu']['largest_pair_other_combinations'] = parity_all
        out['parity_vs_plain_gpu']['max_rel_frob_all'] = max(q['rel_frob'] for q in parity + parity_all)
        out['parity_vs_plain_gpu']['max_marginal_rel_all'] = max(q['marginal_rel'] for q in parity + parity_all)
    return out


def bench_adjoint(args):
    """Gradient of the converged objective w.r.t. 8000 x 30 FP32 source
    embeddings and 8000 FP32 log growth rates against 8000 target cells:
    forward (cost, converged duality-gap solve, FP64 primal) + adjoint kernel.
    Checks: FP64 autograd of the primal at the same plan, and central finite
    differences (h = 1e-4) of the FP64-converged objective on 64 random
    coordinates of a 300 x 360 sub-problem (FD at 8000 x 8000 would need the
    objective to ~1e-13, beyond an FP32 cost stream)."""
    import torch
    import wot_amd.ot as ot
    from wot_amd.synthetic import day_pair
    dp = day_pair(8000, 8000, dim=30, n_components=10, seed=5)
    x = torch.from_numpy(dp.x).float().cuda().requires_grad_(True)
    lg = torch.from_numpy(np.log(dp.growth)).float().cuda().requires_grad_(True)
    y = torch.from_numpy(dp.y).float().cuda()
    if args.warmup:
        xs = x[:500].detach().requires_grad_(True)
        ot.ot_objective(xs, lg[:500].detach(), y[:600]).backward()
    torch.cuda.synchronize()
    times = []
    for _ in range(max(args.steps, 1)):
        x.grad = None
        lg.grad = None
        t0 = time.perf_counter()
        val = ot.ot_objective(x, lg, y)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        val.backward()
        torch.cuda.synchronize()
        times.append((t1 - t0, time.perf_counter() - t1))
    fwd = float(np.mean([t[0] for t in times]))
    bwd = float(np.mean([t[1] for t in times]))
    out = {'metric': 'gradient evaluations/sec of the converged OT objective (forward solve + adjoint)',
           'value': 1.0 / (fwd + bwd), 'unit': 'grads/s', 'n_gpus': 1, 'steps': len(times),
           'warmup': int(bool(args.warmup)), 'ms_per_step': 1000.0 * (fwd + bwd), 'higher_is_better': True,
           'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32 in / f64 adjoint',
           'data': 'synthetic (seeded Gaussian mixture, 30-d)',
           'config': {'workload': 'd OT / d(x, log g), 8000 x 8000, d=30, duality_gap solver'},
           'forward_s': fwd, 'adjoint_s': bwd, 'objective': float(val)}
    out.update(_adjoint_checks(torch, ot, dp, x, lg, y))
    out['cost_split_bf16'] = cost_gemm_report(torch, dp.x, dp.y)
    return out


def _adjoint_checks(torch, ot, dp, x, lg, y):
    from wot_amd.ot import adjoint as adj
    from wot_amd.ot import device as dev
    from oracle import ot_oracle as O
    # FP64 autograd of the primal at the same converged plan (envelope theorem)
    X = x.detach().double()
    Y = y.double()
    G = torch.exp(lg.detach().double())
    cost = ot.DeviceCost.from_coordinates(X, Y)
    prm = dict(adj.DEFAULTS)
    res = dev.solve(cost, G, 'duality_gap', prm, want_plan=False, want_potentials=True)
    eps = res['stats']['epsilon_final']
    gx, glg = adj.adjoint(cost, X, Y, G, res['f'], res['g'], prm['lambda1'], prm['lambda2'], eps)
    xx = X.clone().requires_grad_(True)
    ll = lg.detach().double().clone().requires_grad_(True)
    D = (xx * xx).sum(1)[:, None] + (Y * Y).sum(1)[None, :] - 2.0 * xx @ Y.T
    Dd = dev.sqeuclid_f64(X, Y)
    D = D + (Dd - D).detach()
    mid = cost.median_index.cpu().numpy()
    flat = D.reshape(-1)
    m = 0.5 * (flat[int(mid[0])] + flat[int(mid[1])])
    C = D / m
    n_i, n_j = C.shape
    p = torch.exp(ll)
    q = p.mean()
    with torch.no_grad():
        R = torch.exp((res['f'][:, None] + res['g'][None, :] - C.detach()) / eps)
    K = torch.exp(-C / eps)
    xr, yc = R.sum(1) / n_j, R.sum(0) / n_i
    P = (prm['lambda1'] * (1.0 / n_i * (xr * torch.log(xr / p) - xr + p)).sum()
         + prm['lambda2'] * (1.0 / n_j * (yc * torch.log(yc / q) - yc + q)).sum()
         + (eps * (R * torch.log(R) - R + K).sum() + (R * C).sum()) / (n_i * n_j))
    P.backward()
    e_x = float(torch.linalg.norm(gx - xx.grad) / torch.linalg.norm(xx.grad))
    e_g = float(torch.linalg.norm(glg - ll.grad) / torch.linalg.norm(ll.grad))
    del D, Dd, C, R, K, P
    torch.cuda.empty_cache()
    # central finite differences of the FP64-converged objective (oracle) on a sub-problem
    sub_i, sub_j = 300, 360
    xs, ys, lgs = dp.x[:sub_i], dp.y[:sub_j], np.log(dp.growth[:sub_i])
    fd_prm = dict(lambda1=1.0, lambda2=50.0, epsilon=0.05, scaling_iter=2000, tau=1e4, epsilon0=0.05,
                  extra_iter=0, inner_iter_max=50)
    xt = torch.from_numpy(xs).cuda().requires_grad_(True)
    lt = torch.from_numpy(lgs).cuda().requires_grad_(True)
    ot.ot_objective(xt, lt, torch.from_numpy(ys).cuda(), solver='fixed_iters', **fd_prm).backward()
    gpu_gx = xt.grad.cpu().numpy()
    rng = np.random.default_rng(0)
    coords = [(int(rng.integers(sub_i)), int(rng.integers(xs.shape[1]))) for _ in range(64)]
    h = 1e-4
    fd, an = [], []
    for (i, k) in coords:
        a, b = xs.copy(), xs.copy()
        a[i, k] += h
        b[i, k] -= h
        fa = O.objective_and_grad(a, lgs, ys, **fd_prm)[0]
        fb = O.objective_and_grad(b, lgs, ys, **fd_prm)[0]
        fd.append((fa - fb) / (2 * h))
        an.append(gpu_gx[i, k])
    fd, an = np.array(fd), np.array(an)
    e_fd = float(np.linalg.norm(fd - an) / np.linalg.norm(fd))
    return {'adjoint_vs_fp64_autograd_rel': {'grad_x': e_x, 'grad_log_g': e_g},
            'adjoint_vs_central_fd_rel_64coords': e_fd,
            'fd_setup': '300 x 360 sub-problem, eps0 = eps = 0.05, 2000 iterations, h = 1e-4, FP64 oracle'}


def bench_large(args):
    """One 60,000 x 60,000 pair (50-d, 30 components) solved without a stored
    cost (cost + exp + LSE fused per tile on the matrix cores), 3-term and
    6-term split-BF16 cross terms, full fixed_iters schedule; then a 5-map
    long-range coupling chain (FP32 MFMA GEMMs) with its row-normalised error
    against the FP64 chain."""
    import torch
    import wot_amd.ot as ot
    import wot_amd.tmap as tm
    from wot_amd.ot import device as dev
    from wot_amd.ot import large
    from wot_amd.synthetic import day_pair, time_course
    n = args.large_n
    dp = day_pair(n, n, dim=50, n_components=30, seed=6)
    sched = dict(SOLVER)
    sched.pop('growth_iters')
    if args.iters is not None:
        sched['extra_iter'] = sched['extra_iter'] * args.iters // sched['scaling_iter']
        sched['scaling_iter'] = args.iters
    n_iter = sched['scaling_iter'] + sched['extra_iter']
    out = {'metric': 'transport maps/sec and Sinkhorn iters/sec at \u22641e-5 rel err, fused vs plain GPU',
           'unit': 'maps/s', 'n_gpus': 1, 'steps': 1, 'warmup': 0, 'higher_is_better': True, 'scaling': 'weak',
           'vs_baseline': None, 'dtype': 'bf16 split / f32 acc / f64 potentials',
           'data': 'synthetic (seeded 30-component Gaussian mixture, 50-d)',
           'config': {'workload': '%d x %d pair, cost never stored, fixed_iters' % (n, n), **sched}}
    X = torch.from_numpy(dp.x).cuda()
    Y = torch.from_numpy(dp.y).cuda()
    G = torch.from_numpy(dp.growth).cuda()
    res = {}
    for terms in (3, 6):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = large.transport_stablev2_coords(X, Y, G, terms=terms, **sched)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        res[terms] = r
        sweep = r.stats['sweep_ms'] / 1000.0 / (2 * n_iter)
        out['terms%d' % terms] = {'seconds': el, 'maps_per_s': 1.0 / el, 'sinkhorn_iters_per_s': n_iter / el,
                                  'avg_sweep_us': sweep * 1e6, 'entries_per_s': n * n / sweep,
                                  'mfma_tflops': 2.0 * terms * n * n * (((50 + 15) // 16) * 16) / sweep / 1e12}
        log('terms %d: %.2f s' % (terms, el))
    out['value'] = out['terms3']['maps_per_s']
    out['ms_per_step'] = 1000.0 * out['terms3']['seconds']
    out['sinkhorn_iters_per_s'] = out['terms3']['sinkhorn_iters_per_s']
    # OTModel's default solver on the same pair, cost never stored (6 terms)
    gap_prm = dict(lambda1=SOLVER['lambda1'], lambda2=SOLVER['lambda2'], epsilon=SOLVER['epsilon'], batch_size=5,
                   tolerance=1e-8, tau=1e4, epsilon0=SOLVER['epsilon0'], max_iter=1e7)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rg = large.optimal_transport_duality_gap_coords(X, Y, G, terms=6, **gap_prm)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    out['duality_gap_terms6'] = {'seconds': el, 'maps_per_s': 1.0 / el, 'iterations': rg.stats['iterations'],
                                 'gap_checks': rg.stats['gap_checks'], 'gap': rg.stats['gap'],
                                 'stabilisations': rg.stats['stabilisations']}
    log('duality gap (6 terms): %.2f s, %d iterations' % (el, rg.stats['iterations']))
    del rg
    # full-size cross-check on a short schedule against the stored-cost solver
    short = dict(sched, scaling_iter=100, extra_iter=0)
    cost = ot.DeviceCost.from_coordinates(X, Y)
    ref = dev.solve(cost, G, 'fixed_iters', short, want_plan=False, want_potentials=True)
    rows = torch.arange(0, n, max(n // 512, 1), device='cuda')
    D = dev.sqeuclid_f64(X[rows], Y)
    C = D / cost.median
    eps_s = ref['stats']['epsilon_final']
    Rref = torch.exp((ref['f'][rows][:, None] + ref['g'][None, :] - C) / eps_s)
    for terms in (3, 6):
        r = large.transport_stablev2_coords(X, Y, G, terms=terms, **short)
        R = torch.exp((r.f[rows][:, None] + r.g[None, :] - C) / eps_s)
        out['terms%d' % terms]['rel_frob_vs_stored_cost_100it_512rows'] = float(torch.linalg.norm(R - Rref) /
                                                                                torch.linalg.norm(Rref))
    del cost, D, C, R, Rref
    torch.cuda.empty_cache()
    out['cost_split_bf16_stored_kernel'] = cost_gemm_report(torch, X, Y)
    # 5-map chain: 6 days x 12,000 cells solved with the stored-cost solver
    coords, growth, counts = time_course(dim=50, n_components=30, seed=7, counts=[12000] * 6)
    maps = []
    for t in range(5):
        cst = ot.DeviceCost.from_coordinates(coords[t], coords[t + 1])
        maps.append(ot.transport_stablev2(cst, G=growth[t], device_out=True, **sched))
        del cst
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    chain32 = tm.chain_product([m.float() for m in maps])
    torch.cuda.synchronize()
    t_chain = time.perf_counter() - t0
    ref64 = maps[0]
    for m in maps[1:]:
        ref64 = ref64 @ m
    rn = chain32.double() / chain32.double().sum(1, keepdim=True)
    rr = ref64 / ref64.sum(1, keepdim=True)
    out['chain'] = {'maps': 5, 'cells_per_day': 12000, 'seconds': t_chain,
                    'tflops': 4 * 2.0 * 12000 ** 3 / t_chain / 1e12,
                    'rel_frob_vs_fp64': float(torch.linalg.norm(chain32.double() - ref64) / torch.linalg.norm(ref64)),
                    'row_normalised_max_abs_err': float((rn - rr).abs().max()),
                    'row_normalised_rel_frob': float(torch.linalg.norm(rn - rr) / torch.linalg.norm(rr))}
    return out


TMAP_POPS = 10


def bench_tmap(args):
    """Trajectories (TransportMapModel.trajectories, transport_map_model.py:105-142)
    of 10 populations at the middle day of the 39-day time course (cells per
    day ~ U[1000, 20000], the batched time course's days): pull back through
    19 maps and push forward through 19, every map read once per walk, every
    day's populations written to one (cells x 10) matrix in HBM.  Maps are
    synthetic FP64 (or --tmap-dtype float32) on the device.
    value = trajectory walks per second; roofline per product kernel from the
    library's hipEvent timer: algorithmic bytes = map bytes + 8 k (I + J)."""
    import torch
    import wot_amd.tmap as tm
    from wot_amd.synthetic import time_course
    _, _, counts = time_course(**TC)
    counts = [int(c) for c in counts]
    n_days = len(counts)
    mid = n_days // 2
    tdt = torch.float64 if args.tmap_dtype == 'float64' else torch.float32
    esize = 8 if tdt == torch.float64 else 4
    gen = torch.Generator(device='cuda').manual_seed(7)
    maps = []
    for t in range(n_days - 1):
        m = tm.DeviceMap.empty(counts[t], counts[t + 1], args.tmap_dtype)
        m.X.uniform_(generator=gen)
        maps.append(m)
    k = TMAP_POPS
    P = (torch.rand((counts[mid], k), dtype=torch.float64, device='cuda', generator=gen) > 0.8).double()
    P = P / P.sum(0, keepdim=True)
    offs = np.cumsum([0] + counts)
    X = torch.empty((int(offs[-1]), k), dtype=torch.float64, device='cuda')
    torch.cuda.synchronize()
    t_tr = time.perf_counter()
    for t in range(mid):  # the pull walk streams transposed copies (TransportMapModel default)
        maps[t].make_transposed()
    torch.cuda.synchronize()
    t_tr = time.perf_counter() - t_tr
    back = [maps[t - 1] for t in range(mid, 0, -1)]
    fwd = [maps[t] for t in range(mid, n_days - 1)]
    out_back = [X[offs[t - 1]:offs[t]] for t in range(mid, 0, -1)]
    out_fwd = [X[offs[t + 1]:offs[t + 2]] for t in range(mid, n_days - 1)]
    map_bytes = {'pull': float(sum(esize * m.shape[0] * m.shape[1] for m in back)),
                 'push': float(sum(esize * m.shape[0] * m.shape[1] for m in fwd))}
    alg = {'pull': map_bytes['pull'] + 8.0 * k * sum(m.shape[0] + m.shape[1] for m in back),
           'push': map_bytes['push'] + 8.0 * k * sum(m.shape[0] + m.shape[1] for m in fwd)}

    def walk(stats=None):
        sb, sf = ({}, {}) if stats is not None else (None, None)
        X[offs[mid]:offs[mid + 1]].copy_(P)
        tm.propagate(back, P, tm.PULL_BACK, normalize=True, outs=out_back, stats=sb)
        tm.propagate(fwd, P, tm.PUSH_FORWARD, normalize=True, outs=out_fwd, stats=sf)
        if stats is not None:
            stats['pull'].append(sb['kernel_ms'])
            stats['push'].append(sf['kernel_ms'])

    for _ in range(max(args.warmup, 1)):
        walk()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        walk()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    per = elapsed / args.steps
    kt = {'pull': [], 'push': []}
    for _ in range(args.steps):
        walk(kt)
    kern = {}
    for d, name in (('pull', 'push_partial_kernel (transposed copies)'), ('push', 'push_partial_kernel')):
        ms = float(np.mean(kt[d]))
        n = len(back) if d == 'pull' else len(fwd)
        kern[d] = {'kernel': name, 'launches': n, 'avg_launch_us': 1000.0 * ms / n,
                   'achieved': alg[d] / (ms / 1000.0) / 1e9, 'alg_bytes_per_launch': alg[d] / n}
    dom = min(kern, key=lambda d: kern[d]['achieved'])
    roof = dict(bound='hbm', kernel=kern[dom]['kernel'], achieved=kern[dom]['achieved'], peak=HBM_PEAK_GBS,
                unit='GB/s', frac=kern[dom]['achieved'] / HBM_PEAK_GBS,
                traffic=_traffic('push_partial_kernel_timecourse_trajectories') if esize == 8 else None,
                avg_launch_us=kern[dom]['avg_launch_us'])
    total_bytes = alg['pull'] + alg['push']
    out = {'metric': 'trajectory walks/sec (10 populations, middle day, 38 maps, every map streamed once)',
           'value': 1.0 / per, 'unit': 'walks/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': max(args.warmup, 1),
           'ms_per_step': 1000.0 * per, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
           'dtype': 'f64' if esize == 8 else 'f32 maps / f64 accumulate',
           'data': 'synthetic (uniform random maps on the time course\'s day sizes)',
           'config': {'workload': 'TransportMapModel.trajectories, 39 days, %d populations' % k,
                      'cells_total': int(offs[-1]), 'map_bytes_total': map_bytes['pull'] + map_bytes['push'],
                      'map_dtype': args.tmap_dtype, 'pull_maps': 'transposed copies in HBM',
                      'transpose_once_ms': 1000.0 * t_tr},
           'end_to_end_GBps': total_bytes / per / 1e9,
           'roofline': roof, 'roofline_kernels': kern}
    # parity on the run's own data: the adjoint identity <pull(q), p> = <q, push(p)> for one map
    m = maps[mid]
    pp = torch.rand((m.shape[0], 3), dtype=torch.float64, device='cuda', generator=gen)
    qq = torch.rand((m.shape[1], 3), dtype=torch.float64, device='cuda', generator=gen)
    push = tm.propagate([m], pp, tm.PUSH_FORWARD, normalize=False)
    pull = tm.propagate([m], qq, tm.PULL_BACK, normalize=False)
    lhs, rhs = (pull * pp).sum(0), (qq * push).sum(0)
    out['adjoint_identity_rel'] = float(((lhs - rhs).abs() / rhs.abs()).max())
    if not args.no_baselines:
        out['cpu_baseline'] = _tmap_cpu_baseline(counts, k, total_bytes, per)
    return out


def _tmap_cpu_baseline(counts, k, total_bytes, gpu_per):
    """The numpy oracle's push / pull (one dgemm per map, BLAS threads) on the
    first maps of the same course, FP64, scaled to the full walk by bytes."""
    from oracle import tmap_oracle as O
    try:
        from threadpoolctl import threadpool_info
        cores = max([i.get('num_threads', 1) for i in threadpool_info() if i.get('user_api') == 'blas'] or [1])
    except Exception:
        cores = os.cpu_count()
    rng = np.random.default_rng(0)
    maps = [rng.random((counts[t], counts[t + 1])) for t in range(3)]
    P = rng.random((k, counts[0]))
    Q = rng.random((k, counts[3]))
    nbytes = float(sum(8 * m.size + 8 * k * (m.shape[0] + m.shape[1]) for m in maps))
    t0 = time.perf_counter()
    reps = 0
    while time.perf_counter() - t0 < 10.0:
        O.push_forward(maps, P, 0, 3)
        O.pull_back(maps, Q, 3, 0)
        reps += 1
    dt = (time.perf_counter() - t0) / reps
    rate = 2 * nbytes / dt
    return {'value': rate / total_bytes, 'unit': 'walks/s', 'cores': int(cores), 'kind': 'port',
            'sample': 'push + pull of 10 populations through the first 3 maps (%s cells), %d reps, '
                      'scaled by bytes to the 38-map walk' % ('-'.join(str(c) for c in counts[:4]), reps),
            'GBps': rate / 1e9, 'gpu_over_cpu': (1.0 / gpu_per) / (rate / total_bytes)}


GAP_SOLVER = dict(epsilon=0.05, lambda1=1.0, lambda2=50.0, epsilon0=1.0, tau=1e4, tolerance=1e-8, batch_size=5,
                  max_iter=1e7)


def cpu_gap_sample(x, y, growth, counts):
    """FP64 numpy oracle of the duality-gap solver on the host cores, on a
    bounded sample of the same map: cost + median at full size; one
    exp(-C/eps) pass; two early-returning solves (max_iter 10 and 30) whose
    difference is 20 iterations; one plan build + primal + dual.  Extrapolated
    with the device run's counts: 2 kernel builds per epsilon scaling plus one
    per stabilisation, its iterations, and one primal / dual check per batch
    of the last scaling (the earlier scalings' checks are O(I + J))."""
    from oracle import ot_oracle as O
    try:
        from threadpoolctl import threadpool_info
        threads = max([i.get('num_threads', 1) for i in threadpool_info()] + [1])
    except Exception:  # pragma: no cover
        threads = os.cpu_count()
    t0 = time.perf_counter()
    C = O.default_cost_matrix(x, y)
    t_cost = time.perf_counter() - t0
    eps = GAP_SOLVER['epsilon']
    np.exp(-C[:2] / eps)  # first-call overhead out of the timed pass
    t0 = time.perf_counter()
    Kt = np.exp(-C / eps)
    t_exp = time.perf_counter() - t0
    with np.errstate(all='ignore'):
        t0 = time.perf_counter()
        O.optimal_transport_duality_gap(C, growth, **dict(GAP_SOLVER, max_iter=10))
        t_a = time.perf_counter() - t0
        t0 = time.perf_counter()
        O.optimal_transport_duality_gap(C, growth, **dict(GAP_SOLVER, max_iter=30))
        t_b = time.perf_counter() - t0
        n_i, n_j = C.shape
        dx, dy = np.ones(n_i) / n_i, np.ones(n_j) / n_j
        p, q = growth, np.ones(n_j) * np.average(growth)
        a, b = np.ones(n_i), np.ones(n_j)
        t0 = time.perf_counter()
        R = (Kt.T * a).T * b
        O.primal(C, Kt, R, dx, dy, p, q, a, b, eps, GAP_SOLVER['lambda1'], GAP_SOLVER['lambda2'])
        O.dual(C, Kt, R, dx, dy, p, q, a, b, eps, GAP_SOLVER['lambda1'], GAP_SOLVER['lambda2'])
        t_check = time.perf_counter() - t0
    del C, Kt, R
    t_it = max((t_b - t_a) / 20.0, 0.0)
    n_build = 2 * 6 + counts['stabilisations']
    per_map = t_cost + n_build * t_exp + counts['iterations'] * t_it + counts['final_checks'] * t_check
    return per_map, threads, dict(t_cost_s=t_cost, t_exp_s=t_exp, t_iter_s=t_it, t_check_s=t_check,
                                  sample_wall_s=t_cost + t_exp + t_a + t_b + t_check)


def bench_gap(args):
    """One 15k x 15k map with OTModel's default solver (duality gap,
    tolerance 1e-8, growth_iters=1): cost + median + solve + FP64 plan, inputs
    resident in HBM; plain-GPU restatement and CPU oracle beside it."""
    import torch
    import wot_amd.ot as ot  # noqa: F401  (loads the library)
    from wot_amd.ot import device as dev
    from wot_amd.synthetic import day_pair
    dp = day_pair(CFG['n_source'], CFG['n_target'], dim=CFG['dim'], n_components=CFG['n_components'],
                  seed=CFG['seed'])
    x = torch.from_numpy(dp.x).cuda()
    y = torch.from_numpy(dp.y).cuda()
    g = torch.from_numpy(dp.growth).cuda()
    n_i, n_j = CFG['n_source'], CFG['n_target']

    def one():
        cost = dev.DeviceCost.from_coordinates(x, y, mode=args.cost_mode)
        return cost, dev.solve(cost, g, 'duality_gap', GAP_SOLVER, want_plan=True)

    for w in range(args.warmup):
        log('warmup %d' % w)
        one()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        cost, res = one()
        log('step %d done' % s)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    st = res['stats']
    # the early scalings do not depend on the tolerance: the same solve stopped
    # after one batch of the last scaling gives the last scaling's share
    early = dev.solve(cost, g, 'duality_gap', dict(GAP_SOLVER, tolerance=1e300), want_plan=False)['stats']
    final_iters = st['iterations'] - (early['iterations'] - GAP_SOLVER['batch_size'])
    counts = dict(iterations=st['iterations'], stabilisations=st['stabilisations'], gap_checks=st['gap_checks'],
                  final_iterations=final_iters, final_checks=final_iters // GAP_SOLVER['batch_size'],
                  gap=st['gap'])
    maps_per_s = args.steps / elapsed
    sweep_bytes = 2.0 * st['iterations'] * 4.0 * n_i * n_j
    out = {
        'metric': 'transport maps/sec, duality-gap solver (OTModel default)', 'value': maps_per_s,
        'unit': 'maps/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': 1000.0 * elapsed / args.steps, 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f64/f32', 'data': 'synthetic (seeded Gaussian mixture in 30-d PCA space)',
        'config': {'workload': 'one day pair 15000x15000, d=30, 10 components, duality_gap, growth_iters=1',
                   **GAP_SOLVER, 'cost_mode': args.cost_mode},
        'solver_counts': counts,
        'sinkhorn_iters_per_s': maps_per_s * st['iterations'],
        'sweep_bytes_per_s_end_to_end': sweep_bytes * maps_per_s / 1e9,
    }
    if not args.no_baselines:
        from wot_amd import plain_gpu
        log('plain GPU sample')
        plan = res['plan']
        del res
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        c64 = dev.sqeuclid_f64(x, y)
        v = torch.sort(c64.view(-1)).values
        med = (v[v.numel() // 2 - 1] + v[v.numel() // 2]) / 2
        C = c64 / med
        del v, c64
        torch.cuda.synchronize()
        t_cost = time.perf_counter() - t0
        pst = {}
        t0 = time.perf_counter()
        plain = plain_gpu.optimal_transport_duality_gap(C, g, stats=pst, **GAP_SOLVER)
        torch.cuda.synchronize()
        t_solve = time.perf_counter() - t0
        rel = float(torch.linalg.norm(plan - plain) / torch.linalg.norm(plain))
        del plain, C, plan
        torch.cuda.empty_cache()
        out['plain_gpu'] = {'value': 1.0 / (t_cost + t_solve), 'unit': 'maps/s', 't_cost_s': t_cost,
                            't_solve_s': t_solve, 'counts': pst,
                            'sample': 'torch FP64 dense-K restatement of the duality-gap solver, one full map'}
        out['speedup_vs_plain_gpu'] = maps_per_s * (t_cost + t_solve)
        out['parity_vs_plain_gpu_rel_frob'] = rel
        log('cpu sample')
        per_map_cpu, threads, detail = cpu_gap_sample(dp.x, dp.y, dp.growth, counts)
        out['cpu_baseline'] = {'value': 1.0 / per_map_cpu, 'unit': 'maps/s', 'cores': threads, 'kind': 'port',
                               'sample': 'FP64 numpy oracle at full size: cost+median, one kernel build, 20 '
                                         'iterations (difference of two early-returning solves), one primal/dual '
                                         'check; extrapolated with the device run\'s counts', **detail}
        out['speedup_vs_cpu'] = maps_per_s * per_map_cpu
    return out


PCA_CFG = dict(n_cells=(15000, 15000), genes=2000, rank=40, k=30, seed=11)


def bench_pca(args):
    """Local PCA of one day pair (OTModel.compute_single_transport_map with
    local_pca=30, ot_model.py:296-301): 15,000 + 15,000 cells x 2,000 genes
    (FP32, low-rank signal + noise), k = 30.  value = PCA fits per second
    with the expression matrix resident on the device; roofline on the
    double-centering kernel (2 reads of x + 1 FP64 write of A).  Baseline:
    the reference's own call (sklearn PCA 'auto', random_state 58951) on the
    host cores."""
    import torch
    from wot_amd.ot import pca as gpca
    from oracle import pca_oracle as P
    n1, n2 = PCA_CFG['n_cells']
    g, r, k = PCA_CFG['genes'], PCA_CFG['rank'], PCA_CFG['k']
    rng = np.random.default_rng(PCA_CFG['seed'])
    W = rng.standard_normal((r, g)).astype(np.float32)
    a = (rng.standard_normal((n1, r)).astype(np.float32) @ W + 0.3 * rng.standard_normal((n1, g)).astype(np.float32))
    b = (rng.standard_normal((n2, r)).astype(np.float32) @ W + 0.3 * rng.standard_normal((n2, g)).astype(np.float32)
         + 0.2)
    x = torch.from_numpy(np.vstack([a, b])).cuda()
    for _ in range(max(args.warmup, 1)):
        gpca.fit(x, k)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = gpca.fit(x, k)
    torch.cuda.synchronize()
    per = (time.perf_counter() - t0) / args.steps
    # centering kernel alone, timed with events on the current stream
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    reps = 10
    for _ in range(reps):
        gpca.center(x)
    ev[1].record()
    torch.cuda.synchronize()
    c_us = ev[0].elapsed_time(ev[1]) * 1000.0 / reps
    n = n1 + n2
    alg = 2.0 * 4 * n * g + 8.0 * n * g + 8.0 * (n + g)
    out = {'metric': 'local PCA fits/sec (one day pair, k=30)', 'value': 1.0 / per, 'unit': 'fits/s', 'n_gpus': 1,
           'steps': args.steps, 'warmup': max(args.warmup, 1), 'ms_per_step': 1000.0 * per, 'higher_is_better': True,
           'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32 in / f64', 'data': 'synthetic (rank-40 + noise)',
           'config': {'workload': 'compute_pca, %d + %d cells x %d genes, k=%d, exact Gram method' % (n1, n2, g, k)},
           'roofline': {'bound': 'hbm', 'kernel': 'center_sums_kernel + center_apply_kernel', 'avg_launch_us': c_us,
                        'achieved': alg / (c_us * 1e-6) / 1e9, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                        'frac': alg / (c_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 'traffic': None}}
    if not args.no_baselines:
        import os as _os
        t0 = time.perf_counter()
        r1, r2, rp, _ = P.reference_pca(a, b, k)
        cpu = time.perf_counter() - t0
        Z = res.coordinates().cpu().numpy()
        R = np.vstack([r1, r2]) * rp.singular_values_
        sel = rng.choice(n, 2000, replace=False)
        d_gpu, d_ref = P.sqdist(Z[sel]), P.sqdist(R[sel].astype(np.float64))
        out['parity_vs_reference_sklearn'] = {
            'singular_values_rel': float(np.linalg.norm(res.singular_values_ - rp.singular_values_)
                                         / np.linalg.norm(rp.singular_values_)),
            'sqdist_rel_2000_cells': float(np.linalg.norm(d_gpu - d_ref) / np.linalg.norm(d_ref))}
        out['cpu_baseline'] = {'value': 1.0 / cpu, 'unit': 'fits/s', 'cores': _os.cpu_count(), 'kind': 'reference',
                               'sample': "the reference's call: sklearn PCA(30, random_state=58951).fit(x.T) on the "
                                         'same FP32 matrix, one fit'}
    return out


VAL_CFG = dict(cells=3000, genes=500, days=3, interp_size=3000, local_pca=30, seed=13)


def bench_validation(args):
    """compute_validation_summary (optimal_transport_validation.py:15-196) on
    one day triplet: 3 x 3,000 cells x 500 genes (drifting low-rank signal,
    growth rates, one covariate), local_pca = 30, interp_size = 3,000: local
    PCA + middle-day projection, the (t0, t1) map, three interpolated
    populations (I, R, Rg) and five exact EMDs (I, R, Rg, F, L).  value =
    summaries per second.  The EMDs (network simplex) run on the host; the
    device does PCA, maps, sampling and costs.  The roofline covers the
    sampling-weight kernels (T read, W written).  The EMDs run on a pool of
    host threads; the same summary with every EMD in line is timed beside."""
    import torch
    import pandas as pd
    import wot_amd.ot as ot
    from wot_amd import _anndata
    from wot_amd.ot import validation, emd as emd_mod
    from oracle import validation_oracle as V
    c = VAL_CFG
    rng = np.random.default_rng(c['seed'])
    W = rng.standard_normal((8, c['genes']))
    blocks = [rng.standard_normal((c['cells'], 8)) @ W + 0.3 * t + 0.2 * rng.standard_normal((c['cells'], c['genes']))
              for t in range(c['days'])]
    X = np.vstack(blocks)
    days = np.repeat(np.arange(c['days'], dtype=np.float64), c['cells'])
    obs = pd.DataFrame({'day': days, 'cell_growth_rate': np.exp(0.2 * rng.standard_normal(len(days)))},
                       index=['c%d' % i for i in range(len(days))])
    adata = _anndata.AnnData(X, obs, pd.DataFrame(index=['g%d' % k for k in range(c['genes'])]))
    model = ot.OTModel(adata, local_pca=c['local_pca'])
    np.random.seed(0)
    validation.compute_validation_summary(model, interp_size=200)  # warm-up (small)
    torch.cuda.synchronize()
    t_emd = [0.0]
    real_emd = emd_mod.emd2

    def timed_emd(*a, **k):
        t = time.perf_counter()
        r = real_emd(*a, **k)
        t_emd[0] += time.perf_counter() - t
        return r
    emd_mod.emd2 = timed_emd
    try:
        np.random.seed(1)
        t0 = time.perf_counter()
        df = validation.compute_validation_summary(model, interp_size=c['interp_size'])
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
    finally:
        emd_mod.emd2 = real_emd
    # the same summary with every EMD in line (one host thread), for comparison
    np.random.seed(1)
    t0 = time.perf_counter()
    validation.compute_validation_summary(model, interp_size=c['interp_size'], emd_workers=1)
    torch.cuda.synchronize()
    in_line = time.perf_counter() - t0
    # sampling-weight kernels on one 3000 x 3000 map, hipEvents on the current stream
    tmap = model.compute_transport_map(0.0, 2.0, covariate=(1, 1)).X
    Td = torch.from_numpy(tmap).cuda()
    Wd = torch.empty_like(Td)
    cs = torch.empty(Td.shape[1], dtype=torch.float64, device='cuda')
    from wot_amd import _native
    from wot_amd.ot import device as dev
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(10):
        _native.check(_native.lib().wotk_interp_weights(dev._ptr(Td), Td.stride(0), Td.shape[0], Td.shape[1], 0.5,
                                                        dev._ptr(Wd), dev._ptr(cs), dev._stream(torch)))
    ev[1].record()
    torch.cuda.synchronize()
    w_us = ev[0].elapsed_time(ev[1]) * 100.0
    alg = 3.0 * 8 * Td.numel()  # T read twice (column sums, weights), W written
    out = {'metric': 'validation summaries/sec (one day triplet, 5 exact EMDs)', 'value': 1.0 / elapsed,
           'unit': 'summaries/s', 'n_gpus': 1, 'steps': 1, 'warmup': 1, 'ms_per_step': 1000.0 * elapsed,
           'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64',
           'data': 'synthetic (drifting rank-8 signal, 3 days)',
           'config': {'workload': 'compute_validation_summary, 3 x %d cells x %d genes, local_pca=%d, interp_size=%d'
                                  % (c['cells'], c['genes'], c['local_pca'], c['interp_size']),
                      'rows': list(df['name'])},
           'emd_workers': validation._host_threads(),
           'emd_host_s_summed_over_threads': t_emd[0], 'summary_s_emds_in_line': in_line,
           'roofline': {'bound': 'hbm', 'kernel': 'colsum_partial_kernel + colsum_final_kernel + weights_kernel',
                        'avg_launch_us': w_us,
                        'achieved': alg / (w_us * 1e-6) / 1e9, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                        'frac': alg / (w_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 'traffic': None}}
    if not args.no_baselines:
        p0 = rng.standard_normal((3000, 30))
        p1 = rng.standard_normal((3000, 30))
        t = time.perf_counter()
        np.random.seed(2)
        V.interpolate_with_ot(p0, p1, tmap, 0.5, c['interp_size'])
        cpu = time.perf_counter() - t
        t = time.perf_counter()
        np.random.seed(2)
        validation.interpolate_with_ot(p0, p1, tmap, 0.5, c['interp_size'])
        gpu = time.perf_counter() - t
        out['cpu_baseline'] = {'value': None, 'unit': 'summaries/s', 'cores': 1, 'kind': 'port',
                               'sample': 'interpolate_with_ot on one 3000 x 3000 map, %d samples: numpy %.3f s, '
                                         'device %.3f s (the EMDs run on the host in both)' % (c['interp_size'], cpu,
                                                                                                 gpu),
                               'interp_cpu_s': cpu, 'interp_gpu_s': gpu}
    return out


def bench_sweeps(args):
    """Short, fixed workload for counter and trace runs of the stored-cost
    solver: the 15,000 x 15,000 pair's cost (``--cost-reps`` extra builds in
    each cost mode first), ``--iters`` fixed-iteration Sinkhorn iterations (2
    row log-sum-exp sweeps each, one over C and one over C^T, 4*n*n algorithmic
    bytes per sweep) with the given ``--tile``, then one plain device copy of
    n*n float32 (n*n*4 bytes read and written) that calibrates the FETCH_SIZE
    and WRITE_SIZE counters."""
    import torch
    import wot_amd.ot as ot
    from wot_amd.synthetic import day_pair
    from wot_amd import tune
    n = CFG['n_source']
    iters = args.iters if args.iters is not None else 20
    dp = day_pair(n, n, dim=CFG['dim'], n_components=CFG['n_components'], seed=CFG['seed'])
    if args.tile:
        tune.apply_persisted()
        tune.set_tile(tune.ROW_SWEEP, -1, _parse_tile(args.tile))
    for _ in range(args.cost_reps):
        for mode in ('f64', 'bf16x3', 'bf16x6'):
            ot.DeviceCost.from_coordinates(dp.x, dp.y, mode=mode)
    cost = ot.DeviceCost.from_coordinates(dp.x, dp.y)
    st = {}
    ot.transport_stablev2(cost, G=dp.growth, lambda1=1.0, lambda2=50.0, epsilon=0.05, scaling_iter=iters, tau=1e4,
                          epsilon0=1.0, extra_iter=0, inner_iter_max=50, device_out=True, stats=st)
    src = torch.empty((n, n), dtype=torch.float32, device='cuda').fill_(1.0)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    return {'workload': 'row sweeps of one %d x %d pair' % (n, n), 'iterations': iters, 'tile': args.tile,
            'avg_sweep_us': 1000.0 * st['sweep_ms'] / max(st['sweeps'], 1), 'copy_bytes': src.numel() * 4}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--no-baselines', action='store_true', help='skip the plain-GPU and CPU samples')
    ap.add_argument('--large-n', type=int, default=60000, help='large only: cells per day')
    ap.add_argument('--group-size', type=int, default=None,
                    help='timecourse only: combinations per batched run (default: the whole grid)')
    ap.add_argument('--config', default='timecourse', choices=['pair', 'gap', 'timecourse', 'adjoint', 'large', 'tmap', 'pca',
                                                               'validation', 'sweeps'],
                    help='timecourse: 38 pairs x 9-point grid (default); pair: one 15k x 15k map; '
                         'gap: the same pair with the duality-gap solver; '
                         'adjoint: gradient of the converged objective, 8k x 8k; '
                         'tmap: trajectories through the 38 maps of the time course; '
                         'pca: local PCA of one 15k + 15k cell day pair; '
                         'validation: compute_validation_summary on one day triplet; '
                         'sweeps: a short row-sweep workload for counter runs')
    ap.add_argument('--tmap-dtype', default='float64', choices=['float64', 'float32'],
                    help='tmap only: precision of the stored maps')
    ap.add_argument('--cost-mode', default='bf16x6', choices=['f64', 'bf16x3', 'bf16x6'],
                    help='pair / gap: cost kernel of the timed maps')
    ap.add_argument('--no-tune', action='store_true',
                    help='skip the tile autotuner (use persisted or default tiles)')
    ap.add_argument('--iters', type=int, default=None,
                    help='timecourse / large: override scaling_iter (extra_iter scaled alike) for quick probes')
    ap.add_argument('--tile', default=None,
                    help='timecourse: force the shared-exponential sweep tile ROWSxTHREADS (no tuning)')
    ap.add_argument('--no-parity', action='store_true',
                    help='timecourse: skip the parity and cost-error legs (profiling runs)')
    ap.add_argument('--inner-iter-max', type=int, default=None,
                    help='timecourse: override inner_iter_max (with --iters, a short schedule whose eps converges: '
                         'counter runs of every sweep mode)')
    ap.add_argument('--share-eps', default='on', choices=['on', 'off'],
                    help='timecourse: eps groups in integer k ratios take powers of the largest eps\'s '
                         'exponential once the schedule has converged')
    ap.add_argument('--cost-reps', type=int, default=0,
                    help='sweeps: extra cost builds per cost mode before the sweeps (kernel traces)')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU: libwotk has no multi-GPU path (got --gpus %d)' % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.config == 'gap':
        print(json.dumps(bench_gap(args)), flush=True)
        return
    if args.config == 'timecourse':
        print(json.dumps(bench_timecourse(args)), flush=True)
        return
    if args.config == 'adjoint':
        print(json.dumps(bench_adjoint(args)), flush=True)
        return
    if args.config == 'large':
        print(json.dumps(bench_large(args)), flush=True)
        return
    if args.config == 'tmap':
        print(json.dumps(bench_tmap(args)), flush=True)
        return
    if args.config == 'pca':
        print(json.dumps(bench_pca(args)), flush=True)
        return
    if args.config == 'validation':
        print(json.dumps(bench_validation(args)), flush=True)
        return
    if args.config == 'sweeps':
        print(json.dumps(bench_sweeps(args)), flush=True)
        return

    import torch
    import wot_amd.ot as ot
    from wot_amd.ot import device as dev
    from wot_amd.synthetic import day_pair

    dp = day_pair(CFG['n_source'], CFG['n_target'], dim=CFG['dim'], n_components=CFG['n_components'],
                  seed=CFG['seed'])
    x = torch.from_numpy(dp.x).cuda()
    y = torch.from_numpy(dp.y).cuda()
    g = torch.from_numpy(dp.growth).cuda()
    torch.cuda.synchronize()
    tiles = _tune_pair(args, dev, x, y, g)

    for w in range(args.warmup):
        log('warmup %d' % w)
        one_map(ot, dev, x, y, g, args.cost_mode)
    torch.cuda.synchronize()
    sweep_ms = 0.0
    sweeps = 0
    t0 = time.perf_counter()
    res = cost = None
    for s in range(args.steps):
        cost, res, sms, nsw = one_map(ot, dev, x, y, g, args.cost_mode)
        sweep_ms += sms
        sweeps += nsw
        log('step %d done' % s)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1000.0 * elapsed / args.steps
    maps_per_s = args.steps / elapsed
    iters_per_map = SOLVER['growth_iters'] * (SOLVER['scaling_iter'] + SOLVER['extra_iter'])

    # roofline of the dominant kernel (row LSE sweep): algorithmic bytes per
    # launch = 4 * I * J (one FP32 cost row stream), launches alternate C / C^T
    n_i, n_j = CFG['n_source'], CFG['n_target']
    avg_sweep_s = (sweep_ms / 1000.0) / max(sweeps, 1)
    alg_bytes = 4.0 * n_i * n_j
    achieved = alg_bytes / avg_sweep_s / 1e9

    # size-independent property at full size: column sums of the un-normalised
    # plan equal q * I * exp(-g / lambda2) after the last (b-update) half iteration
    q = float(res['G_last'].mean())
    want = q * n_i * torch.exp(-res['g'] / SOLVER['lambda2'])
    colsum = res['plan'].sum(0) * n_j
    marg_err = float((colsum - want).abs().max() / want.abs().max())

    out = {
        'metric': 'transport maps/sec and Sinkhorn iters/sec at \u22641e-5 rel err, fused vs plain GPU',
        'value': maps_per_s,
        'unit': 'maps/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': ms_per_step,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64/f32',
        'data': 'synthetic (seeded Gaussian mixture in 30-d PCA space)',
        'config': {'workload': 'one day pair 15000x15000, d=30, 10 components, fixed_iters, growth_iters=3',
                   'n_source': n_i, 'n_target': n_j, 'dim': CFG['dim'], **SOLVER, 'cost_mode': args.cost_mode,
                   'tiles': tiles},
        'sinkhorn_iters_per_s': maps_per_s * iters_per_map,
        'marginal_identity_rel_err': marg_err,
        'roofline': {'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': achieved / HBM_PEAK_GBS, 'traffic': _traffic(),
                     'kernel': 'lse_rows_kernel', 'avg_launch_us': avg_sweep_s * 1e6,
                     'alg_bytes_per_launch': alg_bytes},
    }
    del colsum, want, res
    if not args.no_baselines:
        log('plain GPU sample')
        per_map_plain, plain_detail, plan_plain = plain_gpu_sample(torch, dp.x, dp.y, dp.growth)
        # parity at full size: one fused solve with the full default schedule
        # against the FP64 plain-GPU restatement of the reference
        fused = ot.transport_stablev2(cost, G=dp.growth, device_out=True, **SOLVER)
        rel = float(torch.linalg.norm(fused - plan_plain) / torch.linalg.norm(plan_plain))
        del plan_plain, fused
        out['plain_gpu'] = {'value': 1.0 / per_map_plain, 'unit': 'maps/s',
                            'sample': 'torch FP64 dense-K restatement: full-size cost+median and one full '
                                      'fixed_iters solve timed, x3 growth iterations',
                            **plain_detail}
        out['speedup_vs_plain_gpu'] = maps_per_s * per_map_plain
        out['parity_vs_plain_gpu_rel_frob'] = rel
        torch.cuda.empty_cache()
        log('cpu sample')
        per_map_cpu, threads, cpu_detail = cpu_sample(dp.x, dp.y, dp.growth)
        out['cpu_baseline'] = {'value': 1.0 / per_map_cpu, 'unit': 'maps/s', 'cores': threads, 'kind': 'port',
                               'sample': 'FP64 numpy oracle at full size: cost+median, then 50 scheduled '
                                         'iterations (2 kernel builds) and 50 extra iterations timed apart, '
                                         'extrapolated to 3 x (4000 iterations + 61 builds)',
                               **cpu_detail}
        out['speedup_vs_cpu'] = maps_per_s * per_map_cpu
    log('split-BF16 cost report')
    solve = {k: v for k, v in SOLVER.items() if k != 'growth_iters'}
    out['cost_split_bf16'] = cost_gemm_report(torch, x, y, G=g, solve=solve)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
