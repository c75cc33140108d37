"""Benchmark: transport maps/sec (and Sinkhorn iterations/sec) of the fused HIP
solver on one MI355X, versus the plain-GPU torch FP64 restatement and the
FP64 CPU oracle.

Workload (one "step" = one transport map, as OTModel.compute_transport_map
computes it): one day pair of 15,000 x 15,000 cells, 30-d PCA coordinates from
a seeded 10-component Gaussian mixture (target means shifted by N(0, 0.5)),
log-normal growth; cost = squared Euclidean / median (FP64 on the device);
fixed_iters solver (eps=0.05, lambda1=1, lambda2=50, eps0=1, tau=1e4,
scaling_iter=3000, inner_iter_max=50, extra_iter=1000) with growth_iters=3,
i.e. 12,000 Sinkhorn iterations per map; the FP64 plan is materialised in HBM.
Inputs are resident in HBM before the timed region.

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


def _traffic():
    """HBM bytes per sweep launch from the calibrated PMC runs kept in profiles/."""
    path = os.path.join(ROOT, 'profiles', 'lse_rows_kernel_15000x15000.traffic.json')
    try:
        with open(path) as fh:
            return json.load(fh)['traffic_bytes']
    except (OSError, ValueError, KeyError):
        return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def one_map(ot, dev, x, y, growth):
    """cost + median on the device, growth loop, final FP64 plan in HBM."""
    cost = dev.DeviceCost.from_coordinates(x, y)
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--no-baselines', action='store_true', help='skip the plain-GPU and CPU samples')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU: this build has no multi-GPU path (got --gpus %d)' % args.gpus,
              file=sys.stderr)
        sys.exit(2)

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

    for w in range(args.warmup):
        log('warmup %d' % w)
        one_map(ot, dev, x, y, g)
    torch.cuda.synchronize()
    sweep_ms = 0.0
    sweeps = 0
    t0 = time.perf_counter()
    res = cost = None
    for s in range(args.steps):
        cost, res, sms, nsw = one_map(ot, dev, x, y, g)
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
                   'n_source': n_i, 'n_target': n_j, 'dim': CFG['dim'], **SOLVER},
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
                               'sample': 'FP64 numpy oracle: full-size cost+median, 50- and 100-iteration '
                                         'schedules at full size, extrapolated to the full map',
                               **cpu_detail}
        out['speedup_vs_cpu'] = maps_per_s * per_map_cpu
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
