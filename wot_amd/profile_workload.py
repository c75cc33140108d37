"""Short, fixed workloads for hardware-counter and kernel-trace profiling.

    python -m wot_amd.profile_workload [--n 15000] [--iters 20] [--tile 4x512] [--cost-reps 0]

Builds the 15,000 x 15,000 day-pair cost on the device, runs ``--iters``
fixed-iteration Sinkhorn iterations (2 row log-sum-exp sweeps each, one over
C and one over C^T, 4*n*n algorithmic bytes per sweep), then one plain device
copy of a known size (n*n float32, read n*n*4 bytes, written n*n*4 bytes)
that calibrates the FETCH_SIZE / WRITE_SIZE counters.  ``--tile RxT`` runs the
row sweeps with that tile variant; ``--cost-reps k`` first builds the cost k
times in each mode (FP64 direct and split-BF16 MFMA) for kernel traces.
"""
import argparse
import sys


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--n', type=int, default=15000)
    ap.add_argument('--iters', type=int, default=20)
    ap.add_argument('--tile', default=None, help='row-sweep tile as ROWSxTHREADS')
    ap.add_argument('--cost-reps', type=int, default=0)
    args = ap.parse_args()
    import torch
    import wot_amd.ot as ot
    from wot_amd.synthetic import day_pair
    from wot_amd import tune
    dp = day_pair(args.n, args.n, dim=30, n_components=10, seed=1)
    if args.tile:
        r, t = (int(v) for v in args.tile.lower().split('x'))
        tune.apply_persisted()
        tune.set_tile(tune.ROW_SWEEP, -1, (r, t))
    for _ in range(args.cost_reps):
        for mode in ('f64', 'bf16x3', 'bf16x6'):
            ot.DeviceCost.from_coordinates(dp.x, dp.y, mode=mode)
    cost = ot.DeviceCost.from_coordinates(dp.x, dp.y)
    ot.transport_stablev2(cost, G=dp.growth, lambda1=1.0, lambda2=50.0, epsilon=0.05, scaling_iter=args.iters,
                          tau=1e4, epsilon0=1.0, extra_iter=0, inner_iter_max=50, device_out=True)
    src = torch.empty((args.n, args.n), dtype=torch.float32, device='cuda').fill_(1.0)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print('profile workload done: n=%d iters=%d copy_bytes=%d' % (args.n, args.iters, src.numel() * 4),
          file=sys.stderr, flush=True)


if __name__ == '__main__':
    main()
