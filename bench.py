#!/usr/bin/env python3
"""Throughput of the species-rate (wdot) path on one MI355X.

One step = one pass of the fused wdot kernel over the whole state batch,
inputs resident in HBM.  Default workload: the 53-species / 325-reaction
synthetic mechanism over 2^20 seeded FP64 states, FP64 path (the drop-in).

Prints one JSON line: states/s, the dominant kernel's roofline position
(algorithmic flops / its event-timed duration against MI355X's FP64 peak),
and the C oracle's throughput on the host cores over a bounded sample of
the same workload.  The mixed-precision and surrogate paths, when built,
are timed over the same states and reported with their error against the
FP64 path.
"""

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--workload', default='w1_53')
    ap.add_argument('--states', type=int, default=0,
                    help='override the workload\'s state count')
    ap.add_argument('--cpu-sample', type=int, default=1 << 16,
                    help='states in the bounded CPU-baseline sample')
    ap.add_argument('--cpu-threads', type=int, default=0)
    ap.add_argument('--no-cpu', action='store_true')
    return ap.parse_args()


def cpu_baseline(name, T, P, Y, sample, threads):
    """C oracle (restated generated C, gcc -O3 -fopenmp, FP64) on host cores."""
    from oracle.build import Oracle, build, lib_path
    if not os.path.exists(lib_path(name)):
        build(name)
    o = Oracle(name)
    m = min(sample, T.size)
    Ts, Ps = np.ascontiguousarray(T[:m]), np.ascontiguousarray(P[:m])
    Ys = np.ascontiguousarray(Y[:, :m])
    o.eval_wdot(Ts, Ps, Ys, threads)          # warm-up
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        o.eval_wdot(Ts, Ps, Ys, threads)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    return dict(value=m / t, unit='states/s', cores=threads, kind='port',
                sample='{} of the workload\'s states, {} threads, median of 3 '
                       'after 1 warm-up ({:.2f} s each); host {}'.format(
                           m, threads, t, platform.processor() or
                           platform.machine()))


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py runs on one GPU only: every box here has one MI355X '
              'and this path shards trivially over states (no exchange); '
              'multi-GPU runs are not part of this build.', file=sys.stderr)
        sys.exit(2)

    import torch
    from create_rate_subs_amd import build, cost, runtime, states, workloads
    from create_rate_subs_amd.create_rate_subs import load_mechanism

    wl = workloads.WORKLOADS[args.workload]
    n = args.states or wl.n_states
    mpath, tpath = workloads.mechanism_files(wl.mech)
    _e, specs, reacs, _f = load_mechanism(mpath, tpath)
    nsp = len(specs)
    lib = runtime.load(build.lib_path(wl.mech))
    T, P, Y = states.random_states(nsp, n, wl.state_seed)
    Td = torch.from_numpy(T).cuda()
    Pd = torch.from_numpy(P).cuda()
    Yd = torch.from_numpy(Y).cuda()
    Wd = torch.empty_like(Yd)
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        lib.eval_wdot(Td, Pd, Yd, 'fp64', out=Wd)
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record()
        lib.eval_wdot(Td, Pd, Yd, 'fp64', out=Wd)
        ends[k].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
    ms_per_step = wall * 1e3 / args.steps
    value = n / (ms_per_step * 1e-3)

    fl = cost.flops_for(specs, reacs)['total']
    by = cost.bytes_per_state(nsp)
    achieved = fl * n / (kern_ms * 1e-3) / 1e12
    roofline = dict(bound='mfma', achieved=achieved, peak=FP64_PEAK_TFLOPS,
                    unit='TFLOP/s', frac=achieved / FP64_PEAK_TFLOPS,
                    traffic=None, kernel='k_wdot_f64',
                    kernel_ms=kern_ms, flops_per_state=fl,
                    bytes_per_state=by,
                    hbm_GBs=by * n / (kern_ms * 1e-3) / 1e9,
                    hbm_frac=by * n / (kern_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    note='FP64 VALU-bound; peak is MI355X FP64 (vector and '
                         'matrix are both 78.6 TFLOP/s)')

    line = dict(metric='wdot states/s (53sp/325rxn): FP64 vs mixed vs '
                       'surrogate @ stated rel err',
                value=value, unit='states/s', n_gpus=1, steps=args.steps,
                warmup=args.warmup, ms_per_step=ms_per_step,
                higher_is_better=True, scaling='weak', vs_baseline=None,
                dtype='f64', data='synthetic',
                config=dict(workload='{}: {} sp / {} rxn, {} FP64 states, '
                                     'FP64 path'.format(wl.name, nsp,
                                                        len(reacs), n),
                            mechanism=wl.mech, states=n, mode='fp64'),
                roofline=roofline)
    if not args.no_cpu:
        threads = args.cpu_threads or int(os.environ.get(
            'OMP_NUM_THREADS', min(16, os.cpu_count() or 1)))
        line['cpu_baseline'] = cpu_baseline(wl.mech, T, P, Y,
                                            args.cpu_sample, threads)
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
