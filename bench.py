#!/usr/bin/env python3
"""Throughput of the species-rate (wdot) path on one MI355X.

One step = one launch of the fused wdot kernel over the whole state batch,
inputs resident in HBM.  Headline: the 53-species / 325-reaction synthetic
mechanism over 2^20 seeded FP64 states on the FP64 path (the drop-in).

The same JSON line also carries the mixed-precision path on 2^22 states
(plus a 2^18 near-equilibrium slice) with its speed-up over the FP64 kernel
on the same states and its conditioned relative error against FP64, and the
C oracle's throughput on the host cores over a bounded sample.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0      # dense BF16 MFMA, MI355X
HBM_PEAK_GBS = 8000.0
TAU = 5e-2


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=50)
    # untimed launches before each timed window; at least WARM_S seconds
    # of them run whatever the count
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--workload', default='w1_53')
    ap.add_argument('--states', type=int, default=0,
                    help='override the workload\'s state count')
    ap.add_argument('--cpu-sample', type=int, default=1 << 20,
                    help='states in the bounded CPU-baseline sample (the '
                         'whole 2^20-state workload by default: about 20 s '
                         'of CPU work over the timed passes)')
    ap.add_argument('--cpu-threads', type=int, default=0)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-mixed', action='store_true')
    ap.add_argument('--no-surrogate', action='store_true')
    ap.add_argument('--no-large', action='store_true',
                    help='skip the 111-species / 784-reaction workload')
    ap.add_argument('--sur-epochs', type=int, default=120)
    ap.add_argument('--sur-target', type=float, default=1e-2)
    ap.add_argument('--probe', action='store_true',
                    help='one FP64 and one mixed launch plus a plain copy of '
                         'known size, for hardware-counter runs')
    return ap.parse_args()


def log(msg):
    """Progress of the run (stderr; the result line alone goes to stdout)."""
    print('[bench] ' + msg, file=sys.stderr, flush=True)


def cpu_threads(requested=0):
    """Threads for the CPU baseline: the CPUs this process may use, capped by
    OMP_NUM_THREADS when set (a GPU box grants each GPU's job its share of
    the host, 16 CPUs, and sets OMP_NUM_THREADS to it)."""
    if requested:
        return requested
    allowed = len(os.sched_getaffinity(0))
    env = os.environ.get('OMP_NUM_THREADS')
    return min(allowed, int(env)) if env else allowed


def cpu_baseline(workload, sample, threads):
    """The C oracle (restated generated C, gcc -O3 -fopenmp, FP64) on the
    host cores, in a child process (oracle/cpu_baseline.py) so that
    OMP_PROC_BIND=close and OMP_PLACES=cores reach libgomp at its start."""
    import subprocess
    env = dict(os.environ, OMP_PROC_BIND='close', OMP_PLACES='cores',
               OMP_NUM_THREADS=str(threads))
    res = subprocess.run([sys.executable, '-m', 'oracle.cpu_baseline',
                          workload, str(sample), str(threads)], cwd=ROOT,
                         env=env, capture_output=True, text=True,
                         timeout=600)
    if res.returncode != 0:
        raise RuntimeError('CPU baseline failed:\n' + res.stderr[-2000:])
    return json.loads(res.stdout.strip().splitlines()[-1])


def measured_traffic(kernel, workload, n):
    """HBM bytes per launch of ``kernel`` from its committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), scaled to n states;
    None when no summary covers this kernel and workload."""
    import glob
    import re
    here = os.path.dirname(os.path.abspath(__file__))
    for path in sorted(glob.glob(os.path.join(here, 'profiles', '*.md'))):
        text = open(path).read()
        for block in re.findall(r'```json\n(.*?)```', text, re.S):
            try:
                d = json.loads(block)
            except ValueError:
                continue
            if d.get('kernel') == kernel and d.get('workload') == workload:
                per_state = d['traffic_bytes_per_launch'] / d['states']
                return per_state * n, os.path.relpath(path, here)
    return None, None


# Least wall time of untimed warm-up before a timed window.  After the part
# has idled (at the start of the run, after the CPU baseline) its clocks
# take tens of milliseconds of work to ramp: the first ~20 launches of the
# headline kernel average 0.71 ms against 0.62 ms once warm, so a count of
# warm-up launches alone does not reach the steady state.
WARM_S = 0.5


def timed(torch, fn, steps, warmup, info=None):
    """Run fn warmup times, and on (untimed, 8 launches at a time) until
    WARM_S seconds have passed since the first; then time exactly steps
    launches with events and wall.  Returns (wall ms per launch, mean event
    ms per launch); ``info``, a dict, receives the warm-up launch count."""
    t0 = time.perf_counter()
    done = 0
    for _ in range(warmup):
        fn()
        done += 1
    torch.cuda.synchronize()
    while time.perf_counter() - t0 < WARM_S:
        for _ in range(8):
            fn()
        done += 8
        torch.cuda.synchronize()
    if info is not None:
        info['warmup_launches'] = done
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    t0 = time.perf_counter()
    for k in range(steps):
        starts[k].record()
        fn()
        ends[k].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
    return wall * 1e3 / steps, kern


def gross_scale_dev(torch, lib, specs, reacs, T, P, Y, chunk=1 << 19):
    """S_k of every state on the GPU, from the FP64 stage kernels, in chunks."""
    from create_rate_subs_amd import metrics
    n = T.numel()
    S = torch.empty_like(Y)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        Tc, Pc = T[a:b].contiguous(), P[a:b].contiguous()
        C = lib.eval_conc(Tc, Pc, Y[:, a:b].contiguous())['C']
        fwd, rev = lib.eval_rxn_rates(Tc, Pc, C)
        pm = lib.get_rxn_pres_mod(Tc, Pc, C)
        S[:, a:b] = metrics.gross_scale(specs, reacs, fwd, rev, pm)
        del C, fwd, rev, pm
    return S


def err_summary(torch, W, ref, S):
    from create_rate_subs_amd import metrics
    e = metrics.rel_err(W, ref, S, TAU)
    s = metrics.summarize(e)
    ok = (ref.abs() >= TAU * S) & (S > 0)
    pw = ((W - ref).abs() / ref.abs())[ok]
    s['tau'] = TAU
    s['well_conditioned_pointwise_max'] = pw.max().item() if pw.numel() else 0.0
    s['well_conditioned_fraction'] = ok.double().mean().item()
    return s


def large_section(args, torch):
    """111 species / 784 reactions, 2^20 states: FP64 and mixed paths."""
    from create_rate_subs_amd import build, cost, runtime, states, workloads
    from create_rate_subs_amd.create_rate_subs import load_mechanism
    w3 = workloads.WORKLOADS['w3_111']
    mpath, tpath = workloads.mechanism_files(w3.mech)
    _e, specs, reacs, _f = load_mechanism(mpath, tpath)
    lib = runtime.load(build.lib_path(w3.mech))
    nsp = len(specs)
    T, P, Y = states.random_states(nsp, w3.n_states, w3.state_seed)
    Td, Pd, Yd = (torch.from_numpy(a).cuda() for a in (T, P, Y))
    W64 = torch.empty_like(Yd)
    Wmx = torch.empty_like(Yd)
    steps = max(3, args.steps // 2)
    f64_ms, f64_k = timed(torch, lambda: lib.eval_wdot(
        Td, Pd, Yd, 'fp64', out=W64), steps, 1)
    out = dict(workload='w3_111: {} sp / {} rxn, {} states'.format(
        nsp, len(reacs), w3.n_states),
        fp64_states_per_s=w3.n_states / (f64_ms * 1e-3), fp64_kernel_ms=f64_k,
        flops_per_state=cost.flops_for(specs, reacs)['total'],
        bytes_per_state=cost.bytes_per_state(nsp))
    out['fp64_achieved_tflops'] = (out['flops_per_state'] * w3.n_states /
                                   (f64_k * 1e-3) / 1e12)
    out['fp64_frac'] = out['fp64_achieved_tflops'] / FP64_PEAK_TFLOPS
    if lib.modes & 2:
        mx_ms, mx_k = timed(torch, lambda: lib.eval_wdot(
            Td, Pd, Yd, 'mixed', out=Wmx), steps, 1)
        S = gross_scale_dev(torch, lib, specs, reacs, Td, Pd, Yd,
                            chunk=1 << 18)
        out.update(mixed_states_per_s=w3.n_states / (mx_ms * 1e-3),
                   mixed_kernel_ms=mx_k, mixed_speedup_vs_fp64=f64_k / mx_k,
                   mixed_err_vs_fp64=err_summary(torch, Wmx, W64, S))
    return out


def dydt_section(args, torch, lib, T, P, Y):
    """Constant-pressure dydt (the caller of the path) on the headline
    states: the same fused evaluation plus cp, h, the energy equation and
    the mass-fraction rates."""
    n = T.numel()
    nsp = lib.nsp
    y = torch.empty((nsp, n), dtype=torch.float64, device='cuda')
    y[0] = T
    y[1:] = Y[:nsp - 1]
    dy = torch.empty_like(y)
    steps = max(3, args.steps // 2)
    ms, kern = timed(torch, lambda: lib.dydt(P, y, out=dy), steps, 1)
    out = dict(workload='w1_53 states in the reference dydt layout '
                        '(row 0 = T, rows 1..NSP-1 = Y)',
               states_per_s=n / (ms * 1e-3), kernel_ms=kern,
               kernel='k_wdot_dydt + k_dydt_energy')
    if lib.modes & 2:
        mx_ms, mx_kern = timed(
            torch, lambda: lib.dydt(P, y, out=dy, mode='mixed'), steps, 1)
        out['mixed'] = dict(
            kernel='k_wdot_mixed_dydt + k_dydt_energy', kernel_ms=mx_kern,
            states_per_s=n / (mx_ms * 1e-3), speedup_vs_fp64=kern / mx_kern,
            parity='<= 1e-6 conditioned against the oracle '
                   '(tests/test_gpu_dydt.py)')
    return out


def integrate_section(args, torch, lib, specs, n, steps=10):
    """The device-resident time loop (create_rate_subs_amd.integrate): RK4
    steps of the CONP system on n near-equilibrium states of the headline
    mechanism (2^16 distinct element-potential equilibria perturbed by 1e-3,
    repeated; random compositions of the synthetic mechanisms are too stiff
    for any explicit step), the whole loop replayed from one captured HIP
    graph.  The step, 1e-39 s, keeps every state finite (reported); the
    work per step does not depend on it."""
    from create_rate_subs_amd import integrate, states
    nsp = lib.nsp
    m = min(n, 1 << 16)
    Te, Pe, Ye = states.equilibrium_states(specs, m, 2101, perturb=1e-3)
    reps = (n + m - 1) // m
    P = torch.from_numpy(Pe).cuda().repeat(reps)[:n].contiguous()
    y = torch.empty((nsp, n), dtype=torch.float64, device='cuda')
    y[0] = torch.from_numpy(Te).cuda().repeat(reps)[:n]
    y[1:] = torch.from_numpy(Ye[:nsp - 1]).cuda().repeat(1, reps)[:, :n]
    h = 1.0e-39
    out = integrate.integrate(lib, P, y, 0.0, steps * h, h)
    finite = float(torch.isfinite(y).all(0).double().mean())
    return dict(workload='{} near-equilibrium states of the w1_53 '
                         'mechanism, {} RK4 steps (4 wdot + 4 energy-and-stage kernels '
                         'each) from one graph'.format(n, out['steps']),
                steps=out['steps'], h=h, ms_per_step=out['ms_per_step'],
                state_steps_per_s=n * out['steps'] / (out['ms_total'] * 1e-3),
                finite_fraction=finite)


def _quantiles(torch, x, qs=(0.01, 0.1, 0.5, 0.9, 0.99)):
    """Quantiles of a per-state error vector (strided sample if large)."""
    x = x[::max(1, x.numel() >> 20)].double()
    x = torch.where(torch.isfinite(x), x, torch.full_like(x, float('inf')))
    return {str(q): x.quantile(q).item() for q in qs}


def surrogate_section(args, torch, lib, specs, reacs):
    """Train on 2^22 exact-labelled states, calibrate the gate on held-out
    states, then time gated inference with exact fallback on fresh states."""
    from create_rate_subs_amd import cost, metrics, states, workloads
    from create_rate_subs_amd.surrogate import Surrogate
    w4 = workloads.WORKLOADS['w4_53']
    nsp = len(specs)
    n = w4.n_states

    def labelled(count, seed):
        T, P, Y = states.random_states(nsp, count, seed)
        Td, Pd, Yd = (torch.from_numpy(a).cuda() for a in (T, P, Y))
        return Td, Pd, Yd, lib.eval_wdot(Td, Pd, Yd, 'fp64')

    out = dict(workload='w4_53: train {} states, test {} states'.format(n, n))
    Tt, Pt, Yt, Wt = labelled(n, w4.state_seed)
    Tv, Pv, Yv, Wv = labelled(1 << 18, w4.state_seed + 100)
    net = Surrogate(lib, specs, reacs, seed=w4.state_seed)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    log('surrogate: training {} epochs on {} states'.format(args.sur_epochs,
                                                             n))
    hist = net.train(Tt, Pt, Yt, Wt, epochs=args.sur_epochs, batch=16384)
    torch.cuda.synchronize()
    ttrain = time.perf_counter() - t0
    out['train'] = dict(states=n, epochs=args.sur_epochs, seconds=ttrain,
                        state_epochs_per_s=n * args.sur_epochs / ttrain,
                        mse_first=hist[0], mse_last=hist[-1],
                        train_mse_by_epoch=hist[::max(1, len(hist) // 12)])
    # a second member from another initialisation, trained with the
    # validation loss per epoch: the learning curve, and the disagreement
    # of the two as a gate signal
    log('surrogate: second network (validation curve, ensemble signal)')
    net2 = Surrogate(lib, specs, reacs, seed=w4.state_seed + 1000)
    hist2 = net2.train(Tt, Pt, Yt, Wt, epochs=args.sur_epochs, batch=16384,
                       val=(Tv, Pv, Yv, Wv))
    step = max(1, len(hist2) // 12)
    out['learning_curve'] = dict(
        epochs=list(range(1, len(hist2) + 1))[::step] + [len(hist2)],
        train_mse=hist2[::step] + [hist2[-1]],
        val_mse=net2.val_history[::step] + [net2.val_history[-1]],
        schedule='cosine from 2e-3 to 1e-4 over the run')
    del Tt, Pt, Yt, Wt
    # gate calibration on held-out states
    Sv = gross_scale_dev(torch, lib, specs, reacs, Tv, Pv, Yv)
    # how well each gate signal ranks the states by error
    Wpv = torch.empty_like(Wv)
    net.predict(Tv, Pv, Yv, Wpv)
    e_nw = metrics.normwise_err(Wpv, Wv)
    e_c = metrics.rel_err(Wpv, Wv, Sv, TAU).amax(dim=0)
    sig = dict(balance_residual=net.residuals(Wpv),
               ensemble_disagreement=net.disagreement(net2, Tv, Pv, Yv))
    from create_rate_subs_amd.surrogate import gate_threshold
    gates = {}
    for sname, sv in sig.items():
        g = dict(spearman_vs_normwise=metrics.spearman(sv, e_nw),
                 spearman_vs_conditioned=metrics.spearman(sv, e_c))
        for mname, e in (('normwise', e_nw), ('conditioned', e_c)):
            g['admitted_' + mname] = {
                str(t): gate_threshold(e, sv, t)[1] for t in (1e-2, 1e-1)}
        # admitting the lowest-signal fraction f of the states: the
        # normwise error of what is admitted
        order = torch.argsort(torch.where(torch.isfinite(sv), sv,
                                          torch.full_like(sv, 1e300)))
        tab = []
        for f in (0.001, 0.01, 0.1, 0.5):
            k = max(1, int(f * sv.numel()))
            ea = e_nw[order[:k]].double()
            tab.append(dict(fraction=f, normwise_max=ea.max().item(),
                            normwise_p99=ea.quantile(0.99).item(),
                            normwise_median=ea.median().item()))
        g['lowest_signal_fraction'] = tab
        gates[sname] = g
    out['gate_signals'] = gates
    del Wpv, net2
    _r, _a, curve_nw = net.calibrate_gate(Tv, Pv, Yv, Wv, Sv, args.sur_target,
                                          TAU, metric='normwise')
    rmax, admitted, curve = net.calibrate_gate(Tv, Pv, Yv, Wv, Sv,
                                               args.sur_target, TAU)
    out['gate'] = dict(target=args.sur_target, tau=TAU, rmax=rmax,
                       validation_admitted=admitted,
                       tradeoff_on_validation=curve,
                       tradeoff_normwise_on_validation=curve_nw)
    del Tv, Pv, Yv, Wv, Sv
    # fresh test states
    Te, Pe, Ye, We = labelled(n, w4.state_seed + 200)
    Wout = torch.empty_like(We)
    steps = 3
    f64_ms, _k = timed(torch, lambda: lib.eval_wdot(Te, Pe, Ye, 'fp64',
                                                     out=Wout), steps, 1)
    pred_ms, pred_kern_ms = timed(
        torch, lambda: net.predict(Te, Pe, Ye, Wout), steps, 1)
    fb = []

    def gated():
        # one C-ABI call (crs_mlp_eval_gated): inference, gate and exact
        # fallback on the stream, the fallback count left on the device
        fb.append(net.evaluate(Te, Pe, Ye, Wout, read_count=False))
    gated_ms, _k = timed(torch, gated, steps, 1)
    S = gross_scale_dev(torch, lib, specs, reacs, Te, Pe, Ye)
    e = metrics.rel_err(Wout, We, S, TAU)
    per_state = e.max(dim=0).values
    nw = metrics.normwise_err(Wout, We)
    # the network alone, before the gate (fused kernel, all states)
    Wsur = torch.empty_like(We)
    net.predict(Te, Pe, Ye, Wsur)
    e_raw = metrics.rel_err(Wsur, We, S, TAU).max(dim=0).values
    nw_raw = metrics.normwise_err(Wsur, We)
    del Wsur
    flops = cost.surrogate_flops_per_state(net.nin, net.hid,
                                           len(net.shapes) - 1, net.nout)
    achieved = flops * n / (pred_ms * 1e-3) / 1e12
    out['kernel'] = dict(
        name='k_mlp_infer', ms=pred_ms, flops_per_state=flops,
        achieved_tflops=achieved, peak_tflops=BF16_PEAK_TFLOPS,
        frac=achieved / BF16_PEAK_TFLOPS,
        note='fused features + 55-256-256-256-53 GELU MLP (BF16 MFMA, FP32 '
             'accumulate) + decode + gate, one launch; flops are the '
             'unpadded 2*MACs')
    out['network_alone'] = dict(
        conditioned_state_quantiles=_quantiles(torch, e_raw),
        normwise_state_quantiles=_quantiles(torch, nw_raw),
        frac_conditioned_le_target=(e_raw <= args.sur_target).double().mean().item(),
        frac_normwise_le_target=(nw_raw <= args.sur_target).double().mean().item(),
        frac_normwise_le_0p1=(nw_raw <= 0.1).double().mean().item())
    out['test'] = dict(
        states=n, gate_admits_none=rmax < 0,
        gated_path=('the calibrated gate admits no state (rmax < 0): the '
                    'gated call skips the network and evaluates every '
                    'state exactly' if rmax < 0 else
                    'network, gate and exact fallback in one call'),
        gated_states_per_s=n / (gated_ms * 1e-3),
        surrogate_only_states_per_s=n / (pred_ms * 1e-3),
        fp64_states_per_s=n / (f64_ms * 1e-3),
        speedup_vs_fp64=f64_ms / gated_ms,
        fallback_rate=int(fb[-1].item()) / n,
        post_fallback_err=dict(max=e.max().item(),
                               state_median=per_state.median().item(),
                               state_p99=per_state[::16].quantile(0.99).item(),
                               tau=TAU, normwise_max=nw.max().item()),
        audit_1pct=net.audit(Te, Pe, Ye, Wout, frac=0.01))
    return out


def integrate_nonstiff_section(torch, n=1 << 20, steps=20, h=1.0e-5):
    """The same RK4 harness with a step that moves the states: mech9ns (9
    species / 14 reactions, a synthetic mechanism generated non-stiff, every
    Jacobian eigenvalue below ~2e4 1/s at 1200-1500 K and 1 atm), n states
    at 1200-1500 K and 1 atm, `steps` steps of h seconds from one captured
    HIP graph.  That the harness converges at fourth order on this
    mechanism is a test of its own; here the throughput and how far the
    temperatures moved are reported."""
    from create_rate_subs_amd import build, integrate, runtime, states
    lib = runtime.load(build.lib_path('mech9ns'))
    nsp = lib.nsp
    T, _P, Y = states.random_states(nsp, n, 3101, t_range=(1200.0, 1500.0))
    P = torch.full((n,), 101325.0, dtype=torch.float64, device='cuda')
    y = torch.empty((nsp, n), dtype=torch.float64, device='cuda')
    y[0] = torch.from_numpy(T).cuda()
    y[1:] = torch.from_numpy(Y[:nsp - 1]).cuda()
    T0 = y[0].clone()
    out = integrate.integrate(lib, P, y, 0.0, steps * h, h)
    dT = (y[0] - T0).abs()
    return dict(workload='mech9ns (9 sp / 14 rxn, non-stiff), {} states at '
                         '1200-1500 K and 1 atm, {} RK4 steps of {:g} s from '
                         'one graph'.format(n, out['steps'], h),
                steps=out['steps'], h=h, ms_per_step=out['ms_per_step'],
                state_steps_per_s=n * out['steps'] / (out['ms_total'] * 1e-3),
                max_abs_dT_K=dT.max().item(),
                median_abs_dT_K=dT.median().item(),
                finite_fraction=float(torch.isfinite(y).all(0).double()
                                      .mean()))


def main():
    args = parse()
    if args.gpus != 1:
        print('bench.py measures one GPU only: the path shards trivially '
              'over states (no exchange between GPUs), so N GPUs run N '
              'independent copies of the one-GPU workload; no multi-GPU '
              'run is made.', file=sys.stderr)
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

    if args.probe:
        # counter calibration: a device copy of exactly 256 MiB read and
        # written, next to one launch of each fused kernel on the workload
        src = torch.ones(1 << 25, dtype=torch.float64, device='cuda')
        dst = torch.empty_like(src)
        dst.copy_(src)
        lib.eval_wdot(Td, Pd, Yd, 'fp64', out=Wd)
        if lib.modes & 2:
            lib.eval_wdot(Td, Pd, Yd, 'mixed', out=Wd)
        # one CONP dydt launch on the same states (row 0 = T)
        yd = torch.cat([Td[None], Yd[:nsp - 1]]).contiguous()
        lib.dydt(Pd, yd, out=torch.empty_like(yd))
        del yd
        sur = None
        if nsp + 2 <= 64:
            # one fused surrogate launch (untrained weights: traffic and
            # time do not depend on the weights' values)
            from create_rate_subs_amd.surrogate import Surrogate
            sur = Surrogate(lib, specs, reacs, seed=1)
            sur.fit_normalisation(Td, Pd, Yd, Wd)
            sur.gate.rmax = 1e300
            sur.pack()
            torch.cuda.synchronize()
            Ws = torch.empty_like(Wd)
            sur.infer(Td, Pd, Yd, Ws)
        torch.cuda.synchronize()
        print(json.dumps(dict(probe=True, states=n, copy_bytes=8 << 25,
                              surrogate=sur is not None)))
        return

    warm = {}
    ms_per_step, kern_ms = timed(
        torch, lambda: lib.eval_wdot(Td, Pd, Yd, 'fp64', out=Wd),
        args.steps, args.warmup, info=warm)
    value = n / (ms_per_step * 1e-3)
    log('fp64 {}: {:.3e} states/s'.format(wl.name, value))

    fl = cost.flops_for(specs, reacs)['total']
    by = cost.bytes_per_state(nsp)
    achieved = fl * n / (kern_ms * 1e-3) / 1e12
    traffic, traffic_src = measured_traffic('k_wdot_f64', wl.name, n)
    roofline = dict(bound='fp64_valu', achieved=achieved,
                    peak=FP64_PEAK_TFLOPS,
                    unit='TFLOP/s', frac=achieved / FP64_PEAK_TFLOPS,
                    traffic=traffic, traffic_source=traffic_src,
                    kernel='k_wdot_f64',
                    kernel_ms=kern_ms, flops_per_state=fl,
                    bytes_per_state=by,
                    hbm_GBs=by * n / (kern_ms * 1e-3) / 1e9,
                    hbm_frac=by * n / (kern_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    note='FP64 VALU-bound; peak is MI355X FP64 (vector and '
                         'matrix both 78.6 TFLOP/s); flops counted as in the '
                         'reference\'s emitted expressions, exp/log = 1')
    line = dict(metric='wdot states/s (53sp/325rxn): FP64 vs mixed vs '
                       'surrogate @ stated rel err',
                value=value, unit='states/s', n_gpus=1, steps=args.steps,
                warmup=args.warmup, ms_per_step=ms_per_step,
                warmup_launches=warm['warmup_launches'], warmup_min_s=WARM_S,
                higher_is_better=True, scaling='weak', vs_baseline=None,
                dtype='f64', data='synthetic',
                config=dict(workload='{}: {} sp / {} rxn, {} FP64 states, '
                                     'FP64 path'.format(wl.name, nsp,
                                                        len(reacs), n),
                            mechanism=wl.mech, states=n, mode='fp64'),
                roofline=roofline)
    if not args.no_cpu:
        line['cpu_baseline'] = cpu_baseline(wl.name, args.cpu_sample,
                                            cpu_threads(args.cpu_threads))
        log('cpu baseline: {:.3e} states/s'.format(
            line['cpu_baseline']['value']))
    line['dydt'] = dydt_section(args, torch, lib, Td, Pd, Yd)
    log('dydt: {:.3e} states/s'.format(line['dydt']['states_per_s']))
    line['integrate'] = integrate_section(args, torch, lib, specs, n)
    log('integrate: {:.3f} ms per RK4 step'.format(
        line['integrate']['ms_per_step']))
    line['integrate_nonstiff'] = integrate_nonstiff_section(torch)
    log('integrate (non-stiff): {:.3f} ms per RK4 step, max |dT| {:.1f} K'
        .format(line['integrate_nonstiff']['ms_per_step'],
                line['integrate_nonstiff']['max_abs_dT_K']))
    del Td, Pd, Yd, Wd

    if not args.no_mixed and (lib.modes & 2):
        w2 = workloads.WORKLOADS['w2_53']
        T2, P2, Y2 = states.random_states(nsp, w2.n_states, w2.state_seed)
        T2d = torch.from_numpy(T2).cuda()
        P2d = torch.from_numpy(P2).cuda()
        Y2d = torch.from_numpy(Y2).cuda()
        W64 = torch.empty_like(Y2d)
        Wmx = torch.empty_like(Y2d)
        steps = max(3, args.steps // 2)
        f64_ms, f64_k = timed(torch, lambda: lib.eval_wdot(
            T2d, P2d, Y2d, 'fp64', out=W64), steps, 1)
        mx_ms, mx_k = timed(torch, lambda: lib.eval_wdot(
            T2d, P2d, Y2d, 'mixed', out=Wmx), steps, 1)
        S = gross_scale_dev(torch, lib, specs, reacs, T2d, P2d, Y2d)
        mixed = dict(workload='w2_53: {} states'.format(w2.n_states),
                     arithmetic='FP64 Kc (G_k), net rate and sums; Arrhenius '
                                'and falloff transcendentals at ~1e-8 '
                                '(degree-4 exponential)',
                     states_per_s=w2.n_states / (mx_ms * 1e-3),
                     fp64_states_per_s=w2.n_states / (f64_ms * 1e-3),
                     kernel_ms=mx_k, fp64_kernel_ms=f64_k,
                     speedup_vs_fp64=f64_k / mx_k,
                     err_vs_fp64=err_summary(torch, Wmx, W64, S))
        # every counted operation of the default split is FP64 arithmetic
        # (the shorter exponentials are degree-4 FP64 polynomials; the 52
        # FP32 log2 seeds per state are not counted flops), so the
        # time-weighted FP32/FP64 peak is the FP64 one
        mx_tf = fl * w2.n_states / (mx_k * 1e-3) / 1e12
        mixed['roofline'] = dict(
            bound='fp64_valu', achieved=mx_tf, peak=FP64_PEAK_TFLOPS,
            unit='TFLOP/s', frac=mx_tf / FP64_PEAK_TFLOPS,
            flops_per_state=fl, fp32_flops_per_state=0,
            note='the same counted flops as the FP64 kernel, all in FP64 '
                 'arithmetic in this split')
        # the FP32 variant (emitter option mixed_arith='fp32', built beside
        # the default library): hardware FP32 exponential cores, FP32
        # falloff and third-body blocks
        from create_rate_subs_amd import tune
        vpath = tune.variant_lib(wl.mech, 'fp32')
        if os.path.exists(vpath):
            vlib = runtime.load(vpath)
            v_ms, v_k = timed(torch, lambda: vlib.eval_wdot(
                T2d, P2d, Y2d, 'mixed', out=Wmx), steps, 1)
            mixed['fp32_variant'] = dict(
                states_per_s=w2.n_states / (v_ms * 1e-3), kernel_ms=v_k,
                speedup_vs_fp64=f64_k / v_k,
                err_vs_fp64=err_summary(torch, Wmx, W64, S))
        del T2d, P2d, Y2d, W64, Wmx, S
        Te, Pe, Ye = states.equilibrium_states(specs, w2.near_eq_states,
                                               w2.state_seed + 100)
        Ted = torch.from_numpy(Te).cuda()
        Ped = torch.from_numpy(Pe).cuda()
        Yed = torch.from_numpy(Ye).cuda()
        We64 = lib.eval_wdot(Ted, Ped, Yed, 'fp64')
        Wemx = lib.eval_wdot(Ted, Ped, Yed, 'mixed')
        Se = gross_scale_dev(torch, lib, specs, reacs, Ted, Ped, Yed)
        mixed['near_eq'] = dict(states=w2.near_eq_states,
                                err_vs_fp64=err_summary(torch, Wemx, We64, Se))
        line['mixed'] = mixed
        log('mixed: {:.3e} states/s, speed-up {:.2f}'.format(
            mixed['states_per_s'], mixed['speedup_vs_fp64']))
    if not args.no_large:
        line['large'] = large_section(args, torch)
        log('large: {:.3e} states/s fp64'.format(
            line['large']['fp64_states_per_s']))
    if not args.no_surrogate:
        line['surrogate'] = surrogate_section(args, torch, lib, specs, reacs)
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
