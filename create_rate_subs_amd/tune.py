"""Measure emitter variants of the fused wdot kernels.

Every layout and arithmetic choice of the generated kernels (where the
per-species arrays live, the Kc form, the reaction order, the block size)
is an emitter option; the defaults are the measured winners.
This module builds a variant next to the default library and times both on
the same states, scoring each against the default FP64 kernel with the
conditioned error of create_rate_subs_amd.metrics:

    # on a machine with hipcc (no GPU needed): emit and compile
    python -m create_rate_subs_amd.tune build mech53 b128 '{"block": 128}'
    # on the GPU: time the default and the named variants, 2^20 states
    python -m create_rate_subs_amd.tune run mech53 b128
    # on the GPU: the default library against the batch size
    python -m create_rate_subs_amd.tune sweep mech53
    # on the GPU: plain FP64 FMA chains, the VALU issue ceiling
    python -m create_rate_subs_amd.tune ceiling
    # on the GPU: plain copy and the energy stage's row pattern, HBM rate
    python -m create_rate_subs_amd.tune bandwidth

Variants live under _build/variants/<mechanism>/<name>/ with the options
they were built from (opts.json).  One JSON line per (variant, mode).
"""

import json
import os
import sys
import time

from . import build, hip_emitter, workloads

__all__ = ['variant_dir', 'variant_lib', 'build_variant', 'run',
           'PARITY_VARIANTS']

TAU = 5e-2

# Non-default layouts that the GPU parity suite checks against the oracle
# like the default libraries (tests/test_gpu_block_option.py); build()
# compiles them: (mechanism, variant name) -> emitter options.
PARITY_VARIANTS = {
    ('mech13', 'b128'): {'block': 128},
    ('mech53', 'b128'): {'block': 128},
    # the mixed kernel's FP32 variant (hardware exponential cores, FP32
    # falloff and third-body blocks): faster, about 3e-6 instead of 1e-8
    ('mech53', 'fp32'): {'mixed_arith': 'fp32'},
}


def variant_lib(mech, name):
    return os.path.join(variant_dir(mech, name), 'libcrs_{}.so'.format(mech))


def variant_dir(mech, name):
    return os.path.join(workloads.BUILD_ROOT, 'variants', mech, name)


def default_groups(mech):
    """Reactions per basic block the default library of ``mech`` was built
    with (build.build_with_groups' choice), as emitter options."""
    path = os.path.join(os.path.dirname(build.lib_path(mech)), 'groups.json')
    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        choice = json.load(fh).get('choice', {})
    return {'split_' + k: v for k, v in choice.items()}


def build_variant(mech, name, opts):
    """Emit and compile mechanism ``mech`` with emitter options ``opts``
    (the key 'hipcc_flags', a list, is passed to hipcc instead).  Unless
    ``opts`` sets the group size, the variant keeps the default library's
    reactions per basic block, so that it differs from the default only in
    what ``opts`` names (build the default library first)."""
    from .create_rate_subs import load_mechanism
    mpath, tpath = workloads.mechanism_files(mech)
    _e, specs, reacs, fwd = load_mechanism(mpath, tpath)
    out = variant_dir(mech, name)
    opts = dict(opts)
    flags = opts.pop('hipcc_flags', [])
    groups = ({} if any(k.startswith('split_') for k in opts) else
              default_groups(mech))
    full = dict(groups, **opts)
    src = hip_emitter.write_hip(out, specs, reacs, name=mech, spec_order=fwd,
                                opts=full)
    lib = build.compile_library(src, out, extra_flags=flags)
    # 'opts' are the variant's own options, 'groups' the group sizes taken
    # over from the default library
    with open(os.path.join(out, 'opts.json'), 'w') as fh:
        json.dump(dict(opts=opts, groups=groups, hipcc_flags=flags,
                       resolved=hip_emitter.resolve_opts(len(specs), full)),
                  fh, indent=1)
    return lib


def run(mech, names, n=1 << 20, steps=10, seed=2001, rounds=5):
    """Time the default library and each named variant; print JSON lines."""
    import torch
    from . import metrics, runtime, states
    from .create_rate_subs import load_mechanism
    mpath, tpath = workloads.mechanism_files(mech)
    _e, specs, reacs, _f = load_mechanism(mpath, tpath)
    T, P, Y = states.random_states(len(specs), n, seed)
    Td, Pd, Yd = (torch.from_numpy(a).cuda() for a in (T, P, Y))
    base = runtime.load(build.lib_path(mech))
    ref = base.eval_wdot(Td, Pd, Yd, 'fp64')
    C = base.eval_conc(Td, Pd, Yd)['C']
    fwd, rev = base.eval_rxn_rates(Td, Pd, C)
    pm = base.get_rxn_pres_mod(Td, Pd, C)
    S = metrics.gross_scale(specs, reacs, fwd, rev, pm)
    del C, fwd, rev, pm
    libs = [('default', base)]
    for name in names:
        path = os.path.join(variant_dir(mech, name),
                            'libcrs_{}.so'.format(mech))
        libs.append((name, runtime.MechLib(path)))
    out = torch.empty_like(Yd)
    # the CONP right-hand side on the same states (row 0 = T, rows
    # 1..NSP-1 = Y), scored against the default library's
    yd = torch.cat([Td[None], Yd[:-1]]).contiguous()
    dref = base.dydt(Pd, yd)
    dout = torch.empty_like(yd)

    def call(lib, mode):
        if mode == 'dydt':
            lib.dydt(Pd, yd, out=dout)
        else:
            lib.eval_wdot(Td, Pd, Yd, mode, out=out)
    runs = [(name, lib, mode) for name, lib in libs
            for mode in ('fp64', 'mixed', 'dydt')
            if mode != 'mixed' or lib.modes & 2]
    for name, lib, mode in runs:
        for _ in range(2):
            call(lib, mode)
    # rounds interleave the libraries so that clock drift over the run
    # biases none of them; the median round is reported
    times = {(name, mode): [] for name, _lib, mode in runs}
    for _ in range(rounds):
        for name, lib, mode in runs:
            torch.cuda.synchronize()
            t0 = torch.cuda.Event(enable_timing=True)
            t1 = torch.cuda.Event(enable_timing=True)
            t0.record()
            for _ in range(steps):
                call(lib, mode)
            t1.record()
            torch.cuda.synchronize()
            times[(name, mode)].append(t0.elapsed_time(t1) / steps)
    for name, lib, mode in runs:
        call(lib, mode)
        ms = sorted(times[(name, mode)])[rounds // 2]
        if mode == 'dydt':
            finite = torch.isfinite(dout)
            e = ((dout - dref).abs() /
                 dref.abs().clamp_min(1e-300)).where(finite, 0.0)
            e = e.where(dref.abs() > 1e-30, 0.0)
        else:
            finite = torch.isfinite(out)
            e = metrics.rel_err(out, ref, S, TAU)
            e = torch.where(finite, e, torch.zeros_like(e))
        print(json.dumps(dict(
            mechanism=mech, variant=name, mode=mode, states=n, ms=ms,
            ms_rounds=times[(name, mode)], states_per_s=n / (ms * 1e-3),
            nonfinite_fraction=1.0 - finite.double().mean().item(),
            max_err_vs_default_fp64=e.max().item(), tau=TAU)), flush=True)


def sweep(mech, exps=range(10, 23, 2), steps=10, seed=2001, rounds=3):
    """States/s of the default library's fused kernels against the batch
    size (2^k states for k in exps): where the launch fills the chip.
    One JSON line per (size, mode)."""
    import torch
    from . import runtime, states
    from .create_rate_subs import load_mechanism
    mpath, tpath = workloads.mechanism_files(mech)
    _e, specs, _r, _f = load_mechanism(mpath, tpath)
    lib = runtime.load(build.lib_path(mech))
    nmax = 1 << max(exps)
    T, P, Y = states.random_states(len(specs), nmax, seed)
    Td, Pd, Yd = (torch.from_numpy(a).cuda() for a in (T, P, Y))
    # the part's clocks ramp for tens of milliseconds after idling: warm
    # up on the largest batch for half a second before the first size
    warm = torch.empty_like(Yd)
    t_end = time.perf_counter() + 0.5
    while time.perf_counter() < t_end:
        lib.eval_wdot(Td, Pd, Yd, 'fp64', out=warm)
        torch.cuda.synchronize()
    del warm
    for k in exps:
        n = 1 << k
        Tn, Pn, Yn = Td[:n], Pd[:n], Yd[:, :n].contiguous()
        out = torch.empty_like(Yn)
        for mode in ('fp64', 'mixed'):
            if mode == 'mixed' and not lib.modes & 2:
                continue
            for _ in range(2):
                lib.eval_wdot(Tn, Pn, Yn, mode, out=out)
            best = []
            for _ in range(rounds):
                torch.cuda.synchronize()
                t0 = torch.cuda.Event(enable_timing=True)
                t1 = torch.cuda.Event(enable_timing=True)
                t0.record()
                for _ in range(steps):
                    lib.eval_wdot(Tn, Pn, Yn, mode, out=out)
                t1.record()
                torch.cuda.synchronize()
                best.append(t0.elapsed_time(t1) / steps)
            ms = sorted(best)[rounds // 2]
            print(json.dumps(dict(mechanism=mech, states=n, mode=mode, ms=ms,
                                  states_per_s=n / (ms * 1e-3),
                                  blocks=(n + 255) // 256)), flush=True)


def ceiling(ks=(1, 2, 4, 8), waves=(1, 2, 4), blocks=256, n=8192,
            launches=5):
    """FP64 FMA rate of plain FMA chains (csrc/fma_probe.hip) at 1, 2 and
    4 waves per SIMD and 1-8 independent chains per lane, one block per CU:
    the VALU issue ceiling the fused kernels' VALU share is read against.
    One JSON line per (waves, chains)."""
    import ctypes
    import torch
    lib = ctypes.CDLL(build.PROBE_LIB)
    lib.crs_probe_fma.restype = ctypes.c_int
    lib.crs_probe_fma.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p] * 2
    stream = torch.cuda.current_stream()
    for w in waves:
        out = torch.empty(blocks * 256 * w, dtype=torch.float64,
                          device='cuda')
        for k in ks:
            m = max(1, n // k)

            def launch():
                rc = lib.crs_probe_fma(k, w, blocks, m, out.data_ptr(),
                                       stream.cuda_stream)
                if rc:
                    raise RuntimeError('crs_probe_fma failed')
            launch()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(launches):
                launch()
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / launches
            flops = 2.0 * blocks * 256 * w * m * 16 * k
            print(json.dumps(dict(waves_per_simd=w, chains=k, ms=ms,
                                  tflops=flops / (ms * 1e-3) / 1e12,
                                  frac_of_78p6=flops / (ms * 1e-3) / 78.6e12)),
                  flush=True)


def bandwidth(n=1 << 20, rows=53, launches=20):
    """HBM streaming rates (csrc/fma_probe.hip): a plain copy of rows x n
    doubles, and the dydt energy stage's row pattern (rows read from Y and
    from W, rows written to dy, per state) with 1, 2 and 4 states per lane.
    One JSON line each; TB/s counts every byte once."""
    import ctypes
    import torch
    lib = ctypes.CDLL(build.PROBE_LIB)
    lib.crs_probe_copy.restype = ctypes.c_int
    lib.crs_probe_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_int64, ctypes.c_void_p]
    lib.crs_probe_rows.restype = ctypes.c_int
    lib.crs_probe_rows.argtypes = [ctypes.c_void_p] * 3 + [
        ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p]
    stream = torch.cuda.current_stream()
    Y, W, dy = (torch.zeros((rows, n), dtype=torch.float64, device='cuda')
                for _ in range(3))
    nbytes = Y.numel() * 8

    def timed(launch):
        launch()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(launches):
            launch()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / launches

    def copy():
        if lib.crs_probe_copy(Y.data_ptr(), dy.data_ptr(), nbytes,
                              stream.cuda_stream):
            raise RuntimeError('crs_probe_copy failed')
    ms = timed(copy)
    print(json.dumps(dict(kernel='copy', bytes=2 * nbytes, ms=ms,
                          TBs=2 * nbytes / (ms * 1e-3) / 1e12)), flush=True)
    for per_lane in (1, 2, 4):
        def pattern():
            if lib.crs_probe_rows(Y.data_ptr(), W.data_ptr(), dy.data_ptr(),
                                  n, n, per_lane, stream.cuda_stream):
                raise RuntimeError('crs_probe_rows failed')
        ms = timed(pattern)
        print(json.dumps(dict(kernel='rows', states_per_lane=per_lane,
                              bytes=3 * nbytes, ms=ms,
                              TBs=3 * nbytes / (ms * 1e-3) / 1e12)),
              flush=True)


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) >= 4 and argv[0] == 'build':
        print(build_variant(argv[1], argv[2], json.loads(argv[3])))
    elif len(argv) >= 2 and argv[0] == 'run':
        run(argv[1], argv[2:])
    elif len(argv) >= 2 and argv[0] == 'sweep':
        sweep(argv[1])
    elif argv and argv[0] == 'ceiling':
        ceiling()
    elif argv and argv[0] == 'bandwidth':
        bandwidth()
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == '__main__':
    sys.exit(main())
