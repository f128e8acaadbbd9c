"""Compile emitted HIP sources into mechanism libraries for gfx950.

``compile_library(src, out_dir)`` runs hipcc on one ``crs_<name>.hip`` and
writes ``libcrs_<name>.so`` next to it.  ``build_mechanism(name)`` emits and
compiles one of the synthetic mechanisms into ``_build/<name>/``;
``build_with_groups`` picks, per fused kernel, the reaction-group size the
register allocator spills least at (see its docstring).
"""

import hashlib
import json
import os
import re
import subprocess
import tempfile
import time
from concurrent.futures import ThreadPoolExecutor

from . import hip_emitter, workloads
from .create_rate_subs import load_mechanism

__all__ = ['HIPCC_FLAGS', 'compile_library', 'build_mechanism', 'lib_path',
           'INCLUDE_DIRS', 'build_mlp', 'build_with_groups',
           'parse_resource_usage', 'GROUP_CANDIDATES']

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INCLUDE_DIRS = [os.path.join(ROOT, 'include'), os.path.join(HERE, 'csrc')]

# -packed-fp32-ops: keep the compiler from forming v_pk_{add,mul,fma}_f32.
HIPCC_FLAGS = ['-O3', '--offload-arch=gfx950', '-std=c++17', '-fPIC',
               '-shared', '-Xclang', '-target-feature', '-Xclang',
               '-packed-fp32-ops']


def hipcc():
    rocm = os.environ.get('ROCM_PATH', '/opt/rocm')
    return os.path.join(rocm, 'bin', 'hipcc')


_REMARK = '-Rpass-analysis=kernel-resource-usage'


def compile_library(src, out_dir=None, extra_flags=(), verbose=False,
                    usage_out=None):
    """hipcc ``src`` -> ``libcrs_<name>.so``; return the library path.
    With a dict ``usage_out``, the compiler's per-kernel resource report
    (parse_resource_usage) is stored in it."""
    out_dir = out_dir or os.path.dirname(src)
    base = os.path.splitext(os.path.basename(src))[0]
    lib = os.path.join(out_dir, 'lib{}.so'.format(base))
    cmd = [hipcc()] + HIPCC_FLAGS + list(extra_flags)
    if usage_out is not None:
        cmd.append(_REMARK)
    for d in INCLUDE_DIRS:
        cmd += ['-I', d]
    cmd += ['-o', lib + '.tmp', src]
    t0 = time.time()
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError('hipcc failed for {}:\n{}'.format(
            src, res.stderr[-4000:]))
    os.replace(lib + '.tmp', lib)
    if usage_out is not None:
        usage_out.update(parse_resource_usage(res.stderr))
    if verbose:
        print('built {} in {:.1f} s'.format(lib, time.time() - t0))
    return lib


def parse_resource_usage(text):
    """{kernel symbol: {field: int}} from hipcc's kernel-resource-usage
    remarks (fields such as 'VGPRs', 'VGPRs Spill', 'ScratchSize
    [bytes/lane]')."""
    usage, cur = {}, None
    for line in text.splitlines():
        if 'remark:' not in line:
            continue
        body = line.split('remark:', 1)[1].split(' [-Rpass-analysis')[0]
        body = body.strip()
        if body.startswith('Function Name:'):
            cur = usage.setdefault(body.split(':', 1)[1].strip(), {})
        elif cur is not None and ':' in body:
            key, val = body.rsplit(':', 1)
            try:
                cur[key.strip()] = int(val)
            except ValueError:
                pass
    return usage


# a function's entry label (local labels start with '.')
_LABEL = re.compile(r'^([A-Za-z_][\w$]*):')


def count_scratch_ops(asm):
    """{kernel symbol: number of scratch load/store instructions} in gfx950
    assembly text (for straight-line kernels, also the number each wave
    executes)."""
    counts, cur = {}, None
    for line in asm.splitlines():
        lab = _LABEL.match(line)
        if lab:
            cur = lab.group(1)
            counts.setdefault(cur, 0)
        elif cur is not None and line.lstrip().startswith('scratch_'):
            counts[cur] += 1
    return counts


def device_usage(src, extra_flags=()):
    """Resource report of ``src``'s device code alone (no library written),
    with each kernel's scratch instruction count as 'Scratch ops'."""
    with tempfile.TemporaryDirectory() as tmp:
        asm = os.path.join(tmp, 'dev.s')
        cmd = ([hipcc(), '--cuda-device-only', '-S', '-o', asm, _REMARK] +
               [f for f in HIPCC_FLAGS if f not in ('-shared', '-fPIC')] +
               list(extra_flags))
        for d in INCLUDE_DIRS:
            cmd += ['-I', d]
        cmd.append(src)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError('hipcc failed for {}:\n{}'.format(
                src, res.stderr[-4000:]))
        with open(asm) as fh:
            ops = count_scratch_ops(fh.read())
    usage = parse_resource_usage(res.stderr)
    for fn, u in usage.items():
        u['Scratch ops'] = ops.get(fn, 0)
    return usage


# the fused kernels (emitter name -> symbol tag) and the reaction-group
# sizes build_with_groups tries for them
FUSED_KERNELS = (('f64', 'k_wdot_f64E'), ('mixed', 'k_wdot_mixedE'),
                 ('f64dydt', 'k_wdot_dydtE'),
                 ('mixeddydt', 'k_wdot_mixed_dydtE'),
                 ('dydt', 'k_dydt_f64E'))   # the last with dydt='fused'
GROUP_CANDIDATES = (32,) + tuple(range(36, 61)) + (64,)
_DEFAULT_GROUP = 32


def _fused_usage(usage):
    """(scratch instructions, scratch bytes per lane, spilled VGPRs) of each
    fused kernel (instructions 0 where the report has none)."""
    out = {}
    for short, tag in FUSED_KERNELS:
        for fn, u in usage.items():
            if tag in fn:
                out[short] = (u.get('Scratch ops', 0),
                              u.get('ScratchSize [bytes/lane]', 0),
                              u.get('VGPRs Spill', 0))
    return out


def _groups_key(src_text, extra_flags, candidates):
    """Hash of everything the group-size choice depends on: the emitted
    source, the device headers it includes, the flags and the candidates."""
    h = hashlib.sha256(src_text.encode())
    for d in INCLUDE_DIRS:
        for f in sorted(os.listdir(d)):
            if f.endswith('.h'):
                with open(os.path.join(d, f), 'rb') as fh:
                    h.update(fh.read())
    h.update(repr((HIPCC_FLAGS, list(extra_flags), list(candidates)))
             .encode())
    return h.hexdigest()


def build_with_groups(write, out_dir, extra_flags=(), verbose=False,
                      candidates=GROUP_CANDIDATES):
    """Emit and compile a mechanism library, choosing per fused kernel the
    number of reactions per basic block (emitter options split_f64,
    split_mixed, split_f64dydt, split_mixeddydt, split_dydt).

    ``write(opts, out_dir)`` emits the source with extra emitter options
    and returns its path.  The library is first built with 32 reactions
    per group.  Where a fused kernel spills, every candidate size (32,
    36-60, 64) is compiled (device code only, in a temporary directory) and
    each kernel takes the size with the fewest scratch instructions (in
    these straight-line kernels also the number each wave executes), then
    the least scratch per lane, the fewest spilled VGPRs and the size
    nearest 32.  Where the register allocator spills depends on the group
    boundaries, and every spilled value costs HBM traffic and a reload
    wait: k_wdot_f64 at 53 species has 124 B/lane of scratch, 56 scratch
    instructions and runs 0.650 ms at 32 reactions per group; 76 B/lane,
    25 instructions and 0.630 ms at 56; 44 B/lane and 15 instructions at
    47 (the count does not vary smoothly with the size).  The choice, every
    candidate's figures and a hash of the inputs are written to groups.json
    beside the library; a later build of the same source, headers and
    flags takes the recorded choice instead of compiling the candidates
    again (the library itself is always compiled).
    """
    usage = {}
    src = write({}, out_dir)
    with open(src) as fh:
        key = _groups_key(fh.read(), extra_flags, candidates)
    record = os.path.join(out_dir, 'groups.json')
    if os.path.exists(record):
        with open(record) as fh:
            old = json.load(fh)
        if old.get('key') == key:
            choice = old['choice']
            if any(v != _DEFAULT_GROUP for v in choice.values()):
                src = write({'split_' + k: v for k, v in choice.items()},
                            out_dir)
            return compile_library(src, out_dir, extra_flags, verbose)
    lib = compile_library(src, out_dir, extra_flags, verbose,
                          usage_out=usage)
    base = _fused_usage(usage)
    table = {}
    choice = {k: _DEFAULT_GROUP for k in base}
    if any(u[1] for u in base.values()):
        with tempfile.TemporaryDirectory() as tmp:
            # candidates whose source is the same text (a layout whose
            # reaction blocks do not group) are compiled once
            srcs, texts = {}, {}
            for c in candidates:
                path = write({'split_blocks': c}, os.path.join(tmp, str(c)))
                with open(path) as fh:
                    srcs[c] = texts.setdefault(fh.read(), path)
            with ThreadPoolExecutor(min(len(texts),
                                        os.cpu_count() or 4)) as pool:
                done = dict(zip(texts.values(), pool.map(
                    lambda p: _fused_usage(device_usage(p, extra_flags)),
                    texts.values())))
            table = {c: done[srcs[c]] for c in candidates}
        for k in base:
            choice[k] = min(table, key=lambda c: (table[c][k],
                                                  abs(c - _DEFAULT_GROUP), c))
        if any(v != _DEFAULT_GROUP for v in choice.values()):
            src = write({'split_' + k: v for k, v in choice.items()},
                        out_dir)
            lib = compile_library(src, out_dir, extra_flags, verbose)
    with open(record, 'w') as fh:
        json.dump(dict(key=key, choice=choice, candidates={
            str(c): {k: dict(scratch_ops=v[0], scratch_bytes_per_lane=v[1],
                             spilled_vgprs=v[2]) for k, v in u.items()}
            for c, u in sorted(table.items())}), fh, indent=1)
    return lib


MLP_SRC = os.path.join(HERE, 'csrc', 'mlp.hip')


def build_mlp(verbose=False):
    """Compile the mechanism-independent surrogate library."""
    out_dir = workloads.BUILD_ROOT
    os.makedirs(out_dir, exist_ok=True)
    lib = os.path.join(out_dir, 'libcrs_mlp.so')
    cmd = [hipcc()] + HIPCC_FLAGS
    for d in INCLUDE_DIRS:
        cmd += ['-I', d]
    cmd += ['-o', lib + '.tmp', MLP_SRC]
    t0 = time.time()
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError('hipcc failed for mlp.hip:\n' + res.stderr[-4000:])
    os.replace(lib + '.tmp', lib)
    if verbose:
        print('built {} in {:.1f} s'.format(lib, time.time() - t0))
    return lib


PROBE_SRC = os.path.join(HERE, 'csrc', 'fma_probe.hip')
PROBE_LIB = os.path.join(workloads.BUILD_ROOT, 'libcrs_probe.so')


def build_probe(verbose=False):
    """Compile the ceiling probes (tune.ceiling, tune.bandwidth)."""
    os.makedirs(workloads.BUILD_ROOT, exist_ok=True)
    cmd = [hipcc()] + HIPCC_FLAGS + ['-o', PROBE_LIB + '.tmp', PROBE_SRC]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError('hipcc failed for fma_probe.hip:\n' +
                           res.stderr[-4000:])
    os.replace(PROBE_LIB + '.tmp', PROBE_LIB)
    if verbose:
        print('built', PROBE_LIB)
    return PROBE_LIB


def lib_path(name):
    return os.path.join(workloads.BUILD_ROOT, name, 'libcrs_{}.so'.format(name))


def build_mechanism(name, verbose=False, extra_flags=()):
    """Emit and compile synthetic mechanism ``name``; return the .so path."""
    mpath, tpath = workloads.mechanism_files(name)
    _e, specs, reacs, fwd = load_mechanism(
        mpath, tpath, last_spec=workloads.last_species(name))
    out_dir = os.path.join(workloads.BUILD_ROOT, name)

    def write(opts, d):
        return hip_emitter.write_hip(d, specs, reacs, name=name,
                                     spec_order=fwd, opts=opts)
    return build_with_groups(write, out_dir, extra_flags=extra_flags,
                             verbose=verbose)
