"""Compile emitted HIP sources into mechanism libraries for gfx950.

``compile_library(src, out_dir)`` runs hipcc on one ``crs_<name>.hip`` and
writes ``libcrs_<name>.so`` next to it.  ``build_mechanism(name)`` emits and
compiles one of the synthetic mechanisms into ``_build/<name>/``.
"""

import os
import subprocess
import time

from . import hip_emitter, workloads
from .create_rate_subs import load_mechanism

__all__ = ['HIPCC_FLAGS', 'compile_library', 'build_mechanism', 'lib_path',
           'INCLUDE_DIRS', 'build_mlp']

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


def compile_library(src, out_dir=None, extra_flags=(), verbose=False):
    """hipcc ``src`` -> ``libcrs_<name>.so``; return the library path."""
    out_dir = out_dir or os.path.dirname(src)
    base = os.path.splitext(os.path.basename(src))[0]
    lib = os.path.join(out_dir, 'lib{}.so'.format(base))
    cmd = [hipcc()] + HIPCC_FLAGS + list(extra_flags)
    for d in INCLUDE_DIRS:
        cmd += ['-I', d]
    cmd += ['-o', lib + '.tmp', src]
    t0 = time.time()
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError('hipcc failed for {}:\n{}'.format(
            src, res.stderr[-4000:]))
    os.replace(lib + '.tmp', lib)
    if verbose:
        print('built {} in {:.1f} s'.format(lib, time.time() - t0))
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


def lib_path(name):
    return os.path.join(workloads.BUILD_ROOT, name, 'libcrs_{}.so'.format(name))


def build_mechanism(name, verbose=False, extra_flags=()):
    """Emit and compile synthetic mechanism ``name``; return the .so path."""
    mpath, tpath = workloads.mechanism_files(name)
    _e, specs, reacs, fwd = load_mechanism(
        mpath, tpath, last_spec=workloads.last_species(name))
    out_dir = os.path.join(workloads.BUILD_ROOT, name)
    src = hip_emitter.write_hip(out_dir, specs, reacs, name=name,
                                spec_order=fwd)
    return compile_library(src, out_dir, extra_flags=extra_flags,
                           verbose=verbose)
