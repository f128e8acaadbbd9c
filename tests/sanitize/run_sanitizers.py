"""Build and run the host sanitizer checks (test infrastructure).

Two programs, each built with AddressSanitizer and UndefinedBehaviorSanitizer
and run on the CPU (no GPU needed):

* ``oracle_san``: the CPU oracle of one mechanism (its emitted C and
  oracle/driver.c, gcc -fsanitize=address,undefined -fopenmp) driven by
  ``oracle_san.c`` (every oracle entry point, canary guard bands around and
  inside every buffer);
* ``abi_san``: the C ABI shims of the mechanism library and the surrogate
  library (hipcc with ``-Xarch_host -fsanitize=address,undefined``: the
  host code instrumented, the device code as in the product build; no call
  made here reaches a kernel) driven by ``abi_san.cpp`` (argument
  validation, error codes and messages of every entry point of
  include/crs_hip.h and include/crs_mlp.h, per-thread and per-library
  error text).

Outputs go to ``tests/sanitize/_build/<mech>/``.  Run:
``python tests/sanitize/run_sanitizers.py [mech ...]`` (default mech13 and
mech19fx); exit status 0 when every program prints "ok".
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from create_rate_subs_amd import build as pkg_build, hip_emitter, workloads  # noqa: E402
from create_rate_subs_amd.create_rate_subs import load_mechanism  # noqa: E402
from oracle import emit_c  # noqa: E402

SAN = ['-fsanitize=address,undefined', '-fno-sanitize-recover=all',
       '-fno-omit-frame-pointer', '-g']
ROCM = os.environ.get('ROCM_PATH', '/opt/rocm')


def out_dir(name):
    d = os.path.join(HERE, '_build', name)
    os.makedirs(d, exist_ok=True)
    return d


def _run(cmd, what):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError('{} failed ({}):\n{}\n{}'.format(
            what, ' '.join(cmd[:3]), res.stdout[-2000:], res.stderr[-4000:]))
    return res


def _mechanism(name):
    mpath, tpath = workloads.mechanism_files(name)
    return load_mechanism(mpath, tpath, last_spec=workloads.last_species(name))


def build_oracle_san(name, plant=None):
    """gcc the mechanism's oracle C with the sanitizers; return the exe.
    plant: 'PAD_WRITE' or 'OOB_WRITE' builds a copy with that mistake
    planted in the driver (oracle_san.c)."""
    d = out_dir(name)
    _e, specs, reacs, fwd = _mechanism(name)
    srcs = emit_c.emit_c(specs, reacs, os.path.join(d, 'oracle_src'), fwd)
    exe = os.path.join(d, 'oracle_san' + ('_' + plant.lower() if plant
                                           else ''))
    defs = ['-DPLANT_' + plant] if plant else []
    cmd = (['gcc', '-O1', '-fopenmp', '-std=gnu99', '-ffp-contract=off'] +
           SAN + defs + ['-I', os.path.join(d, 'oracle_src'), '-o', exe] +
           srcs +
           [os.path.join(ROOT, 'oracle', 'driver.c'),
            os.path.join(HERE, 'oracle_san.c'), '-lm'])
    _run(cmd, 'oracle sanitizer build')
    return exe


def _host_flags():
    # the library flags, then the sanitizers for the host pass only
    flags = list(pkg_build.HIPCC_FLAGS)
    for f in SAN:
        flags += ['-Xarch_host', f]
    for inc in pkg_build.INCLUDE_DIRS:
        flags += ['-I', inc]
    return flags


def build_abi_san(name):
    """Host-only sanitized builds of libcrs_<name>.so and libcrs_mlp.so and
    the abi_san driver linked against them; return the exe."""
    d = out_dir(name)
    _e, specs, reacs, fwd = _mechanism(name)
    src = hip_emitter.write_hip(d, specs, reacs, name=name, spec_order=fwd)
    hipcc = pkg_build.hipcc()
    lib = os.path.join(d, 'libcrs_{}.so'.format(name))
    _run([hipcc] + _host_flags() + ['-o', lib, src], 'hipcc ' + name)
    mlp = os.path.join(d, 'libcrs_mlp.so')
    _run([hipcc] + _host_flags() + ['-o', mlp, pkg_build.MLP_SRC],
         'hipcc mlp')
    exe = os.path.join(d, 'abi_san')
    clang = os.path.join(ROCM, 'llvm', 'bin', 'clang++')
    _run([clang, '-std=c++17', '-O1'] + SAN +
         ['-I', os.path.join(ROOT, 'include'), '-I',
          os.path.join(ROOT, 'create_rate_subs_amd', 'csrc'), '-o', exe,
          os.path.join(HERE, 'abi_san.cpp'), '-L', d,
          '-lcrs_{}'.format(name), '-lcrs_mlp', '-pthread',
          '-Wl,-rpath,' + d, '-Wl,-rpath,' + os.path.join(ROCM, 'lib')],
         'abi_san link')
    return exe


def run(exe, *args):
    """Run a sanitizer program; (returncode, stdout + stderr)."""
    env = dict(os.environ,
               ASAN_OPTIONS='detect_leaks=1:abort_on_error=0:exitcode=23',
               UBSAN_OPTIONS='print_stacktrace=1:halt_on_error=1:exitcode=24',
               OMP_NUM_THREADS='2')
    res = subprocess.run([exe] + [str(a) for a in args], capture_output=True,
                         text=True, env=env, timeout=300)
    return res.returncode, res.stdout + res.stderr


def main(names):
    status = 0
    for name in names:
        for build, args in ((build_oracle_san, (37, 2)),
                            (build_abi_san, ())):
            exe = build(name)
            rc, out = run(exe, *args)
            print('{} {}: exit {}\n{}'.format(name, os.path.basename(exe),
                                              rc, out.strip()[-3000:]))
            status |= rc != 0 or not out.rstrip().endswith('ok')
    return int(status)


if __name__ == '__main__':
    sys.exit(main(sys.argv[1:] or ['mech13', 'mech19fx']))
