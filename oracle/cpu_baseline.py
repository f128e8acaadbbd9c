"""CPU baseline of the wdot path: the oracle on the host cores.

Test and measurement infrastructure (run by bench.py as a child process, so
that the OpenMP placement variables below reach libgomp before it starts,
and no GPU runtime threads share the cores).  Protocol: the restated
generated C (oracle/emit_c.py + driver.c, gcc -O3 -fopenmp
-ffp-contract=off, FP64), `#pragma omp parallel for schedule(static)` over
states, OMP_PROC_BIND=close and OMP_PLACES=cores, one warm-up pass, then
the median of five passes timed with omp_get_wtime.

    python -m oracle.cpu_baseline <workload> <states> <threads>

prints one JSON object: states/s, the thread count, the host's CPU model,
physical core count and logical CPUs, the CPUs this process may run on,
the compiler and its flags, the FMA-contraction mode and the glibc version.
"""

import json
import os
import platform
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

REPS = 5


def host_info():
    model, cores = None, set()
    phys = core = None
    try:
        with open('/proc/cpuinfo') as fh:
            for line in fh:
                k, _, v = line.partition(':')
                k, v = k.strip(), v.strip()
                if k == 'model name' and model is None:
                    model = v
                elif k == 'physical id':
                    phys = v
                elif k == 'core id':
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    try:
        gcc = subprocess.run(['gcc', '--version'], capture_output=True,
                             text=True).stdout.splitlines()[0]
    except (OSError, IndexError):
        gcc = None
    libc = ' '.join(platform.libc_ver()) or None
    try:
        libc = os.confstr('CS_GNU_LIBC_VERSION')
    except (ValueError, OSError, AttributeError):
        pass
    return dict(cpu_model=model or platform.processor() or platform.machine(),
                physical_cores=len(cores) or None,
                logical_cpus=os.cpu_count(),
                cpus_allowed=len(os.sched_getaffinity(0)),
                compiler=gcc, glibc=libc)


def main(argv):
    workload, m, threads = argv[0], int(argv[1]), int(argv[2])
    # before libgomp starts: with OMP_PROC_BIND it pins this thread to one
    # place, after which the affinity mask shows a single CPU
    info = host_info()
    import numpy as np
    from create_rate_subs_amd import states, workloads
    from oracle.build import CFLAGS, Oracle, build, lib_path
    wl = workloads.WORKLOADS[workload]
    if not os.path.exists(lib_path(wl.mech)):
        build(wl.mech)
    o = Oracle(wl.mech)
    # the first m states of the workload's seeded batch
    T, P, Y = states.random_states(o.nsp, wl.n_states, wl.state_seed)
    m = min(m, wl.n_states)
    T, P, Y = T[:m].copy(), P[:m].copy(), np.ascontiguousarray(Y[:, :m])
    times = o.time_wdot(T, P, Y, threads, REPS)
    t = float(np.median(times))
    out = dict(value=m / t, unit='states/s', cores=threads, kind='port',
               seconds_per_pass=t, passes=[float(x) for x in times],
               sample='{} of the {} states of {} ({} sp), {} OpenMP threads '
                      '(OMP_PROC_BIND={}, OMP_PLACES={}), median of {} '
                      'passes after 1 warm-up, {:.3f} s each ({:.1f} s of '
                      'CPU work timed)'.format(
                          m, wl.n_states, workload, o.nsp, threads,
                          os.environ.get('OMP_PROC_BIND'),
                          os.environ.get('OMP_PLACES'), REPS, t,
                          REPS * t * threads),
               cflags=' '.join(CFLAGS), fp_contract='off', timer='omp_get_wtime')
    out.update(info)
    print(json.dumps(out))


if __name__ == '__main__':
    main(sys.argv[1:])
