"""Build and load the CPU oracle (test infrastructure only).

For a mechanism: parse it with the package's parser, emit the reference's
C subroutines with ``oracle/emit_c.py`` and compile them together with
``oracle/driver.c`` into ``oracle/_build/<name>/liboracle.so`` with
``gcc -O3 -fopenmp`` in FP64 (no fast-math, no FMA: the default x86-64
target has none, so every operation rounds as written).  ``Oracle`` wraps
the library with ctypes.

Only tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may use this module.
"""

import ctypes
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from create_rate_subs_amd import workloads                      # noqa: E402
from create_rate_subs_amd.create_rate_subs import load_mechanism  # noqa: E402
from oracle import emit_c                                        # noqa: E402

BUILD_ROOT = os.path.join(HERE, '_build')
# -ffp-contract=off: no fused multiply-add anywhere (the default x86-64
# target has no FMA instruction either), so every operation rounds as the
# reference's C writes it
CFLAGS = ['-O3', '-fopenmp', '-fPIC', '-shared', '-std=gnu99',
          '-ffp-contract=off']


def lib_path(name):
    return os.path.join(BUILD_ROOT, name, 'liboracle.so')


def build(name, mech_path=None, therm_path=None, force=True):
    """Emit and compile the oracle for mechanism ``name``; return the .so path.

    Without explicit paths ``name`` must be one of the synthetic mechanisms
    (``workloads.mechanism_files``).
    """
    if mech_path is None:
        mech_path, therm_path = workloads.mechanism_files(name)
    out = os.path.join(BUILD_ROOT, name)
    so = os.path.join(out, 'liboracle.so')
    if not force and os.path.exists(so):
        return so
    _e, specs, reacs, fwd = load_mechanism(
        mech_path, therm_path, last_spec=workloads.last_species(name))
    srcs = emit_c.emit_c(specs, reacs, os.path.join(out, 'src'), fwd)
    cmd = (['gcc'] + CFLAGS + ['-I', os.path.join(out, 'src'), '-o',
                                so + '.tmp'] +
           srcs + [os.path.join(HERE, 'driver.c'), '-lm'])
    subprocess.run(cmd, check=True)
    os.replace(so + '.tmp', so)
    return so


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _arr(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle(object):
    """ctypes view of one mechanism's oracle library."""

    def __init__(self, name, path=None):
        self.name = name
        self.path = path or lib_path(name)
        if not os.path.exists(self.path):
            raise FileNotFoundError('oracle library not built: ' + self.path)
        self.lib = ctypes.CDLL(self.path)
        dims = (ctypes.c_int * 4)()
        self.lib.oracle_dims(dims)
        self.nsp, self.nr, self.nrev, self.npd = (int(x) for x in dims)
        P = ctypes.POINTER(ctypes.c_double)
        i64 = ctypes.c_int64
        self.lib.oracle_eval_wdot.argtypes = [P, P, P, P, i64, i64,
                                              ctypes.c_int]
        self.lib.oracle_eval_stages.argtypes = ([P] * 11 +
                                                [i64, i64, ctypes.c_int])
        self.lib.oracle_spec_rates.argtypes = [P, P, P, P, i64, i64,
                                               ctypes.c_int]
        self.lib.oracle_dydt.argtypes = [P, P, P, i64, i64, ctypes.c_int]
        self.lib.oracle_thermo.argtypes = [P, P, P, i64, i64, ctypes.c_int]
        self.lib.oracle_time_wdot.argtypes = [P, P, P, P, i64, i64,
                                              ctypes.c_int, ctypes.c_int, P]

    @staticmethod
    def default_threads():
        return int(os.environ.get('OMP_NUM_THREADS', os.cpu_count() or 1))

    def eval_wdot(self, T, P, Y, nthreads=None):
        T, P, Y = _arr(T), _arr(P), _arr(Y)
        n = T.size
        ld = Y.shape[1]
        W = np.empty((self.nsp, ld))
        self.lib.oracle_eval_wdot(_dptr(T), _dptr(P), _dptr(Y), _dptr(W), n,
                                  ld, nthreads or self.default_threads())
        return W

    def time_wdot(self, T, P, Y, nthreads, reps=5):
        """Wall seconds of ``reps`` timed passes over all states after one
        warm-up pass (omp_get_wtime inside the library)."""
        T, P, Y = _arr(T), _arr(P), _arr(Y)
        W = np.empty((self.nsp, Y.shape[1]))
        times = np.zeros(reps)
        self.lib.oracle_time_wdot(_dptr(T), _dptr(P), _dptr(Y), _dptr(W),
                                  T.size, Y.shape[1], nthreads, reps,
                                  _dptr(times))
        return times

    def eval_stages(self, T, P, Y, nthreads=None):
        """Every intermediate: dict of y_N, mw, rho, C, fwd, rev, pm, wdot."""
        T, P, Y = _arr(T), _arr(P), _arr(Y)
        n = T.size
        ld = Y.shape[1]
        out = dict(y_N=np.empty(n), mw=np.empty(n), rho=np.empty(n),
                   C=np.empty((self.nsp, ld)), fwd=np.empty((self.nr, ld)),
                   rev=np.empty((max(self.nrev, 1), ld)),
                   pm=np.empty((max(self.npd, 1), ld)),
                   wdot=np.empty((self.nsp, ld)))
        self.lib.oracle_eval_stages(
            _dptr(T), _dptr(P), _dptr(Y), _dptr(out['y_N']), _dptr(out['mw']),
            _dptr(out['rho']), _dptr(out['C']), _dptr(out['fwd']),
            _dptr(out['rev']), _dptr(out['pm']), _dptr(out['wdot']), n, ld,
            nthreads or self.default_threads())
        out['rev'] = out['rev'][:self.nrev]
        out['pm'] = out['pm'][:self.npd]
        return out

    def spec_rates(self, fwd, rev, pm, nthreads=None):
        fwd = _arr(fwd)
        ld = fwd.shape[1]
        rev = _arr(rev) if self.nrev else np.zeros((1, ld))
        pm = _arr(pm) if self.npd else np.zeros((1, ld))
        W = np.empty((self.nsp, ld))
        self.lib.oracle_spec_rates(_dptr(fwd), _dptr(rev), _dptr(pm),
                                   _dptr(W), ld, ld,
                                   nthreads or self.default_threads())
        return W


    def dydt(self, P, y, nthreads=None):
        """Constant-pressure RHS: y [>= NSP][ld] (row 0 T, rows 1..NSP-1
        mass fractions) -> dy [NSP][ld]."""
        P, y = _arr(P), _arr(y)
        n = P.size
        ld = y.shape[1]
        dy = np.empty((self.nsp, ld))
        self.lib.oracle_dydt(_dptr(P), _dptr(y), _dptr(dy), n, ld,
                             nthreads or self.default_threads())
        return dy

    def thermo(self, T, nthreads=None):
        """(cp [NSP][n] J/kg/K, h [NSP][n] J/kg) at temperatures T."""
        T = _arr(T)
        n = T.size
        cp = np.empty((self.nsp, n))
        h = np.empty((self.nsp, n))
        self.lib.oracle_thermo(_dptr(T), _dptr(cp), _dptr(h), n, n,
                               nthreads or self.default_threads())
        return cp, h


def build_all(force=True):
    return [build(name, force=force) for name in workloads.LIBRARIES]


if __name__ == '__main__':
    names = sys.argv[1:] or list(workloads.LIBRARIES)
    for nm in names:
        print(build(nm))
