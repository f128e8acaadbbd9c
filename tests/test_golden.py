"""Known-answer vectors (tests/golden/wdot_kat.json, written by
tests/golden/make_wdot_kat.py): wdot and the constant-pressure dydt of five
mechanisms that hold every rate form of the path, eight seeded states
each, evaluated in mpmath at 50 digits from the closed forms.  The C oracle
(CPU) and the HIP kernels (GPU) must reproduce them under the conditioned
error metric.
"""

import json
import os
import re
import subprocess

import numpy as np
import pytest

from create_rate_subs_amd import metrics
from oracle import emit_c, numpy_ref
from oracle.build import CFLAGS, HERE as ORACLE_DIR, Oracle

from _helpers import hip_lib, mechanism, oracle, to_dev, to_host

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, 'golden', 'wdot_kat.json')) as _fh:
    KAT = json.load(_fh)['mechanisms']
NAMES = sorted(KAT)
# Conditioned-error bound against the exact closed forms.  Chebyshev
# reactions are evaluated, as the reference emits them, with the
# reduced-coordinate constants printed to 9 significant digits
# (rate_subs.py:138-223, {:.8e}), which moves their rates by up to ~1e-8
# from the unrounded textbook value; every other form agrees to rounding.
TOL = {'mech13': 1e-11, 'mech13co': 1e-11, 'mech17tm': 1e-11,
       'mech19fx': 1e-11, 'mech21pc': 2e-8}


def _case(name):
    d = KAT[name]
    return (np.array(d['T']), np.array(d['P']), np.array(d['Y']),
            np.array(d['wdot']), np.array(d['S']))


def _dydt_case(name):
    """y in the reference's layout (row 0 = T, rows 1..NSP-1 = Y), P, the
    expected dydt and its scale."""
    d = KAT[name]
    T, P, Y = (np.array(d[k]) for k in ('T', 'P', 'Y'))
    y = np.empty((d['nsp'], T.size))
    y[0] = T
    y[1:] = Y[:d['nsp'] - 1]
    return y, P, np.array(d['dydt']), np.array(d['dydt_scale'])


@pytest.mark.parametrize('name', NAMES)
def test_fixture_matches_its_generator(name):
    """The committed vectors are what mpmath gives for their states (the
    first state of each mechanism, recomputed)."""
    specs, reacs, _ = mechanism(name)
    T, P, Y, W, _S = _case(name)
    exact = np.array([float(v) for v in numpy_ref.mp_state(
        specs, reacs, T[0], P[0], Y[:, 0])])
    scale = np.abs(W[:, 0]).max()
    assert np.abs(exact - W[:, 0]).max() <= 1e-15 * scale


@pytest.mark.parametrize('name', NAMES)
def test_oracle_matches_golden(name):
    T, P, Y, W, S = _case(name)
    got = oracle(name).eval_stages(T, P, Y)['wdot']
    e = metrics.rel_err(got, W, S, 1e-3)
    assert e.max() <= TOL[name], e.max()


@pytest.mark.parametrize('name', NAMES)
def test_oracle_dydt_matches_golden(name):
    y, P, D, DS = _dydt_case(name)
    got = oracle(name).dydt(P, y)
    e = metrics.rel_err(got, D, DS, 1e-3)
    assert e.max() <= TOL[name], e.max()


@pytest.mark.gpu
@pytest.mark.parametrize('name', NAMES)
def test_hip_dydt_matches_golden(name):
    y, P, D, DS = _dydt_case(name)
    Pd, yd = to_dev(P, y)
    got = to_host(hip_lib(name).dydt(Pd, yd))
    e = metrics.rel_err(got, D, DS, 1e-3)
    print(name, 'dydt vs golden', e.max())
    assert e.max() <= TOL[name], e.max()


@pytest.mark.gpu
@pytest.mark.parametrize('name', NAMES)
def test_hip_matches_golden(name):
    T, P, Y, W, S = _case(name)
    lib = hip_lib(name)
    Td, Pd, Yd = to_dev(T, P, Y)
    got = to_host(lib.eval_wdot(Td, Pd, Yd, 'fp64'))
    e = metrics.rel_err(got, W, S, 1e-3)
    print(name, 'fp64 vs golden', e.max())
    assert e.max() <= TOL[name], e.max()
    if lib.modes & 2:
        got = to_host(lib.eval_wdot(Td, Pd, Yd, 'mixed'))
        e = metrics.rel_err(got, W, S, 0.05)
        print(name, 'mixed vs golden', e.max())
        assert np.isfinite(got).all() and e.max() <= 1e-4, e.max()


# mistakes planted in a copy of the oracle's C for mech19fx, each in a
# branch only these vectors reach: (file, pattern, replacement)
MUTATIONS = {
    # SRI's d T^e factor dropped
    'sri_dte_dropped': ('rxn_rates_pres_mod.c',
                        r'\* \S+ \* pow\(T, \S+\) ', ''),
    # a half-order concentration taken to the first power
    'half_order_as_one': ('rxn_rates.c', r'pow\((C\[\d+\]), 0\.5\)',
                          r'\1'),
}


@pytest.mark.parametrize('mutation', sorted(MUTATIONS))
def test_golden_vectors_catch_planted_mistake(tmp_path, mutation):
    name = 'mech19fx'
    specs, reacs, fwd = mechanism(name)
    src = tmp_path / 'src'
    cfiles = emit_c.emit_c(specs, reacs, str(src), fwd)
    fname, pat, repl = MUTATIONS[mutation]
    path = src / fname
    text, count = re.subn(pat, repl, path.read_text())
    assert count > 0, mutation
    path.write_text(text)
    so = str(tmp_path / 'liboracle.so')
    subprocess.run(['gcc'] + CFLAGS + ['-I', str(src), '-o', so] + cfiles +
                   [os.path.join(ORACLE_DIR, 'driver.c'), '-lm'], check=True)
    T, P, Y, W, S = _case(name)
    got = Oracle(name, so).eval_stages(T, P, Y)['wdot']
    e = metrics.rel_err(got, W, S, 1e-3)
    assert e.max() > 1e-4, (mutation, e.max())
