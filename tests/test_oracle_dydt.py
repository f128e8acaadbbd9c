"""Pins of the oracle's constant-pressure dydt, eval_h and eval_cp (CPU).

The reference ships no test for these either.  What fixes them:

* h against the Gibbs function: d(s/R - h/RT)/dT = h W / (RU T^2), with
  s/R - h/RT from chem.calc_spec_smh (chem_utilities.py:251-293, its own
  range selection), so a swapped NASA range or a wrong polynomial
  coefficient in eval_h shows;
* cp = dh/dT (central differences), so eval_cp must be the derivative of
  the h just pinned;
* species rows: dY_k/dt = wdot_k W_k / rho with wdot and rho from the
  oracle's eval_stages (pinned in test_oracle_pins.py);
* temperature row through enthalpy conservation at constant pressure:
  cp_avg dT/dt + sum_k h_k dY_k/dt = 0 over all species, the implied last
  one included (dY_N/dt = wdot_N W_N / rho), with cp_avg = sum_k cp_k Y_k
  formed here from the pinned cp;

and each check is shown to catch a mistake planted in a copy of the oracle.
"""

import os
import re
import subprocess

import numpy as np
import pytest

from create_rate_subs_amd import chem, states, workloads
from create_rate_subs_amd.create_rate_subs import load_mechanism
from oracle import emit_c
from oracle.build import CFLAGS, HERE as ORACLE_DIR, Oracle

from _helpers import mechanism, oracle

MECHS = ['mech13', 'mech53', 'mech111', 'mech17tm', 'mech19fx']


def _state_vector(specs, n, seed):
    """(P, y) in the reference's dydt layout: y[0] = T, y[1+k] = Y_k."""
    T, P, Y = states.random_states(len(specs), n, seed)
    y = np.empty((len(specs), n))
    y[0] = T
    y[1:] = Y[:len(specs) - 1]
    return T, P, Y, y


def _temperatures(specs, n=64):
    """Temperatures on both sides of every T_mid, away from it."""
    lo = min(sp.Trange[0] for sp in specs)
    hi = max(sp.Trange[2] for sp in specs)
    mids = sorted({sp.Trange[1] for sp in specs})
    T = np.linspace(lo + 10.0, hi - 10.0, n)
    keep = np.ones(n, bool)
    for tm in mids:
        keep &= np.abs(T - tm) > 5.0
    return T[keep]


@pytest.mark.parametrize('name', MECHS)
def test_h_is_the_gibbs_function_derivative(name):
    specs, _r, _f = mechanism(name)
    T = _temperatures(specs)
    _cp, h = oracle(name).thermo(T)
    d = 1e-3
    for k, sp in enumerate(specs):
        up = np.array([chem.calc_spec_smh(t + d, [sp])[0] for t in T])
        dn = np.array([chem.calc_spec_smh(t - d, [sp])[0] for t in T])
        dsmh = (up - dn) / (2 * d)
        want = h[k] * sp.mw / (chem.RU * T * T)
        err = np.abs(dsmh - want) / np.maximum(np.abs(want), 1e-3)
        assert err.max() < 1e-6, (name, k, err.max())


@pytest.mark.parametrize('name', MECHS)
def test_cp_is_dh_dT(name):
    specs, _r, _f = mechanism(name)
    T = _temperatures(specs)
    o = oracle(name)
    cp, _h = o.thermo(T)
    d = 1e-2
    _c, hu = o.thermo(T + d)
    _c, hd = o.thermo(T - d)
    err = np.abs((hu - hd) / (2 * d) - cp) / np.abs(cp)
    assert err.max() < 1e-7, err.max()


def _check_dydt(o, specs, n=512, seed=90):
    """Max violations of the species-row and energy pins."""
    T, P, Y, y = _state_vector(specs, n, seed)
    dy = o.dydt(P, y)
    st = o.eval_stages(T, P, Y)
    mw = np.array([sp.mw for sp in specs])
    rows = st['wdot'][:-1] * mw[:-1, None] / st['rho']
    scale = np.maximum(np.abs(rows), 1e-300)
    species_err = (np.abs(dy[1:] - rows) / scale).max()
    cp, h = o.thermo(T)
    Yall = np.vstack([Y[:-1], 1.0 - Y[:-1].sum(0)])
    cp_avg = (cp * Yall).sum(0)
    dY = np.vstack([dy[1:], st['wdot'][-1:] * mw[-1] / st['rho']])
    flux = h * dY
    resid = cp_avg * dy[0] + flux.sum(0)
    energy_err = (np.abs(resid) / np.abs(flux).sum(0)).max()
    return species_err, energy_err


@pytest.mark.parametrize('name', MECHS)
def test_dydt_species_rows_and_energy_balance(name):
    specs, _r, _f = mechanism(name)
    species_err, energy_err = _check_dydt(oracle(name), specs)
    print(name, 'species rows', species_err, 'energy balance', energy_err)
    assert species_err <= 4e-16 * 4
    assert energy_err <= 1e-13


MUTATIONS = {
    # energy equation with the wrong sign
    'energy_sign': ('dydt.c', r'\(-1\.0 / \(rho \* cp_avg\)\)',
                    '(1.0 / (rho * cp_avg))'),
    # the implied species' term dropped from the energy equation
    'dyN_dropped': ('dydt.c', r' \+ \(dy_N \* h\[\d+\] \* [0-9.e+-]+\)', ''),
    # lower and upper NASA ranges swapped in eval_h and eval_cp
    'thermo_range_swap': ('thermo.c', r'if \(T <= ', 'if (T > '),
    # mass-fraction rates multiplied by rho instead of divided
    'mass_rate': ('dydt.c', r' / rho\);', ' * rho);'),
}


# The synthetic mechanisms' bath gas (the implied last species) is inert;
# to exercise the dy_N terms the checks below also run with a reactive
# species moved last (the reference's -ls option).
REACTIVE_LAST = ('mech13', 'CO')


def _reordered(name, last):
    mpath, tpath = workloads.mechanism_files(name)
    _e, specs, reacs, fwd = load_mechanism(mpath, tpath, last_spec=last)
    assert specs[-1].name == last
    assert any(len(specs) - 1 in r.reac + r.prod for r in reacs)
    return specs, reacs, fwd


def _build(tmp_path, name, last, mutation=None):
    specs, reacs, fwd = _reordered(name, last)
    src = tmp_path / 'src'
    cfiles = emit_c.emit_c(specs, reacs, str(src), fwd)
    if mutation:
        fname, pat, rep = MUTATIONS[mutation]
        p = src / fname
        text, count = re.subn(pat, rep, p.read_text())
        assert count > 0, mutation
        p.write_text(text)
    so = str(tmp_path / 'liboracle.so')
    subprocess.run(['gcc'] + CFLAGS + ['-I', str(src), '-o', so] + cfiles +
                   [os.path.join(ORACLE_DIR, 'driver.c'), '-lm'], check=True)
    return specs, Oracle(name, so)


def test_dydt_with_reactive_last_species(tmp_path):
    specs, o = _build(tmp_path, *REACTIVE_LAST)
    species_err, energy_err = _check_dydt(o, specs)
    print('reactive last species', species_err, energy_err)
    assert species_err <= 4e-16 * 4
    assert energy_err <= 1e-13


@pytest.mark.parametrize('mutation', sorted(MUTATIONS))
def test_pins_catch_planted_mistake(tmp_path, mutation):
    specs, o = _build(tmp_path, *REACTIVE_LAST, mutation=mutation)
    species_err, energy_err = _check_dydt(o, specs, n=128)
    T = _temperatures(specs, 16)
    _cp, h = o.thermo(T)
    d = 1e-3
    gibbs_err = 0.0
    for k, sp in enumerate(specs):
        up = np.array([chem.calc_spec_smh(t + d, [sp])[0] for t in T])
        dn = np.array([chem.calc_spec_smh(t - d, [sp])[0] for t in T])
        want = h[k] * sp.mw / (chem.RU * T * T)
        gibbs_err = max(gibbs_err, float(
            (np.abs((up - dn) / (2 * d) - want) /
             np.maximum(np.abs(want), 1e-3)).max()))
    caught = species_err > 1e-6 or energy_err > 1e-6 or gibbs_err > 1e-3
    assert caught, (mutation, species_err, energy_err, gibbs_err)
