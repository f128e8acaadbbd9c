"""Pins of the C oracle (CPU).

The reference has no tests, fixtures or golden vectors, so the oracle is
pinned by what the mathematics fixes:

* an independent NumPy evaluator written from the textbook forms
  (oracle/numpy_ref.py) on every stage,
* mpmath at 50 digits on single states (closed forms, brute force),
* the Kc closed form through the reference's unused calc_spec_smh,
* ideal-gas identities, mass and element conservation,
* detailed balance at element-potential equilibrium states,
* the Lindemann limits,

and each pin is shown to catch a planted mistake (a flipped sign, a
swapped NASA range, a dropped prefactor, an inverted Kc) in a mutated copy
of the oracle.
"""

import math
import os
import re
import subprocess

import numpy as np
import pytest

from create_rate_subs_amd import chem, metrics, states
from oracle import emit_c, numpy_ref
from oracle.build import CFLAGS, HERE as ORACLE_DIR, Oracle

from _helpers import mechanism, oracle

MECHS = ['mech13', 'mech53', 'mech111', 'mech17tm', 'mech19fx']


def _rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)


@pytest.mark.parametrize('name', MECHS)
def test_stages_match_numpy(name):
    specs, reacs, _ = mechanism(name)
    T, P, Y = states.random_states(len(specs), 2048, 31)
    st = oracle(name).eval_stages(T, P, Y)
    ref = numpy_ref.evaluate(specs, reacs, T, P, Y)
    assert _rel(st['C'], ref['C']).max() <= 1e-15
    assert _rel(st['fwd'], ref['fwd']).max() <= 1e-13
    assert _rel(st['rev'], ref['rev']).max() <= 5e-13
    assert _rel(st['pm'], ref['pm']).max() <= 1e-13
    S = metrics.gross_scale(specs, reacs, st['fwd'], st['rev'], st['pm'])
    e = metrics.rel_err(st['wdot'], ref['wdot'], S, 1e-3)
    assert e.max() <= 1e-10, e.max()
    assert np.median(e) <= 1e-14


@pytest.mark.parametrize('name,count', [('mech13', 12), ('mech53', 3)])
def test_wdot_matches_mpmath(name, count):
    specs, reacs, _ = mechanism(name)
    T, P, Y = states.random_states(len(specs), count, 32)
    st = oracle(name).eval_stages(T, P, Y)
    S = metrics.gross_scale(specs, reacs, st['fwd'], st['rev'], st['pm'])
    for i in range(count):
        exact = np.array([float(v) for v in numpy_ref.mp_state(
            specs, reacs, T[i], P[i], Y[:, i])])
        e = metrics.rel_err(st['wdot'][:, i], exact, S[:, i], 1e-3)
        assert e.max() <= 1e-11, (i, e.max())


@pytest.mark.parametrize('name', MECHS)
def test_kc_closed_form(name):
    """Kc recovered from the oracle's f/r equals exp(sum nu smh)(PA/RU T)^dnu."""
    specs, reacs, _ = mechanism(name)
    T, P, Y = states.random_states(len(specs), 64, 33)
    st = oracle(name).eval_stages(T, P, Y)
    rev = [j for j, r in enumerate(reacs) if r.rev]
    worst = 0.0
    for i in (0, 7, 63):
        smh = chem.calc_spec_smh(T[i], specs)
        C = st['C'][:, i]
        for ri, j in enumerate(rev):
            r = reacs[j]
            cf = np.prod([C[k] ** v for k, v in zip(r.reac, r.reac_nu)])
            cr = np.prod([C[k] ** v for k, v in zip(r.prod, r.prod_nu)])
            kc_oracle = (st['fwd'][j, i] / cf) / (st['rev'][ri, i] / cr)
            dnu = sum(r.prod_nu) - sum(r.reac_nu)
            arg = (sum(v * smh[k] for k, v in zip(r.prod, r.prod_nu)) -
                   sum(v * smh[k] for k, v in zip(r.reac, r.reac_nu)))
            kc = math.exp(arg) * (chem.PA / (chem.RU * T[i])) ** dnu
            worst = max(worst, abs(kc_oracle / kc - 1.0))
    assert worst <= 1e-12, worst


@pytest.mark.parametrize('name', MECHS)
def test_ideal_gas_and_conservation(name):
    specs, reacs, _ = mechanism(name)
    T, P, Y = states.random_states(len(specs), 4096, 34)
    st = oracle(name).eval_stages(T, P, Y)
    W = np.array([s.mw for s in specs])
    ctot = st['C'].sum(axis=0)
    assert _rel(ctot, P / (chem.RU * T)).max() <= 1e-13
    assert _rel((st['C'] * W[:, None]).sum(axis=0), st['rho']).max() <= 1e-13
    S = metrics.gross_scale(specs, reacs, st['fwd'], st['rev'], st['pm'])
    # mass: sum_k W_k wdot_k = 0 to the rounding of the terms
    mass = (W[:, None] * st['wdot']).sum(axis=0)
    scale = (W[:, None] * S).sum(axis=0)
    assert (np.abs(mass) / scale).max() <= 1e-13
    # elements
    elems = sorted({e.upper() for s in specs for e, _ in s.elem})
    for el in elems:
        a = np.array([sum(c for e, c in s.elem if e.upper() == el)
                      for s in specs], dtype=float)
        bal = (a[:, None] * st['wdot']).sum(axis=0)
        sc = (a[:, None] * S).sum(axis=0)
        ok = sc > 0
        if ok.any():
            assert (np.abs(bal[ok]) / sc[ok]).max() <= 1e-13, el


@pytest.mark.parametrize('name', ['mech13', 'mech53'])
def test_detailed_balance_at_equilibrium(name):
    specs, reacs, _ = mechanism(name)
    T, P, Y = states.equilibrium_states(specs, 512, 35, perturb=0.0)
    st = oracle(name).eval_stages(T, P, Y)
    rev = [j for j, r in enumerate(reacs) if r.rev]
    f = st['fwd'][rev]
    r = st['rev']
    ok = (f + r) > 1e-280            # away from the subnormal range
    assert ok.mean() > 0.5
    assert (np.abs(f - r)[ok] / (f + r)[ok]).max() <= 1e-11
    # and so wdot vanishes relative to the gross rates
    S = metrics.gross_scale(specs, reacs, st['fwd'], st['rev'], st['pm'])
    irr = [j for j, x in enumerate(reacs) if not x.rev]
    if not irr:
        ok = S > 0
        assert (np.abs(st['wdot'][ok]) / S[ok]).max() <= 1e-11


def test_lindemann_limits():
    specs, reacs, _ = mechanism('mech53')
    lind = [j for j, r in enumerate(reacs)
            if r.pdep and not r.troe and not r.sri]
    pd = [j for j, r in enumerate(reacs) if r.thd_body or r.pdep]
    assert lind
    T, _P, Y = states.random_states(len(specs), 16, 36)
    o = oracle('mech53')
    for P, check in ((np.full(16, 1e-6), 'low'), (np.full(16, 1e15), 'high')):
        st = o.eval_stages(T, P, Y)
        ref = numpy_ref.evaluate(specs, reacs, T, P, Y)
        for j in lind:
            pm = st['pm'][pd.index(j)]
            r = reacs[j]
            k0 = r.low[0] * T ** r.low[1] * np.exp(-r.low[2] / T)
            kinf = r.A * T ** r.b * np.exp(-r.E / T)
            M = ref['C'].sum(axis=0) + sum((e - 1.0) * ref['C'][k]
                                           for k, e in r.thd_body_eff)
            Pr = k0 * M / kinf
            if check == 'low':
                ok = Pr < 1e-8
                assert np.all(np.abs(pm[ok] / Pr[ok] - 1.0) <= 1e-7)
            else:
                ok = Pr > 1e8
                assert np.all(np.abs(pm[ok] - 1.0) <= 1e-7)


def test_third_body_unit_efficiencies_equal_m():
    """[M] with no enhanced species is exactly m = P/(RU T)."""
    specs, reacs, _ = mechanism('mech13')
    T, P, Y = states.random_states(len(specs), 64, 37)
    st = oracle('mech13').eval_stages(T, P, Y)
    pd = [j for j, r in enumerate(reacs) if r.thd_body or r.pdep]
    m = P / (8.31446210e+03 * T)
    # pres_mod - sum (eff-1) C_k == m for every third-body reaction
    for p, j in enumerate(pd):
        r = reacs[j]
        extra = sum((float(e) - 1.0) * st['C'][k] for k, e in r.thd_body_eff)
        assert _rel(st['pm'][p] - extra, m).max() <= 1e-12


# ---------------------------------------------------------------------------
# planted mistakes: a pin that cannot fail pins nothing
# ---------------------------------------------------------------------------

MUTATIONS = {
    # E/T enters with the wrong sign in every forward rate constant
    'arrhenius_sign': ('rxn_rates.c', r' - \((\S+) / T\)\)', r' + (\1 / T))'),
    # lower and upper NASA ranges swapped in Kc
    'nasa_range_swap': ('rxn_rates.c', r'if \(T <= ', r'if (T > '),
    # (PA/RU)^dnu prefactor dropped
    'kc_prefactor': ('rxn_rates.c', r'Kc = \S+ \* exp\(Kc\)', r'Kc = exp(Kc)'),
    # reverse rate multiplied instead of divided by Kc
    'kc_inverted': ('rxn_rates.c', r'kf / Kc', r'kf * Kc'),
    # efficiency offsets applied with the wrong sign
    'efficiency_sign': ('rxn_rates_pres_mod.c', r'= m \+ ', r'= m - '),
}


def _mutant(tmp_path, name, mutation):
    specs, reacs, fwd = mechanism(name)
    src = tmp_path / 'src'
    cfiles = emit_c.emit_c(specs, reacs, str(src), fwd)
    fname, pat, rep = MUTATIONS[mutation]
    p = src / fname
    text, count = re.subn(pat, rep, p.read_text())
    assert count > 0, mutation
    p.write_text(text)
    so = str(tmp_path / 'liboracle.so')
    subprocess.run(['gcc'] + CFLAGS + ['-I', str(src), '-o', so] + cfiles +
                   [os.path.join(ORACLE_DIR, 'driver.c'), '-lm'], check=True)
    return Oracle(name, so)


@pytest.mark.parametrize('mutation', sorted(MUTATIONS))
def test_pins_catch_planted_mistake(tmp_path, mutation):
    name = 'mech13'
    specs, reacs, _ = mechanism(name)
    o = _mutant(tmp_path, name, mutation)
    T, P, Y = states.random_states(len(specs), 256, 38)
    st = o.eval_stages(T, P, Y)
    ref = numpy_ref.evaluate(specs, reacs, T, P, Y)
    S = metrics.gross_scale(specs, reacs, ref['fwd'], ref['rev'], ref['pm'])
    e = metrics.rel_err(st['wdot'], ref['wdot'], S, 1e-3)
    assert e.max() > 1e-3, (mutation, e.max())
