"""The Cantera-format reader (create_rate_subs_amd.mech_ct, after
mech_interpret.py:866-1071) against the Chemkin reader (CPU).

cantera is not installed, so no file cantera wrote can be read here and the
reference's own conversion cannot be run.  What pins the reader: the same
mechanism written as Cantera YAML (here, from the Chemkin reader's output,
in two unit systems: SI/kmol/J and cm/mol/cal) must come back as the data
the Chemkin reader gives for the Chemkin text, field by field: A in
m/kmol/s (with the reaction-order-dependent unit conversion, three-body and
low-pressure limits one order up), activation temperatures, third-body
efficiencies, the falloff collider, Troe/SRI parameters, PLOG rows and
Chebyshev coefficients (the log10 unit shift lands on the constant term),
NASA-7 ranges and molecular weights.  Planted mistakes in the written file
(a wrong length unit, an order dropped from a three-body rate) are caught.
"""

import math

import numpy as np
import pytest
import yaml

from create_rate_subs_amd import chem, mech_ct, workloads
from create_rate_subs_amd import mech_interpret as mi
from create_rate_subs_amd.create_rate_subs import load_mechanism

MECHS = ['mech21pc', 'mech53', 'mech111', 'mech13co', 'mech17tm',
         'mech19fx']

CAL = 4184.0                 # J/kcal: J/kmol per cal/mol
SI = dict(length='m', quantity='kmol', E='J/kmol', Ef=1.0, conc=1.0,
          P='Pa', Pf=1.0)
CGS = dict(length='cm', quantity='mol', E='cal/mol', Ef=CAL, conc=1.0e-3 / 1e-6,
           P='atm', Pf=101325.0)


def _raw(name):
    mpath, tpath = workloads.mechanism_files(name)
    return mi.read_mech(mpath, tpath)


def _side(names, nus):
    terms = []
    for n, nu in zip(names, nus):
        terms.append(n if nu == 1 else '{} {}'.format(
            int(nu) if nu == int(nu) else nu, n))
    return ' + '.join(terms)


def _arr(A, b, E_K, order, u, E_unit=None):
    """Arrhenius in the file's units: A from kmol/m^3 concentrations,
    E from an activation temperature."""
    conc = u['conc']          # file concentration unit in kmol/m^3
    A_file = A / conc ** (1.0 - order)
    E_jkmol = E_K * chem.RU_JOUL * 1000.0
    return {'A': A_file, 'b': b, 'Ea': E_jkmol / u['Ef']}


def write_yaml(name, path, u, plant=None):
    elems, specs, reacs = _raw(name)
    doc = {'units': {'length': u['length'], 'quantity': u['quantity'],
                     'activation-energy': u['E'], 'pressure': u['P']},
           'phases': [{'name': 'gas', 'thermo': 'ideal-gas',
                       'elements': [e.capitalize() if len(e) > 1 else e.upper()
                                    for e in elems],
                       'species': [sp.name for sp in specs]}],
           'species': [], 'reactions': []}
    for sp in specs:
        doc['species'].append({
            'name': sp.name,
            'composition': {e.capitalize() if len(e) > 1 else e.upper(): n
                            for e, n in sp.elem},
            'thermo': {'model': 'NASA7',
                       'temperature-ranges': [float(t) for t in sp.Trange],
                       'data': [[float(c) for c in sp.lo],
                                [float(c) for c in sp.hi]]}})
    for x in reacs:
        order = sum(x.reac_nu)
        lhs, rhs = _side(x.reac, x.reac_nu), _side(x.prod, x.prod_nu)
        arrow = ' <=> ' if x.rev else ' => '
        r = {}
        if x.plog:
            r['equation'] = lhs + arrow + rhs
            r['type'] = 'pressure-dependent-Arrhenius'
            r['rate-constants'] = []
            for p, A, b, E in x.plog_par:
                k = _arr(A, b, E, order, u)
                k['P'] = '{!r} {}'.format(p / u['Pf'], u['P'])
                r['rate-constants'].append(k)
        elif x.cheb:
            r['equation'] = lhs + arrow + rhs
            r['type'] = 'Chebyshev'
            r['temperature-range'] = [float(t) for t in x.cheb_tlim]
            r['pressure-range'] = ['{!r} {}'.format(p / u['Pf'], u['P'])
                                   for p in x.cheb_plim]
            data = np.array(x.cheb_par, dtype=float).copy()
            data[0, 0] -= math.log10(u['conc'] ** (1.0 - order))
            r['data'] = data.tolist()
        elif x.pdep:
            coll = x.pdep_sp or 'M'
            r['equation'] = '{} (+{}){}{} (+{})'.format(lhs, coll, arrow, rhs,
                                                        coll)
            r['type'] = 'falloff' if x.low else 'chemically-activated'
            if x.low:
                r['high-P-rate-constant'] = _arr(x.A, x.b, x.E, order, u)
                r['low-P-rate-constant'] = _arr(*x.low, order + 1, u)
            else:
                r['low-P-rate-constant'] = _arr(x.A, x.b, x.E, order + 1, u)
                r['high-P-rate-constant'] = _arr(*x.high, order, u)
            if x.thd_body_eff:
                r['efficiencies'] = {s: e for s, e in x.thd_body_eff}
            if x.troe:
                t = dict(zip(['A', 'T3', 'T1', 'T2'], x.troe_par))
                r['Troe'] = t
            elif x.sri:
                r['SRI'] = dict(zip(['A', 'B', 'C', 'D', 'E'], x.sri_par))
        elif x.thd_body:
            r['equation'] = lhs + ' + M' + arrow + rhs + ' + M'
            r['type'] = 'three-body'
            o = order if plant == 'thd_order' else order + 1
            r['rate-constant'] = _arr(x.A, x.b, x.E, o, u)
            if x.thd_body_eff:
                r['efficiencies'] = {s: e for s, e in x.thd_body_eff}
        else:
            r['equation'] = lhs + arrow + rhs
            r['rate-constant'] = _arr(x.A, x.b, x.E, order, u)
        if x.dup:
            r['duplicate'] = True
        doc['reactions'].append(r)
    if plant == 'length':
        doc['units']['length'] = 'mm'
    with open(path, 'w') as fh:
        yaml.safe_dump(doc, fh, sort_keys=False)
    return specs, reacs


def _close(a, b, rtol=1e-13):
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    return a.shape == b.shape and np.allclose(a, b, rtol=rtol, atol=0.0)


def _compare(specs0, reacs0, specs1, reacs1):
    """Names of the first mismatching fields (empty when all agree)."""
    bad = []
    assert [s.name for s in specs0] == [s.name for s in specs1]
    for s0, s1 in zip(specs0, specs1):
        if not _close(s0.mw, s1.mw, 1e-12):
            bad.append(('mw', s0.name))
        if not (_close(s0.hi, s1.hi, 0) and _close(s0.lo, s1.lo, 0) and
                _close(s0.Trange, s1.Trange, 0)):
            bad.append(('thermo', s0.name))
        if sorted((e.lower(), n) for e, n in s0.elem) != sorted(
                (e.lower(), n) for e, n in s1.elem):
            bad.append(('elem', s0.name))
    assert len(reacs0) == len(reacs1)
    for j, (x, y) in enumerate(zip(reacs0, reacs1)):
        if (list(x.reac), list(x.reac_nu), list(x.prod), list(x.prod_nu),
                x.rev, x.dup) != (list(y.reac), list(y.reac_nu), list(y.prod),
                                  list(y.prod_nu), y.rev, y.dup):
            bad.append(('species/flags', j))
        if not (x.plog or x.cheb) and not _close([x.A, x.b, x.E],
                                                 [y.A, y.b, y.E]):
            bad.append(('arrhenius', j))
        if (x.thd_body, x.pdep, x.pdep_sp) != (y.thd_body, y.pdep, y.pdep_sp):
            bad.append(('third body', j))
        if dict(x.thd_body_eff) != dict(y.thd_body_eff):
            bad.append(('efficiencies', j))
        for f in ('low', 'high', 'troe_par', 'sri_par'):
            if not _close(getattr(x, f), getattr(y, f)):
                bad.append((f, j))
        if (x.troe, x.sri, x.plog, x.cheb) != (y.troe, y.sri, y.plog, y.cheb):
            bad.append(('forms', j))
        if x.plog and not _close(x.plog_par, y.plog_par):
            bad.append(('plog', j))
        if x.cheb and not (_close(x.cheb_par, y.cheb_par, 1e-12) and
                           _close(x.cheb_plim, y.cheb_plim) and
                           _close(x.cheb_tlim, y.cheb_tlim) and
                           (x.cheb_n_temp, x.cheb_n_pres) ==
                           (y.cheb_n_temp, y.cheb_n_pres)):
            bad.append(('cheb', j))
    return bad


@pytest.mark.parametrize('units', ['si', 'cgs'])
@pytest.mark.parametrize('name', MECHS)
def test_yaml_round_trip_matches_chemkin_reader(tmp_path, name, units):
    path = tmp_path / (name + '.yaml')
    specs0, reacs0 = write_yaml(name, str(path), SI if units == 'si' else CGS)
    _e, specs1, reacs1 = mech_ct.read_mech_ct(str(path))
    assert _compare(specs0, reacs0, specs1, reacs1) == []


def test_mechanism_covers_every_form():
    both = []
    for name in MECHS:
        both += _raw(name)[2]
    assert any(x.plog for x in both) and any(x.cheb for x in both)
    assert any(x.troe for x in both) and any(x.sri for x in both)
    assert any(x.thd_body for x in both)
    assert any(x.pdep and not (x.troe or x.sri) for x in both)


@pytest.mark.parametrize('plant', ['length', 'thd_order'])
def test_round_trip_catches_planted_mistake(tmp_path, plant):
    path = tmp_path / 'm.yaml'
    specs0, reacs0 = write_yaml('mech21pc', str(path), CGS, plant=plant)
    _e, specs1, reacs1 = mech_ct.read_mech_ct(str(path))
    assert _compare(specs0, reacs0, specs1, reacs1) != []


def test_named_collider_and_pipeline(tmp_path):
    """'(+AR)' becomes pdep_sp (default efficiency 0, :952-954), and a YAML
    mechanism runs through load_mechanism's reordering like Chemkin."""
    doc = {
        'units': {'length': 'cm', 'quantity': 'mol',
                  'activation-energy': 'cal/mol'},
        'phases': [{'name': 'gas', 'elements': ['H', 'O', 'Ar'],
                    'species': ['H', 'O2', 'HO2', 'AR']}],
        'species': [
            {'name': n, 'composition': c,
             'thermo': {'model': 'NASA7',
                        'temperature-ranges': [300.0, 1000.0, 5000.0],
                        'data': [[2.5, 0, 0, 0, 0, 100.0, 4.0]] * 2}}
            for n, c in (('H', {'H': 1}), ('O2', {'O': 2}),
                         ('HO2', {'H': 1, 'O': 2}), ('AR', {'Ar': 1}))],
        'reactions': [{
            'equation': 'H + O2 (+AR) <=> HO2 (+AR)',
            'type': 'falloff',
            'high-P-rate-constant': {'A': 4.65e12, 'b': 0.44, 'Ea': 0.0},
            'low-P-rate-constant': {'A': 9.04e19, 'b': -1.5, 'Ea': 490.0},
            'Troe': {'A': 0.5, 'T3': 1.0e-30, 'T1': 1.0e30}}]}
    path = tmp_path / 'ar.yaml'
    with open(path, 'w') as fh:
        yaml.safe_dump(doc, fh)
    _e, specs, reacs = mech_ct.read_mech_ct(str(path))
    x = reacs[0]
    assert x.pdep and x.pdep_sp == 'AR' and x.thd_body_eff == []
    assert x.A == pytest.approx(4.65e12 * 1e-3, rel=1e-15)
    assert x.low[0] == pytest.approx(9.04e19 * 1e-6, rel=1e-15)
    assert x.low[2] == pytest.approx(490.0 * CAL / (chem.RU_JOUL * 1000.0),
                                     rel=1e-15)
    assert x.troe_par == [0.5, 1.0e-30, 1.0e30]
    _e, specs2, reacs2, _fwd = load_mechanism(str(path))
    assert specs2[-1].name == 'AR'
    assert reacs2[0].pdep_sp == 3


def test_cti_without_cantera_is_a_clear_error(tmp_path):
    path = tmp_path / 'm.cti'
    path.write_text('# not read\n')
    with pytest.raises(mi.MechanismError, match='cantera'):
        mech_ct.read_mech_ct(str(path))
