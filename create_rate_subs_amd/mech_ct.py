"""Cantera-format mechanisms (mech_interpret.py:866-1071, read_mech_ct).

The reference reads a Cantera ``Solution`` (from a .cti/.xml file through
the cantera package, or one the caller already holds) and copies its
elements, species and reactions into SpecInfo/ReacInfo.  Cantera keeps every
quantity in SI units with kmol (A in m, kmol, s; activation energies in
J/kmol; pressures in Pa); the reference then turns the activation energy
into an activation temperature with the 'joules/kmole' factor.

cantera is not installed here, so this module has two halves:

* ``read_mech_ct(filename=None, gas=None)``: the reference's conversion,
  over any object with the Solution attributes it reads (``element_names``,
  ``atomic_weights``, ``species_names``, ``molecular_weights``,
  ``species(i)``, ``reactions()``).  Reaction kinds are told apart by class
  name (ThreeBodyReaction, FalloffReaction, ChemicallyActivatedReaction,
  PlogReaction, ChebyshevReaction, ElementaryReaction), as the cantera 2.x
  classes the reference tests with isinstance;
* ``load_yaml(path)``: a reader of Cantera's YAML mechanism format (the
  format cantera converts .cti/.xml files to) that builds such an object,
  with the file's ``units`` section and per-value units converted to
  cantera's SI internals.

Only what the reference reads is modelled: NASA-7 thermo (two ranges),
elementary, three-body, falloff and chemically-activated (Lindemann, Troe,
SRI), PLOG and Chebyshev reactions.  Anything else raises MechanismError,
where the reference prints an error and exits.
"""

import math
import re

import numpy as np

from . import chem
from .mech_interpret import MechanismError, act_energy_fact

__all__ = ['read_mech_ct', 'load_yaml', 'CtSolution']

# cantera's gas constant [J/kmol/K] (CODATA 2018), used when a file gives
# activation energies in K
_CT_GAS_CONSTANT = 8314.462618
_AVOGADRO = 6.02214076e26          # per kmol

# factors to cantera's internal units
_LENGTH = {'m': 1.0, 'cm': 0.01, 'mm': 0.001}
_QUANTITY = {'kmol': 1.0, 'mol': 0.001, 'molec': 1.0 / _AVOGADRO}
_ENERGY_PER_QUANTITY = {          # to J/kmol
    'J/kmol': 1.0, 'J/mol': 1000.0, 'kJ/mol': 1.0e6, 'cal/mol': 4184.0,
    'kcal/mol': 4.184e6, 'eV': 96485332.12331, 'K': _CT_GAS_CONSTANT,
}
_PRESSURE = {'Pa': 1.0, 'atm': 101325.0, 'bar': 1.0e5, 'kPa': 1000.0,
             'MPa': 1.0e6, 'dyn/cm^2': 0.1, 'torr': 101325.0 / 760.0}
_DEFAULT_UNITS = {'length': 'm', 'quantity': 'kmol',
                  'activation-energy': 'J/kmol', 'pressure': 'Pa'}


# ---------------------------------------------------------------------------
# The reference's conversion (mech_interpret.py:866-1071)
# ---------------------------------------------------------------------------

def _arrhenius(rate):
    return (rate.pre_exponential_factor, rate.temperature_exponent,
            rate.activation_energy)


def read_mech_ct(filename=None, gas=None, elem_wt=None):
    """Elements, species and reactions of a Cantera mechanism.

    Returns ``(elems, specs, reacs)`` like mech_interpret.read_mech: A in
    m/kmol/s, E as an activation temperature [K] (the 'joules/kmole' factor
    on cantera's J/kmol, mech_interpret.py:937-938), species and reaction
    lists ready for the same reordering.  ``filename`` may be a YAML
    mechanism (read by load_yaml); .cti/.xml files need cantera itself.
    """
    if filename:
        if str(filename).endswith(('.yaml', '.yml')):
            gas = load_yaml(filename)
        else:
            try:
                import cantera as ct
            except ImportError:
                raise MechanismError(
                    'cantera is not installed: cannot read {}; convert it to '
                    'YAML (cantera ck2yaml / cti2yaml) or Chemkin'.format(
                        filename))
            gas = ct.Solution(filename)
    elif gas is None:
        raise MechanismError('need either a filename or a Solution object')
    if elem_wt is None:
        elem_wt = chem.get_elem_wt()

    # elements (:898-902): new elements' weights join the table
    elems = list(gas.element_names)
    for e, wt in zip(elems, gas.atomic_weights):
        if e.lower() not in elem_wt:
            elem_wt[e.lower()] = wt

    # species (:904-927)
    specs = []
    for i, name in enumerate(gas.species_names):
        spec = chem.SpecInfo(name)
        spec.mw = gas.molecular_weights[i]
        species = gas.species(i)
        for e in species.composition:
            spec.elem.append([e, species.composition[e]])
        thermo = species.thermo
        if type(thermo).__name__ != 'NasaPoly2':
            raise MechanismError('unsupported thermo form for species ' + name)
        coeffs = thermo.coeffs
        spec.Trange = [thermo.min_temp, coeffs[0], thermo.max_temp]
        spec.hi = coeffs[1:8]
        spec.lo = coeffs[8:15]
        specs.append(spec)

    # reactions (:929-1069)
    E_fac = act_energy_fact['joules/kmole']
    reacs = []
    for rxn in gas.reactions():
        kind = type(rxn).__name__
        names_r, nu_r = list(rxn.reactants.keys()), list(rxn.reactants.values())
        names_p, nu_p = list(rxn.products.keys()), list(rxn.products.values())

        def make(A, b, E):
            return chem.ReacInfo(rxn.reversible, names_r, nu_r, names_p, nu_p,
                                 A, b, E)
        if kind == 'ThreeBodyReaction':
            A, b, E = _arrhenius(rxn.rate)
            reac = make(A, b, E * E_fac)
            reac.thd_body = True
            for sp in rxn.efficiencies:
                reac.thd_body_eff.append([sp, rxn.efficiencies[sp]])
        elif kind in ('FalloffReaction', 'ChemicallyActivatedReaction'):
            main, other = (('high_rate', 'low_rate') if kind == 'FalloffReaction'
                           else ('low_rate', 'high_rate'))
            A, b, E = _arrhenius(getattr(rxn, main))
            reac = make(A, b, E * E_fac)
            reac.pdep = True
            if rxn.default_efficiency == 0.0:
                # a single species is the third body
                reac.pdep_sp = list(rxn.efficiencies.keys())[0]
            else:
                for sp in rxn.efficiencies:
                    reac.thd_body_eff.append([sp, rxn.efficiencies[sp]])
            A2, b2, E2 = _arrhenius(getattr(rxn, other))
            if kind == 'FalloffReaction':
                reac.low = [A2, b2, E2 * E_fac]
            else:
                reac.high = [A2, b2, E2 * E_fac]
            if rxn.falloff.type == 'Troe':
                reac.troe = True
                reac.troe_par = list(rxn.falloff.parameters)
            elif rxn.falloff.type == 'SRI':
                reac.sri = True
                reac.sri_par = list(rxn.falloff.parameters)
        elif kind == 'PlogReaction':
            reac = make(0.0, 0.0, 0.0)
            reac.plog = True
            reac.plog_par = []
            for p, rate in rxn.rates:
                A, b, E = _arrhenius(rate)
                reac.plog_par.append([p, A, b, E * E_fac])
        elif kind == 'ChebyshevReaction':
            reac = make(0.0, 0.0, 0.0)
            reac.cheb = True
            reac.cheb_n_temp = rxn.nTemperature
            reac.cheb_n_pres = rxn.nPressure
            reac.cheb_plim = [rxn.Pmin, rxn.Pmax]
            reac.cheb_tlim = [rxn.Tmin, rxn.Tmax]
            reac.cheb_par = rxn.coeffs
        elif kind == 'ElementaryReaction':
            A, b, E = _arrhenius(rxn.rate)
            reac = make(A, b, E * E_fac)
        else:
            raise MechanismError('unsupported reaction: ' + kind)
        reac.dup = rxn.duplicate
        reacs.append(reac)
    return elems, specs, reacs


# ---------------------------------------------------------------------------
# A Solution-like object from Cantera's YAML format
# ---------------------------------------------------------------------------

class Arrhenius(object):
    def __init__(self, A, b, E):
        self.pre_exponential_factor = A
        self.temperature_exponent = b
        self.activation_energy = E


class NasaPoly2(object):
    """cantera's two-range NASA-7 layout: coeffs = [T_mid, high(7), low(7)]."""

    def __init__(self, tmin, tmid, tmax, low, high):
        self.min_temp = tmin
        self.max_temp = tmax
        self.coeffs = np.array([tmid] + list(high) + list(low), dtype=float)


class Species(object):
    def __init__(self, name, composition, thermo):
        self.name = name
        self.composition = composition
        self.thermo = thermo


class Falloff(object):
    def __init__(self, kind, parameters):
        self.type = kind
        self.parameters = np.array(parameters, dtype=float)


class _Reaction(object):
    def __init__(self, equation, reactants, products, reversible, duplicate):
        self.equation = equation
        self.reactants = reactants
        self.products = products
        self.reversible = reversible
        self.duplicate = duplicate


class ElementaryReaction(_Reaction):
    pass


class ThreeBodyReaction(_Reaction):
    pass


class FalloffReaction(_Reaction):
    pass


class ChemicallyActivatedReaction(_Reaction):
    pass


class PlogReaction(_Reaction):
    pass


class ChebyshevReaction(_Reaction):
    pass


class CtSolution(object):
    """The parts of a cantera Solution that read_mech_ct reads."""

    def __init__(self, elements, atomic_weights, species, reactions):
        self.element_names = elements
        self.atomic_weights = atomic_weights
        self._species = species
        self._reactions = reactions
        self.species_names = [sp.name for sp in species]
        wt = dict(zip(elements, atomic_weights))
        self.molecular_weights = np.array(
            [sum(n * wt[e] for e, n in sp.composition.items())
             for sp in species])

    def species(self, i):
        return self._species[i]

    def reactions(self):
        return list(self._reactions)


_QTY = re.compile(r'^\s*([-+0-9.eE]+)\s*(.*?)\s*$')


def _value(x, table, default_unit):
    """A YAML number, or a 'value unit' string, in cantera's SI."""
    if isinstance(x, (int, float)):
        return float(x) * table[default_unit]
    mo = _QTY.match(str(x))
    if not mo:
        raise MechanismError('cannot read quantity {!r}'.format(x))
    unit = mo.group(2) or default_unit
    if unit not in table:
        raise MechanismError('unsupported unit {!r}'.format(unit))
    return float(mo.group(1)) * table[unit]


def _conc_unit(units):
    """Factor of one concentration unit (quantity / length^3) to kmol/m^3."""
    return _QUANTITY[units['quantity']] / _LENGTH[units['length']] ** 3


def _rate_const(d, units, order, where):
    """An Arrhenius mapping or [A, b, Ea] list; A of a rate of the given
    reaction order (concentration^(1 - order) / s) in kmol, m, s."""
    if isinstance(d, dict):
        A, b, Ea = d.get('A'), d.get('b', 0.0), d.get('Ea', 0.0)
    else:
        A, b, Ea = d
    if A is None:
        raise MechanismError('rate constant without A in ' + where)
    if isinstance(A, str):
        # an explicit unit on A is not interpreted; only plain numbers
        raise MechanismError('units on A are not supported (' + where + ')')
    A = float(A) * _conc_unit(units) ** (1.0 - order)
    Ea = _value(Ea, _ENERGY_PER_QUANTITY, units['activation-energy'])
    return Arrhenius(A, float(b), Ea)


_MARKER = re.compile(r'\(\s*\+\s*([^()\s]+)\s*\)')


def _side(text):
    """Species -> stoichiometric coefficient of one side of an equation
    (cantera separates terms with ' + ')."""
    out = {}
    third = False
    for term in text.split(' + '):
        term = term.strip()
        if not term:
            raise MechanismError('empty term in ' + repr(text))
        parts = term.split(None, 1)
        if len(parts) == 2 and re.match(r'^[0-9.]+$', parts[0]):
            nu, name = float(parts[0]), parts[1].strip()
            nu = int(nu) if nu == int(nu) else nu
        else:
            nu, name = 1, term
        if name == 'M':
            third = True
            continue
        out[name] = out.get(name, 0) + nu
    return out, third


def _equation(eq):
    """(reactants, products, reversible, collider) of a cantera equation;
    collider is None, 'M' for '+ M' / '(+M)', or the species of '(+SP)'."""
    for arrow, rev in (('<=>', True), ('=>', False), ('=', True)):
        if arrow in eq:
            lhs, rhs = eq.split(arrow, 1)
            break
    else:
        raise MechanismError('no arrow in equation ' + repr(eq))
    collider = None
    ml, mr = _MARKER.search(lhs), _MARKER.search(rhs)
    if ml or mr:
        collider = (ml or mr).group(1)
        lhs, rhs = _MARKER.sub('', lhs), _MARKER.sub('', rhs)
    reac, tl = _side(lhs)
    prod, tr = _side(rhs)
    if tl or tr:
        collider = 'M'
    return reac, prod, rev, collider


def load_yaml(path):
    """A CtSolution from a Cantera YAML mechanism file (the first phase's
    elements, species and reactions; everything in cantera's SI units)."""
    import yaml
    with open(path) as fh:
        doc = yaml.safe_load(fh)
    units = dict(_DEFAULT_UNITS)
    for k, v in (doc.get('units') or {}).items():
        units[k] = v
    for key, table in (('length', _LENGTH), ('quantity', _QUANTITY),
                       ('activation-energy', _ENERGY_PER_QUANTITY),
                       ('pressure', _PRESSURE)):
        if units[key] not in table:
            raise MechanismError('unsupported {} unit {!r}'.format(
                key, units[key]))
    phase = (doc.get('phases') or [{}])[0]
    species_doc = doc.get('species') or []
    names = phase.get('species') or [s['name'] for s in species_doc]
    if isinstance(names, str) and names == 'all':
        names = [s['name'] for s in species_doc]
    by_name = {s['name']: s for s in species_doc}

    wt_table = chem.get_elem_wt()
    elements = phase.get('elements') or []
    species = []
    for name in names:
        if name not in by_name:
            raise MechanismError('species {} has no entry'.format(name))
        s = by_name[name]
        comp = {e: float(n) for e, n in (s.get('composition') or {}).items()}
        for e in comp:
            if e not in elements:
                elements.append(e)
        th = s.get('thermo') or {}
        if th.get('model') != 'NASA7':
            raise MechanismError('unsupported thermo form for species ' + name)
        tr = [float(t) for t in th['temperature-ranges']]
        data = th['data']
        if len(tr) != 3 or len(data) != 2:
            raise MechanismError('species {}: two NASA-7 ranges expected'
                                 .format(name))
        species.append(Species(name, comp, NasaPoly2(
            tr[0], tr[1], tr[2], [float(c) for c in data[0]],
            [float(c) for c in data[1]])))
    weights = []
    for e in elements:
        if e.lower() not in wt_table:
            raise MechanismError('unknown element ' + e)
        weights.append(wt_table[e.lower()])

    reactions = []
    for r in doc.get('reactions') or []:
        eq = r['equation']
        reac, prod, rev, collider = _equation(eq)
        order = sum(reac.values())
        kind = r.get('type', 'elementary')
        dup = bool(r.get('duplicate', False))
        if r.get('orders') or r.get('negative-A'):
            raise MechanismError('unsupported reaction options in ' + eq)
        if kind == 'elementary' and collider is None:
            rx = ElementaryReaction(eq, reac, prod, rev, dup)
            rx.rate = _rate_const(r['rate-constant'], units, order, eq)
        elif kind in ('three-body', 'elementary') and collider == 'M':
            rx = ThreeBodyReaction(eq, reac, prod, rev, dup)
            rx.rate = _rate_const(r['rate-constant'], units, order + 1, eq)
            rx.efficiencies = dict(r.get('efficiencies') or {})
            rx.default_efficiency = float(r.get('default-efficiency', 1.0))
        elif kind in ('falloff', 'chemically-activated'):
            cls = (FalloffReaction if kind == 'falloff'
                   else ChemicallyActivatedReaction)
            rx = cls(eq, reac, prod, rev, dup)
            rx.high_rate = _rate_const(r['high-P-rate-constant'], units,
                                       order, eq)
            rx.low_rate = _rate_const(r['low-P-rate-constant'], units,
                                      order + 1, eq)
            if collider not in (None, 'M'):
                rx.efficiencies = {collider: 1.0}
                rx.default_efficiency = 0.0
            else:
                rx.efficiencies = dict(r.get('efficiencies') or {})
                rx.default_efficiency = float(r.get('default-efficiency', 1.0))
            if 'Troe' in r:
                t = r['Troe']
                pars = [t['A'], t['T3'], t['T1']] + (
                    [t['T2']] if 'T2' in t else [])
                rx.falloff = Falloff('Troe', pars)
            elif 'SRI' in r:
                t = r['SRI']
                pars = [t['A'], t['B'], t['C']] + (
                    [t['D'], t['E']] if 'D' in t else [])
                rx.falloff = Falloff('SRI', pars)
            else:
                rx.falloff = Falloff('Lindemann', [])
        elif kind == 'pressure-dependent-Arrhenius':
            rx = PlogReaction(eq, reac, prod, rev, dup)
            rx.rates = []
            for k in r['rate-constants']:
                p = _value(k['P'], _PRESSURE, units['pressure'])
                rx.rates.append((p, _rate_const(k, units, order, eq)))
        elif kind == 'Chebyshev':
            rx = ChebyshevReaction(eq, reac, prod, rev, dup)
            rx.Tmin, rx.Tmax = (float(t) for t in r['temperature-range'])
            rx.Pmin, rx.Pmax = (_value(p, _PRESSURE, units['pressure'])
                                for p in r['pressure-range'])
            coeffs = np.array(r['data'], dtype=float)
            # log10 k: the rate's concentration unit enters the constant term
            coeffs[0, 0] += math.log10(_conc_unit(units) ** (1.0 - order))
            rx.nTemperature, rx.nPressure = coeffs.shape
            rx.coeffs = coeffs
        else:
            raise MechanismError('unsupported reaction type {!r} in {}'.format(
                kind, eq))
        reactions.append(rx)
    return CtSolution(elements, weights, species, reactions)
