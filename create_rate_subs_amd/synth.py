"""Seeded synthetic mechanisms, written as real Chemkin and NASA-7 text.

The reference ships no mechanism, so every workload is generated.  The
generator writes mechanism/thermo *files* so that the parser (and with it
every unit conversion) is exercised end to end, exactly as a user's
mechanism would be.

Species are C/H/O compositions drawn from a small lattice (so that many
element-balanced reactions exist) plus N2 as the bath gas in the last
position; N2 takes part in reactions only as a collider.  Every reaction is
element-balanced, which the mass- and element-conservation invariants and
the equilibrium construction of near-equilibrium states rely on.

Thermodynamics: cp/R(T) = 2.5 + c (1 - exp(-T/theta)) with c ~ U(0.5, 8)
and theta ~ U(500, 3000) K is least-squares fitted by a quartic on
[300, 1000] and [1000, 5000] K; a5/a6 of the low range reproduce
h298 ~ N(0, 150 kJ/mol) and s298 ~ U(150, 300) J/mol/K, those of the high
range make H and S continuous at T_mid = 1000 K.

Rates: log10 A ~ U(8, 16) (cgs/mole), b ~ U(-2, 3), Ea ~ U(0, 60) kcal/mol
(written in cal/mole, the Chemkin default).  Printed parameters are
quantised (b to 3 decimals, efficiencies to 2, Troe/SRI to 6 significant
digits) so that Python 2 and Python 3 spellings of every emitted constant
coincide.
"""

import math
import os
from dataclasses import dataclass, field

import numpy as np

__all__ = ['MechConfig', 'MECHS', 'generate_mechanism', 'write_mechanism']

_ELEMS = ('C', 'H', 'O')


@dataclass(frozen=True)
class MechConfig:
    """Recipe for one synthetic mechanism."""
    name: str
    n_species: int
    n_reactions: int
    frac_rev: float
    frac_thd: float
    n_enhanced: int
    frac_troe: float = 0.0
    frac_lind: float = 0.0
    frac_sri: float = 0.0
    frac_plog: float = 0.0
    frac_cheb: float = 0.0
    seed: int = 0
    # species k with k % 3 == 1 get this T_mid instead of 1000 K (0: none),
    # so that a mechanism has two T_mid groups
    tmid_alt: float = 0.0
    # the remaining rate-form branches (from a separate random stream, so
    # the draws above are unchanged): every 2nd Troe reaction without T**,
    # every 2nd SRI one with d and e, every 3rd falloff reaction with a
    # single species as collider, every 2nd Lindemann one chemically
    # activated (HIGH instead of LOW), every 5th reversible elementary one
    # with explicit reverse parameters (REV), every 3rd irreversible
    # elementary one with all coefficients halved (non-integer orders)
    form_variants: bool = False
    # thermochemistry and rate-constant draws (the defaults are the recipe
    # of the workloads; narrower ones give a non-stiff mechanism): h298 ~
    # N(0, h298_sd) J/mol, s298 ~ U(s298_range) J/mol/K, log10 A ~
    # U(logA_range) + logA_per_order (order - 1) in cgs units, with order
    # the reactant coefficient sum (+1 for a third body), b ~ U(b_range),
    # Ea ~ U(Ea_range) kcal/mol
    h298_sd: float = 150.0e3
    s298_range: tuple = (150.0, 300.0)
    logA_range: tuple = (8.0, 16.0)
    logA_per_order: float = 0.0
    b_range: tuple = (-2.0, 3.0)
    Ea_range: tuple = (0.0, 60.0)


# The three mechanisms the workloads are built on.
MECHS = {
    'mech13': MechConfig('mech13', 13, 27, 0.85, 0.15, 3, seed=1000),
    'mech53': MechConfig('mech53', 53, 325, 0.80, 0.10, 5, frac_troe=0.08,
                         frac_lind=0.02, seed=1001),
    'mech111': MechConfig('mech111', 111, 784, 0.80, 0.12, 5, frac_troe=0.10,
                          frac_sri=0.03, seed=1003),
    # pressure-dependent rate forms (PLOG, Chebyshev) beside the others
    'mech21pc': MechConfig('mech21pc', 21, 60, 0.80, 0.10, 3, frac_troe=0.05,
                           frac_plog=0.15, frac_cheb=0.15, seed=1005),
    # two T_mid groups (1000 K and 1300 K): the per-group range masks
    'mech17tm': MechConfig('mech17tm', 17, 50, 0.85, 0.12, 3, frac_troe=0.08,
                           frac_lind=0.04, seed=1007, tmid_alt=1300.0),
    # the rarer rate forms: 3-parameter Troe, 5-parameter SRI, a species
    # as falloff collider, chemically activated (HIGH), explicit REV
    'mech19fx': MechConfig('mech19fx', 19, 80, 0.80, 0.10, 3, frac_troe=0.10,
                           frac_lind=0.10, frac_sri=0.08, seed=1009,
                           form_variants=True),
    # non-stiff: Gibbs energies within a few kJ/mol of each other (moderate
    # equilibrium constants), pre-exponential factors scaled by reaction
    # order and activation energies of 18-24 kcal/mol, so that every time
    # scale at 1200-1500 K and 1 atm is longer than about 0.1 ms: an
    # explicit integrator converges at steps of microseconds
    'mech9ns': MechConfig('mech9ns', 9, 14, 0.80, 0.20, 2, seed=1011,
                          h298_sd=4.0e3, s298_range=(195.0, 225.0),
                          logA_range=(5.0, 6.0), logA_per_order=5.5,
                          b_range=(0.0, 0.5), Ea_range=(22.0, 28.0)),
}

# PLOG table pressures [atm] and Chebyshev validity box of the generator
PLOG_ATM = (0.1, 1.0, 10.0, 100.0)
CHEB_T = (300.0, 3000.0)
CHEB_P_ATM = (0.01, 200.0)


def _species_name(comp):
    name = ''
    for el, n in zip(_ELEMS, comp):
        if n:
            name += el + (str(n) if n > 1 else '')
    return name


def _pick_compositions(rng, count):
    """``count`` distinct (nC, nH, nO) compositions, smallest molecules first."""
    lattice = [(c, h, o) for c in range(5) for h in range(9)
               for o in range(5) if c + h + o > 0]
    # order by atom count, random within a shell
    keys = rng.random(len(lattice))
    lattice = [x for _, x in sorted(zip(keys, lattice))]
    lattice.sort(key=lambda x: sum(x))
    return lattice[:count]


def _fit_nasa(rng, tm=1000.0, h_sd=150.0e3, s_range=(150.0, 300.0)):
    """Random smooth cp/R -> NASA-7 (lo, hi) coefficient lists, ranges
    [300, tm] and [tm, 5000] K."""
    c = rng.uniform(0.5, 8.0)
    theta = rng.uniform(500.0, 3000.0)
    h298 = rng.normal(0.0, h_sd)              # J/mol
    s298 = rng.uniform(*s_range)              # J/mol/K
    R = 8.3144621

    def cp(T):
        return 2.5 + c * (1.0 - np.exp(-T / theta))

    def fit(t0, t1):
        T = np.linspace(t0, t1, 200)
        Ts = T / 1000.0
        V = np.vander(Ts, 5, increasing=True)
        a, *_ = np.linalg.lstsq(V, cp(T), rcond=None)
        return [a[i] / 1000.0 ** i for i in range(5)]

    lo = fit(300.0, tm)
    hi = fit(tm, 5000.0)

    def h_poly(a, T):      # H/R without a5
        return (a[0] * T + a[1] * T ** 2 / 2 + a[2] * T ** 3 / 3 +
                a[3] * T ** 4 / 4 + a[4] * T ** 5 / 5)

    def s_poly(a, T):      # S/R without a6
        return (a[0] * math.log(T) + a[1] * T + a[2] * T ** 2 / 2 +
                a[3] * T ** 3 / 3 + a[4] * T ** 4 / 4)

    T0, Tm = 298.15, tm
    lo5 = h298 / R - h_poly(lo, T0)
    lo6 = s298 / R - s_poly(lo, T0)
    hi5 = h_poly(lo, Tm) + lo5 - h_poly(hi, Tm)
    hi6 = s_poly(lo, Tm) + lo6 - s_poly(hi, Tm)
    return lo + [lo5, lo6], hi + [hi5, hi6]


def _fmt_coef(x):
    return '{:15.8E}'.format(x)


def _thermo_entry(name, elem_counts, lo, hi, tm=1000.0):
    """Four fixed-column NASA-7 lines (columns as read by read_thermo)."""
    elem = ''
    for el, n in elem_counts:
        elem += '{:<2}{:>3}'.format(el, n)
    elem = '{:<20}'.format(elem)
    l1 = '{:<18}{:<6}{}G{:>10.3f}{:>10.3f}{:>8.2f}'.format(
        name, 'SYN', elem, 300.0, 5000.0, tm)
    l1 = '{:<79}1'.format(l1)
    l2 = ''.join(_fmt_coef(x) for x in hi[0:5]) + '    2'
    l3 = ''.join(_fmt_coef(x) for x in (hi[5], hi[6], lo[0], lo[1], lo[2]))
    l3 += '    3'
    l4 = ''.join(_fmt_coef(x) for x in lo[3:7]) + '{:>15}'.format('') + '    4'
    return '\n'.join([l1, l2, l3, l4]) + '\n'


@dataclass
class _Rxn:
    reac: list
    prod: list
    rev: bool
    kind: str = 'elem'          # elem | thd | troe | lind | sri | plog | cheb
    arr: tuple = ()
    low: tuple = ()
    troe: tuple = ()
    sri: tuple = ()
    eff: list = field(default_factory=list)
    plog: list = field(default_factory=list)   # (P atm, A, b, Ea) rows
    cheb: list = field(default_factory=list)   # [n_temp][n_pres] coeffs
    high: tuple = ()            # chemically activated: HIGH instead of LOW
    coll: int = -1              # falloff collider species (-1: M)
    rev_par: tuple = ()         # explicit reverse Arrhenius (REV)


def _side_text(terms, names):
    out = []
    for sp, nu in terms:
        if nu == 1:
            coef = ''
        elif nu == int(nu):
            coef = str(int(nu))
        else:
            coef = '{:g}'.format(nu)
        out.append(coef + names[sp])
    return ' + '.join(out)


def _arrhenius(rng, cfg=None, order=1.0):
    """(A, b, Ea) drawn as MechConfig describes (the workloads' ranges
    without ``cfg``)."""
    la, br, er = ((cfg.logA_range, cfg.b_range, cfg.Ea_range) if cfg else
                  ((8.0, 16.0), (-2.0, 3.0), (0.0, 60.0)))
    logA = rng.uniform(*la)
    if cfg is not None:
        logA += cfg.logA_per_order * (order - 1.0)
    A = float('%.6E' % (10.0 ** logA))
    b = round(rng.uniform(*br), 3)
    Ea = round(rng.uniform(*er) * 1000.0, 3)    # cal/mol
    return A, b, Ea


def _form_variants(rxns, nsp, rng):
    """Turn some reactions into the rarer rate forms (see MechConfig)."""
    n_troe = n_sri = n_fall = n_lind = n_rev = n_irr = 0
    for r in rxns:
        if r.kind == 'troe':
            n_troe += 1
            if n_troe % 2 == 0:
                r.troe = r.troe[:3]
        if r.kind == 'sri':
            n_sri += 1
            if n_sri % 2 == 0:
                r.sri = r.sri + (float('%.6g' % rng.uniform(0.5, 2.0)),
                                 float('%.4f' % rng.uniform(-0.5, 0.5)))
        if r.kind in ('troe', 'lind', 'sri'):
            n_fall += 1
            if n_fall % 3 == 0:
                inside = {k for k, _nu in r.reac + r.prod}
                free = [k for k in range(nsp) if k not in inside]
                r.coll = int(free[int(rng.integers(len(free)))])
                r.eff = []
        if r.kind == 'lind':
            n_lind += 1
            if n_lind % 2 == 0:
                # main line = low-pressure limit, HIGH = high-pressure one
                A, b, Ea = r.low
                r.arr, r.low = (A, b, Ea), ()
                r.high = (float('%.6E' % (A * 10.0 ** -rng.uniform(2.0, 6.0))),
                          round(rng.uniform(-2.0, 3.0), 3),
                          round(rng.uniform(0.0, 60.0) * 1000.0, 3))
        if r.kind == 'elem' and r.rev:
            n_rev += 1
            if n_rev % 5 == 0:
                r.rev_par = _arrhenius(rng)
        if r.kind == 'elem' and not r.rev:
            n_irr += 1
            if n_irr % 3 == 0:
                # still balanced: both sides scaled by 1/2
                r.reac = [(k, nu / 2.0) for k, nu in r.reac]
                r.prod = [(k, nu / 2.0) for k, nu in r.prod]


def generate_mechanism(cfg):
    """Build the mechanism described by ``cfg``; return (mech, thermo) text."""
    rng = np.random.default_rng(cfg.seed)
    comps = _pick_compositions(rng, cfg.n_species - 1)
    names = [_species_name(c) for c in comps] + ['N2']
    nsp = len(names)
    vec = [np.array(c) for c in comps]

    # index every 1- and 2-term product side by its element vector
    sides = {}
    ns = nsp - 1
    for i in range(ns):
        for nu in (1, 2):
            sides.setdefault(tuple(vec[i] * nu), []).append(((i, nu),))
        for j in range(i + 1, ns):
            for nu_i in (1, 2):
                for nu_j in (1, 2):
                    key = tuple(vec[i] * nu_i + vec[j] * nu_j)
                    sides.setdefault(key, []).append(((i, nu_i), (j, nu_j)))

    def coef():
        return 2 if rng.random() < 0.1 else 1

    seen = set()
    balanced = []
    tries = 0
    while len(balanced) < cfg.n_reactions:
        tries += 1
        if tries > 200000:
            raise RuntimeError('could not draw enough balanced reactions')
        n_r = 1 + int(rng.random() < 0.6)
        picks = rng.choice(ns, size=n_r, replace=False)
        reac = tuple(sorted((int(i), coef()) for i in picks))
        key = tuple(sum(vec[i] * nu for i, nu in reac))
        cands = sides.get(key, [])
        rs = {i for i, _ in reac}
        cands = [p for p in cands if not ({i for i, _ in p} & rs)]
        if not cands:
            continue
        prod = cands[int(rng.integers(len(cands)))]
        sig = (reac, tuple(prod))
        if sig in seen or (tuple(prod), reac) in seen:
            continue
        seen.add(sig)
        balanced.append((list(reac), list(prod)))

    nr = cfg.n_reactions
    n_rev = round(cfg.frac_rev * nr)
    n_thd = round(cfg.frac_thd * nr)
    n_troe = round(cfg.frac_troe * nr)
    n_lind = round(cfg.frac_lind * nr)
    n_sri = round(cfg.frac_sri * nr)
    n_plog = round(cfg.frac_plog * nr)
    n_cheb = round(cfg.frac_cheb * nr)
    kinds = (['thd'] * n_thd + ['troe'] * n_troe + ['lind'] * n_lind +
             ['sri'] * n_sri + ['plog'] * n_plog + ['cheb'] * n_cheb)
    kinds += ['elem'] * (nr - len(kinds))
    kinds = [kinds[i] for i in rng.permutation(nr)]
    rev_flags = np.zeros(nr, dtype=bool)
    rev_flags[rng.permutation(nr)[:n_rev]] = True

    rxns = []
    for j, (reac, prod) in enumerate(balanced):
        r = _Rxn(reac=reac, prod=prod, rev=bool(rev_flags[j]), kind=kinds[j])
        r.arr = _arrhenius(rng, cfg, sum(nu for _k, nu in reac) +
                           (1 if kinds[j] == 'thd' else 0))
        if r.kind == 'plog':
            # ln k varies smoothly with ln P: a shifted A per pressure
            A, b, Ea = r.arr
            slope = rng.uniform(-0.5, 1.0)
            for p in PLOG_ATM:
                la = math.log10(A) + slope * math.log10(p) + rng.uniform(-0.2, 0.2)
                r.plog.append((p, float('%.6E' % 10.0 ** la),
                               round(b + rng.uniform(-0.2, 0.2), 3),
                               round(Ea + rng.uniform(-500.0, 500.0), 3)))
        if r.kind == 'cheb':
            nt, npr = (6, 4) if rng.random() < 0.5 else (5, 3)
            c = np.zeros((nt, npr))
            c[0, 0] = round(rng.uniform(9.0, 13.0), 4)
            c[1, 0] = round(-rng.uniform(0.2, 3.0), 4)
            for i in range(nt):
                for jp in range(npr):
                    if (i, jp) not in ((0, 0), (1, 0)):
                        c[i, jp] = round(rng.normal(0.0, 0.3 / (1 + i + jp)), 4)
            r.cheb = c.tolist()
        if r.kind not in ('elem', 'plog', 'cheb'):
            picks = rng.choice(nsp, size=cfg.n_enhanced, replace=False)
            r.eff = [(int(k), round(rng.uniform(0.5, 12.0), 2))
                     for k in sorted(picks)]
        if r.kind in ('troe', 'lind', 'sri'):
            A, b, Ea = r.arr
            logA0 = math.log10(A) + rng.uniform(2.0, 6.0)
            r.low = (float('%.6E' % 10.0 ** logA0),
                     round(rng.uniform(-2.0, 3.0), 3),
                     round(rng.uniform(0.0, 60.0) * 1000.0, 3))
        if r.kind == 'troe':
            r.troe = (float('%.4f' % rng.uniform(0.2, 0.9)),
                      float('%.6g' % rng.uniform(50.0, 500.0)),
                      float('%.6g' % rng.uniform(500.0, 3000.0)),
                      float('%.6g' % rng.uniform(3000.0, 8000.0)))
        if r.kind == 'sri':
            r.sri = (float('%.6g' % rng.uniform(0.1, 10.0)),
                     float('%.6g' % rng.uniform(100.0, 2000.0)),
                     float('%.6g' % rng.uniform(100.0, 2000.0)))
        rxns.append(r)
    if cfg.form_variants:
        _form_variants(rxns, nsp, np.random.default_rng(cfg.seed + 31337))

    # thermo
    thermo = ['THERMO\n', '   300.000  1000.000  5000.000\n']
    for i, comp in enumerate(comps):
        tm = cfg.tmid_alt if (cfg.tmid_alt and i % 3 == 1) else 1000.0
        lo, hi = _fit_nasa(rng, tm, cfg.h298_sd, cfg.s298_range)
        counts = [(el, n) for el, n in zip(_ELEMS, comp) if n]
        thermo.append(_thermo_entry(names[i], counts, lo, hi, tm))
    # bath gas: drawn like the others from its own stream; it only ever
    # acts as a collider, so its thermo never enters an equilibrium constant
    lo, hi = _fit_nasa(np.random.default_rng(cfg.seed + 7919), 1000.0,
                       cfg.h298_sd, cfg.s298_range)
    thermo.append(_thermo_entry('N2', [('N', 2)], lo, hi))
    thermo.append('END\n')

    # mechanism text
    mech = ['ELEMENTS\nC H O N\nEND\n', 'SPECIES\n']
    for k in range(0, nsp, 8):
        mech.append(' '.join(names[k:k + 8]) + '\n')
    mech.append('END\n')
    mech.append('REACTIONS\n')
    for r in rxns:
        lhs = _side_text(r.reac, names)
        rhs = _side_text(r.prod, names)
        if r.kind == 'thd':
            lhs += ' + M'
            rhs += ' + M'
        elif r.kind in ('troe', 'lind', 'sri'):
            coll = 'M' if r.coll < 0 else names[r.coll]
            lhs += ' (+{})'.format(coll)
            rhs += ' (+{})'.format(coll)
        arrow = '<=>' if r.rev else '=>'
        A, b, Ea = r.arr
        mech.append('{} {} {}   {:.6E} {:.3f} {:.3f}\n'.format(
            lhs, arrow, rhs, A, b, Ea))
        if r.low:
            mech.append('    LOW / {:.6E} {:.3f} {:.3f} /\n'.format(*r.low))
        if r.high:
            mech.append('    HIGH / {:.6E} {:.3f} {:.3f} /\n'.format(*r.high))
        if r.rev_par:
            mech.append('    REV / {:.6E} {:.3f} {:.3f} /\n'.format(
                *r.rev_par))
        if r.troe:
            mech.append('    TROE / {} /\n'.format(
                ' '.join('%.6g' % x for x in r.troe)))
        if r.sri:
            mech.append('    SRI / {} /\n'.format(
                ' '.join('%.6g' % x for x in r.sri)))
        if r.eff:
            mech.append('    ' + ' '.join('{}/{:.2f}/'.format(names[k], e)
                                          for k, e in r.eff) + '\n')
        for p, pa, pb, pe in r.plog:
            mech.append('    PLOG / {:.6E} {:.6E} {:.3f} {:.3f} /\n'.format(
                p, pa, pb, pe))
        if r.cheb:
            nt, npr = len(r.cheb), len(r.cheb[0])
            mech.append('    TCHEB / {:.1f} {:.1f} /\n'.format(*CHEB_T))
            mech.append('    PCHEB / {:.4f} {:.1f} /\n'.format(*CHEB_P_ATM))
            flat = ['{:.4f}'.format(v) for row in r.cheb for v in row]
            mech.append('    CHEB / {} {} {} /\n'.format(nt, npr,
                                                          ' '.join(flat[:4])))
            for k in range(4, len(flat), 6):
                mech.append('    CHEB / {} /\n'.format(' '.join(flat[k:k + 6])))
    mech.append('END\n')
    return ''.join(mech), ''.join(thermo)


def write_mechanism(cfg, directory):
    """Write ``<name>.dat`` and ``<name>_therm.dat``; return both paths."""
    os.makedirs(directory, exist_ok=True)
    mech, thermo = generate_mechanism(cfg)
    mpath = os.path.join(directory, cfg.name + '.dat')
    tpath = os.path.join(directory, cfg.name + '_therm.dat')
    with open(mpath, 'w') as fh:
        fh.write(mech)
    with open(tpath, 'w') as fh:
        fh.write(thermo)
    return mpath, tpath
