"""HIP emitter: mechanism -> CDNA4 kernels + C-ABI library source.

The ``hip`` language plugin of the generator.  It takes the parsed
mechanism in the reference's data model (``SpecInfo``/``ReacInfo`` with
integer species indices after the bath-gas reordering, exactly what the
reference's emitters receive at create_rate_subs.py:531-552) and writes one
self-contained ``crs_<name>.hip``: batched kernels for the four functions of
the wdot path, a fused kernel per precision mode, and the ``extern "C"``
entry points of ``include/crs_hip.h``.

Execution model: one thermochemical state per lane, wave64, 256 threads per
block, states contiguous in every structure-of-arrays row so that each
load/store of a wave is one coalesced 512-byte access.  Species indices are
compile-time constants (as in the reference's unrolled code), so per-state
concentrations and species-rate accumulators live in registers; numeric
mechanism constants live in one ``__constant__`` table read with
wave-uniform scalar loads.

Numerics of the FP64 kernels follow the reference's emitted expressions
(rate_subs.py:20-135, 226-704, 707-1073, 1076-1235, 1292-1391) term for
term, with every constant equal to the value of the literal the reference
would print (``utils.emitted``).  The stage kernels for eval_conc and
eval_spec_rates disable FMA contraction so that they round exactly as the
C subroutines do.

Mixed precision (``CRS_MODE_MIXED``) evaluates Arrhenius, falloff and
third-body terms in FP32 and keeps equilibrium constants, the net-rate
subtraction and the species sums in FP64.
"""

import math
import os
import re
import struct
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

from . import chem
from .utils import (emitted, fmt_e16, fmt_e8, fmt_g6, fmt_w3, get_nu,
                    is_integer, py2_str)

__all__ = ['MechIR', 'build_ir', 'write_hip', 'emit_hip_source']


# ---------------------------------------------------------------------------
# Mechanism tables ("IR"): every number the reference would print, as the
# value its literal denotes.
# ---------------------------------------------------------------------------

@dataclass
class ArrIR:
    """Rate-constant expression, rate_subs.py:20-135.

    value = pre * T^nT                     if not has_exp
          = [pre *] exp([logA] [+ b logT] [- E/T])   otherwise
    """
    pre: Optional[float]
    nT: int
    has_exp: bool
    logA: Optional[float]
    b: Optional[float]
    E: Optional[float]
    # raw (unrounded) parameters, for the mixed path's own evaluation
    A_raw: float = 0.0
    b_raw: float = 0.0
    E_raw: float = 0.0


def arr_ir(A, b, E):
    """Build the rate-constant expression the reference would emit."""
    if A > 0:
        logA = emitted(fmt_e16(math.log(A)))
        if not E:
            if not b:
                return ArrIR(emitted(py2_str(A)), 0, False, None, None, None,
                             A, b, E)
            if isinstance(b, int):
                return ArrIR(emitted(py2_str(A)), max(b, 0), False, None,
                             None, None, A, b, E)
            return ArrIR(None, 0, True, logA, emitted(py2_str(b)), None,
                         A, b, E)
        Ek = emitted(fmt_e16(E))
        if not b:
            return ArrIR(None, 0, True, logA, None, Ek, A, b, E)
        return ArrIR(None, 0, True, logA, emitted(py2_str(b)), Ek, A, b, E)
    if A < 0:
        if not E:
            if not b:
                return ArrIR(emitted(py2_str(A)), 0, False, None, None, None,
                             A, b, E)
            if is_integer(b):
                return ArrIR(emitted(py2_str(A)), max(int(b), 0), False,
                             None, None, None, A, b, E)
            return ArrIR(emitted(fmt_e16(A)), 0, True, None,
                         emitted(py2_str(b)), None, A, b, E)
        Ek = emitted(fmt_e16(E))
        if not b:
            return ArrIR(emitted(fmt_e16(A)), 0, True, None, None, Ek, A, b, E)
        return ArrIR(emitted(fmt_e16(A)), 0, True, None, emitted(py2_str(b)),
                     Ek, A, b, E)
    raise NotImplementedError('zero pre-exponential factor')


@dataclass
class RxnIR:
    index: int
    reac: List[Tuple[int, float]]
    prod: List[Tuple[int, float]]
    arr: ArrIR
    rev: bool = False
    rev_index: Optional[int] = None
    rev_arr: Optional[ArrIR] = None          # explicit REV parameters
    kc_groups: List[Tuple[float, list, list]] = field(default_factory=list)
    kc_pref: float = 1.0
    kc_dnu: float = 0.0
    pd_index: Optional[int] = None
    thd: bool = False                        # plain third-body
    falloff: Optional[str] = None            # 'troe' | 'sri' | 'lind'
    eff: List[Tuple[int, float]] = field(default_factory=list)  # (k, eff-1)
    pdep_sp: Optional[int] = None
    low: Optional[ArrIR] = None
    high: Optional[ArrIR] = None
    falloff_low: bool = False                # LOW given (unimolecular)
    falloff_high: bool = False               # HIGH given (chem. activated)
    troe: Optional[dict] = None
    sri: Optional[dict] = None
    plog: Optional[dict] = None              # rate_subs.py:446-485
    cheb: Optional[dict] = None              # rate_subs.py:138-223
    spec: List[Tuple[int, float]] = field(default_factory=list)  # (k, nu)


@dataclass
class MechIR:
    name: str
    nsp: int
    nr: int
    nrev: int
    npd: int
    inv_mw: List[float]
    rxns: List[RxnIR]
    species_names: List[str]
    nasa: List[Tuple[float, list, list]] = field(default_factory=list)
    mw: List[float] = field(default_factory=list)        # {:.16e} literals
    ru_mw: List[float] = field(default_factory=list)     # RU / mw, {:.16e}


def _eff_terms(rxn):
    """(species, signed eff-1) as the reference prints them (:901-915)."""
    out = []
    for isp, eff in rxn.thd_body_eff:
        if eff == 1.0:
            continue
        if eff > 1.0:
            out.append((isp, emitted(py2_str(eff - 1.0))))
        else:
            out.append((isp, -emitted(py2_str(1.0 - eff))))
    return out


def _kc_ir(rxn, specs):
    """Lumped NASA-7 Kc coefficients (rate_subs.py:410-431, 513-662)."""
    groups = {}
    order = []
    sum_nu = 0

    def add(sp, scale):
        def one(a):
            raw = [a[6] - a[0], a[0] - 1.0, a[1] / 2.0, a[2] / 6.0,
                   a[3] / 12.0, a[4] / 20.0, a[5]]
            return [x * scale for x in raw]
        lo, hi = one(sp.lo), one(sp.hi)
        key = sp.Trange[1]
        if key in groups:
            plo, phi = groups[key]
            lo = [lo[i] + plo[i] for i in range(7)]
            hi = [hi[i] + phi[i] for i in range(7)]
        else:
            order.append(key)
        groups[key] = (lo, hi)

    for i, isp in enumerate(rxn.prod):
        if isp in rxn.reac:
            nu = rxn.prod_nu[i] - rxn.reac_nu[rxn.reac.index(isp)]
        else:
            nu = rxn.prod_nu[i]
        if nu == 0:
            continue
        sum_nu += nu
        add(specs[isp], nu * 1.0)
    for i, isp in enumerate(rxn.reac):
        if isp in rxn.prod:
            continue
        nu = rxn.reac_nu[i]
        sum_nu -= nu
        add(specs[isp], nu * -1.0)
    out = []
    for key in order:
        lo, hi = groups[key]
        out.append((emitted(py2_str(key)),
                    [emitted(fmt_e16(x)) for x in lo],
                    [emitted(fmt_e16(x)) for x in hi]))
    return out, emitted(fmt_e16((chem.PA / chem.RU) ** sum_nu)), sum_nu


def build_ir(specs, reacs, name='mech'):
    """Tables of the mechanism as the reference's emitters would print them."""
    nsp = len(specs)
    rev_reacs = [j for j, r in enumerate(reacs) if r.rev]
    pd_reacs = [j for j, r in enumerate(reacs) if r.thd_body or r.pdep]
    rxns = []
    for j, r in enumerate(reacs):
        x = RxnIR(index=j, reac=list(zip(r.reac, r.reac_nu)),
                  prod=list(zip(r.prod, r.prod_nu)),
                  arr=arr_ir(r.A, r.b, r.E))
        if r.plog:
            rows = r.plog_par
            x.plog = dict(
                p=[emitted('{:.4e}'.format(v[0])) for v in rows],
                arr=[arr_ir(v[1], v[2], v[3]) for v in rows],
                lnp=[emitted(fmt_e16(math.log(v[0]))) for v in rows],
                dlnp=[emitted(fmt_e16(math.log(rows[i + 1][0]) -
                                      math.log(rows[i][0])))
                      for i in range(len(rows) - 1)])
        if r.cheb:
            t0, t1 = r.cheb_tlim
            p0, p1 = r.cheb_plim
            x.cheb = dict(
                nt=r.cheb_n_temp, np=r.cheb_n_pres,
                par=[[emitted(fmt_e8(float(v))) for v in row]
                     for row in r.cheb_par],
                tsum=emitted(fmt_e8(1.0 / t0 + 1.0 / t1)),
                tsub=emitted(fmt_e8(1.0 / t1 - 1.0 / t0)),
                psum=emitted(fmt_e8(math.log10(p0) + math.log10(p1))),
                psub=emitted(fmt_e8(math.log10(p1) - math.log10(p0))))
        if r.rev:
            x.rev = True
            x.rev_index = rev_reacs.index(j)
            if r.rev_par:
                x.rev_arr = arr_ir(*r.rev_par[0:3])
            else:
                x.kc_groups, x.kc_pref, x.kc_dnu = _kc_ir(r, specs)
        if r.thd_body or r.pdep:
            x.pd_index = pd_reacs.index(j)
            x.eff = _eff_terms(r)
        if r.thd_body:
            x.thd = True
        elif r.pdep:
            if r.pdep_sp != '':
                x.pdep_sp = int(r.pdep_sp)
            x.low = arr_ir(*r.low[0:3]) if r.low else None
            x.high = arr_ir(*r.high[0:3]) if r.high else None
            x.falloff_low = bool(r.low)
            x.falloff_high = bool(r.high) and not r.low
            if r.troe:
                tp = r.troe_par
                d = dict(one_minus_a=emitted(fmt_e8(1.0 - tp[0])),
                         a=emitted(fmt_e8(tp[0])),
                         t3=emitted(fmt_e8(abs(tp[1]))), t3_pos=tp[1] > 0.0,
                         t1=emitted(fmt_e8(abs(tp[2]))), t1_pos=tp[2] > 0.0,
                         t2=None, t2_pos=True)
                if len(tp) == 4 and tp[3] != 0.0:
                    d['t2'] = emitted(fmt_e8(abs(tp[3])))
                    d['t2_pos'] = tp[3] > 0.0
                x.troe = d
                x.falloff = 'troe'
            elif r.sri:
                sp_ = r.sri_par
                d = dict(a=emitted(fmt_g6(sp_[0])),
                         b=emitted(fmt_g6(abs(sp_[1]))), b_pos=sp_[1] > 0.0,
                         c=emitted(fmt_g6(abs(sp_[2]))), c_pos=sp_[2] > 0.0,
                         d=None, e=None)
                if len(sp_) == 5 and sp_[3] != 1.0 and sp_[4] != 0.0:
                    d['d'] = emitted(fmt_e8(sp_[3]))
                    d['e'] = emitted(fmt_g6(sp_[4]))
                x.sri = d
                x.falloff = 'sri'
            else:
                x.falloff = 'lind'
        for k in sorted(set(r.reac + r.prod)):
            nu = get_nu(k, r)
            if nu != 0:
                x.spec.append((k, nu))
        rxns.append(x)
    return MechIR(name=name, nsp=nsp, nr=len(reacs), nrev=len(rev_reacs),
                  npd=len(pd_reacs),
                  inv_mw=[emitted(fmt_e16(1.0 / sp.mw)) for sp in specs],
                  rxns=rxns, species_names=[sp.name for sp in specs],
                  nasa=[(emitted(py2_str(sp.Trange[1])),
                         [float(v) for v in sp.lo], [float(v) for v in sp.hi])
                        for sp in specs],
                  mw=[emitted(fmt_e16(sp.mw)) for sp in specs],
                  ru_mw=[emitted(fmt_e16(chem.RU / sp.mw)) for sp in specs])


# ---------------------------------------------------------------------------
# Source generation
# ---------------------------------------------------------------------------

def _lit(x):
    """C++ literal that parses to exactly ``x``."""
    if x == 0.0:
        return '0.0'
    r = repr(float(x))
    if 'e' not in r and '.' not in r and 'n' not in r:
        r += '.0'
    return r


class _Consts(object):
    """Where mechanism numbers live in the kernel text.

    ``table``: one ``__constant__`` array read with wave-uniform scalar
    loads; ``literal``: immediates in the instruction stream.
    """

    def __init__(self, mode='table', symbol='crs_K'):
        self.vals = []
        self.index = {}
        self.symbol = symbol
        self.mode = mode
        # kernels read the table through a restrict pointer argument so
        # the compiler keeps them as (batched) scalar loads
        self.ref = 'K'

    def __call__(self, x):
        if self.mode == 'literal':
            return _lit(x)
        key = float(x).hex()
        if key not in self.index:
            self.index[key] = len(self.vals)
            self.vals.append(float(x))
        return '{}[{}]'.format(self.ref, self.index[key])

    def definition(self):
        if self.mode == 'literal':
            return '__constant__ double {}[1] = {{0.0}};\n'.format(self.symbol)
        body = ',\n  '.join(_lit(v) for v in self.vals) or '0.0'
        return ('__constant__ double {}[{}] = {{\n  {}\n}};\n'.format(
            self.symbol, max(len(self.vals), 1), body))


def _arr_expr(a, K, logT='logT', T='T'):
    """FP64 expression of a rate constant, term order as the C text."""
    if not a.has_exp:
        s = K(a.pre)
        for _ in range(a.nT):
            s = '({} * {})'.format(s, T)
        return s
    terms = []
    if a.logA is not None:
        terms.append(K(a.logA))
    if a.b is not None:
        bt = '{} * {}'.format(K(a.b), logT)
        terms.append(bt)
    inner = ' + '.join(terms)
    if a.E is not None:
        et = '({} / {})'.format(K(a.E), T)
        inner = '{} - {}'.format(inner, et) if inner else '-' + et
    s = 'exp({})'.format(inner)
    if a.pre is not None:
        s = '{} * {}'.format(K(a.pre), s)
    return s


def _conc_prod(terms, C='C'):
    """Concentration product prefix 'C[a] * C[a] * C[b] * ' (left to right)."""
    s = ''
    for k, nu in terms:
        if is_integer(nu):
            s += '{}[{}] * '.format(C, k) * int(nu)
        else:
            s += 'crs::pow64({}[{}], {}) * '.format(
                C, k, _lit(emitted(py2_str(nu))))
    return s


def _kc_expr(x, K, logT='logT', T='T'):
    """Lines computing Kc for reaction x, as rate_subs.py:593-662."""
    lines = []
    first = True
    for tmid, lo, hi in x.kc_groups:
        def poly(c):
            return ('({} + {} * {lt} + {T} * ({} + {T} * ({} + {T} * ({} + {} '
                    '* {T}))) - {} / {T})'.format(*[K(v) for v in c], lt=logT,
                                                  T=T))
        sel = '({T} <= {tm}) ? {lo} : {hi}'.format(T=T, tm=_lit(tmid),
                                                   lo=poly(lo), hi=poly(hi))
        lines.append(('Kc = ' if first else 'Kc += ') + '(' + sel + ');')
        first = False
    lines.append('Kc = {} * exp(Kc);'.format(K(x.kc_pref)))
    return lines


def _eff_expr(x, K, base='m', C='C'):
    s = base
    for k, e in x.eff:
        if e >= 0:
            s += ' + {} * {}[{}]'.format(K(e), C, k)
        else:
            s += ' - {} * {}[{}]'.format(K(-e), C, k)
    return s


def _pres_mod_lines(x, K, kf_main='kf', C='C'):
    """Lines ending in 'pm = ...;' for reaction x (rate_subs.py:883-1064)."""
    out = []
    if x.thd:
        out.append('pm = {};'.format(_eff_expr(x, K, C=C)))
        return out
    if x.pdep_sp is None:
        out.append('const double thd = {};'.format(_eff_expr(x, K, C=C)))
        coll = 'thd'
    else:
        coll = '{}[{}]'.format(C, x.pdep_sp)
    k0 = _arr_expr(x.low, K) if x.low is not None else kf_main
    kinf = _arr_expr(x.high, K) if x.high is not None else kf_main
    out.append('const double k0 = {};'.format(k0))
    out.append('const double kinf = {};'.format(kinf))
    out.append('const double Pr = k0 * {} / kinf;'.format(coll))
    tail = ''
    if x.falloff_low:
        tail = ' * Pr / (1.0 + Pr)'
    elif x.falloff_high:
        tail = ' / (1.0 + Pr)'
    if x.falloff == 'troe':
        t = x.troe
        e1 = ('exp(-T / {})' if t['t3_pos'] else 'exp(T / {})').format(K(t['t3']))
        e2 = ('exp(-T / {})' if t['t1_pos'] else 'exp(T / {})').format(K(t['t1']))
        fc = '{} * {} + {} * {}'.format(K(t['one_minus_a']), e1, K(t['a']), e2)
        if t['t2'] is not None:
            fc += ' + ' + ('exp(-{} / T)' if t['t2_pos'] else
                           'exp({} / T)').format(K(t['t2']))
        out.append('const double logFcent = log10(fmax({}, 1.0e-300));'.format(fc))
        out.append('const double lPr = log10(fmax(Pr, 1.0e-300));')
        out.append('const double A = lPr - 0.67 * logFcent - 0.4;')
        out.append('const double B = 0.806 - 1.1762 * logFcent - 0.14 * lPr;')
        out.append('pm = exp10(logFcent / (1.0 + A * A / (B * B))){};'.format(tail))
    elif x.falloff == 'sri':
        s = x.sri
        out.append('const double lPr = log10(fmax(Pr, 1.0e-300));')
        out.append('const double X = 1.0 / (1.0 + lPr * lPr);')
        e1 = ('exp(-{} / T)' if s['b_pos'] else 'exp({} / T)').format(K(s['b']))
        e2 = ('exp(-T / {})' if s['c_pos'] else 'exp(T / {})').format(K(s['c']))
        f = 'pow({} * {} + {}, X)'.format(K(s['a']), e1, e2)
        if s['d'] is not None:
            f += ' * {} * pow(T, {})'.format(K(s['d']), K(s['e']))
        out.append('pm = {}{};'.format(f, tail))
    else:
        if x.falloff_low:
            out.append('pm = Pr / (1.0 + Pr);')
        elif x.falloff_high:
            out.append('pm = 1.0 / (1.0 + Pr);')
        else:
            out.append('pm = 1.0;')
    return out


def _plog_lines(x, K, P='P'):
    """PLOG kf as the reference writes it (rate_subs.py:446-485)."""
    pl = x.plog
    n = len(pl['p'])
    out = ['double kf;', 'double kf2;',
           'if ({} <= {}) {{'.format(P, _lit(pl['p'][0])),
           '    kf = {};'.format(_arr_expr(pl['arr'][0], K))]
    for i in range(n - 1):
        out.append('}} else if (({P} > {a}) && ({P} <= {b})) {{'.format(
            P=P, a=_lit(pl['p'][i]), b=_lit(pl['p'][i + 1])))
        out.append('    kf = log({});'.format(_arr_expr(pl['arr'][i], K)))
        out.append('    kf2 = log({});'.format(_arr_expr(pl['arr'][i + 1], K)))
        out.append('    kf = exp(kf + (kf2 - kf) * (log({}) - {}) / {});'.format(
            P, K(pl['lnp'][i]), K(pl['dlnp'][i])))
    out.append('}} else if ({} > {}) {{'.format(P, _lit(pl['p'][-1])))
    out.append('    kf = {};'.format(_arr_expr(pl['arr'][-1], K)))
    out.append('}')
    out.append('(void)kf2;')
    return out


def _cheb_lines(x, K, P='P'):
    """Chebyshev kf as the reference writes it (rate_subs.py:138-223),
    with the CUDA branch's exp10 in place of C's pow(10.0, .)."""
    c = x.cheb
    out = ['double kf;',
           'const double Tred = ((2.0 / T) - {}) / {};'.format(K(c['tsum']),
                                                               K(c['tsub'])),
           'const double Pred = (2.0 * log10({}) - {}) / {};'.format(
               P, K(c['psum']), K(c['psub'])),
           'double cheb_temp_0 = 1;', 'double cheb_temp_1 = Pred;',
           'double dot_prod[{}];'.format(c['nt'])]
    for i in range(c['nt']):
        out.append('dot_prod[{}] = {} + Pred * {};'.format(
            i, K(c['par'][i][0]), K(c['par'][i][1])))
    one = True
    for j in range(2, c['np']):
        new, old = (1, 0) if one else (0, 1)
        out.append('cheb_temp_{o} = 2 * Pred * cheb_temp_{n} - cheb_temp_{o};'
                   .format(o=old, n=new))
        for i in range(c['nt']):
            out.append('dot_prod[{}] += {} * cheb_temp_{};'.format(
                i, K(c['par'][i][j]), old))
        one = not one
    out += ['cheb_temp_0 = 1;', 'cheb_temp_1 = Tred;',
            'kf = dot_prod[0] + Tred * dot_prod[1];']
    one = True
    for i in range(2, c['nt']):
        new, old = (1, 0) if one else (0, 1)
        out.append('cheb_temp_{o} = 2 * Tred * cheb_temp_{n} - cheb_temp_{o};'
                   .format(o=old, n=new))
        out.append('kf += dot_prod[{}] * cheb_temp_{};'.format(i, old))
        one = not one
    out.append('kf = exp10(kf);')
    return out


def _fast_exponent(a, K, log2=False):
    """ln of an Arrhenius expression (lnA + b lnT - E/T with 1/T
    precomputed; the log of the literal forms without an exponential);
    log2 of it with ``log2``."""
    if not a.has_exp or a.pre is not None:
        return ('log2({})' if log2 else 'log({})').format(_arr_expr(a, K))
    lines, xname = _fast_arr(a, K, 'tmp', log2=log2)
    return lines[0][len('const double x_tmp = '):-1]


def _plog_fast_lines(x, K, expfn):
    """PLOG kf for the fused kernels: the bracketing rows' ln k are
    interpolated in ln P directly (no exp/log round trip) and one
    exponential is taken, same thresholds as the reference."""
    pl = x.plog
    n = len(pl['p'])
    # with the branch-free exponential the rows' exponents are base 2 (the
    # interpolation is linear, so xk is too) and kf = crs::exp2_64(xk)
    base2 = expfn == 'crs::exp64'
    if base2:
        expfn = 'crs::exp2_64'
    xs = [_fast_exponent(a, K, log2=base2) for a in pl['arr']]
    out = ['double xk;',
           'if (P <= {}) {{'.format(_lit(pl['p'][0])),
           '    xk = {};'.format(xs[0])]
    for i in range(n - 1):
        out.append('}} else if (P <= {}) {{'.format(_lit(pl['p'][i + 1])))
        out.append('    const double x0 = {};'.format(xs[i]))
        out.append('    const double x1 = {};'.format(xs[i + 1]))
        out.append('    xk = x0 + (x1 - x0) * ((logP - {}) * {});'.format(
            K(pl['lnp'][i]), K(1.0 / pl['dlnp'][i])))
    out += ['} else {', '    xk = {};'.format(xs[-1]), '}',
            'const double kf = {}(xk);'.format(expfn)]
    return out


def _cheb_fast_lines(x, K, expfn):
    """Chebyshev kf for the fused kernels: reduced coordinates from 1/T and
    log10 P computed once per state, 10^s as exp(s ln 10)."""
    c = x.cheb
    out = ['const double Tred = (2.0 * iT - {}) * {};'.format(
               K(c['tsum']), K(1.0 / c['tsub'])),
           'const double Pred = (2.0 * log10P - {}) * {};'.format(
               K(c['psum']), K(1.0 / c['psub']))]
    # sum_i T_i(Tred) sum_j c_ij T_j(Pred), recurrences as the reference
    out.append('double tp0 = 1.0, tp1 = Pred;')
    dots = ['{} + Pred * {}'.format(K(c['par'][i][0]), K(c['par'][i][1]))
            for i in range(c['nt'])]
    out += ['double d{} = {};'.format(i, d) for i, d in enumerate(dots)]
    one = True
    for j in range(2, c['np']):
        new, old = (1, 0) if one else (0, 1)
        out.append('tp{o} = 2.0 * Pred * tp{n} - tp{o};'.format(o=old, n=new))
        for i in range(c['nt']):
            out.append('d{i} += {c} * tp{o};'.format(i=i, c=K(c['par'][i][j]),
                                                     o=old))
        one = not one
    out += ['double tt0 = 1.0, tt1 = Tred;', 'double lk = d0 + Tred * d1;']
    one = True
    for i in range(2, c['nt']):
        new, old = (1, 0) if one else (0, 1)
        out.append('tt{o} = 2.0 * Tred * tt{n} - tt{o};'.format(o=old, n=new))
        out.append('lk += d{} * tt{};'.format(i, old))
        one = not one
    if expfn == 'crs::exp64':
        out.append('const double kf = crs::exp2_64(lk * {});'.format(_LOG2_10))
    else:
        out.append('const double kf = {}(lk * 2.302585092994046);'.format(
            expfn))
    return out


def _rate_lines(x, K, C='C'):
    """Lines for kf, f and (if reversible) Kc and r of reaction x."""
    if x.plog:
        out = _plog_lines(x, K)
    elif x.cheb:
        out = _cheb_lines(x, K)
    else:
        out = ['const double kf = {};'.format(_arr_expr(x.arr, K))]
    out.append('const double f = {}kf;'.format(_conc_prod(x.reac, C)))
    if x.rev:
        if x.rev_arr is not None:
            out.append('const double r = {}{};'.format(
                _conc_prod(x.prod, C), _arr_expr(x.rev_arr, K)))
        else:
            out.append('double Kc;')
            out.extend(_kc_expr(x, K))
            out.append('const double r = {}kf / Kc;'.format(
                _conc_prod(x.prod, C)))
    return out


def _spec_lines(x, seen, W='w', f='f', r='r', pm='pm'):
    """Species-rate updates of reaction x (rate_subs.py:1162-1217)."""
    out = []
    for k, nu in x.spec:
        anu = abs(nu)
        coef = ''
        if anu != 1.0:
            # integral: '{}'.format(float(nu)); else '{:3}' (rate_subs.py:1189-1193)
            coef = '{} * '.format(_lit(emitted(py2_str(float(anu)))))
        term = '({} - {})'.format(f, r) if x.rev else f
        if x.pd_index is not None:
            term += ' * ' + pm
        if seen[k]:
            op = '-' if nu < 0 else '+'
            out.append('{W}[{k}] = {W}[{k}] {op} {c}{t};'.format(
                W=W, k=k, op=op, c=coef, t=term))
        else:
            neg = '-' if nu < 0 else ''
            if coef:
                out.append('{W}[{k}] = {n}{c}{t};'.format(
                    W=W, k=k, n=neg, c=coef, t=term))
            else:
                out.append('{W}[{k}] = {n}{t};'.format(W=W, k=k, n=neg, t=term))
            seen[k] = True
    return out


# marks the end of the concentration lines in a fused body
_CONC_DONE = '// concentrations complete'


def _conc_lines(ir, K, y='y', C='C'):
    """eval_conc body (rate_subs.py:1292-1391), same operation order."""
    nsp = ir.nsp
    out = []
    s = ' + '.join('{}[{}]'.format(y, k) for k in range(nsp - 1))
    out.append('const double y_N = 1.0 - ({});'.format(s))
    s = ' + '.join('({}[{}] * {})'.format(y, k, K(ir.inv_mw[k]))
                   for k in range(nsp - 1))
    s += ' + (y_N * {})'.format(K(ir.inv_mw[nsp - 1]))
    out.append('double mw_avg = {};'.format(s))
    out.append('mw_avg = 1.0 / mw_avg;')
    out.append('const double rho = P * mw_avg / ({} * T);'.format(
        _lit(emitted(fmt_e8(chem.RU)))))
    out.append('double {}[{}];'.format(C, nsp))
    for k in range(nsp - 1):
        out.append('{C}[{k}] = rho * {y}[{k}] * {w};'.format(
            C=C, k=k, y=y, w=K(ir.inv_mw[k])))
    out.append('{C}[{k}] = rho * y_N * {w};'.format(
        C=C, k=nsp - 1, w=K(ir.inv_mw[nsp - 1])))
    return out


def _indent(lines, n):
    pad = ' ' * n
    return ''.join(pad + l + '\n' for l in lines)


# Measured on MI355X (2^20 states):
# * immediates beat a scalar-loaded table 2x at one wave per SIMD (each
#   s_load is waited for at once, nothing hides it);
# * species-level Kc removes one exponential per reversible reaction;
# * with every per-species array in registers the kernels spill (388-584
#   bytes per lane at 53 species, 3.5-4 KB at 111).  Parking one array in
#   per-lane LDS rows removes the spills at 53 species (FP64 1.54 -> 1.37 ms,
#   mixed 1.07 -> 0.88 ms with the H/G net-rate form below) and cuts them
#   by 60 % at 111 species (FP64 7.2 -> 4.2 ms, mixed 7.9 -> 3.0 ms);
# * the retiring reaction order and the branch-free exp64 add a few %;
# * base-2 Arrhenius exponents (exp2_64 / exp2_mixed) cut 2 % of the
#   instructions, within timing noise;
# * waves spent 27 % of their time in s_waitcnt (SQ_WAIT_ANY), mostly on
#   each reaction block's per-lane LDS reads; issuing them lds_prefetch
#   blocks ahead (_prefetch_lane_lds) fixes most of it;
# * the library log10 / exp10 / exp / pow of the falloff blocks were ~20 %
#   of the FP64 instructions (log10 alone ~90 each); crs::log64, exp64,
#   exp2_64 and rcp64 instead: FP64 1.23 -> 1.05 ms at 53 species, 4.05 ->
#   3.20 ms at 111 (the mixed kernel already used FP32 hardware there);
# * the NASA coefficients as a per-block LDS table selected per lane
#   (nasa='lds': 7 FMAs and 4 LDS reads per species instead of 14 FMAs and
#   28 scalar moves), read 4 species ahead: the default where the H/G form
#   fits (at 111 species it costs concentration rows: 3.22 -> 3.45 ms);
# * the T-only work (NASA basis, the G_k) ahead of eval_conc's sums, which
#   wait for the mass-fraction loads (conc_late): on where the H/G form
#   fits (at 111 species FP64 3.20 -> 3.40 ms: off there).
#   Interleaved A/B at 53 species (5 rounds, median; FP64 / mixed ms):
#   defaults 0.941 / 0.667; without the LDS prefetch 1.024 / 0.774, with 1
#   block instead of 3 0.944 / 0.681 (4-8 blocks within 1 %); without the
#   NASA table 0.976 / 0.704; without conc_late 0.944 / 0.680;
# * exp64 / exp2_64 with Estrin's scheme (exp_poly='estrin', dependency
#   chain 5 instead of 10, 3 more multiplies): 0.920 -> 0.967 ms FP64 at 53
#   species, 3.30 -> 3.49 ms at 111; the scheduler already interleaves the
#   reactions' Horner chains, so the extra instructions only cost;
# * the reaction-block literals as a wave-uniform LDS table
#   (consts_rx='lds' / 'lds_fp64', _block_consts_to_lds) remove 2/3 of the
#   scalar moves and nearly all scalar spills but cost as many LDS reads
#   and AGPR copies: 4 % faster FP64 before the changes above, 1-5 %
#   slower after them; not the default.
# * two waves per SIMD (waves_per_simd=2): the kernel needs 200 VGPRs and
#   no AGPRs, so only the LDS rows (115.7 KB per block at 53 species) held
#   it to one.  39 G rows per wave (2 blocks of 4 waves in 160 KB, no NASA
#   table), the G_k of the 14 species that retire first in registers
#   (g_pick='retire'), the LDS species' G_k read back for H_k
#   (g_reload), rows read 1 block ahead: 138 VGPRs spill (144 scratch
#   accesses per wave in 24k instructions), yet the second wave hides the
#   latencies one wave could not.  Interleaved at 53 species, FP64 /
#   mixed ms: one wave 0.919 / 0.649 against 0.757-0.780 / 0.568-0.579;
#   36 rows with the NASA table 0.822 / 0.656; three waves (25 rows, 520
#   VGPRs spilled) 1.546 / 1.508; dydt at two waves 1.21 against 1.13 (it
#   stays at one).  At 111 species two waves are 1.7x slower.
# 'auto' picks the H/G form with G in LDS when every species fits the LDS
# budget (78 rows), otherwise species Kc with 78 concentration rows in LDS;
# waves_per_simd='auto' takes two waves where 40-56 species would fill the
# LDS at one (the range measured: 53 species).
DEFAULT_OPTS = dict(consts='literal', sched_barrier=False, early_store=True,
                    fused='fast', kc='auto', exp='crs', order='retire',
                    nasa='auto', waves_per_simd='auto')

_LDS_PER_CU = 160 * 1024


def _lane_rows(block, waves_per_simd, table_bytes):
    """Per-lane FP64 LDS rows per wave that let the wdot kernels reach
    ``waves_per_simd`` waves per SIMD with ``block``-thread blocks.

    Every wave of a block owns rows x 512 bytes; a block also holds
    ``table_bytes`` of shared tables and 512 bytes of slack.  The CU's
    160 KB is split over the blocks that must be resident together: enough
    for 4 x waves_per_simd waves, and at least one block."""
    wpb = block // 64
    waves_cu = max(4 * waves_per_simd, wpb)
    blocks_cu = max(1, waves_cu // wpb)
    per_block = _LDS_PER_CU // blocks_cu - table_bytes - 512
    return max(0, min(78, per_block // (wpb * 512)))


def resolve_opts(nsp, opts=None):
    """Emitter options with 'auto' choices made for an NSP-species mechanism.

    LDS budget at one wave per SIMD: 160 KB per CU for 4 waves.  Per-lane
    rows take 512 B per wave each; the NASA coefficient table (option
    nasa='lds') takes 128 B per species per block.  nasa='auto' takes the
    table where the H/G form fits beside it.  waves_per_simd='auto' gives
    40-56 species two waves per SIMD instead: 39 rows per wave, the other
    G_k in registers, no table.  Row counts follow the wdot block size
    (option ``block``, 256 threads by default; _lane_rows).

    The block-wide LDS layouts of earlier measurements (g_store='lds' or
    'lds_sync', c_store='lds': one [rows][block] array shared by the block)
    are refused: they need a barrier after the per-state early exits of
    partial blocks and of the subset kernel, and they lost to the per-lane
    rows.
    """
    o = dict(DEFAULT_OPTS, **(opts or {}))
    if o.get('contraction', 'sparse') != 'sparse':
        raise ValueError(
            "contraction={!r}: only the sparse per-lane accumulation is "
            "built.  The FP32-accumulate MFMA GEMM over nu^T was measured "
            "2.8x slower and overflows FP32 (reverse rates reach 4e56), so it "
            "was removed".format(o.get('contraction')))
    if o.get('g_store') in ('lds', 'lds_sync') or o.get('c_store') == 'lds':
        raise ValueError("block-wide LDS layouts (g_store='lds'/'lds_sync', "
                         "c_store='lds') are not supported; use 'lane_lds'")
    block = int(o.get('block') or 256)
    if block % 64 or not 64 <= block <= 1024:
        raise ValueError('block must be a multiple of 64 in [64, 1024]')

    def rows(table, wps=1):
        return _lane_rows(block, wps, table)
    if o.get('waves_per_simd') == 'auto':
        # only the per-lane G-row layout, and no NASA table (it would not
        # leave room for a second block)
        rows2 = rows(0, 2)
        two = (rows2 < nsp <= 56 and 'lds_species' not in o and
               o.get('kc') in ('auto', 'species_hg') and
               o.get('g_store', 'lane_lds') == 'lane_lds' and
               o.get('nasa') != 'lds')
        o['waves_per_simd'] = 2 if two else None
        if two:
            o['lds_species'] = rows2
            o['kc'] = 'species_hg'
            o.setdefault('g_store', 'lane_lds')
            o.setdefault('g_pick', 'retire')
            o.setdefault('g_reload', True)
            o.setdefault('lds_prefetch', 1)
            if o.get('nasa') == 'auto':
                o['nasa'] = None
    wps = int(o.get('waves_per_simd') or 1)
    if o.get('nasa') == 'auto':
        fits = nsp <= rows(128 * nsp, wps) and o.get('kc') in (
            'auto', 'species_hg')
        o['nasa'] = 'lds' if fits else None
    if 'lds_species' not in o:
        o['lds_species'] = rows(128 * nsp if o.get('nasa') == 'lds' else 0,
                                wps)
    if o.get('kc') == 'auto':
        if nsp <= o['lds_species']:
            o['kc'] = 'species_hg'
            o.setdefault('g_store', 'lane_lds')
        else:
            o['kc'] = 'species'
            o.setdefault('c_store', 'lane_lds')
    o.setdefault('lds_prefetch', 3 if o.get('g_store') == 'lane_lds' else 1)
    o.setdefault('conc_late', o.get('g_store') == 'lane_lds')
    if o.get('nasa') == 'lds':
        o.setdefault('nasa_prefetch', 4)
    return o


def _last_touch(ir, rxns=None):
    last = {}
    for x in (ir.rxns if rxns is None else rxns):
        for k, _nu in x.spec:
            last[k] = x.index
    return last


def _species_of(x):
    return ({k for k, _nu in x.spec} | {k for k, _nu in x.reac} |
            {k for k, _nu in x.prod} | {k for k, _e in x.eff} |
            ({x.pdep_sp} if x.pdep_sp is not None else set()))


def _retiring_order(ir):
    """Reaction order that retires species early (fused kernels only).

    Greedy: repeatedly take the species, among those already touched
    (then any), with the fewest reactions left, and emit all of them.  A
    species' accumulator is live from its first to its last reaction and
    its concentration until its last use, so finishing species early cuts
    the mean live set by about a third on the synthetic mechanisms.  The
    sums of each species change order, nothing else.
    """
    left = set(range(len(ir.rxns)))
    by_sp = {}
    for j, x in enumerate(ir.rxns):
        for k in _species_of(x):
            by_sp.setdefault(k, set()).add(j)
    order, active = [], set()
    while left:
        cand = [k for k in sorted(by_sp) if by_sp[k] & left]
        k = min(cand, key=lambda k: (k not in active, len(by_sp[k] & left), k))
        for j in sorted(by_sp[k] & left):
            order.append(ir.rxns[j])
            left.discard(j)
            active |= _species_of(ir.rxns[j])
    return order


def _fused_f64_lines(ir, K, opts):
    """Per-state body of the fused FP64 kernel.

    Reaction by reaction: rate constant, forward and reverse rates,
    pressure modification and the species updates, in the reference's
    reaction order so every species sum keeps the reference's term order.
    A species' rate is stored as soon as its last reaction is done and a
    scheduling barrier closes every reaction, which bounds the live set to
    the concentrations plus the species still being accumulated.
    """
    nsp = ir.nsp
    ru = _lit(emitted(fmt_e8(chem.RU)))
    lines = ['const double T = T_[i];', 'const double P = P_[i];',
             'double y[{}];'.format(max(nsp - 1, 1))]
    lines += ['y[{k}] = Y[{k} * ld + i];'.format(k=k) for k in range(nsp - 1)]
    lines += _conc_lines(ir, K)
    lines += ['const double logT = log(T);',
              'const double m = P / ({} * T);'.format(ru),
              '(void)m;',
              'double w[{}];'.format(nsp)]
    last = _last_touch(ir)
    seen = [False] * nsp
    for k in range(nsp):
        if k not in last:
            lines.append('W[{k} * ld + i] = 0.0;'.format(k=k))
    for x in ir.rxns:
        inner = _rate_lines(x, K)
        if x.pd_index is not None:
            inner.append('double pm;')
            inner += _pres_mod_lines(x, K)
        inner += _spec_lines(x, seen)
        lines.append('{  // rxn %d' % x.index)
        lines += ['    ' + l for l in inner]
        lines.append('}')
        if opts.get('early_store'):
            for k, _nu in x.spec:
                if last.get(k) == x.index:
                    lines.append('W[{k} * ld + i] = w[{k}];'.format(k=k))
        if opts.get('sched_barrier'):
            lines.append('__builtin_amdgcn_sched_barrier(0);')
    if not opts.get('early_store'):
        lines += ['W[{k} * ld + i] = w[{k}];'.format(k=k)
                  for k in range(nsp) if k in last]
    return lines


_BASIS = ('1.0', 'logT', 'T', 'T2', 'T3', 'T4', 'iT')


_LOG2E = 1.4426950408889634


def _fast_arr(a, K, name, expfn='exp', log2=False):
    """Lines defining rate constant ``name`` (and its exponent ``x_<name>``).

    Returns (lines, exponent_or_None).  The exponent is lnA + b logT - E/T
    with 1/T precomputed once per state; forms without an exponential
    (literal A, A*T^n) are kept as the reference prints them.  With
    ``log2`` the exponent is formed in base 2 (log2 A + b log2 T -
    (E log2 e) / T, constants scaled once here) for an exponential that
    takes a base-2 argument, saving its conversion multiply.
    """
    if not a.has_exp:
        return ['const double {} = {};'.format(name, _arr_expr(a, K))], None
    sc = _LOG2E if log2 else 1.0
    lt = 'log2T' if log2 else 'logT'
    # the T-dependent terms first and ln A added last: every instruction
    # then takes one constant (a VALU operation reads one scalar operand;
    # fma(b, log T, ln A) would first copy ln A into a VGPR pair)
    x = ''
    if a.b is not None:
        x = '{} * {}'.format(K(a.b), lt)
    if a.E is not None:
        x = ('{} - {} * iT'.format(x, K(a.E * sc)) if x else
             '-{} * iT'.format(K(a.E * sc)))
    if a.logA is not None:
        x = '({}) + {}'.format(x, K(a.logA * sc)) if x else K(a.logA * sc)
    out = ['const double x_{} = {};'.format(name, x)]
    if a.pre is not None:
        out.append('const double {n} = {p} * {e}(x_{n});'.format(
            n=name, p=K(a.pre), e=expfn))
        return out, None
    out.append('const double {n} = {e}(x_{n});'.format(n=name, e=expfn))
    return out, 'x_' + name


def _fast_lnkc(x, K):
    """ln Kc as a sum over the per-state masked NASA basis (no selects)."""
    terms = []
    for g, (tmid, lo, hi) in enumerate(x.kc_groups):
        for half, c in (('l', lo), ('h', hi)):
            coef = list(c[0:6]) + [-c[6]]
            if g == 0:
                coef[0] = coef[0] + math.log(x.kc_pref)
            for bi, cv in enumerate(coef):
                if cv != 0.0:
                    terms.append('{} * p{}{}_{}'.format(K(cv), half, g, bi))
    return ' + '.join(terms) if terms else '0.0'


_LOG10E = '0.4342944819032518'
_LOG2_10 = '3.321928094887362'
_LOG10_2 = '0.3010299956639812'


def _exp_t(pos, c, K):
    """exp(-T c) (exp(T c) when not pos): crs::exp2_64 on the constant
    pre-scaled by log2 e while |T c| stays below 1e8 for T up to 10^4 K
    (the reduction is exact there and 2^t over/underflows like exp beyond
    709), the library exp otherwise (e.g. T*** = 1e-30 K)."""
    if abs(c) * 1.0e4 <= 1.0e8:
        return 'crs::exp2_64({}T * {})'.format('-' if pos else '',
                                               K(c * _LOG2E))
    return 'exp({}T * {})'.format('-' if pos else '', K(c))


def _exp_it(pos, c, K):
    """exp(-c / T) as 2^(-(c log2 e) iT); crs::exp2_64 while |c| / T stays
    below 1e8 for T down to 1 K (and below 7e4 so that the term cannot
    underflow to zero at T >= 100 K), the library exp otherwise."""
    if abs(c) <= 7.0e4:
        return 'crs::exp2_64({}{} * iT)'.format('-' if pos else '',
                                                K(c * _LOG2E))
    return 'exp({}{} * iT)'.format('-' if pos else '', K(c))


def _fast_pm_lines(x, K, kf_x, arr2=None):
    """Pressure modification of reaction x into 'pm' (fast FP64 form).

    With ``arr2`` (a base-2 exponential) the Arrhenius exponents, kf's
    included, are base 2."""
    out = []
    if x.thd:
        out.append('const double pm = {};'.format(_eff_expr(x, K)))
        return out
    if x.pdep_sp is None:
        out.append('const double thd = {};'.format(_eff_expr(x, K)))
        coll = 'thd'
    else:
        coll = 'C[{}]'.format(x.pdep_sp)
    # k0 / kinf: whichever limit is not given is the main Arrhenius (kf)
    if x.low is not None:
        l0, x0 = (_fast_arr(x.low, K, 'k0', arr2, log2=True) if arr2 else
                  _fast_arr(x.low, K, 'k0'))
        out += l0
        num, xnum = 'k0', x0
    else:
        num, xnum = 'kf', kf_x
    if x.high is not None:
        l1, x1 = (_fast_arr(x.high, K, 'kinf', arr2, log2=True) if arr2 else
                  _fast_arr(x.high, K, 'kinf'))
        out += l1
        den, xden = 'kinf', x1
    else:
        den, xden = 'kf', kf_x
    if xnum is not None and xden is not None:
        out.append('const double Pr = {} * {}({} - {});'.format(
            coll, arr2 or 'exp', xnum, xden))
    else:
        out.append('const double Pr = {} * {} / {};'.format(num, coll, den))
    tail = ''
    if x.falloff_low:
        tail = ' * Pr * crs::rcp64(1.0 + Pr)'
    elif x.falloff_high:
        tail = ' * crs::rcp64(1.0 + Pr)'
    if x.falloff == 'troe':
        t = x.troe
        e1 = _exp_t(t['t3_pos'], 1.0 / t['t3'], K)
        e2 = _exp_t(t['t1_pos'], 1.0 / t['t1'], K)
        fc = '{} * {} + {} * {}'.format(K(t['one_minus_a']), e1, K(t['a']), e2)
        if t['t2'] is not None:
            fc += ' + ' + _exp_it(t['t2_pos'], t['t2'], K)
        # Fcent <= 3 is finite; an infinite Pr makes pm NaN in the reference
        # too (inf / inf in A^2 / B^2), so neither log needs the inf guard.
        # Both logs in base 2 (crs::log2_64: the hardware seed refined on the
        # exp2 polynomial, absolute error ~1e-15, which is what A and B
        # need), and 10^(logFcent / (1 + A^2 / B^2)) =
        # 2^(log2 Fcent B^2 / (B^2 + A^2))
        out.append('const double l2Fc = crs::log2_64(fmax({}, 1.0e-300));'
                   .format(fc))
        out.append('const double logFcent = l2Fc * {};'.format(_LOG10_2))
        out.append('const double lPr = crs::log2_64(fmax(Pr, 1.0e-300)) * '
                   '{};'.format(_LOG10_2))
        out.append('const double A = lPr - 0.67 * logFcent - 0.4;')
        out.append('const double B = 0.806 - 1.1762 * logFcent - 0.14 * lPr;')
        out.append('const double pm = crs::exp2_64(l2Fc * (B * B) * '
                   'crs::rcp64(B * B + A * A)){};'.format(tail))
    elif x.falloff == 'sri':
        sr = x.sri
        out.append('const double lPr = crs::log64(fmax(Pr, 1.0e-300)) * '
                   '{};'.format(_LOG10E))
        out.append('const double X = crs::rcp64(1.0 + lPr * lPr);')
        e1 = _exp_it(sr['b_pos'], sr['b'], K)
        e2 = _exp_t(sr['c_pos'], 1.0 / sr['c'], K)
        base = '{} * {} + {}'.format(K(sr['a']), e1, e2)
        if sr['a'] > 0.0 and e1.startswith('crs::') and e2.startswith('crs::'):
            # a e^(-b/T) > 0 at every T >= 100 K here, so ln(base) is finite
            f = 'crs::exp64(X * crs::log64({}))'.format(base)
        elif sr['a'] > 0.0:
            f = 'exp(X * log({}))'.format(base)
        else:
            f = 'pow({}, X)'.format(base)
        if sr['d'] is not None:
            f += ' * {} * crs::exp64({} * logT)'.format(K(sr['d']), K(sr['e']))
        out.append('const double pm = {}{};'.format(f, tail))
    else:
        if x.falloff_low:
            out.append('const double pm = Pr * crs::rcp64(1.0 + Pr);')
        elif x.falloff_high:
            out.append('const double pm = crs::rcp64(1.0 + Pr);')
        else:
            out.append('const double pm = 1.0;')
    return out


def _species_kc_ok(x):
    """Species-level Kc applies when every net coefficient is integral."""
    return (x.rev and x.rev_arr is None and
            all(float(nu).is_integer() for _k, nu in x.spec))


def _species_kc_inverse(x, store_inverse=False, gname=None):
    """1/Kc from per-species G_k = exp(S_k/R - H_k/RT) and (RU T / PA)^dnu.

    Kc = prod_k G_k^nu_k (PA/(RU T))^dnu with net coefficients nu_k; the
    products' factors are inverted once per reaction (one reciprocal) unless
    per-species inverses are stored.
    """
    num, den = [], []
    dnu = 0
    for k, nu in x.spec:
        n = int(abs(nu))
        dnu += int(nu)
        g = gname(k) if gname else 'G{}'.format(k)
        if nu > 0:
            den += [g] * n
        else:
            num += [g] * n
    if store_inverse:
        num += ['Gi' + d[1:] for d in den]
        den = []
    if dnu:
        num.append('cpow{}{}'.format('m' if dnu < 0 else 'p', abs(dnu)))
    expr = ' * '.join(num) if num else '1.0'
    if den:
        expr = '{} * crs::rcp64({})'.format(expr, ' * '.join(den))
    return expr


def _gibbs_species(ir):
    """Species whose G_k is needed by some species-level Kc, in order."""
    need = set()
    for x in ir.rxns:
        if _species_kc_ok(x):
            need.update(k for k, _nu in x.spec)
    return sorted(need)


def _nasa_table_values(ir, scale=1.0):
    """Per species, per range (lower, upper): the 7 coefficients of
    s/R - h/RT on [1, ln T, T, T^2, T^3, T^4, 1/T], padded to 8, each
    multiplied by ``scale`` (log2 e for a base-2 exponential)."""
    vals = []
    for _tm, lo, hi in ir.nasa:
        for a in (lo, hi):
            vals += [v * scale for v in
                     (a[6] - a[0], a[0], a[1] / 2.0, a[2] / 6.0, a[3] / 12.0,
                      a[4] / 20.0, -a[5])] + [0.0]
    return vals


def _species_gibbs_lines(ir, K, tmids, store_inverse=False,
                         every=False, expfn='exp', nasa_lds=False):
    """Per-state G_k = exp(S_k/R - H_k/RT) and its inverse, once per species.

    With them Kc = prod_k G_k^nu_k (PA / (RU T))^dnu for every reversible
    reaction, so a reaction needs a few products instead of its own
    exponential of a lumped NASA polynomial.  With the branch-free
    exponential (expfn crs::exp64) the NASA coefficients are pre-scaled by
    log2 e and G_k = crs::exp2_64(smh_k): no conversion multiply per species
    (the LDS table, nasa_lds, holds the scaled values likewise).
    """
    base2 = expfn == 'crs::exp64'
    sc = _LOG2E if base2 else 1.0
    if base2:
        expfn = 'crs::exp2_64'
    need_g, need_gi, dnus = set(), set(), set()
    for x in ir.rxns:
        if not _species_kc_ok(x):
            continue
        d = 0
        for k, nu in x.spec:
            need_g.add(k)
            if nu > 0 and store_inverse:
                need_gi.add(k)
            d += int(nu)
        if d:
            dnus.add(d)
    if every:
        need_g = set(range(ir.nsp))
    out = []
    if not (need_g or need_gi):
        return out
    out.append('const double ctr = {} * T / {};'.format(
        _lit(chem.RU), _lit(chem.PA)))
    out.append('const double ictr = 1.0 / ctr;')
    for d in sorted(dnus):
        base = 'ctr' if d > 0 else 'ictr'
        expr = ' * '.join([base] * abs(d))
        out.append('const double cpow{}{} = {};'.format(
            'p' if d > 0 else 'm', abs(d), expr))
    order = sorted(need_g | need_gi)

    def nasa_reads(k):
        # this lane's range of species k's row of the coefficient table
        g = tmids.index(ir.nasa[k][0]) if ir.nasa[k][0] in tmids else None
        return (['CRS_LDS double *const np{k} = crs::opaque_lds((CRS_LDS '
                 'double *)(crs_nasa_lds + {o} + nsel{g}));'.format(
                     k=k, o=16 * k, g=g)] +
                ['const double nc{k}_{i} = np{k}[{i}];'.format(k=k, i=i)
                 for i in range(7)])
    for pos, k in enumerate(order):
        sp = ir.nasa[k]
        gi = tmids.index(sp[0]) if sp[0] in tmids else None
        if nasa_lds:
            # rows are read nasa_lds blocks ahead of their use (an int >= 1)
            # so that the LDS latency overlaps the previous exponentials
            ahead = 0 if nasa_lds is True else int(nasa_lds)
            if pos == 0:
                for kk in order[:ahead]:
                    out += nasa_reads(kk)
            if pos + ahead < len(order):
                out += nasa_reads(order[pos + ahead])
            out.append('const double smh{k} = nc{k}_0 + nc{k}_1 * logT + '
                       'nc{k}_2 * T + nc{k}_3 * T2 + nc{k}_4 * T3 + '
                       'nc{k}_5 * T4 + nc{k}_6 * iT;'.format(k=k))
        else:
            terms = []
            for half, a in (('l', sp[1]), ('h', sp[2])):
                coef = [v * sc for v in
                        (a[6] - a[0], a[0], a[1] / 2.0, a[2] / 6.0,
                         a[3] / 12.0, a[4] / 20.0, -a[5])]
                for bi, cv in enumerate(coef):
                    if cv != 0.0:
                        terms.append('{} * p{}{}_{}'.format(K(cv), half, gi,
                                                             bi))
            out.append('const double smh{} = {};'.format(k, ' + '.join(terms)))
        out.append('const double G{k} = {e}(smh{k});'.format(k=k, e=expfn))
        if k in need_gi:
            out.append('const double Gi{k} = 1.0 / G{k};'.format(k=k))
    return out


def _hg_ok(x):
    """Reaction whose rates take the H/G form: species-level Kc and integral
    stoichiometric coefficients on both sides."""
    return _species_kc_ok(x) and all(is_integer(nu)
                                     for _k, nu in x.reac + x.prod)


def _hg_net_rate(x, gref, pm):
    """Net rate of a reversible reaction from G_k and H_k = C_k / G_k.

    f = kf prod_r C_r and r = kf / Kc prod_p C_p with
    1/Kc = prod_r G_r / prod_p G_p (RU T / PA)^dnu, hence
    q = kf prod_r G_r (prod_r H_r - (RU T / PA)^dnu prod_p H_p):
    no reciprocal per reaction, and the forward-minus-reverse difference
    is one FMA.
    """
    def prod(terms, fmt):
        f = []
        for k, nu in terms:
            f += [fmt(k)] * int(nu)
        return ' * '.join(f) if f else '1.0'
    dnu = sum(int(nu) for _k, nu in x.spec)
    gr = prod(x.reac, gref)
    hr = prod(x.reac, lambda k: 'H{}'.format(k))
    hp = prod(x.prod, lambda k: 'H{}'.format(k))
    out = ['const double kg = kf * ({});'.format(gr)]
    if dnu:
        cp = 'cpow{}{}'.format('m' if dnu < 0 else 'p', abs(dnu))
        out.append('const double q = kg * __builtin_fma(-{}, {}, {}){};'.format(
            cp, hp, hr, pm))
    else:
        out.append('const double q = kg * (({}) - ({})){};'.format(hr, hp, pm))
    return out


def _litf(x):
    """C++ float literal nearest to ``x``."""
    import numpy as np
    v = float(np.float32(x))
    r = repr(v)
    if 'e' not in r and '.' not in r and 'n' not in r:
        r += '.0'
    return r + 'f'


def _eff_expr_f(x, C='C'):
    s = 'mf'
    for k, e in x.eff:
        if e >= 0:
            s += ' + {} * (float){}[{}]'.format(_litf(e), C, k)
        else:
            s += ' - {} * (float){}[{}]'.format(_litf(-e), C, k)
    return s


def _mixed_pm_lines(x, K, kf_x):
    """Pressure modification in mixed precision: third-body sums, falloff
    blending and the exponential cores in FP32; the Pr/(1+Pr) ratio, which
    spans the whole double range, in FP64."""
    out = []
    if x.thd:
        out.append('const double pm = (double)({});'.format(_eff_expr_f(x)))
        return out
    # a single species as collider: its concentration can be far below the
    # FP32 range (or zero), so Pr and log10 Pr stay FP64 there, clamped as
    # the reference clamps (fmax(Pr, 1e-300))
    named = x.pdep_sp is not None
    if named:
        out.append('const double Md = C[{}];'.format(x.pdep_sp))
    else:
        out.append('const float Mf = {};'.format(_eff_expr_f(x)))
    coll = 'Md' if named else '(double)Mf'
    # every exponent here is base 2 (the mixed kernel's Arrhenius form)
    if x.low is not None:
        l0, x0 = _fast_arr(x.low, K, 'k0', 'crs::exp2_mixed', log2=True)
        out += l0
        num, xnum = 'k0', x0
    else:
        num, xnum = 'kf', kf_x
    if x.high is not None:
        l1, x1 = _fast_arr(x.high, K, 'kinf', 'crs::exp2_mixed', log2=True)
        out += l1
        den, xden = 'kinf', x1
    else:
        den, xden = 'kf', kf_x
    if xnum is not None and xden is not None:
        out.append('const double xd = {} - {};'.format(xnum, xden))
        out.append('const double Pr = {} * crs::exp2_mixed(xd);'.format(coll))
        lpr = ('crs::log10f_hw(Mf) + (float)(xd * 0.30102999566398120)')
    else:
        out.append('const double Pr = {} * {} / {};'.format(num, coll, den))
        lpr = '(float)log10(fmax(Pr, 1.0e-300))'
    if named:
        lpr = ('(float)(crs::log64_finite(fmax(Pr, 1.0e-300)) * '
               '0.4342944819032518)')
    ratio = ''
    if x.falloff_low:
        ratio = ' * (Pr * crs::rcp64(1.0 + Pr))'
    elif x.falloff_high:
        ratio = ' * crs::rcp64(1.0 + Pr)'
    if x.falloff in ('troe', 'sri'):
        out.append('const float lPr = {};'.format(lpr))
    if x.falloff == 'troe':
        t = x.troe
        e1 = 'crs::expf_mixed({}T * {})'.format('-' if t['t3_pos'] else '',
                                                K(1.0 / t['t3']))
        e2 = 'crs::expf_mixed({}T * {})'.format('-' if t['t1_pos'] else '',
                                                K(1.0 / t['t1']))
        fc = '{} * {} + {} * {}'.format(_litf(t['one_minus_a']), e1,
                                        _litf(t['a']), e2)
        if t['t2'] is not None:
            fc += ' + crs::expf_mixed({}{} * iT)'.format(
                '-' if t['t2_pos'] else '', K(t['t2']))
        out.append('const float logFcent = crs::log10f_hw(fmaxf({}, '
                   '1.0e-30f));'.format(fc))
        out.append('const float A = lPr - 0.67f * logFcent - 0.4f;')
        out.append('const float B = 0.806f - 1.1762f * logFcent - 0.14f * lPr;')
        out.append('const float F = crs::exp10f_hw(logFcent * crs::rcpf_hw('
                   '1.0f + A * A * crs::rcpf_hw(B * B)));')
        out.append('const double pm = (double)F{};'.format(ratio))
    elif x.falloff == 'sri':
        sr = x.sri
        out.append('const float X = crs::rcpf_hw(1.0f + lPr * lPr);')
        e1 = 'crs::expf_mixed({}{} * iT)'.format('-' if sr['b_pos'] else '',
                                                 K(sr['b']))
        e2 = 'crs::expf_mixed({}T * {})'.format('-' if sr['c_pos'] else '',
                                                K(1.0 / sr['c']))
        base = '{} * {} + {}'.format(_litf(sr['a']), e1, e2)
        f = '__builtin_amdgcn_exp2f(X * __builtin_amdgcn_logf({}))'.format(base)
        if sr['d'] is not None:
            f += (' * {} * __builtin_amdgcn_exp2f({} * __builtin_amdgcn_logf('
                  'Tf))'.format(_litf(sr['d']), _litf(sr['e'])))
        out.append('const float F = {};'.format(f))
        out.append('const double pm = (double)F{};'.format(ratio))
    else:
        if x.falloff_low:
            out.append('const double pm = Pr * crs::rcp64(1.0 + Pr);')
        elif x.falloff_high:
            out.append('const double pm = crs::rcp64(1.0 + Pr);')
        else:
            out.append('const double pm = 1.0;')
    return out


_C_REF = re.compile(r'\bC\[(\d+)\]')


def _concentrations_to_lane_lds(lines, nsp, nl, block='crs::kBlock'):
    """Park the concentrations of species 0..nl-1 in LDS, one slot per lane.

    Each wave owns an [nl][64] FP64 array (a species row is 512 bytes, so
    every access is a ds_read_b64 with an immediate offset and no bank
    conflict).  Every reaction block re-reads the concentrations it needs
    through an opaque copy of the lane's base pointer, taken at the top of
    the block; the values are therefore live only inside the block instead
    of across the whole reaction list.  No barrier is needed: a lane reads
    only what it wrote itself.
    """
    # the nl species whose concentration is needed longest go to LDS
    last_use = {}
    for pos, l in enumerate(lines):
        for mo in _C_REF.finditer(l):
            last_use[int(mo.group(1))] = pos
    parked = sorted(last_use, key=lambda k: (-last_use[k], k))[:nl]
    slot = {k: s for s, k in enumerate(sorted(parked))}
    out = []
    published = False
    for l in lines:
        if not published and l == _CONC_DONE:
            out.append('__shared__ double sCl[{} / 64][{}][64];'
                       .format(block, max(len(slot), 1)))
            out.append('CRS_LDS double *const cb = (CRS_LDS double *)'
                       '&sCl[threadIdx.x >> 6][0][threadIdx.x & 63];')
            out += ['cb[{s} * 64] = C[{k}];'.format(s=s, k=k)
                    for k, s in sorted(slot.items())]
            published = True
            out.append(l)
            continue
        if not published:
            out.append(l)
            continue
        l = _C_REF.sub(lambda mo: ('cl[{} * 64]'.format(slot[int(mo.group(1))])
                                   if int(mo.group(1)) in slot
                                   else mo.group(0)), l)
        out.append(l)
        if l.startswith('{  // rxn'):
            out.append('    CRS_LDS double *const cl = crs::opaque_lds(cb);')
    return out


def _prefetch_lane_lds(lines, sources, dist):
    """Issue each reaction block's LDS reads ``dist`` blocks early.

    Input: blocks ``{  // rxn ..`` .. ``}`` whose first lines take
    ``ref = crs::opaque_lds(base)`` for every (base, ref, per_lane) of
    ``sources`` and which read ``ref[k * 64]`` (per-lane rows) or
    ``ref[k]`` (the block-wide constant table).  The reads of block j move
    to the outer scope in front of block j - dist, each group behind its
    own opaque copy of the base pointer, and the block uses the loaded
    values.  With one wave per SIMD nothing else covers an LDS read's
    latency; issued a block ahead, the reads complete while the previous
    reaction's exponential runs.  Values stay live for ``dist`` blocks (a
    few registers).
    """
    srcs = []
    for src in sources:
        base, ref, per_lane = src[:3]
        d = src[3] if len(src) > 3 else dist
        take = '    CRS_LDS double *const {} = crs::opaque_lds({});'.format(
            ref, base)
        pat = re.compile(r'\b{}\[(\d+){}\]'.format(
            re.escape(ref), r' \* 64' if per_lane else ''))
        srcs.append((base, ref, per_lane, take, pat, d))
    takes = {t for _b, _r, _l, t, _p, _d in srcs}
    starts = [p for p, l in enumerate(lines) if l.startswith('{  // rxn')]
    blocks = []   # (start, end, refs per source)
    for s in starts:
        e = s + 1
        while lines[e] != '}':
            e += 1
        refs = [sorted({int(mo.group(1)) for l in lines[s:e]
                        for mo in pat.finditer(l)})
                for _b, _r, _l, _t, pat, _d in srcs]
        blocks.append((s, e, refs))

    def group(j, d):
        """Reads of block j for the sources prefetched d blocks ahead."""
        out = []
        for (base, ref, per_lane, _t, _p, sd), refs in zip(srcs,
                                                           blocks[j][2]):
            if not refs or sd != d:
                continue
            p = '{}p{}'.format(ref, j)
            out.append('CRS_LDS double *const {} = crs::opaque_lds({});'
                       .format(p, base))
            out += ['const double {r}{j}_{k} = {p}[{k}{m}];'.format(
                r=ref, j=j, k=k, p=p, m=' * 64' if per_lane else '')
                    for k in refs]
        return out

    dists = sorted({sd for *_x, sd in srcs}, reverse=True)
    out = []
    prev = 0
    for j, (s, e, _refs) in enumerate(blocks):
        out += lines[prev:s]
        for d in dists:
            if j == 0:
                for jj in range(min(d, len(blocks))):
                    out += group(jj, d)
            if j + d < len(blocks):
                out += group(j + d, d)
        for l in lines[s:e + 1]:
            if l in takes:
                continue
            for _b, ref, _l, _t, pat, _d in srcs:
                l = pat.sub(lambda mo, r=ref: '{}{}_{}'.format(
                    r, j, mo.group(1)), l)
            out.append(l)
        prev = e + 1
    out += lines[prev:]
    return out


# FP64 operands that need no scalar moves: the hardware's inline
# constants, and values whose low 32 bits are zero (one 32-bit literal).
_INLINE_F64 = {0.0, 0.5, 1.0, 2.0, 4.0}
_FLT_LIT = re.compile(r'(?<![\w.])(\d+\.\d+(?:e[+-]?\d+)?|\d+e[+-]?\d+)'
                      r'(?![\w.])')


def _cheap_f64(v):
    if v in _INLINE_F64:
        return True
    return struct.unpack('<Q', struct.pack('<d', v))[0] & 0xffffffff == 0


def _block_consts_to_lds(lines, table, base='ktb', ref='kt'):
    """Move the FP64 literals of each reaction block into an LDS table.

    A literal costs two scalar moves; at one wave per SIMD those take issue
    slots from the wave's own vector work, and thousands of them also
    exhaust the scalar registers (spilled to vector lanes).  Every literal
    inside a block that is not an inline constant becomes a slot of
    ``table`` (appended to; slots of one block are contiguous, a value
    repeated within a block shares its slot), read as a wave-uniform LDS
    broadcast through ``ref = crs::opaque_lds(base)`` at the top of the
    block.  Values are unchanged (the table holds the same doubles).
    """
    out = []
    inside = False
    slots = {}
    for l in lines:
        if l.startswith('{  // rxn'):
            inside = True
            slots = {}
            out.append(l)
            out.append('    CRS_LDS double *const {} = crs::opaque_lds({});'
                       .format(ref, base))
            continue
        if inside and l == '}':
            inside = False
        if inside and not l.lstrip().startswith('CRS_LDS'):
            def sub(mo):
                v = float(mo.group(1))
                if _cheap_f64(v):
                    return mo.group(0)
                key = v.hex()
                if key not in slots:
                    slots[key] = len(table)
                    table.append(v)
                return '{}[{}]'.format(ref, slots[key])
            l = _FLT_LIT.sub(sub, l)
        out.append(l)
    return out


def _species_updates(x, seen):
    """w_k +/-= |nu| q for the species of reaction x (first touch assigns)."""
    out = []
    for k, nu in x.spec:
        anu = abs(nu)
        term = 'q' if anu == 1.0 else '{} * q'.format(
            _lit(emitted(py2_str(float(anu)))))
        if seen[k]:
            out.append('w[{k}] = w[{k}] {op} {t};'.format(
                k=k, op='-' if nu < 0 else '+', t=term))
        else:
            out.append('w[{k}] = {n}{t};'.format(
                k=k, n='-' if nu < 0 else '', t=term))
            seen[k] = True
    return out


def _nasa_masked(ir, K, k, tmids, coef_of, powers):
    """sum_b c_b x_b over the masked basis of species k's T_mid group;
    coef_of(a) gives the coefficients for NASA set a, powers the basis
    indices they multiply."""
    sp = ir.nasa[k]
    gi = tmids.index(sp[0])
    terms = []
    for half, a in (('l', sp[1]), ('h', sp[2])):
        for bi, cv in zip(powers, coef_of(a)):
            if cv != 0.0:
                terms.append('{} * p{}{}_{}'.format(K(cv), half, gi, bi))
    return ' + '.join(terms) if terms else '0.0'


# basis indices of 1, T, T^2, T^3, T^4, T^5 in the masked NASA basis
_B1, _BT, _BT2, _BT3, _BT4, _BT5 = 0, 2, 3, 4, 5, 7


def _dydt_cp_lines(ir, K, tmids):
    """cp_avg = sum_k cp_k Y_k with cp_k = RU/W_k (a0 + a1 T + ... + a4 T^4)
    (rate_subs.py:1707-1776 and the cp_avg sum of dydt, :1890-1905)."""
    out = ['double cp_avg = 0.0;']
    for k in range(ir.nsp):
        poly = _nasa_masked(ir, K, k, tmids, lambda a: a[0:5],
                            (_B1, _BT, _BT2, _BT3, _BT4))
        yk = 'y[{}]'.format(k) if k < ir.nsp - 1 else 'y_N'
        out.append('cp_avg = cp_avg + {} * ({}) * {};'.format(
            K(ir.ru_mw[k]), poly, yk))
    return out


def _dydt_finish_lines(ir, K, tmids, k, wk, store='W[{r} * ld + i] = {v};'):
    """Species k's rate wk is final: its term of the energy equation
    (w_k h_k W_k, h_k = RU/W_k (a5 + a0 T + a1 T^2/2 + ... + a4 T^5/5),
    rate_subs.py:1495-1565) and, for k < NSP-1, dY_k/dt = w_k W_k / rho,
    written as ``store`` formats it (row r, value v)."""
    poly = _nasa_masked(
        ir, K, k, tmids,
        lambda a: [a[5], a[0], a[1] / 2.0, a[2] / 3.0, a[3] / 4.0, a[4] / 5.0],
        (_B1, _BT, _BT2, _BT3, _BT4, _BT5))
    out = ['esum = esum + ({w}) * ({ru} * ({p})) * {mw};'.format(
        w=wk, ru=K(ir.ru_mw[k]), p=poly, mw=K(ir.mw[k]))]
    if k < ir.nsp - 1:
        out.append(store.format(r=k + 1, v='({w}) * ({mw} * irho)'.format(
            w=wk, mw=K(ir.mw[k]))))
    return out


def _dydt_energy_lines(ir, K, chunk=8):
    """Body of the energy stage of the split CONP dydt (rate_subs.py:
    1829-1946 after eval_spec_rates).  In: the state Y (row 0 = T, rows
    1..NSP-1 the mass fractions) and S, the species rates as the wdot
    kernel stores them for the dydt (option dydt_rates): row k + 1 holds
    w_k W_k / rho = dY_k/dt for k < NSP-1, row 0 the bath gas's
    w_N W_N / rho.  Out, through the output object (``out.fetch(r)`` before
    ``out.put(r, value)``): rows 1..NSP-1 get dY_k/dt, row 0 last gets
    dT/dt = -sum_k w_k h_k W_k / (rho cp_avg) = -sum_k h_k S_k / cp_avg
    (h_k per unit mass; rho cancels, so neither rho nor P is needed).

    Two passes keep the live set small (the kernel is memory bound and wants
    occupancy): cp_avg over the mass fractions as they load (the
    reference's order, the bath gas's share from y_N last), then the S rows
    in chunks of ``chunk`` species, every row of a chunk loaded before its
    outputs, the energy sum in species order."""
    nsp = ir.nsp
    tmids = []
    for tm, _lo, _hi in ir.nasa:
        if tm not in tmids:
            tmids.append(tm)
    # rows are reached through a uniform row base and the lane's 32-bit
    # byte offset io (crs::ld_row): no 64-bit address per row to keep live
    lines = ['const uint32_t io = (uint32_t)i * 8u;',
             'const double T = crs::ld_row(Y, io);',
             'const double T2 = T * T;', 'const double T3 = T2 * T;',
             'const double T4 = T3 * T;', 'const double T5 = T4 * T;']
    vals = {_B1: '1.0', _BT: 'T', _BT2: 'T2', _BT3: 'T3', _BT4: 'T4',
            _BT5: 'T5'}
    for gi, tm in enumerate(tmids):
        lines.append('const bool lo{g} = T <= {tm};'.format(g=gi,
                                                           tm=_lit(tm)))
        for bi, v in sorted(vals.items()):
            lines.append('const double pl{g}_{b} = lo{g} ? {v} : 0.0;'.format(
                g=gi, b=bi, v=v))
            lines.append('const double ph{g}_{b} = lo{g} ? 0.0 : {v};'.format(
                g=gi, b=bi, v=v))
    # pass 1: cp_avg over the mass fractions (y_N = 1 - their sum)
    lines += ['double ysum = 0.0;', 'double cp_avg = 0.0;']
    cps = [_nasa_masked(ir, K, k, tmids, lambda a: a[0:5],
                        (_B1, _BT, _BT2, _BT3, _BT4)) for k in range(nsp)]
    for k in range(nsp - 1):
        lines.append('{{ const double yk = crs::ld_row(Y + {r} * ld, io);'
                     .format(r=k + 1))
        lines.append('  ysum = {}yk;'.format('ysum + ' if k else ''))
        lines.append('  cp_avg = cp_avg + {} * ({}) * yk; }}'.format(
            K(ir.ru_mw[k]), cps[k]))
    lines += ['const double y_N = 1.0 - ysum;',
              'cp_avg = cp_avg + {} * ({}) * y_N;'.format(
                  K(ir.ru_mw[nsp - 1]), cps[nsp - 1]),
              'double esum = 0.0;']
    # pass 2: the scaled rates in species order, h_k S_k summed
    # (h_k = RU/W_k (a5 + a0 T + a1 T^2/2 + ... + a4 T^5/5),
    # rate_subs.py:1495-1565)
    row = [k + 1 if k < nsp - 1 else 0 for k in range(nsp)]
    for k0 in range(0, nsp, chunk):
        ks = list(range(k0, min(k0 + chunk, nsp)))
        lines.append('{')
        lines += ['  const double s{k} = crs::ld_row(S + {r} * ld, io);'
                  .format(k=k, r=row[k]) for k in ks]
        lines += ['  out.fetch({});'.format(row[k]) for k in ks
                  if row[k] > 0]
        for k in ks:
            poly = _nasa_masked(
                ir, K, k, tmids,
                lambda a: [a[5], a[0], a[1] / 2.0, a[2] / 3.0, a[3] / 4.0,
                           a[4] / 5.0],
                (_B1, _BT, _BT2, _BT3, _BT4, _BT5))
            lines.append('  esum = esum + s{k} * ({ru} * ({p}));'.format(
                k=k, ru=K(ir.ru_mw[k]), p=poly))
            if row[k] > 0:
                lines.append('  out.put({r}, s{k});'.format(r=row[k], k=k))
        lines.append('}')
    # energy equation, rate_subs.py:1910-1925
    lines += ['out.fetch(0);', 'out.put(0, (-1.0 / cp_avg) * esum);']
    return lines


def _fused_fast_lines(ir, K, opts, mixed=False, dydt=False, ktab=None,
                      ktab_name=None):
    """Per-state body of the fused FP64 kernel, arranged for the GPU.

    Same physics and reaction order as the reference, with the arithmetic
    rearranged so that a state needs fewer and cheaper operations and no
    per-reaction data-dependent selects:

    * 1/T, T^2..T^4 once per state; E/T becomes E * (1/T);
    * Kc from the NASA basis [1, lnT, T, T^2, T^3, T^4, 1/T] pre-masked
      once per state for each T_mid group (lower/upper range), so ln Kc is
      a fixed sum of products with constant coefficients;
    * r = prod C'' * exp(ln kf - ln Kc): one exponential instead of
      exp(Kc) and a division;
    * Pr = [M] exp(ln k0 - ln kinf) for falloff (k_inf reuses kf);
    * each species rate is stored as soon as its last reaction is done.
    """
    nsp = ir.nsp
    ru = _lit(emitted(fmt_e8(chem.RU)))
    # preload: the kernel loaded T and P before its LDS table fill
    tp = (['const double T = T_pre;', 'const double P = P_pre;']
          if opts.get('preload') else None)
    if dydt:
        # the reference's state vector: y[0] = T, y[1 + k] = Y_k
        lines = (tp or ['const double T = Y[i];', 'const double P = P_[i];']) + [
            'double y[{}];'.format(max(nsp - 1, 1))]
        lines += ['y[{k}] = Y[{r} * ld + i];'.format(k=k, r=k + 1)
                  for k in range(nsp - 1)]
    else:
        lines = (tp or ['const double T = T_[i];', 'const double P = P_[i];']) + [
            'double y[{}];'.format(max(nsp - 1, 1))]
        lines += ['y[{k}] = Y[{k} * ld + i];'.format(k=k)
                  for k in range(nsp - 1)]
    # conc_late: the T-only work (basis, G_k) runs while the mass fractions
    # are still in flight; eval_conc's sums wait for them afterwards
    # dydt_rates: each species' rate is stored as the CONP dydt's
    # w_k W_k / rho, species k in row k + 1 and the bath gas in row 0
    rates = bool(opts.get('dydt_rates'))
    conc = (_conc_lines(ir, K) +
            (['const double irho_r = 1.0 / rho;'] if rates else []) +
            [_CONC_DONE])
    late = opts.get('conc_late', False)
    if not late:
        lines += conc
    lines += ['const double logT = crs::log64_finite(T);',
              'const double iT = 1.0 / T;',
              'const double T2 = T * T;', 'const double T3 = T2 * T;',
              'const double T4 = T3 * T;', 'const double T5 = T4 * T;',
              'const double m = P / ({} * T);'.format(ru),
              '(void)m; (void)T2; (void)T3; (void)T4; (void)T5;']
    if any(x.plog or x.cheb for x in ir.rxns):
        lines += ['const double logP = crs::log64_finite(P);',
                  'const double log10P = logP * 0.43429448190325182;',
                  '(void)log10P;']
    lines += ['const double log2T = logT * 1.4426950408889634;',
              '(void)log2T;']
    if mixed:
        lines += ['const float mf = (float)m;', 'const float Tf = (float)T;',
                  '(void)mf; (void)Tf;']
    fexp = 'crs::exp64' if opts.get('exp') == 'crs' else 'exp'
    # the mixed kernel's arithmetic (mixed_arith): 'accurate' (default)
    # keeps Kc (the G_k), the net rate and the sums in full FP64 and takes
    # the Arrhenius and falloff exponentials at about 27 bits (degree-4
    # exponential, crs::exp2_m64, 1.0e-8 relative; falloff and third-body
    # blocks in FP64 arithmetic);
    # 'fp32' is the FP32 variant (hardware v_exp_f32 cores, FP32 falloff
    # and third-body blocks), faster and about 20x less accurate
    m32 = mixed and opts.get('mixed_arith', 'accurate') == 'fp32'
    expfn = 'crs::exp_mixed' if m32 else fexp
    # Arrhenius exponentials in base 2 (constants pre-scaled by log2 e):
    # exp2_64 in the FP64 kernel, exp2_m64 or exp2_mixed in the mixed one;
    # mixed_exp='fp64' keeps the mixed kernel's Arrhenius exponentials in
    # full FP64 (to measure how much of its error they carry)
    if mixed and opts.get('mixed_exp') == 'fp64':
        arr2 = 'crs::exp2_64'
    elif mixed:
        arr2 = 'crs::exp2_mixed' if m32 else 'crs::exp2_m64'
    else:
        arr2 = 'crs::exp2_64' if opts.get('exp') == 'crs' else None
    tmids = []
    for x in ir.rxns:
        for tm, _lo, _hi in x.kc_groups:
            if tm not in tmids:
                tmids.append(tm)
    if opts.get('kc') in ('species', 'species_hg') or dydt:
        for tm, _lo, _hi in ir.nasa:
            if tm not in tmids:
                tmids.append(tm)
    # group index of a reaction's g-th Tmid group -> global mask index
    basis = _BASIS + (('T5',) if dydt else ())
    for gi, tm in enumerate(tmids):
        lines.append('const bool lo{g} = T <= {tm};'.format(g=gi, tm=_lit(tm)))
        for bi, bname in enumerate(basis):
            lines.append('const double pl{g}_{b} = lo{g} ? {v} : 0.0;'.format(
                g=gi, b=bi, v=bname))
            lines.append('const double ph{g}_{b} = lo{g} ? 0.0 : {v};'.format(
                g=gi, b=bi, v=bname))
    dydt_lines = []
    if dydt:
        dydt_lines = _dydt_cp_lines(ir, K, tmids) + [
            'const double irho = 1.0 / rho;', 'double esum = 0.0;']
        if not late:
            lines += dydt_lines
    hg = opts.get('kc') == 'species_hg'
    species_kc = opts.get('kc') in ('species', 'species_hg')
    # mixed kernel: FP32 falloff / third-body blocks in the 'fp32' variant
    # (or with mixed_pm='fp32'), FP64 ones otherwise
    mixed_pm32 = mixed and opts.get('mixed_pm',
                                    'fp32' if m32 else 'fp64') == 'fp32'
    nasa_lds = opts.get('nasa') == 'lds'
    if nasa_lds:
        nasa_lds = max(1, int(opts.get('nasa_prefetch', 2))) if opts.get(
            'nasa_prefetch', 2) else True
        lines += ['const int nsel{g} = lo{g} ? 0 : 8;'.format(g=gi)
                  for gi in range(len(tmids))]
    if species_kc:
        lines += _species_gibbs_lines(ir, K, tmids,
                                      opts.get('kc_inverse') == 'store',
                                      every=hg, expfn=fexp, nasa_lds=nasa_lds)
    g_lane = hg and opts.get('g_store') == 'lane_lds'
    # LDS budget: one wave per SIMD leaves 40 KB per wave, 80 FP64 rows
    n_gl = min(nsp, opts.get('lds_species', 78)) if g_lane else 0
    rxns = _retiring_order(ir) if opts.get('order') == 'retire' else ir.rxns
    # which species get an LDS row when not all fit: the first n_gl by
    # index, or (g_pick='retire') the n_gl whose last reaction comes latest,
    # so the G_k left in registers are the ones that die first
    if opts.get('g_pick') == 'retire' and n_gl < nsp:
        end = {k: p for p, x in enumerate(rxns) for k, _nu in x.spec}
        by_end = sorted(range(nsp), key=lambda k: (-end.get(k, -1), k))
        gl_slot = {k: s for s, k in enumerate(sorted(by_end[:n_gl]))}
    else:
        gl_slot = {k: k for k in range(n_gl)}
    gref = (lambda k: 'gl[{} * 64]'.format(gl_slot[k]) if k in gl_slot else
            'G{}'.format(k))
    if hg and g_lane:
        lines.append('__shared__ double sGl[{} / 64][{}][64];'
                     .format(opts.get('block_expr', 'crs::kBlock'), n_gl))
        lines.append('CRS_LDS double *const gb = (CRS_LDS double *)'
                     '&sGl[threadIdx.x >> 6][0][threadIdx.x & 63];')
        lines += ['gb[{s} * 64] = G{k};'.format(s=s, k=k)
                  for k, s in sorted(gl_slot.items())]
    if late:
        lines += conc + dydt_lines
    if hg:
        # H_k = C_k / G_k; from here on C_k is G_k H_k (no third array).
        # g_reload: G_k of the LDS species read back through an opaque
        # pointer, so its register dies at the store instead of living
        # across eval_conc
        reload = g_lane and opts.get('g_reload')
        if reload:
            lines.append('CRS_LDS double *const gh = crs::opaque_lds(gb);')
        lines += ['const double H{k} = C[{k}] * crs::rcp64({g});'.format(
                      k=k, g='gh[{} * 64]'.format(gl_slot[k])
                      if reload and k in gl_slot else 'G{}'.format(k))
                  for k in range(nsp)]
    rx_start = len(lines)
    last = _last_touch(ir, rxns)
    seen = [False] * nsp
    lines.append('double w[{}];'.format(nsp))
    for k in range(nsp):
        if k not in last:
            if dydt:
                if k < nsp - 1:
                    lines.append('W[{r} * ld + i] = 0.0;'.format(r=k + 1))
            elif rates:
                lines.append('W[{r} * ld + i] = 0.0;'.format(
                    r=k + 1 if k < nsp - 1 else 0))
            else:
                lines.append('W[{k} * ld + i] = 0.0;'.format(k=k))

    def w_store(k):
        if rates:
            return 'W[{r} * ld + i] = (w[{k}]) * ({mw} * irho_r);'.format(
                r=k + 1 if k < nsp - 1 else 0, k=k, mw=K(ir.mw[k]))
        return 'W[{k} * ld + i] = w[{k}];'.format(k=k)
    for pos, x in enumerate(rxns):
        if x.plog:
            inner, xf = _plog_fast_lines(x, K, expfn), None
        elif x.cheb:
            inner, xf = _cheb_fast_lines(x, K, expfn), None
        elif arr2:
            inner, xf = _fast_arr(x.arr, K, 'kf', arr2, log2=True)
        else:
            inner, xf = _fast_arr(x.arr, K, 'kf', expfn)
        if hg and _hg_ok(x):
            pm = ''
            if x.pd_index is not None:
                inner += (_mixed_pm_lines(x, K, xf) if mixed_pm32 else
                          _fast_pm_lines(x, K, xf, arr2))
                pm = ' * pm'
            inner += _hg_net_rate(x, gref, pm)
            inner += _species_updates(x, seen)
            lines.append('{  // rxn %d' % x.index)
            lines += ['    ' + l for l in inner]
            lines.append('}')
            for k, _nu in x.spec:
                if last.get(k) == x.index:
                    lines += (_dydt_finish_lines(ir, K, tmids, k,
                                                 'w[{}]'.format(k))
                              if dydt else [w_store(k)])
            continue
        inner.append('const double f = {}kf;'.format(_conc_prod(x.reac)))
        if x.rev:
            if x.rev_arr is not None:
                la, _xr = _fast_arr(x.rev_arr, K, 'kr', expfn)
                inner += la
                inner.append('const double r = {}kr;'.format(
                    _conc_prod(x.prod)))
            elif species_kc and _species_kc_ok(x):
                # 1/Kc first: its factors are O(1) combined, so the
                # concentration product cannot underflow prematurely
                inner.append('const double r = {}(kf * {});'.format(
                    _conc_prod(x.prod), _species_kc_inverse(
                        x, opts.get('kc_inverse') == 'store')))
            else:
                # remap this reaction's Tmid groups onto the global masks
                groups = []
                for tm, lo, hi in x.kc_groups:
                    groups.append((tmids.index(tm), lo, hi))
                terms = []
                for n_g, (gi, lo, hi) in enumerate(groups):
                    for half, c in (('l', lo), ('h', hi)):
                        coef = list(c[0:6]) + [-c[6]]
                        if n_g == 0:
                            coef[0] = coef[0] + math.log(x.kc_pref)
                        for bi, cv in enumerate(coef):
                            if cv != 0.0:
                                terms.append('{} * p{}{}_{}'.format(
                                    K(cv), half, gi, bi))
                lnkc = ' + '.join(terms) if terms else '0.0'
                inner.append('const double lnKc = {};'.format(lnkc))
                if mixed:
                    # Kc stays FP64; only the Arrhenius factor is mixed
                    inner.append('const double r = {}(kf * exp(-lnKc));'.format(
                        _conc_prod(x.prod)))
                elif xf is not None and arr2:
                    inner.append('const double r = {}{}({} - lnKc * '
                                 '1.4426950408889634);'.format(
                                     _conc_prod(x.prod), arr2, xf))
                elif xf is not None:
                    inner.append('const double r = {}exp({} - lnKc);'.format(
                        _conc_prod(x.prod), xf))
                else:
                    inner.append('const double r = {}kf * exp(-lnKc);'.format(
                        _conc_prod(x.prod)))
        q = '(f - r)' if x.rev else 'f'
        if x.pd_index is not None:
            inner += (_mixed_pm_lines(x, K, xf) if mixed_pm32 else
                      _fast_pm_lines(x, K, xf, arr2))
            q = q + ' * pm'
        inner.append('const double q = {};'.format(q))
        inner += _species_updates(x, seen)
        lines.append('{  // rxn %d' % x.index)
        lines += ['    ' + l for l in inner]
        lines.append('}')
        for k, _nu in x.spec:
            if last.get(k) == x.index:
                lines += (_dydt_finish_lines(ir, K, tmids, k,
                                             'w[{}]'.format(k))
                          if dydt else [w_store(k)])
        if opts.get('sched_barrier'):
            lines.append('__builtin_amdgcn_sched_barrier(0);')
    if mixed and not mixed_pm32:
        # the falloff blocks' exponentials at the mixed kernel's accuracy
        # (the G_k, before rx_start, keep exp64: Kc stays full FP64)
        lines = lines[:rx_start] + [
            l.replace('crs::exp64(', 'crs::exp_m64(')
             .replace('crs::exp2_64(', 'crs::exp2_m64(')
             .replace('crs::log2_64(', 'crs::log2_m64(')
            for l in lines[rx_start:]]
    if dydt:
        # energy equation, rate_subs.py:1910-1925
        lines.append('W[i] = (-1.0 / (rho * cp_avg)) * esum;')
    if hg:
        # concentrations after the H_k lines: C_k = G_k * H_k
        def c_of(mo):
            k = int(mo.group(1))
            return '({} * H{})'.format(gref(k), k)
        rx = [_C_REF.sub(c_of, l) for l in lines[rx_start:]]
        if g_lane:
            out = []
            for l in rx:
                out.append(l)
                if l.startswith('{  // rxn'):
                    out.append('    CRS_LDS double *const gl = '
                               'crs::opaque_lds(gb);')
            rx = out
        lines = lines[:rx_start] + rx
    sources = [('gb', 'gl', True)] if hg and g_lane else []
    if opts.get('c_store') == 'lane_lds':
        lines = _concentrations_to_lane_lds(
            lines, ir.nsp, min(ir.nsp, opts.get('lds_species', 78)),
            opts.get('block_expr', 'crs::kBlock'))
        sources.append(('cb', 'cl', True))
    if ktab is not None:
        first = next(p for p, l in enumerate(lines)
                     if l.startswith('{  // rxn'))
        lines = (lines[:first] +
                 ['CRS_LDS double *const ktb = (CRS_LDS double *){};'.format(
                     ktab_name)] +
                 _block_consts_to_lds(lines[first:], ktab))
        sources.append(('ktb', 'kt', False,
                        opts.get('consts_prefetch', opts.get('lds_prefetch'))))
    if opts.get('lds_prefetch') and sources:
        lines = _prefetch_lane_lds(lines, sources, opts['lds_prefetch'])
    if opts.get('exp_poly') == 'estrin':
        lines = [l.replace('crs::exp2_64(', 'crs::exp2_64_estrin(')
                  .replace('crs::exp64(', 'crs::exp64_estrin(') for l in lines]
    if opts.get('launder', False):
        lines = _launder_blocks(lines)
    if opts.get('row_offsets', True):
        lines = _row_offsets(lines)
    split = opts.get('split_blocks', 32)
    if split:
        lines = _split_blocks(lines, int(split))
    return lines


_DECL = re.compile(r'^(const double|CRS_LDS double \*const) (\w+) = ')


def _split_blocks(lines, per):
    """Make every ``per`` consecutive reaction blocks one basic block.

    Straight-line code is one basic block, in which the code generator
    materialises each distinct 32-bit half of a literal once and keeps it
    in an SGPR until its last use anywhere in the kernel; at 53 species
    that spilled ~350 SGPRs into VGPR lanes.  Wrapping a group of blocks in
    `if (crs::go<N>()) { ... }` (a uniform true the compiler cannot see
    through: a scalar branch that is always taken) confines that sharing to
    the group.  Values a group declares for a later one (LDS reads issued
    ahead) are declared before the group and assigned inside it.  The
    exponential's constants are taken once, in registers (crs::ExpK),
    instead of per group.  In the concentration-row layout (c_store
    'lane_lds', 111 species) the rows read ahead between reactions end a
    group, so each reaction is its own basic block there: measured, that
    spills least (207 scratch instructions in k_wdot_f64 at 111 species
    against 394-549 with 32-64 reactions per group).
    """
    lines = ['const crs::ExpK ek = crs::exp_consts();'] + [
        l.replace('crs::exp2_64(', 'crs::exp2_64k(ek, ')
         .replace('crs::exp64(', 'crs::exp64k(ek, ')
         .replace('crs::log2_64(', 'crs::log2_64k(ek, ')
         .replace('crs::exp2_m64(', 'crs::exp2_m64k(ekm, ')
         .replace('crs::exp_m64(', 'crs::exp_m64k(ekm, ')
         .replace('crs::log2_m64(', 'crs::log2_m64k(ekm, ') for l in lines]
    if any('ekm, ' in l for l in lines):
        lines.insert(1, 'const crs::ExpMK ekm = crs::exp_m_consts();')
    # segments: ('plain', [lines]) outside groups, ('group', [lines])
    segs = []
    cur = None
    count = 0
    for l in lines:
        if l.startswith('{  // rxn'):
            if count % per == 0:
                cur = []
                segs.append(('group', cur))
            count += 1
            cur.append(l)
        elif cur is not None and (l.startswith((' ', '}', 'CRS_LDS',
                                                'const double gl',
                                                'crs::st_row')) or
                                  l.startswith('if (')):
            cur.append(l)
        else:
            if segs and segs[-1][0] == 'plain':
                segs[-1][1].append(l)
            else:
                segs.append(('plain', [l]))
            cur = None
            count = 0
    out = []
    groups = 0
    for n, (kind, body) in enumerate(segs):
        if kind == 'plain':
            out += body
            continue
        groups += 1
        later = '\n'.join(l for _k, b in segs[n + 1:] for l in b)
        hoist = []
        inner = []
        for l in body:
            mo = _DECL.match(l)
            if mo and re.search(r'\b{}\b'.format(mo.group(2)), later):
                hoist.append(('double' if mo.group(1) == 'const double'
                              else 'CRS_LDS double *', mo.group(2)))
                l = l.replace('{} {} ='.format(mo.group(1), mo.group(2)),
                              '{} ='.format(mo.group(2)), 1)
            inner.append(l)
        out += ['{} {};'.format(t, h) for t, h in hoist]
        out.append('if (crs::go<{}>()) {{'.format(groups))
        out += inner
        out.append('}')
    return out


_ROW_LOAD = re.compile(r'\b(Y|T_|P_)\[(?:(\d+) \* ld \+ )?i\]')
_ROW_STORE = re.compile(r'^(\s*)W\[(\d+) \* ld \+ i\] = (.*);$')


def _row_offsets(lines):
    """Rewrite the fused kernels' per-state global accesses as uniform row
    pointers plus one 32-bit lane offset (crs::ld_row / st_row): every row
    address is then an SGPR base and the shared offset io = 8 i, instead of
    a 64-bit VGPR pair per row.  Needs 8 i < 2^32 (ld below 2^29, checked
    by the entry points)."""
    def load(mo):
        row = mo.group(1) if mo.group(2) is None else '{} + {} * ld'.format(
            mo.group(1), mo.group(2))
        return 'crs::ld_row({}, io)'.format(row)
    out = ['const uint32_t io = (uint32_t)i * 8u;']
    tag = 0
    for l in lines:
        mo = _ROW_STORE.match(l)
        if mo:
            # each store forms its own row address from a fresh copy of
            # ld: k ld is not shared with the load of row k at the start
            tag += 1
            l = '{}crs::st_row(W + {} * crs::fresh_s<{}>(ld), io, {});'.format(
                mo.group(1), mo.group(2), tag, mo.group(3))
        out.append(_ROW_LOAD.sub(load, l))
    return out


def _launder_blocks(lines):
    """Give every reaction block its own copies of the register-held
    per-state values it reads (H_k, G_k, log2 T, 1/T) through crs::fresh.

    Without them the compiler merges equal subexpressions of different
    blocks (a squared H_k, a collider's G_k H_k, b log2 T for two reactions
    with the same b) into one value that stays live from the first block to
    the last: at 53 species that held ~50 extra FP64 values across the
    reaction list.  The copies emit no instruction.
    """
    out = []
    tag = 0
    i = 0
    while i < len(lines):
        l = lines[i]
        if not l.startswith('{  // rxn'):
            out.append(l)
            i += 1
            continue
        j = i + 1
        while lines[j] != '}':
            j += 1
        body = lines[i + 1:j]
        declared = set()
        for b in body:
            mo = re.match(r'\s*const (?:double|float) (\w+) =', b)
            if mo:
                declared.add(mo.group(1))
        used = []
        for b in body:
            for mo in _FRESH_NAMES.finditer(b):
                n = mo.group(1)
                if n not in declared and n not in used:
                    used.append(n)
        tag += 1
        out.append(l)
        out += ['    const double {n}_f = crs::fresh<{t}>({n});'.format(
            n=n, t=tag) for n in used]
        ren = set(used)
        out += [_FRESH_NAMES.sub(lambda mo: mo.group(1) + '_f'
                                 if mo.group(1) in ren else mo.group(1), b)
                for b in body]
        out.append('}')
        i = j + 1
    return out


def _idx_kernel_tail(fill, call, preload):
    """Body of a fused kernel over a batch (idx == nullptr) or a subset:
    the block's LDS table fill, then one state per thread.  The subset size
    is m, or min(m, *m_dev) when the count lives on the device (m_dev not
    null: a fallback list whose length no host code has read).  With
    preload the thread's T and P are loaded before the fill, so their
    latency overlaps the fill's and its barrier instead of following them.
    The block-wide fill and its barrier come before any thread leaves."""
    count = ('    const int64_t mm = m_dev ? (m < (int64_t)*m_dev ? m : '
             '(int64_t)*m_dev) : m;\n')
    if not (preload and fill):
        return (fill +
                '    const int64_t t = (int64_t)blockIdx.x * blockDim.x + '
                'threadIdx.x;\n' + count +
                '    if (t >= mm) return;\n'
                '    const int64_t i = idx ? (int64_t)idx[t] : t;\n'
                '    if (i < 0 || i >= ld) return;\n'
                '    ' + call.format(pre='') + '\n}\n\n')
    return ('    const int64_t t = (int64_t)blockIdx.x * blockDim.x + '
            'threadIdx.x;\n' + count +
            '    const int64_t i = t < mm ? (idx ? (int64_t)idx[t] : t) : -1;\n'
            '    const bool live = i >= 0 && i < ld;\n'
            '    const double T_pre = live ? T_[i] : 0.0;\n'
            '    const double P_pre = live ? P_[i] : 0.0;\n' + fill +
            '    if (!live) return;\n'
            '    ' + call.format(pre=', T_pre, P_pre') + '\n}\n\n')


def emit_hip_source(ir, opts=None):
    """Full HIP source text for mechanism ``ir``."""
    raw = dict(opts or {})
    opts = resolve_opts(ir.nsp, raw)
    # the CONP dydt kernel keeps one wave per SIMD unless asked otherwise:
    # its cp and h terms add live registers that two waves would spill
    dopts = (resolve_opts(ir.nsp, dict(raw, waves_per_simd=None))
             if raw.get('waves_per_simd', 'auto') == 'auto' else opts)
    # dydt='split' (default): crs_dydt runs the wdot kernel and a light
    # energy-equation kernel (k_dydt_energy), crs_rk4 the wdot kernel and
    # an energy + RK4-stage kernel; dydt='fused': one kernel, k_dydt_f64
    dydt_split = opts.get('dydt', 'split') == 'split'
    if opts.get('dydt', 'split') not in ('split', 'fused'):
        raise ValueError("dydt must be 'split' or 'fused'")
    if dydt_split and opts['fused'] != 'fast':
        raise ValueError("dydt='split' needs fused='fast' (its rate kernel "
                         "is that body storing dydt-scaled rates)")
    if dydt_split:
        dopts = opts
    K = _Consts(opts['consts'])
    nsp, nr, nrev, npd = ir.nsp, ir.nr, ir.nrev, ir.npd
    ru = _lit(emitted(fmt_e8(chem.RU)))
    body = []
    # fused kernels: waves per SIMD the register allocation must leave room
    # for (HIP's second launch bound)
    lb = (', {}'.format(int(opts['waves_per_simd']))
          if opts.get('waves_per_simd') else '')
    dlb = (', {}'.format(int(dopts['waves_per_simd']))
           if dopts.get('waves_per_simd') else '')
    # threads per block of k_wdot_f64 / k_wdot_mixed (a multiple of 64)
    wblock = int(opts.get('block') or 256)
    if wblock % 64 or not 64 <= wblock <= 1024:
        raise ValueError('block must be a multiple of 64 in [64, 1024]')
    body.append('constexpr int kWdotBlock = {};\n\n'.format(wblock))

    # ---- stage kernel: eval_conc -------------------------------------
    lines = ['const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;',
             'if (i >= n) return;',
             'const double T = T_[i];', 'const double P = P_[i];',
             'double y[{}];'.format(max(nsp - 1, 1))]
    lines += ['y[{k}] = Y[{k} * ld + i];'.format(k=k) for k in range(nsp - 1)]
    lines += _conc_lines(ir, K)
    lines += ['y_N_out[i] = y_N;', 'mw_out[i] = mw_avg;', 'rho_out[i] = rho;']
    lines += ['Cout[{k} * ld + i] = C[{k}];'.format(k=k) for k in range(nsp)]
    body.append(
        '__global__ __launch_bounds__(crs::kBlock) void k_eval_conc(\n'
        '    const double *__restrict__ K,\n'
        '    const double *__restrict__ T_, const double *__restrict__ P_,\n'
        '    const double *__restrict__ Y, double *__restrict__ y_N_out,\n'
        '    double *__restrict__ mw_out, double *__restrict__ rho_out,\n'
        '    double *__restrict__ Cout, int64_t n, int64_t ld)\n{\n'
        '#pragma clang fp contract(off)\n' + _indent(lines, 4) + '}\n\n')

    # ---- stage kernel: eval_rxn_rates ---------------------------------
    lines = ['const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;',
             'if (i >= n) return;',
             'const double T = T_[i];',
             'const double P = P_[i];',
             '(void)P;',
             'const double logT = log(T);',
             'double C[{}];'.format(nsp)]
    lines += ['C[{k}] = Cin[{k} * ld + i];'.format(k=k) for k in range(nsp)]
    for x in ir.rxns:
        lines.append('{  // rxn %d' % x.index)
        inner = _rate_lines(x, K)
        inner.append('fwd[{} * ld + i] = f;'.format(x.index))
        if x.rev:
            inner.append('rev[{} * ld + i] = r;'.format(x.rev_index))
        lines += ['    ' + l for l in inner]
        lines.append('}')
    body.append(
        '__global__ __launch_bounds__(crs::kBlock) void k_eval_rxn_rates(\n'
        '    const double *__restrict__ K,\n'
        '    const double *__restrict__ T_, const double *__restrict__ P_,\n'
        '    const double *__restrict__ Cin,\n'
        '    double *__restrict__ fwd, double *__restrict__ rev, int64_t n,\n'
        '    int64_t ld)\n{\n' + _indent(lines, 4) + '}\n\n')

    # ---- stage kernel: get_rxn_pres_mod -------------------------------
    if npd:
        lines = ['const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;',
                 'if (i >= n) return;',
                 'const double T = T_[i];', 'const double P = P_[i];',
                 'const double logT = log(T);',
                 'const double m = P / ({} * T);'.format(ru),
                 'double C[{}];'.format(nsp)]
        lines += ['C[{k}] = Cin[{k} * ld + i];'.format(k=k) for k in range(nsp)]
        for x in ir.rxns:
            if x.pd_index is None:
                continue
            lines.append('{  // rxn %d' % x.index)
            inner = ['double pm;']
            need_kf = (not x.thd) and (x.low is None or x.high is None)
            if need_kf:
                inner.append('const double kf = {};'.format(_arr_expr(x.arr, K)))
            inner += _pres_mod_lines(x, K)
            inner.append('pres_mod[{} * ld + i] = pm;'.format(x.pd_index))
            lines += ['    ' + l for l in inner]
            lines.append('}')
        lines.append('(void)logT; (void)m;')
        body.append(
            '__global__ __launch_bounds__(crs::kBlock) void k_get_rxn_pres_mod(\n'
        '    const double *__restrict__ K,\n'
            '    const double *__restrict__ T_, const double *__restrict__ P_,\n'
            '    const double *__restrict__ Cin, double *__restrict__ pres_mod,\n'
            '    int64_t n, int64_t ld)\n{\n' + _indent(lines, 4) + '}\n\n')

    # ---- stage kernel: eval_spec_rates --------------------------------
    lines = ['const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;',
             'if (i >= n) return;',
             'double w[{}];'.format(nsp)]
    seen = [False] * nsp
    for x in ir.rxns:
        inner = ['const double f = fwd[{} * ld + i];'.format(x.index)]
        if x.rev:
            inner.append('const double r = rev[{} * ld + i];'.format(x.rev_index))
        if x.pd_index is not None:
            inner.append('const double pm = pres_mod[{} * ld + i];'.format(x.pd_index))
        inner += _spec_lines(x, seen)
        lines.append('{  // rxn %d' % x.index)
        lines += ['    ' + l for l in inner]
        lines.append('}')
    for k in range(nsp):
        if not seen[k]:
            lines.append('w[{}] = 0.0;'.format(k))
    lines += ['sp_rates[{k} * ld + i] = w[{k}];'.format(k=k)
              for k in range(nsp - 1)]
    lines.append('dy_N[i] = w[{}];'.format(nsp - 1))
    body.append(
        '__global__ __launch_bounds__(crs::kBlock) void k_eval_spec_rates(\n'
        '    const double *__restrict__ K,\n'
        '    const double *__restrict__ fwd, const double *__restrict__ rev,\n'
        '    const double *__restrict__ pres_mod, double *__restrict__ sp_rates,\n'
        '    double *__restrict__ dy_N, int64_t n, int64_t ld)\n{\n'
        '#pragma clang fp contract(off)\n' + _indent(lines, 4) + '}\n\n')

    # ---- NASA coefficient table in LDS (option nasa='lds') -------------
    nasa_fill = ''
    if 'lds' in (opts.get('nasa'), dopts.get('nasa')):
        # the table holds s/R - h/RT in the units of the exponential that
        # reads it: log2 e times the coefficients for crs::exp2_64
        base2 = {o.get('exp') == 'crs' for o in (opts, dopts)
                 if o.get('nasa') == 'lds'}
        if len(base2) > 1:
            raise ValueError("the NASA table is shared by the wdot and dydt "
                             "kernels: give both the same 'exp' option")
        vals = _nasa_table_values(ir, _LOG2E if base2.pop() else 1.0)
        body.append('__device__ const double crs_nasa_tab[{}] = {{\n'.format(
            len(vals)) + ''.join('{},{}'.format(_lit(v), '\n' if q % 8 == 7
                                               else ' ')
                                 for q, v in enumerate(vals)) + '};\n')
        body.append('__shared__ __align__(16) double crs_nasa_lds[{}];\n\n'.format(
            len(vals)))
        nasa_fill = ('    for (int q = threadIdx.x; q < {n}; q += blockDim.x)\n'
                     '        crs_nasa_lds[q] = crs_nasa_tab[q];\n'
                     '    __syncthreads();\n'.format(n=len(vals)))

    def fused_state(name, o, tab=None, **kw):
        """Lines of one fused per-state body and the fill code of its
        kernel; with consts_rx='lds' (every fused kernel) or 'lds_fp64'
        (the FP64 ones) the body's reaction-block literals live in an LDS
        table (declared here, filled by the kernel; named after ``tab``,
        by default ``name``, which also picks the group size)."""
        mode = opts.get('consts_rx')
        tab = tab or name
        nfill = nasa_fill if o.get('nasa') == 'lds' else ''
        # reactions per basic block, per kernel (split_f64, split_mixed,
        # split_f64dydt, split_mixeddydt, split_dydt) over the common
        # split_blocks
        if 'split_' + name in o:
            o = dict(o, split_blocks=o['split_' + name])
        if mode not in ('lds', 'lds_fp64') or (mode == 'lds_fp64' and
                                                name == 'mixed'):
            return _fused_fast_lines(ir, K, o, **kw), nfill
        tab = []
        vals = []
        lines = _fused_fast_lines(ir, K, o, ktab=vals,
                                  ktab_name='crs_kt_' + tab, **kw)
        n = max(len(vals), 1)
        body.append('__device__ const double crs_kt_{}_src[{}] = {{\n'.format(
            tab, n) + ''.join('{},{}'.format(_lit(v), '\n' if q % 8 == 7
                                            else ' ')
                              for q, v in enumerate(vals or [0.0])) + '};\n')
        body.append('__shared__ __align__(16) double crs_kt_{}[{}];\n\n'.format(
            tab, n))
        fill = ('    for (int q = threadIdx.x; q < {n}; q += blockDim.x)\n'
                '        crs_kt_{name}[q] = crs_kt_{name}_src[q];\n'
                '    __syncthreads();\n'.format(n=n, name=tab))
        return lines, nfill + fill

    def fused_kernel(state, kernel, lines, fill, note=''):
        """A fused per-state body as ``state`` and its batch/subset kernel
        ``kernel`` (the signature every fused kernel shares)."""
        body.append(
            note + '__device__ __forceinline__ void ' + state + '(\n'
            '    const double *__restrict__ K,\n'
            '    const double *__restrict__ T_, const double *__restrict__ P_,\n'
            '    const double *__restrict__ Y, double *__restrict__ W,\n'
            '    int64_t ld, int64_t i' + pre_arg + ')\n{\n' +
            _indent(lines, 4) + '}\n\n')
        body.append(
            '__global__ __launch_bounds__(kWdotBlock' + lb + ') void ' +
            kernel + '(\n'
            '    const double *__restrict__ K,\n'
            '    const double *__restrict__ T_, const double *__restrict__ P_,\n'
            '    const double *__restrict__ Y, double *__restrict__ W,\n'
            '    const int32_t *__restrict__ idx, int64_t m,\n'
            '    const int32_t *__restrict__ m_dev, int64_t ld)\n{\n' +
            _idx_kernel_tail(fill, state + '(K, T_, P_, Y, W, ld, i{pre});',
                             preload))

    rates_note = (
        '// The same body storing each species\' rate as the CONP dydt\'s\n'
        '// w_k W_k / rho, species k in row k + 1 and the bath gas in row\n'
        '// 0: the rates crs_dydt and crs_rk4 hand to their energy stage.\n')

    # ---- fused FP64 per-state body ------------------------------------
    # preload needs a fill to overlap with: the NASA table's
    preload = (bool(opts.get('preload')) and opts['fused'] == 'fast' and
               opts.get('nasa') == 'lds' and dopts.get('nasa') == 'lds')
    if opts['fused'] == 'fast':
        f64_lines, f64_fill = fused_state('f64', dict(
            opts, preload=preload, block_expr='kWdotBlock'))
    else:
        f64_lines, f64_fill = _fused_f64_lines(ir, K, opts), (
            nasa_fill if opts.get('nasa') == 'lds' else '')
    pre_arg = ', double T_pre, double P_pre' if preload else ''
    body.append(
        '__device__ __forceinline__ void wdot_f64_state(\n'
        '    const double *__restrict__ K,\n'
        '    const double *__restrict__ T_, const double *__restrict__ P_,\n'
        '    const double *__restrict__ Y, double *__restrict__ W,\n'
        '    int64_t ld, int64_t i' + pre_arg + ')\n{\n' +
        _indent(f64_lines, 4) + '}\n\n')
    # one kernel serves the full batch (idx == nullptr) and subsets
    body.append(
        '__global__ __launch_bounds__(kWdotBlock' + lb + ') void k_wdot_f64(\n'
        '    const double *__restrict__ K,\n'
        '    const double *__restrict__ T_, const double *__restrict__ P_,\n'
        '    const double *__restrict__ Y, double *__restrict__ W,\n'
        '    const int32_t *__restrict__ idx, int64_t m,\n'
        '    const int32_t *__restrict__ m_dev, int64_t ld)\n{\n' +
        _idx_kernel_tail(f64_fill,
                         'wdot_f64_state(K, T_, P_, Y, W, ld, i{pre});',
                         preload))
    if dydt_split:
        r_lines, r_fill = fused_state('f64dydt', dict(
            opts, preload=preload, block_expr='kWdotBlock', dydt_rates=True),
            tab='f64_dydt')
        fused_kernel('wdot_dydt_state', 'k_wdot_dydt', r_lines, r_fill,
                     rates_note)

    # ---- constant-pressure dydt (FP64), the caller of the path -------
    if dydt_split:
        body.append(
            '// Energy stage of the CONP dydt: the state Y (row 0 = T, rows\n'
            '// 1..NSP-1 the mass fractions) and the rates S the wdot kernel\n'
            '// stored scaled for the dydt (row k + 1 = dY_k/dt, row 0 the\n'
            '// bath gas\'s w W / rho) give dT/dt and dY_k/dt, handed to the\n'
            '// output object (fetch(row) ahead of a chunk, put(row, value)).\n'
            '// No contraction: the values are the same bits in\n'
            '// k_dydt_energy and in k_rk4_stage.\n'
            'template <class Out>\n'
            '__device__ __forceinline__ void dydt_energy_state(\n'
            '    const double *Y, const double *S, int64_t ld, int64_t i,\n'
            '    Out &out)\n{\n'
            '#pragma clang fp contract(off)\n' +
            _indent(_dydt_energy_lines(ir, K), 4) + '}\n\n')
        body.append(
            '// crs_dydt\'s output: the wdot kernel wrote dY_k/dt into dy\n'
            '// rows 1..NSP-1 already; only row 0 changes, from the bath\n'
            '// gas\'s scaled rate (read first) to dT/dt.\n'
            'struct DydtOut {\n'
            '    double *dy;\n'
            '    int64_t ld;\n'
            '    uint32_t io;\n'
            '    __device__ __forceinline__ void fetch(int) {}\n'
            '    __device__ __forceinline__ void put(int r, double v)\n'
            '    {\n'
            '        if (r == 0) crs::st_row(dy, io, v);\n'
            '    }\n'
            '};\n\n'
            '__global__ __launch_bounds__(crs::kBlock) void k_dydt_energy(\n'
            '    const double *__restrict__ Y, double *S, int64_t n,\n'
            '    int64_t ld)\n{\n'
            '    const int64_t i = (int64_t)blockIdx.x * blockDim.x + '
            'threadIdx.x;\n'
            '    if (i >= n) return;\n'
            '    DydtOut out{S, ld, (uint32_t)i * 8u};\n'
            '    dydt_energy_state(Y, S, ld, i, out);\n}\n\n')
        body.append(
            '// One RK4 stage: k = dydt(ys) from the scaled rates the wdot\n'
            '// kernel wrote for the stage state ys (the energy stage in\n'
            '// registers),\n'
            '// then, without contraction, the textbook combination:\n'
            '//   1: acc = k,      ytmp = y + h/2 k\n'
            '//   2: acc += 2 k,   ytmp = y + h/2 k\n'
            '//   3: acc += 2 k,   ytmp = y + h k\n'
            '//   4: y += h/6 (acc + k); a state whose T first reaches\n'
            '//      T0 + dT_ign gets t_ign = t_step, the time at the start of\n'
            '//      this step, as the reference records t before its t += h\n'
            '//      (create_rate_subs.py:267-276).  NaN marks a state that has\n'
            '//      not ignited yet, so any finite time (a negative t0 too) is\n'
            '//      kept once set.\n'
            '// fetch(row) loads the row\'s y and acc ahead of the chunk\'s\n'
            '// stores.  No pointer is restrict-qualified on purpose: the\n'
            '// possible aliasing keeps each chunk\'s loads behind the\n'
            '// previous chunk\'s stores, so the compiler cannot hoist every\n'
            '// row of the state into registers (it would spill).\n'
            'template <int STAGE>\n'
            'struct Rk4Out {\n'
            '    double *y, *ytmp, *acc;\n'
            '    int64_t ld;\n'
            '    uint32_t io;\n'
            '    double h;\n'
            '    double yv[NSP], av[NSP];\n'
            '    __device__ __forceinline__ void fetch(int r)\n'
            '    {\n'
            '        yv[r] = crs::ld_row(y + r * ld, io);\n'
            '        av[r] = STAGE > 1 ? crs::ld_row(acc + r * ld, io) : 0.0;\n'
            '    }\n'
            '    __device__ __forceinline__ void put(int r, double kv)\n'
            '    {\n'
            '#pragma clang fp contract(off)\n'
            '        if (STAGE == 1) {\n'
            '            crs::st_row(acc + r * ld, io, kv);\n'
            '            crs::st_row(ytmp + r * ld, io, yv[r] + 0.5 * h * kv);\n'
            '        } else if (STAGE == 2) {\n'
            '            crs::st_row(acc + r * ld, io, av[r] + 2.0 * kv);\n'
            '            crs::st_row(ytmp + r * ld, io, yv[r] + 0.5 * h * kv);\n'
            '        } else if (STAGE == 3) {\n'
            '            crs::st_row(acc + r * ld, io, av[r] + 2.0 * kv);\n'
            '            crs::st_row(ytmp + r * ld, io, yv[r] + h * kv);\n'
            '        } else {\n'
            '            crs::st_row(y + r * ld, io,\n'
            '                        yv[r] + (h / 6.0) * (av[r] + kv));\n'
            '        }\n'
            '    }\n'
            '};\n\n'
            '// (at least four waves per SIMD: the kernel is memory bound)\n'
            'template <int STAGE>\n'
            '__global__ __launch_bounds__(crs::kBlock, 4) void k_rk4_stage(\n'
            '    const double *wk, double *y, double *ytmp, double *acc,\n'
            '    int64_t n, int64_t ld, double h, double t_step,\n'
            '    const double *__restrict__ T0, double *__restrict__ t_ign,\n'
            '    double dT_ign)\n{\n'
            '    const int64_t i = (int64_t)blockIdx.x * blockDim.x + '
            'threadIdx.x;\n'
            '    if (i >= n) return;\n'
            '    Rk4Out<STAGE> out{y, ytmp, acc, ld, (uint32_t)i * 8u, h, {}, {}};\n'
            '    dydt_energy_state(STAGE == 1 ? y : ytmp, wk, ld, i, out);\n'
            '    if (STAGE == 4 && t_ign != nullptr && t_ign[i] != t_ign[i] &&\n'
            '        y[i] >= T0[i] + dT_ign)\n'
            '        t_ign[i] = t_step;\n}\n\n')
    else:
        dydt_lines, dydt_fill = fused_state(
            'dydt', dict(dopts, preload=preload), dydt=True)
        body.append(
            '__device__ __forceinline__ void dydt_f64_state(\n'
            '    const double *__restrict__ K, const double *__restrict__ P_,\n'
            '    const double *__restrict__ Y, double *__restrict__ W,\n'
            '    int64_t ld, int64_t i' + pre_arg + ')\n{\n' +
            _indent(dydt_lines, 4) + '}\n\n')
        if preload and dydt_fill:
            dydt_tail = (
                '    const int64_t i = (int64_t)blockIdx.x * blockDim.x + '
                'threadIdx.x;\n'
                '    const bool live = i < n;\n'
                '    const double T_pre = live ? Y[i] : 0.0;\n'
                '    const double P_pre = live ? P_[i] : 0.0;\n' + dydt_fill +
                '    if (!live) return;\n'
                '    dydt_f64_state(K, P_, Y, W, ld, i, T_pre, P_pre);\n}\n\n')
        else:
            dydt_tail = (
                dydt_fill +
                '    const int64_t i = (int64_t)blockIdx.x * blockDim.x + '
                'threadIdx.x;\n'
                '    if (i >= n) return;\n'
                '    dydt_f64_state(K, P_, Y, W, ld, i);\n}\n\n')
        body.append(
            '__global__ __launch_bounds__(crs::kBlock' + dlb + ') void k_dydt_f64(\n'
            '    const double *__restrict__ K, const double *__restrict__ P_,\n'
            '    const double *__restrict__ Y, double *__restrict__ W, int64_t n,\n'
            '    int64_t ld)\n{\n' + dydt_tail)

        # ---- RK4 stage combination for the batched integrator ------------
        body.append(
            '// One RK4 stage on the state rows 0..NSP-1 (row 0 = T), written\n'
            '// without contraction so that it is the textbook sequence exactly:\n'
            '//   1: acc = k,      ytmp = y + h/2 k\n'
            '//   2: acc += 2 k,   ytmp = y + h/2 k\n'
            '//   3: acc += 2 k,   ytmp = y + h k\n'
            '//   4: y += h/6 (acc + k); a state whose T first reaches\n'
            '//      T0 + dT_ign gets t_ign = t_step, the time at the start of\n'
            '//      this step, as the reference records t before its t += h\n'
            '//      (create_rate_subs.py:267-276).  NaN marks a state that has\n'
            '//      not ignited yet, so any finite time (a negative t0 too) is\n'
            '//      kept once set.\n'
            '__global__ __launch_bounds__(crs::kBlock) void k_rk4_combine(\n'
            '    double *__restrict__ y, double *__restrict__ ytmp,\n'
            '    double *__restrict__ acc, const double *__restrict__ k,\n'
            '    int64_t n, int64_t ld, int stage, double h, double t_step,\n'
            '    const double *__restrict__ T0, double *__restrict__ t_ign,\n'
            '    double dT_ign)\n{\n'
            '#pragma clang fp contract(off)\n'
            '    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;\n'
            '    if (i >= n) return;\n'
            '    for (int r = 0; r < NSP; ++r) {\n'
            '        const int64_t o = (int64_t)r * ld + i;\n'
            '        const double kv = k[o];\n'
            '        if (stage == 1) {\n'
            '            acc[o] = kv;\n'
            '            ytmp[o] = y[o] + 0.5 * h * kv;\n'
            '        } else if (stage == 2) {\n'
            '            acc[o] += 2.0 * kv;\n'
            '            ytmp[o] = y[o] + 0.5 * h * kv;\n'
            '        } else if (stage == 3) {\n'
            '            acc[o] += 2.0 * kv;\n'
            '            ytmp[o] = y[o] + h * kv;\n'
            '        } else {\n'
            '            y[o] += (h / 6.0) * (acc[o] + kv);\n'
            '        }\n'
            '    }\n'
            '    if (stage == 4 && t_ign != nullptr && t_ign[i] != t_ign[i] &&\n'
            '        y[i] >= T0[i] + dT_ign)\n'
            '        t_ign[i] = t_step;\n}\n\n')

    # ---- fused mixed-precision per-state body -------------------------
    if opts.get('mixed', True):
        mixed_lines, mixed_fill = fused_state(
            'mixed', dict(opts, preload=preload, block_expr='kWdotBlock'),
            mixed=True)
        body.append(
            '__device__ __forceinline__ void wdot_mixed_state(\n'
            '    const double *__restrict__ K,\n'
            '    const double *__restrict__ T_, const double *__restrict__ P_,\n'
            '    const double *__restrict__ Y, double *__restrict__ W,\n'
            '    int64_t ld, int64_t i' + pre_arg + ')\n{\n' +
            _indent(mixed_lines, 4) + '}\n\n')
        body.append(
            '__global__ __launch_bounds__(kWdotBlock' + lb + ') void k_wdot_mixed(\n'
            '    const double *__restrict__ K,\n'
            '    const double *__restrict__ T_, const double *__restrict__ P_,\n'
            '    const double *__restrict__ Y, double *__restrict__ W,\n'
            '    const int32_t *__restrict__ idx, int64_t m,\n'
            '    const int32_t *__restrict__ m_dev, int64_t ld)\n{\n' +
            _idx_kernel_tail(mixed_fill,
                             'wdot_mixed_state(K, T_, P_, Y, W, ld, i{pre});',
                             preload))
        if dydt_split:
            r_lines, r_fill = fused_state(
                'mixeddydt', dict(opts, preload=preload,
                                  block_expr='kWdotBlock', dydt_rates=True),
                tab='mixed_dydt', mixed=True)
            fused_kernel('wdot_mixed_dydt_state', 'k_wdot_mixed_dydt',
                         r_lines, r_fill, rates_note)

    header = (
        '// Generated by create_rate_subs_amd.hip_emitter for mechanism '
        '"{name}".\n'
        '// {nsp} species, {nr} reactions ({nrev} reversible, {npd} '
        'third-body/falloff).\n'
        '#include "crs_common.h"\n\n'
        'namespace {{\n\n'
        'constexpr int NSP = {nsp};\nconstexpr int NR = {nr};\n'
        'constexpr int NREV = {nrev};\nconstexpr int NPD = {npd};\n\n'
        .format(name=ir.name, nsp=nsp, nr=nr, nrev=nrev, npd=npd))
    consts = K.definition() + '\n'
    abi = _abi_text(ir, opts.get('mixed', True), dydt_split)
    return header + consts + ''.join(body) + '}  // namespace\n\n' + abi


# crs_dydt and crs_rk4: the split form (wdot kernel, then the energy stage or
# the energy stage with the RK4 combination) and the fused one (k_dydt_f64,
# then k_rk4_combine)
_DYDT_RK4_SPLIT = '''int crs_dydt(const double *P, const double *y, double *dy, int64_t n,
             int64_t ld, int mode, void *stream)
{{
    const char *fn = "crs_dydt";
    CRS_TRY(crs::check_sizes(fn, n, ld));
    CRS_TRY(check_mode(fn, mode));
    CRS_TRY(crs::check_ptr(fn, P, "P", n));
    CRS_TRY(crs::check_ptr(fn, y, "y", n));
    CRS_TRY(crs::check_ptr(fn, dy, "dy", n));
    // dy is written while y and P are still read
    CRS_TRY(crs::check_disjoint(fn, y, NSP, "y", dy, NSP, "dy", n, ld));
    CRS_TRY(crs::check_disjoint(fn, P, 1, "P", dy, NSP, "dy", n, ld));
    if (n == 0) return CRS_OK;
    hipStream_t s = (hipStream_t)stream;
    const double *K = nullptr;
    CRS_TRY(table(fn, &K));
    // the mode's fused kernel stores the state's (T = row 0, Y = rows
    // 1..) dY_k/dt = w_k W_k / rho into dy rows 1.. and the bath gas's
    // w W / rho into row 0; the FP64 energy stage then writes dT/dt there
    for (int64_t c0 = 0; c0 < n; c0 += crs::kMaxChunk) {{
        const int64_t c = n - c0 < crs::kMaxChunk ? n - c0 : crs::kMaxChunk;
@MIXED_RATES@        k_wdot_dydt<<<crs::grid_for(c, kWdotBlock), kWdotBlock, 0, s>>>(
            K, y + c0, P + c0, y + ld + c0, dy + c0, nullptr, c, nullptr, ld);
        k_dydt_energy<<<crs::grid_for(c), crs::kBlock, 0, s>>>(
            y + c0, dy + c0, c, ld);
    }}
    return crs::launched(fn);
}}

int crs_rk4(const double *P, double *y, double *work, int64_t n, int64_t ld,
            double t0, double h, int nsteps, const double *T0, double *t_ign,
            double dT_ign, void *stream)
{{
    const char *fn = "crs_rk4";
    CRS_TRY(crs::check_sizes(fn, n, ld));
    if (nsteps < 0)
        return crs::fail(CRS_ERR_ARG, "%s: nsteps must be >= 0", fn);
    if (!(h > 0.0))
        return crs::fail(CRS_ERR_ARG, "%s: h must be > 0", fn);
    CRS_TRY(crs::check_ptr(fn, P, "P", n));
    CRS_TRY(crs::check_ptr(fn, y, "y", n));
    CRS_TRY(crs::check_ptr(fn, work, "work", n));
    if (n > crs::kMaxChunk)
        return crs::fail(CRS_ERR_ARG, "%s: n = %lld exceeds %lld states per "
                         "call", fn, (long long)n, (long long)crs::kMaxChunk);
    if ((T0 == nullptr) != (t_ign == nullptr))
        return crs::fail(CRS_ERR_ARG, "%s: T0 and t_ign go together", fn);
    // y, work and t_ign are written while P, T0 and each other are read:
    // no two of them may share an entry
    CRS_TRY(crs::check_disjoint(fn, y, NSP, "y", work, 3 * NSP, "work", n, ld));
    CRS_TRY(crs::check_disjoint(fn, P, 1, "P", y, NSP, "y", n, ld));
    CRS_TRY(crs::check_disjoint(fn, P, 1, "P", work, 3 * NSP, "work", n, ld));
    if (t_ign) {{
        CRS_TRY(crs::check_disjoint(fn, t_ign, 1, "t_ign", y, NSP, "y", n, ld));
        CRS_TRY(crs::check_disjoint(fn, t_ign, 1, "t_ign", work, 3 * NSP,
                                    "work", n, ld));
        CRS_TRY(crs::check_disjoint(fn, t_ign, 1, "t_ign", P, 1, "P", n, ld));
        CRS_TRY(crs::check_disjoint(fn, t_ign, 1, "t_ign", T0, 1, "T0", n, ld));
        CRS_TRY(crs::check_disjoint(fn, T0, 1, "T0", y, NSP, "y", n, ld));
        CRS_TRY(crs::check_disjoint(fn, T0, 1, "T0", work, 3 * NSP, "work",
                                    n, ld));
    }}
    if (n == 0 || nsteps == 0) return CRS_OK;
    hipStream_t s = (hipStream_t)stream;
    const double *K = nullptr;
    CRS_TRY(table(fn, &K));
    double *ytmp = work;
    double *acc = work + (int64_t)NSP * ld;
    double *wk = work + 2 * (int64_t)NSP * ld;
    const dim3 grid = crs::grid_for(n);
    const dim3 wgrid = crs::grid_for(n, kWdotBlock);
    double t = t0;
    for (int step = 0; step < nsteps; ++step) {{
        for (int stage = 1; stage <= 4; ++stage) {{
            const double *ys = stage == 1 ? y : ytmp;
            k_wdot_dydt<<<wgrid, kWdotBlock, 0, s>>>(K, ys, P, ys + ld, wk,
                                                      nullptr, n, nullptr, ld);
            if (stage == 1)
                k_rk4_stage<1><<<grid, crs::kBlock, 0, s>>>(
                    wk, y, ytmp, acc, n, ld, h, t, T0, t_ign, dT_ign);
            else if (stage == 2)
                k_rk4_stage<2><<<grid, crs::kBlock, 0, s>>>(
                    wk, y, ytmp, acc, n, ld, h, t, T0, t_ign, dT_ign);
            else if (stage == 3)
                k_rk4_stage<3><<<grid, crs::kBlock, 0, s>>>(
                    wk, y, ytmp, acc, n, ld, h, t, T0, t_ign, dT_ign);
            else
                k_rk4_stage<4><<<grid, crs::kBlock, 0, s>>>(
                    wk, y, ytmp, acc, n, ld, h, t, T0, t_ign, dT_ign);
        }}
        t += h;
    }}
    return crs::launched(fn);
}}

'''

_DYDT_RK4_FUSED = '''int crs_dydt(const double *P, const double *y, double *dy, int64_t n,
             int64_t ld, int mode, void *stream)
{{
    const char *fn = "crs_dydt";
    CRS_TRY(crs::check_sizes(fn, n, ld));
    if (mode != CRS_MODE_FP64)
        return crs::fail(mode == CRS_MODE_MIXED ? CRS_ERR_UNSUPPORTED
                                                : CRS_ERR_ARG,
                         "%s: mode %d not available (FP64 only)", fn, mode);
    CRS_TRY(crs::check_ptr(fn, P, "P", n));
    CRS_TRY(crs::check_ptr(fn, y, "y", n));
    CRS_TRY(crs::check_ptr(fn, dy, "dy", n));
    // dy is written while y and P are still read
    CRS_TRY(crs::check_disjoint(fn, y, NSP, "y", dy, NSP, "dy", n, ld));
    CRS_TRY(crs::check_disjoint(fn, P, 1, "P", dy, NSP, "dy", n, ld));
    if (n == 0) return CRS_OK;
    hipStream_t s = (hipStream_t)stream;
    const double *K = nullptr;
    CRS_TRY(table(fn, &K));
    for (int64_t c0 = 0; c0 < n; c0 += crs::kMaxChunk) {{
        const int64_t c = n - c0 < crs::kMaxChunk ? n - c0 : crs::kMaxChunk;
        k_dydt_f64<<<crs::grid_for(c), crs::kBlock, 0, s>>>(
            K, P + c0, y + c0, dy + c0, c, ld);
    }}
    return crs::launched(fn);
}}

int crs_rk4(const double *P, double *y, double *work, int64_t n, int64_t ld,
            double t0, double h, int nsteps, const double *T0, double *t_ign,
            double dT_ign, void *stream)
{{
    const char *fn = "crs_rk4";
    CRS_TRY(crs::check_sizes(fn, n, ld));
    if (nsteps < 0)
        return crs::fail(CRS_ERR_ARG, "%s: nsteps must be >= 0", fn);
    if (!(h > 0.0))
        return crs::fail(CRS_ERR_ARG, "%s: h must be > 0", fn);
    CRS_TRY(crs::check_ptr(fn, P, "P", n));
    CRS_TRY(crs::check_ptr(fn, y, "y", n));
    CRS_TRY(crs::check_ptr(fn, work, "work", n));
    if (n > crs::kMaxChunk)
        return crs::fail(CRS_ERR_ARG, "%s: n = %lld exceeds %lld states per "
                         "call", fn, (long long)n, (long long)crs::kMaxChunk);
    if ((T0 == nullptr) != (t_ign == nullptr))
        return crs::fail(CRS_ERR_ARG, "%s: T0 and t_ign go together", fn);
    // y, work and t_ign are written while P, T0 and each other are read:
    // no two of them may share an entry
    CRS_TRY(crs::check_disjoint(fn, y, NSP, "y", work, 3 * NSP, "work", n, ld));
    CRS_TRY(crs::check_disjoint(fn, P, 1, "P", y, NSP, "y", n, ld));
    CRS_TRY(crs::check_disjoint(fn, P, 1, "P", work, 3 * NSP, "work", n, ld));
    if (t_ign) {{
        CRS_TRY(crs::check_disjoint(fn, t_ign, 1, "t_ign", y, NSP, "y", n, ld));
        CRS_TRY(crs::check_disjoint(fn, t_ign, 1, "t_ign", work, 3 * NSP,
                                    "work", n, ld));
        CRS_TRY(crs::check_disjoint(fn, t_ign, 1, "t_ign", P, 1, "P", n, ld));
        CRS_TRY(crs::check_disjoint(fn, t_ign, 1, "t_ign", T0, 1, "T0", n, ld));
        CRS_TRY(crs::check_disjoint(fn, T0, 1, "T0", y, NSP, "y", n, ld));
        CRS_TRY(crs::check_disjoint(fn, T0, 1, "T0", work, 3 * NSP, "work",
                                    n, ld));
    }}
    if (n == 0 || nsteps == 0) return CRS_OK;
    hipStream_t s = (hipStream_t)stream;
    const double *K = nullptr;
    CRS_TRY(table(fn, &K));
    double *ytmp = work;
    double *acc = work + (int64_t)NSP * ld;
    double *k = work + 2 * (int64_t)NSP * ld;
    const dim3 grid = crs::grid_for(n);
    double t = t0;
    for (int step = 0; step < nsteps; ++step) {{
        for (int stage = 1; stage <= 4; ++stage) {{
            k_dydt_f64<<<grid, crs::kBlock, 0, s>>>(K, P, stage == 1 ? y : ytmp,
                                                    k, n, ld);
            k_rk4_combine<<<grid, crs::kBlock, 0, s>>>(
                y, ytmp, acc, k, n, ld, stage, h, t, T0, t_ign, dT_ign);
        }}
        t += h;
    }}
    return crs::launched(fn);
}}

'''


def _abi_text(ir, mixed=True, dydt_split=True):
    npd = ir.npd
    modes = '(1 << CRS_MODE_FP64)' + (' | (1 << CRS_MODE_MIXED)' if mixed else '')
    if mixed:
        launch_full = '''        if (mode == CRS_MODE_MIXED)
            k_wdot_mixed<<<crs::grid_for(c, kWdotBlock), kWdotBlock, 0, s>>>(
                K, Tc, Pc, Yc, Wc, nullptr, c, nullptr, ld);
        else
'''
        launch_sub = '''    if (mode == CRS_MODE_MIXED)
        k_wdot_mixed<<<crs::grid_for(m, kWdotBlock), kWdotBlock, 0, s>>>(
            K, T, P, Y, wdot, idx, m, m_dev, ld);
    else
'''
        mixed_ok = 'mode == CRS_MODE_MIXED'
        mixed_rates = '''        if (mode == CRS_MODE_MIXED)
            k_wdot_mixed_dydt<<<crs::grid_for(c, kWdotBlock), kWdotBlock, 0,
                                s>>>(K, y + c0, P + c0, y + ld + c0, dy + c0,
                                     nullptr, c, nullptr, ld);
        else
'''
    else:
        launch_full = launch_sub = mixed_rates = ''
        mixed_ok = 'false'
    pm_launch = ('    k_get_rxn_pres_mod<<<crs::grid_for(n), crs::kBlock, 0, s>>>'
                 '(K, T, P, C, pres_mod, n, ld);\n' if npd else '')
    return '''// Device address of the constant table, looked up once per device.
static int table(const char *fn, const double **out)
{{
    static std::mutex lock;
    static const double *addr[64] = {{nullptr}};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
        return crs::fail(CRS_ERR_HIP, "%s: no current HIP device", fn);
    std::lock_guard<std::mutex> g(lock);
    if (addr[dev] == nullptr) {{
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(crs_K));
        if (e != hipSuccess)
            return crs::fail(CRS_ERR_HIP, "%s: constant table: %s", fn,
                             hipGetErrorString(e));
        addr[dev] = (const double *)p;
    }}
    *out = addr[dev];
    return CRS_OK;
}}

extern "C" {{

int crs_get_mech_info(crs_mech_info *out)
{{
    if (out == nullptr)
        return crs::fail(CRS_ERR_ARG, "crs_get_mech_info: out is NULL");
    out->nsp = NSP;
    out->nr = NR;
    out->nrev = NREV;
    out->npd = NPD;
    out->modes = {modes};
    out->name = "{name}";
    return CRS_OK;
}}

const char *crs_last_error(void) {{ return crs::error_buffer(); }}

int crs_eval_conc(const double *T, const double *P, const double *Y,
                  double *y_N, double *mw_avg, double *rho, double *conc,
                  int64_t n, int64_t ld, void *stream)
{{
    const char *fn = "crs_eval_conc";
    CRS_TRY(crs::check_sizes(fn, n, ld));
    CRS_TRY(crs::check_ptr(fn, T, "T", n));
    CRS_TRY(crs::check_ptr(fn, P, "P", n));
    CRS_TRY(crs::check_ptr(fn, Y, "Y", n));
    CRS_TRY(crs::check_ptr(fn, y_N, "y_N", n));
    CRS_TRY(crs::check_ptr(fn, mw_avg, "mw_avg", n));
    CRS_TRY(crs::check_ptr(fn, rho, "rho", n));
    CRS_TRY(crs::check_ptr(fn, conc, "conc", n));
    if (n == 0) return CRS_OK;
    hipStream_t s = (hipStream_t)stream;
    const double *K = nullptr;
    CRS_TRY(table(fn, &K));
    k_eval_conc<<<crs::grid_for(n), crs::kBlock, 0, s>>>(K, T, P, Y, y_N, mw_avg,
                                                         rho, conc, n, ld);
    return crs::launched(fn);
}}

int crs_eval_rxn_rates(const double *T, const double *P, const double *C,
                       double *fwd, double *rev, int64_t n, int64_t ld,
                       void *stream)
{{
    const char *fn = "crs_eval_rxn_rates";
    CRS_TRY(crs::check_sizes(fn, n, ld));
    CRS_TRY(crs::check_ptr(fn, T, "T", n));
    CRS_TRY(crs::check_ptr(fn, P, "P", n));
    CRS_TRY(crs::check_ptr(fn, C, "C", n));
    CRS_TRY(crs::check_ptr(fn, fwd, "fwd", n));
    CRS_TRY(crs::check_ptr(fn, rev, "rev", NREV > 0 ? n : 0));
    if (n == 0) return CRS_OK;
    hipStream_t s = (hipStream_t)stream;
    const double *K = nullptr;
    CRS_TRY(table(fn, &K));
    k_eval_rxn_rates<<<crs::grid_for(n), crs::kBlock, 0, s>>>(K, T, P, C, fwd,
                                                              rev, n, ld);
    return crs::launched(fn);
}}

int crs_get_rxn_pres_mod(const double *T, const double *P, const double *C,
                         double *pres_mod, int64_t n, int64_t ld,
                         void *stream)
{{
    const char *fn = "crs_get_rxn_pres_mod";
    CRS_TRY(crs::check_sizes(fn, n, ld));
    if (NPD == 0) return CRS_OK;
    CRS_TRY(crs::check_ptr(fn, T, "T", n));
    CRS_TRY(crs::check_ptr(fn, P, "P", n));
    CRS_TRY(crs::check_ptr(fn, C, "C", n));
    CRS_TRY(crs::check_ptr(fn, pres_mod, "pres_mod", n));
    if (n == 0) return CRS_OK;
    hipStream_t s = (hipStream_t)stream;
    const double *K = nullptr;
    CRS_TRY(table(fn, &K));
{pm_launch}    return crs::launched(fn);
}}

int crs_eval_spec_rates(const double *fwd, const double *rev,
                        const double *pres_mod, double *sp_rates,
                        double *dy_N, int64_t n, int64_t ld, void *stream)
{{
    const char *fn = "crs_eval_spec_rates";
    CRS_TRY(crs::check_sizes(fn, n, ld));
    CRS_TRY(crs::check_ptr(fn, fwd, "fwd", n));
    CRS_TRY(crs::check_ptr(fn, rev, "rev", NREV > 0 ? n : 0));
    CRS_TRY(crs::check_ptr(fn, pres_mod, "pres_mod", NPD > 0 ? n : 0));
    CRS_TRY(crs::check_ptr(fn, sp_rates, "sp_rates", NSP > 1 ? n : 0));
    CRS_TRY(crs::check_ptr(fn, dy_N, "dy_N", n));
    if (n == 0) return CRS_OK;
    hipStream_t s = (hipStream_t)stream;
    const double *K = nullptr;
    CRS_TRY(table(fn, &K));
    k_eval_spec_rates<<<crs::grid_for(n), crs::kBlock, 0, s>>>(
        K, fwd, rev, pres_mod, sp_rates, dy_N, n, ld);
    return crs::launched(fn);
}}

static int check_mode(const char *fn, int mode)
{{
    if (mode == CRS_MODE_FP64 || {mixed_ok}) return CRS_OK;
    if (mode == CRS_MODE_MIXED)
        return crs::fail(CRS_ERR_UNSUPPORTED,
                         "%s: mode %d not compiled into this library", fn, mode);
    return crs::fail(CRS_ERR_ARG, "%s: unknown mode %d", fn, mode);
}}

int crs_eval_wdot(const double *T, const double *P, const double *Y,
                  double *wdot, int64_t n, int64_t ld, int mode, void *stream)
{{
    const char *fn = "crs_eval_wdot";
    CRS_TRY(crs::check_sizes(fn, n, ld));
    CRS_TRY(check_mode(fn, mode));
    CRS_TRY(crs::check_ptr(fn, T, "T", n));
    CRS_TRY(crs::check_ptr(fn, P, "P", n));
    CRS_TRY(crs::check_ptr(fn, Y, "Y", n));
    CRS_TRY(crs::check_ptr(fn, wdot, "wdot", n));
    if (n == 0) return CRS_OK;
    hipStream_t s = (hipStream_t)stream;
    const double *K = nullptr;
    CRS_TRY(table(fn, &K));
    // the fused kernels address a state with a 32-bit byte offset from its
    // row: batches beyond kMaxChunk states go in chunks of shifted rows
    for (int64_t c0 = 0; c0 < n; c0 += crs::kMaxChunk) {{
        const int64_t c = n - c0 < crs::kMaxChunk ? n - c0 : crs::kMaxChunk;
        const double *Tc = T + c0, *Pc = P + c0, *Yc = Y + c0;
        double *Wc = wdot + c0;
{launch_full_c}        k_wdot_f64<<<crs::grid_for(c, kWdotBlock), kWdotBlock, 0, s>>>(
            K, Tc, Pc, Yc, Wc, nullptr, c, nullptr, ld);
    }}
    return crs::launched(fn);
}}

{dydt_rk4}// Shared by crs_eval_wdot_subset (host count m, m_dev NULL) and
// crs_eval_wdot_subset_dev (at most m states, the count read on the device).
static int wdot_subset(const char *fn, const double *T, const double *P,
                       const double *Y, double *wdot, const int32_t *idx,
                       int64_t m, const int32_t *m_dev, int64_t ld, int mode,
                       void *stream)
{{
    if (m < 0)
        return crs::fail(CRS_ERR_ARG, "%s: m must be >= 0", fn);
    if (ld < 0)
        return crs::fail(CRS_ERR_ARG, "%s: ld must be >= 0", fn);
    CRS_TRY(check_mode(fn, mode));
    CRS_TRY(crs::check_ptr(fn, T, "T", m));
    CRS_TRY(crs::check_ptr(fn, P, "P", m));
    CRS_TRY(crs::check_ptr(fn, Y, "Y", m));
    CRS_TRY(crs::check_ptr(fn, wdot, "wdot", m));
    CRS_TRY(crs::check_ptr(fn, idx, "idx", m));
    if (ld > crs::kMaxChunk)
        return crs::fail(CRS_ERR_ARG, "%s: ld = %lld exceeds %lld (state "
                         "indices are byte offsets of 32 bits)", fn,
                         (long long)ld, (long long)crs::kMaxChunk);
    if (m == 0) return CRS_OK;
    hipStream_t s = (hipStream_t)stream;
    const double *K = nullptr;
    CRS_TRY(table(fn, &K));
{launch_sub}    k_wdot_f64<<<crs::grid_for(m, kWdotBlock), kWdotBlock, 0, s>>>(
        K, T, P, Y, wdot, idx, m, m_dev, ld);
    return crs::launched(fn);
}}

int crs_eval_wdot_subset(const double *T, const double *P, const double *Y,
                         double *wdot, const int32_t *idx, int64_t m,
                         int64_t ld, int mode, void *stream)
{{
    return wdot_subset("crs_eval_wdot_subset", T, P, Y, wdot, idx, m,
                       nullptr, ld, mode, stream);
}}

int crs_eval_wdot_subset_dev(const double *T, const double *P,
                             const double *Y, double *wdot,
                             const int32_t *idx, const int32_t *count,
                             int64_t m_max, int64_t ld, int mode, void *stream)
{{
    const char *fn = "crs_eval_wdot_subset_dev";
    CRS_TRY(crs::check_ptr(fn, count, "count", m_max));
    return wdot_subset(fn, T, P, Y, wdot, idx, m_max, count, ld, mode,
                       stream);
}}

}}  // extern "C"
'''.format(name=ir.name, pm_launch=pm_launch, modes=modes,
           launch_full_c=launch_full, launch_sub=launch_sub, mixed_ok=mixed_ok,
           dydt_rk4=(_DYDT_RK4_SPLIT.replace('@MIXED_RATES@', mixed_rates)
                     if dydt_split else _DYDT_RK4_FUSED)
           .replace('{{', '{').replace('}}', '}'))


def write_hip(out_dir, specs, reacs, name='mech', spec_order=None, opts=None,
              filename=None):
    """Emit ``crs_<name>.hip`` into ``out_dir``; return its path."""
    ir = build_ir(specs, reacs, name=name)
    os.makedirs(out_dir, exist_ok=True)
    path = os.path.join(out_dir, filename or 'crs_{}.hip'.format(name))
    text = emit_hip_source(ir, opts)
    old = None
    if os.path.exists(path):
        with open(path) as fh:
            old = fh.read()
    if old != text:
        with open(path, 'w') as fh:
            fh.write(text)
    return path
