"""C emitter for the CPU oracle (test infrastructure, never shipped).

Restates the C branch of the reference's generator for the four functions
on the wdot path, formula for formula and constant format for constant
format, so that compiling its output gives the subroutines the reference
would have generated for the same mechanism:

* ``eval_conc``          rate_subs.py:1292-1391 (chem_utils.c)
* ``eval_rxn_rates``     rate_subs.py:226-704, Arrhenius text :20-135
* ``get_rxn_pres_mod``   rate_subs.py:707-1073
* ``eval_spec_rates``    rate_subs.py:1076-1235
* ``header.h``           create_rate_subs.py:332-427 (the parts the four
  functions need)
* ``eval_h``, ``eval_cp`` rate_subs.py:1495-1565, 1707-1776 (chem_utils.c)
* ``dydt`` (constant pressure) rate_subs.py:1829-1946, the caller of the
  four functions: energy equation and mass-fraction rates (emitted without
  the reference's ``#if defined(CONP)`` guard)

The reference is Python 2; every float it prints with ``str()`` or an empty
format spec is spelled here with ``utils.py2_str`` (12 significant digits),
and every other literal with the same ``{:.16e}``/``{:.8e}``/``{:.6}``
format the reference uses.  It works directly on the reference's data model
(``SpecInfo``/``ReacInfo`` attributes), not on the HIP emitter's tables, so
that the two code paths share nothing after parsing.

PLOG and Chebyshev rate forms follow rate_subs.py:446-485 and :138-223.
"""

import math
import os

from create_rate_subs_amd import chem
from create_rate_subs_amd.utils import (py2_str, fmt_e16, fmt_e8, fmt_g6,
                                        fmt_w3, is_integer, get_nu)

__all__ = ['arrhenius_text', 'emit_c', 'mech_dims']


def arrhenius_text(A, b, E):
    """Right-hand side for a rate constant, as rate_subs.py:20-135 spells it."""
    if A > 0:
        logA = math.log(A)
        if not E:
            if not b:
                return py2_str(A)
            if isinstance(b, int):
                return py2_str(A) + ' * T' * max(b, 0)
            s = 'exp({}'.format(fmt_e16(logA))
            s += (' + ' + py2_str(b)) if b > 0 else (' - ' + py2_str(abs(b)))
            return s + ' * logT)'
        if not b:
            return 'exp({} - ({} / T))'.format(fmt_e16(logA), fmt_e16(E))
        s = 'exp({}'.format(fmt_e16(logA))
        s += (' + ' + py2_str(b)) if b > 0 else (' - ' + py2_str(abs(b)))
        return s + ' * logT - ({} / T))'.format(fmt_e16(E))
    if A < 0:
        if not E:
            if not b:
                return py2_str(A)
            if is_integer(b):
                return py2_str(A) + ' * T' * max(int(b), 0)
            s = '{} * exp('.format(fmt_e16(A))
            s += py2_str(b) if b > 0 else '-' + py2_str(abs(b))
            return s + ' * logT)'
        if not b:
            return '{} * exp(-({} / T))'.format(fmt_e16(A), fmt_e16(E))
        s = '{} * exp('.format(fmt_e16(A))
        s += py2_str(b) if b > 0 else '-' + py2_str(abs(b))
        return s + ' * logT - ({} / T))'.format(fmt_e16(E))
    raise NotImplementedError('zero pre-exponential factor')


def _conc_product(species, nus, trailing_space):
    """Concentration product for a rate line (rate_subs.py:487-499, :667-680)."""
    s = ''
    for isp, nu in zip(species, nus):
        if is_integer(nu):
            s += 'C[{}] * '.format(isp) * int(nu)
        else:
            s += 'pow(C[{}], {}) *'.format(isp, py2_str(nu))
            if trailing_space:
                s += ' '
    return s


def _gibbs_terms(sp, scale):
    """Lumped Kc coefficients of one species (rate_subs.py:410-431).

    ``scale`` is the signed stoichiometric factor; the seven numbers
    multiply [1, logT, T, T^2, T^3, T^4, -1/T] in the emitted polynomial.
    """
    def one(a):
        raw = [a[6] - a[0], a[0] - 1.0, a[1] / 2.0, a[2] / 6.0,
               a[3] / 12.0, a[4] / 20.0, a[5]]
        return [x * scale for x in raw]
    return one(sp.lo), one(sp.hi)


def _kc_lines(rxn, specs):
    groups = {}
    order = []
    sum_nu = 0

    def add(sp, scale):
        lo, hi = _gibbs_terms(sp, scale)
        key = sp.Trange[1]
        if key not in groups:
            groups[key] = (lo, hi)
            order.append(key)
        else:
            plo, phi = groups[key]
            groups[key] = ([lo[i] + plo[i] for i in range(7)],
                           [hi[i] + phi[i] for i in range(7)])

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

    def poly(c):
        return ('({} + {} * logT + T * ({} + T * ({} + T * ({} + {} * T))) - '
                '{} / T)'.format(*[fmt_e16(x) for x in c]))

    out = ''
    first = True
    for t_mid in order:
        lo, hi = groups[t_mid]
        op = '    Kc = ' if first else '    Kc += '
        out += '  if (T <= {}) {{\n'.format(py2_str(t_mid))
        out += op + poly(lo) + ';\n'
        out += '  } else {\n'
        out += op + poly(hi) + ';\n'
        out += '  }\n\n'
        first = False
    out += '  Kc = {} * exp(Kc);\n'.format(fmt_e16((chem.PA / chem.RU) ** sum_nu))
    return out


def _cheb_text(rxn):
    """Chebyshev kf, rate_subs.py:138-223 (C: pow(10.0, ...))."""
    t0, t1 = rxn.cheb_tlim
    p0, p1 = rxn.cheb_plim
    tsum = 1.0 / t0 + 1.0 / t1
    tsub = 1.0 / t1 - 1.0 / t0
    psum = math.log10(p0) + math.log10(p1)
    psub = math.log10(p1) - math.log10(p0)
    par = rxn.cheb_par
    out = ['  Tred = ((2.0 / T) - {}) / {};\n'.format(fmt_e8(tsum), fmt_e8(tsub)),
           '  Pred = (2.0 * log10(pres) - {}) / {};\n'.format(fmt_e8(psum),
                                                             fmt_e8(psub)),
           '  cheb_temp_0 = 1;\n', '  cheb_temp_1 = Pred;\n']
    for i in range(rxn.cheb_n_temp):
        out.append('  dot_prod[{}]= {} + Pred * {};\n'.format(
            i, fmt_e8(par[i][0]), fmt_e8(par[i][1])))
    one = True
    for jj in range(2, rxn.cheb_n_pres):
        new, old = (1, 0) if one else (0, 1)
        out.append('  cheb_temp_{o} = 2 * Pred * cheb_temp_{n} - cheb_temp_{o};'
                   '\n'.format(o=old, n=new))
        for i in range(rxn.cheb_n_temp):
            out.append('  dot_prod[{}] += {} * cheb_temp_{};\n'.format(
                i, fmt_e8(par[i][jj]), old))
        one = not one
    out += ['  cheb_temp_0 = 1;\n', '  cheb_temp_1 = Tred;\n',
            '  kf = dot_prod[0] + Tred * dot_prod[1];\n']
    one = True
    for i in range(2, rxn.cheb_n_temp):
        new, old = (1, 0) if one else (0, 1)
        out.append('  cheb_temp_{o} = 2 * Tred * cheb_temp_{n} - cheb_temp_{o};'
                   '\n'.format(o=old, n=new))
        out.append('  kf += dot_prod[{}] * cheb_temp_{};\n'.format(i, old))
        one = not one
    out.append('  kf = pow(10.0, kf);\n')
    return ''.join(out)


def _plog_text(rxn):
    """PLOG kf, rate_subs.py:446-485: the bracketing pair interpolated
    linearly in ln k over ln P, the end sets held constant outside."""
    rows = rxn.plog_par
    v = rows[0]
    out = ['  if (pres <= {:.4e}) {{\n'.format(v[0]) +
           '    kf = ' + arrhenius_text(v[1], v[2], v[3]) + ';\n']
    for idx, v in enumerate(rows[:-1]):
        v2 = rows[idx + 1]
        out.append('  }} else if ((pres > {:.4e}) '.format(v[0]) +
                   '&& (pres <= {:.4e})) {{\n'.format(v2[0]))
        out.append('    kf = log(' + arrhenius_text(v[1], v[2], v[3]) + ');\n')
        out.append('    kf2 = log(' + arrhenius_text(v2[1], v2[2], v2[3]) +
                   ');\n')
        out.append('    kf = exp(kf + (kf2 - kf) * (log(pres) - '
                   '{}) / {});\n'.format(
                       fmt_e16(math.log(v[0])),
                       fmt_e16(math.log(v2[0]) - math.log(v[0]))))
    v = rows[-1]
    out.append('  }} else if (pres > {:.4e}) {{\n'.format(v[0]) +
               '    kf = ' + arrhenius_text(v[1], v[2], v[3]) + ';\n' +
               '  }\n')
    return ''.join(out)


def _write_rxn_rates(specs, reacs):
    rev_reacs = [i for i, r in enumerate(reacs) if r.rev]
    body = ['#include "rates.h"\n',
            'void eval_rxn_rates (const double T, const double pres, '
            'const double * C, double * fwd_rxn_rates, '
            'double * rev_rxn_rates) {\n',
            '  double logT = log(T);\n', '\n']
    kf_declared = False
    if rev_reacs and any(not r.rev_par for r in reacs):
        body.append('  double kf;\n  double Kc;\n')
        kf_declared = True
    if any(r.cheb for r in reacs):
        if not kf_declared:
            body.append('  double kf;\n')
            kf_declared = True
        body.append('  double Tred;\n  double Pred;\n')
        body.append('  double cheb_temp_0, cheb_temp_1;\n')
        body.append('  double dot_prod[{}];\n'.format(
            max(r.cheb_n_temp for r in reacs if r.cheb)))
    if any(r.plog for r in reacs):
        if not kf_declared:
            body.append('  double kf;\n')
        body.append('  double kf2;\n')
    body.append('\n')
    for j, rxn in enumerate(reacs):
        body.append('  //rxn {}\n'.format(j))
        uses_kf = (rxn.rev and not rxn.rev_par) or rxn.plog or rxn.cheb
        if rxn.cheb:
            body.append(_cheb_text(rxn))
        elif rxn.plog:
            body.append(_plog_text(rxn))
        elif uses_kf:
            body.append('  kf = ' + arrhenius_text(rxn.A, rxn.b, rxn.E) +
                        ';\n')
        line = '  fwd_rxn_rates[{}] = '.format(j)
        line += _conc_product(rxn.reac, rxn.reac_nu, trailing_space=False)
        line += 'kf' if uses_kf else arrhenius_text(rxn.A, rxn.b, rxn.E)
        body.append(line + ';\n')
        if rxn.rev:
            if not rxn.rev_par:
                body.append(_kc_lines(rxn, specs))
            line = '  rev_rxn_rates[{}] = '.format(rev_reacs.index(j))
            line += _conc_product(rxn.prod, rxn.prod_nu, trailing_space=True)
            if rxn.rev_par:
                line += arrhenius_text(*rxn.rev_par[0:3])
            else:
                line += 'kf / Kc'
            body.append(line + ';\n')
            body.append('\n')
    body.append('} // end eval_rxn_rates\n\n')
    return ''.join(body)


def _efficiency_terms(rxn):
    s = ''
    for isp, eff in rxn.thd_body_eff:
        if eff == 1.0:
            continue
        if eff > 1.0:
            s += ' + {}'.format(py2_str(eff - 1.0))
        else:
            s += ' - {}'.format(py2_str(1.0 - eff))
        s += ' * C[{}]'.format(isp)
    return s


def _write_pres_mod(specs, reacs):
    thd_flag = any(r.thd_body for r in reacs)
    pdep_flag = any(r.pdep and not r.thd_body for r in reacs)
    troe_flag = any(r.pdep and r.troe for r in reacs if not r.thd_body)
    sri_flag = any(r.pdep and r.sri for r in reacs if not r.thd_body)
    out = ['#include <math.h>\n#include "header.h"\n\n',
           'void get_rxn_pres_mod (const double T, const double pres, '
           'const double * C, double * pres_mod) {\n']
    if thd_flag or pdep_flag:
        out.append('  // third body variable declaration\n  double thd;\n\n')
    if pdep_flag:
        out.append('  // pressure dependence variable declarations\n'
                   '  double k0;\n  double kinf;\n  double Pr;\n\n')
        if troe_flag:
            out.append('  // troe variable declarations\n'
                       '  double logFcent;\n  double A;\n  double B;\n\n')
        if sri_flag:
            out.append('  // sri variable declarations\n  double X;\n\n')
    out.append('  double logT = log(T);\n'
               '  double m = pres / ({} * T);\n'.format(fmt_e8(chem.RU)))
    out.append('\n')
    p = 0
    for j, rxn in enumerate(reacs):
        if not (rxn.pdep or rxn.thd_body):
            continue
        out.append('  // reaction {};\n'.format(j))
        if rxn.thd_body:
            out.append('  pres_mod[{}] = m{};\n'.format(
                p, _efficiency_terms(rxn)))
        if rxn.pdep:
            if rxn.pdep_sp == '':
                out.append('  thd = m{};\n'.format(_efficiency_terms(rxn)))
            low = rxn.low if rxn.low else [rxn.A, rxn.b, rxn.E]
            high = rxn.high if rxn.high else [rxn.A, rxn.b, rxn.E]
            out.append('  k0 = ' + arrhenius_text(*low[0:3]) + ';\n')
            out.append('  kinf = ' + arrhenius_text(*high[0:3]) + ';\n')
            if rxn.pdep_sp != '':
                out.append('  Pr = k0 * C[{}] / kinf;\n'.format(rxn.pdep_sp))
            else:
                out.append('  Pr = k0 * thd / kinf;\n')
            simple = False
            if rxn.troe:
                tp = rxn.troe_par
                s = '  logFcent = log10( fmax({} * '.format(fmt_e8(1.0 - tp[0]))
                s += ('exp(-T / {})'.format(fmt_e8(tp[1])) if tp[1] > 0.0
                      else 'exp(T / {})'.format(fmt_e8(abs(tp[1]))))
                s += ' + {} * '.format(fmt_e8(tp[0]))
                s += ('exp(-T / {})'.format(fmt_e8(tp[2])) if tp[2] > 0.0
                      else 'exp(T / {})'.format(fmt_e8(abs(tp[2]))))
                if len(tp) == 4 and tp[3] != 0.0:
                    s += ' + '
                    s += ('exp(-{} / T)'.format(fmt_e8(tp[3])) if tp[3] > 0.0
                          else 'exp({} / T)'.format(fmt_e8(abs(tp[3]))))
                s += ', 1.0e-300));\n'
                out.append(s)
                out.append('  A = log10(fmax(Pr, 1.0e-300)) - 0.67 * '
                           'logFcent - 0.4;\n')
                out.append('  B = 0.806 - 1.1762 * logFcent - 0.14 * '
                           'log10(fmax(Pr, 1.0e-300));\n')
                line = ('  pres_mod[{}] = pow(10.0, logFcent / '
                        '(1.0 + A * A / (B * B))) '.format(p))
            elif rxn.sri:
                sp_ = rxn.sri_par
                out.append('  X = 1.0 / (1.0 + log10(fmax(Pr, 1.0e-300)) * '
                           'log10(fmax(Pr, 1.0e-300)));\n')
                line = '  pres_mod[{}] = pow({} * '.format(p, fmt_g6(sp_[0]))
                line += ('exp(-{} / T)'.format(fmt_g6(sp_[1])) if sp_[1] > 0.0
                         else 'exp({} / T)'.format(fmt_g6(abs(sp_[1]))))
                line += (' + exp(-T / {}), X) '.format(fmt_g6(sp_[2]))
                         if sp_[2] > 0.0 else
                         ' + exp(T / {}), X) '.format(fmt_g6(abs(sp_[2]))))
                if len(sp_) == 5 and sp_[3] != 1.0 and sp_[4] != 0.0:
                    line += '* {} * pow(T, {}) '.format(fmt_e8(sp_[3]),
                                                        fmt_g6(sp_[4]))
            else:
                simple = True
                line = '  pres_mod[{}] = '.format(p)
                if rxn.low:
                    line += ' Pr / (1.0 + Pr)'
                elif rxn.high:
                    line += '1.0 / (1.0 + Pr)'
            if not simple:
                if rxn.low:
                    line += '* Pr / (1.0 + Pr)'
                elif rxn.high:
                    line += '/ (1.0 + Pr)'
            out.append(line + ';\n')
        out.append('\n')
        p += 1
    out.append('} // end get_rxn_pres_mod\n\n')
    return ''.join(out)


def _write_spec_rates(specs, reacs, spec_order):
    rev_reacs = [i for i, r in enumerate(reacs) if r.rev]
    pdep_reacs = [i for i, r in enumerate(reacs) if r.thd_body or r.pdep]
    nsp = len(specs)
    out = ['#include "header.h"\n\n',
           'void eval_spec_rates (const double * fwd_rates, '
           'const double * rev_rates, const double * pres_mod, '
           'double * sp_rates, double * dy_N) {\n']

    def target(k):
        return '  (*dy_N)' if k + 1 == nsp else '  sp_rates[{}]'.format(k)

    seen = [False] * nsp
    for j, rxn in enumerate(reacs):
        out.append('  //rxn {}\n'.format(j))
        # each species of a reaction writes a different accumulator, so the
        # order within a reaction does not change any value
        for k in sorted(set(rxn.reac + rxn.prod)):
            nu = get_nu(k, rxn)
            if nu == 0.0:
                continue
            out.append('  //sp {}\n'.format(spec_order[k]))
            sign = '-' if nu < 0 else '+'
            line = target(k) + ' {}= '.format(sign if seen[k] else '')
            if not seen[k] and nu < 0:
                line += sign
            anu = abs(nu)
            if anu != 1.0:
                if is_integer(anu):
                    line += '{} * '.format(py2_str(float(anu)))
                else:
                    line += '{} * '.format(fmt_w3(anu))
            if rxn.rev:
                term = '(fwd_rates[{}] - rev_rates[{}])'.format(
                    j, rev_reacs.index(j))
            else:
                term = 'fwd_rates[{}]'.format(j)
            if rxn.thd_body or rxn.pdep:
                term += ' * pres_mod[{}]'.format(pdep_reacs.index(j))
            seen[k] = True
            out.append(line + term + ';\n')
        out.append('\n')
    for k in range(nsp):
        if not seen[k]:
            out.append('  //sp {}\n'.format(spec_order[k]))
            out.append(target(k) + ' = 0.0;\n')
    out.append('} // end eval_spec_rates\n\n')
    return ''.join(out)


def _write_eval_conc(specs):
    nsp = len(specs)
    out = ['#include "header.h"\n\n',
           'void eval_conc (const double T, const double pres, '
           'const double * y, double * y_N, double * mw_avg, '
           'double * rho, double * conc) {\n\n',
           '  // mass fraction of final species\n']
    line = '  *y_N = 1.0 - ('
    for k in range(nsp - 1):
        if len(line) > 70:
            out.append(line + '\n')
            line = '               '
        if k:
            line += ' + '
        line += 'y[{}]'.format(k)
    out.append(line + ');\n')
    line = '  *mw_avg = '
    for k in range(nsp - 1):
        if len(line) > 70:
            out.append(line + '\n')
            line = '     '
        if k:
            line += ' + '
        line += '(y[{}] * {})'.format(k, fmt_e16(1.0 / specs[k].mw))
    line += ' + ((*y_N) * {});\n'.format(fmt_e16(1.0 / specs[-1].mw))
    out.append(line)
    out.append('  *mw_avg = 1.0 / *mw_avg;\n')
    out.append('  // mass-averaged density\n')
    out.append('  *rho = pres * (*mw_avg) / ({} * T);\n'.format(fmt_e8(chem.RU)))
    for k in range(nsp - 1):
        out.append('  conc[{0}] = (*rho) * y[{0}] * {1};\n'.format(
            k, fmt_e16(1.0 / specs[k].mw)))
    out.append('  conc[{}] = (*rho) * (*y_N) * {};\n\n'.format(
        nsp - 1, fmt_e16(1.0 / specs[-1].mw)))
    out.append('} // end eval_conc\n\n')
    return ''.join(out)


def _nasa_if(sp):
    return '  if (T <= {})'.format(py2_str(sp.Trange[1]))


def _write_thermo(specs):
    """eval_h and eval_cp, rate_subs.py:1495-1565 and :1707-1776."""
    out = ['#include "header.h"\n\n',
           'void eval_h (const double T, double * h) {\n\n']
    for k, sp in enumerate(specs):
        out.append(_nasa_if(sp) + ' {\n')
        for half, a in (('lo', sp.lo), ('hi', sp.hi)):
            out.append('    h[{}] = {} * ({} + T * ({} + T * ({} + T * ({} + '
                       'T * ({} + {} * T)))));\n'.format(
                           k, fmt_e16(chem.RU / sp.mw), fmt_e16(a[5]),
                           fmt_e16(a[0]), fmt_e16(a[1] / 2.0),
                           fmt_e16(a[2] / 3.0), fmt_e16(a[3] / 4.0),
                           fmt_e16(a[4] / 5.0)))
            if half == 'lo':
                out.append('  } else {\n')
        out.append('  }\n\n')
    out.append('} // end eval_h\n\n')
    out.append('void eval_cp (const double T, double * cp) {\n\n')
    for k, sp in enumerate(specs):
        out.append(_nasa_if(sp) + ' {\n')
        for half, a in (('lo', sp.lo), ('hi', sp.hi)):
            out.append('    cp[{}] = {} * ({} + T * ({} + T * ({} + T * ({} + '
                       '{} * T))));\n'.format(
                           k, fmt_e16(chem.RU / sp.mw), fmt_e16(a[0]),
                           fmt_e16(a[1]), fmt_e16(a[2]), fmt_e16(a[3]),
                           fmt_e16(a[4])))
            if half == 'lo':
                out.append('  } else {\n')
        out.append('  }\n\n')
    out.append('} // end eval_cp\n\n')
    return ''.join(out)


def _write_dydt(specs, reacs):
    """Constant-pressure dydt, rate_subs.py:1829-1946 (C branch).

    y[0] = T, y[1..NSP-1] = mass fractions of species 0..NSP-2; dy alike;
    the last species is implied and its slot is not written.
    """
    nsp = len(specs)
    nr = len(reacs)
    nrev = sum(1 for r in reacs if r.rev)
    npd = sum(1 for r in reacs if r.thd_body or r.pdep)
    out = ['#include "header.h"\n#include "chem_utils.h"\n#include "rates.h"\n\n',
           'void dydt (const double t, const double pres, '
           'const double * y, double * dy) {\n\n',
           '  // species molar concentrations\n',
           '  double conc[{}];\n'.format(nsp),
           '  double y_N;\n  double mw_avg;\n  double rho;\n',
           '  eval_conc (y[0], pres, &y[1], &y_N, &mw_avg, &rho, conc);\n\n',
           '  // local arrays holding reaction rates\n',
           '  double fwd_rates[{}];\n'.format(nr)]
    if nrev:
        out.append('  double rev_rates[{}];\n'.format(nrev))
    else:
        out.append('  double* rev_rates = 0;\n')
    out.append('  eval_rxn_rates (y[0], pres, conc, fwd_rates, rev_rates);\n\n')
    if npd:
        out.append('  // get pressure modifications to reaction rates\n'
                   '  double pres_mod[{}];\n'.format(npd) +
                   '  get_rxn_pres_mod (y[0], pres, conc, pres_mod);\n')
    else:
        out.append('  double* pres_mod = 0;\n')
    out.append('\n')
    out.append('  // evaluate species molar net production rates\n'
               '  double dy_N;\n'
               '  eval_spec_rates (fwd_rates, rev_rates, pres_mod, &dy[1], '
               '&dy_N);\n\n')
    out.append('  // local array holding constant pressure specific heat\n'
               '  double cp[{}];\n'.format(nsp) +
               '  eval_cp (y[0], cp);\n\n')
    out.append('  // constant pressure mass-average specific heat\n')
    terms = ['(cp[{}] * y[{}])'.format(k, k + 1) for k in range(nsp - 1)]
    terms.append('(cp[{}] * y_N)'.format(nsp - 1))
    out.append('  double cp_avg = ' + ' + '.join(terms) + ';\n\n')
    out.append('  // local array for species enthalpies\n'
               '  double h[{}];\n'.format(nsp) + '  eval_h(y[0], h);\n')
    out.append('  // rate of change of temperature\n')
    terms = ['(dy[{}] * h[{}] * {})'.format(k + 1, k, fmt_e16(specs[k].mw))
             for k in range(nsp - 1)]
    terms.append('(dy_N * h[{}] * {})'.format(nsp - 1, fmt_e16(specs[-1].mw)))
    out.append('  dy[0] = (-1.0 / (rho * cp_avg)) * (' + ' + '.join(terms) +
               ');\n\n')
    out.append('  // calculate rate of change of species mass fractions\n')
    for k in range(nsp - 1):
        out.append('  dy[{}] *= ({} / rho);\n'.format(k + 1,
                                                      fmt_e16(specs[k].mw)))
    out.append('\n} // end dydt\n\n')
    return ''.join(out)


def _write_headers(specs, reacs):
    nsp = len(specs)
    pdep = any(r.thd_body or r.pdep for r in reacs)
    header = ('#include <stdlib.h>\n#include <stdio.h>\n#include <math.h>\n'
              '#include <string.h>\n#include <stdbool.h>\n\n'
              '#define DOUBLE\n#define Real double\n\n'
              '/** Number of species.\n')
    for k, sp in enumerate(specs):
        header += ' * {} {}\n'.format(k, sp.name)
    header += ' */\n#define NSP {}\n#define NN {}\n\n'.format(nsp, nsp + 1)
    rates = ('#ifndef RATES_HEAD\n#define RATES_HEAD\n\n#include "header.h"\n\n'
             'void eval_rxn_rates (const double, const double, const double*, '
             'double*, double*);\n'
             'void eval_spec_rates (const double*, const double*, '
             'const double*, double*, double*);\n')
    if pdep:
        rates += ('void get_rxn_pres_mod (const double, const double, '
                  'const double*, double*);\n')
    rates += '\n#endif\n'
    chem_h = ('#ifndef CHEM_UTILS_HEAD\n#define CHEM_UTILS_HEAD\n\n'
              '#include "header.h"\n\n'
              'void eval_conc (const double, const double, const double*, '
              'double*, double*, double*, double*);\n'
              'void eval_h (const double, double*);\n'
              'void eval_cp (const double, double*);\n'
              'void dydt (const double, const double, const double*, '
              'double*);\n\n#endif\n')
    return header, rates, chem_h


def mech_dims(specs, reacs):
    """(NSP, NR, NREV, NPD) of a mechanism."""
    return (len(specs), len(reacs), sum(1 for r in reacs if r.rev),
            sum(1 for r in reacs if r.thd_body or r.pdep))


def emit_c(specs, reacs, outdir, spec_order=None):
    """Write the oracle's C sources for the mechanism into ``outdir``.

    ``specs``/``reacs`` are in final order with integer species indices
    (after the bath-gas reordering).  ``spec_order[k]`` is the original
    index of species k, used only in comments.  Returns the list of C files.
    """
    if spec_order is None:
        spec_order = list(range(len(specs)))
    os.makedirs(outdir, exist_ok=True)
    header, rates_h, chem_h = _write_headers(specs, reacs)
    nsp, nr, nrev, npd = mech_dims(specs, reacs)
    dims = ('#ifndef MECH_DIMS_HEAD\n#define MECH_DIMS_HEAD\n'
            '#define MECH_NR {}\n#define MECH_NREV {}\n#define MECH_NPD {}\n'
            '#endif\n'.format(nr, nrev, npd))
    files = {
        'header.h': header, 'rates.h': rates_h, 'chem_utils.h': chem_h,
        'mech_dims.h': dims,
        'chem_utils.c': _write_eval_conc(specs),
        'rxn_rates.c': _write_rxn_rates(specs, reacs),
        'spec_rates.c': _write_spec_rates(specs, reacs, spec_order),
        'thermo.c': _write_thermo(specs),
        'dydt.c': _write_dydt(specs, reacs),
    }
    if npd:
        files['rxn_rates_pres_mod.c'] = _write_pres_mod(specs, reacs)
    for name, text in files.items():
        with open(os.path.join(outdir, name), 'w') as fh:
            fh.write(text)
    return [os.path.join(outdir, n) for n in files if n.endswith('.c')]
