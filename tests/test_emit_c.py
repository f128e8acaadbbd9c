"""Text of the oracle's C emitter against hand-composed expected lines.

Each expected string is composed by hand from the reference's format
strings (rate_subs.py:20-135 Arrhenius, :593-662 Kc, :901-1057 pressure
modification, :1162-1226 species rates, :1319-1386 eval_conc) with Python
2.7 float spellings (12 significant digits for str()).
"""

import math
import os
import sys

import pytest

from create_rate_subs_amd import chem
from create_rate_subs_amd.utils import fmt_g6, py2_str, fmt_w3
from oracle import emit_c


def test_py2_float_spelling():
    assert py2_str(0.30000000000000004) == '0.3'
    assert py2_str(2.0) == '2.0'
    assert py2_str(1e16) == '1e+16'
    assert py2_str(-1.991) == '-1.991'
    assert py2_str(123456789012345.0) == '1.23456789012e+14'
    assert py2_str(1.0 / 3.0) == '0.333333333333'
    assert py2_str(3) == '3'
    assert fmt_g6(1000.0) == '1000.0'
    assert fmt_g6(1234567.0) == '1.23457e+06'
    assert fmt_w3(0.5) == '0.5' and fmt_w3(1.25) == '1.25'


def test_arrhenius_forms():
    A, b, E = 2.5e10, -1.5, 1234.5
    la = '{:.16e}'.format(math.log(A))
    assert (emit_c.arrhenius_text(A, b, E) ==
            'exp(' + la + ' - 1.5 * logT - (1.2345000000000000e+03 / T))')
    assert (emit_c.arrhenius_text(A, 2.0, E) ==
            'exp(' + la + ' + 2.0 * logT - (1.2345000000000000e+03 / T))')
    assert (emit_c.arrhenius_text(A, 0.0, E) ==
            'exp(' + la + ' - (1.2345000000000000e+03 / T))')
    assert emit_c.arrhenius_text(A, 0.7, 0.0) == 'exp(' + la + ' + 0.7 * logT)'
    assert emit_c.arrhenius_text(A, 0.0, 0.0) == '25000000000.0'
    assert emit_c.arrhenius_text(1.0 / 3.0, 0.0, 0.0) == '0.333333333333'
    # negative A (rate_subs.py:97-130)
    assert (emit_c.arrhenius_text(-2.0, 0.5, 10.0) ==
            '-2.0000000000000000e+00 * exp(0.5 * logT - '
            '(1.0000000000000000e+01 / T))')
    assert (emit_c.arrhenius_text(-2.0, 0.0, 10.0) ==
            '-2.0000000000000000e+00 * exp(-(1.0000000000000000e+01 / T))')
    assert emit_c.arrhenius_text(-2.0, 2.0, 0.0) == '-2.0 * T * T'
    with pytest.raises(NotImplementedError):
        emit_c.arrhenius_text(0.0, 1.0, 1.0)


class _Sp(object):
    def __init__(self, name, mw, lo, hi, tmid=1000.0):
        self.name, self.mw = name, mw
        self.lo, self.hi = lo, hi
        self.Trange = [300.0, tmid, 5000.0]


class _Rx(object):
    def __init__(self, reac, reac_nu, prod, prod_nu, A, b, E, rev=True):
        self.reac, self.reac_nu, self.prod, self.prod_nu = (reac, reac_nu,
                                                            prod, prod_nu)
        self.A, self.b, self.E, self.rev = A, b, E, rev
        self.rev_par = []
        self.thd_body = False
        self.thd_body_eff = []
        self.pdep = False
        self.pdep_sp = ''
        self.low, self.high = [], []
        self.troe = self.sri = self.cheb = self.plog = False
        self.troe_par, self.sri_par = [], []


def _tiny():
    lo = [1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0]
    hi = [0.5, 0.25, 0.125, 0.0625, 0.03125, 8.0, 9.0]
    specs = [_Sp('A', 2.0, lo, hi), _Sp('B', 4.0, hi, lo),
             _Sp('N2', 28.0, lo, lo)]
    r0 = _Rx([0], [2], [1], [1], 1e6, 0.0, 100.0)
    r1 = _Rx([1], [1], [0], [2], 5.0, 0.0, 0.0, rev=False)
    r1.thd_body = True
    r1.thd_body_eff = [(0, 0.7), (2, 1.5), (1, 1.0)]
    return specs, [r0, r1]


def test_eval_conc_text(tmp_path):
    specs, reacs = _tiny()
    emit_c.emit_c(specs, reacs, str(tmp_path))
    text = (tmp_path / 'chem_utils.c').read_text()
    assert '  *y_N = 1.0 - (y[0] + y[1]);\n' in text
    assert ('  *mw_avg = (y[0] * 5.0000000000000000e-01) + (y[1] * '
            '2.5000000000000000e-01) + ((*y_N) * 3.5714285714285712e-02);\n'
            in text)
    assert '  *mw_avg = 1.0 / *mw_avg;\n' in text
    assert '  *rho = pres * (*mw_avg) / (8.31446210e+03 * T);\n' in text
    assert '  conc[1] = (*rho) * y[1] * 2.5000000000000000e-01;\n' in text
    assert '  conc[2] = (*rho) * (*y_N) * 3.5714285714285712e-02;\n' in text


def test_rxn_rates_text(tmp_path):
    specs, reacs = _tiny()
    emit_c.emit_c(specs, reacs, str(tmp_path))
    text = (tmp_path / 'rxn_rates.c').read_text()
    la = '{:.16e}'.format(math.log(1e6))
    assert ('  kf = exp(' + la + ' - (1.0000000000000000e+02 / T));\n'
            in text)
    assert '  fwd_rxn_rates[0] = C[0] * C[0] * kf;\n' in text
    # Kc coefficients: products (+1) then reactants (-2), summed
    lo_a = [7.0 - 1.0, 1.0 - 1.0, 2.0 / 2, 3.0 / 6, 4.0 / 12, 5.0 / 20, 6.0]
    hi_a = [9.0 - 0.5, 0.5 - 1.0, 0.25 / 2, 0.125 / 6, 0.0625 / 12,
            0.03125 / 20, 8.0]
    # species B uses lo<->hi swapped sets
    c_lo = [lo_a[i] * -2.0 + hi_a[i] * 1.0 for i in range(7)]
    expect = ('    Kc = ({} + {} * logT + T * ({} + T * ({} + T * ({} + {} * T))) '
              '- {} / T);\n'.format(*['{:.16e}'.format(x) for x in c_lo]))
    assert '  if (T <= 1000.0) {\n' + expect in text
    assert ('  Kc = {:.16e} * exp(Kc);\n'.format((chem.PA / chem.RU) ** -1)
            in text)
    assert '  rev_rxn_rates[0] = C[1] * kf / Kc;\n' in text
    assert '  fwd_rxn_rates[1] = C[1] * 5.0;\n' in text


def test_pres_mod_and_spec_rates_text(tmp_path):
    specs, reacs = _tiny()
    emit_c.emit_c(specs, reacs, str(tmp_path))
    pm = (tmp_path / 'rxn_rates_pres_mod.c').read_text()
    # eff 0.7 -> '- 0.3' (Py2 str of 0.30000000000000004), 1.0 skipped
    assert '  pres_mod[0] = m - 0.3 * C[0] + 0.5 * C[2];\n' in pm
    assert '  double m = pres / (8.31446210e+03 * T);\n' in pm
    sp = (tmp_path / 'spec_rates.c').read_text()
    assert '  sp_rates[0] = -2.0 * (fwd_rates[0] - rev_rates[0]);\n' in sp
    assert '  sp_rates[1] = (fwd_rates[0] - rev_rates[0]);\n' in sp
    assert '  sp_rates[0] += 2.0 * fwd_rates[1] * pres_mod[0];\n' in sp
    assert '  sp_rates[1] -= fwd_rates[1] * pres_mod[0];\n' in sp
    assert '  (*dy_N) = 0.0;\n' in sp


def test_troe_sri_text():
    r = _Rx([0], [1], [1], [1], 1e3, 0.0, 0.0, rev=False)
    r.pdep = True
    r.low = [2e5, 0.0, 0.0]
    r.troe = True
    r.troe_par = [0.25, 100.0, -2000.0, 500.0]
    specs, _ = _tiny()
    text = emit_c._write_pres_mod(specs, [r])
    assert ('  logFcent = log10( fmax(7.50000000e-01 * exp(-T / 1.00000000e+02)'
            ' + 2.50000000e-01 * exp(T / 2.00000000e+03) + exp(-5.00000000e+02'
            ' / T), 1.0e-300));\n' in text)
    assert ('  pres_mod[0] = pow(10.0, logFcent / (1.0 + A * A / (B * B))) '
            '* Pr / (1.0 + Pr);\n' in text)
    r.troe = False
    r.sri = True
    r.sri_par = [1.1, 300.0, 900.0, 2.0, 0.5]
    text = emit_c._write_pres_mod(specs, [r])
    assert ('  pres_mod[0] = pow(1.1 * exp(-300.0 / T) + exp(-T / 900.0), X) '
            '* 2.00000000e+00 * pow(T, 0.5) * Pr / (1.0 + Pr);\n' in text)
    r.sri = False
    text = emit_c._write_pres_mod(specs, [r])
    assert '  pres_mod[0] =  Pr / (1.0 + Pr);\n' in text
