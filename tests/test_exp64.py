"""crs::exp64 (the branch-free FP64 exponential of the fused kernels).

The coefficients in crs_common.h must be the ones exp64_coeffs.py derives
(Chebyshev interpolant of (e^r - 1 - r)/r^2), and the device sequence,
emulated with one rounding per FMA, must stay within 1.5 units of 2^-53
relative of e^x (mpmath, 60 digits) over |x| <= 700.
"""

import importlib.util
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, 'create_rate_subs_amd', 'csrc')


def _coeffs_module():
    spec = importlib.util.spec_from_file_location(
        'exp64_coeffs', os.path.join(CSRC, 'exp64_coeffs.py'))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _device_constants(fn='double exp64(double x)'):
    text = open(os.path.join(CSRC, 'crs_common.h')).read()
    body = text[text.index(fn):]
    body = body[:body.index('\n}\n')]
    return [float(v) for v in re.findall(r'(-?\d+\.\d+(?:e-?\d+)?)', body)]


def test_header_holds_derived_coefficients():
    mod = _coeffs_module()
    consts = _device_constants()
    hi, lo = mod.ln2_split()
    assert consts[0] == 1.4426950408889634
    assert consts[1:3] == [hi, lo]
    assert consts[3:13] == mod.coefficients()


def test_emulated_sequence_error_bound():
    mod = _coeffs_module()
    assert mod.max_error(samples=3000, seed=7) <= 1.5


def test_exp2_64_shares_the_polynomial_and_error_bound():
    mod = _coeffs_module()
    consts = _device_constants('double exp2_64(double t)')
    hi, _lo = mod.ln2_split()
    assert consts[0] == hi
    assert consts[1:11] == mod.coefficients()
    assert mod.max_error2(samples=3000, seed=8) <= 1.5


def test_estrin_form_shares_coefficients_and_bound():
    """exp_q_estrin holds the same ten coefficients (paired low/high power)
    and the Estrin evaluation keeps the 1.5 ulp bound."""
    mod = _coeffs_module()
    consts = _device_constants('double exp_q_estrin(double r)')
    assert sorted(consts) == sorted(mod.coefficients())
    assert mod.max_error2_estrin(samples=3000, seed=9) <= 1.5


def test_mixed_exponential_coefficients_and_bound():
    """crs::exp2_m64k (the mixed kernel's exponential) holds the degree-4
    interpolant's coefficients and stays within 1.2e-8 relative."""
    mod = _coeffs_module()
    text = open(os.path.join(CSRC, 'crs_common.h')).read()
    body = text[text.index('ExpMK exp_m_consts()'):]
    body = body[:body.index('return k;')]
    consts = [float(v) for v in re.findall(r'const\((-?\d+\.\d+(?:e-?\d+)?)\)',
                                           body)]
    hi, _lo = mod.ln2_split()
    assert consts[:5] == mod.coefficients(mod.MIXED_DEGREE)
    assert consts[5:] == [hi, 1.4426950408889634]
    assert mod.max_error2_mixed(samples=3000, seed=11) <= 1.2e-8
