"""crs::exp64 (the branch-free FP64 exponential of the fused kernels).

The coefficients in crs_common.h must be the ones exp64_coeffs.py derives
(Chebyshev interpolants of Q(r) = (e^r - 1)/r for exp64, of
Q2(s) = (2^s - 1)/s for exp2_64 and the fused kernels' register-held ExpK,
of q2(s) = (2^s - 1 - s ln2)/s^2 for the mixed kernel's exp2_m64k, of
q(r) = (e^r - 1 - r)/r^2 for the Estrin form), and the device sequences,
emulated with one rounding per FMA, must stay within 7 units of 2^-53
relative of e^x and 2^t (mpmath, 60 digits) over |x| <= 700, |t| <= 1000
(1.5 units for the q form, 107 for exp64k's conversion at |x| <= 100,
1.2e-8 for the mixed form).  A planted mistake (a dropped or perturbed
coefficient) breaks the bound.
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
    assert consts[3:13] == mod.coefficients_e()
    assert consts[13:] == [1.0]


def test_register_constants_match_the_literal_form():
    """ExpK (the fused kernels' register-held constants) holds exp2_64's
    base-2 coefficients in the same order, and log2 e for exp64k."""
    mod = _coeffs_module()
    text = open(os.path.join(CSRC, 'crs_common.h')).read()
    body = text[text.index('ExpK exp_consts()'):]
    body = body[:body.index('return k;')]
    consts = [float(v) for v in re.findall(r'const\((-?\d+\.\d+(?:e-?\d+)?)\)',
                                           body)]
    assert consts[:10] == mod.coefficients_b2()
    assert consts[10:] == [1.4426950408889634, 1.4142135623730951]


def test_emulated_sequence_error_bound():
    mod = _coeffs_module()
    assert mod.max_error_e(samples=3000, seed=7) <= 7.0


def test_exp2_64_base2_polynomial_and_error_bound():
    """exp2_64 holds the base-2 interpolant's coefficients and, emulated
    with one rounding per operation, stays within 7 ulp of 2^t."""
    mod = _coeffs_module()
    consts = _device_constants('double exp2_64(double t)')
    assert consts[:10] == mod.coefficients_b2()
    assert consts[10:] == [1.0]
    assert mod.max_error_b2(samples=3000, seed=8) <= 7.0


def test_exp64k_conversion_error_bound():
    """exp64k = exp2_64k(x log2 e): the rounded product adds |x| 2^-53, so
    at |x| <= 100 the bound is 7 + 100 ulp."""
    mod = _coeffs_module()
    assert mod.max_error_e_b2(samples=3000, seed=10, xmax=100.0) <= 107.0


def test_planted_base2_coefficient_mistakes_break_the_bound():
    """A dropped top coefficient or a 1e-12 change of the linear one in the
    base-2 polynomial exceeds 7 ulp."""
    import mpmath as mp
    mod = _coeffs_module()
    good = mod.coefficients_b2()
    for cs in ([0.0] + good[1:], good[:-1] + [good[-1] * (1 + 1e-12)]):
        worst = 0.0
        for t in (-0.49, -0.3, 0.21, 0.4999, 7.4, -18.6):
            ex = mp.power(2, mp.mpf(t))
            worst = max(worst, float(abs(mod.emulate_b2(t, cs) - ex) / ex))
        assert worst / 2.0 ** -53 > 7.0


def test_planted_coefficient_mistakes_break_the_bound():
    """The bound is tight enough to see a wrong coefficient: dropping the
    top one, or perturbing the linear one by 1e-12, exceeds 7 ulp."""
    import mpmath as mp
    mod = _coeffs_module()
    hi, lo = mod.ln2_split()
    good = mod.coefficients_e()
    for cs in ([0.0] + good[1:], good[:-2] + [good[-2] * (1 + 1e-12),
                                              good[-1]]):
        worst = 0.0
        for x in (-0.34, -0.2, 0.17, 0.3465, 5.3, -12.9):
            ex = mp.exp(mp.mpf(x))
            worst = max(worst, float(abs(mod.emulate_e(x, cs, hi, lo) - ex) /
                                     ex))
        assert worst / 2.0 ** -53 > 7.0


def test_estrin_form_shares_coefficients_and_bound():
    """exp_q_estrin holds the q form's ten coefficients (paired low/high
    power) and the Estrin evaluation keeps the 1.5 ulp bound."""
    mod = _coeffs_module()
    consts = _device_constants('double exp_q_estrin(double r)')
    assert sorted(consts) == sorted(mod.coefficients())
    assert mod.max_error2_estrin(samples=3000, seed=9) <= 1.5


def test_mixed_exponential_coefficients_and_bound():
    """crs::exp2_m64k (the mixed kernel's exponential) holds the base-2
    degree-4 interpolant's coefficients and stays within 1.2e-8 relative;
    a dropped top coefficient breaks that bound."""
    import mpmath as mp
    mod = _coeffs_module()
    text = open(os.path.join(CSRC, 'crs_common.h')).read()
    body = text[text.index('ExpMK exp_m_consts()'):]
    body = body[:body.index('return k;')]
    consts = [float(v) for v in re.findall(r'const\((-?\d+\.\d+(?:e-?\d+)?)\)',
                                           body)]
    hi, _lo = mod.ln2_split()
    assert consts[:5] == mod.coefficients_b2m()
    assert consts[5:] == [hi, 1.4426950408889634, 1.4142135623730951]
    assert mod.max_error_b2m(samples=3000, seed=11) <= 1.2e-8
    bad = [0.0] + consts[1:5]
    worst = max(float(abs(mod.emulate_b2m(t, bad, hi) - mp.power(2, t)) /
                      mp.power(2, t)) for t in (-0.49, 0.45, 3.47))
    assert worst > 1.2e-8
