"""crs::log64 (the branch-free FP64 logarithm of the fused kernels).

The coefficients in crs_common.h must be the ones log64_coeffs.py derives
(Chebyshev interpolant of (2 atanh(sqrt z) - 2 sqrt z) / z^(3/2)), and the
device sequence, emulated with one rounding per operation and a reciprocal
seed 2^-23 off in either direction, must stay within 2 units of 2^-53
relative of ln x (mpmath, 60 digits) from 1e-300 to 1e300 and next to 1.
A dropped series term or a wrong ln2 split fails the bound.
"""

import importlib.util
import os
import re

import mpmath as mp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, 'create_rate_subs_amd', 'csrc')


def _coeffs_module():
    spec = importlib.util.spec_from_file_location(
        'log64_coeffs', os.path.join(CSRC, 'log64_coeffs.py'))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _device_constants():
    text = open(os.path.join(CSRC, 'crs_common.h')).read()
    body = text[text.index('double log64_finite(double x)'):]
    body = body[:body.index('\n}\n')]
    return [float(v) for v in re.findall(r'(-?\d+\.\d+(?:e-?\d+)?)', body)]


def test_header_holds_derived_coefficients():
    mod = _coeffs_module()
    consts = _device_constants()
    hi, lo = mod.ln2_split()
    # sqrt(1/2) threshold, the 1.0 / 2.0 of f and 2 + f, P, the ln2 split
    assert consts[0] == mod.SQRT_HALF
    cs = mod.coefficients()
    k = consts.index(cs[0])
    assert consts[k:k + len(cs)] == cs
    assert consts[-2:] == [hi, lo]


def test_emulated_sequence_error_bound():
    mod = _coeffs_module()
    assert mod.max_error(samples=1200, seed=11) <= 2.0


def test_bound_catches_a_dropped_term():
    mod = _coeffs_module()
    cs = mod.coefficients()
    hi, lo = mod.ln2_split()
    x = 1.37
    good = mod.emulate(x, cs, hi, lo)
    bad = mod.emulate(x, cs[:-1] + [0.6], hi, lo)
    ex = mp.log(mp.mpf(x))
    assert abs(good - ex) / ex < 2 * 2.0 ** -53
    assert abs(bad - ex) / ex > 1e-6


def _exp_module():
    spec = importlib.util.spec_from_file_location(
        'exp64_coeffs', os.path.join(CSRC, 'exp64_coeffs.py'))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_pow64_composition_bound():
    """crs::pow64 = exp64(y * log64_finite(x)) for the non-integer orders
    of a reaction: the emulated composition stays within 1e-13 relative of
    x^y (mpmath) over concentrations 1e-300 .. 1e3 and orders 0.25 .. 2.5."""
    lm = _coeffs_module()
    em = _exp_module()
    lcs, (lhi, llo) = lm.coefficients(), lm.ln2_split()
    ecs, (ehi, elo) = em.coefficients_e(), em.ln2_split()
    import random
    rng = random.Random(5)
    worst = 0.0
    for _ in range(300):
        y = rng.choice([0.25, 0.5, 1.5, 2.5])
        # results stay normal doubles (x^y >= 1e-300)
        x = 10.0 ** rng.uniform(max(-300.0, -300.0 / y), 3)
        ln = float(lm.emulate(x, lcs, lhi, llo))
        got = em.emulate_e(float(mp.mpf(y) * mp.mpf(ln)), ecs, ehi, elo)
        ex = mp.power(mp.mpf(x), mp.mpf(y))
        worst = max(worst, float(abs(got - ex) / ex))
    assert worst < 1e-13


def _fn_body(signature):
    text = open(os.path.join(CSRC, 'crs_common.h')).read()
    body = text[text.index(signature):]
    return body[:body.index('\n}\n')]


def test_log2_64k_uses_the_exp2_polynomial_and_bound():
    """crs::log2_64k (the falloff blocks' logarithm) refines the hardware
    FP32 log2 seed with the register-held exp2 polynomial; emulated with
    the seed exact or 2^-20 off, it stays within 12 units of 2^-53 (absolute
    below |log2 x| = 1, relative above)."""
    mod = _coeffs_module()
    body = _fn_body('double log2_64k(const ExpK &k, double x)')
    assert ['k.b{}'.format(i) in body for i in range(10)] == [True] * 10
    assert 'k.sqrt2' in body and 'k.log2e' in body
    assert '__builtin_amdgcn_logf' in body
    text = open(os.path.join(CSRC, 'crs_common.h')).read()
    consts = text[text.index('ExpK exp_consts()'):]
    assert 'k.sqrt2 = sgpr_const(1.4142135623730951);' in consts
    assert mod.SQRT2 == float(mp.sqrt(2))
    assert mod.max_error_log2(samples=600, seed=13) <= 12 * 2.0 ** -53


def test_log2_m64k_bound():
    """The mixed kernel's log2 (degree-4 exp2 form) stays within 2e-8."""
    mod = _coeffs_module()
    body = _fn_body('double log2_m64k(const ExpMK &k, double x)')
    assert 'k.c4' in body and 'k.sqrt2' in body
    assert mod.max_error_log2(samples=600, seed=14, mixed=True) <= 2e-8


def test_log2_planted_mistakes_break_the_bound():
    """Without the t^2/2 term, or with sqrt 2 off in its 9th digit, the
    correction no longer absorbs a seed 2^-20 off."""
    mod = _coeffs_module()
    assert mod.max_error_log2(samples=60, seed=15,
                              second_order=False) > 12 * 2.0 ** -53
    good = mod.SQRT2
    try:
        mod.SQRT2 = 1.41421356
        assert mod.max_error_log2(samples=60, seed=16) > 12 * 2.0 ** -53
    finally:
        mod.SQRT2 = good
