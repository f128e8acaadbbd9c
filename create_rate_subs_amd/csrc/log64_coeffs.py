"""Coefficients of crs::log64 (crs_common.h) and their accuracy check.

x = 2^e m with m in [sqrt(1/2), sqrt(2)), f = m - 1 (exact), s = f / (2 + f)
and ln m = 2 atanh(s) = 2 s + s^3 P(s^2), where
P(z) = (2 atanh(sqrt z) - 2 sqrt z) / z^(3/2) = 2/3 + 2z/5 + 2z^2/7 + ...
is replaced by its Chebyshev interpolant on z in [0, (3 - 2 sqrt 2)^2]
(mpmath.chebyfit, 60 digits).  s is formed with one reciprocal (hardware
seed and two Newton steps) and one correction step that also takes up the
rounding of 2 + f, so s carries about half an ulp; ln x = e ln2 + ln m with
ln2 split into a double and its remainder.

The check emulates the device sequence with every operation rounded once
to double; the reciprocal seed is perturbed by 2^-23 relative (the
hardware seed's accuracy is not assumed better than that).
Run: python log64_coeffs.py [samples]
"""

import random
import sys

import mpmath as mp

SQRT_HALF = 0.7071067811865476


def z_max():
    mp.mp.dps = 60
    return (3 - 2 * mp.sqrt(2)) ** 2


def coefficients(degree=7):
    """P's coefficients, highest power first, rounded to double."""
    mp.mp.dps = 60

    def P(z):
        if z < mp.mpf('1e-30'):
            return mp.mpf(2) / 3
        s = mp.sqrt(z)
        return (2 * mp.atanh(s) - 2 * s) / (s * z)

    poly, _err = mp.chebyfit(P, [0, z_max() * mp.mpf('1.0001')], degree + 1,
                             error=True)
    return [float(c) for c in poly]


def ln2_split():
    mp.mp.dps = 60
    hi = float(mp.log(2))
    return hi, float(mp.log(2) - mp.mpf(hi))


def emulate(x, cs, hi, lo, seed_err=2.0 ** -23):
    """The device sequence with each operation rounded to double."""
    def rd(v):
        return mp.mpf(float(v))

    def fma(a, b, c):
        return rd(mp.mpf(a) * mp.mpf(b) + mp.mpf(c))

    x = mp.mpf(x)
    m, e = mp.frexp(x)                 # m in [0.5, 1)
    if m < SQRT_HALF:
        m, e = m * 2, e - 1
    f = rd(m - 1)
    d = rd(2 + f)
    # rcp64: hardware seed, two Newton steps
    r = rd((1 / d) * (1 + seed_err))
    t = fma(-d, r, 1.0)
    r = fma(r, t, r)
    t = fma(-d, r, 1.0)
    r = fma(r, t, r)
    q = rd(f * r)
    dl = rd(f - rd(d - 2))             # rounding error of 2 + f, exact
    res = fma(-q, d, f)                # f - q d, exact
    res = fma(-q, dl, res)
    s = fma(res, r, q)
    z = rd(s * s)
    p = mp.mpf(cs[0])
    for c in cs[1:]:
        p = fma(p, z, c)
    lm = fma(rd(s * z), p, rd(2 * s))
    de = mp.mpf(e)
    return fma(de, hi, fma(de, lo, lm))


def max_error(samples=20000, seed=3):
    """Max relative error of the emulated sequence, in units of 2^-53."""
    cs = coefficients()
    hi, lo = ln2_split()
    rng = random.Random(seed)
    worst = 0.0
    for i in range(samples):
        k = i % 4
        if k == 0:
            x = 10.0 ** rng.uniform(-300.0, 300.0)
        elif k == 1:
            x = rng.uniform(0.5, 2.0)
        elif k == 2:
            x = 1.0 + rng.uniform(-1e-6, 1e-6)
        else:
            x = rng.uniform(1e-3, 1e3)
        if x == 1.0:
            continue
        ex = mp.log(mp.mpf(x))
        for se in (2.0 ** -23, -(2.0 ** -23), 0.0):
            got = emulate(x, cs, hi, lo, se)
            worst = max(worst, float(abs(got - ex) / abs(ex)))
    return worst / 2.0 ** -53


def _exp_coeffs():
    """The exp2 polynomials of exp64_coeffs.py beside this file."""
    import importlib.util
    import os
    spec = importlib.util.spec_from_file_location(
        'exp64_coeffs', os.path.join(os.path.dirname(os.path.abspath(
            __file__)), 'exp64_coeffs.py'))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


_POLY = {}


def _exp2_poly(mixed):
    """exp2_64's (mixed: exp2_m64k's) coefficients, derived once."""
    if mixed not in _POLY:
        ec = _exp_coeffs()
        _POLY[mixed] = (ec.coefficients_b2m() if mixed else
                        ec.coefficients_b2())
    return _POLY[mixed]


SQRT2 = 1.4142135623730951
LOG2E = 1.4426950408889634


def emulate_log2(x, seed_off=0.0, mixed=False, cs=None, second_order=True):
    """crs::log2_64k (mixed: crs::log2_m64k) with every double operation
    rounded once.  The hardware FP32 log2 of the mantissa is modelled as
    the FP32 rounding of log2(fl32(m)) moved by ``seed_off`` (absolute) and
    rounded to FP32 again: the sequence must not depend on that seed's last
    bits."""
    import numpy as np

    def rd(v):
        return mp.mpf(float(v))

    def fma(a, b, c):
        return rd(mp.mpf(a) * mp.mpf(b) + mp.mpf(c))

    if cs is None:
        cs = _exp2_poly(mixed)
    m, e = mp.frexp(mp.mpf(x))
    mf = float(np.float32(float(m)))
    y0 = float(np.float32(float(mp.log(mf, 2))))
    y0 = mp.mpf(float(np.float32(y0 + seed_off)))
    s = -y0 - mp.mpf(0.5)                      # exact in double
    q = mp.mpf(cs[0])
    for c in cs[1:]:
        q = fma(q, s, c)
    if mixed:
        p = fma(s, fma(s, q, float(mp.log(2))), 1.0)
    else:
        p = fma(s, q, 1.0)
    t = fma(rd(m * SQRT2), p, -1.0)
    u = fma(rd(-0.5 * t), t, t) if second_order else t
    return rd(fma(u, LOG2E, y0) + e)


def max_error_log2(samples=4000, seed=5, mixed=False, cs=None,
                   second_order=True):
    """Max error of the emulated log2_64k (log2_m64k) against log2 x over
    1e-300 .. 1e300 and next to powers of two, with the FP32 seed exact and
    2^-20 off in either direction: absolute where |log2 x| < 1, relative
    beyond (the final sum e + log2 m rounds to the ulp of its size)."""
    rng = random.Random(seed)
    worst = 0.0
    for i in range(samples):
        k = i % 3
        if k == 0:
            x = 10.0 ** rng.uniform(-300.0, 300.0)
        elif k == 1:
            x = rng.uniform(0.25, 4.0)
        else:
            x = 2.0 ** rng.randint(-60, 60) * (1.0 + rng.uniform(-1e-6, 1e-6))
        ex = mp.log(mp.mpf(x), 2)
        for so in (0.0, 2.0 ** -20, -(2.0 ** -20)):
            got = emulate_log2(x, so, mixed, cs, second_order)
            worst = max(worst, float(abs(got - ex) / max(abs(ex), 1)))
    return worst


if __name__ == '__main__':
    print('P coefficients:', [repr(c) for c in coefficients()])
    print('ln2 split:', [repr(v) for v in ln2_split()])
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 20000
    print('log64 max error (units of 2^-53):', max_error(n))
    print('log2_64k max absolute error:', max_error_log2(n // 5))
    print('log2_m64k max absolute error:', max_error_log2(n // 5, mixed=True))
