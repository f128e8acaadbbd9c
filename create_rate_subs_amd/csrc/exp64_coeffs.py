"""Coefficients of crs::exp64 / crs::exp2_64 (crs_common.h) and their
accuracy check.

e^r = 1 + r Q(r) on |r| <= ln2/2, with Q(r) = (e^r - 1) / r approximated by
its degree-9 Chebyshev interpolant (mpmath.chebyfit, 60 digits): ten
coefficients, nine Horner FMAs and one final FMA.  The earlier form,
e^r = 1 + r (1 + r q(r)) with q(r) = (e^r - 1 - r) / r^2 of degree 9
(``coefficients``), is more accurate (1.2 ulp against 6.3) at one FMA more
per exponential; it remains in crs::exp_q_estrin.  The checks emulate the
device sequences exactly: every FMA is rounded once to double, the
reduction uses ln2 split into a double and its remainder.
Run: python exp64_coeffs.py [samples]
"""

import random
import sys

import mpmath as mp

LOG2E = 1.4426950408889634


def coefficients(degree=9):
    """q's coefficients, highest power first, rounded to double."""
    mp.mp.dps = 60
    a = mp.log(2) / 2

    def q(r):
        if abs(r) < mp.mpf('1e-20'):
            return mp.mpf(1) / 2
        return (mp.exp(r) - 1 - r) / r ** 2

    poly, _err = mp.chebyfit(q, [-a * 1.0001, a * 1.0001], degree + 1,
                             error=True)
    return [float(c) for c in poly]


def coefficients_e(degree=9):
    """Q's coefficients (Q(r) = (e^r - 1) / r), highest power first,
    rounded to double: those of crs::exp64 and crs::exp2_64."""
    mp.mp.dps = 60
    a = mp.log(2) / 2

    def Q(r):
        if abs(r) < mp.mpf('1e-25'):
            return mp.mpf(1)
        return (mp.exp(r) - 1) / r

    poly, _err = mp.chebyfit(Q, [-a * 1.0001, a * 1.0001], degree + 1,
                             error=True)
    return [float(c) for c in poly]


def _tail_e(r, cs, fma):
    q = mp.mpf(cs[0])
    for c in cs[1:]:
        q = fma(q, r, c)
    return fma(r, q, 1.0)


def emulate_e(x, cs, hi, lo):
    """crs::exp64: Cody-Waite reduction, then 1 + r Q(r), each operation
    rounded to double."""
    def rd(v):
        return mp.mpf(float(v))

    def fma(a, b, c):
        return rd(mp.mpf(a) * mp.mpf(b) + mp.mpf(c))

    n = mp.nint(rd(mp.mpf(x) * LOG2E))
    r = fma(-n, hi, x)
    r = fma(-n, lo, r)
    return rd(mp.ldexp(_tail_e(r, cs, fma), int(n)))


def emulate2_e(t, cs, hi):
    """crs::exp2_64: n = rint(t), r = (t - n) ln2, then 1 + r Q(r)."""
    def rd(v):
        return mp.mpf(float(v))

    def fma(a, b, c):
        return rd(mp.mpf(a) * mp.mpf(b) + mp.mpf(c))

    n = mp.nint(t)
    r = rd((mp.mpf(t) - n) * hi)
    return rd(mp.ldexp(_tail_e(r, cs, fma), int(n)))


def max_error_e(samples=20000, seed=1):
    """Max relative error of crs::exp64 against e^x, units of 2^-53."""
    cs = coefficients_e()
    hi, lo = ln2_split()
    rng = random.Random(seed)
    worst = 0.0
    for i in range(samples):
        x = rng.uniform(-700.0, 700.0) if i % 2 else rng.uniform(-2.0, 2.0)
        ex = mp.exp(mp.mpf(x))
        worst = max(worst, float(abs(emulate_e(x, cs, hi, lo) - ex) / ex))
    return worst / 2.0 ** -53


def max_error2_e(samples=20000, seed=2):
    """Max relative error of crs::exp2_64 against 2^t, units of 2^-53."""
    cs = coefficients_e()
    hi, _lo = ln2_split()
    rng = random.Random(seed)
    worst = 0.0
    for i in range(samples):
        t = rng.uniform(-1000.0, 1000.0) if i % 2 else rng.uniform(-3.0, 3.0)
        ex = mp.power(2, mp.mpf(t))
        worst = max(worst, float(abs(emulate2_e(t, cs, hi) - ex) / ex))
    return worst / 2.0 ** -53


def ln2_split():
    mp.mp.dps = 60
    hi = float(mp.log(2))
    return hi, float(mp.log(2) - mp.mpf(hi))


def emulate(x, cs, hi, lo):
    """The device sequence with each operation rounded to double."""
    def rd(v):
        return mp.mpf(float(v))

    def fma(a, b, c):
        return rd(mp.mpf(a) * mp.mpf(b) + mp.mpf(c))

    n = mp.nint(rd(mp.mpf(x) * LOG2E))
    r = fma(-n, hi, x)
    r = fma(-n, lo, r)
    q = mp.mpf(cs[0])
    for c in cs[1:]:
        q = fma(q, r, c)
    p = fma(r, fma(r, q, 1.0), 1.0)
    return rd(mp.ldexp(p, int(n)))


def emulate2(t, cs, hi):
    """crs::exp2_64: n = rint(t), r = (t - n) ln2 (t - n exact)."""
    def rd(v):
        return mp.mpf(float(v))

    def fma(a, b, c):
        return rd(mp.mpf(a) * mp.mpf(b) + mp.mpf(c))

    n = mp.nint(t)
    r = rd((mp.mpf(t) - n) * hi)
    q = mp.mpf(cs[0])
    for c in cs[1:]:
        q = fma(q, r, c)
    p = fma(r, fma(r, q, 1.0), 1.0)
    return rd(mp.ldexp(p, int(n)))


def emulate2_estrin(t, cs, hi):
    """crs::exp2_64_estrin: as emulate2 with q(r) by Estrin's scheme."""
    def rd(v):
        return mp.mpf(float(v))

    def fma(a, b, c):
        return rd(mp.mpf(a) * mp.mpf(b) + mp.mpf(c))

    a = cs[::-1]                          # a0 .. a9
    n = mp.nint(t)
    r = rd((mp.mpf(t) - n) * hi)
    r2 = rd(r * r)
    r4 = rd(r2 * r2)
    r8 = rd(r4 * r4)
    p01, p23 = fma(a[1], r, a[0]), fma(a[3], r, a[2])
    p45, p67 = fma(a[5], r, a[4]), fma(a[7], r, a[6])
    p89 = fma(a[9], r, a[8])
    q = fma(p89, r8, fma(fma(p67, r2, p45), r4, fma(p23, r2, p01)))
    p = fma(r, fma(r, q, 1.0), 1.0)
    return rd(mp.ldexp(p, int(n)))


def max_error2_estrin(samples=20000, seed=4):
    """Max relative error of crs::exp2_64_estrin, units of 2^-53."""
    cs = coefficients()
    hi, _lo = ln2_split()
    rng = random.Random(seed)
    worst = 0.0
    for i in range(samples):
        t = rng.uniform(-1000.0, 1000.0) if i % 2 else rng.uniform(-3.0, 3.0)
        ex = mp.power(2, mp.mpf(t))
        worst = max(worst, float(abs(emulate2_estrin(t, cs, hi) - ex) / ex))
    return worst / 2.0 ** -53


def max_error2(samples=20000, seed=2):
    """Max relative error of the q-form 2^t (1 + r (1 + r q(r))) against
    2^t, units of 2^-53."""
    cs = coefficients()
    hi, _lo = ln2_split()
    rng = random.Random(seed)
    worst = 0.0
    for i in range(samples):
        t = rng.uniform(-1000.0, 1000.0) if i % 2 else rng.uniform(-3.0, 3.0)
        ex = mp.power(2, mp.mpf(t))
        worst = max(worst, float(abs(emulate2(t, cs, hi) - ex) / ex))
    return worst / 2.0 ** -53


def max_error(samples=20000, seed=1):
    """Max relative error of the q-form e^x (1 + r (1 + r q(r))), units of
    2^-53."""
    cs = coefficients()
    hi, lo = ln2_split()
    rng = random.Random(seed)
    worst = 0.0
    for i in range(samples):
        x = rng.uniform(-700.0, 700.0) if i % 2 else rng.uniform(-2.0, 2.0)
        ex = mp.exp(mp.mpf(x))
        worst = max(worst, float(abs(emulate(x, cs, hi, lo) - ex) / ex))
    return worst / 2.0 ** -53


MIXED_DEGREE = 4


def max_error2_mixed(samples=20000, seed=5, degree=MIXED_DEGREE):
    """Max relative error of crs::exp2_m64k (the mixed kernel's exponential:
    the exp2_64 sequence with q of lower degree), plain relative units."""
    cs = coefficients(degree)
    hi, _lo = ln2_split()
    rng = random.Random(seed)
    worst = 0.0
    for i in range(samples):
        t = rng.uniform(-1000.0, 1000.0) if i % 2 else rng.uniform(-3.0, 3.0)
        ex = mp.power(2, mp.mpf(t))
        worst = max(worst, float(abs(emulate2(t, cs, hi) - ex) / ex))
    return worst


if __name__ == '__main__':
    print('Q coefficients (exp64, exp2_64):',
          [repr(c) for c in coefficients_e()])
    print('q coefficients (exp_q_estrin):', [repr(c) for c in coefficients()])
    print('ln2 split:', [repr(v) for v in ln2_split()])
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 20000
    print('exp64 max error (units of 2^-53):', max_error_e(n))
    print('exp2_64 max error (units of 2^-53):', max_error2_e(n))
    print('q-form exp / exp2 max error (units of 2^-53):', max_error(n),
          max_error2(n))
    print('mixed q coefficients:', [repr(c) for c in
                                    coefficients(MIXED_DEGREE)])
    print('exp2_m64 max relative error:', max_error2_mixed(n))
