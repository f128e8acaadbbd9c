"""Coefficients of the branch-free FP64 exponentials of crs_common.h and
their accuracy checks.

* crs::exp64: Cody-Waite reduction x = n ln2 + r, |r| <= ln2/2, ln2 split
  in two; e^r = 1 + r Q(r), Q(r) = (e^r - 1) / r of degree 9
  (``coefficients_e``), 6.3 ulp.
* crs::exp2_64 and the fused kernels' exp2_64k / exp64k: n = rint(t),
  s = t - n exact; 2^s = 1 + s Q2(s), Q2(s) = (2^s - 1) / s of degree 9 on
  [-1/2, 1/2] (``coefficients_b2``), 5.9 ulp; exp64k takes 2^(x log2 e),
  whose rounded product adds |x| ulp.
* crs::exp2_m64k (the mixed kernel): 2^s = 1 + s (ln2 + s q2(s)), q2 of
  degree 4 (``coefficients_b2m``), 1.0e-8 relative.
* crs::exp_q_estrin: e^r = 1 + r (1 + r q(r)), q(r) = (e^r - 1 - r) / r^2
  of degree 9 (``coefficients``), by Estrin's scheme, 1.2 ulp.

Coefficients are Chebyshev interpolants (mpmath.chebyfit, 60 digits)
rounded to double.  The checks emulate the device sequences exactly: every
operation is rounded once to double.
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


def coefficients_b2(degree=9):
    """Q2's coefficients (Q2(s) = (2^s - 1) / s on |s| <= 1/2), highest
    power first, rounded to double: those of crs::exp2_64 and of the fused
    kernels' register-held ExpK.  2^s = 1 + s Q2(s) needs no conversion of
    the reduced argument to base e (Q2(s) = ln2 Q(s ln2))."""
    mp.mp.dps = 60
    ln2 = mp.log(2)

    def Q2(s):
        if abs(s) < mp.mpf('1e-25'):
            return ln2
        return (mp.power(2, s) - 1) / s

    poly, _err = mp.chebyfit(Q2, [-0.5 * 1.0001, 0.5 * 1.0001], degree + 1,
                             error=True)
    return [float(c) for c in poly]


def coefficients_b2m(degree=4):
    """q2's coefficients (q2(s) = (2^s - 1 - s ln2) / s^2 on |s| <= 1/2),
    highest power first: the mixed kernel's 2^s = 1 + s (ln2 + s q2(s))."""
    mp.mp.dps = 60
    ln2 = mp.log(2)

    def q2(s):
        if abs(s) < mp.mpf('1e-12'):
            return ln2 ** 2 / 2 + s * ln2 ** 3 / 6
        return (mp.power(2, s) - 1 - s * ln2) / s ** 2

    poly, _err = mp.chebyfit(q2, [-0.5 * 1.0001, 0.5 * 1.0001], degree + 1,
                             error=True)
    return [float(c) for c in poly]


def _rd(v):
    return mp.mpf(float(v))


def _fma(a, b, c):
    return _rd(mp.mpf(a) * mp.mpf(b) + mp.mpf(c))


def emulate_b2(t, cs):
    """crs::exp2_64 / exp2_64k: n = rint(t), s = t - n (exact), then
    2^s = 1 + s Q2(s) by Horner, scaled by 2^n."""
    n = mp.nint(t)
    s = mp.mpf(t) - n
    return _rd(mp.ldexp(_tail_e(s, cs, _fma), int(n)))


def emulate_b2m(t, cs, ln2):
    """crs::exp2_m64k: n = rint(t), s = t - n, 2^s = 1 + s (ln2 + s q2(s))."""
    n = mp.nint(t)
    s = mp.mpf(t) - n
    q = mp.mpf(cs[0])
    for c in cs[1:]:
        q = _fma(q, s, c)
    p = _fma(s, _fma(s, q, ln2), 1.0)
    return _rd(mp.ldexp(p, int(n)))


def emulate_e_b2(x, cs, log2e=LOG2E):
    """crs::exp64k: e^x as 2^(x log2 e), the product rounded once."""
    return emulate_b2(_rd(mp.mpf(x) * log2e), cs)


def max_error_b2(samples=20000, seed=2):
    """Max relative error of crs::exp2_64 against 2^t, units of 2^-53."""
    cs = coefficients_b2()
    rng = random.Random(seed)
    worst = 0.0
    for i in range(samples):
        t = rng.uniform(-1000.0, 1000.0) if i % 2 else rng.uniform(-3.0, 3.0)
        ex = mp.power(2, mp.mpf(t))
        worst = max(worst, float(abs(emulate_b2(t, cs) - ex) / ex))
    return worst / 2.0 ** -53


def max_error_e_b2(samples=20000, seed=3, xmax=100.0):
    """Max relative error of crs::exp64k against e^x for |x| <= xmax, in
    units of 2^-53: the polynomial's error plus |x| 2^-53 from the rounded
    conversion product."""
    cs = coefficients_b2()
    rng = random.Random(seed)
    worst = 0.0
    for i in range(samples):
        x = rng.uniform(-xmax, xmax) if i % 2 else rng.uniform(-2.0, 2.0)
        ex = mp.exp(mp.mpf(x))
        worst = max(worst, float(abs(emulate_e_b2(x, cs) - ex) / ex))
    return worst / 2.0 ** -53


def max_error_b2m(samples=20000, seed=5, degree=4):
    """Max relative error of crs::exp2_m64k against 2^t (plain units)."""
    cs = coefficients_b2m(degree)
    ln2 = float(mp.log(2))
    rng = random.Random(seed)
    worst = 0.0
    for i in range(samples):
        t = rng.uniform(-1000.0, 1000.0) if i % 2 else rng.uniform(-3.0, 3.0)
        ex = mp.power(2, mp.mpf(t))
        worst = max(worst, float(abs(emulate_b2m(t, cs, ln2) - ex) / ex))
    return worst


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


def emulate2_estrin(t, cs, hi):
    """crs::exp2_64_estrin: n = rint(t), r = (t - n) ln2, then
    1 + r (1 + r q(r)) with q(r) by Estrin's scheme."""
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


if __name__ == '__main__':
    print('Q coefficients (exp64):', [repr(c) for c in coefficients_e()])
    print('Q2 coefficients (exp2_64, ExpK):',
          [repr(c) for c in coefficients_b2()])
    print('q2 coefficients (exp2_m64k):',
          [repr(c) for c in coefficients_b2m()])
    print('q coefficients (exp_q_estrin):', [repr(c) for c in coefficients()])
    print('ln2 split:', [repr(v) for v in ln2_split()])
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 20000
    print('exp64 max error (units of 2^-53):', max_error_e(n))
    print('exp2_64 max error (units of 2^-53):', max_error_b2(n))
    print('exp64k at |x| <= 100 (units of 2^-53):', max_error_e_b2(n))
    print('q-form exp / Estrin exp2 max error (units of 2^-53):', max_error(n),
          max_error2_estrin(n))
    print('exp2_m64k max relative error:', max_error_b2m(n))
