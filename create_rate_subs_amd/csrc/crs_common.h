// Runtime support shared by every emitted mechanism library.
//
// Included once by each generated crs_<mech>.hip: per-thread error text,
// argument checks and launch helpers behind the C ABI of include/crs_hip.h.
// Device code here is mechanism independent.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <mutex>

#include "crs_hip.h"
#include "crs_overlap.h"

namespace crs {

// Threads per block for the per-state kernels: four wave64 wavefronts.
constexpr int kBlock = 256;

// Most states one launch of a fused kernel addresses: a state's byte offset
// within a row, 8 x index, is a 32-bit value (ld_row / st_row below).
constexpr int64_t kMaxChunk = (int64_t)1 << 29;

// The host helpers below have internal linkage: every library that
// includes this file (each mechanism library, the surrogate library) keeps
// its own error text, so one library's failure never overwrites the
// message another library's crs_last_error / crs_mlp_last_error reports.
namespace {

inline char *error_buffer()
{
    static thread_local char buf[512] = {0};
    return buf;
}

inline int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(error_buffer(), 512, fmt, ap);
    va_end(ap);
    return code;
}

inline void clear_error() { error_buffer()[0] = '\0'; }

// Checks shared by all batched entry points.  A null pointer is only an
// error when there is work to do.
inline int check_sizes(const char *fn, int64_t n, int64_t ld)
{
    if (n < 0)
        return fail(CRS_ERR_ARG, "%s: n must be >= 0 (got %lld)", fn,
                    (long long)n);
    if (ld < n)
        return fail(CRS_ERR_ARG, "%s: ld must be >= n (ld=%lld, n=%lld)", fn,
                    (long long)ld, (long long)n);
    return CRS_OK;
}

inline int check_ptr(const char *fn, const void *p, const char *what, int64_t n)
{
    if (n > 0 && p == nullptr)
        return fail(CRS_ERR_ARG, "%s: %s is NULL", fn, what);
    return CRS_OK;
}

inline int check_disjoint(const char *fn, const void *a, int64_t ra,
                          const char *wa, const void *b, int64_t rb,
                          const char *wb, int64_t n, int64_t ld)
{
    if (rows_overlap(a, ra, b, rb, n, ld))
        return fail(CRS_ERR_ARG, "%s: %s and %s overlap", fn, wa, wb);
    return CRS_OK;
}

inline int launched(const char *fn)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(CRS_ERR_HIP, "%s: kernel launch failed: %s", fn,
                    hipGetErrorString(e));
    return CRS_OK;
}

inline dim3 grid_for(int64_t n, int block = kBlock)
{
    return dim3((unsigned)((n + block - 1) / block));
}

}  // namespace

// 1/x to within an ulp: hardware reciprocal estimate refined by two
// Newton-Raphson steps (each doubles the correct bits).  For positive
// normal x away from the overflow boundary.
__device__ __forceinline__ double rcp64(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

// e^x in FP64 to about 6 ulp (max relative error 6.3 * 2^-53 measured by
// exact emulation of the FMA sequence).  Cody-Waite reduction
// x = n ln2 + r, |r| <= ln2/2, with ln2 split into a double and its
// remainder; e^r = 1 + r Q(r) where Q(r) = (e^r - 1) / r is the degree-9
// Chebyshev interpolant on [-ln2/2, ln2/2] (coefficients from
// exp64_coeffs.py beside this file: ten coefficients, nine Horner FMAs and
// one final FMA); the result is scaled by 2^n with ldexp, which also
// delivers overflow to +inf and underflow to 0.  It takes no branch and no
// range test, unlike the library exp: finite arguments and NaN only
// (exp(-inf) would give NaN, which no rate expression produces).
__device__ __forceinline__ double exp64(double x)
{
    const double n = __builtin_rint(x * 1.4426950408889634);
    double r = __builtin_fma(-n, 0.6931471805599453, x);
    r = __builtin_fma(-n, 2.3190468138462996e-17, r);
    double q = 2.7620088445409746e-07;
    q = __builtin_fma(q, r, 2.763265575933203e-06);
    q = __builtin_fma(q, r, 2.4801521295954376e-05);
    q = __builtin_fma(q, r, 0.0001984119061583585);
    q = __builtin_fma(q, r, 0.0013888888917213717);
    q = __builtin_fma(q, r, 0.008333333367332009);
    q = __builtin_fma(q, r, 0.04166666666662413);
    q = __builtin_fma(q, r, 0.16666666666615607);
    q = __builtin_fma(q, r, 0.5000000000000001);
    q = __builtin_fma(q, r, 1.0000000000000013);
    return __builtin_ldexp(__builtin_fma(r, q, 1.0), (int)n);
}

// 2^t in FP64 to about 6 ulp (5.9 * 2^-53 max by exact emulation): t = n +
// s with n = rint(t) and s = t - n exact, |s| <= 1/2; 2^s = 1 + s Q2(s)
// where Q2(s) = (2^s - 1) / s is the degree-9 Chebyshev interpolant on
// [-1/2, 1/2] (exp64_coeffs.py coefficients_b2), so the reduced argument
// needs no conversion to base e.  For exponents formed in base 2 (constants
// pre-scaled by log2 e): no conversion multiply, no two-step reduction.
__device__ __forceinline__ double exp2_64(double t)
{
    const double n = __builtin_rint(t);
    const double s = t - n;
    double q = 7.070981083365768e-09;
    q = __builtin_fma(q, s, 1.0205910984889862e-07);
    q = __builtin_fma(q, s, 1.3215451619146607e-06);
    q = __builtin_fma(q, s, 1.52526729004753e-05);
    q = __builtin_fma(q, s, 0.0001540353042479538);
    q = __builtin_fma(q, s, 0.001333355820082724);
    q = __builtin_fma(q, s, 0.009618129107618658);
    q = __builtin_fma(q, s, 0.05550410866465154);
    q = __builtin_fma(q, s, 0.24022650695910075);
    q = __builtin_fma(q, s, 0.6931471805599462);
    return __builtin_ldexp(__builtin_fma(s, q, 1.0), (int)n);
}

// The constants of exp2_64 as run-time values held in registers.  A fused
// kernel takes them once at its start (ExpK k = exp_consts()) and passes
// them to exp2_64k / exp64k: each reaction block is its own basic block,
// and with literal coefficients every block would materialise all of them
// again (two scalar moves per constant, plus a vector move for the one that
// must sit in a VGPR: a VALU instruction takes one scalar operand).  The
// values pass through empty asm so the compiler cannot fold them back into
// literals; b8 is kept in a VGPR pair for the first Horner step.
struct ExpK {
    double b9, b8, b7, b6, b5, b4, b3, b2, b1, b0;
    double log2e, sqrt2;
};

__device__ __forceinline__ double sgpr_const(double x)
{
    asm("; const" : "+s"(x));
    return x;
}

__device__ __forceinline__ double vgpr_const(double x)
{
    asm("; const" : "+v"(x));
    return x;
}

__device__ __forceinline__ ExpK exp_consts()
{
    ExpK k;
    k.b9 = sgpr_const(7.070981083365768e-09);
    k.b8 = vgpr_const(1.0205910984889862e-07);
    k.b7 = sgpr_const(1.3215451619146607e-06);
    k.b6 = sgpr_const(1.52526729004753e-05);
    k.b5 = sgpr_const(0.0001540353042479538);
    k.b4 = sgpr_const(0.001333355820082724);
    k.b3 = sgpr_const(0.009618129107618658);
    k.b2 = sgpr_const(0.05550410866465154);
    k.b1 = sgpr_const(0.24022650695910075);
    k.b0 = sgpr_const(0.6931471805599462);
    k.log2e = sgpr_const(1.4426950408889634);
    k.sqrt2 = sgpr_const(1.4142135623730951);
    return k;
}

// exp2_64 on the register-held constants: the same operations in the same
// order, so the same results bit for bit.
__device__ __forceinline__ double exp2_64k(const ExpK &k, double t)
{
    const double n = __builtin_rint(t);
    const double s = t - n;
    double q = __builtin_fma(k.b9, s, k.b8);
    q = __builtin_fma(q, s, k.b7);
    q = __builtin_fma(q, s, k.b6);
    q = __builtin_fma(q, s, k.b5);
    q = __builtin_fma(q, s, k.b4);
    q = __builtin_fma(q, s, k.b3);
    q = __builtin_fma(q, s, k.b2);
    q = __builtin_fma(q, s, k.b1);
    q = __builtin_fma(q, s, k.b0);
    return __builtin_ldexp(__builtin_fma(s, q, 1.0), (int)n);
}

// e^x in the fused kernels as 2^(x log2 e): the rounding of the product
// adds |x| 2^-53 to the relative error (102 * 2^-53 at |x| <= 100, exact
// emulation in exp64_coeffs.py).  The emitter pre-scales the arguments it
// forms itself (NASA Gibbs sums, Troe and Arrhenius exponents) and calls
// exp2_64k directly; this form is left for the rarer rate terms (SRI
// blending, d T^e).
__device__ __forceinline__ double exp64k(const ExpK &k, double x)
{
    return exp2_64k(k, x * k.log2e);
}

// log2 x for positive normal x (callers clamp with fmax first), to about
// 1e-15 absolute (exact emulation in log64_coeffs.py with the hardware
// seed taken 2^-20 off).  x = 2^e m, m in [1/2, 1); the hardware FP32
// log2 of m gives y0 within about 2^-22; s = -y0 - 1/2 (exact, |s| <= 1/2),
// 2^-y0 = sqrt2 * 2^s from the exp2_64 polynomial, and the residual
// t = m 2^-y0 - 1 (|t| < 1e-6) corrects y0 by log2(1 + t) =
// (t - t^2/2) log2 e to within t^3.  About 23 VALU instructions against
// log64's 34; the error is absolute (the falloff blocks' log10 Pr and
// log10 Fcent enter sums, where absolute error is what counts), so it is
// not a substitute for log64 near x = 1 where a relative bound matters.
// +inf gives NaN.
__device__ __forceinline__ double log2_64k(const ExpK &k, double x)
{
    const double m = __builtin_amdgcn_frexp_mant(x);
    const int e = __builtin_amdgcn_frexp_exp(x);
    const double y0 = (double)__builtin_amdgcn_logf((float)m);
    const double s = -y0 - 0.5;
    double q = __builtin_fma(k.b9, s, k.b8);
    q = __builtin_fma(q, s, k.b7);
    q = __builtin_fma(q, s, k.b6);
    q = __builtin_fma(q, s, k.b5);
    q = __builtin_fma(q, s, k.b4);
    q = __builtin_fma(q, s, k.b3);
    q = __builtin_fma(q, s, k.b2);
    q = __builtin_fma(q, s, k.b1);
    q = __builtin_fma(q, s, k.b0);
    const double p = __builtin_fma(s, q, 1.0);
    const double t = __builtin_fma(m * k.sqrt2, p, -1.0);
    const double u = __builtin_fma(-0.5 * t, t, t);
    return __builtin_fma(u, k.log2e, y0) + (double)e;
}

__device__ __forceinline__ double log2_64(double x)
{
    return log2_64k(exp_consts(), x);
}

// Exponentials of the mixed-precision kernel: the same reduction as
// exp2_64 (n = rint(t), s = t - n exact), then 2^s = 1 + s (ln2 + s q2(s))
// with q2(s) = (2^s - 1 - s ln2) / s^2 of degree 4 (exp64_coeffs.py
// coefficients_b2m): relative error 1.0e-8 by exact emulation, five FMAs
// fewer than exp2_64.  FP64 arithmetic at about 27 bits of accuracy, better
// than the FP32 hardware exponential (about 1e-7, which a species sum with
// cancellation amplifies up to 20-fold) and about as cheap once the
// conversions and the quarter issue rate of v_exp_f32 are counted.
struct ExpMK {
    double c4, c3, c2, c1, c0, ln2, log2e, sqrt2;
};

__device__ __forceinline__ ExpMK exp_m_consts()
{
    ExpMK k;
    k.c4 = sgpr_const(0.000154448922353153);
    k.c3 = vgpr_const(0.001338131210261626);
    k.c2 = sgpr_const(0.00961810324267985);
    k.c1 = sgpr_const(0.05550381001841135);
    k.c0 = sgpr_const(0.24022650695910072);
    k.ln2 = sgpr_const(0.6931471805599453);
    k.log2e = sgpr_const(1.4426950408889634);
    k.sqrt2 = sgpr_const(1.4142135623730951);
    return k;
}

__device__ __forceinline__ double exp2_m64k(const ExpMK &k, double t)
{
    const double n = __builtin_rint(t);
    const double s = t - n;
    double q = __builtin_fma(k.c4, s, k.c3);
    q = __builtin_fma(q, s, k.c2);
    q = __builtin_fma(q, s, k.c1);
    q = __builtin_fma(q, s, k.c0);
    const double p = __builtin_fma(s, __builtin_fma(s, q, k.ln2), 1.0);
    return __builtin_ldexp(p, (int)n);
}

// e^x = 2^(x log2 e) for the mixed kernel (one more rounding, of the
// product: |x| 2^-53 relative, negligible beside 1.0e-8 for |x| < 1e4).
__device__ __forceinline__ double exp_m64k(const ExpMK &k, double x)
{
    return exp2_m64k(k, x * k.log2e);
}

__device__ __forceinline__ double exp2_m64(double t)
{
    return exp2_m64k(exp_m_consts(), t);
}

__device__ __forceinline__ double exp_m64(double x)
{
    return exp_m64k(exp_m_consts(), x);
}

// log2_64k at the mixed kernel's accuracy: the same residual correction on
// the degree-4 exp2 form, about 1.5e-8 absolute (log64_coeffs.py), five
// FMAs fewer.
__device__ __forceinline__ double log2_m64k(const ExpMK &k, double x)
{
    const double m = __builtin_amdgcn_frexp_mant(x);
    const int e = __builtin_amdgcn_frexp_exp(x);
    const double y0 = (double)__builtin_amdgcn_logf((float)m);
    const double s = -y0 - 0.5;
    double q = __builtin_fma(k.c4, s, k.c3);
    q = __builtin_fma(q, s, k.c2);
    q = __builtin_fma(q, s, k.c1);
    q = __builtin_fma(q, s, k.c0);
    const double p = __builtin_fma(s, __builtin_fma(s, q, k.ln2), 1.0);
    const double t = __builtin_fma(m * k.sqrt2, p, -1.0);
    const double u = __builtin_fma(-0.5 * t, t, t);
    return __builtin_fma(u, k.log2e, y0) + (double)e;
}

__device__ __forceinline__ double log2_m64(double x)
{
    return log2_m64k(exp_m_consts(), x);
}

// exp2_64 / exp64 in the q form, e^r = 1 + r (1 + r q(r)) with
// q(r) = (e^r - 1 - r) / r^2 of degree 9 (exp64_coeffs.py coefficients()),
// q by Estrin's scheme: 3 more multiplies than Horner's, a dependency
// chain of 5 operations instead of 10 (1.20 ulp max by exact emulation in
// exp64_coeffs.py).  An emitter option (exp_poly), measured slower.
__device__ __forceinline__ double exp_q_estrin(double r)
{
    const double r2 = r * r;
    const double r4 = r2 * r2;
    const double r8 = r4 * r4;
    const double p01 = __builtin_fma(0.16666666666666669, r, 0.5000000000000001);
    const double p23 = __builtin_fma(0.008333333333330062, r, 0.04166666666662413);
    const double p45 = __builtin_fma(0.00019841269863053618, r, 0.0013888888917213717);
    const double p67 = __builtin_fma(2.7557268459997064e-06, r, 2.4801521295954376e-05);
    const double p89 = __builtin_fma(2.510038549551032e-08, r, 2.7620088445409746e-07);
    const double p03 = __builtin_fma(p23, r2, p01);
    const double p47 = __builtin_fma(p67, r2, p45);
    return __builtin_fma(p89, r8, __builtin_fma(p47, r4, p03));
}

__device__ __forceinline__ double exp2_64_estrin(double t)
{
    const double n = __builtin_rint(t);
    const double r = (t - n) * 0.6931471805599453;
    const double q = exp_q_estrin(r);
    const double p = __builtin_fma(r, __builtin_fma(r, q, 1.0), 1.0);
    return __builtin_ldexp(p, (int)n);
}

__device__ __forceinline__ double exp64_estrin(double x)
{
    const double n = __builtin_rint(x * 1.4426950408889634);
    double r = __builtin_fma(-n, 0.6931471805599453, x);
    r = __builtin_fma(-n, 2.3190468138462996e-17, r);
    const double q = exp_q_estrin(r);
    const double p = __builtin_fma(r, __builtin_fma(r, q, 1.0), 1.0);
    return __builtin_ldexp(p, (int)n);
}

// ln x in FP64 to about 2 ulp (max relative error 1.81 * 2^-53 measured by
// exact emulation with a reciprocal seed 2^-23 off).  x = 2^e m with m in
// [sqrt(1/2), sqrt(2)), f = m - 1 (exact), s = f / (2 + f) from the
// reciprocal with one correction step that also takes up the rounding of
// 2 + f, and ln m = 2 s + s^3 P(s^2) with P the degree-7 Chebyshev
// interpolant of (2 atanh(sqrt z) - 2 sqrt z) / z^(3/2) on
// [0, (3 - 2 sqrt 2)^2] (coefficients from log64_coeffs.py beside this
// file); ln x = e ln2 + ln m with ln2 split in two.  About 35 VALU
// instructions against 90 for the library log.  Positive normal or
// subnormal finite x (callers clamp with fmax first); +inf gives NaN, which
// log64 below turns into +inf.
__device__ __forceinline__ double log64_finite(double x)
{
    double m = __builtin_amdgcn_frexp_mant(x);
    int e = __builtin_amdgcn_frexp_exp(x);
    const bool lo = m < 0.7071067811865476;
    m = lo ? m + m : m;
    e = lo ? e - 1 : e;
    const double f = m - 1.0;
    const double d = 2.0 + f;
    const double r = rcp64(d);
    const double q = f * r;
    const double dl = f - (d - 2.0);
    double res = __builtin_fma(-q, d, f);
    res = __builtin_fma(-q, dl, res);
    const double s = __builtin_fma(res, r, q);
    const double z = s * s;
    double p = 0.13086767009948175;
    p = __builtin_fma(p, z, 0.13268759668560612);
    p = __builtin_fma(p, z, 0.15386240205813478);
    p = __builtin_fma(p, z, 0.18181795630884753);
    p = __builtin_fma(p, z, 0.2222222239180047);
    p = __builtin_fma(p, z, 0.2857142857080323);
    p = __builtin_fma(p, z, 0.4000000000000088);
    p = __builtin_fma(p, z, 0.6666666666666666);
    const double lm = __builtin_fma(s * z, p, 2.0 * s);
    const double de = (double)e;
    return __builtin_fma(de, 0.6931471805599453,
                         __builtin_fma(de, 2.3190468138462996e-17, lm));
}

// log64_finite with ln(+inf) = +inf, for arguments that may overflow.
__device__ __forceinline__ double log64(double x)
{
    const double y = log64_finite(x);
    return x == __builtin_inf() ? x : y;
}

// x^y for a concentration x and a non-integer reaction order y > 0, as
// exp(y ln x) with the functions above (relative error about
// (1.2 + 1.8 |y ln x|) ulp, under 1e-13 for every x above 1e-300 and
// y <= 4).  Matches the library pow's special cases that occur here:
// 0^y = 0, a negative base gives NaN, NaN stays NaN.
__device__ __forceinline__ double pow64(double x, double y)
{
    const double r = exp64(y * log64_finite(fmax(x, 4.9406564584124654e-324)));
    return x > 0.0 ? r : (x == 0.0 ? 0.0 : __builtin_nan(""));
}

// Per-lane LDS storage.  A copy of an LDS pointer taken through an empty asm
// is a value the compiler cannot see through: loads made through it can be
// neither hoisted above the point where the copy is taken nor forwarded from
// the stores that filled the array, so data parked in LDS stays there until
// the code that needs it.
#define CRS_LDS __attribute__((address_space(3)))

__device__ __forceinline__ CRS_LDS double *opaque_lds(CRS_LDS double *p)
{
    asm volatile("" : "+v"(p));
    return p;
}

// Row access of the [rows][ld] arrays with a uniform row pointer and a
// 32-bit per-lane byte offset: the address is (row in SGPRs) + zext(off),
// which the hardware takes directly (global_load/store with an SGPR base),
// instead of one 64-bit VGPR address per row.  off = 8 * state index.
__device__ __forceinline__ double ld_row(const double *__restrict__ row,
                                         uint32_t off)
{
    return *(const double *)((const char *)row + off);
}

__device__ __forceinline__ void st_row(double *__restrict__ row, uint32_t off,
                                       double v)
{
    *(double *)((char *)row + off) = v;
}

// A copy of x the compiler cannot prove equal to x or to any copy with
// another Tag (the tag is an operand of the empty asm).  Each reaction block
// of the fused kernels reads its per-state inputs through its own copies, so
// equal products in two blocks (H_k H_k, G_k H_k, b log2 T) are not merged
// into one value kept live between them.  No instruction is emitted.
template <int Tag>
__device__ __forceinline__ double fresh(double x)
{
    asm("; fresh %1" : "+v"(x) : "n"(Tag));
    return x;
}

// The same for a wave-uniform integer held in SGPRs (a row stride): the
// row address of a late store is then formed where it is used instead of
// being shared with an early load's and kept in SGPRs in between.
template <int Tag>
__device__ __forceinline__ int64_t fresh_s(int64_t x)
{
    asm("; fresh %1" : "+s"(x) : "n"(Tag));
    return x;
}

// A wave-uniform true the compiler cannot see through, one per Tag.  The
// fused kernels open every reaction block with `if (crs::go<N>())`: the
// block becomes its own basic block, so the code generator materialises the
// block's literal constants inside it instead of sharing equal 32-bit halves
// of constants across the whole straight-line kernel (kept in SGPRs from
// the first use to the last, they were spilled to VGPR lanes).  Costs one
// scalar move, compare and branch per block.
template <int Tag>
__device__ __forceinline__ bool go()
{
    int v;
    asm("s_mov_b32 %0, 1 ; go %1" : "=s"(v) : "n"(Tag));
    return v != 0;
}

// s/R - h/RT of one species from its 7 coefficients (in LDS) on the basis
// [1, ln T, T, T^2, T^3, T^4, 1/T].
__device__ __forceinline__ double nasa_smh(CRS_LDS const double *c,
                                           double logT, double T, double T2,
                                           double T3, double T4, double iT)
{
    return c[0] + c[1] * logT + c[2] * T + c[3] * T2 + c[4] * T3 +
           c[5] * T4 + c[6] * iT;
}

// ---------------------------------------------------------------------------
// Mixed-precision building blocks.  The argument of an exponential is formed
// and range-reduced in FP64 (x log2 e = n + f, |f| <= 1/2, exact since
// n = rint(x log2 e)); only 2^f is taken on the FP32 hardware exponential
// (v_exp_f32), then rescaled by 2^n.  The relative error is that of v_exp_f32
// on [-1/2, 1/2] plus the FP32 rounding of f, about 1e-7, independent of |x|.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double exp_mixed(double x)
{
    const double t = x * 1.4426950408889634;   // log2(e)
    const double n = __builtin_rint(t);
    const float f = (float)(t - n);
    return __builtin_ldexp((double)__builtin_amdgcn_exp2f(f), (int)n);
}

// 2^t with the same split (t = n + f, only 2^f on v_exp_f32): the base-2
// form of exp_mixed for exponents already scaled by log2 e.
__device__ __forceinline__ double exp2_mixed(double t)
{
    const double n = __builtin_rint(t);
    const float f = (float)(t - n);
    return __builtin_ldexp((double)__builtin_amdgcn_exp2f(f), (int)n);
}

// Same reduction, FP32 result (for arguments whose result fits FP32).
__device__ __forceinline__ float expf_mixed(double x)
{
    const double t = x * 1.4426950408889634;
    const double n = __builtin_rint(t);
    const float f = (float)(t - n);
    return __builtin_ldexpf(__builtin_amdgcn_exp2f(f), (int)n);
}

// log10 of a positive normal FP32 value on the hardware log2.
__device__ __forceinline__ float log10f_hw(float x)
{
    return __builtin_amdgcn_logf(x) * 0.30102999566398120f;
}

// 10^x for |x| of order one on the hardware exp2.
__device__ __forceinline__ float exp10f_hw(float x)
{
    return __builtin_amdgcn_exp2f(x * 3.3219280948873623f);
}

__device__ __forceinline__ float rcpf_hw(float x)
{
    return __builtin_amdgcn_rcpf(x);
}

}  // namespace crs

#define CRS_TRY(expr)                 \
    do {                              \
        int crs_rc_ = (expr);         \
        if (crs_rc_ != CRS_OK)        \
            return crs_rc_;           \
    } while (0)
