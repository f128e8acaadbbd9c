/*
 * Host AddressSanitizer / UndefinedBehaviorSanitizer run of the CPU oracle
 * (test infrastructure).
 *
 * Linked with one mechanism's emitted oracle sources and oracle/driver.c,
 * all built with -fsanitize=address,undefined.  Every oracle entry point is
 * called on a small seeded batch whose leading dimension exceeds the state
 * count, with every buffer allocated to its exact size plus guard words:
 *   - out-of-bounds accesses in the emitted subroutines (their stack arrays
 *     of species, reactions, reverse and falloff terms) and in the batched
 *     driver are ASan errors;
 *   - guard words before and after each array, the padding columns
 *     n..ld-1 of every row and the unread last row of Y hold a NaN canary:
 *     a write there is found by comparing the canary afterwards, a read of
 *     it by the non-finite output it produces;
 *   - signed overflow, bad shifts, misaligned or null accesses are UBSan
 *     errors (-fno-sanitize-recover: the run stops at the first).
 * Exit status 0 and a final "ok" line when every check passes.
 * Built with -DPLANT_PAD_WRITE or -DPLANT_OOB_WRITE it plants one such
 * mistake after the first oracle call, so the tests can show both checks
 * catch it.
 * Usage: oracle_san [n] [nthreads]
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int oracle_dims(int *out);
void oracle_eval_wdot(const double *T, const double *P, const double *Y,
                      double *W, int64_t n, int64_t ld, int nthreads);
void oracle_eval_stages(const double *T, const double *P, const double *Y,
                        double *y_N, double *mw, double *rho, double *C,
                        double *fwd, double *rev, double *pm, double *W,
                        int64_t n, int64_t ld, int nthreads);
void oracle_spec_rates(const double *fwd, const double *rev, const double *pm,
                       double *W, int64_t n, int64_t ld, int nthreads);
void oracle_dydt(const double *P, const double *y, double *dy, int64_t n,
                 int64_t ld, int nthreads);
void oracle_thermo(const double *T, double *cp, double *h, int64_t n,
                   int64_t ld, int nthreads);

enum { GUARD = 16 };
static const uint64_t CANARY_BITS = 0x7ff4dead5eedf00dULL; /* a quiet NaN */
static int failures = 0;

typedef struct {
    double *base; /* allocation start (guard words first) */
    double *p;    /* first element */
    int rows;
    int64_t ld;
} buf;

static double canary(void)
{
    double d;
    memcpy(&d, &CANARY_BITS, sizeof d);
    return d;
}

static int is_canary(double d)
{
    uint64_t u;
    memcpy(&u, &d, sizeof u);
    return u == CANARY_BITS;
}

static buf alloc_rows(int rows, int64_t ld)
{
    buf b;
    const int64_t total = (int64_t)rows * ld + 2 * GUARD;
    b.base = (double *)malloc((size_t)total * sizeof(double));
    if (!b.base) {
        fprintf(stderr, "out of memory\n");
        exit(2);
    }
    for (int64_t i = 0; i < total; ++i)
        b.base[i] = canary();
    b.p = b.base + GUARD;
    b.rows = rows;
    b.ld = ld;
    return b;
}

/* Guard words and padding columns untouched, the n live columns of every
 * row written (no canary left) and finite. */
static void check_out(const char *what, buf b, int64_t n, int live_rows)
{
    int bad = 0;
    for (int g = 0; g < GUARD; ++g) {
        if (!is_canary(b.base[g]))
            ++bad;
        if (!is_canary(b.p[(int64_t)b.rows * b.ld + g]))
            ++bad;
    }
    if (bad) {
        fprintf(stderr, "%s: %d guard words overwritten\n", what, bad);
        ++failures;
    }
    bad = 0;
    int unwritten = 0, nonfinite = 0;
    for (int r = 0; r < b.rows; ++r) {
        for (int64_t i = 0; i < b.ld; ++i) {
            const double v = b.p[(int64_t)r * b.ld + i];
            if (i >= n || r >= live_rows) {
                if (!is_canary(v))
                    ++bad;
            } else if (is_canary(v)) {
                ++unwritten;
            } else if (!isfinite(v)) {
                ++nonfinite;
            }
        }
    }
    if (bad || unwritten || nonfinite) {
        fprintf(stderr, "%s: %d padding entries written, %d live entries "
                "unwritten, %d non-finite\n", what, bad, unwritten, nonfinite);
        ++failures;
    }
}

static void check_same(const char *what, buf a, buf b, int64_t n, int rows)
{
    for (int r = 0; r < rows; ++r)
        for (int64_t i = 0; i < n; ++i)
            if (memcmp(&a.p[(int64_t)r * a.ld + i], &b.p[(int64_t)r * b.ld + i],
                       sizeof(double)) != 0) {
                fprintf(stderr, "%s: row %d state %lld differs\n", what, r,
                        (long long)i);
                ++failures;
                return;
            }
}

/* xorshift64*, uniform in [0, 1) */
static uint64_t rng_state = 0x9e3779b97f4a7c15ULL;
static double uniform(void)
{
    rng_state ^= rng_state >> 12;
    rng_state ^= rng_state << 25;
    rng_state ^= rng_state >> 27;
    return (double)((rng_state * 0x2545f4914f6cdd1dULL) >> 11) * 0x1.0p-53;
}

int main(int argc, char **argv)
{
    const int64_t n = argc > 1 ? atoll(argv[1]) : 37;
    const int nt = argc > 2 ? atoi(argv[2]) : 2;
    const int64_t ld = n + 5;
    int dims[4];
    oracle_dims(dims);
    const int nsp = dims[0], nr = dims[1], nrev = dims[2], npd = dims[3];
    const int rev_rows = nrev > 0 ? nrev : 1, pm_rows = npd > 0 ? npd : 1;
    printf("mechanism: %d species, %d reactions, %d reversible, %d "
           "pressure-dependent; n = %lld, ld = %lld, %d threads\n", nsp, nr,
           nrev, npd, (long long)n, (long long)ld, nt);

    /* inputs: T in [800, 2500] K, P log-uniform in 1-50 atm, Y_k (k <
     * NSP-1) a skewed draw scaled to sum 0.4; row NSP-1 of Y stays canary
     * (the oracle must never read it: it forms the bath gas as 1 - sum) */
    buf T = alloc_rows(1, ld), P = alloc_rows(1, ld), Y = alloc_rows(nsp, ld);
    for (int64_t i = 0; i < n; ++i) {
        T.p[i] = 800.0 + 1700.0 * uniform();
        P.p[i] = 101325.0 * exp(log(50.0) * uniform());
        double s = 0.0;
        for (int k = 0; k < nsp - 1; ++k) {
            const double u = uniform();
            Y.p[(int64_t)k * ld + i] = u * u * u * u + 1e-12;
            s += Y.p[(int64_t)k * ld + i];
        }
        for (int k = 0; k < nsp - 1; ++k)
            Y.p[(int64_t)k * ld + i] *= 0.4 / s;
    }

    buf W = alloc_rows(nsp, ld);
    oracle_eval_wdot(T.p, P.p, Y.p, W.p, n, ld, nt);
#ifdef PLANT_PAD_WRITE
    W.p[n] = 0.0; /* planted: a write into a padding column */
#endif
#ifdef PLANT_OOB_WRITE
    W.p[(int64_t)nsp * ld + GUARD] = 0.0; /* planted: past the allocation */
#endif
    check_out("oracle_eval_wdot", W, n, nsp);

    buf yN = alloc_rows(1, ld), mw = alloc_rows(1, ld), rho = alloc_rows(1, ld);
    buf C = alloc_rows(nsp, ld), fwd = alloc_rows(nr, ld);
    buf rev = alloc_rows(rev_rows, ld), pm = alloc_rows(pm_rows, ld);
    buf W2 = alloc_rows(nsp, ld);
    oracle_eval_stages(T.p, P.p, Y.p, yN.p, mw.p, rho.p, C.p, fwd.p, rev.p,
                       pm.p, W2.p, n, ld, 1);
    check_out("stages y_N", yN, n, 1);
    check_out("stages mw", mw, n, 1);
    check_out("stages rho", rho, n, 1);
    check_out("stages conc", C, n, nsp);
    check_out("stages fwd", fwd, n, nr);
    check_out("stages rev", rev, n, nrev);
    check_out("stages pres_mod", pm, n, npd);
    check_out("stages wdot", W2, n, nsp);
    check_same("wdot, 1 thread against several", W, W2, n, nsp);

    /* every output optional */
    buf W3 = alloc_rows(nsp, ld);
    oracle_eval_stages(T.p, P.p, Y.p, NULL, NULL, NULL, NULL, NULL, NULL,
                       NULL, W3.p, n, ld, nt);
    check_same("wdot with the other outputs NULL", W, W3, n, nsp);

    buf W4 = alloc_rows(nsp, ld);
    oracle_spec_rates(fwd.p, rev.p, pm.p, W4.p, n, ld, nt);
    check_out("oracle_spec_rates", W4, n, nsp);
    check_same("spec_rates against the fused path", W, W4, n, nsp);

    /* dydt in the reference's state layout: row 0 = T, rows 1..NSP-1 =
     * mass fractions of species 0..NSP-2 */
    buf y = alloc_rows(nsp, ld), dy = alloc_rows(nsp, ld);
    for (int64_t i = 0; i < n; ++i) {
        y.p[i] = T.p[i];
        for (int k = 0; k < nsp - 1; ++k)
            y.p[(int64_t)(k + 1) * ld + i] = Y.p[(int64_t)k * ld + i];
    }
    oracle_dydt(P.p, y.p, dy.p, n, ld, nt);
    check_out("oracle_dydt", dy, n, nsp);

    buf cp = alloc_rows(nsp, ld), h = alloc_rows(nsp, ld);
    oracle_thermo(T.p, cp.p, h.p, n, ld, nt);
    check_out("oracle_thermo cp", cp, n, nsp);
    check_out("oracle_thermo h", h, n, nsp);

    /* an empty batch touches nothing */
    buf W5 = alloc_rows(nsp, ld);
    oracle_eval_wdot(T.p, P.p, Y.p, W5.p, 0, ld, nt);
    check_out("oracle_eval_wdot, n = 0", W5, 0, nsp);

    buf all[] = {T, P, Y, W, yN, mw, rho, C, fwd, rev, pm, W2, W3, W4,
                 y, dy, cp, h, W5};
    for (size_t i = 0; i < sizeof all / sizeof all[0]; ++i)
        free(all[i].base);
    if (failures) {
        fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    printf("ok\n");
    return 0;
}
