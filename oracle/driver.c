/*
 * Batched OpenMP driver for the CPU oracle (test infrastructure only).
 *
 * Compiled together with the C emitted by oracle/emit_c.py.  Each state is
 * gathered from structure-of-arrays input into stack arrays, run through
 * eval_conc -> eval_rxn_rates -> get_rxn_pres_mod -> eval_spec_rates exactly
 * as the reference's generated dydt does (rate_subs.py:1845-1880), and its
 * rates scattered back.  The state loop is the reference's host "thread"
 * loop (create_rate_subs.py:60-63) with `#pragma omp parallel for` on it.
 *
 * Layout: T[n], P[n] (Pa), Y[NSP][ld] with only rows 0..NSP-2 read,
 * output rows [NSP][ld] (row NSP-1 receives dy_N).
 */
#include <omp.h>
#include <stdint.h>

#include "header.h"
#include "chem_utils.h"
#include "rates.h"
#include "mech_dims.h"

#define NREV_ALLOC (MECH_NREV > 0 ? MECH_NREV : 1)
#define NPD_ALLOC (MECH_NPD > 0 ? MECH_NPD : 1)

int oracle_dims(int *out)
{
    out[0] = NSP;
    out[1] = MECH_NR;
    out[2] = MECH_NREV;
    out[3] = MECH_NPD;
    return 0;
}

static void one_state(double T, double P, const double *Y, int64_t ld,
                      int64_t i, double *conc, double *fwd, double *rev,
                      double *pm, double *sp, double *y_N, double *mw,
                      double *rho)
{
    double y[NSP];
    for (int k = 0; k < NSP - 1; ++k)
        y[k] = Y[(int64_t)k * ld + i];
    eval_conc(T, P, y, y_N, mw, rho, conc);
    eval_rxn_rates(T, P, conc, fwd, rev);
#if MECH_NPD > 0
    get_rxn_pres_mod(T, P, conc, pm);
#endif
    eval_spec_rates(fwd, rev, pm, sp, &sp[NSP - 1]);
}

/* wdot[NSP][ld] for n states with nthreads OpenMP threads. */
void oracle_eval_wdot(const double *T, const double *P, const double *Y,
                      double *W, int64_t n, int64_t ld, int nthreads)
{
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t i = 0; i < n; ++i) {
        double conc[NSP], fwd[MECH_NR], rev[NREV_ALLOC], pm[NPD_ALLOC];
        double sp[NSP], y_N, mw, rho;
        one_state(T[i], P[i], Y, ld, i, conc, fwd, rev, pm, sp, &y_N, &mw,
                  &rho);
        for (int k = 0; k < NSP; ++k)
            W[(int64_t)k * ld + i] = sp[k];
    }
}

/*
 * CPU baseline timing: one untimed warm-up pass of oracle_eval_wdot, then
 * `reps` timed passes; times_out[r] is pass r's wall time from
 * omp_get_wtime (the reference's generated main times its loop with
 * clock(), create_rate_subs.py:202, :285-302).
 */
void oracle_time_wdot(const double *T, const double *P, const double *Y,
                      double *W, int64_t n, int64_t ld, int nthreads,
                      int reps, double *times_out)
{
    oracle_eval_wdot(T, P, Y, W, n, ld, nthreads);
    for (int r = 0; r < reps; ++r) {
        const double t0 = omp_get_wtime();
        oracle_eval_wdot(T, P, Y, W, n, ld, nthreads);
        times_out[r] = omp_get_wtime() - t0;
    }
}

/*
 * Every intermediate of the path for n states: y_N/mw/rho [n],
 * conc [NSP][ld], fwd [NR][ld], rev [NREV][ld], pres_mod [NPD][ld],
 * wdot [NSP][ld].  Any output pointer may be NULL.
 */
void oracle_eval_stages(const double *T, const double *P, const double *Y,
                        double *y_N_out, double *mw_out, double *rho_out,
                        double *C_out, double *fwd_out, double *rev_out,
                        double *pm_out, double *W, int64_t n, int64_t ld,
                        int nthreads)
{
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t i = 0; i < n; ++i) {
        double conc[NSP], fwd[MECH_NR], rev[NREV_ALLOC], pm[NPD_ALLOC];
        double sp[NSP], y_N, mw, rho;
        one_state(T[i], P[i], Y, ld, i, conc, fwd, rev, pm, sp, &y_N, &mw,
                  &rho);
        if (y_N_out) y_N_out[i] = y_N;
        if (mw_out) mw_out[i] = mw;
        if (rho_out) rho_out[i] = rho;
        if (C_out)
            for (int k = 0; k < NSP; ++k) C_out[(int64_t)k * ld + i] = conc[k];
        if (fwd_out)
            for (int j = 0; j < MECH_NR; ++j) fwd_out[(int64_t)j * ld + i] = fwd[j];
        if (rev_out)
            for (int j = 0; j < MECH_NREV; ++j) rev_out[(int64_t)j * ld + i] = rev[j];
        if (pm_out)
            for (int j = 0; j < MECH_NPD; ++j) pm_out[(int64_t)j * ld + i] = pm[j];
        if (W)
            for (int k = 0; k < NSP; ++k) W[(int64_t)k * ld + i] = sp[k];
    }
}

/* eval_spec_rates alone on given fwd/rev/pres_mod rows (stage parity). */
void oracle_spec_rates(const double *fwd_in, const double *rev_in,
                       const double *pm_in, double *W, int64_t n, int64_t ld,
                       int nthreads)
{
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t i = 0; i < n; ++i) {
        double fwd[MECH_NR], rev[NREV_ALLOC], pm[NPD_ALLOC], sp[NSP];
        for (int j = 0; j < MECH_NR; ++j) fwd[j] = fwd_in[(int64_t)j * ld + i];
        for (int j = 0; j < MECH_NREV; ++j) rev[j] = rev_in[(int64_t)j * ld + i];
        for (int j = 0; j < MECH_NPD; ++j) pm[j] = pm_in[(int64_t)j * ld + i];
        eval_spec_rates(fwd, rev, pm, sp, &sp[NSP - 1]);
        for (int k = 0; k < NSP; ++k) W[(int64_t)k * ld + i] = sp[k];
    }
}

/*
 * Constant-pressure right-hand side for n states: y and dy are [rows][ld]
 * with row 0 = T, rows 1..NSP-1 = mass fractions of species 0..NSP-2 (the
 * reference's y / dy, rate_subs.py:1829-1946); P[n] in Pa.  Rows beyond
 * NSP-1 are neither read nor written.
 */
void oracle_dydt(const double *P, const double *y, double *dy, int64_t n,
                 int64_t ld, int nthreads)
{
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t i = 0; i < n; ++i) {
        double ys[NSP], dys[NSP];
        for (int k = 0; k < NSP; ++k)
            ys[k] = y[(int64_t)k * ld + i];
        dydt(0.0, P[i], ys, dys);
        for (int k = 0; k < NSP; ++k)
            dy[(int64_t)k * ld + i] = dys[k];
    }
}

/* Species cp [J/kg/K] and h [J/kg], [NSP][ld], at n temperatures. */
void oracle_thermo(const double *T, double *cp, double *h, int64_t n,
                   int64_t ld, int nthreads)
{
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int64_t i = 0; i < n; ++i) {
        double c[NSP], e[NSP];
        eval_cp(T[i], c);
        eval_h(T[i], e);
        for (int k = 0; k < NSP; ++k) {
            cp[(int64_t)k * ld + i] = c[k];
            h[(int64_t)k * ld + i] = e[k];
        }
    }
}
