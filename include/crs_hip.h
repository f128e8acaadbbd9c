/*
 * C ABI of a create_rate_subs_amd mechanism library (libcrs_<mech>.so).
 *
 * One shared library is emitted and built per mechanism, the way the
 * reference emits one set of C/CUDA subroutines per mechanism.  Every
 * library exports the symbols below with plain C types only; HIP streams
 * travel as `void *` (a hipStream_t, NULL for the default stream).
 *
 * Data layout (device memory, FP64, structure of arrays): T[n] in K,
 * P[n] in Pa, and every per-species / per-reaction quantity as rows of a
 * [rows][ld] array, element (row r, state i) at r*ld + i, ld >= n.
 * Species are in the library's order (bath gas last).  Mass-fraction input
 * Y has NSP rows of which rows 0..NSP-2 are read; the bath-gas fraction is
 * 1 - sum(others), as in the reference's eval_conc.
 *
 * Errors: every entry point returns CRS_OK (0) or a nonzero CRS_ERR_* code
 * and never aborts; crs_last_error() gives a message for the calling
 * thread.  Work is enqueued asynchronously on the given stream; results
 * are ready when the stream is synchronised.  The library holds no mutable
 * state besides per-thread error text and per-device constant-table
 * addresses, so concurrent calls from several threads are safe.
 */
#ifndef CRS_HIP_H
#define CRS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CRS_OK 0
#define CRS_ERR_ARG 1         /* bad pointer, size, leading dimension or mode */
#define CRS_ERR_HIP 2         /* a HIP runtime call or kernel launch failed */
#define CRS_ERR_UNSUPPORTED 3 /* mode not compiled into this library */

/* Precision modes of crs_eval_wdot. */
#define CRS_MODE_FP64 0  /* FP64 throughout, the drop-in path */
/* Reduced-precision rate constants, FP64 Kc / net rate / sums: the
 * Arrhenius and falloff exponentials to 1.0e-8 relative (a degree-4 FP64
 * polynomial), everything else FP64; max conditioned error 1.9e-7 against
 * CRS_MODE_FP64 (53 species / 325 reactions, 2^22 random states, tau 0.05),
 * 1.09-1.15x its throughput.  A library emitted with mixed_arith='fp32' takes
 * FP32 exponential cores, falloff and third-body terms here instead:
 * 1.28-1.36x, max 3.8e-6. */
#define CRS_MODE_MIXED 1

typedef struct crs_mech_info {
    int32_t nsp;  /* species (bath gas last) */
    int32_t nr;   /* reactions */
    int32_t nrev; /* reversible reactions (rows of rev_rates) */
    int32_t npd;  /* third-body + falloff reactions (rows of pres_mod) */
    int32_t modes; /* bit (1 << mode) set for every mode compiled in */
    const char *name;
} crs_mech_info;

/* Sizes of this library's mechanism.  Host-only; no device work. */
int crs_get_mech_info(crs_mech_info *out);

/* Message for the last failing call on this thread ("" if none). */
const char *crs_last_error(void);

/*
 * Batched eval_conc: per state, the implied bath-gas fraction y_N, the mean
 * molecular weight mw_avg, the density rho and concentrations conc[NSP].
 * Replaces `void eval_conc(const double T, const double pres,
 * const double *y, double *y_N, double *mw_avg, double *rho, double *conc)`
 * emitted by rate_subs.py:1292-1391 (prototype rate_subs.py:1267-1282).
 * y_N, mw_avg, rho: [n]; conc: [NSP][ld].
 */
int crs_eval_conc(const double *T, const double *P, const double *Y,
                  double *y_N, double *mw_avg, double *rho, double *conc,
                  int64_t n, int64_t ld, void *stream);

/*
 * Batched eval_rxn_rates: forward rates fwd[NR][ld] and reverse rates of
 * the reversible reactions rev[NREV][ld] (compact order, as the reference).
 * Replaces `void eval_rxn_rates(const double T, const double pres,
 * const double *C, double *fwd_rxn_rates, double *rev_rxn_rates)`
 * emitted by rate_subs.py:226-704 (prototype rate_subs.py:259-272).
 */
int crs_eval_rxn_rates(const double *T, const double *P, const double *C,
                       double *fwd, double *rev, int64_t n, int64_t ld,
                       void *stream);

/*
 * Batched get_rxn_pres_mod: third-body concentrations and falloff factors
 * pres_mod[NPD][ld].  Replaces `void get_rxn_pres_mod(const double T,
 * const double pres, const double *C, double *pres_mod)` emitted by
 * rate_subs.py:707-1073 (prototype rate_subs.py:274-277).  With NPD == 0
 * it validates its arguments and does nothing, as the reference emits no
 * such function then.
 */
int crs_get_rxn_pres_mod(const double *T, const double *P, const double *C,
                         double *pres_mod, int64_t n, int64_t ld,
                         void *stream);

/*
 * Batched eval_spec_rates: sp_rates[NSP-1][ld] and the bath gas' rate
 * dy_N[n].  Replaces `void eval_spec_rates(const double *fwd_rates,
 * const double *rev_rates, const double *pres_mod, double *sp_rates,
 * double *dy_N)` emitted by rate_subs.py:1076-1235 (prototype
 * rate_subs.py:268-269).  Passing dy_N = sp_rates + (NSP-1)*ld gives the
 * [NSP][ld] layout of crs_eval_wdot.  rev / pres_mod may be NULL when the
 * mechanism has no such reactions.
 */
int crs_eval_spec_rates(const double *fwd, const double *rev,
                        const double *pres_mod, double *sp_rates,
                        double *dy_N, int64_t n, int64_t ld, void *stream);

/*
 * Fused conc -> rates -> pres_mod -> species rates: wdot[NSP][ld], with
 * row NSP-1 the bath gas (the reference's dy_N).  The per-state call chain
 * it replaces is the reference's dydt body, rate_subs.py:1845-1880.
 * mode: CRS_MODE_FP64 or CRS_MODE_MIXED.
 */
int crs_eval_wdot(const double *T, const double *P, const double *Y,
                  double *wdot, int64_t n, int64_t ld, int mode,
                  void *stream);

/*
 * crs_eval_wdot on a subset: for t < m, state idx[t] of the [.][ld]
 * arrays is evaluated and written to column idx[t] of wdot; other columns
 * are untouched.  Used to re-evaluate gated states exactly.  Indices must
 * lie in [0, ld); idx is a device array; ld is at most 2^29 (the kernels
 * address a state by a 32-bit byte offset).  No counterpart in the
 * reference (its states are all evaluated every call).
 */
int crs_eval_wdot_subset(const double *T, const double *P, const double *Y,
                         double *wdot, const int32_t *idx, int64_t m,
                         int64_t ld, int mode, void *stream);

/*
 * crs_eval_wdot_subset with the subset's length in device memory: states
 * idx[t] for t < min(*count, m_max), where count is a device int32 written
 * by earlier work on the same stream (the surrogate's fallback list), so
 * no host code has to read it.  The launch covers m_max states; those
 * beyond *count leave at once.  The signature is crs_subset_fn of
 * include/crs_mlp.h: the gated surrogate calls it for its exact fallback.
 * No counterpart in the reference (its states are all evaluated every
 * call).
 */
int crs_eval_wdot_subset_dev(const double *T, const double *P,
                             const double *Y, double *wdot,
                             const int32_t *idx, const int32_t *count,
                             int64_t m_max, int64_t ld, int mode,
                             void *stream);

/*
 * Batched constant-pressure right-hand side.  Replaces `void dydt(const
 * double t, const double pres, const double *y, double *dy)` emitted by
 * rate_subs.py:1829-1946 (CONP branch; prototype rate_subs.py:1808-1816):
 * y and dy are [rows][ld] with row 0 = T [K] and rows 1..NSP-1 the mass
 * fractions of species 0..NSP-2 (the last species is implied); P[n] [Pa].
 * dy row 0 = dT/dt = -sum_k wdot_k h_k W_k / (rho cp), rows 1..NSP-1 =
 * dY_k/dt = wdot_k W_k / rho.  Any row beyond NSP-1 (an integrator's NN-th
 * slot) is neither read nor written, as in the reference.  The time
 * argument of the reference is unused there and omitted here.  Two kernels
 * per call: the fused wdot kernel of the mode writes dY_k/dt = wdot_k W_k /
 * rho into dy rows 1..NSP-1 (and the bath gas's wdot W / rho into row 0),
 * then an FP64 energy kernel replaces row 0 with dT/dt.  mode CRS_MODE_FP64, or CRS_MODE_MIXED where the library has
 * it (crs_get_mech_info modes; the rates as crs_eval_wdot computes them in
 * that mode, CRS_ERR_UNSUPPORTED otherwise).  dy therefore must
 * share no entry with y or P (they may share an allocation, e.g. side by
 * side within rows); an overlap returns CRS_ERR_ARG before any launch.
 */
int crs_dydt(const double *P, const double *y, double *dy, int64_t n,
             int64_t ld, int mode, void *stream);

/*
 * Device-resident batched integrator: nsteps classical RK4 steps of size h
 * (seconds) of the constant-pressure system crs_dydt defines, starting at
 * time t0, updating y [rows][ld] in place.  Replaces the time loop of the
 * reference's generated main (create_rate_subs.py:210-283: intDriver
 * per step, with the state copied host <-> device around every step and
 * the integrator left as an INTEGRATOR macro): here the state never leaves
 * the device and the step is RK4.  work: 3 * NSP rows of stride ld
 * (caller-owned scratch).  Optional ignition bookkeeping as the reference's
 * IGN block (:267-273): with T0[n] and t_ign[n] (initialise t_ign to NaN,
 * the "not ignited" marker), a state whose T first reaches T0 + dT_ign at
 * the end of a step gets t_ign = the time at the START of that step (the
 * reference sets t_ign = t at :271 before its `t += h` at :276); step k
 * starts at t0 plus k additions of h, as the reference accumulates t.
 * n is at most 2^29 states per call.  y, work, P, T0 and t_ign must share
 * no entry (CRS_ERR_ARG otherwise, before any launch).
 * Enqueues 8 kernels per step on the stream, per RK4 stage the fused wdot
 * kernel and one energy-and-stage kernel (capturable in a graph; no
 * memset, no allocation).
 */
int crs_rk4(const double *P, double *y, double *work, int64_t n, int64_t ld,
            double t0, double h, int nsteps, const double *T0, double *t_ign,
            double dT_ign, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* CRS_HIP_H */
