/*
 * C ABI of the wdot surrogate library (libcrs_mlp.so).
 *
 * Mechanism-independent kernels for a small MLP that maps per-state
 * features (ln T, ln P, ln Y_k) to arcsinh(wdot_k / s_k): BF16 MFMA matrix
 * products with FP32 accumulation, GELU, MSE gradient, Adam with FP32
 * master weights, feature/target/decode transforms and the fallback gate.
 * No counterpart in the reference (its rates are always evaluated exactly);
 * callers are the surrogate training/inference driver
 * (create_rate_subs_amd/surrogate.py).
 *
 * Conventions: activations are state-major [rows][ld] BF16 (raw 16-bit
 * bf16 in `void *`), weights [out][in] BF16 with in/out padded to multiples
 * of 64 (padding zero); the reduction length of every product is a
 * multiple of 32.  All pointers are device pointers; work is enqueued on
 * `stream` (a hipStream_t, NULL = default).  Return 0 or a CRS_ERR_* code;
 * crs_mlp_last_error() gives the message.
 */
#ifndef CRS_MLP_H
#define CRS_MLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct crs_mlp_gate_params {
    double t_lo, t_hi; /* training box in T [K] */
    double p_lo, p_hi; /* training box in P [Pa] */
    double rmax;       /* max element / mass balance residual, relative;
                          < 0: the gate admits no state */
} crs_mlp_gate_params;

const char *crs_mlp_last_error(void);

/* Z = X W^T + b; act != 0: A_out = gelu(Z) (bf16), Z_out (bf16, optional),
 * and (optional) AT_out = A_out transposed, [N][ldat] with the states
 * contiguous and zero padding (ldat >= M rounded up to 32); act == 0:
 * Y_out = Z (fp32) for columns < n_valid.  X [M][ldx], W [N][ldw],
 * outputs [M][ldo].  N % 64 == 0, K % 32 == 0. */
int crs_mlp_linear_fwd(const void *X, int64_t ldx, const void *W,
                       int64_t ldw, const float *bias, int64_t M, int N,
                       int K, int act, void *A_out, void *Z_out, float *Y_out,
                       int64_t ldo, int n_valid, void *AT_out, int64_t ldat,
                       void *stream);

/* dZprev = (dZ W) * gelu'(Zprev): dZ [M][ldz] (N wide), W given
 * transposed, WT [K][ldw] (N wide; crs_mlp_transpose_bf16 of W [N][K]),
 * Zprev and dZprev [M][ldo] (K wide, bf16); (optional) dZprevT = dZprev
 * transposed, [K][ldt] as AT_out of crs_mlp_linear_fwd. */
int crs_mlp_linear_bwd_data(const void *dZ, int64_t ldz, const void *WT,
                            int64_t ldw, int64_t M, int N, int K,
                            const void *Zprev, void *dZprev, int64_t ldo,
                            void *dZprevT, int64_t ldt, void *stream);

/* dW[N][K] += dZ^T X and db[N] += colsum(dZ), from the transposed operands
 * dZT [N][ldz] and XT [K][ldx] (states contiguous, as crs_mlp_transpose_bf16
 * writes them; ldz and ldx at least M rounded up to 32, the padding zero).
 * FP32 atomics into dW and db: zero them first. */
int crs_mlp_linear_bwd_weight(const void *dZT, int64_t ldz, const void *XT,
                              int64_t ldx, int64_t M, int N, int K, float *dW,
                              float *db, void *stream);

/* dst[c][r] = src[r][c] for r < rows, c < cols (bf16); dst [cols][ld_dst]
 * with its columns rows..ld_dst-1 set to zero. */
int crs_mlp_transpose_bf16(const void *src, int64_t ld_src, int64_t rows,
                           int cols, void *dst, int64_t ld_dst, void *stream);

/* X[t][0..kp) = normalised [ln T, ln P, ln max(Y_k, 1e-30) for k < nsp]
 * of state idx[t] (or t when idx is NULL), the last species taken as
 * 1 - sum(others). */
int crs_mlp_features(const double *T, const double *P, const double *Y,
                     const int32_t *idx, int64_t n, int64_t ld, int nsp,
                     const float *mu, const float *sd, void *X, int kp,
                     void *stream);

/* Yt[t][k] = (asinh(W[k][t] / scale[k]) - mu[k]) / sd[k]. */
int crs_mlp_targets(const double *W, int64_t n, int64_t ld, int nsp,
                    const double *scale, const float *mu, const float *sd,
                    float *Yt, int ldt, void *stream);

/* W[k][idx[t]] = scale[k] sinh(mu[k] + sd[k] Yh[t][k]) (FP64 rows). */
int crs_mlp_decode(const float *Yh, int ldy, const int32_t *idx, int64_t n,
                   int64_t ld, int nsp, const double *scale, const float *mu,
                   const float *sd, double *W, void *stream);

/* dY = 2 (Yh - Yt) / (n nv) as bf16 [n][ldd] (zero beyond nv); adds the
 * sum of squared errors to *loss when loss is not NULL. */
int crs_mlp_mse_grad(const float *Yh, const float *Yt, int64_t n, int nv,
                     int ldy, int ldt, void *dY, int ldd, float *loss,
                     void *stream);

/* One Adam step (bias-corrected, decoupled weight decay wd) on n FP32
 * parameters; refreshes the BF16 copy pb. */
int crs_mlp_adam(float *p, const float *g, float *m1, float *m2, void *pb,
                 int64_t n, float lr, float b1, float b2, float eps, int step,
                 float wd, void *stream);

/* b = bf16(a). */
int crs_mlp_to_bf16(const float *a, void *b, int64_t n, void *stream);

/* Per-state fallback decision; flagged state indices appended to fb_idx,
 * their count added to *fb_count (zero it first).  elem [n_elem][nsp]
 * element counts, mw [nsp] molecular weights. */
int crs_mlp_gate(const double *T, const double *P, const double *Y,
                 const double *W, int64_t n, int64_t ld, int nsp,
                 crs_mlp_gate_params gp, const double *elem, int n_elem,
                 const double *mw, int32_t *flags, int32_t *fb_idx,
                 int32_t *fb_count, void *stream);

/* Rearrange one bf16 weight matrix W [n_out][n_in] (row major) into the
 * fused inference kernel's MFMA tile order: tiles of 16 rows x 32 inputs,
 * row block major, each as 64 lanes x 8 bf16.  perm != 0 for every layer
 * fed by a hidden layer (its input index follows the MFMA output layout).
 * n_out % 16 == 0, n_in % 32 == 0. */
int crs_mlp_pack_tiles(const void *W, int n_out, int n_in, int perm,
                       void *out, void *stream);

/* Fused surrogate evaluation of n states: features, the 64 -> 3 x 256 (GELU)
 * -> 64 network from the packed tiles (layers back to back) and biases
 * (fp32, 256 + 256 + 256 + 64, 16-byte aligned), decode
 * wdot_k = scale_k sinh(ymu_k + ysd_k y_k) into W [nsp][ld], and the gate of
 * crs_mlp_gate (flags, fallback list, count added to *fb_count).  Only the
 * 64/256/3/64 shape is built (CRS_ERR_UNSUPPORTED otherwise); nsp <= 62,
 * n_elem <= 8. */
int crs_mlp_infer(const double *T, const double *P, const double *Y,
                  int64_t n, int64_t ld, int nsp, int kin, int hidden,
                  int layers, int kout, const void *tiles, const float *bias,
                  const float *mu, const float *sd, const double *scale,
                  const float *ymu, const float *ysd, double *W,
                  crs_mlp_gate_params gp, const double *elem, int n_elem,
                  const double *mw, int32_t *flags, int32_t *fb_idx,
                  int32_t *fb_count, void *stream);

/* The exact evaluation a gated call falls back to: crs_eval_wdot_subset_dev
 * of a mechanism library (include/crs_hip.h), which evaluates states idx[t]
 * for t < min(*count, m_max) and reads the count on the device. */
typedef int (*crs_subset_fn)(const double *T, const double *P,
                             const double *Y, double *wdot,
                             const int32_t *idx, const int32_t *count,
                             int64_t m_max, int64_t ld, int mode,
                             void *stream);

/* A trained network for the gated entry points: the arguments of
 * crs_mlp_infer after the state arrays, gathered (device pointers; the
 * struct itself is read on the host during the call only). */
typedef struct crs_mlp_net {
    int nsp, kin, hidden, layers, kout;
    const void *tiles;   /* crs_mlp_pack_tiles order, 16-byte aligned */
    const float *bias;   /* 256 + 256 + 256 + 64, 16-byte aligned */
    const float *mu, *sd;            /* feature normalisation [nsp + 2] */
    const double *scale;             /* s_k [nsp] */
    const float *ymu, *ysd;          /* output normalisation [nsp] */
    crs_mlp_gate_params gate;
    const double *elem;              /* [n_elem][nsp] */
    int n_elem;
    const double *mw;                /* [nsp] */
} crs_mlp_net;

/* Gated surrogate evaluation of n states into W [nsp][ld]: fused inference
 * and gate (crs_mlp_infer), then `exact` over the states the gate sent back,
 * all enqueued on `stream` with no host synchronisation (the fallback count
 * stays on the device).  W then holds the surrogate's rates for admitted
 * states and the exact kernel's for every other state (bit-identical to
 * crs_eval_wdot in mode exact_mode).  Caller-owned workspace: fb_idx [n]
 * (the fallback states, in no particular order) and *fb_count (their
 * number, zeroed by this call); flags [n] (1 = sent back) may be NULL.
 * A gate with rmax < 0 admits no state (what calibration sets when no
 * residual threshold meets the error target): the network is then not
 * evaluated, fb_idx gets 0..n-1, *fb_count n, flags all 1, and W is the
 * exact kernel's throughout.
 * This is the reference's eval_wdot chain (rate_subs.py:1845-1880) with a
 * learned evaluation in front of it; the reference has no surrogate. */
int crs_mlp_eval_gated(const crs_mlp_net *net, const double *T,
                       const double *P, const double *Y, double *W, int64_t n,
                       int64_t ld, crs_subset_fn exact, int exact_mode,
                       int32_t *flags, int32_t *fb_idx, int32_t *fb_count,
                       void *stream);

/* crs_mlp_eval_gated without caller workspace and without allocating:
 * the fallback list and its count live in a static device array of this
 * library (16.8 MB per GPU, held until the process ends), so the call makes
 * no hipMalloc/hipFree and no allocating library call; n is at most
 * 4194304 (2^22) states, larger n is refused with CRS_ERR_ARG before any
 * work is enqueued.  No flags are written.  Every call on one GPU shares
 * that GPU's list (the count is cleared by each call): calls that overlap
 * on one GPU (other streams or threads) race.  No counterpart in the
 * reference; for callers that cannot allocate (code under graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). */
int crs_mlp_eval_gated_noalloc(const crs_mlp_net *net, const double *T,
                               const double *P, const double *Y, double *W,
                               int64_t n, int64_t ld, crs_subset_fn exact,
                               int exact_mode, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* CRS_MLP_H */
