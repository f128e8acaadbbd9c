// Learned wdot surrogate: MLP kernels for MI355X (gfx950).
//
// A small MLP maps per-state features (ln T, ln P, ln Y_k) to arcsinh-scaled
// species rates.  Every matrix product runs on BF16 MFMA with FP32
// accumulation (v_mfma_f32_16x16x32_bf16); activations between layers are
// BF16, master weights and optimiser state FP32.  The library is independent
// of the mechanism; the gate uses the mechanism's element matrix and molecular
// weights passed in as data.
//
// Layout: row-major [rows][ld] with the state index as the row ("state
// major") for activations, [out][in] for weights.  Padded dimensions are
// multiples of 64 and padding columns hold zeros.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "crs_common.h"
#include "crs_mlp.h"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

constexpr int TM = 128;   // block tile rows (states)
constexpr int TN = 64;    // block tile columns
constexpr int TR = 32;    // reduction step (one MFMA K)
constexpr int LDS_PAD = 8;
constexpr int LROW = TR + LDS_PAD;   // LDS row length in bf16 (80 bytes)

__device__ __forceinline__ float bf2f(__bf16 x) { return (float)x; }
__device__ __forceinline__ __bf16 f2bf(float x) { return (__bf16)x; }

// GELU, tanh form: 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3))).
__device__ __forceinline__ float gelu_t(float z, float *dgelu)
{
    const float c = 0.7978845608028654f;
    const float u = c * (z + 0.044715f * z * z * z);
    // tanh(u) = 1 - 2 / (exp(2u) + 1), exp on the hardware exp2
    const float e = __builtin_amdgcn_exp2f(2.0f * u * 1.4426950408889634f);
    const float t = 1.0f - 2.0f / (e + 1.0f);
    if (dgelu)
        *dgelu = 0.5f * (1.0f + t) +
                 0.5f * z * (1.0f - t * t) * c * (1.0f + 3.0f * 0.044715f * z * z);
    return 0.5f * z * (1.0f + t);
}

enum Epi {
    EPI_BIAS_GELU = 0,   // Z = acc + b (fp32, optional), A = gelu(Z) (bf16)
    EPI_BIAS_F32 = 1,    // out = acc + b (fp32), columns < n_valid
    EPI_DGELU_BF16 = 2,  // out = acc * gelu'(Zprev) (bf16)
    EPI_ATOMIC_F32 = 3,  // out += acc (fp32 atomics; split reduction)
};

// C[m][n] = sum_r A(m, r) B(n, r)
//   A(m, r) = A[m * lda + r]   if A_RC (reduction contiguous) else A[r * lda + m]
//   B(n, r) = B[n * ldb + r]   if B_RC                        else B[r * ldb + n]
// Rows m < M are valid (others read as zero / not stored); N and R are
// padded to multiples of TN and TR.  The reduction range of this block is
// [r0, r1) (split reduction for weight gradients).
template <bool A_RC, bool B_RC, int EPI>
__global__ __launch_bounds__(256) void k_gemm(
    const __bf16 *__restrict__ A, int64_t lda, const __bf16 *__restrict__ B,
    int64_t ldb, int64_t M, int N, int64_t R, int64_t r_chunk,
    const float *__restrict__ bias, float *__restrict__ out_f32, int64_t ldo,
    __bf16 *__restrict__ out_bf16, float *__restrict__ zsave,
    const float *__restrict__ zprev, int n_valid)
{
    __shared__ __bf16 As[TM * LROW];
    __shared__ __bf16 Bs[TN * LROW];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int64_t m0 = (int64_t)blockIdx.x * TM;
    const int n0 = blockIdx.y * TN;
    const int64_t r0 = (int64_t)blockIdx.z * r_chunk;
    const int64_t r1 = (r0 + r_chunk < R) ? r0 + r_chunk : R;

    floatx4 acc[2][4];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 4; ++j)
            acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    for (int64_t rk = r0; rk < r1; rk += TR) {
        // ---- stage A tile [TM][TR] ----
        if (A_RC) {
            // 128 rows x 32 bf16 = 512 16-byte chunks, 2 per thread
            for (int c = tid; c < TM * TR / 8; c += 256) {
                const int row = c >> 2, seg = (c & 3) * 8;
                const int64_t m = m0 + row;
                bf16x8 v = {};
                if (m < M && rk + seg < r1)
                    v = *(const bf16x8 *)(A + m * lda + rk + seg);
                *(bf16x8 *)(As + row * LROW + seg) = v;
            }
        } else {
            // reduction-major: contiguous along m; 32 r x 128 m
            for (int c = tid; c < TM * TR / 8; c += 256) {
                const int r = c >> 4, mseg = (c & 15) * 8;
                const int64_t rr = rk + r;
                bf16x8 v = {};
                if (rr < r1) {
                    if (m0 + mseg + 8 <= M) {
                        v = *(const bf16x8 *)(A + rr * lda + m0 + mseg);
                    } else {
                        for (int e = 0; e < 8; ++e)
                            if (m0 + mseg + e < M)
                                v[e] = A[rr * lda + m0 + mseg + e];
                    }
                }
                for (int e = 0; e < 8; ++e)
                    As[(mseg + e) * LROW + r] = v[e];
            }
        }
        // ---- stage B tile [TN][TR] ----
        if (B_RC) {
            for (int c = tid; c < TN * TR / 8; c += 256) {
                const int col = c >> 2, seg = (c & 3) * 8;
                bf16x8 v = {};
                if (rk + seg < r1)
                    v = *(const bf16x8 *)(B + (int64_t)(n0 + col) * ldb + rk + seg);
                *(bf16x8 *)(Bs + col * LROW + seg) = v;
            }
        } else {
            for (int c = tid; c < TN * TR / 8; c += 256) {
                const int r = c >> 3, nseg = (c & 7) * 8;
                bf16x8 v = {};
                if (rk + r < r1)
                    v = *(const bf16x8 *)(B + (rk + r) * ldb + n0 + nseg);
                for (int e = 0; e < 8; ++e)
                    Bs[(nseg + e) * LROW + r] = v[e];
            }
        }
        __syncthreads();
        // ---- MFMA: wave owns rows [32 wave, 32 wave + 32) x all 64 cols ----
        bf16x8 af[2], bfr[4];
        for (int i = 0; i < 2; ++i)
            af[i] = *(const bf16x8 *)(As + (wave * 32 + i * 16 + (lane & 15)) * LROW +
                                      (lane >> 4) * 8);
        for (int j = 0; j < 4; ++j)
            bfr[j] = *(const bf16x8 *)(Bs + (j * 16 + (lane & 15)) * LROW +
                                       (lane >> 4) * 8);
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 4; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    af[i], bfr[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }

    // ---- epilogue: lane holds rows (lane>>4)*4 + e, column lane & 15 ----
    for (int i = 0; i < 2; ++i) {
        for (int j = 0; j < 4; ++j) {
            const int col = n0 + j * 16 + (lane & 15);
            for (int e = 0; e < 4; ++e) {
                const int64_t m = m0 + wave * 32 + i * 16 + (lane >> 4) * 4 + e;
                if (m >= M)
                    continue;
                const float a = acc[i][j][e];
                if (EPI == EPI_BIAS_GELU) {
                    const float z = a + bias[col];
                    if (zsave)
                        zsave[m * ldo + col] = z;
                    out_bf16[m * ldo + col] = f2bf(gelu_t(z, nullptr));
                } else if (EPI == EPI_BIAS_F32) {
                    if (col < n_valid)
                        out_f32[m * ldo + col] = a + bias[col];
                } else if (EPI == EPI_DGELU_BF16) {
                    float d;
                    gelu_t(zprev[m * ldo + col], &d);
                    out_bf16[m * ldo + col] = f2bf(a * d);
                } else {
                    atomicAdd(&out_f32[m * ldo + col], a);
                }
            }
        }
    }
}

// Features: x = [(ln T - mu0)/sd0, (ln P - mu1)/sd1, (ln max(Y_k,1e-30) - mu)/sd]
// for k < nsp (the last species is 1 - sum of the others), padded with zeros.
__global__ __launch_bounds__(256) void k_features(
    const double *__restrict__ T, const double *__restrict__ P,
    const double *__restrict__ Y, const int32_t *__restrict__ idx, int64_t n,
    int64_t ld, int nsp, const float *__restrict__ mu,
    const float *__restrict__ sd, __bf16 *__restrict__ X, int kp)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    const int64_t i = idx ? (int64_t)idx[t] : t;
    __bf16 *x = X + t * kp;
    x[0] = f2bf(((float)log(T[i]) - mu[0]) / sd[0]);
    x[1] = f2bf(((float)log(P[i]) - mu[1]) / sd[1]);
    double s = 0.0;
    for (int k = 0; k < nsp - 1; ++k) {
        const double y = Y[(int64_t)k * ld + i];
        s += y;
        const float ly = (float)log(y > 1e-30 ? y : 1e-30);
        x[2 + k] = f2bf((ly - mu[2 + k]) / sd[2 + k]);
    }
    const double yN = 1.0 - s;
    const float lyN = (float)log(yN > 1e-30 ? yN : 1e-30);
    x[1 + nsp] = f2bf((lyN - mu[1 + nsp]) / sd[1 + nsp]);
    for (int c = 2 + nsp; c < kp; ++c)
        x[c] = f2bf(0.0f);
}

// Targets: y[m][k] = (arcsinh(w_k / s_k) - mu_k) / sd_k, fp32, state major.
__global__ __launch_bounds__(256) void k_targets(
    const double *__restrict__ W, int64_t n, int64_t ld, int nsp,
    const double *__restrict__ scale, const float *__restrict__ mu,
    const float *__restrict__ sd, float *__restrict__ Yt, int ldt)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    for (int k = 0; k < nsp; ++k)
        Yt[t * ldt + k] =
            ((float)asinh(W[(int64_t)k * ld + t] / scale[k]) - mu[k]) / sd[k];
}

// Decode: wdot[k][i] = s_k sinh(mu_k + sd_k yhat[t][k]) for state
// i = idx[t] (or t).
__global__ __launch_bounds__(256) void k_decode(
    const float *__restrict__ Yh, int ldy, const int32_t *__restrict__ idx,
    int64_t n, int64_t ld, int nsp, const double *__restrict__ scale,
    const float *__restrict__ mu, const float *__restrict__ sd,
    double *__restrict__ W)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    const int64_t i = idx ? (int64_t)idx[t] : t;
    for (int k = 0; k < nsp; ++k)
        W[(int64_t)k * ld + i] =
            scale[k] * sinh((double)mu[k] + (double)sd[k] * (double)Yh[t * ldy + k]);
}

// MSE gradient in arcsinh space: dY = 2 (yhat - y) / (n * nv), bf16, padded
// columns zero; loss sum accumulated into *loss.
__global__ __launch_bounds__(256) void k_mse_grad(
    const float *__restrict__ Yh, const float *__restrict__ Yt, int64_t n,
    int nv, int ldy, int ldt, float scale, __bf16 *__restrict__ dY, int ldd,
    float *__restrict__ loss)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float part = 0.f;
    if (t < n) {
        for (int k = 0; k < ldd; ++k) {
            float g = 0.f;
            if (k < nv) {
                const float d = Yh[t * ldy + k] - Yt[t * ldt + k];
                part += d * d;
                g = 2.f * d * scale;
            }
            dY[t * ldd + k] = f2bf(g);
        }
    }
    // wave reduction then one atomic per wave
    for (int off = 32; off > 0; off >>= 1)
        part += __shfl_down(part, off, 64);
    if ((threadIdx.x & 63) == 0 && loss)
        atomicAdd(loss, part);
}

// Column sums (bias gradient): db[c] += sum_m dZ[m][c], one thread per column
// per chunk of rows.
__global__ __launch_bounds__(256) void k_colsum(
    const __bf16 *__restrict__ dZ, int64_t M, int N, int64_t ld,
    int64_t rows_per_block, float *__restrict__ db)
{
    const int c = threadIdx.x;
    if (c >= N)
        return;
    const int64_t a = (int64_t)blockIdx.x * rows_per_block;
    const int64_t b = (a + rows_per_block < M) ? a + rows_per_block : M;
    float s = 0.f;
    for (int64_t m = a; m < b; ++m)
        s += bf2f(dZ[m * ld + c]);
    atomicAdd(&db[c], s);
}

// Adam with FP32 master weights; writes the BF16 working copy.
__global__ __launch_bounds__(256) void k_adam(
    float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m1,
    float *__restrict__ m2, __bf16 *__restrict__ pb, int64_t n, float lr,
    float b1, float b2, float eps, float c1, float c2, float wd)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float gi = g[i] + wd * p[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float v = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = v;
    const float upd = lr * (a / c1) / (sqrtf(v / c2) + eps);
    p[i] -= upd;
    pb[i] = f2bf(p[i]);
}

__global__ __launch_bounds__(256) void k_to_bf16(const float *__restrict__ a,
                                                 __bf16 *__restrict__ b,
                                                 int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        b[i] = f2bf(a[i]);
}

// Gate: a state falls back to the exact kernel when it is outside the
// training box, has an invalid composition, a non-finite prediction, or a
// prediction whose element / mass balance residual exceeds rmax.  Flagged
// state indices are appended to fb_idx (order is arbitrary); flags[i] = 1.
__global__ __launch_bounds__(256) void k_gate(
    const double *__restrict__ T, const double *__restrict__ P,
    const double *__restrict__ Y, const double *__restrict__ W, int64_t n,
    int64_t ld, int nsp, const crs_mlp_gate_params gp,
    const double *__restrict__ elem, int n_elem, const double *__restrict__ mw,
    int32_t *__restrict__ flags, int32_t *__restrict__ fb_idx,
    int32_t *__restrict__ fb_count)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    int bad = 0;
    const double t = T[i], p = P[i];
    if (!(t >= gp.t_lo && t <= gp.t_hi && p >= gp.p_lo && p <= gp.p_hi))
        bad = 1;
    double s = 0.0;
    for (int k = 0; k < nsp - 1; ++k) {
        const double y = Y[(int64_t)k * ld + i];
        if (!(y >= 0.0 && y <= 1.0))
            bad = 1;
        s += y;
    }
    if (!(s <= 1.0 + 1e-12))
        bad = 1;
    double mres = 0.0, mscale = 0.0;
    for (int k = 0; k < nsp; ++k) {
        const double w = W[(int64_t)k * ld + i];
        if (!isfinite(w))
            bad = 1;
        mres += mw[k] * w;
        mscale += mw[k] * fabs(w);
    }
    if (mscale > 0.0 && fabs(mres) > gp.rmax * mscale)
        bad = 1;
    for (int e = 0; e < n_elem && !bad; ++e) {
        double r = 0.0, sc = 0.0;
        for (int k = 0; k < nsp; ++k) {
            const double a = elem[e * nsp + k];
            if (a != 0.0) {
                const double w = W[(int64_t)k * ld + i];
                r += a * w;
                sc += a * fabs(w);
            }
        }
        if (sc > 0.0 && fabs(r) > gp.rmax * sc)
            bad = 1;
    }
    flags[i] = bad;
    if (bad) {
        const int32_t slot = atomicAdd(fb_count, 1);
        fb_idx[slot] = (int32_t)i;
    }
}

int check_gemm(const char *fn, int64_t M, int N, int64_t R)
{
    if (M < 0 || N <= 0 || R <= 0 || N % TN != 0 || R % TR != 0)
        return crs::fail(CRS_ERR_ARG,
                         "%s: bad sizes M=%lld N=%d R=%lld (N %% %d, R %% %d)",
                         fn, (long long)M, N, (long long)R, TN, TR);
    return CRS_OK;
}

}  // namespace

extern "C" {

const char *crs_mlp_last_error(void) { return crs::error_buffer(); }

int crs_mlp_linear_fwd(const void *X, int64_t ldx, const void *Wb, int64_t ldw,
                       const float *bias, int64_t M, int N, int K, int act,
                       void *A_out, float *Z_out, float *Y_out, int64_t ldo,
                       int n_valid, void *stream)
{
    const char *fn = "crs_mlp_linear_fwd";
    CRS_TRY(check_gemm(fn, M, N, K));
    if (M == 0)
        return CRS_OK;
    if (!X || !Wb || !bias)
        return crs::fail(CRS_ERR_ARG, "%s: NULL input", fn);
    dim3 grid((unsigned)((M + TM - 1) / TM), N / TN, 1);
    hipStream_t s = (hipStream_t)stream;
    if (act) {
        if (!A_out)
            return crs::fail(CRS_ERR_ARG, "%s: NULL activation output", fn);
        k_gemm<true, true, EPI_BIAS_GELU><<<grid, 256, 0, s>>>(
            (const __bf16 *)X, ldx, (const __bf16 *)Wb, ldw, M, N, K, K, bias,
            nullptr, ldo, (__bf16 *)A_out, Z_out, nullptr, N);
    } else {
        if (!Y_out)
            return crs::fail(CRS_ERR_ARG, "%s: NULL output", fn);
        k_gemm<true, true, EPI_BIAS_F32><<<grid, 256, 0, s>>>(
            (const __bf16 *)X, ldx, (const __bf16 *)Wb, ldw, M, N, K, K, bias,
            Y_out, ldo, nullptr, nullptr, nullptr, n_valid);
    }
    return crs::launched(fn);
}

int crs_mlp_linear_bwd_data(const void *dZ, int64_t ldz, const void *Wb,
                            int64_t ldw, int64_t M, int N, int K,
                            const float *Zprev, void *dZprev, int64_t ldo,
                            void *stream)
{
    const char *fn = "crs_mlp_linear_bwd_data";
    // dX[M][K] = dZ[M][N] W[N][K]: reduction over N; columns K
    CRS_TRY(check_gemm(fn, M, K, N));
    if (M == 0)
        return CRS_OK;
    if (!dZ || !Wb || !Zprev || !dZprev)
        return crs::fail(CRS_ERR_ARG, "%s: NULL pointer", fn);
    dim3 grid((unsigned)((M + TM - 1) / TM), K / TN, 1);
    k_gemm<true, false, EPI_DGELU_BF16><<<grid, 256, 0, (hipStream_t)stream>>>(
        (const __bf16 *)dZ, ldz, (const __bf16 *)Wb, ldw, M, K, N, N, nullptr,
        nullptr, ldo, (__bf16 *)dZprev, nullptr, Zprev, K);
    return crs::launched(fn);
}

int crs_mlp_linear_bwd_weight(const void *dZ, int64_t ldz, const void *X,
                              int64_t ldx, int64_t M, int N, int K, float *dW,
                              float *db, void *stream)
{
    const char *fn = "crs_mlp_linear_bwd_weight";
    // dW[N][K] = sum_m dZ[m][n] X[m][k]: rows N, columns K, reduction M
    if (M <= 0 || N <= 0 || K <= 0 || N % TN != 0 || K % TN != 0)
        return crs::fail(CRS_ERR_ARG, "%s: bad sizes", fn);
    if (!dZ || !X || !dW || !db)
        return crs::fail(CRS_ERR_ARG, "%s: NULL pointer", fn);
    const int64_t Mr = ((M + TR - 1) / TR) * TR;
    // split the reduction so the grid covers the GPU
    int64_t chunk = 4096;
    const int64_t nz = (Mr + chunk - 1) / chunk;
    dim3 grid((unsigned)((N + TM - 1) / TM), K / TN, (unsigned)nz);
    hipStream_t s = (hipStream_t)stream;
    // rows of the output are n: A(n, m) = dZ[m * ldz + n] (reduction-major)
    k_gemm<false, false, EPI_ATOMIC_F32><<<grid, 256, 0, s>>>(
        (const __bf16 *)dZ, ldz, (const __bf16 *)X, ldx, N, K, M, chunk,
        nullptr, dW, K, nullptr, nullptr, nullptr, K);
    CRS_TRY(crs::launched(fn));
    const int64_t rows = 512;
    k_colsum<<<(unsigned)((M + rows - 1) / rows), 256, 0, s>>>(
        (const __bf16 *)dZ, M, N, ldz, rows, db);
    return crs::launched(fn);
}

int crs_mlp_features(const double *T, const double *P, const double *Y,
                     const int32_t *idx, int64_t n, int64_t ld, int nsp,
                     const float *mu, const float *sd, void *X, int kp,
                     void *stream)
{
    const char *fn = "crs_mlp_features";
    if (n < 0 || ld < 0 || nsp < 2 || kp < nsp + 2)
        return crs::fail(CRS_ERR_ARG, "%s: bad sizes", fn);
    if (n == 0)
        return CRS_OK;
    if (!T || !P || !Y || !mu || !sd || !X)
        return crs::fail(CRS_ERR_ARG, "%s: NULL pointer", fn);
    k_features<<<crs::grid_for(n), crs::kBlock, 0, (hipStream_t)stream>>>(
        T, P, Y, idx, n, ld, nsp, mu, sd, (__bf16 *)X, kp);
    return crs::launched(fn);
}

int crs_mlp_targets(const double *W, int64_t n, int64_t ld, int nsp,
                    const double *scale, const float *mu, const float *sd,
                    float *Yt, int ldt, void *stream)
{
    const char *fn = "crs_mlp_targets";
    if (n < 0 || ldt < nsp)
        return crs::fail(CRS_ERR_ARG, "%s: bad sizes", fn);
    if (n == 0)
        return CRS_OK;
    k_targets<<<crs::grid_for(n), crs::kBlock, 0, (hipStream_t)stream>>>(
        W, n, ld, nsp, scale, mu, sd, Yt, ldt);
    return crs::launched(fn);
}

int crs_mlp_decode(const float *Yh, int ldy, const int32_t *idx, int64_t n,
                   int64_t ld, int nsp, const double *scale, const float *mu,
                   const float *sd, double *W, void *stream)
{
    const char *fn = "crs_mlp_decode";
    if (n < 0 || ldy < nsp)
        return crs::fail(CRS_ERR_ARG, "%s: bad sizes", fn);
    if (n == 0)
        return CRS_OK;
    k_decode<<<crs::grid_for(n), crs::kBlock, 0, (hipStream_t)stream>>>(
        Yh, ldy, idx, n, ld, nsp, scale, mu, sd, W);
    return crs::launched(fn);
}

int crs_mlp_mse_grad(const float *Yh, const float *Yt, int64_t n, int nv,
                     int ldy, int ldt, void *dY, int ldd, float *loss,
                     void *stream)
{
    const char *fn = "crs_mlp_mse_grad";
    if (n <= 0 || nv <= 0 || ldd < nv)
        return crs::fail(CRS_ERR_ARG, "%s: bad sizes", fn);
    const float scale = 1.0f / ((float)n * (float)nv);
    k_mse_grad<<<crs::grid_for(n), crs::kBlock, 0, (hipStream_t)stream>>>(
        Yh, Yt, n, nv, ldy, ldt, scale, (__bf16 *)dY, ldd, loss);
    return crs::launched(fn);
}

int crs_mlp_adam(float *p, const float *g, float *m1, float *m2, void *pb,
                 int64_t n, float lr, float b1, float b2, float eps, int step,
                 float wd, void *stream)
{
    const char *fn = "crs_mlp_adam";
    if (n <= 0 || step <= 0)
        return crs::fail(CRS_ERR_ARG, "%s: bad sizes", fn);
    const float c1 = 1.f - powf(b1, (float)step);
    const float c2 = 1.f - powf(b2, (float)step);
    k_adam<<<crs::grid_for(n), crs::kBlock, 0, (hipStream_t)stream>>>(
        p, g, m1, m2, (__bf16 *)pb, n, lr, b1, b2, eps, c1, c2, wd);
    return crs::launched(fn);
}

int crs_mlp_to_bf16(const float *a, void *b, int64_t n, void *stream)
{
    if (n <= 0)
        return CRS_OK;
    k_to_bf16<<<crs::grid_for(n), crs::kBlock, 0, (hipStream_t)stream>>>(
        a, (__bf16 *)b, n);
    return crs::launched("crs_mlp_to_bf16");
}

int crs_mlp_gate(const double *T, const double *P, const double *Y,
                 const double *W, int64_t n, int64_t ld, int nsp,
                 crs_mlp_gate_params gp, const double *elem,
                 int n_elem, const double *mw, int32_t *flags,
                 int32_t *fb_idx, int32_t *fb_count, void *stream)
{
    const char *fn = "crs_mlp_gate";
    if (n < 0 || nsp < 2 || n_elem < 0)
        return crs::fail(CRS_ERR_ARG, "%s: bad sizes", fn);
    if (n == 0)
        return CRS_OK;
    if (!T || !P || !Y || !W || !mw || !flags || !fb_idx || !fb_count ||
        (n_elem > 0 && !elem))
        return crs::fail(CRS_ERR_ARG, "%s: NULL pointer", fn);
    k_gate<<<crs::grid_for(n), crs::kBlock, 0, (hipStream_t)stream>>>(
        T, P, Y, W, n, ld, nsp, gp, elem, n_elem, mw, flags, fb_idx, fb_count);
    return crs::launched(fn);
}

}  // extern "C"
