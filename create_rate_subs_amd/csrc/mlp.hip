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

// GELU, tanh form: 0.5 z (1 + tanh(u)), u = sqrt(2/pi) (z + 0.044715 z^3),
// evaluated as z sigma(2u) = z / (1 + 2^(z (k0 + k1 z^2))) with
// k0 = -2 sqrt(2/pi) log2(e), k1 = 0.044715 k0: one hardware exp2 and one
// hardware reciprocal.  With s = sigma(2u), tanh(u) = 2s - 1 and
// gelu'(z) = s + 2 z s (1 - s) sqrt(2/pi) (1 + 3 * 0.044715 z^2).
__device__ __forceinline__ float gelu_t(float z, float *dgelu)
{
    const float k0 = -2.3022082f, k1 = -0.10294324f;
    const float z2 = z * z;
    const float e = __builtin_amdgcn_exp2f(z * __builtin_fmaf(k1, z2, k0));
    const float sg = __builtin_amdgcn_rcpf(1.0f + e);
    if (dgelu)
        *dgelu = sg + 2.0f * z * sg * (1.0f - sg) * 0.7978845608028654f *
                          (1.0f + 0.134145f * z2);
    return z * sg;
}

enum Epi {
    EPI_BIAS_GELU = 0,   // Z = acc + b (bf16, optional), A = gelu(Z) (bf16)
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
    __bf16 *__restrict__ out_bf16, __bf16 *__restrict__ zsave,
    const __bf16 *__restrict__ zprev, int n_valid,
    __bf16 *__restrict__ out_t = nullptr, int64_t ldt = 0)
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
            // bf16 outputs also go, when asked, to the transposed array
            // out_t[col][m] (states contiguous: the weight gradient's
            // operand), four consecutive states in one 8-byte store, zero
            // for the padding rows M <= m < ldt
            __bf16 tv[4];
            const int64_t mb = m0 + wave * 32 + i * 16 + (lane >> 4) * 4;
            for (int e = 0; e < 4; ++e) {
                const int64_t m = m0 + wave * 32 + i * 16 + (lane >> 4) * 4 + e;
                if (m >= M)
                    continue;
                const float a = acc[i][j][e];
                if (EPI == EPI_BIAS_GELU) {
                    const float z = a + bias[col];
                    if (zsave)
                        zsave[m * ldo + col] = f2bf(z);
                    tv[e] = f2bf(gelu_t(z, nullptr));
                    out_bf16[m * ldo + col] = tv[e];
                } else if (EPI == EPI_BIAS_F32) {
                    if (col < n_valid)
                        out_f32[m * ldo + col] = a + bias[col];
                } else if (EPI == EPI_DGELU_BF16) {
                    float d;
                    gelu_t(bf2f(zprev[m * ldo + col]), &d);
                    tv[e] = f2bf(a * d);
                    out_bf16[m * ldo + col] = tv[e];
                } else {
                    atomicAdd(&out_f32[m * ldo + col], a);
                }
            }
            if ((EPI == EPI_BIAS_GELU || EPI == EPI_DGELU_BF16) && out_t &&
                mb < ldt) {
                for (int e = 0; e < 4; ++e)
                    if (mb + e >= M)
                        tv[e] = f2bf(0.0f);
                // ldt and mb are multiples of 4: an aligned 8-byte store
                typedef __attribute__((__vector_size__(4 * sizeof(__bf16))))
                    __bf16 bf16x4;
                *(bf16x4 *)(out_t + (int64_t)col * ldt + mb) =
                    bf16x4{tv[0], tv[1], tv[2], tv[3]};
            }
        }
    }
}

// ln of a positive value on the FP32 hardware log2 (inputs are clamped to
// at least 1e-30, inside the FP32 normal range).  Used for every feature so
// that training and fused inference see identical inputs.
__device__ __forceinline__ float ln_feature(double v)
{
    const float x = (float)(v > 1e-30 ? v : 1e-30);
    return __builtin_amdgcn_logf(x) * 0.69314718055994531f;
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
    x[0] = f2bf((ln_feature(T[i]) - mu[0]) / sd[0]);
    x[1] = f2bf((ln_feature(P[i]) - mu[1]) / sd[1]);
    double s = 0.0;
    for (int k = 0; k < nsp - 1; ++k) {
        const double y = Y[(int64_t)k * ld + i];
        s += y;
        x[2 + k] = f2bf((ln_feature(y) - mu[2 + k]) / sd[2 + k]);
    }
    x[1 + nsp] = f2bf((ln_feature(1.0 - s) - mu[1 + nsp]) / sd[1 + nsp]);
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
// columns zero; loss sum accumulated into *loss.  A block covers
// kMseRows states: threads run along the columns (coalesced loads and
// stores), the loss is reduced in the block and added with one atomic.
constexpr int kMseRows = 32;

__global__ __launch_bounds__(256) void k_mse_grad(
    const float *__restrict__ Yh, const float *__restrict__ Yt, int64_t n,
    int nv, int ldy, int ldt, float scale, __bf16 *__restrict__ dY, int ldd,
    float *__restrict__ loss)
{
    const int64_t t0 = (int64_t)blockIdx.x * kMseRows;
    float part = 0.f;
    for (int e = threadIdx.x; e < kMseRows * ldd; e += 256) {
        const int64_t t = t0 + e / ldd;
        const int k = e % ldd;
        if (t >= n)
            break;
        float g = 0.f;
        if (k < nv) {
            const float d = Yh[t * ldy + k] - Yt[t * ldt + k];
            part += d * d;
            g = 2.f * d * scale;
        }
        dY[t * ldd + k] = f2bf(g);
    }
    for (int off = 32; off > 0; off >>= 1)
        part += __shfl_down(part, off, 64);
    __shared__ float wsum[4];
    if ((threadIdx.x & 63) == 0)
        wsum[threadIdx.x >> 6] = part;
    __syncthreads();
    if (threadIdx.x == 0 && loss)
        atomicAdd(loss, wsum[0] + wsum[1] + wsum[2] + wsum[3]);
}

// dst[c][r] = src[r][c] for r < rows, c < cols (bf16), through a padded
// 64 x 64 LDS tile (coalesced reads and writes, no bank conflicts); columns
// r in [rows, ld_dst) of every dst row are written as zero, so a reduction
// over a padded length reads zeros there.
__global__ __launch_bounds__(256) void k_transpose_bf16(
    const __bf16 *__restrict__ src, int64_t ld_src, int64_t rows, int cols,
    __bf16 *__restrict__ dst, int64_t ld_dst)
{
    __shared__ __bf16 tile[64][66];
    const int64_t r0 = (int64_t)blockIdx.x * 64;
    const int c0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int j = ty; j < 64; j += 4) {
        const int64_t r = r0 + j;
        const int c = c0 + tx;
        tile[j][tx] = (r < rows && c < cols) ? src[r * ld_src + c]
                                             : f2bf(0.0f);
    }
    __syncthreads();
    for (int j = ty; j < 64; j += 4) {
        const int c = c0 + j;
        const int64_t r = r0 + tx;
        if (c < cols && r < ld_dst)
            dst[(int64_t)c * ld_dst + r] = tile[tx][j];
    }
}

// Row sums of a bf16 [rows][ld] array (the bias gradient from dZ^T):
// one block per row, 256 threads, 8-element vector loads.
__global__ __launch_bounds__(256) void k_rowsum(
    const __bf16 *__restrict__ a, int64_t ld, int64_t len,
    float *__restrict__ out)
{
    const __bf16 *row = a + (int64_t)blockIdx.x * ld;
    float s = 0.f;
    for (int64_t m = (int64_t)threadIdx.x * 8; m < len; m += 256 * 8) {
        if (m + 8 <= len) {
            const bf16x8 v = *(const bf16x8 *)(row + m);
#pragma unroll
            for (int e = 0; e < 8; ++e)
                s += bf2f(v[e]);
        } else {
            for (int64_t e = m; e < len; ++e)
                s += bf2f(row[e]);
        }
    }
    for (int off = 32; off > 0; off >>= 1)
        s += __shfl_down(s, off, 64);
    __shared__ float part[4];
    if ((threadIdx.x & 63) == 0)
        part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0)
        out[blockIdx.x] += part[0] + part[1] + part[2] + part[3];
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

// ---------------------------------------------------------------------------
// Fused inference: features -> 3 x (256, GELU) -> 64 outputs -> decode -> gate
// in one kernel, activations never leaving registers.
//
// A wave owns 32 states as two 16-state column blocks.  Every product is
// D = W_tile (16 x 32, from LDS) x act (32 x 16 states) on
// v_mfma_f32_16x16x32_bf16, so a layer's output lands with lane l holding
// units 16r + 4(l/16) + t (t < 4) of state l % 16.  Two consecutive row
// blocks hold exactly the 8 values lane l must supply as the next layer's
// B operand for one 32-deep step, provided the next layer's reduction index
// is permuted the same way: step s, lane group g, element e reads unit
// 32 s + 16 (e / 4) + 4 g + e % 4.  crs_mlp_pack writes every weight matrix
// in that tile order (64 lanes x 8 bf16 per 16 x 32 tile), so activations
// pass from layer to layer with no shuffle and no LDS round trip.
//
// The block (4 waves, 128 states) stages the tiles of the first and output
// layers whole and those of each 256 x 256 layer in two halves of 8 row
// blocks, 64 KB of LDS at a time; each tile read feeds two MFMAs (one per
// column block).  Two blocks share a CU and run apart, so one block's
// staging overlaps the other's products and epilogues.
// Features go through LDS once (per-lane state -> B layout), outputs once
// (D layout -> per-lane state), then each lane decodes its state's rates to
// FP64, applies the gate and appends fallback states with one atomic per
// wave.
// ---------------------------------------------------------------------------
constexpr int IH = 256;           // hidden width
constexpr int IKIN = 64;          // padded input width
constexpr int IKOUT = 64;         // padded output width
constexpr int INC = 2;            // 16-state column blocks per wave
constexpr int IWAVES = 4;         // 128 states per block; two blocks per
                                  // CU (two waves per SIMD: one multiplies
                                  // while the other runs its GELU epilogue)
                                  // run apart, so one block's tile staging
                                  // overlaps the other block's products
constexpr int ITHREADS = 64 * IWAVES;
constexpr int ISTATES = 16 * INC * IWAVES;
constexpr int IFEAT_LD = 72;      // bf16 per state row (144 B, 16 B aligned)
constexpr int IOUT_LD = 68;       // floats per state row
constexpr int ILDS = 64 * 1024;   // half a hidden layer of tiles
constexpr int IPIECES = 2;        // a 256 x 256 layer is staged in halves

static_assert(2 * ISTATES == ITHREADS, "two threads per state in the state stages");
static_assert(ISTATES * IFEAT_LD * 2 + ITHREADS * 12 <= ILDS,
              "feature stage and its exchange fit");
static_assert(ISTATES * IOUT_LD * 4 <= ILDS, "output stage fits");
static_assert(IH * IH * 2 / IPIECES <= ILDS, "half a hidden layer fits");
static_assert(IH * IKIN * 2 <= ILDS && IKOUT * IH * 2 <= ILDS,
              "first and output layers fit whole");

// Tile element (lane l = i + 16 g, element e) of row block r, step s.
__device__ __forceinline__ int tile_k(int perm, int s, int g, int e)
{
    return perm ? 32 * s + 16 * (e >> 2) + 4 * g + (e & 3)
                : 32 * s + 8 * g + e;
}

__global__ __launch_bounds__(256) void k_pack_tiles(
    const __bf16 *__restrict__ W, int n_out, int n_in, int perm,
    __bf16 *__restrict__ out)
{
    const int ns = n_in / 32;
    const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)(n_out / 16) * ns * 64;
    if (u >= total)
        return;
    const int lane = (int)(u & 63);
    const int64_t tile = u >> 6;
    const int r = (int)(tile / ns), s = (int)(tile % ns);
    const int i = lane & 15, g = lane >> 4;
    for (int e = 0; e < 8; ++e)
        out[u * 8 + e] = W[(int64_t)(16 * r + i) * n_in + tile_k(perm, s, g, e)];
}

// Copy N16 16-byte words global -> LDS with the whole block, 8 loads in
// flight per thread before their LDS writes (a plain copy loop waits for
// every load before the next one is issued).
template <int N16>
__device__ __forceinline__ void stage_tiles(const uint4 *__restrict__ src,
                                            uint4 *dst)
{
    constexpr int BATCH = N16 / ITHREADS < 8 ? N16 / ITHREADS : 8;
    static_assert(BATCH > 0 && N16 % (BATCH * ITHREADS) == 0, "whole batches");
#pragma unroll
    for (int base = 0; base < N16; base += BATCH * ITHREADS) {
        uint4 v[BATCH];
#pragma unroll
        for (int u = 0; u < BATCH; ++u)
            v[u] = src[base + u * ITHREADS + threadIdx.x];
#pragma unroll
        for (int u = 0; u < BATCH; ++u)
            dst[base + u * ITHREADS + threadIdx.x] = v[u];
    }
}

// One layer on the wave's column blocks: NS reduction steps of 32, NRB
// output row blocks of 16.  The bias seeds the accumulator.  Hidden layers
// write bf16 GELU activations in the next layer's B layout; the output
// layer keeps FP32 results.  Tiles of row block r + 1 are read from LDS
// while row block r multiplies.
typedef __attribute__((__vector_size__(2 * sizeof(__bf16)))) __bf16 bf16x2;
typedef __attribute__((__vector_size__(4 * sizeof(unsigned)))) unsigned u32x4;

// Two GELU activations packed as one bf16 pair, materialised here: the
// empty asm pins the packing to this point (otherwise the multiplies and
// conversions sink to the next layer and keep two FP32 values live per
// activation across the LDS staging).
__device__ __forceinline__ unsigned gelu_pair(float z0, float z1)
{
    bf16x2 p = {f2bf(gelu_t(z0, nullptr)), f2bf(gelu_t(z1, nullptr))};
    unsigned w = __builtin_bit_cast(unsigned, p);
    asm volatile("" : "+v"(w));
    return w;
}

template <int NS, int NRB, bool HIDDEN, int RB0 = 0>
__device__ __forceinline__ void infer_layer(const bf16x8 *wt,
                                            const float *__restrict__ bias,
                                            const u32x4 (&in)[INC][8],
                                            u32x4 (&nxt)[INC][8],
                                            floatx4 (&res)[INC][4], int lane)
{
    const int g = lane >> 4;
    bf16x8 a_cur[NS], a_nxt[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s)
        a_cur[s] = wt[s * 64 + lane];
#pragma unroll
    for (int r = 0; r < NRB; ++r) {
        if (r + 1 < NRB) {
#pragma unroll
            for (int s = 0; s < NS; ++s)
                a_nxt[s] = wt[((r + 1) * NS + s) * 64 + lane];
        }
        const floatx4 bv = *(const floatx4 *)(bias + 16 * (RB0 + r) + 4 * g);
        floatx4 acc[INC];
#pragma unroll
        for (int b = 0; b < INC; ++b)
            acc[b] = bv;
#pragma unroll
        for (int s = 0; s < NS; ++s) {
#pragma unroll
            for (int b = 0; b < INC; ++b)
                acc[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    a_cur[s], __builtin_bit_cast(bf16x8, in[b][s]), acc[b],
                    0, 0, 0);
        }
#pragma unroll
        for (int b = 0; b < INC; ++b) {
            // row block RB0 + r of the layer
            if (HIDDEN) {
                nxt[b][(RB0 + r) >> 1][((RB0 + r) & 1) * 2] =
                    gelu_pair(acc[b][0], acc[b][1]);
                nxt[b][(RB0 + r) >> 1][((RB0 + r) & 1) * 2 + 1] =
                    gelu_pair(acc[b][2], acc[b][3]);
            } else {
                res[b][RB0 + r] = acc[b];
            }
        }
#pragma unroll
        for (int s = 0; s < NS; ++s)
            a_cur[s] = a_nxt[s];
        // one row block at a time: left free, the scheduler interleaves the
        // epilogues of many row blocks and runs out of registers; the other
        // wave on the SIMD overlaps this wave's epilogue instead
        __builtin_amdgcn_sched_barrier(0);
    }
}

// sinh for the decode, FP32 arithmetic with an FP64 result (the decoded
// rate carries the surrogate's own error, far above FP32 rounding): odd
// Taylor series below 1/2, exponentials up to 20, and above that 0.5 e^|x|
// formed as 2^64 * (0.5 e^|x| 2^-64) so the FP32 exponential stays finite up
// to |x| = 132 (beyond, +-inf, which the gate sends to the exact kernel).
__device__ __forceinline__ double sinh_decode(float x)
{
    const float a = fabsf(x);
    if (a < 0.5f) {
        const float x2 = x * x;
        return (double)(x * (1.0f + x2 * (1.0f / 6.0f + x2 * (1.0f / 120.0f +
                                                               x2 * (1.0f / 5040.0f)))));
    }
    if (a < 20.0f) {
        const float e = __builtin_amdgcn_exp2f(a * 1.4426950408889634f);
        return (double)copysignf(0.5f * (e - __builtin_amdgcn_rcpf(e)), x);
    }
    const float e = __builtin_amdgcn_exp2f(a * 1.4426950408889634f - 64.0f);
    return copysign(__builtin_ldexp((double)(0.5f * e), 64), (double)x);
}

// Per-species constants of the state stages, staged once per block.
struct InferConsts {
    float fmu[IKIN], fsd[IKIN];      // feature normalisation
    float ymu[IKOUT], ysd[IKOUT];    // output standardisation
    double scale[IKOUT], mw[IKOUT];  // arcsinh scale, molecular weight
    float elem[8][IKOUT];            // element counts
};

// Partial gate sums of the second half of a state's species.
struct GatePart {
    double mres, mscale, er[8], es[8];
    int bad;
};

static_assert(ISTATES * IOUT_LD * 4 + ISTATES * (int)sizeof(GatePart) <= ILDS,
              "output stage and gate partials fit");

__global__ __launch_bounds__(ITHREADS) void k_mlp_infer(
    const double *__restrict__ T, const double *__restrict__ P,
    const double *__restrict__ Y, int64_t n, int64_t ld, int nsp,
    const __bf16 *__restrict__ wt, const float *__restrict__ bias,
    const float *__restrict__ mu, const float *__restrict__ sd,
    const double *__restrict__ scale, const float *__restrict__ ymu,
    const float *__restrict__ ysd, double *__restrict__ W,
    const crs_mlp_gate_params gp, const double *__restrict__ elem, int n_elem,
    const double *__restrict__ mw, int32_t *__restrict__ flags,
    int32_t *__restrict__ fb_idx, int32_t *__restrict__ fb_count)
{
    __shared__ __attribute__((aligned(16))) unsigned char smem[ILDS];
    __shared__ InferConsts cst;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int g = lane >> 4, j = lane & 15;
    // state stages: thread t works on state t % ISTATES, species of parity
    // t / ISTATES (two threads per state)
    const int st = threadIdx.x % ISTATES;
    const int half = threadIdx.x / ISTATES;
    const int64_t m = (int64_t)blockIdx.x * ISTATES + st;
    const bool valid = m < n;

    for (int c = threadIdx.x; c < IKIN; c += ITHREADS) {
        const bool in = c < nsp + 2;
        cst.fmu[c] = in ? mu[c] : 0.0f;
        cst.fsd[c] = in ? sd[c] : 1.0f;
    }
    for (int k = threadIdx.x; k < IKOUT; k += ITHREADS) {
        const bool in = k < nsp;
        cst.ymu[k] = in ? ymu[k] : 0.0f;
        cst.ysd[k] = in ? ysd[k] : 0.0f;
        cst.scale[k] = in ? scale[k] : 0.0;
        cst.mw[k] = in ? mw[k] : 0.0;
        for (int e = 0; e < 8; ++e)
            cst.elem[e][k] = (in && e < n_elem) ? (float)elem[e * nsp + k] : 0.0f;
    }
    __syncthreads();

    // ---- features of the state, composition checks ---------------------
    // The state's two threads split its species by parity; each issues all
    // of its (at most 32) loads before using any, then the halves exchange
    // their partial mass-fraction sums and composition checks through LDS.
    // Every element 0..IKIN-1 of the state's row is written exactly once.
    int bad = 0;
    double *xsum = (double *)(smem + ISTATES * IFEAT_LD * 2);
    int *xbad = (int *)(xsum + ITHREADS);
    {
        __bf16 *frow = (__bf16 *)smem + (int64_t)st * IFEAT_LD;
        if (half == 0)
            for (int c = nsp + 2; c < IKIN; ++c)
                frow[c] = f2bf(0.0f);
        double y[32];
#pragma unroll
        for (int u = 0; u < 32; ++u) {
            const int k = 2 * u + half;
            y[u] = (valid && k < nsp - 1) ? Y[(int64_t)k * ld + m] : 0.0;
        }
        double psum = 0.0;
        int pbad = 0;
#pragma unroll
        for (int u = 0; u < 32; ++u) {
            const int k = 2 * u + half;
            if (k < nsp - 1) {
                if (!(y[u] >= 0.0 && y[u] <= 1.0))
                    pbad = 1;
                psum += y[u];
                frow[2 + k] = valid ? f2bf((ln_feature(y[u]) - cst.fmu[2 + k]) /
                                           cst.fsd[2 + k])
                                    : f2bf(0.0f);
            }
        }
        xsum[threadIdx.x] = psum;
        xbad[threadIdx.x] = pbad;
        if (valid) {
            const double t = T[m], p = P[m];
            if (!(t >= gp.t_lo && t <= gp.t_hi && p >= gp.p_lo && p <= gp.p_hi))
                pbad = 1;
            frow[half] = f2bf((ln_feature(half ? p : t) - cst.fmu[half]) /
                              cst.fsd[half]);
        } else {
            frow[half] = f2bf(0.0f);
        }
        __syncthreads();
        const int other = threadIdx.x ^ ISTATES;
        const double ssum = psum + xsum[other];
        bad = pbad | xbad[other];
        if (!(ssum <= 1.0 + 1e-12))
            bad = 1;
        if (((nsp - 1) & 1) == half)
            frow[1 + nsp] = valid ? f2bf((ln_feature(1.0 - ssum) -
                                          cst.fmu[1 + nsp]) / cst.fsd[1 + nsp])
                                  : f2bf(0.0f);
    }
    __syncthreads();
    u32x4 act[INC][8], nxt[INC][8];
    floatx4 res[INC][4];
    {
        const __bf16 *fw =
            (const __bf16 *)smem + (int64_t)(wave * 16 * INC) * IFEAT_LD;
#pragma unroll
        for (int b = 0; b < INC; ++b)
#pragma unroll
            for (int s = 0; s < IKIN / 32; ++s)
                act[b][s] = *(const u32x4 *)(fw + (16 * b + j) * IFEAT_LD +
                                             32 * s + 8 * g);
    }

    // ---- layers --------------------------------------------------------
    const bf16x8 *lt = (const bf16x8 *)smem;
    const uint4 *src = (const uint4 *)wt;
    constexpr int n0 = IH * IKIN * 2 / 16, n1 = IH * IH * 2 / 16;
    constexpr int n3 = IKOUT * IH * 2 / 16;
    __syncthreads();
    stage_tiles<n0>(src, (uint4 *)smem);
    __syncthreads();
    infer_layer<IKIN / 32, IH / 16, true>(lt, bias, act, nxt, res, lane);
    src += n0;
    // the 256 x 256 layers in two halves of 8 row blocks each
    constexpr int nh = n1 / IPIECES;
    static_assert(IPIECES == 2, "two halves per hidden layer");
    __syncthreads();
    stage_tiles<nh>(src, (uint4 *)smem);
    __syncthreads();
    infer_layer<IH / 32, IH / 32, true, 0>(lt, bias + IH, nxt, act, res, lane);
    __syncthreads();
    stage_tiles<nh>(src + nh, (uint4 *)smem);
    __syncthreads();
    infer_layer<IH / 32, IH / 32, true, IH / 32>(lt, bias + IH, nxt, act, res,
                                                 lane);
    src += n1;
    __syncthreads();
    stage_tiles<nh>(src, (uint4 *)smem);
    __syncthreads();
    infer_layer<IH / 32, IH / 32, true, 0>(lt, bias + 2 * IH, act, nxt, res,
                                           lane);
    __syncthreads();
    stage_tiles<nh>(src + nh, (uint4 *)smem);
    __syncthreads();
    infer_layer<IH / 32, IH / 32, true, IH / 32>(lt, bias + 2 * IH, act, nxt,
                                                 res, lane);
    src += n1;
    __syncthreads();
    stage_tiles<n3>(src, (uint4 *)smem);
    __syncthreads();
    infer_layer<IH / 32, IKOUT / 16, false>(lt, bias + 3 * IH, nxt, act, res,
                                            lane);
    __syncthreads();

    // ---- outputs to per-state rows -------------------------------------
    {
        float *ow = (float *)smem + (int64_t)(wave * 16 * INC) * IOUT_LD;
#pragma unroll
        for (int b = 0; b < INC; ++b)
#pragma unroll
            for (int r = 0; r < IKOUT / 16; ++r)
                *(floatx4 *)(ow + (16 * b + j) * IOUT_LD + 16 * r + 4 * g) =
                    res[b][r];
    }
    __syncthreads();
    GatePart *parts = (GatePart *)(smem + ISTATES * IOUT_LD * 4);
    GatePart gpart = {};
    if (valid) {
        // decode this thread's species (parity = half), partial balances
        const float *orow = (const float *)smem + (int64_t)st * IOUT_LD;
#pragma unroll 2
        for (int k = half; k < nsp; k += 2) {
            const float ya = cst.ymu[k] + cst.ysd[k] * orow[k];
            const double w = sinh_decode(ya) * cst.scale[k];
            W[(int64_t)k * ld + m] = w;
            if (!isfinite(w))
                gpart.bad = 1;
            gpart.mres += cst.mw[k] * w;
            gpart.mscale += cst.mw[k] * fabs(w);
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const double a = (double)cst.elem[e][k];
                gpart.er[e] += a * w;
                gpart.es[e] += a * fabs(w);
            }
        }
        if (half == 1)
            parts[st] = gpart;
    }
    __syncthreads();
    if (half == 0) {
        if (valid) {
            const GatePart &o = parts[st];
            bad |= gpart.bad | o.bad;
            const double mres = gpart.mres + o.mres;
            const double mscale = gpart.mscale + o.mscale;
            if (mscale > 0.0 && fabs(mres) > gp.rmax * mscale)
                bad = 1;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const double r = gpart.er[e] + o.er[e];
                const double sc = gpart.es[e] + o.es[e];
                if (sc > 0.0 && fabs(r) > gp.rmax * sc)
                    bad = 1;
            }
            if (flags)
                flags[m] = bad;
        }
        // ---- fallback list: one atomic per wave ------------------------
        const uint64_t ballot = __ballot(valid && bad);
        if (ballot) {
            const int leader = __ffsll((unsigned long long)ballot) - 1;
            int base = 0;
            if (lane == leader)
                base = atomicAdd(fb_count, (int)__popcll(ballot));
            base = __shfl(base, leader);
            if (valid && bad)
                fb_idx[base + __popcll(ballot & ((1ull << lane) - 1ull))] =
                    (int32_t)m;
        }
    }
}

// Most states one call of crs_mlp_eval_gated_noalloc takes.
constexpr int64_t kNoAllocStates = 1 << 22;

// *p = 0 (a fill kernel rather than a memset, so that a captured graph
// replays it as a kernel node).
__global__ void k_zero_i32(int32_t *__restrict__ p)
{
    if (threadIdx.x == 0)
        *p = 0;
}

// Every state to the exact fallback: flags 1 (when given), idx[t] = t,
// *count = n.  The gated call's path for a gate that admits no state.
__global__ void k_all_fallback(int32_t *__restrict__ flags,
                               int32_t *__restrict__ idx,
                               int32_t *__restrict__ count, int32_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    idx[t] = (int32_t)t;
    if (flags) flags[t] = 1;
    if (t == 0) *count = n;
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
                       void *A_out, void *Z_out, float *Y_out, int64_t ldo,
                       int n_valid, void *AT_out, int64_t ldat, void *stream)
{
    const char *fn = "crs_mlp_linear_fwd";
    CRS_TRY(check_gemm(fn, M, N, K));
    if (AT_out && (ldat < ((M + 31) / 32) * 32 || ldat % 4))
        return crs::fail(CRS_ERR_ARG, "%s: ldat must be M rounded up to "
                         "32 or more, a multiple of 4", fn);
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
            nullptr, ldo, (__bf16 *)A_out, (__bf16 *)Z_out, nullptr, N,
            (__bf16 *)AT_out, ldat);
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
                            const void *Zprev, void *dZprev, int64_t ldo,
                            void *dZprevT, int64_t ldt, void *stream)
{
    const char *fn = "crs_mlp_linear_bwd_data";
    // dX[M][K] = dZ[M][N] W[N][K]: reduction over N; columns K
    CRS_TRY(check_gemm(fn, M, K, N));
    if (dZprevT && (ldt < ((M + 31) / 32) * 32 || ldt % 4))
        return crs::fail(CRS_ERR_ARG, "%s: ldt must be M rounded up to 32 "
                         "or more, a multiple of 4", fn);
    if (M == 0)
        return CRS_OK;
    if (!dZ || !Wb || !Zprev || !dZprev)
        return crs::fail(CRS_ERR_ARG, "%s: NULL pointer", fn);
    dim3 grid((unsigned)((M + TM - 1) / TM), K / TN, 1);
    // B(k, n) = W^T[k][n]: reduction-contiguous rows of the transposed W
    k_gemm<true, true, EPI_DGELU_BF16><<<grid, 256, 0, (hipStream_t)stream>>>(
        (const __bf16 *)dZ, ldz, (const __bf16 *)Wb, ldw, M, K, N, N, nullptr,
        nullptr, ldo, (__bf16 *)dZprev, nullptr, (const __bf16 *)Zprev, K,
        (__bf16 *)dZprevT, ldt);
    return crs::launched(fn);
}

int crs_mlp_linear_bwd_weight(const void *dZT, int64_t ldz, const void *XT,
                              int64_t ldx, int64_t M, int N, int K, float *dW,
                              float *db, void *stream)
{
    const char *fn = "crs_mlp_linear_bwd_weight";
    // dW[N][K] = sum_m dZT[n][m] XT[k][m]: rows N, columns K, reduction M
    // along the contiguous dimension of both operands
    if (M <= 0 || N <= 0 || K <= 0 || N % TN != 0 || K % TN != 0)
        return crs::fail(CRS_ERR_ARG, "%s: bad sizes", fn);
    const int64_t Mr = ((M + TR - 1) / TR) * TR;
    if (ldz < Mr || ldx < Mr)
        return crs::fail(CRS_ERR_ARG, "%s: leading dimensions must cover M "
                         "rounded up to %d (zero padded)", fn, TR);
    if (!dZT || !XT || !dW || !db)
        return crs::fail(CRS_ERR_ARG, "%s: NULL pointer", fn);
    // split the reduction so that the grid covers the GPU: 256+ blocks
    const int64_t tiles = (int64_t)((N + TM - 1) / TM) * (K / TN);
    int64_t nz = (256 + tiles - 1) / tiles;
    int64_t chunk = ((Mr + nz - 1) / nz + TR - 1) / TR * TR;
    nz = (Mr + chunk - 1) / chunk;
    dim3 grid((unsigned)((N + TM - 1) / TM), K / TN, (unsigned)nz);
    hipStream_t s = (hipStream_t)stream;
    k_gemm<true, true, EPI_ATOMIC_F32><<<grid, 256, 0, s>>>(
        (const __bf16 *)dZT, ldz, (const __bf16 *)XT, ldx, N, K, Mr, chunk,
        nullptr, dW, K, nullptr, nullptr, nullptr, K);
    CRS_TRY(crs::launched(fn));
    k_rowsum<<<(unsigned)N, 256, 0, s>>>((const __bf16 *)dZT, ldz, M, db);
    return crs::launched(fn);
}

int crs_mlp_transpose_bf16(const void *src, int64_t ld_src, int64_t rows,
                           int cols, void *dst, int64_t ld_dst, void *stream)
{
    const char *fn = "crs_mlp_transpose_bf16";
    if (rows < 0 || cols < 0 || ld_src < cols || ld_dst < rows)
        return crs::fail(CRS_ERR_ARG, "%s: bad sizes", fn);
    if (rows == 0 || cols == 0)
        return CRS_OK;
    if (!src || !dst)
        return crs::fail(CRS_ERR_ARG, "%s: NULL pointer", fn);
    dim3 grid((unsigned)((ld_dst + 63) / 64), (unsigned)((cols + 63) / 64));
    k_transpose_bf16<<<grid, 256, 0, (hipStream_t)stream>>>(
        (const __bf16 *)src, ld_src, rows, cols, (__bf16 *)dst, ld_dst);
    return crs::launched(fn);
}

int crs_mlp_features(const double *T, const double *P, const double *Y,
                     const int32_t *idx, int64_t n, int64_t ld, int nsp,
                     const float *mu, const float *sd, void *X, int kp,
                     void *stream)
{
    const char *fn = "crs_mlp_features";
    // without idx, states 0..n-1 are read: ld must cover them
    if (n < 0 || ld < 0 || (!idx && ld < n) || nsp < 2 || kp < nsp + 2)
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
    if (n < 0 || ld < n || nsp < 1 || ldt < nsp)
        return crs::fail(CRS_ERR_ARG, "%s: bad sizes", fn);
    if (n == 0)
        return CRS_OK;
    if (!W || !scale || !mu || !sd || !Yt)
        return crs::fail(CRS_ERR_ARG, "%s: NULL pointer", fn);
    k_targets<<<crs::grid_for(n), crs::kBlock, 0, (hipStream_t)stream>>>(
        W, n, ld, nsp, scale, mu, sd, Yt, ldt);
    return crs::launched(fn);
}

int crs_mlp_decode(const float *Yh, int ldy, const int32_t *idx, int64_t n,
                   int64_t ld, int nsp, const double *scale, const float *mu,
                   const float *sd, double *W, void *stream)
{
    const char *fn = "crs_mlp_decode";
    if (n < 0 || ld < 0 || (!idx && ld < n) || nsp < 1 || ldy < nsp)
        return crs::fail(CRS_ERR_ARG, "%s: bad sizes", fn);
    if (n == 0)
        return CRS_OK;
    if (!Yh || !scale || !mu || !sd || !W)
        return crs::fail(CRS_ERR_ARG, "%s: NULL pointer", fn);
    k_decode<<<crs::grid_for(n), crs::kBlock, 0, (hipStream_t)stream>>>(
        Yh, ldy, idx, n, ld, nsp, scale, mu, sd, W);
    return crs::launched(fn);
}

int crs_mlp_mse_grad(const float *Yh, const float *Yt, int64_t n, int nv,
                     int ldy, int ldt, void *dY, int ldd, float *loss,
                     void *stream)
{
    const char *fn = "crs_mlp_mse_grad";
    if (n <= 0 || nv <= 0 || ldy < nv || ldt < nv || ldd < nv)
        return crs::fail(CRS_ERR_ARG, "%s: bad sizes", fn);
    if (!Yh || !Yt || !dY)
        return crs::fail(CRS_ERR_ARG, "%s: NULL pointer", fn);
    const float scale = 1.0f / ((float)n * (float)nv);
    k_mse_grad<<<(unsigned)((n + kMseRows - 1) / kMseRows), 256, 0,
                 (hipStream_t)stream>>>(Yh, Yt, n, nv, ldy, ldt, scale,
                                        (__bf16 *)dY, ldd, loss);
    return crs::launched(fn);
}

int crs_mlp_adam(float *p, const float *g, float *m1, float *m2, void *pb,
                 int64_t n, float lr, float b1, float b2, float eps, int step,
                 float wd, void *stream)
{
    const char *fn = "crs_mlp_adam";
    if (n <= 0 || step <= 0)
        return crs::fail(CRS_ERR_ARG, "%s: bad sizes", fn);
    if (!p || !g || !m1 || !m2 || !pb)
        return crs::fail(CRS_ERR_ARG, "%s: NULL pointer", fn);
    const float c1 = 1.f - powf(b1, (float)step);
    const float c2 = 1.f - powf(b2, (float)step);
    k_adam<<<crs::grid_for(n), crs::kBlock, 0, (hipStream_t)stream>>>(
        p, g, m1, m2, (__bf16 *)pb, n, lr, b1, b2, eps, c1, c2, wd);
    return crs::launched(fn);
}

int crs_mlp_to_bf16(const float *a, void *b, int64_t n, void *stream)
{
    if (n < 0)
        return crs::fail(CRS_ERR_ARG, "crs_mlp_to_bf16: n must be >= 0");
    if (n == 0)
        return CRS_OK;
    if (!a || !b)
        return crs::fail(CRS_ERR_ARG, "crs_mlp_to_bf16: NULL pointer");
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
    if (n < 0 || ld < n || nsp < 2 || n_elem < 0)
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

int crs_mlp_pack_tiles(const void *W, int n_out, int n_in, int perm,
                       void *out, void *stream)
{
    const char *fn = "crs_mlp_pack_tiles";
    if (n_out <= 0 || n_in <= 0 || n_out % 16 || n_in % 32)
        return crs::fail(CRS_ERR_ARG, "%s: bad shape %d x %d", fn, n_out, n_in);
    if (!W || !out)
        return crs::fail(CRS_ERR_ARG, "%s: NULL pointer", fn);
    const int64_t total = (int64_t)n_out / 16 * (n_in / 32) * 64;
    k_pack_tiles<<<crs::grid_for(total), crs::kBlock, 0, (hipStream_t)stream>>>(
        (const __bf16 *)W, n_out, n_in, perm, (__bf16 *)out);
    return crs::launched(fn);
}

}  // extern "C"

namespace {

// Checks and launch of k_mlp_infer, shared by crs_mlp_infer and the gated
// entry points.  flags may be NULL when need_flags is false.
int launch_infer(const char *fn, const double *T, const double *P,
                 const double *Y, int64_t n, int64_t ld, int nsp, int kin,
                 int hidden, int layers, int kout, const void *tiles,
                 const float *bias, const float *mu, const float *sd,
                 const double *scale, const float *ymu, const float *ysd,
                 double *W, crs_mlp_gate_params gp, const double *elem,
                 int n_elem, const double *mw, int32_t *flags, bool need_flags,
                 int32_t *fb_idx, int32_t *fb_count, hipStream_t s)
{
    if (kin != IKIN || hidden != IH || layers != 3 || kout != IKOUT)
        return crs::fail(CRS_ERR_UNSUPPORTED,
                         "%s: fused kernel is built for %d -> 3 x %d -> %d "
                         "(got %d -> %d x %d -> %d)", fn, IKIN, IH, IKOUT, kin,
                         layers, hidden, kout);
    if (n < 0 || nsp < 2 || nsp + 2 > IKIN || nsp > IKOUT || ld < n ||
        n_elem < 0 || n_elem > 8)
        return crs::fail(CRS_ERR_ARG, "%s: bad sizes n=%lld ld=%lld nsp=%d "
                         "n_elem=%d", fn, (long long)n, (long long)ld, nsp,
                         n_elem);
    if (n == 0)
        return CRS_OK;
    if (!T || !P || !Y || !tiles || !bias || !mu || !sd || !scale || !ymu ||
        !ysd || !W || !mw || (need_flags && !flags) || !fb_idx ||
        !fb_count || (n_elem > 0 && !elem))
        return crs::fail(CRS_ERR_ARG, "%s: NULL pointer", fn);
    if (((uintptr_t)tiles & 15) || ((uintptr_t)bias & 15))
        return crs::fail(CRS_ERR_ARG, "%s: tiles and bias need 16-byte "
                         "alignment", fn);
    const dim3 grid((unsigned)((n + ISTATES - 1) / ISTATES));
    k_mlp_infer<<<grid, ITHREADS, 0, s>>>(
        T, P, Y, n, ld, nsp, (const __bf16 *)tiles, bias, mu, sd, scale, ymu,
        ysd, W, gp, elem, n_elem, mw, flags, fb_idx, fb_count);
    return crs::launched(fn);
}

// Gated evaluation on one stream: count = 0, fused inference + gate
// (appending fallback states to fb_idx and counting them in *fb_count),
// then the mechanism library's exact subset kernel over fb_idx, bounded by
// the device-side count.  Nothing is read back to the host.
int gated(const char *fn, const crs_mlp_net *net, const double *T,
          const double *P, const double *Y, double *W, int64_t n, int64_t ld,
          crs_subset_fn exact, int exact_mode, int32_t *flags,
          int32_t *fb_idx, int32_t *fb_count, hipStream_t s)
{
    if (!net || !exact)
        return crs::fail(CRS_ERR_ARG, "%s: NULL network or exact function",
                         fn);
    if (n < 0 || ld < n)
        return crs::fail(CRS_ERR_ARG, "%s: bad sizes n=%lld ld=%lld", fn,
                         (long long)n, (long long)ld);
    if (n == 0)
        return CRS_OK;
    if (!fb_idx || !fb_count)
        return crs::fail(CRS_ERR_ARG, "%s: NULL fallback list", fn);
    if (!(net->gate.rmax >= 0.0)) {
        // a gate that admits no state (rmax < 0: no residual threshold met
        // the error target): the network is not evaluated, every state
        // goes to the exact kernel
        if (n > INT32_MAX)
            return crs::fail(CRS_ERR_ARG, "%s: n = %lld exceeds the int32 "
                             "fallback list", fn, (long long)n);
        k_all_fallback<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(
            flags, fb_idx, fb_count, (int32_t)n);
        CRS_TRY(crs::launched(fn));
    } else {
        k_zero_i32<<<1, 64, 0, s>>>(fb_count);
        CRS_TRY(crs::launched(fn));
        CRS_TRY(launch_infer(fn, T, P, Y, n, ld, net->nsp, net->kin,
                             net->hidden, net->layers, net->kout, net->tiles,
                             net->bias, net->mu, net->sd, net->scale,
                             net->ymu, net->ysd, W, net->gate, net->elem,
                             net->n_elem, net->mw, flags, false, fb_idx,
                             fb_count, s));
    }
    const int rc = exact(T, P, Y, W, fb_idx, fb_count, n, ld, exact_mode,
                         (void *)s);
    if (rc != CRS_OK)
        return crs::fail(rc, "%s: exact fallback failed with code %d (the "
                         "mechanism library's crs_last_error has the "
                         "message)", fn, rc);
    return CRS_OK;
}

}  // namespace

// Fallback list of crs_mlp_eval_gated_noalloc: the count in element 0
// (elements 1..63 pad the list to a 256-byte boundary), then up to
// kNoAllocStates state indices.  16.8 MB on every GPU the library runs
// on, held until the process ends.  Host code never uses this name as an
// address (on the host it is only a stand-in): it uses the per-device
// address scratch_address() takes from hipGetSymbolAddress.
static __device__ __align__(256) int32_t scratch[64 + kNoAllocStates];

namespace {

// Device address of `scratch` on the current device, looked up once per
// device (each GPU holds its own copy at its own address).
int scratch_address(const char *fn, int32_t **out)
{
    static std::mutex lock;
    static int32_t *addr[64] = {nullptr};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
        return crs::fail(CRS_ERR_HIP, "%s: no current HIP device", fn);
    std::lock_guard<std::mutex> g(lock);
    if (addr[dev] == nullptr) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            return crs::fail(CRS_ERR_HIP, "%s: static fallback list: %s",
                             fn, hipGetErrorString(e));
        addr[dev] = (int32_t *)p;
    }
    *out = addr[dev];
    return CRS_OK;
}

}  // namespace

extern "C" {

int crs_mlp_eval_gated_noalloc(const crs_mlp_net *net, const double *T,
                               const double *P, const double *Y, double *W,
                               int64_t n, int64_t ld, crs_subset_fn exact,
                               int exact_mode, void *stream)
{
    const char *fn = "crs_mlp_eval_gated_noalloc";
    if (n > kNoAllocStates)
        return crs::fail(CRS_ERR_ARG, "%s: n = %lld exceeds the static "
                         "fallback list (%lld states)", fn, (long long)n,
                         (long long)kNoAllocStates);
    if (n <= 0)
        return gated(fn, net, T, P, Y, W, n, ld, exact, exact_mode, nullptr,
                     nullptr, nullptr, (hipStream_t)stream);
    int32_t *list = nullptr;
    CRS_TRY(scratch_address(fn, &list));
    return gated(fn, net, T, P, Y, W, n, ld, exact, exact_mode, nullptr,
                 list + 64, list, (hipStream_t)stream);
}

int crs_mlp_infer(const double *T, const double *P, const double *Y,
                  int64_t n, int64_t ld, int nsp, int kin, int hidden,
                  int layers, int kout, const void *tiles, const float *bias,
                  const float *mu, const float *sd, const double *scale,
                  const float *ymu, const float *ysd, double *W,
                  crs_mlp_gate_params gp, const double *elem, int n_elem,
                  const double *mw, int32_t *flags, int32_t *fb_idx,
                  int32_t *fb_count, void *stream)
{
    return launch_infer("crs_mlp_infer", T, P, Y, n, ld, nsp, kin, hidden,
                        layers, kout, tiles, bias, mu, sd, scale, ymu, ysd, W,
                        gp, elem, n_elem, mw, flags, true, fb_idx, fb_count,
                        (hipStream_t)stream);
}

int crs_mlp_eval_gated(const crs_mlp_net *net, const double *T,
                       const double *P, const double *Y, double *W, int64_t n,
                       int64_t ld, crs_subset_fn exact, int exact_mode,
                       int32_t *flags, int32_t *fb_idx, int32_t *fb_count,
                       void *stream)
{
    return gated("crs_mlp_eval_gated", net, T, P, Y, W, n, ld, exact,
                 exact_mode, flags, fb_idx, fb_count, (hipStream_t)stream);
}

}  // extern "C"
