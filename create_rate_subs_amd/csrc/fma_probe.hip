// Ceilings of the part that the kernels' figures are read against, timed
// by the caller on its stream:
// * FP64 FMA issue (create_rate_subs_amd.tune ceiling): K independent FMA
//   chains per lane, one block of 256 x w threads per CU (w waves per
//   SIMD), so that the fused wdot kernels' VALU share can be read against
//   what plain FP64 FMA code reaches at the same occupancy;
// * HBM streaming (create_rate_subs_amd.tune bandwidth): a plain copy, and
//   the row pattern of the dydt energy and RK4 stage kernels (per state 53
//   rows read, 53 more read, 53 written, row stride ld) with 1, 2 or 4
//   states per lane.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

template <int K>
__global__ void k_fma_chains(double *__restrict__ out, double b, double c,
                             int n)
{
    double a[K];
#pragma unroll
    for (int k = 0; k < K; ++k) a[k] = (double)(threadIdx.x + k);
    for (int i = 0; i < n; ++i) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
#pragma unroll
            for (int k = 0; k < K; ++k) a[k] = __builtin_fma(a[k], b, c);
        }
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) s += a[k];
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

typedef double d2 __attribute__((ext_vector_type(2)));
typedef double d4 __attribute__((ext_vector_type(4)));

// b[t] = a[t], 16 bytes per lane, grid-stride
__global__ __launch_bounds__(256) void k_copy16(const d2 *__restrict__ a,
                                                d2 *__restrict__ b, int64_t m)
{
    for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < m;
         t += (int64_t)gridDim.x * 256)
        b[t] = a[t];
}

// The energy stage's rows: a sum over the rows of Y, then the rows of W in
// chunks of 8 from the last one down, each scaled by the sum into dy.  V
// holds the lane's 1, 2 or 4 consecutive states; m = n / states per lane.
template <class V, int R>
__global__ __launch_bounds__(256) void k_rows(const double *__restrict__ Y,
                                              const double *W, double *dy,
                                              int64_t m, int64_t ld)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= m) return;
    const int64_t i = t * (int64_t)(sizeof(V) / sizeof(double));
    V s = (V)(0.0);
    for (int r = 0; r < R; ++r)
        s = s + *(const V *)(Y + r * ld + i) * (1.0 + r * 1e-3);
    for (int c = R - 8; c > -8; c -= 8) {
        V w[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (c + k >= 0) w[k] = *(const V *)(W + (c + k) * ld + i);
#pragma unroll
        for (int k = 7; k >= 0; --k)
            if (c + k >= 0) *(V *)(dy + (c + k) * ld + i) = w[k] * s;
    }
}

}  // namespace

// Copy nbytes (a multiple of 16, both pointers 16-byte aligned) from a to
// b with 4096 blocks of 256 threads.  Returns 0, or -1 for a bad argument
// or a failed launch.
extern "C" int crs_probe_copy(const void *a, void *b, int64_t nbytes,
                              void *stream)
{
    if (a == nullptr || b == nullptr || nbytes <= 0 || nbytes % 16 ||
        ((uintptr_t)a | (uintptr_t)b) % 16)
        return -1;
    k_copy16<<<4096, 256, 0, (hipStream_t)stream>>>(
        (const d2 *)a, (d2 *)b, nbytes / 16);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// The energy-stage row pattern on 53 rows of n states at stride ld in Y, W
// and dy (each 53 x ld doubles), lanes taking 1, 2 or 4 consecutive
// states.  n must be a multiple of the states per lane and ld a multiple of
// 4; the arrays 32-byte aligned.  Returns 0, or -1 for a bad argument or a
// failed launch.
extern "C" int crs_probe_rows(const double *Y, const double *W, double *dy,
                              int64_t n, int64_t ld, int per_lane,
                              void *stream)
{
    if (Y == nullptr || W == nullptr || dy == nullptr || n <= 0 ||
        ld < n || ld % 4 || n % per_lane ||
        ((uintptr_t)Y | (uintptr_t)W | (uintptr_t)dy) % 32)
        return -1;
    const int64_t m = n / per_lane;
    const dim3 grid((unsigned)((m + 255) / 256));
    hipStream_t s = (hipStream_t)stream;
    switch (per_lane) {
    case 1: k_rows<double, 53><<<grid, 256, 0, s>>>(Y, W, dy, m, ld); break;
    case 2: k_rows<d2, 53><<<grid, 256, 0, s>>>(Y, W, dy, m, ld); break;
    case 4: k_rows<d4, 53><<<grid, 256, 0, s>>>(Y, W, dy, m, ld); break;
    default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Launch blocks x (256 x waves_per_simd) threads, each running k chains of
// n x 16 dependent FMAs (2 flops each); out holds blocks x threads doubles.
// Returns 0, or -1 for an argument outside k in {1, 2, 4, 8},
// waves_per_simd in 1..4, blocks >= 1, n >= 1, or a failed launch.
extern "C" int crs_probe_fma(int k, int waves_per_simd, int blocks, int n,
                             double *out, void *stream)
{
    if (waves_per_simd < 1 || waves_per_simd > 4 || blocks < 1 || n < 1 ||
        out == nullptr)
        return -1;
    const dim3 grid(blocks), block(256 * waves_per_simd);
    hipStream_t s = (hipStream_t)stream;
    switch (k) {
    case 1: k_fma_chains<1><<<grid, block, 0, s>>>(out, 0.999, 1e-3, n); break;
    case 2: k_fma_chains<2><<<grid, block, 0, s>>>(out, 0.999, 1e-3, n); break;
    case 4: k_fma_chains<4><<<grid, block, 0, s>>>(out, 0.999, 1e-3, n); break;
    case 8: k_fma_chains<8><<<grid, block, 0, s>>>(out, 0.999, 1e-3, n); break;
    default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
