// FP64 FMA issue ceiling of the part: K independent FMA chains per lane,
// one block of 256 x w threads per CU (w waves per SIMD), so that the fused
// wdot kernels' VALU share can be read against what plain FP64 FMA code
// reaches at the same occupancy.  Timed by the caller on its stream
// (create_rate_subs_amd.tune ceiling).
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

}  // namespace

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
