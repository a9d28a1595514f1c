// synth.hip -- seeded synthetic triangles and right-hand sides generated in
// place in device memory.
//
// Same counter-based SplitMix64 streams as rocblas_amd/synth.py (Steele, Lea &
// Flood, OOPSLA 2014): value c of stream `key` is mix(key + (c+1)*GAMMA), a
// uniform double takes its top 53 bits, element (i, j) of a k x k triangle
// draws counter i + j*k and element (i, j) of an m x n right-hand side counter
// i + j*m.  The uniform distributions are bit-identical to the host generator
// (only exact operations and one correctly rounded division); the normal one
// uses the device log1p/cos and agrees to a few ulps.  Factors of the serving
// loop and the large benchmark inputs are drawn here, so they never cross
// PCIe.
#include "device_common.hpp"

#include "../../include/rocblas_amd_ext.h"

#include <new>

namespace rbamd
{
namespace
{
constexpr uint64_t kGamma   = 0x9E3779B97F4A7C15ull;
constexpr uint64_t kTagMult = 0xD1B54A32D192ED03ull;

__host__ __device__ inline uint64_t splitmix_mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ inline double splitmix_u01(uint64_t key, uint64_t ctr)
{
    const uint64_t z = splitmix_mix(key + (ctr + 1) * kGamma);
    return double(z >> 11) * (1.0 / 9007199254740992.0);
}

inline uint64_t stream_key(uint64_t seed, uint64_t tag)
{
    return splitmix_mix(seed + tag * kTagMult);
}

// 0: diagonal U[1,2], off-diagonal U[-1,1]/k; 1: off-diagonal N(0,1)/sqrt(k),
// normal number k*k + i + j*k of the stream, i.e. Box-Muller on uniform
// counters 2(k*k + i + j*k) and that + 1.
template <typename T>
__global__ void __launch_bounds__(256) synth_tri_kernel(T* A, int k, int lda, bool upper,
                                                         int dist, uint64_t key, double scale,
                                                         double poison)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if(i >= lda)
        return;
    const double inv_sqrt_k = 1.0 / sqrt(double(k));
    for(int j = blockIdx.y; j < k; j += gridDim.y)
    {
        double     v   = poison;
        const bool tri = i < k && (i == j || (upper ? i < j : i > j));
        if(tri)
        {
            const uint64_t c = uint64_t(i) + uint64_t(j) * uint64_t(k);
            if(i == j)
                v = (1.0 + splitmix_u01(key, c)) * scale;
            else if(dist == 0)
                v = ((2.0 * splitmix_u01(key, c) - 1.0) / double(k)) * scale;
            else
            {
                const uint64_t base = 2 * (uint64_t(k) * uint64_t(k) + c);
                const double   u1   = splitmix_u01(key, base);
                const double   u2   = splitmix_u01(key, base + 1);
                v = sqrt(-2.0 * log1p(-u1)) * cos(2.0 * M_PI * u2) * inv_sqrt_k * scale;
            }
        }
        A[int64_t(i) + int64_t(j) * lda] = T(v);
    }
}

// U[-1,1] in rows < m, zero in the padding rows m..ldb-1
template <typename T>
__global__ void __launch_bounds__(256) synth_rhs_kernel(T* B, int m, int n, int ldb,
                                                         uint64_t key)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if(i >= ldb)
        return;
    for(int j = blockIdx.y; j < n; j += gridDim.y)
    {
        double v = 0.0;
        if(i < m)
            v = -1.0 + 2.0 * splitmix_u01(key, uint64_t(i) + uint64_t(j) * uint64_t(m));
        B[int64_t(i) + int64_t(j) * ldb] = T(v);
    }
}

template <typename T>
rocblas_status synth_tri_entry(rocblas_handle h, rocblas_fill uplo, rocblas_amd_synth_dist dist,
                               rocblas_int k, uint64_t seed, double scale, double poison, T* A,
                               rocblas_int lda)
{
    if(!h)
        return rocblas_status_invalid_handle;
    if(uplo != rocblas_fill_upper && uplo != rocblas_fill_lower)
        return rocblas_status_invalid_value;
    if(dist != rocblas_amd_synth_uniform && dist != rocblas_amd_synth_normal)
        return rocblas_status_invalid_value;
    if(k < 0 || lda < k || lda < 1)
        return rocblas_status_invalid_size;
    if(!k)
        return rocblas_status_success;
    if(!A)
        return rocblas_status_invalid_pointer;
    DeviceGuard g(h->device);
    const dim3 grid((lda + 255) / 256, std::min(k, 65535));
    hipLaunchKernelGGL(synth_tri_kernel<T>, grid, dim3(256), 0, h->stream, A, k, lda,
                       uplo == rocblas_fill_upper, int(dist), stream_key(seed, 1), scale, poison);
    return launch_status();
}

template <typename T>
rocblas_status synth_rhs_entry(rocblas_handle h, rocblas_int m, rocblas_int n, uint64_t seed,
                               rocblas_int tag, T* B, rocblas_int ldb)
{
    if(!h)
        return rocblas_status_invalid_handle;
    if(m < 0 || n < 0 || ldb < m || ldb < 1)
        return rocblas_status_invalid_size;
    if(!m || !n)
        return rocblas_status_success;
    if(!B)
        return rocblas_status_invalid_pointer;
    DeviceGuard g(h->device);
    const dim3 grid((ldb + 255) / 256, std::min(n, 65535));
    hipLaunchKernelGGL(synth_rhs_kernel<T>, grid, dim3(256), 0, h->stream, B, m, n, ldb,
                       stream_key(seed, uint64_t(tag)));
    return launch_status();
}
} // namespace
} // namespace rbamd

using namespace rbamd;

extern "C" {
rocblas_status rocblas_amd_dsynth_tri(rocblas_handle         handle,
                                      rocblas_fill           uplo,
                                      rocblas_amd_synth_dist dist,
                                      rocblas_int            k,
                                      uint64_t               seed,
                                      double                 scale,
                                      double                 poison,
                                      double*                A,
                                      rocblas_int            lda)
try
{
    return synth_tri_entry<double>(handle, uplo, dist, k, seed, scale, poison, A, lda);
}
catch(...)
{
    return rocblas_status_internal_error;
}

rocblas_status rocblas_amd_ssynth_tri(rocblas_handle         handle,
                                      rocblas_fill           uplo,
                                      rocblas_amd_synth_dist dist,
                                      rocblas_int            k,
                                      uint64_t               seed,
                                      double                 scale,
                                      double                 poison,
                                      float*                 A,
                                      rocblas_int            lda)
try
{
    return synth_tri_entry<float>(handle, uplo, dist, k, seed, scale, poison, A, lda);
}
catch(...)
{
    return rocblas_status_internal_error;
}

rocblas_status rocblas_amd_dsynth_rhs(rocblas_handle handle,
                                      rocblas_int    m,
                                      rocblas_int    n,
                                      uint64_t       seed,
                                      rocblas_int    tag,
                                      double*        B,
                                      rocblas_int    ldb)
try
{
    return synth_rhs_entry<double>(handle, m, n, seed, tag, B, ldb);
}
catch(...)
{
    return rocblas_status_internal_error;
}

rocblas_status rocblas_amd_ssynth_rhs(rocblas_handle handle,
                                      rocblas_int    m,
                                      rocblas_int    n,
                                      uint64_t       seed,
                                      rocblas_int    tag,
                                      float*         B,
                                      rocblas_int    ldb)
try
{
    return synth_rhs_entry<float>(handle, m, n, seed, tag, B, ldb);
}
catch(...)
{
    return rocblas_status_internal_error;
}
}
