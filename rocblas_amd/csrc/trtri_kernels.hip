// trtri_kernels.hip -- kernels of the triangular inversion rocblas_?trtri.
//
// trtri_diag_kernel inverts the diagonal blocks (order <= KMAX) of a batch of
// triangles, one wavefront per (block, instance): the block is staged in LDS
// as a canonical lower triangle (an upper triangle is reversed: the inverse
// of the reversal is the reversal of the inverse) with its reciprocal
// diagonal, lane j solves L x = e_j with the unit vector generated in
// registers (nothing is read from invA), and the block of the inverse is
// written once, the opposite triangle of the block as zeros.  Entries past
// the block's order are 0 (1 on the diagonal), so the fully unrolled
// substitution carries no bounds.  Lane j's x_i stay exactly zero for i < j,
// so the block's strictly-opposite triangle is written as exact zeros.
//
// trtri_zero_opposite_kernel clears the strictly-opposite triangle of invA
// outside the diagonal blocks (the doubling steps use it as scratch).
//
// The reference inverts its diagonal blocks with trtri_small / the 16 -> 32
// doubling kernels of library/src/blas3/rocblas_trtri.hpp:65-253 and fills the
// opposite triangle with rocblas_trtri_fill (:393-426); these kernels are this
// library's own design for the same roles.
#include "device_common.hpp"
#include "internal.hpp"
#include "trtri_kernels.hpp"

namespace rbamd
{
namespace
{

template <typename T, int KMAX>
__global__ void __launch_bounds__(WAVE) trtri_diag_kernel(const TrtriDiagArgs<T> a)
{
    constexpr int LD = KMAX + 1; // LDS row stride (odd: rows fall in different banks)
    __shared__ T  Ls[KMAX * LD];
    const int     lane = threadIdx.x;
    const int     r0   = blockIdx.x * a.nb;
    const int     K    = min(a.nb, a.n - r0); // order of this diagonal block (uniform)
    const bool    up   = a.upper;

    for(int inst = blockIdx.y; inst < a.batch; inst += gridDim.y)
    {
        const T* Ap = a.A.get(inst) + (r0 + int64_t(r0) * a.lda);
        T*       Xp = a.X.get(inst) + (r0 + int64_t(r0) * a.ldx);

        // canonical (i, j) <- real (p, q): consecutive lanes walk down a
        // column of A (i, and so p, consecutive), so the loads coalesce
        for(int e = lane; e < KMAX * KMAX; e += WAVE)
        {
            const int i = e % KMAX, j = e / KMAX;
            const int p = up ? K - 1 - i : i;
            const int q = up ? K - 1 - j : j;
            T         v = T(0);
            if(i < K && j < K)
            {
                if(i == j)
                    v = a.unit ? T(1) : T(1) / gload(Ap + (p + int64_t(q) * a.lda));
                else if(j < i)
                    v = -gload(Ap + (p + int64_t(q) * a.lda));
            }
            else if(i == j)
                v = T(1);
            Ls[i * LD + j] = v;
        }
        __syncthreads();

        // column `lane` of the inverse: L x = e_lane, by columns of L
        T x[KMAX];
#pragma unroll
        for(int i = 0; i < KMAX; ++i)
            x[i] = i == lane ? T(1) : T(0);
#pragma unroll
        for(int j = 0; j < KMAX; ++j)
        {
            x[j] *= Ls[j * LD + j];
            // column j in chunks of CH entries; the empty asm after a chunk
            // takes its results and clobbers memory, so the compiler can
            // neither hoist later LDS reads above it nor sink the chunk's
            // multiply-adds below it (hoisting every read of the unrolled
            // solve would need thousands of live registers)
            constexpr int CH = 16;
#pragma unroll
            for(int i0 = j + 1; i0 < KMAX; i0 += CH)
            {
                T l[CH];
#pragma unroll
                for(int t = 0; t < CH; ++t)
                    if(i0 + t < KMAX)
                        l[t] = Ls[(i0 + t) * LD + j];
#pragma unroll
                for(int t = 0; t < CH; ++t)
                    if(i0 + t < KMAX)
                    {
                        x[i0 + t] = fma(l[t], x[j], x[i0 + t]);
                        asm volatile("" : "+v"(x[i0 + t]) : : "memory");
                    }
            }
        }
        __syncthreads(); // every read of L is done: reuse Ls for the result

        // stage X (canonical row i, column lane), then write the K x K block
        // coalesced: consecutive lanes take consecutive real rows
        if(lane < KMAX)
        {
#pragma unroll
            for(int i = 0; i < KMAX; ++i)
                Ls[i * LD + lane] = x[i];
        }
        __syncthreads();
        for(int e = lane; e < K * K; e += WAVE)
        {
            const int p = e % K, q = e / K; // real row and column of the block
            const int i = up ? K - 1 - p : p;
            const int j = up ? K - 1 - q : q;
            gstore(Xp + (p + int64_t(q) * a.ldx), Ls[i * LD + j]);
        }
        __syncthreads(); // LDS reuse by the next instance
    }
}

// Strictly-opposite triangle of invA outside the nb x nb diagonal blocks := 0.
template <typename T>
__global__ void __launch_bounds__(256) trtri_zero_opposite_kernel(const TrtriDiagArgs<T> a)
{
    const int64_t total = int64_t(a.n) * a.n;
    for(int inst = blockIdx.y; inst < a.batch; inst += gridDim.y)
    {
        T* Xp = a.X.get(inst);
        for(int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
            e += int64_t(gridDim.x) * blockDim.x)
        {
            const int  p   = int(e % a.n), q = int(e / a.n);
            const bool opp = a.upper ? p > q : p < q;
            if(opp && p / a.nb != q / a.nb)
                gstore(Xp + (p + int64_t(q) * a.ldx), T(0));
        }
    }
}

} // namespace

template <typename T>
rocblas_status trtri_diag_launch(rocblas_handle h, const TrtriDiagArgs<T>& a)
{
    const int  nblk = (a.n + a.nb - 1) / a.nb;
    const dim3 grid(unsigned(nblk), grid_batch(a.batch));
    DeviceGuard g(h->device);
    if(a.nb > kTrtriNB)
        return rocblas_status_internal_error;
    if(a.nb <= 16)
        hipLaunchKernelGGL((trtri_diag_kernel<T, 16>), grid, dim3(WAVE), 0, h->stream, a);
    else
        hipLaunchKernelGGL((trtri_diag_kernel<T, kTrtriNB>), grid, dim3(WAVE), 0, h->stream, a);
    return launch_status();
}

template <typename T>
rocblas_status trtri_zero_opposite_launch(rocblas_handle h, const TrtriDiagArgs<T>& a)
{
    const int64_t  total = int64_t(a.n) * a.n;
    const unsigned bx    = unsigned(std::min<int64_t>((total + 255) / 256, 4096));
    DeviceGuard    g(h->device);
    hipLaunchKernelGGL(trtri_zero_opposite_kernel<T>, dim3(bx, grid_batch(a.batch)), dim3(256), 0,
                       h->stream, a);
    return launch_status();
}

template rocblas_status trtri_diag_launch<float>(rocblas_handle, const TrtriDiagArgs<float>&);
template rocblas_status trtri_diag_launch<double>(rocblas_handle, const TrtriDiagArgs<double>&);
template rocblas_status trtri_zero_opposite_launch<float>(rocblas_handle,
                                                          const TrtriDiagArgs<float>&);
template rocblas_status trtri_zero_opposite_launch<double>(rocblas_handle,
                                                           const TrtriDiagArgs<double>&);

} // namespace rbamd
