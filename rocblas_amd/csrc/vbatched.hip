// vbatched.hip -- variable-size batched triangular solve in one persistent
// launch (plus a one-workgroup planning launch), graph-capturable.
//
//   rocblas_amd_{s,d}trsm_vbatched(handle, side, uplo, transA, diag,
//                                  m[], n[], alpha, A[], lda[], B[], ldb[], batch)
//
// m, n, lda, ldb and the A/B pointer arrays live in device memory, so a
// captured graph can be replayed with new sizes.  There is no reference
// counterpart: rocBLAS has no variable-size batched trsm, and its callers loop
// over rocblas_dtrsm (one to thirty launches per instance).
//
// Plan kernel (1 workgroup): counting sort of the instances by cost bucket
// (floor(log2(k^2 * nvec))), largest first, into an order array in the handle
// workspace, and reset of the work counter (a plain store: no memset node in a
// graph).
//
// Solve kernel (persistent, one wavefront per workgroup): each wavefront
// takes the next instance from the counter and solves it on its own, 64
// right-hand-side vectors at a time (lane = vector), with the canonical
// forward-substitution form used by the small kernels (small_trsm.hpp):
// blocked, left-looking over NB-row blocks of the canonical triangle,
//    c_I = alpha * b_I - sum_{J<I} L_IJ y_J ;  y_I = L_II^{-1} c_I
// with each L block staged through LDS (coalesced along A's columns) and
// read as LDS broadcasts; solved blocks y_J are re-read from B (L2-resident).
// NB is 16 or 32 per instance, chosen by the size-bucket table in
// vbatched_tuning.hpp (written by rocblas_amd/autotune_vbatched.py).
// Results do not depend on the order in which wavefronts take instances.
#include "device_common.hpp"
#include "vbatched_tuning.hpp"

#include "../../include/rocblas_amd_ext.h"

#include <new>

namespace rbamd
{

template <typename T>
struct VbArgs
{
    const rocblas_int* m;
    const rocblas_int* n;
    const rocblas_int* lda;
    const rocblas_int* ldb;
    const T* const*    A;
    T* const*          B;
    int                batch;
    int                flags; // SF_TA / SF_UPPER / SF_REV / SF_UNIT / SF_COLS
    T                  alpha_host;
    const T*           alpha_dev;
    int*               order;   // workspace: batch ints
    int*               counter; // workspace: 1 int
    int                nb16_max_k; // k <= this: 16-row blocks, else 32
};

// Process-wide tuning values; start from the generated table and can be
// overridden at run time (rocblas_amd_vbatched_set_tuning) by the autotuner.
struct VbTuning
{
    int nb16_max_k   = kVbNb16MaxK;
    int waves_per_cu = kVbWavesPerCu;
};
static VbTuning g_vb_tuning;

constexpr int VB_BUCKETS = 64;

__device__ __forceinline__ int vb_cost_bucket(long long k, long long nv)
{
    const long long cost = k * k * (nv > 0 ? nv : 1);
    const int       lg   = 63 - __clzll((unsigned long long)(cost | 1));
    return VB_BUCKETS - 1 - min(lg, VB_BUCKETS - 1); // descending cost
}

template <typename T>
__global__ void __launch_bounds__(1024) vb_plan_kernel(const VbArgs<T> a)
{
    __shared__ int hist[VB_BUCKETS];
    __shared__ int offs[VB_BUCKETS];
    const bool     cols = a.flags & SF_COLS;
    for(int i = threadIdx.x; i < VB_BUCKETS; i += blockDim.x)
        hist[i] = 0;
    __syncthreads();
    for(int i = threadIdx.x; i < a.batch; i += blockDim.x)
    {
        const int k  = cols ? a.m[i] : a.n[i];
        const int nv = cols ? a.n[i] : a.m[i];
        atomicAdd(&hist[vb_cost_bucket(k, nv)], 1);
    }
    __syncthreads();
    if(threadIdx.x == 0)
    {
        int s = 0;
        for(int b = 0; b < VB_BUCKETS; ++b)
        {
            offs[b] = s;
            s += hist[b];
        }
        *a.counter = 0;
    }
    __syncthreads();
    for(int i = threadIdx.x; i < a.batch; i += blockDim.x)
    {
        const int k   = cols ? a.m[i] : a.n[i];
        const int nv  = cols ? a.n[i] : a.m[i];
        const int pos = atomicAdd(&offs[vb_cost_bucket(k, nv)], 1);
        a.order[pos]  = i;
    }
}

// One instance, one chunk of up to 64 vectors, block size NB.
template <typename T, int NB>
__device__ void vb_solve_chunk(const VbArgs<T>& a, T* Ls, int K, int nvec, int c0, const T* Ap,
                               int lda, T* Bp, int ldb, T alpha)
{
    constexpr int LDL  = NB + 1; // LDS row stride of a staged block
    const int     lane = threadIdx.x;
    const bool    tA   = a.flags & SF_TA;
    const bool    rev  = a.flags & SF_REV;
    const bool    unit = a.flags & SF_UNIT;
    const bool    cols = a.flags & SF_COLS;
    const int     vec  = c0 + lane;
    const bool    vok  = vec < nvec;
    const int     nblk = (K + NB - 1) / NB;

    // element (canonical row r, vector vec) of B
    auto bidx = [&](int r) -> int64_t {
        const int real = rev ? K - 1 - r : r;
        return cols ? int64_t(real) + int64_t(vec) * ldb : int64_t(vec) + int64_t(real) * ldb;
    };
    // stage canonical block L[I*NB.., J*NB..] (NB x NB, zero outside K) into Ls[i*LDL + j]
    auto stage = [&](int I, int J, bool diagblk) {
        for(int e = lane; e < NB * NB; e += WAVE)
        {
            // walk A's storage order so the loads are coalesced: A[p + q*lda]
            // has consecutive p for consecutive e
            const int sp = e % NB, sq = e / NB;
            int       i, j; // canonical block-local coordinates
            {
                // canonical (gi, gj) <-> real M indices (mi, mj) <-> A (p, q)
                // choose (i, j) such that p varies with sp
                if(!tA)
                {
                    // p = real row mi, contiguous in canonical i (forward or reversed)
                    i = sp;
                    j = sq;
                }
                else
                {
                    // p = real mj
                    i = sq;
                    j = sp;
                }
            }
            const int gi = I * NB + i, gj = J * NB + j;
            T         v  = T(0);
            if(gi < K && gj < K && (!diagblk || gi >= gj))
            {
                const int mi = rev ? K - 1 - gi : gi;
                const int mj = rev ? K - 1 - gj : gj;
                const int p  = tA ? mj : mi;
                const int q  = tA ? mi : mj;
                if(gi == gj)
                    v = unit ? T(1) : T(1) / gload(Ap + (p + int64_t(q) * lda));
                else
                    v = gload(Ap + (p + int64_t(q) * lda));
            }
            else if(diagblk && i == j)
                v = T(1);
            Ls[i * LDL + j] = v;
        }
    };

    for(int I = 0; I < nblk; ++I)
    {
        T c[NB];
#pragma unroll
        for(int r = 0; r < NB; ++r)
        {
            const int  gr = I * NB + r;
            const bool ok = vok && gr < K;
            c[r]          = ok ? alpha * gload(Bp + bidx(gr)) : T(0);
        }
        for(int J = 0; J < I; ++J)
        {
            stage(I, J, false);
            T y[NB];
#pragma unroll
            for(int s = 0; s < NB; ++s)
            {
                const int gs = J * NB + s;
                y[s]         = (vok && gs < K) ? gload(Bp + bidx(gs)) : T(0);
            }
            __syncthreads();
#pragma unroll
            for(int s = 0; s < NB; ++s)
            {
#pragma unroll
                for(int r = 0; r < NB; ++r)
                    c[r] = fma(-Ls[r * LDL + s], y[s], c[r]);
            }
            __syncthreads();
        }
        stage(I, I, true);
        __syncthreads();
#pragma unroll
        for(int j = 0; j < NB; ++j)
        {
            const T x = c[j] * Ls[j * LDL + j];
            c[j]      = x;
#pragma unroll
            for(int i = j + 1; i < NB; ++i)
                c[i] = fma(-Ls[i * LDL + j], x, c[i]);
        }
#pragma unroll
        for(int r = 0; r < NB; ++r)
        {
            const int gr = I * NB + r;
            if(vok && gr < K)
                gstore(Bp + bidx(gr), c[r]);
        }
        __syncthreads();
    }
}

template <typename T>
__global__ void __launch_bounds__(WAVE) vb_solve_kernel(const VbArgs<T> a)
{
    __shared__ T Ls[32 * 33];
    __shared__ int item_s;
    const T    alpha = a.alpha_dev ? *a.alpha_dev : a.alpha_host;
    const bool cols  = a.flags & SF_COLS;

    for(;;)
    {
        if(threadIdx.x == 0)
            item_s = atomicAdd(a.counter, 1);
        __syncthreads();
        const int item = item_s;
        __syncthreads();
        if(item >= a.batch)
            break;
        const int inst = a.order[item];
        const int m    = a.m[inst];
        const int n    = a.n[inst];
        const int K    = cols ? m : n;
        const int nvec = cols ? n : m;
        const int lda = a.lda[inst];
        const int ldb = a.ldb[inst];
        // empty or malformed instances (lda < k, ldb < m) are left untouched
        if(m <= 0 || n <= 0 || lda < K || ldb < m)
            continue;
        const T* Ap = a.A[inst];
        T*       Bp = a.B[inst];
        for(int c0 = 0; c0 < nvec; c0 += WAVE)
        {
            if(alpha == T(0))
            {
                // B := 0, A not referenced
                const int vec = c0 + threadIdx.x;
                if(vec < nvec)
                    for(int r = 0; r < K; ++r)
                        gstore(Bp + (cols ? int64_t(r) + int64_t(vec) * ldb
                                          : int64_t(vec) + int64_t(r) * ldb),
                               T(0));
                continue;
            }
            if(K <= a.nb16_max_k)
                vb_solve_chunk<T, 16>(a, Ls, K, nvec, c0, Ap, lda, Bp, ldb, alpha);
            else
                vb_solve_chunk<T, 32>(a, Ls, K, nvec, c0, Ap, lda, Bp, ldb, alpha);
        }
    }
}

template <typename T>
rocblas_status vbatched_launch(rocblas_handle h, VbArgs<T> a)
{
    DeviceGuard g(h->device);
    hipLaunchKernelGGL(vb_plan_kernel<T>, dim3(1), dim3(1024), 0, h->stream, a);
    rocblas_status st = launch_status();
    if(st != rocblas_status_success)
        return st;
    const int waves = std::min(a.batch, h->num_cus * g_vb_tuning.waves_per_cu);
    hipLaunchKernelGGL(vb_solve_kernel<T>, dim3(waves), dim3(WAVE), 0, h->stream, a);
    return launch_status();
}

namespace
{
template <typename T>
rocblas_status vb_entry(rocblas_handle     handle,
                        rocblas_side       side,
                        rocblas_fill       uplo,
                        rocblas_operation  transA,
                        rocblas_diagonal   diag,
                        const rocblas_int* m,
                        const rocblas_int* n,
                        const T*           alpha,
                        const T* const*    A,
                        const rocblas_int* lda,
                        T* const*          B,
                        const rocblas_int* ldb,
                        rocblas_int        batch_count)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    if(side != rocblas_side_left && side != rocblas_side_right)
        return rocblas_status_invalid_value;
    if(uplo != rocblas_fill_lower && uplo != rocblas_fill_upper)
        return rocblas_status_invalid_value;
    if(transA != rocblas_operation_none && transA != rocblas_operation_transpose
       && transA != rocblas_operation_conjugate_transpose)
        return rocblas_status_invalid_value;
    if(diag != rocblas_diagonal_non_unit && diag != rocblas_diagonal_unit)
        return rocblas_status_invalid_value;
    if(batch_count < 0)
        return rocblas_status_invalid_size;
    const size_t ws = roundup_ws(size_t(batch_count) * sizeof(int)) + 64;
    if(handle->size_query)
    {
        if(!batch_count || roundup_ws(ws) <= handle->query_size)
            return rocblas_status_size_unchanged;
        handle->query_size = roundup_ws(ws);
        return rocblas_status_size_increased;
    }
    if(!batch_count)
        return rocblas_status_success;
    if(!m || !n || !lda || !ldb || !alpha || !A || !B)
        return rocblas_status_invalid_pointer;

    WorkspaceLease lease(handle, ws);
    if(!lease.ptr)
        return rocblas_status_memory_error;

    const bool sideL  = side == rocblas_side_left;
    const bool tA     = sideL ? (transA != rocblas_operation_none)
                              : (transA == rocblas_operation_none);
    const bool upperA = uplo == rocblas_fill_upper;
    const bool lowerM = (!upperA) != tA;

    VbArgs<T> a;
    a.m          = m;
    a.n          = n;
    a.lda        = lda;
    a.ldb        = ldb;
    a.A          = A;
    a.B          = B;
    a.batch      = batch_count;
    a.flags      = (tA ? SF_TA : 0) | (upperA ? SF_UPPER : 0) | (lowerM ? 0 : SF_REV)
              | (diag == rocblas_diagonal_unit ? SF_UNIT : 0) | (sideL ? SF_COLS : 0);
    a.alpha_host = T(1);
    a.alpha_dev  = nullptr;
    if(handle->pointer_mode == rocblas_pointer_mode_host)
        a.alpha_host = *alpha;
    else
        a.alpha_dev = alpha;
    a.nb16_max_k = g_vb_tuning.nb16_max_k;
    a.counter    = static_cast<int*>(lease.ptr);
    a.order   = reinterpret_cast<int*>(static_cast<char*>(lease.ptr) + 64);
    return vbatched_launch<T>(handle, a);
}
} // namespace
} // namespace rbamd

using namespace rbamd;

#define RBAMD_VB(NAME, T)                                                                     \
    rocblas_status NAME(rocblas_handle     handle,                                            \
                        rocblas_side       side,                                              \
                        rocblas_fill       uplo,                                              \
                        rocblas_operation  transA,                                            \
                        rocblas_diagonal   diag,                                              \
                        const rocblas_int* m,                                                 \
                        const rocblas_int* n,                                                 \
                        const T*           alpha,                                             \
                        const T* const     A[],                                               \
                        const rocblas_int* lda,                                               \
                        T* const           B[],                                               \
                        const rocblas_int* ldb,                                               \
                        rocblas_int        batch_count)                                       \
    try                                                                                       \
    {                                                                                         \
        return vb_entry<T>(handle, side, uplo, transA, diag, m, n, alpha, A, lda, B, ldb,     \
                           batch_count);                                                      \
    }                                                                                         \
    catch(const std::bad_alloc&)                                                              \
    {                                                                                         \
        return rocblas_status_memory_error;                                                   \
    }                                                                                         \
    catch(...)                                                                                \
    {                                                                                         \
        return rocblas_status_internal_error;                                                 \
    }

extern "C" {
RBAMD_VB(rocblas_amd_strsm_vbatched, float)
RBAMD_VB(rocblas_amd_dtrsm_vbatched, double)

rocblas_status rocblas_amd_vbatched_set_tuning(rocblas_int nb16_max_k, rocblas_int waves_per_cu)
{
    if(nb16_max_k < 0 || waves_per_cu < 0 || waves_per_cu > 32)
        return rocblas_status_invalid_value;
    const VbTuning def;
    g_vb_tuning.nb16_max_k   = nb16_max_k ? nb16_max_k : def.nb16_max_k;
    g_vb_tuning.waves_per_cu = waves_per_cu ? waves_per_cu : def.waves_per_cu;
    return rocblas_status_success;
}
}
