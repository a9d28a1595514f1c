// vbatched.hip -- variable-size batched triangular solve in one planning
// launch plus two persistent launches, graph-capturable.
//
//   rocblas_amd_{s,d}trsm_vbatched(handle, side, uplo, transA, diag,
//                                  m[], n[], alpha, A[], lda[], B[], ldb[], batch)
//
// m, n, lda, ldb and the A/B pointer arrays live in device memory, so a
// captured graph can be replayed with new sizes.  There is no reference
// counterpart: rocBLAS has no variable-size batched trsm, and its callers loop
// over rocblas_dtrsm (one to thirty launches per instance).
//
// Plan kernel (1 workgroup): every valid instance with k <= 512 becomes up to
// kVbPanelGroups work items, each a strided share of the instance's 16-vector
// panels; the items are counting-sorted by cost bucket
// (floor(log2(k^2 * vectors))), largest first.  Instances with k > 512 go to a
// second list.  Both work counters are reset with plain stores (no memset
// node in a graph).
//
// Panel kernel (persistent, 256 threads, one workgroup per CU): each
// workgroup takes the next item and solves its panels with the MFMA panel
// solve of panel_trsm.hpp.
//
// Wave kernel (persistent, one wavefront per workgroup) for k > 512: each
// wavefront takes the next instance and solves it on its own, 64
// right-hand-side vectors at a time (lane = vector): blocked, left-looking
// over NB-row blocks of the canonical triangle,
//    c_I = alpha * b_I - sum_{J<I} L_IJ y_J ;  y_I = L_II^{-1} c_I
// with each L block staged through LDS and read as LDS broadcasts.  NB is 16
// or 32 by the size-bucket table in vbatched_tuning.hpp (written by
// rocblas_amd/autotune_vbatched.py).
// Results do not depend on the order in which workers take items.
#include "panel_trsm.hpp"
#include "vbatched.hpp"
#include "vbatched_tuning.hpp"

#include "../../include/rocblas_amd_ext.h"

#include <atomic>
#include <mutex>
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
    // per-instance case flags and alpha (device arrays; null: flags / alpha
    // above for every instance)
    const int*         flags_arr;
    const T*           alpha_arr;
    int*               items;   // workspace: panel work items (inst * 8 + group, groups * batch)
    int*               order;   // workspace: k > 512 instances (batch ints)
    int*               counter; // workspace: [0] panel items taken, [1] large taken,
                                //            [2] panel items, [3] large instances
    int                nb16_max_k; // wave kernel: k <= this: 16-row blocks, else 32
    int                groups[kVbPanelBuckets]; // work items per instance of the panel
                                                // kernel (1..8), by size bucket of k
};

// Process-wide tuning values; start from the generated table and can be
// overridden at run time (rocblas_amd_vbatched_set_tuning) by the autotuner.
struct VbTuning
{
    int nb16_max_k       = kVbNb16MaxK;
    int waves_per_cu     = kVbWavesPerCu;
    int groups[kVbPanelBuckets];
    int panel_wgs_per_cu = 2;
    VbTuning()
    {
        for(int b = 0; b < kVbPanelBuckets; ++b)
            groups[b] = kVbPanelGroups[b];
    }
};
static VbTuning g_vb_tuning;

constexpr int VB_BUCKETS = 64;

// size bucket of the panel kernel's tuning table: k <= 64, <= 128, <= 256, <= 512
__host__ __device__ constexpr int vb_panel_bucket(int k)
{
    return k <= 64 ? 0 : k <= 128 ? 1 : k <= 256 ? 2 : 3;
}
static_assert(kVbPanelBuckets == 4, "four size buckets up to kPanelMaxK");

__device__ __forceinline__ int vb_cost_bucket(long long k, long long nv)
{
    const long long cost = k * k * (nv > 0 ? nv : 1);
    const int       lg   = 63 - __clzll((unsigned long long)(cost | 1));
    return VB_BUCKETS - 1 - min(lg, VB_BUCKETS - 1); // descending cost
}

template <typename T>
__device__ __forceinline__ int vb_flags(const VbArgs<T>& a, int i)
{
    return a.flags_arr ? a.flags_arr[i] : a.flags;
}

template <typename T>
__device__ __forceinline__ T vb_alpha(const VbArgs<T>& a, int i)
{
    return a.alpha_arr ? a.alpha_arr[i] : a.alpha_dev ? *a.alpha_dev : a.alpha_host;
}

// instance i: triangle size k, vectors nv, and whether it is solvable
template <typename T>
__device__ __forceinline__ bool vb_shape(const VbArgs<T>& a, int i, int& k, int& nv)
{
    const bool cols = vb_flags(a, i) & SF_COLS;
    const int  m = a.m[i], n = a.n[i];
    k  = cols ? m : n;
    nv = cols ? n : m;
    return m > 0 && n > 0 && a.lda[i] >= k && a.ldb[i] >= m;
}

template <typename T>
__device__ __forceinline__ int vb_groups(const VbArgs<T>& a, int k, int nv)
{
    const int panels = (nv + panel::PNC - 1) / panel::PNC;
    return min(panels, a.groups[vb_panel_bucket(k)]);
}

template <typename T>
__global__ void __launch_bounds__(1024) vb_plan_kernel(const VbArgs<T> a)
{
    __shared__ int hist[VB_BUCKETS];
    __shared__ int offs[VB_BUCKETS];
    __shared__ int nlarge;
    for(int i = threadIdx.x; i < VB_BUCKETS; i += blockDim.x)
        hist[i] = 0;
    if(threadIdx.x == 0)
        nlarge = 0;
    __syncthreads();
    for(int i = threadIdx.x; i < a.batch; i += blockDim.x)
    {
        int k, nv;
        if(!vb_shape(a, i, k, nv))
            continue;
        if(k <= kPanelMaxK)
        {
            const int G   = vb_groups(a, k, nv);
            const int per = ((nv + panel::PNC - 1) / panel::PNC + G - 1) / G;
            atomicAdd(&hist[vb_cost_bucket(k, min(nv, per * panel::PNC))], G);
        }
        else
            a.order[atomicAdd(&nlarge, 1)] = i;
    }
    __syncthreads();
    if(threadIdx.x == 0)
    {
        int t = 0;
        for(int b = 0; b < VB_BUCKETS; ++b)
        {
            offs[b] = t;
            t += hist[b];
        }
        a.counter[0] = 0;
        a.counter[1] = 0;
        a.counter[2] = t;
        a.counter[3] = nlarge;
    }
    __syncthreads();
    for(int i = threadIdx.x; i < a.batch; i += blockDim.x)
    {
        int k, nv;
        if(!vb_shape(a, i, k, nv) || k > kPanelMaxK)
            continue;
        const int G   = vb_groups(a, k, nv);
        const int per = ((nv + panel::PNC - 1) / panel::PNC + G - 1) / G;
        const int pos = atomicAdd(&offs[vb_cost_bucket(k, min(nv, per * panel::PNC))], G);
        for(int g = 0; g < G; ++g)
            a.items[pos + g] = i * 8 + g;
    }
}

// Panel kernel: persistent workgroups over the panel work items.
template <typename T>
__global__ void __launch_bounds__(panel::PNT) __attribute__((amdgpu_waves_per_eu(2)))
    vb_panel_kernel(const VbArgs<T> a)
{
    extern __shared__ __align__(16) unsigned char vb_smem[];
    __shared__ int item_s;
    const int      total = a.counter[2];
    if(int(blockIdx.x) >= total)
        return; // more workgroups than work items: the rest take them all
    for(;;)
    {
        if(threadIdx.x == 0)
            item_s = atomicAdd(&a.counter[0], 1);
        __syncthreads();
        // the work item and the instance's sizes are wave-uniform: keep them
        // in scalar registers (the panel solve's branches then never write
        // EXEC, which its DPP updates rely on)
        const int item = __builtin_amdgcn_readfirstlane(item_s);
        __syncthreads();
        if(item >= total)
            break;
        const int code = __builtin_amdgcn_readfirstlane(a.items[item]);
        const int inst = code / 8;
        const int g    = code % 8;
        int       k, nv;
        vb_shape(a, inst, k, nv);
        k  = __builtin_amdgcn_readfirstlane(k);
        nv = __builtin_amdgcn_readfirstlane(nv);
        const int G      = vb_groups(a, k, nv);
        const int panels = (nv + panel::PNC - 1) / panel::PNC;
        // per-instance values read from device arrays are uniform too
        const T*  Ap     = uniform_ptr(a.A[inst]);
        T*        Bp     = uniform_ptr(a.B[inst]);
        const T   alpha  = uniform_val(vb_alpha(a, inst));
        const int flags  = uniform_int(vb_flags(a, inst));
        const int lda    = uniform_int(a.lda[inst]);
        const int ldb    = uniform_int(a.ldb[inst]);
        for(int p = g; p < panels; p += G)
        {
            if(alpha == T(0))
                panel::panel_zero<T>(Bp, k, nv, ldb, flags, p * panel::PNC);
            else
                panel::panel_solve<T>(Ap, Bp, k, nv, lda, ldb, flags, alpha, p * panel::PNC,
                                      vb_smem);
        }
    }
}

// One instance, one chunk of up to 64 vectors, block size NB.
template <typename T, int NB>
__device__ void vb_solve_chunk(int flags, T* Ls, int K, int nvec, int c0, const T* Ap, int lda,
                               T* Bp, int ldb, T alpha)
{
    constexpr int LDL  = NB + 1; // LDS row stride of a staged block
    const int     lane = threadIdx.x;
    const bool    tA   = flags & SF_TA;
    const bool    rev  = flags & SF_REV;
    const bool    unit = flags & SF_UNIT;
    const bool    cols = flags & SF_COLS;
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

    const int  total = a.counter[3];
    // the first `total` wavefronts take every instance between them; the
    // rest leave without touching the shared counter
    if(int(blockIdx.x) >= total)
        return;
    for(;;)
    {
        if(threadIdx.x == 0)
            item_s = atomicAdd(&a.counter[1], 1);
        __syncthreads();
        const int item = item_s;
        __syncthreads();
        if(item >= total)
            break;
        const int inst  = a.order[item];
        const int flags = vb_flags(a, inst);
        const T   alpha = vb_alpha(a, inst);
        const bool cols = flags & SF_COLS;
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
                vb_solve_chunk<T, 16>(flags, Ls, K, nvec, c0, Ap, lda, Bp, ldb, alpha);
            else
                vb_solve_chunk<T, 32>(flags, Ls, K, nvec, c0, Ap, lda, Bp, ldb, alpha);
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
    const size_t lds = panel::lds_bytes<T>(kPanelMaxK);
    static const hipError_t opt = hipFuncSetAttribute(
        reinterpret_cast<const void*>(&vb_panel_kernel<T>),
        hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    (void)opt;
    (void)hipGetLastError();
    int gmax = 1;
    for(int b = 0; b < kVbPanelBuckets; ++b)
        gmax = std::max(gmax, a.groups[b]);
    const int wgs = int(std::min<int64_t>(int64_t(a.batch) * gmax,
                                          int64_t(h->num_cus) * g_vb_tuning.panel_wgs_per_cu));
    hipLaunchKernelGGL(vb_panel_kernel<T>, dim3(wgs), dim3(panel::PNT), lds, h->stream, a);
    if((st = launch_status()) != rocblas_status_success)
        return st;
    const int waves = std::min(a.batch, h->num_cus * g_vb_tuning.waves_per_cu);
    hipLaunchKernelGGL(vb_solve_kernel<T>, dim3(waves), dim3(WAVE), 0, h->stream, a);
    return launch_status();
}

// ---------------------------------------------------------------------------
// Workspace that needs no allocation (rocblas_amd_dtrsm_vbatched_noalloc):
// the plan's counters, large-instance list and work items for up to
// kVbNoAllocMaxBatch instances, in a file-scope device array.  Host code
// never passes the array's name to a kernel: its address on each device is
// looked up once (hipGetSymbolAddress, with that device current) and kept.
// ---------------------------------------------------------------------------
constexpr int    kVbNoAllocMaxBatch = 65536;
constexpr size_t kVbScratchBytes    = 64 + size_t(kVbNoAllocMaxBatch) * sizeof(int)
                                   + size_t(kVbNoAllocMaxBatch) * 8 * sizeof(int);
static_assert(kVbScratchBytes % 64 == 0, "scratch sections are 64-byte rounded");

static __device__ __align__(256) unsigned char scratch[kVbScratchBytes];

constexpr int               kVbMaxDevices = 64;
static std::atomic<void*>   g_scratch_ptr[kVbMaxDevices];
static std::mutex           g_scratch_mutex;

// Device address of `scratch` on `device` (current when called), or null.
static void* vb_scratch_ptr(int device)
{
    if(device < 0 || device >= kVbMaxDevices)
        return nullptr;
    void* p = g_scratch_ptr[device].load(std::memory_order_acquire);
    if(p)
        return p;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    p = g_scratch_ptr[device].load(std::memory_order_relaxed);
    if(!p)
    {
        if(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
        {
            (void)hipGetLastError();
            return nullptr;
        }
        g_scratch_ptr[device].store(p, std::memory_order_release);
    }
    return p;
}

namespace
{
// Workspace bytes of a batch: counters, then the large-instance list, then
// the panel work items (8 per instance).
inline size_t vb_workspace_bytes(rocblas_int batch_count)
{
    return 64 + roundup_ws(size_t(batch_count) * sizeof(int))
           + roundup_ws(size_t(batch_count) * 8 * sizeof(int));
}

template <typename T>
void vb_carve(VbArgs<T>& a, void* wsp)
{
    a.counter = static_cast<int*>(wsp);
    a.order   = reinterpret_cast<int*>(static_cast<char*>(wsp) + 64);
    a.items   = reinterpret_cast<int*>(static_cast<char*>(wsp) + 64
                                     + roundup_ws(size_t(a.batch) * sizeof(int)));
}

template <typename T>
rocblas_status vb_entry(bool               noalloc,
                        rocblas_handle     handle,
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
    if(batch_count < 0 || batch_count >= (1 << 28))
        return rocblas_status_invalid_size; // work items encode instance * 8 + group
    if(noalloc && batch_count > kVbNoAllocMaxBatch)
        return rocblas_status_invalid_size; // beyond the static scratch array
    const size_t ws = vb_workspace_bytes(batch_count);
    if(handle->size_query)
    {
        if(noalloc || !batch_count || roundup_ws(ws) <= handle->query_size)
            return rocblas_status_size_unchanged;
        handle->query_size = roundup_ws(ws);
        return rocblas_status_size_increased;
    }
    if(!batch_count)
        return rocblas_status_success;
    if(!m || !n || !lda || !ldb || !alpha || !A || !B)
        return rocblas_status_invalid_pointer;

    DeviceGuard    dg(handle->device);
    WorkspaceLease lease(handle, noalloc ? 0 : ws);
    void*          wsp = noalloc ? vb_scratch_ptr(handle->device) : lease.ptr;
    if(!wsp)
        return noalloc ? rocblas_status_internal_error : rocblas_status_memory_error;

    VbArgs<T> a;
    a.m          = m;
    a.n          = n;
    a.lda        = lda;
    a.ldb        = ldb;
    a.A          = A;
    a.B          = B;
    a.batch      = batch_count;
    a.flags      = trsm_case_flags(side, uplo, transA, diag);
    a.alpha_host = T(1);
    a.alpha_dev  = nullptr;
    a.flags_arr  = nullptr;
    a.alpha_arr  = nullptr;
    if(handle->pointer_mode == rocblas_pointer_mode_host)
        a.alpha_host = *alpha;
    else
        a.alpha_dev = alpha;
    a.nb16_max_k = g_vb_tuning.nb16_max_k;
    for(int b = 0; b < kVbPanelBuckets; ++b)
        a.groups[b] = g_vb_tuning.groups[b];
    vb_carve(a, wsp);
    return vbatched_launch<T>(handle, a);
}
} // namespace

int trsm_case_flags(rocblas_side side, rocblas_fill uplo, rocblas_operation trans,
                    rocblas_diagonal diag)
{
    const bool sideL  = side == rocblas_side_left;
    const bool tA     = sideL ? (trans != rocblas_operation_none)
                              : (trans == rocblas_operation_none);
    const bool upperA = uplo == rocblas_fill_upper;
    const bool lowerM = (!upperA) != tA;
    return (tA ? SF_TA : 0) | (upperA ? SF_UPPER : 0) | (lowerM ? 0 : SF_REV)
           | (diag == rocblas_diagonal_unit ? SF_UNIT : 0) | (sideL ? SF_COLS : 0);
}

template <typename T>
rocblas_status vbatched_cases(rocblas_handle     handle,
                              const rocblas_int* m,
                              const rocblas_int* n,
                              const rocblas_int* lda,
                              const rocblas_int* ldb,
                              const T* const*    A,
                              T* const*          B,
                              const int*         flags,
                              const T*           alpha,
                              rocblas_int        batch_count)
{
    if(batch_count < 0 || batch_count >= (1 << 28))
        return rocblas_status_invalid_size;
    if(!batch_count)
        return rocblas_status_success;
    const size_t ws = vb_workspace_bytes(batch_count);
    DeviceGuard    dg(handle->device);
    WorkspaceLease lease(handle, ws);
    if(!lease.ptr)
        return rocblas_status_memory_error;
    VbArgs<T> a{};
    a.m          = m;
    a.n          = n;
    a.lda        = lda;
    a.ldb        = ldb;
    a.A          = A;
    a.B          = B;
    a.batch      = batch_count;
    a.flags      = 0;
    a.alpha_host = T(1);
    a.alpha_dev  = nullptr;
    a.flags_arr  = flags;
    a.alpha_arr  = alpha;
    a.nb16_max_k = g_vb_tuning.nb16_max_k;
    for(int b = 0; b < kVbPanelBuckets; ++b)
        a.groups[b] = g_vb_tuning.groups[b];
    vb_carve(a, lease.ptr);
    return vbatched_launch<T>(handle, a);
}

template rocblas_status vbatched_cases<float>(rocblas_handle, const rocblas_int*,
                                              const rocblas_int*, const rocblas_int*,
                                              const rocblas_int*, const float* const*,
                                              float* const*, const int*, const float*,
                                              rocblas_int);
template rocblas_status vbatched_cases<double>(rocblas_handle, const rocblas_int*,
                                               const rocblas_int*, const rocblas_int*,
                                               const rocblas_int*, const double* const*,
                                               double* const*, const int*, const double*,
                                               rocblas_int);
} // namespace rbamd

using namespace rbamd;

#define RBAMD_VB(NAME, T, NOALLOC)                                                            \
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
        return vb_entry<T>(NOALLOC, handle, side, uplo, transA, diag, m, n, alpha, A, lda, B, \
                           ldb, batch_count);                                                 \
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
RBAMD_VB(rocblas_amd_strsm_vbatched, float, false)
RBAMD_VB(rocblas_amd_dtrsm_vbatched, double, false)
RBAMD_VB(rocblas_amd_dtrsm_vbatched_noalloc, double, true)

rocblas_status rocblas_amd_vbatched_set_bucket_tuning(rocblas_int bucket, rocblas_int panel_groups)
{
    if(bucket < 0 || bucket >= kVbPanelBuckets || panel_groups < 0 || panel_groups > 8)
        return rocblas_status_invalid_value;
    const VbTuning def;
    g_vb_tuning.groups[bucket] = panel_groups ? panel_groups : def.groups[bucket];
    return rocblas_status_success;
}

rocblas_status rocblas_amd_vbatched_set_tuning(rocblas_int nb16_max_k,
                                               rocblas_int waves_per_cu,
                                               rocblas_int panel_groups)
{
    if(nb16_max_k < 0 || waves_per_cu < 0 || waves_per_cu > 32 || panel_groups < 0
       || panel_groups > 8)
        return rocblas_status_invalid_value;
    const VbTuning def;
    g_vb_tuning.nb16_max_k   = nb16_max_k ? nb16_max_k : def.nb16_max_k;
    g_vb_tuning.waves_per_cu = waves_per_cu ? waves_per_cu : def.waves_per_cu;
    for(int b = 0; b < kVbPanelBuckets; ++b)
        g_vb_tuning.groups[b] = panel_groups ? panel_groups : def.groups[b];
    return rocblas_status_success;
}
}
