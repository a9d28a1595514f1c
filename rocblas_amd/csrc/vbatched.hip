// vbatched.hip -- variable-size batched triangular solve in one planning
// launch plus one persistent launch, graph-capturable.
//
//   rocblas_amd_{s,d}trsm_vbatched(handle, side, uplo, transA, diag,
//                                  m[], n[], alpha, A[], lda[], B[], ldb[], batch)
//
// m, n, lda, ldb and the A/B pointer arrays live in device memory, so a
// captured graph can be replayed with new sizes.  There is no reference
// counterpart: rocBLAS has no variable-size batched trsm, and its callers loop
// over rocblas_dtrsm (one to thirty launches per instance).
//
// Plan kernel (1 workgroup): every valid instance becomes up to
// kVbPanelGroups work items, each a strided share of the instance's 16-vector
// panels; the items are counting-sorted by cost bucket
// (floor(log2(k^2 * vectors))), largest first.  The work counter is reset
// with a plain store (no memset node in a graph).
//
// Panel kernel (persistent, 256 threads, two workgroups per CU): each
// workgroup takes the next item and solves its panels with the MFMA panel
// solve of panel_trsm.hpp, whose LDS does not depend on k, so every triangle
// size takes this one path.
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
    int*               counter; // workspace: [0] items taken, [1] items
    int                groups[kVbPanelBuckets]; // work items per instance of the panel
                                                // kernel (1..8), by size bucket of k
};

// Process-wide tuning values; start from the generated table and can be
// overridden at run time (rocblas_amd_vbatched_set_tuning) by the autotuner.
struct VbTuning
{
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

// size bucket of the panel kernel's tuning table: k <= 64, <= 128, <= 256, larger
__host__ __device__ constexpr int vb_panel_bucket(int k)
{
    return k <= 64 ? 0 : k <= 128 ? 1 : k <= 256 ? 2 : 3;
}
static_assert(kVbPanelBuckets == 4, "four size buckets");

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
    for(int i = threadIdx.x; i < VB_BUCKETS; i += blockDim.x)
        hist[i] = 0;
    __syncthreads();
    for(int i = threadIdx.x; i < a.batch; i += blockDim.x)
    {
        int k, nv;
        if(!vb_shape(a, i, k, nv))
            continue;
        const int G   = vb_groups(a, k, nv);
        const int per = ((nv + panel::PNC - 1) / panel::PNC + G - 1) / G;
        atomicAdd(&hist[vb_cost_bucket(k, min(nv, per * panel::PNC))], G);
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
        a.counter[1] = t;
    }
    __syncthreads();
    for(int i = threadIdx.x; i < a.batch; i += blockDim.x)
    {
        int k, nv;
        if(!vb_shape(a, i, k, nv))
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
    const int      total = a.counter[1];
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
    return launch_status();
}

// ---------------------------------------------------------------------------
// Workspace that needs no allocation (rocblas_amd_dtrsm_vbatched_noalloc):
// the plan's counters and work items for up to
// kVbNoAllocMaxBatch instances, in a file-scope device array.  Host code
// never passes the array's name to a kernel: its address on each device is
// looked up once (hipGetSymbolAddress, with that device current) and kept.
// ---------------------------------------------------------------------------
constexpr int    kVbNoAllocMaxBatch = 65536;
constexpr size_t kVbScratchBytes    = 64 + size_t(kVbNoAllocMaxBatch) * 8 * sizeof(int);
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
// Workspace bytes of a batch: counters, then the panel work items (8 per
// instance).
inline size_t vb_workspace_bytes(rocblas_int batch_count)
{
    return 64 + roundup_ws(size_t(batch_count) * 8 * sizeof(int));
}

template <typename T>
void vb_carve(VbArgs<T>& a, void* wsp)
{
    a.counter = static_cast<int*>(wsp);
    a.items   = reinterpret_cast<int*>(static_cast<char*>(wsp) + 64);
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

rocblas_status rocblas_amd_vbatched_set_tuning(rocblas_int panel_groups)
{
    if(panel_groups < 0 || panel_groups > 8)
        return rocblas_status_invalid_value;
    const VbTuning def;
    for(int b = 0; b < kVbPanelBuckets; ++b)
        g_vb_tuning.groups[b] = panel_groups ? panel_groups : def.groups[b];
    return rocblas_status_success;
}
}
