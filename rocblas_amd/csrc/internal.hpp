// internal.hpp -- handle state, workspace and launch helpers shared by the
// C-ABI entry points and the kernel launchers of librocblas_amd.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/rocblas_amd.h"

// Handle state.  Mirrors the subset of the reference's _rocblas_handle that the
// trsm path uses (library/src/include/handle.hpp:223-254, 319-337, 385-456,
// 485-495): bound device, stream, pointer mode, and a device workspace that is
// library-managed (grown on demand), user-managed (fixed size) or user-owned
// (caller memory), plus the device-memory size-query state.
enum class WorkspaceOwner
{
    library_managed,
    user_managed,
    user_owned
};

namespace rbamd
{
struct LogSink;
}

struct _rocblas_handle
{
    int                  device        = 0;
    int                  num_cus       = 256;
    size_t               max_lds       = 65536;
    hipStream_t          stream        = nullptr;
    rocblas_pointer_mode pointer_mode  = rocblas_pointer_mode_host;
    WorkspaceOwner       owner         = WorkspaceOwner::library_managed;
    void*                ws            = nullptr;
    size_t               ws_size       = 0;
    size_t               default_ws    = 0;
    bool                 ws_in_use     = false;
    bool                 size_query    = false;
    size_t               query_size    = 0;
    // completion counters of the chained solve (chain_trsm.hip): one per
    // panel, zero between launches; followed in the same allocation by the
    // tile counters of the split-K GEMM (kSplitCounters, zero between launches)
    void*                chain_flags   = nullptr;
    // host-visible word a kernel raises when one of its bounded waits gave
    // up; read and cleared by the next launching call and by
    // rocblas_amd_get_async_status
    unsigned int*        async_error   = nullptr;
    // test hook: chained launches and split-K GEMMs give up every wait after
    // a few polls
    bool                 chain_debug_timeout = false;
    // ROCBLAS_LAYER bits read at creation and the streams they opened
    int                  layer_mode     = 0;
    rbamd::LogSink*      log_trace_os   = nullptr;
    rbamd::LogSink*      log_bench_os   = nullptr;
    rbamd::LogSink*      log_profile_os = nullptr;
};

namespace rbamd
{
// Bytes of a handle's chain_flags buffer.
constexpr size_t kChainFlagsBytes = 4096;
// Tile counters of the split-K GEMM after them: one per (instance, tile) of a
// split launch.
constexpr int kSplitCounters = 4096;

inline unsigned* split_counters(rocblas_handle h)
{
    return h->chain_flags ? reinterpret_cast<unsigned*>(static_cast<char*>(h->chain_flags)
                                                        + kChainFlagsBytes)
                          : nullptr;
}

// Round a workspace request the way the reference does (64-byte chunks,
// library/src/include/handle.hpp:56-61).
inline size_t roundup_ws(size_t s)
{
    return (s + 63) / 64 * 64;
}

// Default workspace for gfx950: 128 MiB, the size the reference uses for its
// newest CDNA parts (library/src/include/handle.hpp:319-330).
constexpr size_t kDefaultWorkspace = size_t(128) << 20;

rocblas_status hip_to_status(hipError_t e);

// A wait of an earlier launch on this handle gave up (its result is wrong):
// clears the word and returns rocblas_status_internal_error, else success.
inline rocblas_status take_async_error(rocblas_handle h)
{
    if(h->async_error && __atomic_load_n(h->async_error, __ATOMIC_ACQUIRE) != 0u)
    {
        __atomic_store_n(h->async_error, 0u, __ATOMIC_RELEASE);
        return rocblas_status_internal_error;
    }
    return rocblas_status_success;
}

// Temporarily make the handle's device current (restored on destruction).
struct DeviceGuard
{
    int  saved  = -1;
    bool switched = false;
    explicit DeviceGuard(int dev)
    {
        if(hipGetDevice(&saved) == hipSuccess && saved != dev)
            switched = hipSetDevice(dev) == hipSuccess;
    }
    ~DeviceGuard()
    {
        if(switched)
            (void)hipSetDevice(saved);
    }
};

// Borrow `bytes` of the handle's workspace.  Library-managed workspaces grow
// (hipFree + hipMalloc) when nothing is outstanding; user-managed/owned ones
// never grow.  Returns nullptr when the request cannot be met.
void* workspace_acquire(rocblas_handle h, size_t bytes);
void  workspace_release(rocblas_handle h);

struct WorkspaceLease
{
    rocblas_handle h;
    void*          ptr = nullptr;
    WorkspaceLease(rocblas_handle handle, size_t bytes) : h(handle)
    {
        ptr = bytes ? workspace_acquire(h, bytes) : nullptr;
    }
    ~WorkspaceLease()
    {
        if(ptr)
            workspace_release(h);
    }
};

// Device-pointer-or-array batch accessor: instance b lives at
// (arr ? arr[b] : base + b*stride) + offset.
template <typename T>
struct BatchPtr
{
    T* const* arr;
    T*        base;
    int64_t   stride;
    int64_t   offset;

    __host__ __device__ BatchPtr shifted(int64_t off) const
    {
        return BatchPtr{arr, base, stride, offset + off};
    }
    __device__ T* get(int64_t b) const
    {
        return (arr ? arr[b] : base + b * stride) + offset;
    }
};

// Problem description after argument checks, with conjugate-transpose folded
// into transpose (real types) -- library/src/blas3/rocblas_trsm_kernels.hpp:3625-3626.
template <typename T>
struct TrsmProblem
{
    rocblas_side      side;
    rocblas_fill      uplo;
    rocblas_operation trans;
    rocblas_diagonal  diag;
    int               m, n;
    int               lda, ldb;
    BatchPtr<const T> A;
    BatchPtr<T>       B;
    int               batch;
    T                 alpha_host;     // alpha when it lives on the host
    const T*          alpha_dev;      // alpha when it lives on the device (else null)
};

template <typename T>
rocblas_status trsm_launch(rocblas_handle h, const TrsmProblem<T>& p);

// invA := op(A)^-1 for batch_count n x n triangles (no argument checks).  With
// scratch (n * n * batch_count elements) A is copied there first, which makes
// invA == A (in place) safe.  split_ws (split_bytes, from trtri_split_bytes)
// lets the doubling products with few tiles balance their K steps over the
// workgroups; nullptr runs them without.
template <typename T>
rocblas_status trtri_core(rocblas_handle    handle,
                          rocblas_fill      uplo,
                          rocblas_diagonal  diag,
                          int               n,
                          BatchPtr<const T> A,
                          int               lda,
                          BatchPtr<T>       invA,
                          int               ldinvA,
                          int               batch_count,
                          T*                scratch,
                          void*             split_ws,
                          size_t            split_bytes);

// Workspace bytes trtri_core's balanced products use for an order-n
// inversion of `batch` instances (groups_batched: the single-instance form
// batches a level's equal groups, which pointer arrays do not).
template <typename T>
size_t trtri_split_bytes(rocblas_handle h, int n, int batch, bool groups_batched);

template <typename T>
rocblas_status zero_matrix_launch(rocblas_handle h, int m, int n, BatchPtr<T> B, int ldb,
                                  int batch);

template <typename T>
size_t trsm_workspace_bytes(const TrsmProblem<T>& p);

template <typename T>
rocblas_status identity_launch(rocblas_handle h, int n, BatchPtr<T> X, int ldx, int batch);

template <typename T>
rocblas_status copy_launch(rocblas_handle h, int m, int n, BatchPtr<const T> S, int64_t lds,
                           BatchPtr<T> D, int64_t ldd, int batch);

} // namespace rbamd
