// gemm_api.cpp -- the trailing-update GEMM of the solve as a callable
// entry point: rocblas_amd_{s,d}gemm, with rocblas_{s,d}gemm's signature and
// argument rules (rocblas-functions.h:20047-20060).  Host pointer mode only
// reads alpha and beta on the host; device mode needs a synchronising read.
#include "device_common.hpp"

#include "../../include/rocblas_amd_ext.h"

#include <new>

namespace rbamd
{
namespace
{
template <typename T>
rocblas_status gemm_entry(rocblas_handle    h,
                          rocblas_operation transA,
                          rocblas_operation transB,
                          rocblas_int       m,
                          rocblas_int       n,
                          rocblas_int       k,
                          const T*          alpha,
                          const T*          A,
                          rocblas_int       lda,
                          const T*          B,
                          rocblas_int       ldb,
                          const T*          beta,
                          T*                C,
                          rocblas_int       ldc)
{
    if(!h)
        return rocblas_status_invalid_handle;
    auto op_ok = [](rocblas_operation o) {
        return o == rocblas_operation_none || o == rocblas_operation_transpose
               || o == rocblas_operation_conjugate_transpose;
    };
    if(!op_ok(transA) || !op_ok(transB))
        return rocblas_status_invalid_value;
    const bool tA = transA != rocblas_operation_none;
    const bool tB = transB != rocblas_operation_none;
    if(m < 0 || n < 0 || k < 0 || ldc < m || lda < (tA ? k : m) || ldb < (tB ? n : k)
       || lda < 1 || ldb < 1 || ldc < 1)
        return rocblas_status_invalid_size;
    if(!m || !n)
        return h->size_query ? rocblas_status_size_unchanged : rocblas_status_success;
    if(h->size_query)
        return rocblas_status_size_unchanged;
    if(!alpha || !beta || !C)
        return rocblas_status_invalid_pointer;
    T a = T(0), b = T(0);
    DeviceGuard dg(h->device);
    if(h->pointer_mode == rocblas_pointer_mode_host)
    {
        a = *alpha;
        b = *beta;
    }
    else
    {
        T          ab[2];
        hipError_t e = hipMemcpyAsync(&ab[0], alpha, sizeof(T), hipMemcpyDeviceToHost, h->stream);
        if(e == hipSuccess)
            e = hipMemcpyAsync(&ab[1], beta, sizeof(T), hipMemcpyDeviceToHost, h->stream);
        if(e == hipSuccess)
            e = hipStreamSynchronize(h->stream);
        if(e != hipSuccess)
            return hip_to_status(e);
        a = ab[0];
        b = ab[1];
    }
    if(k && a != T(0) && (!A || !B))
        return rocblas_status_invalid_pointer;
    GemmArgs<T> g{};
    g.M     = m;
    g.N     = n;
    g.K     = (a == T(0)) ? 0 : k;
    g.opA   = tA;
    g.opB   = tB;
    g.A     = BatchPtr<const T>{nullptr, A, 0, 0};
    g.lda   = lda;
    g.B     = BatchPtr<const T>{nullptr, B, 0, 0};
    g.ldb   = ldb;
    g.C     = BatchPtr<T>{nullptr, C, 0, 0};
    g.ldc   = ldc;
    g.alpha = a;
    g.beta  = b;
    g.batch = 1;
    return launch_gemm(h, g);
}
} // namespace
} // namespace rbamd

using namespace rbamd;

#define RBAMD_GEMM(NAME, T)                                                                    \
    rocblas_status NAME(rocblas_handle    handle,                                              \
                        rocblas_operation transA,                                              \
                        rocblas_operation transB,                                              \
                        rocblas_int       m,                                                   \
                        rocblas_int       n,                                                   \
                        rocblas_int       k,                                                   \
                        const T*          alpha,                                               \
                        const T*          A,                                                   \
                        rocblas_int       lda,                                                 \
                        const T*          B,                                                   \
                        rocblas_int       ldb,                                                 \
                        const T*          beta,                                                \
                        T*                C,                                                   \
                        rocblas_int       ldc)                                                 \
    try                                                                                        \
    {                                                                                          \
        return gemm_entry<T>(handle, transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, C, \
                             ldc);                                                             \
    }                                                                                          \
    catch(const std::bad_alloc&)                                                               \
    {                                                                                          \
        return rocblas_status_memory_error;                                                    \
    }                                                                                          \
    catch(...)                                                                                 \
    {                                                                                          \
        return rocblas_status_internal_error;                                                  \
    }

extern "C" {
RBAMD_GEMM(rocblas_amd_sgemm, float)
RBAMD_GEMM(rocblas_amd_dgemm, double)
}
