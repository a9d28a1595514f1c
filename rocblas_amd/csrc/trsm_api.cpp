// trsm_api.cpp -- extern "C" rocblas_{s,d}trsm{,_batched,_strided_batched}.
//
// Each entry point runs the reference's checks in the reference's order
// (library/src/blas3/rocblas_trsm.hpp:29-82 via
// library/src/blas3/rocblas_trsm_imp.hpp:54-157 and the batched /
// strided-batched *_imp.hpp files), then hands a TrsmProblem to the launcher
// in trsm_driver.hip.  No exception crosses the ABI.
//
// rocblas_trsm{,_batched,_strided_batched}_ex (library/src/blas3/
// rocblas_trsm.cpp:28-73 and the batched / strided-batched .cpp files) take the
// data type at run time and an optional invA of packed diagonal-block
// inverses.  The solve here forms no inverses, so a supplied invA is accepted
// and not read; the result is the same solve as the typed entry points.
//
// Every entry point logs its call per ROCBLAS_LAYER (logging.hpp) before the
// argument check, as the reference does; the _ex forms log as the typed
// routine they compute (the reference routes them through the same _impl).
#include "internal.hpp"
#include "logging.hpp"

#include <algorithm>
#include <new>

namespace rbamd
{
namespace
{

template <typename T>
rocblas_status trsm_entry(rocblas_handle    handle,
                          rocblas_side      side,
                          rocblas_fill      uplo,
                          rocblas_operation transA,
                          rocblas_diagonal  diag,
                          rocblas_int       m,
                          rocblas_int       n,
                          const T*          alpha,
                          BatchPtr<const T> A,
                          bool              A_null,
                          rocblas_int       lda,
                          BatchPtr<T>       B,
                          bool              B_null,
                          rocblas_int       ldb,
                          rocblas_int       batch_count)
{
    if(!handle)
        return rocblas_status_invalid_handle;

    // enums (rocblas_trsm.hpp:37-51)
    if(side != rocblas_side_left && side != rocblas_side_right)
        return rocblas_status_invalid_value;
    if(uplo != rocblas_fill_lower && uplo != rocblas_fill_upper)
        return rocblas_status_invalid_value;
    if(transA != rocblas_operation_none && transA != rocblas_operation_transpose
       && transA != rocblas_operation_conjugate_transpose)
        return rocblas_status_invalid_value;
    if(diag != rocblas_diagonal_non_unit && diag != rocblas_diagonal_unit)
        return rocblas_status_invalid_value;

    // sizes and leading dimensions (rocblas_trsm.hpp:53-56)
    const rocblas_int k = side == rocblas_side_left ? m : n;
    if(batch_count < 0 || m < 0 || n < 0 || lda < k || ldb < m)
        return rocblas_status_invalid_size;

    // quick return (rocblas_trsm.hpp:58-61)
    if(!m || !n || !batch_count)
        return handle->size_query ? rocblas_status_size_unchanged : rocblas_status_success;

    // pointers (rocblas_trsm.hpp:63-67)
    if(!handle->size_query)
    {
        if(B_null || !alpha
           || (handle->pointer_mode == rocblas_pointer_mode_host && *alpha != T(0) && A_null))
            return rocblas_status_invalid_pointer;
    }

    TrsmProblem<T> p;
    p.side  = side;
    p.uplo  = uplo;
    // conjugate transpose is transpose for real types
    // (library/src/blas3/rocblas_trsm_kernels.hpp:3625-3626)
    p.trans      = transA == rocblas_operation_none ? rocblas_operation_none
                                                    : rocblas_operation_transpose;
    p.diag       = diag;
    p.m          = m;
    p.n          = n;
    p.lda        = lda;
    p.ldb        = ldb;
    p.A          = A;
    p.B          = B;
    p.batch      = batch_count;
    p.alpha_host = T(1);
    p.alpha_dev  = nullptr;

    // device-memory size query: report, do no work
    // (library/src/blas3/rocblas_trsm_kernels.hpp:1898-1910)
    if(handle->size_query)
    {
        size_t need = roundup_ws(trsm_workspace_bytes(p));
        if(!need || need <= handle->query_size)
            return rocblas_status_size_unchanged;
        handle->query_size = need;
        return rocblas_status_size_increased;
    }

    if(handle->pointer_mode == rocblas_pointer_mode_host)
    {
        p.alpha_host = *alpha;
        // alpha == 0: B := 0, A not referenced (rocblas_trsm_imp.hpp:154-157)
        if(p.alpha_host == T(0))
            return zero_matrix_launch<T>(handle, m, n, B, ldb, batch_count);
    }
    else
    {
        // alpha is read on the device by every kernel, so the call never
        // synchronises and can be captured in a HIP graph.
        p.alpha_dev = alpha;
    }
    return trsm_launch<T>(handle, p);
}

// Leading dimensions beyond INT32_MAX (the reference's 64-bit substitution
// fallback, rocblas_trsm_kernels_64.cpp:150-215): such a matrix has few
// columns (A: k of them; B: n, or k for side right), so its used part is
// gathered into a compact workspace copy (leading dimension k, or the rows
// of the chunk), solved there by the 32-bit path and, for B, scattered
// back.  One instance at a time, right-hand sides in chunks that keep the
// copy of B under 2^28 elements.  The copies move O(size) data around the
// usual solve.
template <typename T>
rocblas_status trsm64_compact(rocblas_handle    handle,
                              rocblas_side      side,
                              rocblas_fill      uplo,
                              rocblas_operation transA,
                              rocblas_diagonal  diag,
                              int64_t           m,
                              int64_t           n,
                              const T*          alpha,
                              BatchPtr<const T> A,
                              bool              A_null,
                              int64_t           lda,
                              BatchPtr<T>       B,
                              bool              B_null,
                              int64_t           ldb,
                              int64_t           batch_count)
{
    constexpr int64_t i32max = 0x7fffffff;
    const bool        sideL  = side == rocblas_side_left;
    const int64_t     k      = sideL ? m : n;
    const int64_t     nv     = sideL ? n : m; // right-hand-side vectors
    const int64_t     other  = sideL ? m : n; // length of one vector
    const bool        cpA    = lda > i32max;
    const bool        cpB    = ldb > i32max;
    const int64_t     vstep  = std::max<int64_t>(
        1, std::min<int64_t>(int64_t(1) << 30, (int64_t(1) << 28) / other));
    const int64_t vc_max = std::min(vstep, nv);
    const size_t  a_els  = cpA ? size_t(k) * size_t(k) : 0;
    const size_t  b_els  = cpB ? size_t(vc_max) * size_t(other) : 0;
    const size_t  bytes  = roundup_ws(a_els * sizeof(T)) + roundup_ws(b_els * sizeof(T));
    if(handle->size_query)
    {
        if(!bytes || bytes <= handle->query_size)
            return rocblas_status_size_unchanged;
        handle->query_size = bytes;
        return rocblas_status_size_increased;
    }
    if(B_null || !alpha
       || (handle->pointer_mode == rocblas_pointer_mode_host && *alpha != T(0) && A_null))
        return rocblas_status_invalid_pointer;
    const bool skipA = A_null || (handle->pointer_mode == rocblas_pointer_mode_host
                                  && *alpha == T(0)); // A is not referenced

    DeviceGuard    dg(handle->device);
    WorkspaceLease lease(handle, bytes);
    if(!lease.ptr)
        return rocblas_status_memory_error;
    T* const Ac = static_cast<T*>(lease.ptr);
    T* const Bc = reinterpret_cast<T*>(static_cast<char*>(lease.ptr)
                                       + roundup_ws(a_els * sizeof(T)));
    for(int64_t b = 0; b < batch_count; ++b)
    {
        BatchPtr<const T> Ab = A.arr ? BatchPtr<const T>{A.arr + b, nullptr, 0, A.offset}
                                     : BatchPtr<const T>{nullptr, A.base, 0, A.offset + b * A.stride};
        BatchPtr<T> Bb = B.arr ? BatchPtr<T>{B.arr + b, nullptr, 0, B.offset}
                               : BatchPtr<T>{nullptr, B.base, 0, B.offset + b * B.stride};
        rocblas_status st;
        BatchPtr<const T> Au   = Ab;
        int               ldau = cpA ? int(k) : int(lda);
        if(cpA)
        {
            Au = BatchPtr<const T>{nullptr, Ac, 0, 0};
            if(!skipA
               && (st = copy_launch<T>(handle, int(k), int(k), Ab, lda,
                                       BatchPtr<T>{nullptr, Ac, 0, 0}, k, 1))
                      != rocblas_status_success)
                return st;
        }
        for(int64_t v0 = 0; v0 < nv; v0 += vstep)
        {
            const int64_t vc   = std::min(vstep, nv - v0);
            BatchPtr<T>   Bv   = Bb.shifted(sideL ? v0 * ldb : v0);
            const int     rows = int(sideL ? m : vc), cols = int(sideL ? vc : n);
            BatchPtr<T>   Bu   = Bv;
            int           ldbu = int(std::min(ldb, i32max));
            if(cpB)
            {
                Bu   = BatchPtr<T>{nullptr, Bc, 0, 0};
                ldbu = rows;
                if((st = copy_launch<T>(handle, rows, cols,
                                        BatchPtr<const T>{(const T* const*)Bv.arr, Bv.base,
                                                          Bv.stride, Bv.offset},
                                        ldb, Bu, rows, 1))
                   != rocblas_status_success)
                    return st;
            }
            st = trsm_entry<T>(handle, side, uplo, transA, diag, rows, cols, alpha, Au,
                               A_null && !cpA, ldau, Bu, false, ldbu, 1);
            if(st != rocblas_status_success)
                return st;
            if(cpB
               && (st = copy_launch<T>(handle, rows, cols,
                                       BatchPtr<const T>{nullptr, Bc, 0, 0}, rows, Bv, ldb, 1))
                      != rocblas_status_success)
                return st;
        }
    }
    return rocblas_status_success;
}

// ILP64 forms (library/src/src64/blas3/rocblas_trsm_64.cpp,
// rocblas_trsm_kernels_64.cpp:31-225): the same checks on 64-bit sizes, a
// triangle larger than INT32_MAX is invalid_size as in the reference, and the
// work is handed to the 32-bit path in chunks of right-hand sides and of the
// batch; leading dimensions beyond INT32_MAX go through trsm64_compact.
template <typename T>
rocblas_status trsm64_entry(rocblas_handle    handle,
                            rocblas_side      side,
                            rocblas_fill      uplo,
                            rocblas_operation transA,
                            rocblas_diagonal  diag,
                            int64_t           m,
                            int64_t           n,
                            const T*          alpha,
                            BatchPtr<const T> A,
                            bool              A_null,
                            int64_t           lda,
                            BatchPtr<T>       B,
                            bool              B_null,
                            int64_t           ldb,
                            int64_t           batch_count)
{
    constexpr int64_t i32max = 0x7fffffff;
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
    const bool    sideL = side == rocblas_side_left;
    const int64_t k     = sideL ? m : n;
    if(batch_count < 0 || m < 0 || n < 0 || lda < k || ldb < m)
        return rocblas_status_invalid_size;
    if(!m || !n || !batch_count)
        return handle->size_query ? rocblas_status_size_unchanged : rocblas_status_success;
    if(k > i32max)
        return rocblas_status_invalid_size; // exceeds practical memory (reference behaviour)
    if(lda > i32max || ldb > i32max)
        return trsm64_compact<T>(handle, side, uplo, transA, diag, m, n, alpha, A, A_null, lda, B,
                                 B_null, ldb, batch_count);
    // chunks: right-hand-side vectors (side L: columns, side R: rows) and batch
    const int64_t nv    = sideL ? n : m;
    const int64_t vstep = int64_t(1) << 30;
    const int64_t bstep = int64_t(1) << 30;
    for(int64_t b0 = 0; b0 < batch_count; b0 += bstep)
    {
        const int64_t bc = std::min(bstep, batch_count - b0);
        BatchPtr<const T> Ab = A.arr ? BatchPtr<const T>{A.arr + b0, nullptr, 0, A.offset}
                                     : BatchPtr<const T>{nullptr, A.base, A.stride,
                                                         A.offset + b0 * A.stride};
        BatchPtr<T> Bb = B.arr ? BatchPtr<T>{B.arr + b0, nullptr, 0, B.offset}
                               : BatchPtr<T>{nullptr, B.base, B.stride, B.offset + b0 * B.stride};
        for(int64_t v0 = 0; v0 < nv; v0 += vstep)
        {
            const int64_t vc  = std::min(vstep, nv - v0);
            BatchPtr<T>   Bc  = Bb.shifted(sideL ? v0 * ldb : v0);
            rocblas_status st = trsm_entry<T>(
                handle, side, uplo, transA, diag, sideL ? int(m) : int(vc), sideL ? int(vc) : int(n),
                alpha, Ab, A_null, int(lda), Bc, B_null, int(ldb), int(bc));
            if(st != rocblas_status_success)
                return st;
        }
    }
    return rocblas_status_success;
}

} // namespace
} // namespace rbamd

using namespace rbamd;

#define RBAMD_TRY_BEGIN try {
#define RBAMD_TRY_END                                 \
    }                                                 \
    catch(const std::bad_alloc&)                      \
    {                                                 \
        return rocblas_status_memory_error;           \
    }                                                 \
    catch(...)                                        \
    {                                                 \
        return rocblas_status_internal_error;         \
    }

#define RBAMD_TRSM_PLAIN(NAME, T)                                                               \
    rocblas_status NAME(rocblas_handle    handle,                                               \
                        rocblas_side      side,                                                 \
                        rocblas_fill      uplo,                                                 \
                        rocblas_operation transA,                                               \
                        rocblas_diagonal  diag,                                                 \
                        rocblas_int       m,                                                    \
                        rocblas_int       n,                                                    \
                        const T*          alpha,                                                \
                        const T*          A,                                                    \
                        rocblas_int       lda,                                                  \
                        T*                B,                                                    \
                        rocblas_int       ldb)                                                  \
    {                                                                                           \
        RBAMD_TRY_BEGIN                                                                         \
        log_trsm<T>(handle, TrsmForm::plain, false, side, uplo, transA, diag, m, n, alpha, A, lda, \
                    0, B, ldb, 0, 1);                                                           \
        return trsm_entry<T>(handle, side, uplo, transA, diag, m, n, alpha,                     \
                             BatchPtr<const T>{nullptr, A, 0, 0}, A == nullptr, lda,            \
                             BatchPtr<T>{nullptr, B, 0, 0}, B == nullptr, ldb, 1);              \
        RBAMD_TRY_END                                                                           \
    }

#define RBAMD_TRSM_BATCHED(NAME, T)                                                             \
    rocblas_status NAME(rocblas_handle    handle,                                               \
                        rocblas_side      side,                                                 \
                        rocblas_fill      uplo,                                                 \
                        rocblas_operation transA,                                               \
                        rocblas_diagonal  diag,                                                 \
                        rocblas_int       m,                                                    \
                        rocblas_int       n,                                                    \
                        const T*          alpha,                                                \
                        const T* const    A[],                                                  \
                        rocblas_int       lda,                                                  \
                        T* const          B[],                                                  \
                        rocblas_int       ldb,                                                  \
                        rocblas_int       batch_count)                                          \
    {                                                                                           \
        RBAMD_TRY_BEGIN                                                                         \
        log_trsm<T>(handle, TrsmForm::batched, false, side, uplo, transA, diag, m, n, alpha, A,  \
                    lda, 0, B, ldb, 0, batch_count);                                            \
        return trsm_entry<T>(handle, side, uplo, transA, diag, m, n, alpha,                     \
                             BatchPtr<const T>{A, nullptr, 0, 0}, A == nullptr, lda,            \
                             BatchPtr<T>{B, nullptr, 0, 0}, B == nullptr, ldb, batch_count);    \
        RBAMD_TRY_END                                                                           \
    }

#define RBAMD_TRSM_STRIDED(NAME, T)                                                             \
    rocblas_status NAME(rocblas_handle    handle,                                               \
                        rocblas_side      side,                                                 \
                        rocblas_fill      uplo,                                                 \
                        rocblas_operation transA,                                               \
                        rocblas_diagonal  diag,                                                 \
                        rocblas_int       m,                                                    \
                        rocblas_int       n,                                                    \
                        const T*          alpha,                                                \
                        const T*          A,                                                    \
                        rocblas_int       lda,                                                  \
                        rocblas_stride    stride_a,                                             \
                        T*                B,                                                    \
                        rocblas_int       ldb,                                                  \
                        rocblas_stride    stride_b,                                             \
                        rocblas_int       batch_count)                                          \
    {                                                                                           \
        RBAMD_TRY_BEGIN                                                                         \
        log_trsm<T>(handle, TrsmForm::strided, false, side, uplo, transA, diag, m, n, alpha, A,  \
                    lda, stride_a, B, ldb, stride_b, batch_count);                              \
        return trsm_entry<T>(handle, side, uplo, transA, diag, m, n, alpha,                     \
                             BatchPtr<const T>{nullptr, A, stride_a, 0}, A == nullptr, lda,     \
                             BatchPtr<T>{nullptr, B, stride_b, 0}, B == nullptr, ldb,           \
                             batch_count);                                                      \
        RBAMD_TRY_END                                                                           \
    }

#define RBAMD_TRSM_EX_DISPATCH(CALL)                   \
    switch(compute_type)                                \
    {                                                   \
    case rocblas_datatype_f32_r:                        \
    {                                                   \
        using T = float;                                \
        CALL;                                           \
    }                                                   \
    case rocblas_datatype_f64_r:                        \
    {                                                   \
        using T = double;                               \
        CALL;                                           \
    }                                                   \
    default:                                            \
        return rocblas_status_not_implemented;          \
    }

extern "C" {
rocblas_status rocblas_trsm_ex(rocblas_handle    handle,
                               rocblas_side      side,
                               rocblas_fill      uplo,
                               rocblas_operation transA,
                               rocblas_diagonal  diag,
                               rocblas_int       m,
                               rocblas_int       n,
                               const void*       alpha,
                               const void*       A,
                               rocblas_int       lda,
                               void*             B,
                               rocblas_int       ldb,
                               const void*       invA,
                               rocblas_int       invA_size,
                               rocblas_datatype  compute_type)
{
    (void)invA;
    (void)invA_size;
    RBAMD_TRY_BEGIN
    RBAMD_TRSM_EX_DISPATCH(
        log_trsm<T>(handle, TrsmForm::plain, false, side, uplo, transA, diag, m, n,
                    static_cast<const T*>(alpha), A, lda, 0, B, ldb, 0, 1);
        return trsm_entry<T>(handle, side, uplo, transA, diag, m, n,
                             static_cast<const T*>(alpha),
                             BatchPtr<const T>{nullptr, static_cast<const T*>(A), 0, 0},
                             A == nullptr, lda, BatchPtr<T>{nullptr, static_cast<T*>(B), 0, 0},
                             B == nullptr, ldb, 1))
    RBAMD_TRY_END
}

rocblas_status rocblas_trsm_batched_ex(rocblas_handle    handle,
                                       rocblas_side      side,
                                       rocblas_fill      uplo,
                                       rocblas_operation transA,
                                       rocblas_diagonal  diag,
                                       rocblas_int       m,
                                       rocblas_int       n,
                                       const void*       alpha,
                                       const void*       A,
                                       rocblas_int       lda,
                                       void*             B,
                                       rocblas_int       ldb,
                                       rocblas_int       batch_count,
                                       const void*       invA,
                                       rocblas_int       invA_size,
                                       rocblas_datatype  compute_type)
{
    (void)invA;
    (void)invA_size;
    RBAMD_TRY_BEGIN
    RBAMD_TRSM_EX_DISPATCH(
        log_trsm<T>(handle, TrsmForm::batched, false, side, uplo, transA, diag, m, n,
                    static_cast<const T*>(alpha), A, lda, 0, B, ldb, 0, batch_count);
        return trsm_entry<T>(handle, side, uplo, transA, diag, m, n,
                             static_cast<const T*>(alpha),
                             BatchPtr<const T>{static_cast<const T* const*>(A), nullptr, 0, 0},
                             A == nullptr, lda,
                             BatchPtr<T>{static_cast<T* const*>(B), nullptr, 0, 0},
                             B == nullptr, ldb, batch_count))
    RBAMD_TRY_END
}

rocblas_status rocblas_trsm_strided_batched_ex(rocblas_handle    handle,
                                               rocblas_side      side,
                                               rocblas_fill      uplo,
                                               rocblas_operation transA,
                                               rocblas_diagonal  diag,
                                               rocblas_int       m,
                                               rocblas_int       n,
                                               const void*       alpha,
                                               const void*       A,
                                               rocblas_int       lda,
                                               rocblas_stride    stride_A,
                                               void*             B,
                                               rocblas_int       ldb,
                                               rocblas_stride    stride_B,
                                               rocblas_int       batch_count,
                                               const void*       invA,
                                               rocblas_int       invA_size,
                                               rocblas_stride    stride_invA,
                                               rocblas_datatype  compute_type)
{
    (void)invA;
    (void)invA_size;
    (void)stride_invA;
    RBAMD_TRY_BEGIN
    RBAMD_TRSM_EX_DISPATCH(
        log_trsm<T>(handle, TrsmForm::strided, false, side, uplo, transA, diag, m, n,
                    static_cast<const T*>(alpha), A, lda, stride_A, B, ldb, stride_B,
                    batch_count);
        return trsm_entry<T>(handle, side, uplo, transA, diag, m, n,
                             static_cast<const T*>(alpha),
                             BatchPtr<const T>{nullptr, static_cast<const T*>(A), stride_A, 0},
                             A == nullptr, lda,
                             BatchPtr<T>{nullptr, static_cast<T*>(B), stride_B, 0}, B == nullptr,
                             ldb, batch_count))
    RBAMD_TRY_END
}

#define RBAMD_TRSM64_PLAIN(NAME, T)                                                              \
    rocblas_status NAME(rocblas_handle    handle,                                                \
                        rocblas_side      side,                                                  \
                        rocblas_fill      uplo,                                                  \
                        rocblas_operation transA,                                                \
                        rocblas_diagonal  diag,                                                  \
                        int64_t           m,                                                     \
                        int64_t           n,                                                     \
                        const T*          alpha,                                                 \
                        const T*          A,                                                     \
                        int64_t           lda,                                                   \
                        T*                B,                                                     \
                        int64_t           ldb)                                                   \
    {                                                                                            \
        RBAMD_TRY_BEGIN                                                                          \
        log_trsm<T>(handle, TrsmForm::plain, true, side, uplo, transA, diag, m, n, alpha, A, lda, \
                    0, B, ldb, 0, 1);                                                            \
        return trsm64_entry<T>(handle, side, uplo, transA, diag, m, n, alpha,                    \
                               BatchPtr<const T>{nullptr, A, 0, 0}, A == nullptr, lda,           \
                               BatchPtr<T>{nullptr, B, 0, 0}, B == nullptr, ldb, 1);             \
        RBAMD_TRY_END                                                                            \
    }

#define RBAMD_TRSM64_BATCHED(NAME, T)                                                            \
    rocblas_status NAME(rocblas_handle    handle,                                                \
                        rocblas_side      side,                                                  \
                        rocblas_fill      uplo,                                                  \
                        rocblas_operation transA,                                                \
                        rocblas_diagonal  diag,                                                  \
                        int64_t           m,                                                     \
                        int64_t           n,                                                     \
                        const T*          alpha,                                                 \
                        const T* const    A[],                                                   \
                        int64_t           lda,                                                   \
                        T* const          B[],                                                   \
                        int64_t           ldb,                                                   \
                        int64_t           batch_count)                                           \
    {                                                                                            \
        RBAMD_TRY_BEGIN                                                                          \
        log_trsm<T>(handle, TrsmForm::batched, true, side, uplo, transA, diag, m, n, alpha, A,    \
                    lda, 0, B, ldb, 0, batch_count);                                             \
        return trsm64_entry<T>(handle, side, uplo, transA, diag, m, n, alpha,                    \
                               BatchPtr<const T>{A, nullptr, 0, 0}, A == nullptr, lda,           \
                               BatchPtr<T>{B, nullptr, 0, 0}, B == nullptr, ldb, batch_count);   \
        RBAMD_TRY_END                                                                            \
    }

#define RBAMD_TRSM64_STRIDED(NAME, T)                                                            \
    rocblas_status NAME(rocblas_handle    handle,                                                \
                        rocblas_side      side,                                                  \
                        rocblas_fill      uplo,                                                  \
                        rocblas_operation transA,                                                \
                        rocblas_diagonal  diag,                                                  \
                        int64_t           m,                                                     \
                        int64_t           n,                                                     \
                        const T*          alpha,                                                 \
                        const T*          A,                                                     \
                        int64_t           lda,                                                   \
                        rocblas_stride    stride_a,                                              \
                        T*                B,                                                     \
                        int64_t           ldb,                                                   \
                        rocblas_stride    stride_b,                                              \
                        int64_t           batch_count)                                           \
    {                                                                                            \
        RBAMD_TRY_BEGIN                                                                          \
        log_trsm<T>(handle, TrsmForm::strided, true, side, uplo, transA, diag, m, n, alpha, A,    \
                    lda, stride_a, B, ldb, stride_b, batch_count);                               \
        return trsm64_entry<T>(handle, side, uplo, transA, diag, m, n, alpha,                    \
                               BatchPtr<const T>{nullptr, A, stride_a, 0}, A == nullptr, lda,    \
                               BatchPtr<T>{nullptr, B, stride_b, 0}, B == nullptr, ldb,          \
                               batch_count);                                                     \
        RBAMD_TRY_END                                                                            \
    }

RBAMD_TRSM64_PLAIN(rocblas_strsm_64, float)
RBAMD_TRSM64_PLAIN(rocblas_dtrsm_64, double)
RBAMD_TRSM64_BATCHED(rocblas_strsm_batched_64, float)
RBAMD_TRSM64_BATCHED(rocblas_dtrsm_batched_64, double)
RBAMD_TRSM64_STRIDED(rocblas_strsm_strided_batched_64, float)
RBAMD_TRSM64_STRIDED(rocblas_dtrsm_strided_batched_64, double)

RBAMD_TRSM_PLAIN(rocblas_strsm, float)
RBAMD_TRSM_PLAIN(rocblas_dtrsm, double)
RBAMD_TRSM_BATCHED(rocblas_strsm_batched, float)
RBAMD_TRSM_BATCHED(rocblas_dtrsm_batched, double)
RBAMD_TRSM_STRIDED(rocblas_strsm_strided_batched, float)
RBAMD_TRSM_STRIDED(rocblas_dtrsm_strided_batched, double)
}
