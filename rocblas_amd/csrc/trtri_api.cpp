// trtri_api.cpp -- extern "C" rocblas_{s,d}trtri{,_batched,_strided_batched}.
//
// Checks and quick returns follow library/src/blas3/rocblas_trtri.hpp:31-62
// (the size query is answered first, as in library/src/blas3/rocblas_trtri.cpp:55-59).
// The inverse is computed as the triangular solve op(A) X = I on invA, with
// the same kernels as rocblas_?trsm (small substitution kernels and the
// recursive MFMA-updated driver): invA := I, then solve in place.  The
// triangle opposite `uplo` of invA comes out exactly zero ("may be cleared",
// rocblas-functions.h:19064-19071).  When invA is A itself (in-place call), A
// is first copied to the handle workspace.
#include "internal.hpp"

#include <new>

namespace rbamd
{
namespace
{

template <typename T>
rocblas_status trtri_entry(rocblas_handle    handle,
                           rocblas_fill      uplo,
                           rocblas_diagonal  diag,
                           rocblas_int       n,
                           BatchPtr<const T> A,
                           bool              A_null,
                           rocblas_int       lda,
                           BatchPtr<T>       invA,
                           bool              invA_null,
                           rocblas_int       ldinvA,
                           rocblas_int       batch_count,
                           bool              in_place)
{
    if(!handle)
        return rocblas_status_invalid_handle;

    const size_t ws_bytes
        = in_place && n > 0 && batch_count > 0 ? size_t(batch_count) * n * n * sizeof(T) : 0;
    if(handle->size_query)
    {
        const size_t need = roundup_ws(ws_bytes);
        if(!need || need <= handle->query_size)
            return rocblas_status_size_unchanged;
        handle->query_size = need;
        return rocblas_status_size_increased;
    }

    if(uplo != rocblas_fill_lower && uplo != rocblas_fill_upper)
        return rocblas_status_invalid_value;
    if(diag != rocblas_diagonal_non_unit && diag != rocblas_diagonal_unit)
        return rocblas_status_invalid_value;
    if(batch_count < 0 || n < 0 || lda < n || ldinvA < n)
        return rocblas_status_invalid_size;
    if(!n || !batch_count)
        return rocblas_status_success;
    if(A_null || invA_null)
        return rocblas_status_invalid_pointer;

    WorkspaceLease lease(handle, ws_bytes);
    if(in_place && !lease.ptr)
        return rocblas_status_memory_error;
    return trtri_core<T>(handle, uplo, diag, n, A, lda, invA, ldinvA, batch_count,
                         in_place ? static_cast<T*>(lease.ptr) : nullptr);
}

} // namespace

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
                          T*                scratch)
{
    BatchPtr<const T> Asrc = A;
    int               ldas = lda;
    if(scratch)
    {
        BatchPtr<T> W{nullptr, scratch, int64_t(n) * n, 0};
        rocblas_status st = copy_launch<T>(handle, n, n, A, lda, W, n, batch_count);
        if(st != rocblas_status_success)
            return st;
        Asrc = BatchPtr<const T>{nullptr, W.base, W.stride, 0};
        ldas = n;
    }

    rocblas_status st = identity_launch<T>(handle, n, invA, ldinvA, batch_count);
    if(st != rocblas_status_success)
        return st;

    TrsmProblem<T> p;
    p.side       = rocblas_side_left;
    p.uplo       = uplo;
    p.trans      = rocblas_operation_none;
    p.diag       = diag;
    p.m          = n;
    p.n          = n;
    p.lda        = ldas;
    p.ldb        = ldinvA;
    p.A          = Asrc;
    p.B          = invA;
    p.batch      = batch_count;
    p.alpha_host = T(1);
    p.alpha_dev  = nullptr;
    return trsm_launch<T>(handle, p);
}

template rocblas_status trtri_core<float>(rocblas_handle, rocblas_fill, rocblas_diagonal, int,
                                          BatchPtr<const float>, int, BatchPtr<float>, int, int,
                                          float*);
template rocblas_status trtri_core<double>(rocblas_handle, rocblas_fill, rocblas_diagonal, int,
                                           BatchPtr<const double>, int, BatchPtr<double>, int,
                                           int, double*);

} // namespace rbamd

using namespace rbamd;

#define RBAMD_TRY_BEGIN try {
#define RBAMD_TRY_END                         \
    }                                         \
    catch(const std::bad_alloc&)              \
    {                                         \
        return rocblas_status_memory_error;   \
    }                                         \
    catch(...)                                \
    {                                         \
        return rocblas_status_internal_error; \
    }

// rocblas-functions.h:19081-19097; library/src/blas3/rocblas_trtri.cpp:42-175
#define RBAMD_TRTRI_PLAIN(NAME, T)                                                            \
    rocblas_status NAME(rocblas_handle   handle,                                              \
                        rocblas_fill     uplo,                                                \
                        rocblas_diagonal diag,                                                \
                        rocblas_int      n,                                                   \
                        const T*         A,                                                   \
                        rocblas_int      lda,                                                 \
                        T*               invA,                                                \
                        rocblas_int      ldinvA)                                              \
    {                                                                                         \
        RBAMD_TRY_BEGIN                                                                       \
        return trtri_entry<T>(handle, uplo, diag, n, BatchPtr<const T>{nullptr, A, 0, 0},     \
                              A == nullptr, lda, BatchPtr<T>{nullptr, invA, 0, 0},            \
                              invA == nullptr, ldinvA, 1,                                     \
                              A != nullptr && (const void*)A == (const void*)invA);           \
        RBAMD_TRY_END                                                                         \
    }

// rocblas-functions.h:19159-19180; library/src/blas3/rocblas_trtri_batched.cpp
#define RBAMD_TRTRI_BATCHED(NAME, T)                                                          \
    rocblas_status NAME(rocblas_handle   handle,                                              \
                        rocblas_fill     uplo,                                                \
                        rocblas_diagonal diag,                                                \
                        rocblas_int      n,                                                   \
                        const T* const   A[],                                                 \
                        rocblas_int      lda,                                                 \
                        T* const         invA[],                                              \
                        rocblas_int      ldinvA,                                              \
                        rocblas_int      batch_count)                                         \
    {                                                                                         \
        RBAMD_TRY_BEGIN                                                                       \
        return trtri_entry<T>(handle, uplo, diag, n, BatchPtr<const T>{A, nullptr, 0, 0},     \
                              A == nullptr, lda, BatchPtr<T>{invA, nullptr, 0, 0},            \
                              invA == nullptr, ldinvA, batch_count,                           \
                              A != nullptr && (const void*)A == (const void*)invA);           \
        RBAMD_TRY_END                                                                         \
    }

// rocblas-functions.h:19249-19272; library/src/blas3/rocblas_trtri_strided_batched.cpp
#define RBAMD_TRTRI_STRIDED(NAME, T)                                                          \
    rocblas_status NAME(rocblas_handle   handle,                                              \
                        rocblas_fill     uplo,                                                \
                        rocblas_diagonal diag,                                                \
                        rocblas_int      n,                                                   \
                        const T*         A,                                                   \
                        rocblas_int      lda,                                                 \
                        rocblas_stride   stride_a,                                            \
                        T*               invA,                                                \
                        rocblas_int      ldinvA,                                              \
                        rocblas_stride   stride_invA,                                         \
                        rocblas_int      batch_count)                                         \
    {                                                                                         \
        RBAMD_TRY_BEGIN                                                                       \
        return trtri_entry<T>(handle, uplo, diag, n,                                          \
                              BatchPtr<const T>{nullptr, A, stride_a, 0}, A == nullptr, lda,  \
                              BatchPtr<T>{nullptr, invA, stride_invA, 0}, invA == nullptr,    \
                              ldinvA, batch_count,                                            \
                              A != nullptr && (const void*)A == (const void*)invA);           \
        RBAMD_TRY_END                                                                         \
    }

extern "C" {
RBAMD_TRTRI_PLAIN(rocblas_strtri, float)
RBAMD_TRTRI_PLAIN(rocblas_dtrtri, double)
RBAMD_TRTRI_BATCHED(rocblas_strtri_batched, float)
RBAMD_TRTRI_BATCHED(rocblas_dtrtri_batched, double)
RBAMD_TRTRI_STRIDED(rocblas_strtri_strided_batched, float)
RBAMD_TRTRI_STRIDED(rocblas_dtrtri_strided_batched, double)
}
