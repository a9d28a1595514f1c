// internal_api.cpp -- the C++ template entry points rocSOLVER links against
// (include/rocblas_amd_internal.hpp): rocblas_internal_trsm{,_batched}_template,
// their workspace queries, their ILP64 _64 forms and
// rocblas_internal_trtri{,_batched}_template.
//
// As in the reference (library/src/blas3/rocblas_trsm_kernels.cpp:28-140,
// rocblas_trsm_kernels.hpp:3594-3626) these check no arguments beyond the
// handle: a zero batch, m or n returns success, conjugate transpose is
// transpose for real types, and alpha is read per the handle's pointer mode
// (on the device in device mode, so the call never synchronises).  Offsets and
// strides address each operand as base + offset + b * stride (or
// array[b] + offset for the pointer-array forms).
#include "internal.hpp"

#include <algorithm>
#include <cstdint>

#include "../../include/rocblas_amd_internal.hpp"

namespace rbamd
{
namespace
{

template <typename T>
rocblas_status internal_trsm(rocblas_handle    handle,
                             rocblas_side      side,
                             rocblas_fill      uplo,
                             rocblas_operation transA,
                             rocblas_diagonal  diag,
                             rocblas_int       m,
                             rocblas_int       n,
                             const T*          alpha,
                             BatchPtr<const T> A,
                             rocblas_int       lda,
                             BatchPtr<T>       B,
                             rocblas_int       ldb,
                             rocblas_int       batch_count)
{
    if(!handle)
        return rocblas_status_invalid_handle;
    if(batch_count <= 0 || m <= 0 || n <= 0)
        return rocblas_status_success;

    TrsmProblem<T> p;
    p.side       = side;
    p.uplo       = uplo;
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
    if(handle->pointer_mode == rocblas_pointer_mode_host)
    {
        p.alpha_host = *alpha;
        if(p.alpha_host == T(0))
            return zero_matrix_launch<T>(handle, m, n, B, ldb, batch_count);
    }
    else
        p.alpha_dev = alpha;
    return trsm_launch<T>(handle, p);
}

// ILP64 forms (library/src/src64/blas3/rocblas_trsm_kernels_64.cpp:31-225):
// a triangle beyond INT32_MAX is invalid_size as in the reference; otherwise
// the work goes to the 32-bit path in chunks of right-hand-side vectors and of
// the batch.  Leading dimensions beyond INT32_MAX are not implemented.
template <typename T>
rocblas_status internal_trsm_64(rocblas_handle    handle,
                                rocblas_side      side,
                                rocblas_fill      uplo,
                                rocblas_operation transA,
                                rocblas_diagonal  diag,
                                int64_t           m,
                                int64_t           n,
                                const T*          alpha,
                                BatchPtr<const T> A,
                                int64_t           lda,
                                BatchPtr<T>       B,
                                int64_t           ldb,
                                int64_t           batch_count)
{
    constexpr int64_t i32max = 0x7fffffff;
    if(!handle)
        return rocblas_status_invalid_handle;
    if(m <= 0 || n <= 0 || batch_count <= 0)
        return rocblas_status_success;
    const bool sideL = side == rocblas_side_left;
    if((sideL ? m : n) > i32max)
        return rocblas_status_invalid_size; // exceeds practical memory
    if(lda > i32max || ldb > i32max)
        return rocblas_status_not_implemented;
    const int64_t nv   = sideL ? n : m;
    const int64_t step = int64_t(1) << 30;
    for(int64_t b0 = 0; b0 < batch_count; b0 += step)
    {
        const int64_t     bc = std::min(step, batch_count - b0);
        BatchPtr<const T> Ab = A.arr ? BatchPtr<const T>{A.arr + b0, nullptr, 0, A.offset}
                                     : BatchPtr<const T>{nullptr, A.base, A.stride,
                                                         A.offset + b0 * A.stride};
        BatchPtr<T> Bb = B.arr ? BatchPtr<T>{B.arr + b0, nullptr, 0, B.offset}
                               : BatchPtr<T>{nullptr, B.base, B.stride, B.offset + b0 * B.stride};
        for(int64_t v0 = 0; v0 < nv; v0 += step)
        {
            const int64_t        vc = std::min(step, nv - v0);
            const rocblas_status st = internal_trsm<T>(
                handle, side, uplo, transA, diag, sideL ? int(m) : int(vc),
                sideL ? int(vc) : int(n), alpha, Ab, int(lda), Bb.shifted(sideL ? v0 * ldb : v0),
                int(ldb), int(bc));
            if(st != rocblas_status_success)
                return st;
        }
    }
    return rocblas_status_success;
}

rocblas_status no_workspace(size_t* w_x_tmp_size,
                            size_t* w_x_tmp_arr_size,
                            size_t* w_invA_size,
                            size_t* w_invA_arr_size,
                            size_t* w_x_tmp_size_backup)
{
    if(!w_x_tmp_size || !w_x_tmp_arr_size || !w_invA_size || !w_invA_arr_size
       || !w_x_tmp_size_backup)
        return rocblas_status_invalid_pointer;
    *w_x_tmp_size        = 0;
    *w_x_tmp_arr_size    = 0;
    *w_invA_size         = 0;
    *w_invA_arr_size     = 0;
    *w_x_tmp_size_backup = 0;
    return rocblas_status_continue;
}

} // namespace
} // namespace rbamd

using namespace rbamd;

template <typename T>
rocblas_status rocblas_internal_trsm_workspace_size(rocblas_side,
                                                    rocblas_operation,
                                                    rocblas_int,
                                                    rocblas_int,
                                                    rocblas_int,
                                                    rocblas_int,
                                                    size_t* w_x_tmp_size,
                                                    size_t* w_x_tmp_arr_size,
                                                    size_t* w_invA_size,
                                                    size_t* w_invA_arr_size,
                                                    size_t* w_x_tmp_size_backup)
{
    return no_workspace(w_x_tmp_size, w_x_tmp_arr_size, w_invA_size, w_invA_arr_size,
                        w_x_tmp_size_backup);
}

template <typename T>
rocblas_status rocblas_internal_trsm_batched_workspace_size(rocblas_side,
                                                            rocblas_operation,
                                                            rocblas_int,
                                                            rocblas_int,
                                                            rocblas_int,
                                                            rocblas_int,
                                                            size_t* w_x_tmp_size,
                                                            size_t* w_x_tmp_arr_size,
                                                            size_t* w_invA_size,
                                                            size_t* w_invA_arr_size,
                                                            size_t* w_x_tmp_size_backup)
{
    return no_workspace(w_x_tmp_size, w_x_tmp_arr_size, w_invA_size, w_invA_arr_size,
                        w_x_tmp_size_backup);
}

template <typename T>
rocblas_status rocblas_internal_trsm_template(rocblas_handle    handle,
                                              rocblas_side      side,
                                              rocblas_fill      uplo,
                                              rocblas_operation transA,
                                              rocblas_diagonal  diag,
                                              rocblas_int       m,
                                              rocblas_int       n,
                                              const T*          alpha,
                                              const T*          A,
                                              rocblas_stride    offset_A,
                                              rocblas_int       lda,
                                              rocblas_stride    stride_A,
                                              T*                B,
                                              rocblas_stride    offset_B,
                                              rocblas_int       ldb,
                                              rocblas_stride    stride_B,
                                              rocblas_int       batch_count,
                                              bool,
                                              void*,
                                              void*,
                                              void*,
                                              void*,
                                              const T*,
                                              rocblas_int,
                                              rocblas_stride,
                                              rocblas_stride)
try
{
    return internal_trsm<T>(handle, side, uplo, transA, diag, m, n, alpha,
                            BatchPtr<const T>{nullptr, A, stride_A, offset_A}, lda,
                            BatchPtr<T>{nullptr, B, stride_B, offset_B}, ldb, batch_count);
}
catch(...)
{
    return rocblas_status_internal_error;
}

template <typename T>
rocblas_status rocblas_internal_trsm_batched_template(rocblas_handle    handle,
                                                      rocblas_side      side,
                                                      rocblas_fill      uplo,
                                                      rocblas_operation transA,
                                                      rocblas_diagonal  diag,
                                                      rocblas_int       m,
                                                      rocblas_int       n,
                                                      const T*          alpha,
                                                      const T* const*   A,
                                                      rocblas_stride    offset_A,
                                                      rocblas_int       lda,
                                                      rocblas_stride,
                                                      T* const*      B,
                                                      rocblas_stride offset_B,
                                                      rocblas_int    ldb,
                                                      rocblas_stride,
                                                      rocblas_int batch_count,
                                                      bool,
                                                      void*,
                                                      void*,
                                                      void*,
                                                      void*,
                                                      const T* const*,
                                                      rocblas_int,
                                                      rocblas_stride,
                                                      rocblas_stride)
try
{
    return internal_trsm<T>(handle, side, uplo, transA, diag, m, n, alpha,
                            BatchPtr<const T>{A, nullptr, 0, offset_A}, lda,
                            BatchPtr<T>{B, nullptr, 0, offset_B}, ldb, batch_count);
}
catch(...)
{
    return rocblas_status_internal_error;
}

template <typename T>
rocblas_status rocblas_internal_trsm_workspace_size_64(rocblas_side,
                                                       rocblas_operation,
                                                       int64_t,
                                                       int64_t,
                                                       int64_t,
                                                       int64_t,
                                                       int64_t,
                                                       int64_t,
                                                       size_t* w_x_tmp_size,
                                                       size_t* w_x_tmp_arr_size,
                                                       size_t* w_invA_size,
                                                       size_t* w_invA_arr_size,
                                                       size_t* w_x_tmp_size_backup)
{
    return no_workspace(w_x_tmp_size, w_x_tmp_arr_size, w_invA_size, w_invA_arr_size,
                        w_x_tmp_size_backup);
}

template <typename T>
rocblas_status rocblas_internal_trsm_batched_workspace_size_64(rocblas_side,
                                                               rocblas_operation,
                                                               int64_t,
                                                               int64_t,
                                                               int64_t,
                                                               int64_t,
                                                               int64_t,
                                                               int64_t,
                                                               size_t* w_x_tmp_size,
                                                               size_t* w_x_tmp_arr_size,
                                                               size_t* w_invA_size,
                                                               size_t* w_invA_arr_size,
                                                               size_t* w_x_tmp_size_backup)
{
    return no_workspace(w_x_tmp_size, w_x_tmp_arr_size, w_invA_size, w_invA_arr_size,
                        w_x_tmp_size_backup);
}

// The most any call with m <= M, n <= N needs: nothing, so the reference's
// own check (every smaller size fits within the maximum) holds trivially.
template <typename T>
rocblas_status rocblas_internal_trsm_workspace_max_size_64(rocblas_side,
                                                           int64_t,
                                                           int64_t,
                                                           int64_t,
                                                           size_t* w_x_tmp_size,
                                                           size_t* w_invA_size,
                                                           size_t* w_x_tmp_size_backup)
{
    if(!w_x_tmp_size || !w_invA_size || !w_x_tmp_size_backup)
        return rocblas_status_invalid_pointer;
    *w_x_tmp_size        = 0;
    *w_invA_size         = 0;
    *w_x_tmp_size_backup = 0;
    return rocblas_status_success;
}

template <typename T>
rocblas_status rocblas_internal_trsm_batched_workspace_max_size_64(rocblas_side,
                                                                   int64_t,
                                                                   int64_t,
                                                                   int64_t,
                                                                   size_t* w_x_tmp_size,
                                                                   size_t* w_x_tmp_arr_size,
                                                                   size_t* w_invA_size,
                                                                   size_t* w_invA_arr_size,
                                                                   size_t* w_x_tmp_size_backup)
{
    const rocblas_status st = no_workspace(w_x_tmp_size, w_x_tmp_arr_size, w_invA_size,
                                           w_invA_arr_size, w_x_tmp_size_backup);
    return st == rocblas_status_continue ? rocblas_status_success : st;
}

template <typename T>
rocblas_status rocblas_internal_trsm_template_64(rocblas_handle    handle,
                                                 rocblas_side      side,
                                                 rocblas_fill      uplo,
                                                 rocblas_operation transA,
                                                 rocblas_diagonal  diag,
                                                 int64_t           m,
                                                 int64_t           n,
                                                 const T*          alpha,
                                                 const T*          A,
                                                 rocblas_stride    offset_A,
                                                 int64_t           lda,
                                                 rocblas_stride    stride_A,
                                                 T*                B,
                                                 rocblas_stride    offset_B,
                                                 int64_t           ldb,
                                                 rocblas_stride    stride_B,
                                                 int64_t           batch_count,
                                                 bool,
                                                 void*,
                                                 void*,
                                                 void*,
                                                 void*,
                                                 const T*,
                                                 int64_t,
                                                 rocblas_stride,
                                                 rocblas_stride)
try
{
    return internal_trsm_64<T>(handle, side, uplo, transA, diag, m, n, alpha,
                               BatchPtr<const T>{nullptr, A, stride_A, offset_A}, lda,
                               BatchPtr<T>{nullptr, B, stride_B, offset_B}, ldb, batch_count);
}
catch(...)
{
    return rocblas_status_internal_error;
}

template <typename T>
rocblas_status rocblas_internal_trsm_batched_template_64(rocblas_handle    handle,
                                                         rocblas_side      side,
                                                         rocblas_fill      uplo,
                                                         rocblas_operation transA,
                                                         rocblas_diagonal  diag,
                                                         int64_t           m,
                                                         int64_t           n,
                                                         const T*          alpha,
                                                         const T* const*   A,
                                                         rocblas_stride    offset_A,
                                                         int64_t           lda,
                                                         rocblas_stride,
                                                         T* const*      B,
                                                         rocblas_stride offset_B,
                                                         int64_t        ldb,
                                                         rocblas_stride,
                                                         int64_t batch_count,
                                                         bool,
                                                         void*,
                                                         void*,
                                                         void*,
                                                         void*,
                                                         const T* const*,
                                                         int64_t,
                                                         rocblas_stride,
                                                         rocblas_stride)
try
{
    return internal_trsm_64<T>(handle, side, uplo, transA, diag, m, n, alpha,
                               BatchPtr<const T>{A, nullptr, 0, offset_A}, lda,
                               BatchPtr<T>{B, nullptr, 0, offset_B}, ldb, batch_count);
}
catch(...)
{
    return rocblas_status_internal_error;
}

size_t rocblas_internal_trtri_temp_elements(rocblas_int n, rocblas_int batch_count)
{
    // scratch for an in-place call: one n x n copy of A per matrix
    return n > 0 && batch_count > 0 ? size_t(n) * size_t(n) * size_t(batch_count) : 0;
}

template <typename T>
rocblas_status rocblas_internal_trtri_template(rocblas_handle   handle,
                                               rocblas_fill     uplo,
                                               rocblas_diagonal diag,
                                               rocblas_int      n,
                                               const T*         A,
                                               rocblas_stride   offset_A,
                                               rocblas_int      lda,
                                               rocblas_stride   stride_A,
                                               rocblas_stride   sub_stride_A,
                                               T*               invA,
                                               rocblas_stride   offset_invA,
                                               rocblas_int      ldinvA,
                                               rocblas_stride   stride_invA,
                                               rocblas_stride   sub_stride_invA,
                                               rocblas_int      batch_count,
                                               rocblas_int      sub_batch_count,
                                               T*               w_C_tmp)
try
{
    if(!handle)
        return rocblas_status_invalid_handle;
    if(n <= 0 || batch_count <= 0 || sub_batch_count <= 0)
        return rocblas_status_success;
    const bool in_place = static_cast<const void*>(A + offset_A)
                          == static_cast<const void*>(invA + offset_invA);
    if(in_place && !w_C_tmp)
        return rocblas_status_invalid_pointer;
    // the split products' partial tiles, from the handle's workspace when it
    // has room (without, the products run unsplit)
    const size_t split_bytes
        = std::max({trtri_split_bytes<T>(handle, n, batch_count, true),
                    trtri_split_bytes<T>(handle, n, sub_batch_count, true),
                    trtri_split_bytes<T>(handle, n, 1, true)});
    WorkspaceLease lease(handle, split_bytes);
    void* const    sw = lease.ptr;
    const size_t   sb = lease.ptr ? split_bytes : 0;
    if(sub_batch_count == 1)
        return trtri_core<T>(handle, uplo, diag, n,
                             BatchPtr<const T>{nullptr, A, stride_A, offset_A}, lda,
                             BatchPtr<T>{nullptr, invA, stride_invA, offset_invA}, ldinvA,
                             batch_count, in_place ? w_C_tmp : nullptr, sw, sb);
    for(rocblas_int b = 0; b < batch_count; ++b)
    {
        const rocblas_stride oa = offset_A + b * stride_A, oi = offset_invA + b * stride_invA;
        rocblas_status       st = rocblas_status_success;
        if(!in_place)
            st = trtri_core<T>(handle, uplo, diag, n,
                               BatchPtr<const T>{nullptr, A, sub_stride_A, oa}, lda,
                               BatchPtr<T>{nullptr, invA, sub_stride_invA, oi}, ldinvA,
                               sub_batch_count, nullptr, sw, sb);
        else
            // in place, one matrix at a time through the n x n scratch
            for(rocblas_int s = 0; s < sub_batch_count && st == rocblas_status_success; ++s)
                st = trtri_core<T>(handle, uplo, diag, n,
                                   BatchPtr<const T>{nullptr, A, 0, oa + s * sub_stride_A}, lda,
                                   BatchPtr<T>{nullptr, invA, 0, oi + s * sub_stride_invA},
                                   ldinvA, 1, w_C_tmp, sw, sb);
        if(st != rocblas_status_success)
            return st;
    }
    return rocblas_status_success;
}
catch(...)
{
    return rocblas_status_internal_error;
}

template <typename T>
rocblas_status rocblas_internal_trtri_batched_template(rocblas_handle   handle,
                                                       rocblas_fill     uplo,
                                                       rocblas_diagonal diag,
                                                       rocblas_int      n,
                                                       const T* const*  A,
                                                       rocblas_stride   offset_A,
                                                       rocblas_int      lda,
                                                       rocblas_stride,
                                                       rocblas_stride   sub_stride_A,
                                                       T* const*        invA,
                                                       rocblas_stride   offset_invA,
                                                       rocblas_int      ldinvA,
                                                       rocblas_stride,
                                                       rocblas_stride sub_stride_invA,
                                                       rocblas_int    batch_count,
                                                       rocblas_int    sub_batch_count,
                                                       T* const*)
try
{
    if(!handle)
        return rocblas_status_invalid_handle;
    if(n <= 0 || batch_count <= 0 || sub_batch_count <= 0)
        return rocblas_status_success;
    // the pointer arrays live on the device: an in-place call (the same array,
    // the same offsets) is not supported in this form
    if(static_cast<const void*>(A) == static_cast<const void*>(invA) && offset_A == offset_invA)
        return rocblas_status_not_implemented;
    const size_t   split_bytes = trtri_split_bytes<T>(handle, n, batch_count, false);
    WorkspaceLease lease(handle, split_bytes);
    for(rocblas_int s = 0; s < sub_batch_count; ++s)
    {
        const rocblas_status st = trtri_core<T>(
            handle, uplo, diag, n,
            BatchPtr<const T>{A, nullptr, 0, offset_A + s * sub_stride_A}, lda,
            BatchPtr<T>{invA, nullptr, 0, offset_invA + s * sub_stride_invA}, ldinvA, batch_count,
            nullptr, lease.ptr, lease.ptr ? split_bytes : 0);
        if(st != rocblas_status_success)
            return st;
    }
    return rocblas_status_success;
}
catch(...)
{
    return rocblas_status_internal_error;
}

#define RBAMD_INSTANTIATE_INTERNAL(T_)                                                            \
    template ROCBLAS_AMD_EXPORT rocblas_status rocblas_internal_trsm_workspace_size<T_>(          \
        rocblas_side, rocblas_operation, rocblas_int, rocblas_int, rocblas_int, rocblas_int,      \
        size_t*, size_t*, size_t*, size_t*, size_t*);                                             \
    template ROCBLAS_AMD_EXPORT rocblas_status rocblas_internal_trsm_batched_workspace_size<T_>(  \
        rocblas_side, rocblas_operation, rocblas_int, rocblas_int, rocblas_int, rocblas_int,      \
        size_t*, size_t*, size_t*, size_t*, size_t*);                                             \
    template ROCBLAS_AMD_EXPORT rocblas_status rocblas_internal_trsm_template<T_>(                \
        rocblas_handle, rocblas_side, rocblas_fill, rocblas_operation, rocblas_diagonal,          \
        rocblas_int, rocblas_int, const T_*, const T_*, rocblas_stride, rocblas_int,              \
        rocblas_stride, T_*, rocblas_stride, rocblas_int, rocblas_stride, rocblas_int, bool,      \
        void*, void*, void*, void*, const T_*, rocblas_int, rocblas_stride, rocblas_stride);      \
    template ROCBLAS_AMD_EXPORT rocblas_status rocblas_internal_trsm_batched_template<T_>(        \
        rocblas_handle, rocblas_side, rocblas_fill, rocblas_operation, rocblas_diagonal,          \
        rocblas_int, rocblas_int, const T_*, const T_* const*, rocblas_stride, rocblas_int,       \
        rocblas_stride, T_* const*, rocblas_stride, rocblas_int, rocblas_stride, rocblas_int,     \
        bool, void*, void*, void*, void*, const T_* const*, rocblas_int, rocblas_stride,          \
        rocblas_stride);                                                                          \
    template ROCBLAS_AMD_EXPORT rocblas_status rocblas_internal_trtri_template<T_>(               \
        rocblas_handle, rocblas_fill, rocblas_diagonal, rocblas_int, const T_*, rocblas_stride,   \
        rocblas_int, rocblas_stride, rocblas_stride, T_*, rocblas_stride, rocblas_int,            \
        rocblas_stride, rocblas_stride, rocblas_int, rocblas_int, T_*);                           \
    template ROCBLAS_AMD_EXPORT rocblas_status rocblas_internal_trtri_batched_template<T_>(       \
        rocblas_handle, rocblas_fill, rocblas_diagonal, rocblas_int, const T_* const*,            \
        rocblas_stride, rocblas_int, rocblas_stride, rocblas_stride, T_* const*, rocblas_stride,  \
        rocblas_int, rocblas_stride, rocblas_stride, rocblas_int, rocblas_int, T_* const*);       \
    template ROCBLAS_AMD_EXPORT rocblas_status rocblas_internal_trsm_workspace_size_64<T_>(       \
        rocblas_side, rocblas_operation, int64_t, int64_t, int64_t, int64_t, int64_t, int64_t,    \
        size_t*, size_t*, size_t*, size_t*, size_t*);                                             \
    template ROCBLAS_AMD_EXPORT rocblas_status                                                    \
        rocblas_internal_trsm_batched_workspace_size_64<T_>(rocblas_side, rocblas_operation,      \
                                                            int64_t, int64_t, int64_t, int64_t,   \
                                                            int64_t, int64_t, size_t*, size_t*,   \
                                                            size_t*, size_t*, size_t*);           \
    template ROCBLAS_AMD_EXPORT rocblas_status rocblas_internal_trsm_workspace_max_size_64<T_>(   \
        rocblas_side, int64_t, int64_t, int64_t, size_t*, size_t*, size_t*);                      \
    template ROCBLAS_AMD_EXPORT rocblas_status                                                    \
        rocblas_internal_trsm_batched_workspace_max_size_64<T_>(                                  \
            rocblas_side, int64_t, int64_t, int64_t, size_t*, size_t*, size_t*, size_t*,          \
            size_t*);                                                                             \
    template ROCBLAS_AMD_EXPORT rocblas_status rocblas_internal_trsm_template_64<T_>(             \
        rocblas_handle, rocblas_side, rocblas_fill, rocblas_operation, rocblas_diagonal, int64_t, \
        int64_t, const T_*, const T_*, rocblas_stride, int64_t, rocblas_stride, T_*,              \
        rocblas_stride, int64_t, rocblas_stride, int64_t, bool, void*, void*, void*, void*,       \
        const T_*, int64_t, rocblas_stride, rocblas_stride);                                      \
    template ROCBLAS_AMD_EXPORT rocblas_status rocblas_internal_trsm_batched_template_64<T_>(     \
        rocblas_handle, rocblas_side, rocblas_fill, rocblas_operation, rocblas_diagonal, int64_t, \
        int64_t, const T_*, const T_* const*, rocblas_stride, int64_t, rocblas_stride,            \
        T_* const*, rocblas_stride, int64_t, rocblas_stride, int64_t, bool, void*, void*, void*,  \
        void*, const T_* const*, int64_t, rocblas_stride, rocblas_stride);

RBAMD_INSTANTIATE_INTERNAL(float)
RBAMD_INSTANTIATE_INTERNAL(double)

#undef RBAMD_INSTANTIATE_INTERNAL
