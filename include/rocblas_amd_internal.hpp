/* rocblas_amd_internal.hpp -- the C++ template entry points of the
 * triangular-solve path that rocSOLVER links against (getrs, potrs, trtri and
 * their batched forms), with the reference's names and signatures so the
 * mangled symbols match:
 *
 *   rocblas_internal_trsm{,_batched}_workspace_size<T>
 *       library/src/blas3/rocblas_trsm.hpp:173-199 (instantiated
 *       rocblas_trsm_kernels.cpp:115-140)
 *   rocblas_internal_trsm{,_batched}_template<T>
 *       library/src/blas3/rocblas_trsm.hpp:201-257 (rocblas_trsm_kernels.cpp:28-113)
 *   rocblas_internal_trsm{,_batched}_{workspace_size,workspace_max_size,template}_64<T>
 *       library/src/src64/blas3/rocblas_trsm_64.hpp:72-185 (ILP64 sizes; the
 *       launcher rocblas_trsm_kernels_64.cpp:31-225, max size :433-530)
 *   rocblas_internal_trtri{,_batched}_template<T>, rocblas_internal_trtri_temp_elements
 *       library/src/blas3/rocblas_trtri.hpp:1098-1150 (rocblas_trtri_kernels.cpp:89-120)
 *
 * T is float or double.  Like the reference's, these functions check no
 * arguments (the caller has validated them) and take an element offset and a
 * stride per operand.  Every solve of this library works in place, so the
 * workspace queries report that nothing is needed (all sizes 0, status
 * rocblas_status_continue -- the reference's answer for its substitution
 * paths) and the trsm workspace / invA arguments are accepted and not read.
 * The trtri templates use w_C_tmp only when invA is A itself (in place): it
 * must then hold rocblas_internal_trtri_temp_elements(n, batch_count)
 * elements, n * n per matrix solved at a time.  The _64 forms take 64-bit
 * sizes: a triangle beyond INT32_MAX is rocblas_status_invalid_size (as in the
 * reference), leading dimensions beyond INT32_MAX rocblas_status_not_implemented;
 * their max-size queries report zero bytes with rocblas_status_success.
 */
#ifndef ROCBLAS_AMD_INTERNAL_HPP
#define ROCBLAS_AMD_INTERNAL_HPP

#include "rocblas_amd.h"

#include <cstddef>
#include <cstdint>

template <typename T>
ROCBLAS_AMD_EXPORT rocblas_status
    rocblas_internal_trsm_workspace_size(rocblas_side      side,
                                         rocblas_operation transA,
                                         rocblas_int       m,
                                         rocblas_int       n,
                                         rocblas_int       batch_count,
                                         rocblas_int       supplied_invA_size,
                                         size_t*           w_x_tmp_size,
                                         size_t*           w_x_tmp_arr_size,
                                         size_t*           w_invA_size,
                                         size_t*           w_invA_arr_size,
                                         size_t*           w_x_tmp_size_backup);

template <typename T>
ROCBLAS_AMD_EXPORT rocblas_status
    rocblas_internal_trsm_batched_workspace_size(rocblas_side      side,
                                                 rocblas_operation transA,
                                                 rocblas_int       m,
                                                 rocblas_int       n,
                                                 rocblas_int       batch_count,
                                                 rocblas_int       supplied_invA_size,
                                                 size_t*           w_x_tmp_size,
                                                 size_t*           w_x_tmp_arr_size,
                                                 size_t*           w_invA_size,
                                                 size_t*           w_invA_arr_size,
                                                 size_t*           w_x_tmp_size_backup);

template <typename T>
ROCBLAS_AMD_EXPORT rocblas_status
    rocblas_internal_trsm_template(rocblas_handle    handle,
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
                                   bool              optimal_mem,
                                   void*             w_x_temp,
                                   void*             w_x_temparr,
                                   void*             invA               = nullptr,
                                   void*             invAarr            = nullptr,
                                   const T*          supplied_invA      = nullptr,
                                   rocblas_int       supplied_invA_size = 0,
                                   rocblas_stride    offset_invA        = 0,
                                   rocblas_stride    stride_invA        = 0);

template <typename T>
ROCBLAS_AMD_EXPORT rocblas_status
    rocblas_internal_trsm_batched_template(rocblas_handle    handle,
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
                                           rocblas_stride    stride_A,
                                           T* const*         B,
                                           rocblas_stride    offset_B,
                                           rocblas_int       ldb,
                                           rocblas_stride    stride_B,
                                           rocblas_int       batch_count,
                                           bool              optimal_mem,
                                           void*             w_x_temp,
                                           void*             w_x_temparr,
                                           void*             invA               = nullptr,
                                           void*             invAarr            = nullptr,
                                           const T* const*   supplied_invA      = nullptr,
                                           rocblas_int       supplied_invA_size = 0,
                                           rocblas_stride    offset_invA        = 0,
                                           rocblas_stride    stride_invA        = 0);

template <typename T>
ROCBLAS_AMD_EXPORT rocblas_status
    rocblas_internal_trsm_workspace_size_64(rocblas_side      side,
                                            rocblas_operation transA,
                                            int64_t           m,
                                            int64_t           n,
                                            int64_t           lda,
                                            int64_t           ldb,
                                            int64_t           batch_count,
                                            int64_t           supplied_invA_size,
                                            size_t*           w_x_tmp_size,
                                            size_t*           w_x_tmp_arr_size,
                                            size_t*           w_invA_size,
                                            size_t*           w_invA_arr_size,
                                            size_t*           w_x_tmp_size_backup);

template <typename T>
ROCBLAS_AMD_EXPORT rocblas_status
    rocblas_internal_trsm_batched_workspace_size_64(rocblas_side      side,
                                                    rocblas_operation transA,
                                                    int64_t           m,
                                                    int64_t           n,
                                                    int64_t           lda,
                                                    int64_t           ldb,
                                                    int64_t           batch_count,
                                                    int64_t           supplied_invA_size,
                                                    size_t*           w_x_tmp_size,
                                                    size_t*           w_x_tmp_arr_size,
                                                    size_t*           w_invA_size,
                                                    size_t*           w_invA_arr_size,
                                                    size_t*           w_x_tmp_size_backup);

/* the most any call with m <= M and n <= N needs (rocblas_trsm_kernels_64.cpp:433-530) */
template <typename T>
ROCBLAS_AMD_EXPORT rocblas_status
    rocblas_internal_trsm_workspace_max_size_64(rocblas_side side,
                                                int64_t      M,
                                                int64_t      N,
                                                int64_t      batch_count,
                                                size_t*      w_x_tmp_size,
                                                size_t*      w_invA_size,
                                                size_t*      w_x_tmp_size_backup);

template <typename T>
ROCBLAS_AMD_EXPORT rocblas_status
    rocblas_internal_trsm_batched_workspace_max_size_64(rocblas_side side,
                                                        int64_t      M,
                                                        int64_t      N,
                                                        int64_t      batch_count,
                                                        size_t*      w_x_tmp_size,
                                                        size_t*      w_x_tmp_arr_size,
                                                        size_t*      w_invA_size,
                                                        size_t*      w_invA_arr_size,
                                                        size_t*      w_x_tmp_size_backup);

template <typename T>
ROCBLAS_AMD_EXPORT rocblas_status
    rocblas_internal_trsm_template_64(rocblas_handle    handle,
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
                                      bool              optimal_mem,
                                      void*             w_x_temp,
                                      void*             w_x_temparr,
                                      void*             invA               = nullptr,
                                      void*             invAarr            = nullptr,
                                      const T*          supplied_invA      = nullptr,
                                      int64_t           supplied_invA_size = 0,
                                      rocblas_stride    offset_invA        = 0,
                                      rocblas_stride    stride_invA        = 0);

template <typename T>
ROCBLAS_AMD_EXPORT rocblas_status
    rocblas_internal_trsm_batched_template_64(rocblas_handle    handle,
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
                                              rocblas_stride    stride_A,
                                              T* const*         B,
                                              rocblas_stride    offset_B,
                                              int64_t           ldb,
                                              rocblas_stride    stride_B,
                                              int64_t           batch_count,
                                              bool              optimal_mem,
                                              void*             w_x_temp,
                                              void*             w_x_temparr,
                                              void*             invA               = nullptr,
                                              void*             invAarr            = nullptr,
                                              const T* const*   supplied_invA      = nullptr,
                                              int64_t           supplied_invA_size = 0,
                                              rocblas_stride    offset_invA        = 0,
                                              rocblas_stride    stride_invA        = 0);

/* batch_count instances (stride_A apart) of sub_batch_count matrices each
 * (sub_stride_A apart); rocblas_trtri.hpp:1102-1119 */
template <typename T>
ROCBLAS_AMD_EXPORT rocblas_status
    rocblas_internal_trtri_template(rocblas_handle   handle,
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
                                    T*               w_C_tmp);

/* elements of w_C_tmp a trtri call needs; rocblas_trtri.hpp:1124-1126 */
ROCBLAS_AMD_EXPORT size_t rocblas_internal_trtri_temp_elements(rocblas_int n,
                                                               rocblas_int batch_count);

/* rocblas_trtri.hpp:1132-1150 */
template <typename T>
ROCBLAS_AMD_EXPORT rocblas_status
    rocblas_internal_trtri_batched_template(rocblas_handle   handle,
                                            rocblas_fill     uplo,
                                            rocblas_diagonal diag,
                                            rocblas_int      n,
                                            const T* const*  A,
                                            rocblas_stride   offset_A,
                                            rocblas_int      lda,
                                            rocblas_stride   stride_A,
                                            rocblas_stride   sub_stride_A,
                                            T* const*        invA,
                                            rocblas_stride   offset_invA,
                                            rocblas_int      ldinvA,
                                            rocblas_stride   stride_invA,
                                            rocblas_stride   sub_stride_invA,
                                            rocblas_int      batch_count,
                                            rocblas_int      sub_batch_count,
                                            T* const*        w_C_tmp);

#endif /* ROCBLAS_AMD_INTERNAL_HPP */
