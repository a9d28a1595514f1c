/*
 * rocblas_amd.h -- C ABI of the MI355X-native triangular-solve library
 * (librocblas_amd.so).
 *
 * Every symbol below has the name, signature, enum values and status
 * semantics of the rocBLAS entry point it replaces, so a program written
 * against rocBLAS's trsm path links against this library unchanged.  The
 * comment above each declaration cites the reference declaration
 * (library/include/internal/...) and the implementation whose behaviour it
 * follows.
 */
#ifndef ROCBLAS_AMD_H
#define ROCBLAS_AMD_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef ROCBLAS_AMD_EXPORT
#define ROCBLAS_AMD_EXPORT __attribute__((visibility("default")))
#endif

/* ---- types: rocblas-types.h:63-84 --------------------------------------- */
typedef struct _rocblas_handle* rocblas_handle;    /* rocblas-types.h:63 */
typedef struct ihipStream_t*    hipStream_t;       /* rocblas-types.h:65 */
typedef int32_t                 rocblas_int;       /* rocblas-types.h:77-81 (LP64) */
typedef int64_t                 rocblas_stride;    /* rocblas-types.h:84 */

/* ---- enums: rocblas-types.h:125-157, 205-233 ---------------------------- */
typedef enum rocblas_operation_
{
    rocblas_operation_none                = 111,
    rocblas_operation_transpose           = 112,
    rocblas_operation_conjugate_transpose = 113
} rocblas_operation;

typedef enum rocblas_fill_
{
    rocblas_fill_upper = 121,
    rocblas_fill_lower = 122,
    rocblas_fill_full  = 123
} rocblas_fill;

typedef enum rocblas_diagonal_
{
    rocblas_diagonal_non_unit = 131,
    rocblas_diagonal_unit     = 132
} rocblas_diagonal;

typedef enum rocblas_side_
{
    rocblas_side_left  = 141,
    rocblas_side_right = 142,
    rocblas_side_both  = 143
} rocblas_side;

typedef enum rocblas_status_
{
    rocblas_status_success             = 0,
    rocblas_status_invalid_handle      = 1,
    rocblas_status_not_implemented     = 2,
    rocblas_status_invalid_pointer     = 3,
    rocblas_status_invalid_size        = 4,
    rocblas_status_memory_error        = 5,
    rocblas_status_internal_error      = 6,
    rocblas_status_perf_degraded       = 7,
    rocblas_status_size_query_mismatch = 8,
    rocblas_status_size_increased      = 9,
    rocblas_status_size_unchanged      = 10,
    rocblas_status_invalid_value       = 11,
    rocblas_status_continue            = 12,
    rocblas_status_check_numerics_fail = 13,
    rocblas_status_excluded_from_build = 14,
    rocblas_status_arch_mismatch       = 15
} rocblas_status;

typedef enum rocblas_pointer_mode_
{
    rocblas_pointer_mode_host   = 0,
    rocblas_pointer_mode_device = 1
} rocblas_pointer_mode;

/* rocblas-types.h:165-185 (the real types the _ex entry points accept here;
 * the other values are declared so callers' code compiles) */
typedef enum rocblas_datatype_
{
    rocblas_datatype_f16_r   = 150,
    rocblas_datatype_f32_r   = 151,
    rocblas_datatype_f64_r   = 152,
    rocblas_datatype_f16_c   = 153,
    rocblas_datatype_f32_c   = 154,
    rocblas_datatype_f64_c   = 155,
    rocblas_datatype_i8_r    = 160,
    rocblas_datatype_u8_r    = 161,
    rocblas_datatype_i32_r   = 162,
    rocblas_datatype_u32_r   = 163,
    rocblas_datatype_i8_c    = 164,
    rocblas_datatype_u8_c    = 165,
    rocblas_datatype_i32_c   = 166,
    rocblas_datatype_u32_c   = 167,
    rocblas_datatype_bf16_r  = 168,
    rocblas_datatype_bf16_c  = 169,
    rocblas_datatype_f8_r    = 170,
    rocblas_datatype_bf8_r   = 171,
    rocblas_datatype_invalid = 255
} rocblas_datatype;

/* ---- handle and auxiliary ------------------------------------------------ */

/* rocblas-auxiliary.h:40; library/src/rocblas_auxiliary.cpp:188-204.
 * The handle binds the device current at creation. */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_create_handle(rocblas_handle* handle);

/* rocblas-auxiliary.h:44; rocblas_auxiliary.cpp:211-227 */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_destroy_handle(rocblas_handle handle);

/* rocblas-auxiliary.h:48; rocblas_auxiliary.cpp:234-276 */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_set_stream(rocblas_handle handle, hipStream_t stream);

/* rocblas-auxiliary.h:52; rocblas_auxiliary.cpp:282-296 */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_get_stream(rocblas_handle handle, hipStream_t* stream);

/* rocblas-auxiliary.h:56; rocblas_auxiliary.cpp:68-82 */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_set_pointer_mode(rocblas_handle       handle,
                                                           rocblas_pointer_mode pointer_mode);

/* rocblas-auxiliary.h:60; rocblas_auxiliary.cpp:48-63 */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_get_pointer_mode(rocblas_handle        handle,
                                                           rocblas_pointer_mode* pointer_mode);

/* rocblas-functions.h:24972; rocblas_auxiliary.cpp:783-810 */
ROCBLAS_AMD_EXPORT const char* rocblas_status_to_string(rocblas_status status);

/* rocblas-functions.h:25027; library/src/handle.cpp:449-462 */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_start_device_memory_size_query(rocblas_handle handle);

/* rocblas-functions.h:25039; handle.cpp:468-483 */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_stop_device_memory_size_query(rocblas_handle handle,
                                                                        size_t*        size);

/* rocblas-functions.h:25043; handle.cpp:654-657 */
ROCBLAS_AMD_EXPORT bool rocblas_is_device_memory_size_query(rocblas_handle handle);

/* rocblas-functions.h:25069; handle.cpp:489-500 */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_get_device_memory_size(rocblas_handle handle,
                                                                 size_t*        size);

/* rocblas-functions.h:25087; handle.cpp:541-590 */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_set_device_memory_size(rocblas_handle handle,
                                                                 size_t         size);

/* rocblas-functions.h:25104; handle.cpp:596-624 */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_set_workspace(rocblas_handle handle, void* addr,
                                                        size_t size);

/* rocblas-functions.h:25112; handle.cpp:630-638 */
ROCBLAS_AMD_EXPORT bool rocblas_is_managing_device_memory(rocblas_handle handle);

/* rocblas-functions.h:25122; handle.cpp:643-650 */
ROCBLAS_AMD_EXPORT bool rocblas_is_user_managing_device_memory(rocblas_handle handle);

/* ---- triangular solve ---------------------------------------------------- */
/* op(A) X = alpha B (side left) or X op(A) = alpha B (side right); X overwrites
 * B.  Argument checking order and quick returns follow
 * library/src/blas3/rocblas_trsm.hpp:29-82 and
 * library/src/blas3/rocblas_trsm_imp.hpp:54-157. */

/* rocblas-functions.h:19388; rocblas_trsm_imp.hpp:285-314 */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_strsm(rocblas_handle    handle,
                                                rocblas_side      side,
                                                rocblas_fill      uplo,
                                                rocblas_operation transA,
                                                rocblas_diagonal  diag,
                                                rocblas_int       m,
                                                rocblas_int       n,
                                                const float*      alpha,
                                                const float*      A,
                                                rocblas_int       lda,
                                                float*            B,
                                                rocblas_int       ldb);

/* rocblas-functions.h:19401; rocblas_trsm_imp.hpp:285-314 */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_dtrsm(rocblas_handle    handle,
                                                rocblas_side      side,
                                                rocblas_fill      uplo,
                                                rocblas_operation transA,
                                                rocblas_diagonal  diag,
                                                rocblas_int       m,
                                                rocblas_int       n,
                                                const double*     alpha,
                                                const double*     A,
                                                rocblas_int       lda,
                                                double*           B,
                                                rocblas_int       ldb);

/* rocblas-functions.h:19574; rocblas_trsm_batched_imp.hpp:299-329.
 * A and B are device arrays of batch_count device pointers. */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_strsm_batched(rocblas_handle     handle,
                                                        rocblas_side       side,
                                                        rocblas_fill       uplo,
                                                        rocblas_operation  transA,
                                                        rocblas_diagonal   diag,
                                                        rocblas_int        m,
                                                        rocblas_int        n,
                                                        const float*       alpha,
                                                        const float* const A[],
                                                        rocblas_int        lda,
                                                        float* const       B[],
                                                        rocblas_int        ldb,
                                                        rocblas_int        batch_count);

/* rocblas-functions.h:19588; rocblas_trsm_batched_imp.hpp:299-329 */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_dtrsm_batched(rocblas_handle      handle,
                                                        rocblas_side        side,
                                                        rocblas_fill        uplo,
                                                        rocblas_operation   transA,
                                                        rocblas_diagonal    diag,
                                                        rocblas_int         m,
                                                        rocblas_int         n,
                                                        const double*       alpha,
                                                        const double* const A[],
                                                        rocblas_int         lda,
                                                        double* const       B[],
                                                        rocblas_int         ldb,
                                                        rocblas_int         batch_count);

/* rocblas-functions.h:19774; rocblas_trsm_strided_batched_imp.hpp:312-357 */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_strsm_strided_batched(rocblas_handle    handle,
                                                                rocblas_side      side,
                                                                rocblas_fill      uplo,
                                                                rocblas_operation transA,
                                                                rocblas_diagonal  diag,
                                                                rocblas_int       m,
                                                                rocblas_int       n,
                                                                const float*      alpha,
                                                                const float*      A,
                                                                rocblas_int       lda,
                                                                rocblas_stride    stride_a,
                                                                float*            B,
                                                                rocblas_int       ldb,
                                                                rocblas_stride    stride_b,
                                                                rocblas_int       batch_count);

/* rocblas-functions.h:19790; rocblas_trsm_strided_batched_imp.hpp:312-357 */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_dtrsm_strided_batched(rocblas_handle    handle,
                                                                rocblas_side      side,
                                                                rocblas_fill      uplo,
                                                                rocblas_operation transA,
                                                                rocblas_diagonal  diag,
                                                                rocblas_int       m,
                                                                rocblas_int       n,
                                                                const double*     alpha,
                                                                const double*     A,
                                                                rocblas_int       lda,
                                                                rocblas_stride    stride_a,
                                                                double*           B,
                                                                rocblas_int       ldb,
                                                                rocblas_stride    stride_b,
                                                                rocblas_int       batch_count);

/* ---- triangular inverse --------------------------------------------------- */
/* invA := inverse of the uplo triangle of A (n x n); the opposite triangle of
 * invA is cleared.  invA may be A itself.  Checks follow
 * library/src/blas3/rocblas_trtri.hpp:31-62. */

/* rocblas-functions.h:19081 */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_strtri(rocblas_handle   handle,
                                                 rocblas_fill     uplo,
                                                 rocblas_diagonal diag,
                                                 rocblas_int      n,
                                                 const float*     A,
                                                 rocblas_int      lda,
                                                 float*           invA,
                                                 rocblas_int      ldinvA);

/* rocblas-functions.h:19090 */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_dtrtri(rocblas_handle   handle,
                                                 rocblas_fill     uplo,
                                                 rocblas_diagonal diag,
                                                 rocblas_int      n,
                                                 const double*    A,
                                                 rocblas_int      lda,
                                                 double*          invA,
                                                 rocblas_int      ldinvA);

/* rocblas-functions.h:19159 */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_strtri_batched(rocblas_handle     handle,
                                                         rocblas_fill       uplo,
                                                         rocblas_diagonal   diag,
                                                         rocblas_int        n,
                                                         const float* const A[],
                                                         rocblas_int        lda,
                                                         float* const       invA[],
                                                         rocblas_int        ldinvA,
                                                         rocblas_int        batch_count);

/* rocblas-functions.h:19169 */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_dtrtri_batched(rocblas_handle      handle,
                                                         rocblas_fill        uplo,
                                                         rocblas_diagonal    diag,
                                                         rocblas_int         n,
                                                         const double* const A[],
                                                         rocblas_int         lda,
                                                         double* const       invA[],
                                                         rocblas_int         ldinvA,
                                                         rocblas_int         batch_count);

/* rocblas-functions.h:19249 */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_strtri_strided_batched(rocblas_handle   handle,
                                                                 rocblas_fill     uplo,
                                                                 rocblas_diagonal diag,
                                                                 rocblas_int      n,
                                                                 const float*     A,
                                                                 rocblas_int      lda,
                                                                 rocblas_stride   stride_a,
                                                                 float*           invA,
                                                                 rocblas_int      ldinvA,
                                                                 rocblas_stride   stride_invA,
                                                                 rocblas_int      batch_count);

/* rocblas-functions.h:19261 */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_dtrtri_strided_batched(rocblas_handle   handle,
                                                                 rocblas_fill     uplo,
                                                                 rocblas_diagonal diag,
                                                                 rocblas_int      n,
                                                                 const double*    A,
                                                                 rocblas_int      lda,
                                                                 rocblas_stride   stride_a,
                                                                 double*          invA,
                                                                 rocblas_int      ldinvA,
                                                                 rocblas_stride   stride_invA,
                                                                 rocblas_int      batch_count);

/* ---- ILP64 trsm ---------------------------------------------------------
 * rocblas-functions.h:19441-19466, 19631-19658, 19839-19870; implementation
 * library/src/src64/blas3/rocblas_trsm_64.cpp, rocblas_trsm_kernels_64.cpp:31-225.
 * Same checks on 64-bit sizes; a triangle dimension above INT32_MAX returns
 * rocblas_status_invalid_size (as the reference); leading dimensions above
 * INT32_MAX return rocblas_status_not_implemented; other sizes are solved by
 * the 32-bit path in chunks of right-hand sides and batch entries. */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_strsm_64(rocblas_handle    handle,
                                                   rocblas_side      side,
                                                   rocblas_fill      uplo,
                                                   rocblas_operation transA,
                                                   rocblas_diagonal  diag,
                                                   int64_t           m,
                                                   int64_t           n,
                                                   const float*      alpha,
                                                   const float*      A,
                                                   int64_t           lda,
                                                   float*            B,
                                                   int64_t           ldb);

ROCBLAS_AMD_EXPORT rocblas_status rocblas_strsm_batched_64(rocblas_handle    handle,
                                                           rocblas_side      side,
                                                           rocblas_fill      uplo,
                                                           rocblas_operation transA,
                                                           rocblas_diagonal  diag,
                                                           int64_t           m,
                                                           int64_t           n,
                                                           const float*        alpha,
                                                           const float* const  A[],
                                                           int64_t           lda,
                                                           float* const        B[],
                                                           int64_t           ldb,
                                                           int64_t           batch_count);

ROCBLAS_AMD_EXPORT rocblas_status rocblas_strsm_strided_batched_64(rocblas_handle    handle,
                                                                   rocblas_side      side,
                                                                   rocblas_fill      uplo,
                                                                   rocblas_operation transA,
                                                                   rocblas_diagonal  diag,
                                                                   int64_t           m,
                                                                   int64_t           n,
                                                                   const float*        alpha,
                                                                   const float*        A,
                                                                   int64_t           lda,
                                                                   rocblas_stride    stride_a,
                                                                   float*              B,
                                                                   int64_t           ldb,
                                                                   rocblas_stride    stride_b,
                                                                   int64_t           batch_count);

ROCBLAS_AMD_EXPORT rocblas_status rocblas_dtrsm_64(rocblas_handle    handle,
                                                   rocblas_side      side,
                                                   rocblas_fill      uplo,
                                                   rocblas_operation transA,
                                                   rocblas_diagonal  diag,
                                                   int64_t           m,
                                                   int64_t           n,
                                                   const double*     alpha,
                                                   const double*     A,
                                                   int64_t           lda,
                                                   double*           B,
                                                   int64_t           ldb);

ROCBLAS_AMD_EXPORT rocblas_status rocblas_dtrsm_batched_64(rocblas_handle    handle,
                                                           rocblas_side      side,
                                                           rocblas_fill      uplo,
                                                           rocblas_operation transA,
                                                           rocblas_diagonal  diag,
                                                           int64_t           m,
                                                           int64_t           n,
                                                           const double*        alpha,
                                                           const double* const  A[],
                                                           int64_t           lda,
                                                           double* const        B[],
                                                           int64_t           ldb,
                                                           int64_t           batch_count);

ROCBLAS_AMD_EXPORT rocblas_status rocblas_dtrsm_strided_batched_64(rocblas_handle    handle,
                                                                   rocblas_side      side,
                                                                   rocblas_fill      uplo,
                                                                   rocblas_operation transA,
                                                                   rocblas_diagonal  diag,
                                                                   int64_t           m,
                                                                   int64_t           n,
                                                                   const double*        alpha,
                                                                   const double*        A,
                                                                   int64_t           lda,
                                                                   rocblas_stride    stride_a,
                                                                   double*              B,
                                                                   int64_t           ldb,
                                                                   rocblas_stride    stride_b,
                                                                   int64_t           batch_count);

/* ---- trsm_ex ------------------------------------------------------------
 * rocblas-functions.h:23223-23237, 23404-23419, 23563-23581; implementation
 * rocblas_trsm.cpp:28-73 (and the batched / strided-batched .cpp files).
 * compute_type rocblas_datatype_f32_r or _f64_r selects the element type
 * (others: rocblas_status_not_implemented; the reference also takes the
 * complex types).  invA (packed 128x128 diagonal-block inverses) is an
 * optional hint in the reference; this solve forms no inverses and does not
 * read it, so results equal rocblas_{s,d}trsm's. */
ROCBLAS_AMD_EXPORT rocblas_status rocblas_trsm_ex(rocblas_handle    handle,
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
                                                  rocblas_datatype  compute_type);

ROCBLAS_AMD_EXPORT rocblas_status rocblas_trsm_batched_ex(rocblas_handle    handle,
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
                                                          rocblas_datatype  compute_type);

ROCBLAS_AMD_EXPORT rocblas_status rocblas_trsm_strided_batched_ex(rocblas_handle    handle,
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
                                                                  rocblas_datatype  compute_type);

#ifdef __cplusplus
}
#endif

#endif /* ROCBLAS_AMD_H */
