// vbatched.hpp -- in-library entry of the variable-size batched solve
// (vbatched.hip) with a case and an alpha per instance, used by the serving
// loop to solve requests of every side / uplo / trans / diag and alpha in one
// launch sequence.
#pragma once

#include "internal.hpp"

namespace rbamd
{
// Canonical-form flags (SF_* of device_common.hpp) of one solve case.
int trsm_case_flags(rocblas_side side, rocblas_fill uplo, rocblas_operation trans,
                    rocblas_diagonal diag);

// Instance i solves with flags[i] (trsm_case_flags) and alpha[i]; every
// array is in device memory.  Borrows the handle workspace like
// rocblas_amd_{s,d}trsm_vbatched; no argument checks beyond batch_count.
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
                              rocblas_int        batch_count);
} // namespace rbamd
