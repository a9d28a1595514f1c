// trtri_kernels.hpp -- launchers of the inversion kernels (trtri_kernels.hip)
// used by the blocked inversion in trtri_api.cpp.
#pragma once

#include "internal.hpp"

namespace rbamd
{
// Order of the diagonal blocks the inversion kernel handles for n > 16 (its
// FP64 form holds a 32-entry column per lane without spilling; larger
// blocks are left to the doubling steps, whose GEMMs carry the flops).
constexpr int kTrtriNB = 32;

// Diagonal blocks [b nb, min(n, (b+1) nb)) of n x n triangles: A (lda) and
// the inverse X (ldx), batch instances.
template <typename T>
struct TrtriDiagArgs
{
    BatchPtr<const T> A;
    int               lda;
    BatchPtr<T>       X;
    int               ldx;
    int               n;
    int               nb; // 1..kTrtriNB
    int               upper;
    int               unit;
    int               batch;
};

// X's diagonal blocks := inverses of A's (the opposite triangle of each
// block := 0); X is only written.
template <typename T>
rocblas_status trtri_diag_launch(rocblas_handle h, const TrtriDiagArgs<T>& a);

// X's strictly-opposite triangle outside the diagonal blocks := 0 (A unused).
template <typename T>
rocblas_status trtri_zero_opposite_launch(rocblas_handle h, const TrtriDiagArgs<T>& a);
} // namespace rbamd
