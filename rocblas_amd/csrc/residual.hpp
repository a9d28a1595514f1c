// residual.hpp -- batched normwise backward error of solved instances
// (residual.hip): per-instance sizes, cases (SF_TA / SF_REV / SF_UNIT /
// SF_COLS of the canonical form) and alpha, FP64 accumulation.
#pragma once

#include "internal.hpp"

namespace rbamd
{
template <typename T>
struct EtaArgs
{
    const int*      K;       // triangle order per instance
    const int*      nvec;    // right-hand-side vectors per instance
    const int*      flags;   // canonical-form flags per instance
    const int*      lda;
    const int*      ldx;
    const int*      ldb;     // of B0
    const T* const* A;
    const T* const* X;       // the solution
    const T* const* B0;      // the right-hand side the solve started from
    const T*        alpha;   // per instance
    const int64_t*  acc_off; // per instance: offset of its column sums in acc
    double*         acc;     // eta_acc_doubles() per instance
    int64_t         acc_total;
    const int*      items;   // (instance, panel, row block) triples
    int             nitems;
    int             batch;
    double*         eta;     // per instance, written
};

// Column sums an instance needs in acc, and its work items.
int64_t eta_acc_doubles(int K, int nvec, bool side_left);
int64_t eta_items(int K, int nvec);
// Writes eta_items(K, nvec) triples for instance `inst` to out.
void eta_fill_items(int inst, int K, int nvec, int* out);

// Zeroes acc, accumulates and writes eta (three launches on `stream`).
template <typename T>
rocblas_status eta_launch(hipStream_t stream, const EtaArgs<T>& a);
} // namespace rbamd
