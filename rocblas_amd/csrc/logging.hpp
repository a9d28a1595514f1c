// logging.hpp -- ROCBLAS_LAYER argument logging of the trsm entry points.
//
// The reference reads ROCBLAS_LAYER when a handle is created
// (library/src/handle.cpp:829-877) and logs each call before its argument
// check (library/src/blas3/rocblas_trsm_imp.hpp:71-150 and the batched /
// strided-batched _imp.hpp files) in three forms, chosen by bits of the mask:
//   1  trace:   one comma-separated line of the arguments per call
//   2  bench:   a rocblas-bench command line that replays the call
//   4  profile: a YAML list of the distinct argument sets with call counts,
//               written when the process exits
// (formats: library/src/include/logging.hpp:150-275,
// library/src/include/rocblas_ostream.hpp:325-520).  Each stream goes to the
// file named by ROCBLAS_LOG_{TRACE,BENCH,PROFILE}_PATH, else ROCBLAS_LOG_PATH,
// else stderr.
#pragma once

#include "internal.hpp"

#include <cstdint>

namespace rbamd
{
// A log destination shared by every handle that names the same file.
struct LogSink;

enum class TrsmForm
{
    plain,
    batched,
    strided
};

// Read ROCBLAS_LAYER and open the streams it asks for (handle creation).
void init_logging(rocblas_handle h);

// Log one trsm call if the handle's layer mode asks for it; nothing in a
// device-memory size query.  api64 selects the ILP64 names.
template <typename T>
void log_trsm(rocblas_handle    h,
              TrsmForm          form,
              bool              api64,
              rocblas_side      side,
              rocblas_fill      uplo,
              rocblas_operation transA,
              rocblas_diagonal  diag,
              int64_t           m,
              int64_t           n,
              const T*          alpha,
              const void*       A,
              int64_t           lda,
              rocblas_stride    stride_a,
              const void*       B,
              int64_t           ldb,
              rocblas_stride    stride_b,
              int64_t           batch_count);
} // namespace rbamd
