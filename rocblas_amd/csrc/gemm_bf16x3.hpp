// gemm_bf16x3.hpp -- launcher of the split-BF16x3 update GEMM (gemm_bf16x3.hip).
#pragma once

#include "device_common.hpp"

namespace rbamd
{
// C (FP64) := alpha op(A) op(B) + beta C with every product formed from BF16
// halves of its operands (three BF16 MFMA products accumulated in FP32).
// Dense operands only (a.tri must be 0); any op pair.
template <typename TA, typename TB>
rocblas_status launch_gemm_bf16x3(rocblas_handle h, GemmArgsT<TA, TB, double> a);
} // namespace rbamd
