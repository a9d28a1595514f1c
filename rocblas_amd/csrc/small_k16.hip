// small_k16.hip -- instantiates the k <= 64 solve kernel for KMAX = 16.
#include "small_trsm.hpp"

RBAMD_INSTANTIATE_SMALL(16)
