// small_k64.hip -- instantiates the k <= 64 solve kernel for KMAX = 64.
#include "small_trsm.hpp"

RBAMD_INSTANTIATE_SMALL(64)
