// small_k32.hip -- instantiates the k <= 64 solve kernel for KMAX = 32.
#include "small_trsm.hpp"

RBAMD_INSTANTIATE_SMALL(32)
