// small_k48.hip -- instantiates the k <= 64 solve kernel for KMAX = 48.
#include "small_trsm.hpp"

RBAMD_INSTANTIATE_SMALL(48)
