// small_k8.hip -- instantiates the k <= 64 solve kernel for KMAX = 8.
#include "small_trsm.hpp"

RBAMD_INSTANTIATE_SMALL(8)
