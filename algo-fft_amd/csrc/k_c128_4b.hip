// Power-of-two c128 kernels, variant VAR_4B.
#include "inst.hpp"

AFFT_POW2_SMALL(double, ::afft::VAR_4B)

