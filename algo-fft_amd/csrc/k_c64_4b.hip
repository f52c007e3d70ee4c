// Power-of-two c64 kernels, variant VAR_4B.
#include "inst.hpp"

AFFT_POW2_SMALL(float, ::afft::VAR_4B)
AFFT_POW2_16K(float, ::afft::VAR_4B)
