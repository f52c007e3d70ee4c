// Fused Bluestein kernels (c64) for padded power-of-two lengths held in LDS.
#include "inst.hpp"

AFFT_POW2_SMALL(float, ::afft::VAR_BLUE)
AFFT_POW2_16K(float, ::afft::VAR_BLUE)
