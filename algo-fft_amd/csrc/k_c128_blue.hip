// Fused Bluestein kernels (c128) for padded power-of-two lengths held in LDS.
#include "inst.hpp"

AFFT_POW2_SMALL(double, ::afft::VAR_BLUE)

