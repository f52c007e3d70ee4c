// Fused real-transform kernels (c64): N/2-point power-of-two schedules.
#include "inst.hpp"

AFFT_POW2_SMALL(float, ::afft::VAR_REAL)
AFFT_POW2_16K(float, ::afft::VAR_REAL)
