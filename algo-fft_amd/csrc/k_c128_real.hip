// Fused real-transform kernels (c128): N/2-point power-of-two schedules.
#include "inst.hpp"

AFFT_POW2_SMALL(double, ::afft::VAR_REAL)

