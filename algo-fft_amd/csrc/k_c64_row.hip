// Power-of-two c64 kernels, variant VAR_ROW.
#include "inst.hpp"

AFFT_POW2_SMALL(float, ::afft::VAR_ROW)
AFFT_POW2_16K(float, ::afft::VAR_ROW)
