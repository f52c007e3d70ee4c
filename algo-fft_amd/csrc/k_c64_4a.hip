// Power-of-two c64 kernels, variant VAR_4A.
#include "inst.hpp"

AFFT_POW2_SMALL(float, ::afft::VAR_4A)
AFFT_POW2_16K(float, ::afft::VAR_4A)
