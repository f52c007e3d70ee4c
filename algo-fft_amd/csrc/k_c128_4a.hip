// Power-of-two c128 kernels, variant VAR_4A.
#include "inst.hpp"

AFFT_POW2_SMALL(double, ::afft::VAR_4A)

