// Length-specialised Poisson launchers for transform lengths 192 and 160.
#include "poisson_fixed_launch.h"

PM_FX_INSTANTIATE(192)
PM_FX_INSTANTIATE(160)
