// Length-specialised Poisson launchers for transform lengths 512 and 64.
#include "poisson_fixed_launch.h"

PM_FX_INSTANTIATE(512)
PM_FX_INSTANTIATE(64)
