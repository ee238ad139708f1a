// Length-specialised Poisson launchers for transform lengths 256 and 128.
#include "poisson_fixed_launch.h"

PM_FX_INSTANTIATE(256)
PM_FX_INSTANTIATE(128)
