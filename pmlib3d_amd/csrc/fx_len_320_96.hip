// Length-specialised Poisson launchers for transform lengths 320 and 96.
#include "poisson_fixed_launch.h"

PM_FX_INSTANTIATE(320)
PM_FX_INSTANTIATE(96)
