// Length-specialised Poisson launchers for transform lengths 640 and 48.
#include "poisson_fixed_launch.h"

PM_FX_INSTANTIATE(640)
PM_FX_INSTANTIATE(48)
