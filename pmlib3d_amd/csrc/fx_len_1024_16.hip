// Length-specialised Poisson launchers for transform lengths 1024 and 16.
#include "poisson_fixed_launch.h"

PM_FX_INSTANTIATE(1024)
PM_FX_INSTANTIATE(16)
