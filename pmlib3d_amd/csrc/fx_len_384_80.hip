// Length-specialised Poisson launchers for transform lengths 384 and 80.
#include "poisson_fixed_launch.h"

PM_FX_INSTANTIATE(384)
PM_FX_INSTANTIATE(80)
