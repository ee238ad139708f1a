// Length-specialised Poisson launchers for transform lengths 768 and 32.
#include "poisson_fixed_launch.h"

PM_FX_INSTANTIATE(768)
PM_FX_INSTANTIATE(32)
