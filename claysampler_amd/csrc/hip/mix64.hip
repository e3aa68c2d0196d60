// Mix kernels for 64-tap sinc tables (the templates are in mix.h; one
// translation unit per tap count so they compile in parallel).
#include "mix.h"

CLAY_INSTANTIATE_MIX_WAVES_TAPS(64)
