// Kernel instantiations for the f16 storage path (compute type float).
#define HPXFFT_INSTANTIATE
#include "launch.hpp"

HPXFFT_DEFINE_LAUNCH(f16, float, __half, false)
