// Kernel instantiations for the f32 storage path (compute type float).
#define HPXFFT_INSTANTIATE
#include "launch.hpp"

HPXFFT_DEFINE_LAUNCH(f32, float, float, true)
