// Kernel instantiations for the f64 storage path (compute type double).
#define HPXFFT_INSTANTIATE
#include "launch.hpp"

HPXFFT_DEFINE_LAUNCH(f64, double, double, true)
