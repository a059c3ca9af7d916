// Kernel instantiations for the bf16 storage path (compute type float).
#define HPXFFT_INSTANTIATE
#include "launch.hpp"

HPXFFT_DEFINE_LAUNCH(bf16, float, hpxfft_dev::bf16_t, false)
