// opt.hpp -- hpxfft::shared::opt resolved to the MI355X class hpxfft::gpu::opt,
// so code written against HPX-FFT's core/include/hpxfft/shared/opt.hpp (class
// surface at opt.hpp:23-27) builds unchanged against this include directory and
// links with -lhpxfft_hip.
#ifndef HPXFFT_GPU_SHARED_OPT_HPP
#define HPXFFT_GPU_SHARED_OPT_HPP

#include "../gpu/loop.hpp"

namespace hpxfft::shared
{
using vector_2d = hpxfft::gpu::vector_2d;
using opt = hpxfft::gpu::opt;
}  // namespace hpxfft::shared

#endif
