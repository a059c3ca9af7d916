// loop.hpp -- hpxfft::shared::loop resolved to the MI355X class hpxfft::gpu::loop,
// so code written against HPX-FFT's core/include/hpxfft/shared/loop.hpp (class
// surface at loop.hpp:15-57) builds unchanged against this include directory and
// links with -lhpxfft_hip.
#ifndef HPXFFT_GPU_SHARED_LOOP_HPP
#define HPXFFT_GPU_SHARED_LOOP_HPP

#include "../gpu/loop.hpp"

namespace hpxfft::shared
{
using vector_2d = hpxfft::gpu::vector_2d;
using loop = hpxfft::gpu::loop;
}  // namespace hpxfft::shared

#endif
