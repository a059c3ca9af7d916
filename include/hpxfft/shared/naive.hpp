// naive.hpp -- hpxfft::shared::naive resolved to the MI355X class hpxfft::gpu::naive,
// so code written against HPX-FFT's core/include/hpxfft/shared/naive.hpp (class
// surface at naive.hpp:23-27) builds unchanged against this include directory and
// links with -lhpxfft_hip.
#ifndef HPXFFT_GPU_SHARED_NAIVE_HPP
#define HPXFFT_GPU_SHARED_NAIVE_HPP

#include "../gpu/loop.hpp"

namespace hpxfft::shared
{
using vector_2d = hpxfft::gpu::vector_2d;
using naive = hpxfft::gpu::naive;
}  // namespace hpxfft::shared

#endif
