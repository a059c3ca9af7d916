// sync.hpp -- hpxfft::shared::sync resolved to the MI355X class hpxfft::gpu::sync,
// so code written against HPX-FFT's core/include/hpxfft/shared/sync.hpp (class
// surface at sync.hpp:23-27) builds unchanged against this include directory and
// links with -lhpxfft_hip.
#ifndef HPXFFT_GPU_SHARED_SYNC_HPP
#define HPXFFT_GPU_SHARED_SYNC_HPP

#include "../gpu/loop.hpp"

namespace hpxfft::shared
{
using vector_2d = hpxfft::gpu::vector_2d;
using sync = hpxfft::gpu::sync;
}  // namespace hpxfft::shared

#endif
