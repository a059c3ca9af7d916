// batched.hpp -- host interface of the batched persistent kernels
// (implemented in inst_batched.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace hpxfft_host {

constexpr int BATCH_NCFG = 4;   // candidate configurations per phase
constexpr int BATCH_MIN_SIDE = 64, BATCH_MAX_SIDE = 2048;
constexpr int BATCH_QUEUES = 8;   // ticket counters per phase (one task range per XCD); `ticket` points at them

// phase 0 = rows, 1 = columns
int batched_pts(int phase, int cfg);
int batched_resident_blocks(int phase, int cfg);
hipError_t batched_launch(int phase, int cfg, const void *grids, const void *tasks, int ntasks, unsigned *ticket,
                          const float *tw32, int nblocks, hipStream_t s);
size_t batched_grid_desc_bytes();
size_t batched_task_desc_bytes();
void batched_fill_grid(void *dst, float *data, int nx, int ny);
void batched_fill_task(void *dst, int grid, int first);

}   // namespace hpxfft_host
