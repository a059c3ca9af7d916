// fft_batched.hpp -- many independent in-place 2-D r2c grids of uneven sizes
// (FP32), each dimension 64..2048.
//
// Two persistent launches: the row phase works through the row tasks of all
// grids, then the column phase through the column tasks of all grids (the
// kernel boundary is the only synchronisation).  A task always covers the
// same number of points, PTS: PTS/L rows of a grid with row length L =
// Ny/2, or a strip of PTS/Nx adjacent columns of a grid with Nx rows, so
// uneven grids still give equal-sized work items.  Workgroups take tasks
// from device ticket counters until the task table is exhausted.  Column
// strips of one grid share 128-byte lines with their neighbours, so the
// column table (neighbouring strips next to each other) is cut into 8
// contiguous ranges, one per XCD (workgroup i runs on XCD i mod 8), and
// neighbours run behind one L2; a workgroup whose range is exhausted takes
// tasks from the other ranges.  Row tasks share no lines and take tickets
// from one counter.
//
// Candidate configurations (the autotuner times them on the plan's grids):
//   rows:    CFG 0: PTS 4096, 16 points/thread, 256 threads
//            CFG 1: PTS 2048,  8 points/thread, 256 threads (radix 8)
//            CFG 2: PTS 8192, 16 points/thread, 512 threads
//            CFG 3: PTS 16384, 16 points/thread, 1024 threads
//   columns: CFG 0: PTS 8192, 16 points/thread, 512 threads
//            CFG 1: PTS 4096, 16 points/thread, 256 threads
//            CFG 2: PTS 8192,  8 points/thread, 1024 threads (radix 8)
//            CFG 3: PTS 16384, 16 points/thread, 1024 threads (a 2048-row
//                   grid's strip is 8 columns: 64-byte row segments, not 32)
#pragma once

#include "fft_kernels.hpp"

namespace hpxfft_dev {

struct BGrid
{
    float *data;     // in-place padded layout: row pitch ld floats = 2 (Ny/2+1)
    int nx, ny;
    int lgx, lgL;    // log2 Nx, log2 (Ny/2)
};

struct BTask
{
    int grid;
    int first;       // first row (row phase) or strip index (column phase)
};

constexpr int BATCH_MIN_LG = 6, BATCH_MAX_LG = 11;   // 64 .. 2048 per side

template <int CFG> struct brow_cfg;
template <> struct brow_cfg<0> { static constexpr int PTS = 4096, E = 16, NT = 256; };
template <> struct brow_cfg<1> { static constexpr int PTS = 2048, E = 8, NT = 256; };
template <> struct brow_cfg<2> { static constexpr int PTS = 8192, E = 16, NT = 512; };
template <> struct brow_cfg<3> { static constexpr int PTS = 16384, E = 16, NT = 1024; };
template <int CFG> struct bcol_cfg;
template <> struct bcol_cfg<0> { static constexpr int PTS = 8192, E = 16, NT = 512; };
template <> struct bcol_cfg<1> { static constexpr int PTS = 4096, E = 16, NT = 256; };
template <> struct bcol_cfg<2> { static constexpr int PTS = 8192, E = 8, NT = 1024; };
template <> struct bcol_cfg<3> { static constexpr int PTS = 16384, E = 16, NT = 1024; };

// largest LDS footprint over the row lengths of a configuration (complex elements)
template <int CFG> constexpr int brow_lds()
{
    return brow_cfg<CFG>::PTS + brow_cfg<CFG>::PTS / 16;
}
template <int CFG> constexpr int bcol_lds()
{
    return bcol_cfg<CFG>::PTS + bcol_cfg<CFG>::PTS / 16;
}

constexpr int BATCH_QUEUES = 8;   // ticket counters per phase: one task range per XCD

// the next task for this workgroup, or ntasks when every range is exhausted.
// NQ ranges (1 or BATCH_QUEUES): its own XCD's range first (counter x, tasks
// [ntasks x / NQ, ntasks (x+1) / NQ)), then the others in turn.  A counter
// past its range stays past it, so each workgroup ends after at most one
// miss per range.
template <int NQ>
__device__ __forceinline__ int take_ticket(unsigned *ticket, int ntasks, unsigned *shared_slot)
{
    if (threadIdx.x == 0)
    {
        const int x0 = (int)(blockIdx.x % NQ);
        int t = ntasks;
        for (int d = 0; d < NQ; ++d)
        {
            const int x = (x0 + d) % NQ;
            const int lo = (int)((long long)ntasks * x / NQ);
            const int hi = (int)((long long)ntasks * (x + 1) / NQ);
            if (lo >= hi) continue;
            const unsigned k = atomicAdd(&ticket[x], 1u);
            if (k < (unsigned)(hi - lo))
            {
                t = lo + (int)k;
                break;
            }
        }
        *shared_slot = (unsigned)t;
    }
    __syncthreads();
    const int t = (int)*shared_slot;
    __syncthreads();
    return t;
}

// non-inlined: keeps each size's index arithmetic inside the task loop
// (hoisted out of the persistent loop it would pin registers of all sizes)
template <int CFG, int LGL>
__device__ __attribute__((noinline)) void brow_task(const BGrid &g, int row0, const float *tw32, cpx<float> *lds)
{
    using C = brow_cfg<CFG>;
    constexpr int L = 1 << LGL;
    constexpr int NSEQ = C::PTS / L;
    using EN = Engine<float, L, C::E, NSEQ, false, false>;
    static_assert(EN::NT == C::NT, "thread count");
    static_assert(EN::LDS_ELEMS <= brow_lds<CFG>(), "LDS");
    RowArgs a{};
    a.in = g.data;
    a.out = g.data;
    a.ld_in = 2 * (L + 1);
    a.ld_out = 2 * (L + 1);
    a.nrows = g.nx;
    a.tw = tw32 + 2 * (2 * L);   // w_{2L} table
    row_r2c_body<float, float, float, L, C::E, NSEQ, false>(a, 0, row0, lds, nullptr);
}

template <int CFG>
__global__ void __launch_bounds__(brow_cfg<CFG>::NT)
batched_rows_kernel(const BGrid *__restrict__ grids, const BTask *__restrict__ tasks, int ntasks,
                    unsigned *ticket, const float *__restrict__ tw32)
{
    __shared__ cpx<float> lds[brow_lds<CFG>()];
    __shared__ unsigned slot;
    for (;;)
    {
        const int t = take_ticket<1>(ticket, ntasks, &slot);
        if (t >= ntasks) break;
        const BTask task = tasks[t];
        const BGrid g = grids[task.grid];
        switch (g.lgL)
        {
        case 5: brow_task<CFG, 5>(g, task.first, tw32, lds); break;
        case 6: brow_task<CFG, 6>(g, task.first, tw32, lds); break;
        case 7: brow_task<CFG, 7>(g, task.first, tw32, lds); break;
        case 8: brow_task<CFG, 8>(g, task.first, tw32, lds); break;
        case 9: brow_task<CFG, 9>(g, task.first, tw32, lds); break;
        default: brow_task<CFG, 10>(g, task.first, tw32, lds); break;
        }
        __syncthreads();   // LDS reuse by the next task
    }
}

template <int CFG, int LGX>
__device__ __attribute__((noinline)) void bcol_task(const BGrid &g, int strip, const float *tw32, cpx<float> *lds)
{
    using C = bcol_cfg<CFG>;
    constexpr int N = 1 << LGX;
    constexpr int W = C::PTS / N;
    using EN = Engine<float, N, C::E, W, true, false>;
    static_assert(EN::NT == C::NT, "thread count");
    static_assert(EN::LDS_ELEMS <= bcol_lds<CFG>(), "LDS");
    const int ncols = g.ny / 2 + 1;
    ColArgs a{};
    a.in = g.data;
    a.out = g.data;
    a.pitch = ncols;
    a.ncols = ncols;
    a.tw_L = tw32 + 2 * N;
    col_body<float, float, float, N, C::E, W, false, false>(a, 0, strip, lds, nullptr);
}

template <int CFG>
__global__ void __launch_bounds__(bcol_cfg<CFG>::NT)
batched_cols_kernel(const BGrid *__restrict__ grids, const BTask *__restrict__ tasks, int ntasks,
                    unsigned *ticket, const float *__restrict__ tw32)
{
    __shared__ cpx<float> lds[bcol_lds<CFG>()];
    __shared__ unsigned slot;
    for (;;)
    {
        const int t = take_ticket<BATCH_QUEUES>(ticket, ntasks, &slot);
        if (t >= ntasks) break;
        const BTask task = tasks[t];
        const BGrid g = grids[task.grid];
        switch (g.lgx)
        {
        case 6: bcol_task<CFG, 6>(g, task.first, tw32, lds); break;
        case 7: bcol_task<CFG, 7>(g, task.first, tw32, lds); break;
        case 8: bcol_task<CFG, 8>(g, task.first, tw32, lds); break;
        case 9: bcol_task<CFG, 9>(g, task.first, tw32, lds); break;
        case 10: bcol_task<CFG, 10>(g, task.first, tw32, lds); break;
        default: bcol_task<CFG, 11>(g, task.first, tw32, lds); break;
        }
        __syncthreads();
    }
}

}   // namespace hpxfft_dev
