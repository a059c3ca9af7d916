// Launchers of the batched (uneven grids, FP32) persistent kernels.
#include "fft_batched.hpp"
#include "batched.hpp"

namespace hpxfft_host {

using namespace hpxfft_dev;

static_assert(hpxfft_host::BATCH_QUEUES == hpxfft_dev::BATCH_QUEUES, "ticket counters per phase");

int batched_pts(int phase, int cfg)
{
    if (phase == 0)
        return cfg == 0 ? brow_cfg<0>::PTS : cfg == 1 ? brow_cfg<1>::PTS : cfg == 2 ? brow_cfg<2>::PTS : brow_cfg<3>::PTS;
    return cfg == 0 ? bcol_cfg<0>::PTS : cfg == 1 ? bcol_cfg<1>::PTS : cfg == 2 ? bcol_cfg<2>::PTS : bcol_cfg<3>::PTS;
}

static int nthreads(int phase, int cfg)
{
    if (phase == 0)
        return cfg == 0 ? brow_cfg<0>::NT : cfg == 1 ? brow_cfg<1>::NT : cfg == 2 ? brow_cfg<2>::NT : brow_cfg<3>::NT;
    return cfg == 0 ? bcol_cfg<0>::NT : cfg == 1 ? bcol_cfg<1>::NT : cfg == 2 ? bcol_cfg<2>::NT : bcol_cfg<3>::NT;
}

template <typename K> static int resident(K kern, int nt)
{
    int dev = 0, cus = 0, per_cu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, nt, 0);
    return (cus > 0 ? cus : 1) * (per_cu > 0 ? per_cu : 1);
}

int batched_resident_blocks(int phase, int cfg)
{
    const int nt = nthreads(phase, cfg);
    if (phase == 0)
    {
        if (cfg == 0) return resident(batched_rows_kernel<0>, nt);
        if (cfg == 1) return resident(batched_rows_kernel<1>, nt);
        if (cfg == 2) return resident(batched_rows_kernel<2>, nt);
        return resident(batched_rows_kernel<3>, nt);
    }
    if (cfg == 0) return resident(batched_cols_kernel<0>, nt);
    if (cfg == 1) return resident(batched_cols_kernel<1>, nt);
    if (cfg == 2) return resident(batched_cols_kernel<2>, nt);
    return resident(batched_cols_kernel<3>, nt);
}

hipError_t batched_launch(int phase, int cfg, const void *grids, const void *tasks, int ntasks, unsigned *ticket,
                          const float *tw32, int nblocks, hipStream_t s)
{
    const BGrid *g = reinterpret_cast<const BGrid *>(grids);
    const BTask *t = reinterpret_cast<const BTask *>(tasks);
    const dim3 grid((unsigned)nblocks), block((unsigned)nthreads(phase, cfg));
    if (phase == 0)
    {
        if (cfg == 0) hipLaunchKernelGGL(batched_rows_kernel<0>, grid, block, 0, s, g, t, ntasks, ticket, tw32);
        else if (cfg == 1) hipLaunchKernelGGL(batched_rows_kernel<1>, grid, block, 0, s, g, t, ntasks, ticket, tw32);
        else if (cfg == 2) hipLaunchKernelGGL(batched_rows_kernel<2>, grid, block, 0, s, g, t, ntasks, ticket, tw32);
        else hipLaunchKernelGGL(batched_rows_kernel<3>, grid, block, 0, s, g, t, ntasks, ticket, tw32);
    }
    else
    {
        if (cfg == 0) hipLaunchKernelGGL(batched_cols_kernel<0>, grid, block, 0, s, g, t, ntasks, ticket, tw32);
        else if (cfg == 1) hipLaunchKernelGGL(batched_cols_kernel<1>, grid, block, 0, s, g, t, ntasks, ticket, tw32);
        else if (cfg == 2) hipLaunchKernelGGL(batched_cols_kernel<2>, grid, block, 0, s, g, t, ntasks, ticket, tw32);
        else hipLaunchKernelGGL(batched_cols_kernel<3>, grid, block, 0, s, g, t, ntasks, ticket, tw32);
    }
    return hipGetLastError();
}

size_t batched_grid_desc_bytes() { return sizeof(BGrid); }
size_t batched_task_desc_bytes() { return sizeof(BTask); }

void batched_fill_grid(void *dst, float *data, int nx, int ny)
{
    BGrid g;
    g.data = data;
    g.nx = nx;
    g.ny = ny;
    int lx = 0, ll = 0;
    while ((1 << lx) < nx) ++lx;
    while ((1 << ll) < ny / 2) ++ll;
    g.lgx = lx;
    g.lgL = ll;
    *reinterpret_cast<BGrid *>(dst) = g;
}

void batched_fill_task(void *dst, int grid, int first)
{
    BTask t;
    t.grid = grid;
    t.first = first;
    *reinterpret_cast<BTask *>(dst) = t;
}

}   // namespace hpxfft_host
