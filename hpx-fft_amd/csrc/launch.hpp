// launch.hpp -- host-side dispatch from runtime sizes to kernel instantiations.
#pragma once

#include "fft_kernels.hpp"
#include "fft_row_big.hpp"
#include "fft_row_pair.hpp"
#include "fft_tile.hpp"
#include <algorithm>

namespace hpxfft_host {

using hpxfft_dev::ColArgs;
using hpxfft_dev::RowArgs;

enum prec_id { P_F64 = 0, P_F32 = 1, P_F16 = 2, P_BF16 = 3 };

constexpr int ROW_MAX_LG = 14;       // Ny <= 32768
constexpr int COL_TILE_MIN_LG = 5;   // four-step sub-FFT lengths 32..256
constexpr int COL_TILE_MAX_LG = 8;

struct geom
{
    int nt;     // threads per workgroup
    int nseq;   // rows (row kernels) or columns W (column kernels) per workgroup
    int E;      // elements per thread
    bool split; // LDS exchange by real/imag halves
};

// per precision (defined in inst_<prec>.hip)
#define HPXFFT_DECL_LAUNCH(sfx)                                                                                    \
    hipError_t row_r2c_##sfx(int lg, const RowArgs &a, unsigned gx, unsigned gy, hipStream_t s);                   \
    hipError_t row_r2c_pair_##sfx(int lg, const RowArgs &a, unsigned total, hipStream_t s);                       \
    hipError_t row_c2r_##sfx(int lg, bool halve, const RowArgs &a, unsigned gx, unsigned gy, hipStream_t s);       \
    hipError_t col_##sfx(int lg, int mode, bool conj, const ColArgs &a, unsigned gx, unsigned gy, unsigned gz,    \
                         hipStream_t s);                                                                           \
    geom row_geom_##sfx(int lg);                                                                                   \
    geom col_geom_##sfx(int lg);                                                                                   \
    geom tile_geom_##sfx(int lg);                                                                                  \
    int col_direct_max_lg_##sfx();

HPXFFT_DECL_LAUNCH(f64)
HPXFFT_DECL_LAUNCH(f32)
HPXFFT_DECL_LAUNCH(f16)
HPXFFT_DECL_LAUNCH(bf16)

// rows of 2^13 / 2^14 complex points with lane-pair radix-32 stages
// (fft_row_pair.hpp; FP32 arithmetic: FP32 and 16-bit storage); `total` rows
inline hipError_t row_r2c_pair(int p, int lg, const RowArgs &a, unsigned total, hipStream_t s)
{
    switch (p)
    {
    case P_F32: return row_r2c_pair_f32(lg, a, total, s);
    case P_F16: return row_r2c_pair_f16(lg, a, total, s);
    case P_BF16: return row_r2c_pair_bf16(lg, a, total, s);
    default: return hipErrorInvalidValue;
    }
}
inline bool row_pair_covers(int prec, int lg) { return prec != P_F64 && (lg == 13 || lg == 14); }

inline hipError_t row_r2c(int p, int lg, const RowArgs &a, unsigned gx, unsigned gy, hipStream_t s)
{
    switch (p)
    {
    case P_F64: return row_r2c_f64(lg, a, gx, gy, s);
    case P_F32: return row_r2c_f32(lg, a, gx, gy, s);
    case P_F16: return row_r2c_f16(lg, a, gx, gy, s);
    default: return row_r2c_bf16(lg, a, gx, gy, s);
    }
}
// halve: adjoint (interior bins halved); otherwise the unnormalised inverse c2r
inline hipError_t row_c2r(int p, int lg, bool halve, const RowArgs &a, unsigned gx, unsigned gy, hipStream_t s)
{
    switch (p)
    {
    case P_F64: return row_c2r_f64(lg, halve, a, gx, gy, s);
    case P_F32: return row_c2r_f32(lg, halve, a, gx, gy, s);
    default: return hipErrorInvalidValue;
    }
}
inline hipError_t col(int p, int lg, int mode, bool conj, const ColArgs &a, unsigned gx, unsigned gy, unsigned gz,
                      hipStream_t s)
{
    switch (p)
    {
    case P_F64: return col_f64(lg, mode, conj, a, gx, gy, gz, s);
    case P_F32: return col_f32(lg, mode, conj, a, gx, gy, gz, s);
    case P_F16: return col_f16(lg, mode, conj, a, gx, gy, gz, s);
    default: return col_bf16(lg, mode, conj, a, gx, gy, gz, s);
    }
}
inline geom row_geom(int p, int lg)
{
    switch (p)
    {
    case P_F64: return row_geom_f64(lg);
    case P_F32: return row_geom_f32(lg);
    case P_F16: return row_geom_f16(lg);
    default: return row_geom_bf16(lg);
    }
}
inline geom col_geom(int p, int lg)
{
    switch (p)
    {
    case P_F64: return col_geom_f64(lg);
    case P_F32: return col_geom_f32(lg);
    case P_F16: return col_geom_f16(lg);
    default: return col_geom_bf16(lg);
    }
}
inline geom tile_geom(int p, int lg)
{
    switch (p)
    {
    case P_F64: return tile_geom_f64(lg);
    case P_F32: return tile_geom_f32(lg);
    case P_F16: return tile_geom_f16(lg);
    default: return tile_geom_bf16(lg);
    }
}
inline int col_direct_max_lg(int p)
{
    switch (p)
    {
    case P_F64: return col_direct_max_lg_f64();
    case P_F32: return col_direct_max_lg_f32();
    case P_F16: return col_direct_max_lg_f16();
    default: return col_direct_max_lg_bf16();
    }
}

}   // namespace hpxfft_host

// ---------------------------------------------------------------------------
// instantiation helpers, used by inst_<prec>.hip only
#ifdef HPXFFT_INSTANTIATE
namespace hpxfft_host {
namespace detail {
using namespace hpxfft_dev;

template <typename T, typename S, int LG> constexpr bool use_big_row()
{
    return LG == 11 || LG == 13 || LG == 14;
}
template <typename T, typename S, int LG>
hipError_t launch_row_r2c(const RowArgs &a, unsigned gx, unsigned gy, hipStream_t s)
{
    if constexpr (use_big_row<T, S, LG>())
    {
        (void)gx;
        const unsigned total = (unsigned)a.nrows * gy;
        hipLaunchKernelGGL((row_r2c_big_kernel<T, S, S, LG>), dim3(total), dim3((1 << LG) / (16 * big_nb<T, LG>())), 0,
                           s, a);
    }
    else
    {
        hipLaunchKernelGGL((row_r2c_kernel<T, S, S, LG>), dim3(gx, gy), dim3(cfg<T, S>::row_NT(LG)), 0, s, a);
    }
    return hipGetLastError();
}

template <typename T, typename S>
hipError_t launch_row_r2c_pair(int lg, const RowArgs &a, unsigned total, hipStream_t s)
{
    if constexpr (sizeof(T) == 8) return hipErrorInvalidValue;   // FP64 rows keep the radix-16 kernel
    else if (lg == 13) hipLaunchKernelGGL((row_r2c_pair_kernel<T, S, S, 13>), dim3(total), dim3(256), 0, s, a);
    else if (lg == 14) hipLaunchKernelGGL((row_r2c_pair_kernel<T, S, S, 14>), dim3(total), dim3(512), 0, s, a);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

template <typename T, typename S, int LG>
hipError_t launch_row_c2r(bool halve, const RowArgs &a, unsigned gx, unsigned gy, hipStream_t s)
{
    if constexpr (use_big_row<T, S, LG>())
    {
        // long rows: the radix-16 stages of the forward kernel, one row per workgroup
        (void)gx;
        const unsigned total = (unsigned)a.nrows * gy;
        const dim3 nt((1 << LG) / (16 * big_nb<T, LG>()));
        if (halve) hipLaunchKernelGGL((row_c2r_big_kernel<T, S, LG, true>), dim3(total), nt, 0, s, a);
        else hipLaunchKernelGGL((row_c2r_big_kernel<T, S, LG, false>), dim3(total), nt, 0, s, a);
    }
    else if (halve)
        hipLaunchKernelGGL((row_c2r_kernel<T, S, S, LG, true>), dim3(gx, gy), dim3(cfg<T, S>::row_NT(LG)), 0, s, a);
    else
        hipLaunchKernelGGL((row_c2r_kernel<T, S, S, LG, false>), dim3(gx, gy), dim3(cfg<T, S>::row_NT(LG)), 0, s, a);
    return hipGetLastError();
}
template <typename T, typename S, int LG, int MODE, bool CONJ>
hipError_t launch_col(const ColArgs &a, unsigned gx, unsigned gy, unsigned gz, hipStream_t s)
{
    if constexpr (MODE != 0)
    {
        // four-step tiles: specialised two-stage kernel (scratch in the compute type), gz = grids
        hipLaunchKernelGGL((tile_kernel<T, S, LG, MODE, CONJ>), dim3(gx, gy, gz), dim3((cfg<T, S>::TW << LG) / 16), 0,
                           s, a);
    }
    else
    {
        hipLaunchKernelGGL((col_kernel<T, S, LG, CONJ>), dim3(gx, gy, gz), dim3(cfg<T, S>::col_NT(LG)), 0, s, a);
    }
    return hipGetLastError();
}

template <typename T, typename S, int LG = 0>
hipError_t row_r2c_sw(int lg, const RowArgs &a, unsigned gx, unsigned gy, hipStream_t s)
{
    if constexpr (LG > ROW_MAX_LG) return hipErrorInvalidValue;
    else return lg == LG ? launch_row_r2c<T, S, LG>(a, gx, gy, s) : row_r2c_sw<T, S, LG + 1>(lg, a, gx, gy, s);
}
template <typename T, typename S, int LG = 0>
hipError_t row_c2r_sw(int lg, bool halve, const RowArgs &a, unsigned gx, unsigned gy, hipStream_t s)
{
    if constexpr (LG > ROW_MAX_LG) return hipErrorInvalidValue;
    else
        return lg == LG ? launch_row_c2r<T, S, LG>(halve, a, gx, gy, s)
                        : row_c2r_sw<T, S, LG + 1>(lg, halve, a, gx, gy, s);
}
template <typename T, typename S, bool ADJ, int LG = 0>
hipError_t col_sw(int lg, int mode, bool conj, const ColArgs &a, unsigned gx, unsigned gy, unsigned gz, hipStream_t s)
{
    constexpr int DMAX = cfg<T, S>::col_direct_max_lg;
    if constexpr (LG > DMAX && LG > COL_TILE_MAX_LG) return hipErrorInvalidValue;
    else
    {
        if (lg == LG)
        {
            if (mode == 0)
            {
                if constexpr (LG <= DMAX)
                {
                    if (!conj) return launch_col<T, S, LG, 0, false>(a, gx, gy, gz, s);
                    if constexpr (ADJ) return launch_col<T, S, LG, 0, true>(a, gx, gy, gz, s);
                }
                return hipErrorInvalidValue;
            }
            if constexpr (LG >= COL_TILE_MIN_LG && LG <= COL_TILE_MAX_LG)
            {
                if (!conj)
                    return mode == 1 ? launch_col<T, S, LG, 1, false>(a, gx, gy, gz, s)
                                     : launch_col<T, S, LG, 2, false>(a, gx, gy, gz, s);
                if constexpr (ADJ)
                    return mode == 1 ? launch_col<T, S, LG, 1, true>(a, gx, gy, gz, s)
                                     : launch_col<T, S, LG, 2, true>(a, gx, gy, gz, s);
            }
            return hipErrorInvalidValue;
        }
        return col_sw<T, S, ADJ, LG + 1>(lg, mode, conj, a, gx, gy, gz, s);
    }
}

template <typename T, typename S, bool ADJ>
hipError_t row_c2r_maybe(int lg, bool halve, const RowArgs &a, unsigned gx, unsigned gy, hipStream_t s)
{
    if constexpr (ADJ) return row_c2r_sw<T, S>(lg, halve, a, gx, gy, s);
    else return hipErrorInvalidValue;
}

template <typename T, typename S> geom row_geom_t(int lg)
{
    using C = cfg<T, S>;
    if (lg == 11 || lg == 13 || lg == 14)
    {
        const int nb = sizeof(T) == 4 && lg >= 13 ? 2 : 1;   // as big_nb<T, LG>()
        return geom{(1 << lg) / (16 * nb), 1, 16 * nb, (long)((1 << lg) + (1 << lg) / 16) * 2 * (long)sizeof(T) > 40 * 1024};
    }
    return geom{C::row_NT(lg), C::row_NSEQ(lg), C::row_E(lg), C::row_SPLIT(lg)};
}
template <typename T, typename S> geom col_geom_t(int lg)
{
    using C = cfg<T, S>;
    return geom{C::col_NT(lg), C::W, C::col_E(lg), C::col_SPLIT(lg)};
}
template <typename T, typename S> geom tile_geom_t(int lg)
{
    using C = cfg<T, S>;
    return geom{C::tile_NT(lg), C::TW, C::tile_E(lg), C::tile_SPLIT(lg)};
}
}   // namespace detail
}   // namespace hpxfft_host

#define HPXFFT_DEFINE_LAUNCH(sfx, T, S, ADJ)                                                                         \
    namespace hpxfft_host {                                                                                        \
    hipError_t row_r2c_##sfx(int lg, const RowArgs &a, unsigned gx, unsigned gy, hipStream_t s)                    \
    {                                                                                                              \
        return detail::row_r2c_sw<T, S>(lg, a, gx, gy, s);                                                         \
    }                                                                                                              \
    hipError_t row_r2c_pair_##sfx(int lg, const RowArgs &a, unsigned total, hipStream_t s)                        \
    {                                                                                                              \
        return detail::launch_row_r2c_pair<T, S>(lg, a, total, s);                                                 \
    }                                                                                                              \
    hipError_t row_c2r_##sfx(int lg, bool halve, const RowArgs &a, unsigned gx, unsigned gy, hipStream_t s)        \
    {                                                                                                              \
        return detail::row_c2r_maybe<T, S, ADJ>(lg, halve, a, gx, gy, s);                                          \
    }                                                                                                              \
    hipError_t col_##sfx(int lg, int mode, bool conj, const ColArgs &a, unsigned gx, unsigned gy, unsigned gz,     \
                         hipStream_t s)                                                                            \
    {                                                                                                              \
        return detail::col_sw<T, S, ADJ>(lg, mode, conj, a, gx, gy, gz, s);                                        \
    }                                                                                                              \
    geom row_geom_##sfx(int lg) { return detail::row_geom_t<T, S>(lg); }                                           \
    geom col_geom_##sfx(int lg) { return detail::col_geom_t<T, S>(lg); }                                           \
    geom tile_geom_##sfx(int lg) { return detail::tile_geom_t<T, S>(lg); }                                         \
    int col_direct_max_lg_##sfx() { return hpxfft_dev::cfg<T, S>::col_direct_max_lg; }                             \
    }
#endif
