// fft_tile.hpp -- four-step column tiles for FP64 / FP32 storage.
//
// A tile is W adjacent columns (one 512-byte row segment for FP64) of
// L = 16*R1 rows (L = 32..256), transformed by W*L/16 threads with 16 points
// each: a radix-16 stage in VGPRs, one LDS exchange, then R1-point stages
// (16/R1 butterflies per thread).  LDS layout [e][w]; consecutive threads
// take consecutive columns, so every LDS and global access of a wave covers
// whole row segments.
//   MODE 1 (phase A): rows m1 + P*e of the spectrum -> e-th DFT output k2,
//          times w_Nx^{m1 k2} -> scratch ((strip*Q + k2)*P + m1)*W + w
//   MODE 2 (phase B): scratch ((strip*Q + k2)*P + e)*W + w -> DFT over m1 ->
//          spectrum row Q*k1 + k2
// CONJ computes the inverse transform as conj(DFT(conj(.))).
//
// SS is the spectrum's storage type; the scratch holds the compute type T.
// 16-bit SS (T = float): phase A unscales row m by 2^-row_exp[m] on load and
// stores each column's largest scratch component of the workgroup to
// colpart[m1][col] (float bits); colmax_reduce_kernel folds them into
// colmax[col]; phase B stores X * 2^e per column with
// e = pick_exp(1.5 P colmax), since |X[Q k1 + k2]| <= sum_m1 |Z[m1][k2]|
// <= P sqrt(2) max component, and writes e to out_exp.
#pragma once

#include "fft_kernels.hpp"

namespace hpxfft_dev {

// FP32 tiles: twiddles w^{r j}, r = 1..R-1, formed in FP64 from the FP64
// table values w, w^2, w^4, w^8 (products of at most three) and rounded once
template <typename T, int R>
__device__ __forceinline__ void twiddle_r64(cpx<T> *a, const cpx<double> *__restrict__ tab, int jstep)
{
    auto app = [&](int r, const cpx<double> &w) { a[r] = cmul(a[r], cpx<T>{(T)w.x, (T)w.y}); };
    const cpx<double> w1 = tab[jstep];
    app(1, w1);
    if constexpr (R >= 4)
    {
        const cpx<double> w2 = tab[2 * jstep], w3 = cmul(w2, w1);
        app(2, w2);
        app(3, w3);
        if constexpr (R >= 8)
        {
            const cpx<double> w4 = tab[4 * jstep];
            const cpx<double> w5 = cmul(w4, w1), w6 = cmul(w4, w2), w7 = cmul(w4, w3);
            app(4, w4);
            app(5, w5);
            app(6, w6);
            app(7, w7);
            if constexpr (R == 16)
            {
                const cpx<double> w8 = tab[8 * jstep];
                app(8, w8);
                app(9, cmul(w8, w1));
                app(10, cmul(w8, w2));
                app(11, cmul(w8, w3));
                app(12, cmul(w8, w4));
                app(13, cmul(w8, w5));
                app(14, cmul(w8, w6));
                app(15, cmul(w8, w7));
            }
        }
    }
}

// twiddles w^{r j}, r = 1..R-1 (R <= 16) applied to a[1..R-1]; w^m = tab[m*step]
template <typename T, int R>
__device__ __forceinline__ void twiddle_r(cpx<T> *a, const cpx<T> *__restrict__ tab, int jstep)
{
    if constexpr (sizeof(T) == 8 && R >= 8)
    {
        // FP64: w, w^2, w^4 (and w^8) from the table, the rest as products
        const cpx<T> w1 = tab[jstep], w2 = tab[2 * jstep], w4 = tab[4 * jstep];
        const cpx<T> w3 = cmul(w2, w1);
        a[1] = cmul(a[1], w1);
        a[2] = cmul(a[2], w2);
        a[3] = cmul(a[3], w3);
        a[4] = cmul(a[4], w4);
        a[5] = cmul(a[5], cmul(w4, w1));
        a[6] = cmul(a[6], cmul(w4, w2));
        a[7] = cmul(a[7], cmul(w4, w3));
        if constexpr (R == 16)
        {
            const cpx<T> w8 = tab[8 * jstep];
            a[8] = cmul(a[8], w8);
#pragma unroll
            for (int r = 9; r < 16; ++r)
            {
                const int lo = r - 8;
                const cpx<T> wl = lo == 1 ? w1 : lo == 2 ? w2 : lo == 3 ? w3 : lo == 4 ? w4
                                  : lo == 5 ? cmul(w4, w1) : lo == 6 ? cmul(w4, w2) : cmul(w4, w3);
                a[r] = cmul(a[r], cmul(w8, wl));
            }
        }
    }
    else
    {
#pragma unroll
        for (int r = 1; r < R; ++r) a[r] = cmul(a[r], tab[r * jstep]);
    }
}

template <typename T, int L, int W> struct tileinfo
{
    static constexpr int R1 = L / 16;        // second-stage radix (2, 4, 8, 16)
    static constexpr int NT = W * L / 16;
    static constexpr int JPC = L / 16;       // threads per column
    static constexpr int NB1 = 16 / R1;      // second-stage butterflies per thread
    // FP64 tiles, and tiles whose complex exchange would take 64 KiB or more,
    // exchange real and imaginary halves (half the LDS, more workgroups per
    // CU where registers allow)
    static constexpr bool SPLIT = sizeof(T) == 8 || (long)L * W * 2 * (long)sizeof(T) >= 64 * 1024;
};

template <typename T, typename SS, int LG, int MODE, bool CONJ>
__global__ void __launch_bounds__((cfg<T, SS>::TW << LG) / 16)
tile_kernel(ColArgs a)
{
    constexpr int L = 1 << LG;
    constexpr int W = cfg<T, SS>::TW;
    using TI = tileinfo<T, L, W>;
    constexpr int R1 = TI::R1, JPC = TI::JPC, NB1 = TI::NB1;
    static_assert(R1 >= 2 && R1 <= 16, "L = 32..256");
    constexpr bool SPLIT = TI::SPLIT;
    constexpr bool S16 = is16<SS>();
    using SIN = typename std::conditional<MODE == 2, T, SS>::type;    // scratch is the compute type
    using SOUT = typename std::conditional<MODE == 1, T, SS>::type;
    __shared__ T lds[SPLIT ? L * W : 2 * L * W];
    __shared__ unsigned smax[S16 ? W : 1];

    const int t = threadIdx.x;
    if constexpr (S16 && MODE == 1)
    {
        if (t < W) smax[t] = 0u;   // ordered before the atomics by the exchange barriers
    }
    const int w = t % W, jj = t / W;   // column in the tile, butterfly index in the column
    const blk3 bk = xcd_block();       // neighbouring strips of one sub share an XCD
    const int sub = bk.y;              // m1 (MODE 1) or k2 (MODE 2)
    const int strip = bk.x;
    const long long gb = bk.z;         // grid of a batch
    // this grid's spectrum, scratch slice and exponent / maximum arrays
    const SIN *gin = reinterpret_cast<const SIN *>(a.in) + 2 * gb * (MODE == 1 ? a.dist : a.sdist);
    SOUT *gout = reinterpret_cast<SOUT *>(a.out) + 2 * gb * (MODE == 1 ? a.sdist : a.dist);
    // phase B writes rows Q*k1 + k2; Q is a multiple of the elements per
    // 128-byte line, so all of them share one line offset.  Shifting the
    // tile's columns left by that offset makes every row segment whole lines
    // (no line is shared with the neighbouring tile); the shifted tiles still
    // partition the columns, the first `skew` of them read from the previous
    // strip's scratch block.
    constexpr int ELINE = 128 / (2 * (int)sizeof(SS));
    int skew = 0;
    if constexpr (MODE == 2) skew = (int)((a.out_line_off + gb * a.dist + (long long)sub * a.pitch) % ELINE);
    const int col0 = strip * W - skew;   // first column of the tile
    if (col0 >= a.ncols) return;         // a trailing phase-B tile (uniform across the workgroup)
    const int c = col0 + w;              // this thread's column
    const bool valid = c >= 0 && c < a.ncols;
    const long long sbase = (long long)strip * a.P * a.Q * W;   // phase A's scratch block (within the slice)
    const cpx<T> *twL = reinterpret_cast<const cpx<T> *>(a.tw_L);

    // ---- load: element e = jj + r*JPC of column w (stage-0 butterfly jj).
    // Lanes outside the grid load from an in-bounds column and then take 0:
    // no branch around the loads, so all 16 are in flight before the first
    // wait (a guarded load followed by a conversion cost one round trip each)
    const int cl = valid ? c : (c < 0 ? 0 : a.ncols - 1);
    cpx<T> v[16];
#pragma unroll
    for (int r = 0; r < 16; ++r)
    {
        const int e = jj + r * JPC;
        long long idx;
        if constexpr (MODE == 1) idx = (long long)(sub + a.P * e) * a.pitch + cl;
        else
        {
            const int blk = cl / W;   // scratch block holding column cl
            idx = (((long long)blk * a.Q + sub) * a.P + e) * W + (cl - blk * W);
        }
        cpx<T> z = stor<SIN>::ld(gin + 2 * idx);
        if (!valid) z = cpx<T>{T(0), T(0)};
        if constexpr (CONJ && MODE == 1) z = conjc(z);
        v[r] = z;
    }
    if constexpr (S16 && MODE == 1)
    {
        // unscale row m1 + P e by 2^-row_exp.  W is a multiple of 64, so jj
        // is the same on every lane of a wave: the 16 exponents are
        // wave-uniform reads, issued after (and waited for with) the data
        // loads rather than one round trip per element
        static_assert(W % 64 == 0, "16-bit tiles: one butterfly index per wave");
        if (a.row_exp)
        {
            const int jw = __builtin_amdgcn_readfirstlane(jj);
            const int *rexp = a.row_exp + gb * a.rexp_dist + sub;
            int ex[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) ex[r] = rexp[(long long)a.P * (jw + r * JPC)];
#pragma unroll
            for (int r = 0; r < 16; ++r) v[r] = scal(v[r], (T)ldexpf(1.0f, -ex[r]));
        }
    }
    // ---- stage 0: radix 16, Ns = 1; outputs at e = 16 jj + q
    dft16(v);
    // ---- exchange; stage 1 (radix R1, Ns = 16): butterfly j1 = jj + i*JPC reads e = j1 + 16 r
    if constexpr (!SPLIT)
    {
        cpx<T> *l = reinterpret_cast<cpx<T> *>(lds);
#pragma unroll
        for (int q = 0; q < 16; ++q) l[(16 * jj + q) * W + w] = v[q];
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NB1; ++i)
        {
            const int j1 = jj + i * JPC;
#pragma unroll
            for (int r = 0; r < R1; ++r) v[i * R1 + r] = l[(j1 + 16 * r) * W + w];
        }
    }
    else
    {
#pragma unroll
        for (int q = 0; q < 16; ++q) lds[(16 * jj + q) * W + w] = v[q].x;
        __syncthreads();
        T re[16];
#pragma unroll
        for (int i = 0; i < NB1; ++i)
        {
            const int j1 = jj + i * JPC;
#pragma unroll
            for (int r = 0; r < R1; ++r) re[i * R1 + r] = lds[(j1 + 16 * r) * W + w];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; ++q) lds[(16 * jj + q) * W + w] = v[q].y;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NB1; ++i)
        {
            const int j1 = jj + i * JPC;
#pragma unroll
            for (int r = 0; r < R1; ++r) v[i * R1 + r] = {re[i * R1 + r], lds[(j1 + 16 * r) * W + w]};
        }
    }
#pragma unroll
    for (int i = 0; i < NB1; ++i)
    {
        const int j1 = jj + i * JPC;
        // w_L^{r j1}: FP64 and 16-bit storage from their own tables, FP32 formed in FP64
        if constexpr (sizeof(T) == 8 || S16) twiddle_r<T, R1>(v + i * R1, twL, j1);
        else twiddle_r64<T, R1>(v + i * R1, reinterpret_cast<const cpx<double> *>(a.tw64_L), j1);
        dftR<T, R1>(v + i * R1);
    }
    // outputs: butterfly j1, q -> e = j1 + 16 q
    if constexpr (MODE == 1)
    {
        // post-twiddle w_Nx^{m1 e}, e = j1 + 16 q
        if (sub != 0)
        {
            const cpx<T> *twN = reinterpret_cast<const cpx<T> *>(a.tw_N);
            const int nmask = a.P * a.Q - 1;
            if constexpr (sizeof(T) == 8 || S16)
            {
                // w^{m1 j1} (per lane) * w^{16 m1 q} (uniform across the workgroup), both table values
                // (FP64, or the FP32 table for 16-bit storage)
#pragma unroll
                for (int i = 0; i < NB1; ++i)
                {
                    const int j1 = jj + i * JPC;
                    const cpx<T> wb = twN[sub * j1];
#pragma unroll
                    for (int q = 0; q < R1; ++q)
                    {
                        const cpx<T> wq = twN[(16 * sub * q) & nmask];
                        v[i * R1 + q] = cmul(v[i * R1 + q], q == 0 ? wb : cmul(wb, wq));
                    }
                }
            }
            else
            {
                // FP32: the same two factors from the FP64 table, multiplied in FP64, rounded once
                const cpx<double> *t64 = reinterpret_cast<const cpx<double> *>(a.tw64_N);
#pragma unroll
                for (int i = 0; i < NB1; ++i)
                {
                    const int j1 = jj + i * JPC;
                    const cpx<double> wb = t64[sub * j1];
#pragma unroll
                    for (int q = 0; q < R1; ++q)
                    {
                        const cpx<double> w = q == 0 ? wb : cmul(wb, t64[(16 * sub * q) & nmask]);
                        v[i * R1 + q] = cmul(v[i * R1 + q], cpx<T>{(T)w.x, (T)w.y});
                    }
                }
            }
        }
    }
    if constexpr (S16 && MODE == 1)
    {
        // this thread's column w: largest scratch component -> LDS -> colpart[m1][col]
        float m = 0.0f;
#pragma unroll
        for (int x = 0; x < 16; ++x) m = fmaxf(m, fmaxf(fabsf((float)v[x].x), fabsf((float)v[x].y)));
        atomicMax(&smax[w], __float_as_uint(m));
        __syncthreads();
        if (t < W && t < a.ncols - col0 && a.colpart)
            a.colpart[(gb * a.P + sub) * a.ncols + col0 + t] = smax[t];
    }
    T osc = T(1);
    if constexpr (S16 && MODE == 2)
    {
        if (valid)
        {
            const int ex = pick_exp(__uint_as_float(a.colmax[gb * a.ncols + c]) * 1.5f * (float)a.P);
            osc = (T)ldexpf(1.0f, ex);
            if (jj == 0 && a.out_exp) a.out_exp[gb * a.cexp_dist + c] = ex;   // the thread holding e = 0
        }
    }
#pragma unroll
    for (int i = 0; i < NB1; ++i)
    {
        const int j1 = jj + i * JPC;
#pragma unroll
        for (int q = 0; q < R1; ++q)
        {
            const int e = j1 + 16 * q;
            cpx<T> z = v[i * R1 + q];
            if constexpr (MODE == 1)
            {
                // FP64: streaming stores of the scratch and (below) of the
                // spectrum measured 1.4% faster over the three launches
                if constexpr (sizeof(SS) == 8) st_stream<SOUT>(gout + 2 * (sbase + ((long long)e * a.P + sub) * W + w), z);
                else stor<SOUT>::st(gout + 2 * (sbase + ((long long)e * a.P + sub) * W + w), z);
            }
            else
            {
                if constexpr (CONJ) z = conjc(z);
                if constexpr (S16) z = scal(z, osc);
                if (valid)
                {
                    if constexpr (sizeof(SS) == 8) st_stream<SOUT>(gout + 2 * ((long long)(a.Q * e + sub) * a.pitch + c), z);
                    else stor<SOUT>::st(gout + 2 * ((long long)(a.Q * e + sub) * a.pitch + c), z);
                }
            }
        }
    }
}

}   // namespace hpxfft_dev
