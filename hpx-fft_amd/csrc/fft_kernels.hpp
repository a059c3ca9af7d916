// fft_kernels.hpp -- gfx950 kernels of the 2-D r2c path.
//
//   row_r2c_kernel : phase 1 of hpxfft::shared::loop::fft_2d_r2c_par
//                    (core/src/shared/loop.cpp:61-69): 1-D r2c of length Ny
//                    on every row, as an Ny/2-point complex Stockham FFT of
//                    z[n] = x[2n] + i x[2n+1] plus the r2c split, in VGPRs/LDS.
//   col_kernel     : phases 2-4 (loop.cpp:72-101) for Nx <= 1024: the
//                    transposes are fused into the column kernel's
//                    row-segment loads/stores, and the length-Nx c2c runs in
//                    one workgroup per strip of columns.  Longer columns use
//                    the four-step tiles of fft_tile.hpp.
//   row_c2r_kernel : the row half of the adjoint (no reference counterpart):
//                    halve interior bins, keep Re of DC/Nyquist, unnormalised
//                    c2r of length Ny.
#pragma once

#include "fft_device.hpp"
#include <type_traits>

namespace hpxfft_dev {

// ---------------------------------------------------------------------------
// compile-time launch configurations (shared by host dispatch and kernels)
template <typename T, typename S> struct cfg
{
    static constexpr int lds_budget = 64 * 1024;
    // row kernel, L = 2^lg complex points per row
    static constexpr int row_E(int lg) { return lg >= 4 ? 16 : (1 << lg); }
    static constexpr int row_NT(int lg) { return lg <= 11 ? 128 : (1 << lg) / 16; }
    static constexpr int row_NSEQ(int lg) { return row_NT(lg) * row_E(lg) / (1 << lg); }
    static constexpr int lp(int L) { return L >= 16 ? L + L / 16 : L; }
    static constexpr bool row_SPLIT(int lg)
    {
        return (long)row_NSEQ(lg) * lp(1 << lg) * 2 * (long)sizeof(T) > lds_budget;
    }
    // direct column kernel: W adjacent columns = one 128-byte row segment of storage
    static constexpr int W = 128 / (2 * (int)sizeof(S)) > 32 ? 32 : 128 / (2 * (int)sizeof(S));
    static constexpr int col_direct_max_lg = sizeof(S) == 2 ? 9 : 10;
    static constexpr int col_E(int lg)
    {
        return (1 << lg) < 16 ? (1 << lg) : ((W << lg) <= 1024 && lg <= 8 ? 8 : 16);
    }
    static constexpr int col_NT(int lg) { return (W << lg) / col_E(lg); }
    static constexpr bool col_SPLIT(int lg)
    {
        // LDS layout [e][W]: padded only when a W-row is narrower than 128 B
        return (long)(W << lg) * 2 * (long)sizeof(T) > lds_budget;
    }
    // four-step tiles: TW adjacent columns x L rows (row segments of 512 B for
    // FP64 and FP32 storage, 256 B for 16-bit storage)
    static constexpr int TW = sizeof(S) == 8 ? 32 : 64;
    static constexpr int tile_E(int) { return 16; }
    static constexpr int tile_NT(int lg) { return (TW << lg) / 16; }
    static constexpr bool tile_SPLIT(int lg) { return (long)(TW << lg) * 2 * (long)sizeof(T) > lds_budget; }
};

struct RowArgs
{
    const void *in;
    void *out;
    long long ld_in;     // elements of S between input rows (reals)
    long long ld_out;    // elements of S between output rows (reals; 2 per complex)
    long long dist_in;   // between grids (batch)
    long long dist_out;
    int nrows;
    const void *tw;        // w_{2L}^m, m < 2L  (compute precision)
    const void *tw64;      // the same table in FP64 (twiddle products of FP32 rows are formed in FP64)
    const void *stw64;     // radix-16 stage twiddles of the long-row kernels (stage_tw_off), FP64
    const void *stw32;     // the same in FP32
    const int *in_exp;     // optional per-row input exponent (value = stored * 2^-e)
    int *out_exp;          // 16-bit outputs: per-row exponent written here
    long long exp_dist;    // between grids, for in_exp/out_exp
};

// column-pass arguments; MODE 0 = direct column kernel, 1 = four-step phase A
// (spectrum -> scratch), 2 = four-step phase B (scratch -> spectrum)
struct ColArgs
{
    const void *in;    // spectrum (MODE 0,1) or scratch (MODE 2)
    void *out;         // spectrum (MODE 0,2) or scratch (MODE 1)
    long long pitch;   // complex elements between spectrum rows
    long long dist;    // complex elements between spectrum grids (batch)
    int ncols;         // valid columns of the spectrum (Ny/2+1)
    int P, Q;          // four-step split Nx = P*Q; scratch index ((strip*Q + k2)*P + m1)*TW + w
    int out_line_off;  // four-step phase B: (output address / complex size) mod elements per 128-byte line
    long long sdist;   // four-step: scratch complex elements per grid
    const void *tw_L;  // w_L^m, m < L (sub-FFT length)
    const void *tw_N;  // w_Nx^m, m < Nx (MODE 1 post-twiddle)
    const void *tw64_L;  // the same tables in FP64 (FP32 four-step tiles form their twiddles in FP64)
    const void *tw64_N;
    const int *row_exp;    // 16-bit inputs (MODE 0,1): per-row exponent of the rows
    unsigned *colpart;     // 16-bit four-step phase A: per-workgroup column maxima [m1][ncols] (float bits)
    unsigned *colmax;      // 16-bit four-step phase B: per-column maxima (reduced from colpart)
    int *out_exp;          // 16-bit outputs (MODE 0,2): per-column exponent
    long long rexp_dist;   // between grids for row_exp
    long long cexp_dist;   // between grids for out_exp
};

template <typename S> constexpr bool is16() { return sizeof(S) == 2; }

// exponent e with max * 2^e in [2^14, 2^15): keeps fp16 clear of overflow
__device__ __forceinline__ int pick_exp(float m)
{
    if (!(m > 0.0f) || !isfinite(m)) return 0;
    return 14 - ilogbf(m);
}

// ---------------------------------------------------------------------------
// NSEQ rows [row0, row0+NSEQ) of grid b; the body is shared by the row
// kernel and the batched (uneven grids) row-phase kernel.
template <typename T, typename SI, typename SO, int L, int E, int NSEQ, bool SPLIT>
__device__ __forceinline__ void row_r2c_body(const RowArgs &a, long long b, int row0,
                                             typename Engine<T, L, E, NSEQ, false, SPLIT>::lds_t *lds,
                                             unsigned *smax)
{
    using EN = Engine<T, L, E, NSEQ, false, SPLIT>;
    constexpr int NT = EN::NT;
    constexpr int HK = L / 2 + 1;                        // r2c pair tasks per row
    constexpr int NTASK = (NSEQ * HK + NT - 1) / NT;

    const int t = threadIdx.x;
    const SI *in = reinterpret_cast<const SI *>(a.in) + b * a.dist_in;
    SO *out = reinterpret_cast<SO *>(a.out) + b * a.dist_out;
    const cpx<T> *tw = reinterpret_cast<const cpx<T> *>(a.tw);

    // 16-bit storage: unguarded loads from in-bounds rows (rows past the end
    // read the last row and take 0), so all E are in flight before the first
    // wait; wider storage keeps the guarded loads (as col_body)
    cpx<T> v[E];
#pragma unroll
    for (int x = 0; x < E; ++x)
    {
        int seq, e;
        EN::in_elem(t, x, seq, e);
        const int row = row0 + seq, rowc = row < a.nrows ? row : a.nrows - 1;
        cpx<T> z = {T(0), T(0)};
        if constexpr (is16<SI>())
        {
            z = stor<SI>::ld(in + (long long)rowc * a.ld_in + 2 * e);
            if (row >= a.nrows) z = cpx<T>{T(0), T(0)};
        }
        else if (row < a.nrows)
            z = stor<SI>::ld(in + (long long)row * a.ld_in + 2 * e);
        v[x] = z;
    }
    if (a.in_exp)
    {
#pragma unroll
        for (int x = 0; x < E; ++x)
        {
            int seq, e;
            EN::in_elem(t, x, seq, e);
            const int row = row0 + seq, rowc = row < a.nrows ? row : a.nrows - 1;
            v[x] = scal(v[x], (T)ldexp(1.0, -a.in_exp[b * a.exp_dist + rowc]));
        }
    }

    EN::run(v, lds, tw, 2, t);

    // r2c split: X[k] = Fe + w_{2L}^k Fo, X[L-k] = conj(Fe - w_{2L}^k Fo)
    cpx<T> A[NTASK], B[NTASK];
    if constexpr (!SPLIT)
    {
#pragma unroll
        for (int x = 0; x < E; ++x) { int seq, e; EN::out_elem(t, x, seq, e); lds[EN::addr(seq, e)] = v[x]; }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NTASK; ++i)
        {
            const int tau = t + i * NT;
            const int seq = tau / HK, k = tau - seq * HK;
            if (tau < NSEQ * HK) { A[i] = lds[EN::addr(seq, k & (L - 1))]; B[i] = lds[EN::addr(seq, (L - k) & (L - 1))]; }
        }
    }
    else
    {
#pragma unroll
        for (int x = 0; x < E; ++x) { int seq, e; EN::out_elem(t, x, seq, e); lds[EN::addr(seq, e)] = v[x].x; }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NTASK; ++i)
        {
            const int tau = t + i * NT;
            const int seq = tau / HK, k = tau - seq * HK;
            if (tau < NSEQ * HK) { A[i].x = lds[EN::addr(seq, k & (L - 1))]; B[i].x = lds[EN::addr(seq, (L - k) & (L - 1))]; }
        }
        __syncthreads();
#pragma unroll
        for (int x = 0; x < E; ++x) { int seq, e; EN::out_elem(t, x, seq, e); lds[EN::addr(seq, e)] = v[x].y; }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NTASK; ++i)
        {
            const int tau = t + i * NT;
            const int seq = tau / HK, k = tau - seq * HK;
            if (tau < NSEQ * HK) { A[i].y = lds[EN::addr(seq, k & (L - 1))]; B[i].y = lds[EN::addr(seq, (L - k) & (L - 1))]; }
        }
    }

    cpx<T> X0[NTASK], X1[NTASK];
    float lmax[NTASK];
#pragma unroll
    for (int i = 0; i < NTASK; ++i)
    {
        const int tau = t + i * NT;
        const int k = tau % HK;
        const cpx<T> Ak = A[i], Bk = conjc(B[i]);
        if (k == 0)
        {
            X0[i] = {Ak.x + Ak.y, T(0)};
            X1[i] = {Ak.x - Ak.y, T(0)};
        }
        else
        {
            const cpx<T> Fe = scal(Ak + Bk, T(0.5));
            const cpx<T> D = scal(Ak - Bk, T(0.5));
            const cpx<T> Fo = {D.y, -D.x};
            const cpx<T> w = tw[k];
            const cpx<T> tt = cmul(w, Fo);
            X0[i] = Fe + tt;
            X1[i] = conjc(Fe - tt);
        }
        if constexpr (is16<SO>())
        {
            lmax[i] = (float)fmax(fmax(fabs(X0[i].x), fabs(X0[i].y)), fmax(fabs(X1[i].x), fabs(X1[i].y)));
        }
    }

    if constexpr (is16<SO>())
    {
        for (int s = t; s < NSEQ; s += NT) smax[s] = 0u;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NTASK; ++i)
        {
            const int tau = t + i * NT;
            if (tau < NSEQ * HK) atomicMax(&smax[tau / HK], __float_as_uint(lmax[i]));
        }
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < NTASK; ++i)
    {
        const int tau = t + i * NT;
        const int seq = tau / HK, k = tau - seq * HK;
        const int row = row0 + seq;
        if (tau < NSEQ * HK && row < a.nrows)
        {
            cpx<T> x0 = X0[i], x1 = X1[i];
            if constexpr (is16<SO>())
            {
                const int ex = pick_exp(__uint_as_float(smax[seq]));
                const T s = (T)ldexpf(1.0f, ex);
                x0 = scal(x0, s);
                x1 = scal(x1, s);
                if (k == 0 && a.out_exp) a.out_exp[b * a.exp_dist + row] = ex;
            }
            SO *orow = out + (long long)row * a.ld_out;
            stor<SO>::st(orow + 2 * k, x0);
            if (k == 0) stor<SO>::st(orow + 2 * L, x1);
            else if (k != L - k) stor<SO>::st(orow + 2 * (L - k), x1);
        }
    }
}


template <typename T, typename SI, typename SO, int LG>
__global__ void __launch_bounds__((cfg<T, SI>::row_NT(LG)))
row_r2c_kernel(RowArgs a)
{
    constexpr int L = 1 << LG;
    using C = cfg<T, SI>;
    constexpr int E = C::row_E(LG), NSEQ = C::row_NSEQ(LG);
    constexpr bool SPLIT = C::row_SPLIT(LG);
    using EN = Engine<T, L, E, NSEQ, false, SPLIT>;
    __shared__ typename EN::lds_t lds[EN::LDS_ELEMS];
    __shared__ unsigned smax[is16<SO>() ? NSEQ : 1];
    const blk3 bk = xcd_block();
    row_r2c_body<T, SI, SO, L, E, NSEQ, SPLIT>(a, bk.y, bk.x * NSEQ, lds, smax);
}

// ---------------------------------------------------------------------------
// Direct column pass: one workgroup transforms a strip of W adjacent columns
// (one 128-byte row segment of storage) over all L = Nx rows of grid b, in
// place or from `in` to `out` (same layout).  Shared by the column kernel
// and the batched column-phase kernel.  16-bit storage: rows are unscaled by
// 2^-row_exp[m] on load, and each output column is stored as X * 2^e with e
// from its exact maximum, written to out_exp.
template <typename T, typename SI, typename SO, int L, int E, int W, bool SPLIT, bool CONJ>
__device__ __forceinline__ void col_body(const ColArgs &a, long long b, int strip,
                                         typename Engine<T, L, E, W, true, SPLIT>::lds_t *lds, unsigned *smax)
{
    using EN = Engine<T, L, E, W, true, SPLIT>;
    constexpr bool OUT16 = is16<SO>();
    const int t = threadIdx.x;
    const int col0 = strip * W;
    const int nvalid = min(W, a.ncols - col0);
    const SI *in = reinterpret_cast<const SI *>(a.in) + 2 * b * a.dist;   // dist counts complex elements
    SO *out = reinterpret_cast<SO *>(a.out) + 2 * b * a.dist;

    // 16-bit storage: unguarded loads from in-bounds columns (lanes past the
    // last column read it and take 0), so all E are in flight before the first
    // wait (a guarded load followed by its conversion cost one round trip
    // each).  Wider storage keeps the guarded loads, which the compiler
    // already issues together and which hold fewer registers in the
    // persistent batched kernel built on this body.
    cpx<T> v[E];
#pragma unroll
    for (int x = 0; x < E; ++x)
    {
        int w, e;
        EN::in_elem(t, x, w, e);
        cpx<T> z = {T(0), T(0)};
        if constexpr (is16<SI>())
        {
            const int wc = w < nvalid ? w : nvalid - 1;
            z = stor<SI>::ld(in + 2 * ((long long)e * a.pitch + col0 + wc));
            if (w >= nvalid) z = cpx<T>{T(0), T(0)};
        }
        else if (w < nvalid)
            z = stor<SI>::ld(in + 2 * ((long long)e * a.pitch + col0 + w));
        if constexpr (CONJ) z = conjc(z);
        v[x] = z;
    }
    if constexpr (is16<SI>())
    {
        if (a.row_exp)
        {
#pragma unroll
            for (int x = 0; x < E; ++x)
            {
                int w, e;
                EN::in_elem(t, x, w, e);
                v[x] = scal(v[x], (T)ldexpf(1.0f, -a.row_exp[b * a.rexp_dist + e]));
            }
        }
    }

    EN::run(v, lds, reinterpret_cast<const cpx<T> *>(a.tw_L), 1, t);

    if constexpr (OUT16)
    {
        // per-column maxima for the output exponents
        for (int s = t; s < W; s += EN::NT) smax[s] = 0u;
        __syncthreads();
#pragma unroll
        for (int x = 0; x < E; ++x)
        {
            int w, e;
            EN::out_elem(t, x, w, e);
            const float m = (float)fmax(fabs(v[x].x), fabs(v[x].y));
            atomicMax(&smax[w], __float_as_uint(m));
        }
        __syncthreads();
    }

#pragma unroll
    for (int x = 0; x < E; ++x)
    {
        int w, e;
        EN::out_elem(t, x, w, e);
        if (w < nvalid)
        {
            cpx<T> z = v[x];
            if constexpr (CONJ) z = conjc(z);
            if constexpr (OUT16)
            {
                const int ex = pick_exp(__uint_as_float(smax[w]));
                z = scal(z, (T)ldexpf(1.0f, ex));
                if (e == 0 && a.out_exp) a.out_exp[b * a.cexp_dist + col0 + w] = ex;
            }
            stor<SO>::st(out + 2 * ((long long)e * a.pitch + col0 + w), z);
        }
    }
}

// blockIdx.x = strip, blockIdx.z = grid (XCD-contiguous order)
template <typename T, typename S, int LG, bool CONJ>
__global__ void __launch_bounds__((cfg<T, S>::col_NT(LG)))
col_kernel(ColArgs a)
{
    constexpr int L = 1 << LG;
    using C = cfg<T, S>;
    constexpr int W = C::W, E = C::col_E(LG);
    constexpr bool SPLIT = C::col_SPLIT(LG);
    using EN = Engine<T, L, E, W, true, SPLIT>;
    __shared__ typename EN::lds_t lds[EN::LDS_ELEMS];
    __shared__ unsigned smax[is16<S>() ? W : 1];
    const blk3 bk = xcd_block();
    col_body<T, S, S, L, E, W, SPLIT, CONJ>(a, bk.z, bk.x, lds, smax);
}

// ---------------------------------------------------------------------------
// Row c2r.  Input: half-spectrum rows G[0..L] (complex, SI), output: real
// rows (SO).  out[2n] + i out[2n+1] = IDFT_L(Z),
//   H = G on 0<k<L (halved when HALVE), H[0] = Re G[0], H[L] = Re G[L],
//   Z[k] = (H[k] + conj H[L-k]) + i (H[k] - conj H[L-k]) conj(w_{2L}^k).
// Without halving this is the unnormalised c2r of length 2L (FFTW_BACKWARD,
// Hermitian extension of G); with halving it is the row half of the r2c
// adjoint under Re<A,B> on the half spectrum.  The inverse DFT is computed
// as conj(DFT(conj Z)).
template <typename T, typename SI, typename SO, int LG, bool HALVE>
__global__ void __launch_bounds__((cfg<T, SI>::row_NT(LG)))
row_c2r_kernel(RowArgs a)
{
    constexpr int L = 1 << LG;
    using C = cfg<T, SI>;
    constexpr int E = C::row_E(LG), NSEQ = C::row_NSEQ(LG);
    constexpr bool SPLIT = C::row_SPLIT(LG);
    using EN = Engine<T, L, E, NSEQ, false, SPLIT>;
    constexpr int NT = EN::NT;
    constexpr int HK = L / 2 + 1;
    constexpr int NTASK = (NSEQ * HK + NT - 1) / NT;

    __shared__ typename EN::lds_t lds[EN::LDS_ELEMS];

    const int t = threadIdx.x;
    const blk3 bk = xcd_block();
    const long long b = bk.y;
    const SI *in = reinterpret_cast<const SI *>(a.in) + b * a.dist_in;
    SO *out = reinterpret_cast<SO *>(a.out) + b * a.dist_out;
    const cpx<T> *tw = reinterpret_cast<const cpx<T> *>(a.tw);
    const int row0 = bk.x * NSEQ;

    // loads first, unguarded: tasks past the rows read the last row (an
    // in-bounds address) and take 0, so the 2 NTASK data loads and NTASK
    // twiddle loads are all in flight before the first wait (a guarded load
    // in a divergent branch is waited on inside that branch)
    cpx<T> G0[NTASK], G1[NTASK], W[NTASK];
#pragma unroll
    for (int i = 0; i < NTASK; ++i)
    {
        const int tau = t + i * NT;
        const int seq = tau / HK, k = tau - seq * HK;
        const int row = row0 + seq, rowc = row < a.nrows ? row : a.nrows - 1;
        const SI *irow = in + (long long)rowc * a.ld_in;
        G0[i] = stor<SI>::ld(irow + 2 * k);
        G1[i] = stor<SI>::ld(irow + 2 * (L - k));
        W[i] = tw[k];
    }
    // pre-processing: pairs (k, L-k) -> conj Z[k], conj Z[L-k]
    cpx<T> Z0[NTASK], Z1[NTASK];
#pragma unroll
    for (int i = 0; i < NTASK; ++i)
    {
        const int tau = t + i * NT;
        const int seq = tau / HK, k = tau - seq * HK;
        const int row = row0 + seq;
        const bool live = tau < NSEQ * HK && row < a.nrows;
        const cpx<T> g0 = live ? G0[i] : cpx<T>{T(0), T(0)};
        const cpx<T> g1 = live ? G1[i] : cpx<T>{T(0), T(0)};
        if (k == 0)
        {
            // H0 = Re G0, HL = Re GL: Z0 = (H0 + HL) + i (H0 - HL)
            Z0[i] = conjc(cpx<T>{g0.x + g1.x, g0.x - g1.x});
            Z1[i] = Z0[i];
        }
        else
        {
            const T hs = HALVE ? T(0.5) : T(1);
            const cpx<T> h0 = scal(g0, hs), h1 = scal(g1, hs);
            const cpx<T> w = W[i];
            const cpx<T> wc = conjc(w);
            // Z[k]
            const cpx<T> s0 = h0 + conjc(h1), d0 = cmul(h0 - conjc(h1), wc);
            const cpx<T> zk = {s0.x - d0.y, s0.y + d0.x};   // s0 + i d0
            // Z[L-k]: conj(w_{2L}^{L-k}) = -w_{2L}^k
            const cpx<T> s1 = h1 + conjc(h0), d1 = cmul(h1 - conjc(h0), cpx<T>{-w.x, -w.y});
            const cpx<T> zlk = {s1.x - d1.y, s1.y + d1.x};
            Z0[i] = conjc(zk);
            Z1[i] = conjc(zlk);
        }
    }
    cpx<T> v[E];
    if constexpr (!SPLIT)
    {
#pragma unroll
        for (int i = 0; i < NTASK; ++i)
        {
            const int tau = t + i * NT;
            const int seq = tau / HK, k = tau - seq * HK;
            if (tau < NSEQ * HK)
            {
                lds[EN::addr(seq, k & (L - 1))] = Z0[i];
                if (k != 0) lds[EN::addr(seq, L - k)] = Z1[i];
            }
        }
        __syncthreads();
#pragma unroll
        for (int x = 0; x < E; ++x) { int seq, e; EN::in_elem(t, x, seq, e); v[x] = lds[EN::addr(seq, e)]; }
        __syncthreads();
    }
    else
    {
#pragma unroll
        for (int h = 0; h < 2; ++h)
        {
#pragma unroll
            for (int i = 0; i < NTASK; ++i)
            {
                const int tau = t + i * NT;
                const int seq = tau / HK, k = tau - seq * HK;
                if (tau < NSEQ * HK)
                {
                    lds[EN::addr(seq, k & (L - 1))] = h ? Z0[i].y : Z0[i].x;
                    if (k != 0) lds[EN::addr(seq, L - k)] = h ? Z1[i].y : Z1[i].x;
                }
            }
            __syncthreads();
#pragma unroll
            for (int x = 0; x < E; ++x)
            {
                int seq, e;
                EN::in_elem(t, x, seq, e);
                if (h) v[x].y = lds[EN::addr(seq, e)];
                else v[x].x = lds[EN::addr(seq, e)];
            }
            __syncthreads();
        }
    }

    EN::run(v, lds, tw, 2, t);

#pragma unroll
    for (int x = 0; x < E; ++x)
    {
        int seq, e;
        EN::out_elem(t, x, seq, e);
        const int row = row0 + seq;
        if (row < a.nrows)
        {
            // u = DFT(conj Z); z = conj(u): out[2e] = Re u, out[2e+1] = -Im u
            stor<SO>::st(out + (long long)row * a.ld_out + 2 * e, cpx<T>{v[x].x, -v[x].y});
        }
    }
}

template <int UNUSED = 0> __global__ void zero_u32_kernel(unsigned *p, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// colmax[c] = max over m1 < P of colpart[m1 * ncols + c] (non-negative float
// bit patterns order as unsigned integers), per grid.  256 threads: 64
// columns x 4 slices of the P partial rows, combined through LDS.
template <int UNUSED = 0> __global__ void __launch_bounds__(256)
colmax_reduce_kernel(const unsigned *colpart, unsigned *colmax, int ncols, int P)
{
    __shared__ unsigned part[4][64];
    colpart += (long long)blockIdx.y * P * ncols;   // grid blockIdx.y of a batch
    colmax += (long long)blockIdx.y * ncols;
    const int lc = threadIdx.x & 63, sl = threadIdx.x >> 6;
    const int c = blockIdx.x * 64 + lc;
    unsigned m = 0u;
    if (c < ncols)
        for (int i = sl; i < P; i += 4) m = max(m, colpart[(long long)i * ncols + c]);
    part[sl][lc] = m;
    __syncthreads();
    if (sl == 0 && c < ncols) colmax[c] = max(max(part[0][lc], part[1][lc]), max(part[2][lc], part[3][lc]));
}

}   // namespace hpxfft_dev
