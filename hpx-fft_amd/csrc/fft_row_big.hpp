// fft_row_big.hpp -- row r2c kernel for long rows, one row per workgroup,
// L/16 threads with 16 points each (L = Ny/2 complex points, L = 2^11, 2^13
// or 2^14: 16^k times a final radix RL of 8, 2 or 4).
//
// Stages: radix-16 Stockham stages in VGPRs with one LDS exchange each, then
// the final radix-RL stage fused with the r2c split.  The thread that owns
// butterflies j and NSL-j of the final stage holds both Z[k] and Z[L-k] for
// every output it produces, so the split needs no further LDS round trip.
// LDS addresses are pad(e) = e + e/16; each exchange address is a per-thread
// base plus a compile-time stride (immediate offsets).  FP64 rows exchange
// real and imaginary halves separately so the LDS footprint is 8 bytes per
// point (two workgroups per CU at L = 8192).
//
// 16-bit storage (FP32 arithmetic, stage twiddles formed in FP32): inputs
// are converted on load and scaled by 2^-in_exp[row] when given; the row's
// outputs stay in registers until the workgroup has their largest component,
// then are stored as X * 2^e with e = pick_exp(max) written to out_exp[row].
#pragma once

#include "fft_kernels.hpp"

namespace hpxfft_dev {

// NB radix-16 butterflies per thread per stage: L/(16 NB) threads with 16 NB
// points each (NB = 2 halves the workgroup of the longest FP32 rows so two
// fit on a CU).  Butterfly j = t + b NT of a stage reads d[j + r M], M = L/16.
template <typename T, int L, int NB = 1> struct bigrow
{
    static constexpr int LG = ilog2c(L);
    static constexpr int M = L / 16;                   // radix-16 butterflies per stage
    static constexpr int NT = M / NB;
    static constexpr int NR16 = LG / 4;                // radix-16 stages before the fused one
    static constexpr int RL = 1 << (LG % 4);           // radix of the fused final stage
    static constexpr int NSL = L / RL;                 // its Ns
    static constexpr int HALF = NSL / 2;               // fused-stage tasks j = 0..HALF
    static constexpr int NMAIN = HALF / NT;            // regular tasks per thread (j = t + i NT < HALF)
    static constexpr int LP = L + L / 16;
    // exchange real/imag halves when a complex row would take more than 40 KiB of LDS
    static constexpr bool SPLIT = (long)LP * 2 * (long)sizeof(T) > 40 * 1024;
    static constexpr int ns(int s) { return s == 0 ? 1 : 16 * ns(s - 1); }
    __device__ static __forceinline__ int pad(int e) { return e + (e >> 4); }
    static_assert(LG % 4 != 0 && NR16 >= 2 && NMAIN >= 1 && HALF % NT == 0 && M % 16 == 0,
                  "supported: L = 2^11, 2^13, 2^14");
};

// butterflies per thread: two for FP32 (and 16-bit storage) rows of 2^13 and
// 2^14 points, whose 256- / 512-thread workgroups then need <= 128 VGPRs
// (4 waves per SIMD) to run four / two per CU
template <typename T, int LG> constexpr int big_nb() { return sizeof(T) == 4 && LG >= 13 ? 2 : 1; }
template <typename T, int LG> constexpr int big_min_waves() { return big_nb<T, LG>() == 2 ? 4 : 1; }

// Stage twiddle tables (RowArgs::stw64 / stw32): for each radix-16 stage
// with twiddles, NS = 16 and 256 butterflies per group, four arrays of NS
// entries, w_{16 NS}^{k}, w_{16 NS}^{2k}, w_{16 NS}^{4k}, w_{16 NS}^{8k}
// (k < NS), at stage_tw_off<NS>() entries.  The values are those of the
// w_{2L} table at index m k L / (8 NS); laid out this way, the butterflies
// of a wave (consecutive k) read consecutive entries of 17 KiB of hot table,
// where gathers from the 2L-entry table touch a line per lane and, at
// L = 2^14 (a 512 KiB FP64 table), miss in L2 behind the streaming rows.
constexpr int STAGE_TW_ENTRIES = 4 * 16 + 4 * 256;
template <int NS> constexpr int stage_tw_off()
{
    static_assert(NS == 16 || NS == 256, "stage tables cover NS = 16 and 256");
    return NS == 16 ? 0 : 4 * 16;
}

// stage twiddles w_{16 NS}^{r k}, r = 1..15: table values w, w^2, w^4, w^8
// and products of at most 3 factors.  W = double: the FP64 tables, products
// formed in FP64 and (for FP32 rows) rounded once -- the only FP64 step of the
// FP32 path.  W = T = float (rows stored in 16 bits): the FP32 tables,
// products in FP32 (a few FP32 ulps, far below the 16-bit rounding of the
// outputs) at half the issue cost of the FP64 products.
template <typename T, int NS, typename W = double>
__device__ __forceinline__ void twiddle16(cpx<T> *v, const cpx<W> *__restrict__ stab, int k)
{
    const cpx<W> *s = stab + stage_tw_off<NS>();
    const cpx<W> w1 = s[k], w2 = s[NS + k], w4 = s[2 * NS + k], w8 = s[3 * NS + k];
    const cpx<W> w3 = cmul(w2, w1), w5 = cmul(w4, w1), w6 = cmul(w4, w2), w7 = cmul(w4, w3);
    // each factor is applied as soon as it is formed (fewer live registers)
    auto app = [&](int r, const cpx<W> &w) { v[r] = cmul(v[r], cpx<T>{(T)w.x, (T)w.y}); };
    app(1, w1);
    app(2, w2);
    app(3, w3);
    app(4, w4);
    app(5, w5);
    app(6, w6);
    app(7, w7);
    app(8, w8);
    app(9, cmul(w8, w1));
    app(10, cmul(w8, w2));
    app(11, cmul(w8, w3));
    app(12, cmul(w8, w4));
    app(13, cmul(w8, w5));
    app(14, cmul(w8, w6));
    app(15, cmul(w8, w7));
}

// TW32: stage twiddles from the FP32 stage tables stw32 (16-bit storage),
// else from the FP64 tables stw64
template <typename T, int L, int NB, int S, bool TW32 = false>
__device__ __forceinline__ void radix16_stages(cpx<T> *v, T *lds, const cpx<double> *__restrict__ stw64, int t,
                                               const cpx<T> *__restrict__ stw32 = nullptr)
{
    using B = bigrow<T, L, NB>;
    // stages 0 .. NR16-2 (each followed by an exchange); stage NR16-1 is
    // computed by the caller and handed to the fused stage
    if constexpr (S + 1 < B::NR16)
    {
        constexpr int NS = B::ns(S);
        // butterfly j: outputs e_q = (j/NS) 16 NS + k + q NS -> padded base + q*WS; next inputs j + r M
        constexpr int WS = NS == 1 ? 1 : NS + NS / 16;
        constexpr int RS = B::M + B::M / 16;
        int wb[NB], rb[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b)
        {
            const int j = t + b * B::NT, k = j & (NS - 1);
            if constexpr (NS > 1)
            {
                if constexpr (TW32) twiddle16<T, NS, T>(v + 16 * b, stw32, k);
                else twiddle16<T, NS>(v + 16 * b, stw64, k);
            }
            dft16(v + 16 * b);
            wb[b] = B::pad((j / NS) * 16 * NS + k);
            rb[b] = B::pad(j);
        }
        if constexpr (!B::SPLIT)
        {
            cpx<T> *l = reinterpret_cast<cpx<T> *>(lds);
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int q = 0; q < 16; ++q) l[wb[b] + q * WS] = v[16 * b + q];
            __syncthreads();
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int r = 0; r < 16; ++r) v[16 * b + r] = l[rb[b] + r * RS];
            __syncthreads();
        }
        else
        {
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int q = 0; q < 16; ++q) lds[wb[b] + q * WS] = v[16 * b + q].x;
            __syncthreads();
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int r = 0; r < 16; ++r) v[16 * b + r].x = lds[rb[b] + r * RS];
            __syncthreads();
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int q = 0; q < 16; ++q) lds[wb[b] + q * WS] = v[16 * b + q].y;
            __syncthreads();
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int r = 0; r < 16; ++r) v[16 * b + r].y = lds[rb[b] + r * RS];
            __syncthreads();
        }
        radix16_stages<T, L, NB, S + 1, TW32>(v, lds, stw64, t, stw32);
    }
}

// gather the fused-stage inputs: for task j (butterflies j and NSL-j),
// a[r] = d[j + r NSL], c[r] = d[NSL-j + r NSL]
template <typename T, int L, int NB, bool HALFSEL>
__device__ __forceinline__ void gather_fused(T *lds, const cpx<T> *v,
                                             cpx<T> (&a)[bigrow<T, L, NB>::NMAIN][bigrow<T, L, NB>::RL],
                                             cpx<T> (&c)[bigrow<T, L, NB>::NMAIN][bigrow<T, L, NB>::RL],
                                             cpx<T> (&s)[bigrow<T, L, NB>::RL], int t)
{
    using B = bigrow<T, L, NB>;
    constexpr int RL = B::RL, NSL = B::NSL, NT = B::NT;
    constexpr int NS = B::ns(B::NR16 - 1);
    constexpr int WS = NS == 1 ? 1 : NS + NS / 16;
    constexpr int RS = NSL + NSL / 16;
    // write this thread's last radix-16 outputs (one component when split)
#pragma unroll
    for (int b = 0; b < NB; ++b)
    {
        const int j = t + b * NT, k = j & (NS - 1);
        const int wb = B::pad((j / NS) * 16 * NS + k);
#pragma unroll
        for (int q = 0; q < 16; ++q)
        {
            if constexpr (!B::SPLIT) reinterpret_cast<cpx<T> *>(lds)[wb + q * WS] = v[16 * b + q];
            else lds[wb + q * WS] = HALFSEL ? v[16 * b + q].y : v[16 * b + q].x;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < B::NMAIN; ++i)
    {
        const int j = t + i * NT;   // < HALF
        const int pa = B::pad(j), pc = B::pad(NSL - j);
#pragma unroll
        for (int r = 0; r < RL; ++r)
        {
            const int ia = pa + r * RS, ic = (j == 0 ? pa : pc) + r * RS;
            if constexpr (!B::SPLIT)
            {
                a[i][r] = reinterpret_cast<const cpx<T> *>(lds)[ia];
                c[i][r] = reinterpret_cast<const cpx<T> *>(lds)[ic];
            }
            else if constexpr (HALFSEL)
            {
                a[i][r].y = lds[ia];
                c[i][r].y = lds[ic];
            }
            else
            {
                a[i][r].x = lds[ia];
                c[i][r].x = lds[ic];
            }
        }
    }
    if (t == B::HALF % NT)
    {
        const int ps = B::pad(B::HALF);
#pragma unroll
        for (int r = 0; r < RL; ++r)
        {
            if constexpr (!B::SPLIT) s[r] = reinterpret_cast<const cpx<T> *>(lds)[ps + r * RS];
            else if constexpr (HALFSEL) s[r].y = lds[ps + r * RS];
            else s[r].x = lds[ps + r * RS];
        }
    }
    __syncthreads();
}

// r2c split of one pair: A = Z[k], Bc = Z[L-k];
// X[k] = Fe + w_{2L}^k Fo, X[L-k] = conj(Fe - w_{2L}^k Fo)
template <typename T>
__device__ __forceinline__ void r2c_val(cpx<T> A, cpx<T> Bc, cpx<T> w, cpx<T> &xk, cpx<T> &xlk)
{
    const cpx<T> Bk = conjc(Bc);
    const cpx<T> Fe = scal(A + Bk, T(0.5));
    const cpx<T> D = scal(A - Bk, T(0.5));
    const cpx<T> Fo = {D.y, -D.x};
    const cpx<T> tt = cmul(w, Fo);
    xk = Fe + tt;
    xlk = conjc(Fe - tt);
}

template <typename T, typename SO>
__device__ __forceinline__ void r2c_pair(SO *orow, int L, int k, cpx<T> A, cpx<T> Bc, const cpx<T> *__restrict__ tw)
{
    cpx<T> x0, x1;
    r2c_val(A, Bc, tw[k], x0, x1);
    if constexpr (sizeof(SO) == 8)   // FP64: streaming stores (measured 0.5% faster over the three launches)
    {
        st_stream<SO>(orow + 2 * k, x0);
        st_stream<SO>(orow + 2 * (L - k), x1);
    }
    else
    {
        stor<SO>::st(orow + 2 * k, x0);
        stor<SO>::st(orow + 2 * (L - k), x1);
    }
}

// z * (-i)^k, exact
template <typename T> __device__ __forceinline__ cpx<T> rot_mi(cpx<T> z, int k)
{
    switch (k & 3)
    {
    case 0: return z;
    case 1: return cpx<T>{z.y, -z.x};
    case 2: return cpx<T>{-z.x, -z.y};
    default: return cpx<T>{-z.y, z.x};
    }
}

// final-stage twiddles of butterflies j and jp = NSL - j: w_{2L}^{2 r j} and
// w_{2L}^{2 r jp} = (-i)^{4 r / RL} conj(w_{2L}^{2 r j}); for RL <= 4 the
// second is formed exactly from the first (the table is octant-symmetric, so
// the values are the table's own), saving a load per factor
template <typename T, int L, int RL>
__device__ __forceinline__ void fused_twiddles(cpx<T> *a, cpx<T> *c, int j, const cpx<T> *__restrict__ tw)
{
    constexpr int NSL = L / RL;
    const int jp = NSL - j;
#pragma unroll
    for (int r = 1; r < RL; ++r)
    {
        const cpx<T> wa = tw[2 * r * j];
        a[r] = cmul(a[r], wa);
        if constexpr (RL <= 4) c[r] = cmul(c[r], rot_mi(conjc(wa), 4 * r / RL));
        else c[r] = cmul(c[r], tw[(2 * r * jp) & (2 * L - 1)]);
    }
}

// r2c twiddle w_{2L}^{j + q NSL}: for RL = 2, w^{j + NSL} = -i w^j exactly
template <typename T, int L, int RL>
__device__ __forceinline__ cpx<T> split_twiddle(const cpx<T> &wj, int j, int q, const cpx<T> *__restrict__ tw)
{
    if constexpr (RL == 2) return q == 0 ? wj : rot_mi(wj, 1);
    else return tw[j + q * (L / RL)];
}

template <typename T>
__device__ __forceinline__ float cmax(cpx<T> a, cpx<T> b)
{
    return (float)fmax(fmax(fabs(a.x), fabs(a.y)), fmax(fabs(b.x), fabs(b.y)));
}

// Transform one row held in v[] (stage-0 inputs: v[16 b + r] = z[t + b NT + r M])
// and store its Ny/2+1 outputs to orow (storage SO).  16-bit SO: smax is a
// workgroup slot cleared before the first barrier, oexp the row's exponent.
template <typename T, typename SO, int LG, int NB>
__device__ __forceinline__ void big_row_body(cpx<T> *v, T *lds, const cpx<T> *__restrict__ tw,
                                             const cpx<double> *__restrict__ stw64,
                                             const cpx<T> *__restrict__ stw32, SO *orow, int t,
                                             unsigned *smax, int *oexp)
{
    constexpr int L = 1 << LG;
    using B = bigrow<T, L, NB>;
    constexpr int NT = B::NT, RL = B::RL, NSL = B::NSL, HALF = B::HALF;

    constexpr bool TW32 = is16<SO>();   // 16-bit storage: FP32 twiddle products
    radix16_stages<T, L, NB, 0, TW32>(v, lds, stw64, t, stw32);
    {
        // last radix-16 stage; its outputs are handed to the fused stage through LDS
        constexpr int NS = B::ns(B::NR16 - 1);
#pragma unroll
        for (int b = 0; b < NB; ++b)
        {
            const int k = (t + b * NT) & (NS - 1);
            if constexpr (NS > 1)
            {
                if constexpr (TW32) twiddle16<T, NS, T>(v + 16 * b, stw32, k);
                else twiddle16<T, NS>(v + 16 * b, stw64, k);
            }
            dft16(v + 16 * b);
        }
    }
    cpx<T> a[B::NMAIN][RL], c[B::NMAIN][RL], s[RL];
    gather_fused<T, L, NB, false>(lds, v, a, c, s, t);
    if constexpr (B::SPLIT) gather_fused<T, L, NB, true>(lds, v, a, c, s, t);

    const bool self = t == HALF % NT;
    if (self)
    {
        // self-paired butterfly j = HALF: pairs (q, RL-1-q)
#pragma unroll
        for (int r = 1; r < RL; ++r) s[r] = cmul(s[r], tw[2 * r * HALF]);
        dftR<T, RL>(s);
    }

    if constexpr (!is16<SO>())
    {
        // FP64 / FP32 storage: store as produced.  FP64: the streaming hint on
        // the lanes of whole lines only; a thread's stores are NT or NSL
        // elements apart (multiples of 128 bytes), so one test per direction
        static_assert((NT * 2 * sizeof(SO)) % 128 == 0 && (NSL * 2 * sizeof(SO)) % 128 == 0, "line-aligned strides");
        const bool wf = in_whole_line<SO, 1>(orow + 2 * t, t & 63);
        const bool wb = in_whole_line<SO, -1>(orow + 2 * (L - t), t & 63);
#pragma unroll
        for (int i = 0; i < B::NMAIN; ++i)
        {
            const int j = t + i * NT;   // its mirror NSL - j is NSL for j == 0 (then c == a)
            // final-stage twiddles w_L^{r b} = w_{2L}^{2 r b}
            fused_twiddles<T, L, RL>(a[i], c[i], j, tw);
            dftR<T, RL>(a[i]);
            dftR<T, RL>(c[i]);
            // a[q] = Z[j + q NSL], c[q] = Z[jp + q NSL]
            if (j == 0)
            {
                const cpx<T> z0 = a[i][0];
                stor<SO>::st(orow, cpx<T>{z0.x + z0.y, T(0)});
                stor<SO>::st(orow + 2 * L, cpx<T>{z0.x - z0.y, T(0)});
#pragma unroll
                for (int q = 1; q < RL / 2; ++q) r2c_pair<T, SO>(orow, L, q * NSL, a[i][q], a[i][RL - q], tw);
                stor<SO>::st(orow + 2 * (L / 2), conjc(a[i][RL / 2]));
            }
            else
            {
                const cpx<T> wj = tw[j];
#pragma unroll
                for (int q = 0; q < RL; ++q)
                {
                    cpx<T> x0, x1;
                    r2c_val(a[i][q], c[i][RL - 1 - q], split_twiddle<T, L, RL>(wj, j, q, tw), x0, x1);
                    if constexpr (sizeof(SO) == 8)   // FP64: streaming stores, as in r2c_pair
                    {
                        st_stream_if<SO>(orow + 2 * (j + q * NSL), x0, wf);
                        st_stream_if<SO>(orow + 2 * (L - j - q * NSL), x1, wb);
                    }
                    else
                    {
                        stor<SO>::st(orow + 2 * (j + q * NSL), x0);
                        stor<SO>::st(orow + 2 * (L - j - q * NSL), x1);
                    }
                }
            }
        }
        if (self)
        {
#pragma unroll
            for (int q = 0; q < RL / 2; ++q) r2c_pair<T, SO>(orow, L, HALF + q * NSL, s[q], s[RL - 1 - q], tw);
        }
    }
    else
    {
        // 16-bit storage: outputs X[k] (xa) and X[L-k] (xb), k = j + q NSL,
        // kept until the row maximum is known
        cpx<T> xa[B::NMAIN][RL], xb[B::NMAIN][RL], ya[RL / 2], yb[RL / 2];
        float m = 0.0f;
#pragma unroll
        for (int i = 0; i < B::NMAIN; ++i)
        {
            const int j = t + i * NT;
            fused_twiddles<T, L, RL>(a[i], c[i], j, tw);
            dftR<T, RL>(a[i]);
            dftR<T, RL>(c[i]);
            if (j == 0)
            {
                // slot 0: X[0], X[L]; slots 1..RL/2-1: pairs; slot RL/2: X[L/2]; the rest unused
                const cpx<T> z0 = a[i][0];
                xa[i][0] = cpx<T>{z0.x + z0.y, T(0)};
                xb[i][0] = cpx<T>{z0.x - z0.y, T(0)};
#pragma unroll
                for (int q = 1; q < RL / 2; ++q) r2c_val(a[i][q], a[i][RL - q], tw[q * NSL], xa[i][q], xb[i][q]);
                xa[i][RL / 2] = conjc(a[i][RL / 2]);
                xb[i][RL / 2] = xa[i][RL / 2];
#pragma unroll
                for (int q = RL / 2 + 1; q < RL; ++q) xa[i][q] = xb[i][q] = cpx<T>{T(0), T(0)};
            }
            else
            {
                const cpx<T> wj = tw[j];
#pragma unroll
                for (int q = 0; q < RL; ++q)
                    r2c_val(a[i][q], c[i][RL - 1 - q], split_twiddle<T, L, RL>(wj, j, q, tw), xa[i][q], xb[i][q]);
            }
#pragma unroll
            for (int q = 0; q < RL; ++q) m = fmaxf(m, cmax(xa[i][q], xb[i][q]));
        }
        if (self)
        {
#pragma unroll
            for (int q = 0; q < RL / 2; ++q)
            {
                r2c_val(s[q], s[RL - 1 - q], tw[HALF + q * NSL], ya[q], yb[q]);
                m = fmaxf(m, cmax(ya[q], yb[q]));
            }
        }
        atomicMax(smax, __float_as_uint(m));
        __syncthreads();
        const int ex = pick_exp(__uint_as_float(*smax));
        const T sc = (T)ldexpf(1.0f, ex);
        if (t == 0 && oexp) *oexp = ex;
#pragma unroll
        for (int i = 0; i < B::NMAIN; ++i)
        {
            const int j = t + i * NT;
#pragma unroll
            for (int q = 0; q < RL; ++q)
            {
                const int k = j + q * NSL;
                if (j != 0 || q <= RL / 2)
                {
                    stor<SO>::st(orow + 2 * k, scal(xa[i][q], sc));
                    if (j != 0 || q != RL / 2) stor<SO>::st(orow + 2 * (L - k), scal(xb[i][q], sc));
                }
            }
        }
        if (self)
        {
#pragma unroll
            for (int q = 0; q < RL / 2; ++q)
            {
                const int k = HALF + q * NSL;
                stor<SO>::st(orow + 2 * k, scal(ya[q], sc));
                stor<SO>::st(orow + 2 * (L - k), scal(yb[q], sc));
            }
        }
    }
}

// Forward r2c of long rows: workgroup g transforms row g of the nrows x
// batch rows.  SI / SO: input / output storage (T, or 16-bit with T = float).
template <typename T, typename SI, typename SO, int LG>
__global__ void __launch_bounds__((1 << LG) / (16 * big_nb<T, LG>()), (big_min_waves<T, LG>()))
row_r2c_big_kernel(RowArgs a)
{
    constexpr int L = 1 << LG;
    constexpr int NB = big_nb<T, LG>();
    using B = bigrow<T, L, NB>;
    constexpr int NT = B::NT;
    __shared__ T lds[B::SPLIT ? B::LP : 2 * B::LP];
    __shared__ unsigned smax;
    const int t = threadIdx.x;
    if (is16<SO>() && t == 0) smax = 0u;   // ordered before its use by the stage barriers
    const int g = xcd_order((int)blockIdx.x, (int)gridDim.x);   // consecutive rows share an XCD
    const int b = g / a.nrows, row = g - b * a.nrows;
    const SI *irow = reinterpret_cast<const SI *>(a.in) + (long long)b * a.dist_in + (long long)row * a.ld_in;
    cpx<T> v[16 * NB];
    // FP32 / FP64 rows: streaming loads of the input (read once; measured
    // 6% faster for FP32 rows, 0.5% over the three FP64 launches).  FP64
    // rows take the hint only on the lanes whose 128-byte line lies inside
    // the wave's access, and load the line it shares with its neighbour
    // plainly: the row pass's fetch fell to its algorithmic bytes and the
    // 2^14 x 2^14 launch from 0.78 to 0.74 ms.  A thread's loads are NT or M
    // elements apart (multiples of 128 bytes), so one test holds for all of
    // them.  FP32 rows keep the hint everywhere: at their 128-VGPR bound the
    // test spilled registers and made them slower (0.41 -> 0.49 ms).
    static_assert(sizeof(SI) != 8 || ((NT * 2 * sizeof(SI)) % 128 == 0 && (B::M * 2 * sizeof(SI)) % 128 == 0),
                  "line-aligned strides");
    bool whole = true;
    if constexpr (sizeof(SI) == 8) whole = in_whole_line<SI, 1>(irow + 2 * t, t & 63);
#pragma unroll
    for (int u = 0; u < 16 * NB; ++u)
    {
        const SI *src = irow + 2 * (t + (u / 16) * NT + (u % 16) * B::M);
        if constexpr (sizeof(SI) >= 4) v[u] = ld_stream_if<SI>(src, whole);
        else v[u] = stor<SI>::ld(src);
    }
    if constexpr (is16<SI>())
    {
        if (a.in_exp)
        {
            const T sc = (T)ldexpf(1.0f, -a.in_exp[(long long)b * a.exp_dist + row]);
#pragma unroll
            for (int u = 0; u < 16 * NB; ++u) v[u] = scal(v[u], sc);
        }
    }
    SO *orow = reinterpret_cast<SO *>(a.out) + (long long)b * a.dist_out + (long long)row * a.ld_out;
    int *oexp = (is16<SO>() && a.out_exp) ? a.out_exp + (long long)b * a.exp_dist + row : nullptr;
    big_row_body<T, SO, LG, NB>(v, lds, reinterpret_cast<const cpx<T> *>(a.tw),
                                reinterpret_cast<const cpx<double> *>(a.stw64),
                                reinterpret_cast<const cpx<T> *>(a.stw32), orow, t, &smax, oexp);
}

// Backward rows of 2^11, 2^13 and 2^14 complex points on the radix-16 stages
// of the forward kernel: the row half of the adjoint (HALVE) and the inverse
// c2r.  From the half spectrum G[0..L] of a row:
//   H = G on 0<k<L (halved when HALVE), H[0] = Re G[0], H[L] = Re G[L],
//   Z[k] = (H[k] + conj H[L-k]) + i (H[k] - conj H[L-k]) conj(w_{2L}^k),
// and the output row is conj(DFT(conj Z)) as interleaved reals.  Each thread
// builds the pairs (k, L-k) of its tasks into LDS (loads of G, G[L-k] and the
// twiddle all issued before the first wait), the stages run as in the forward
// kernel, and the last stage is a plain radix-RL Stockham stage whose outputs
// j + q L/RL are stored as they are produced.  Workgroup g takes row g.
// waves per SIMD of the backward long rows: as the forward rows, and FP64
// rows of 8192 points held to 128 VGPRs so that two workgroups share a CU
template <typename T, int LG> constexpr int c2r_min_waves()
{
    return sizeof(T) == 8 && LG == 13 ? 4 : big_min_waves<T, LG>();
}

template <typename T, typename S, int LG, bool HALVE>
__global__ void __launch_bounds__((1 << LG) / (16 * big_nb<T, LG>()), (c2r_min_waves<T, LG>()))
row_c2r_big_kernel(RowArgs a)
{
    constexpr int L = 1 << LG;
    constexpr int NB = big_nb<T, LG>();
    using B = bigrow<T, L, NB>;
    constexpr int NT = B::NT, RL = B::RL, NSL = B::NSL, M = B::M;
    constexpr int HK = L / 2 + 1;                      // pair tasks k = 0 .. L/2
    constexpr int NTASK = (HK + NT - 1) / NT;
    constexpr int NF = NSL / NT;                       // last-stage butterflies per thread
    static_assert(NSL % NT == 0, "last-stage butterflies divide among the threads");
    __shared__ T lds[B::SPLIT ? B::LP : 2 * B::LP];
    const int t = threadIdx.x;
    const int g = xcd_order((int)blockIdx.x, (int)gridDim.x);   // consecutive rows share an XCD
    const int bat = g / a.nrows, row = g - bat * a.nrows;
    const S *irow = reinterpret_cast<const S *>(a.in) + (long long)bat * a.dist_in + (long long)row * a.ld_in;
    S *orow = reinterpret_cast<S *>(a.out) + (long long)bat * a.dist_out + (long long)row * a.ld_out;
    const cpx<T> *tw = reinterpret_cast<const cpx<T> *>(a.tw);
    const cpx<double> *stw64 = reinterpret_cast<const cpx<double> *>(a.stw64);

    // pair tasks in chunks of CH: each chunk's loads of G[k], G[L-k] and the
    // twiddle are issued together (tasks past L/2 read the pair L/2 and write
    // nothing); split rows build one component per pass, reading G again
    constexpr int CH = 4;
    auto pair = [&](int k, const cpx<T> &g0, const cpx<T> &g1, const cpx<T> &w, cpx<T> &z0, cpx<T> &z1) {
        if (k == 0)
        {
            // H0 = Re G0, HL = Re GL: Z0 = (H0 + HL) + i (H0 - HL)
            z0 = conjc(cpx<T>{g0.x + g1.x, g0.x - g1.x});
            z1 = z0;
        }
        else
        {
            const T hs = HALVE ? T(0.5) : T(1);
            const cpx<T> h0 = scal(g0, hs), h1 = scal(g1, hs);
            const cpx<T> s0 = h0 + conjc(h1), d0 = cmul(h0 - conjc(h1), conjc(w));
            const cpx<T> s1 = h1 + conjc(h0), d1 = cmul(h1 - conjc(h0), cpx<T>{-w.x, -w.y});   // conj w^{L-k} = -w^k
            z0 = conjc(cpx<T>{s0.x - d0.y, s0.y + d0.x});
            z1 = conjc(cpx<T>{s1.x - d1.y, s1.y + d1.x});
        }
    };
    cpx<T> v[16 * NB];
#pragma unroll
    for (int h = 0; h < (B::SPLIT ? 2 : 1); ++h)
    {
#pragma unroll 1
        for (int c0 = 0; c0 < NTASK; c0 += CH)
        {
            cpx<T> G0[CH], G1[CH], W[CH];
#pragma unroll
            for (int i = 0; i < CH; ++i)
            {
                const int k = min(t + (c0 + i) * NT, L / 2);
                G0[i] = stor<S>::ld(irow + 2 * k);
                G1[i] = stor<S>::ld(irow + 2 * (L - k));
                W[i] = tw[k];
            }
#pragma unroll
            for (int i = 0; i < CH; ++i)
            {
                const int k = t + (c0 + i) * NT;
                if (c0 + i >= NTASK || k > L / 2) continue;
                cpx<T> z0, z1;
                pair(k, G0[i], G1[i], W[i], z0, z1);
                if constexpr (!B::SPLIT)
                {
                    cpx<T> *l = reinterpret_cast<cpx<T> *>(lds);
                    l[B::pad(k)] = z0;
                    if (k != 0) l[B::pad(L - k)] = z1;
                }
                else
                {
                    lds[B::pad(k)] = h ? z0.y : z0.x;
                    if (k != 0) lds[B::pad(L - k)] = h ? z1.y : z1.x;
                }
            }
        }
        __syncthreads();
        // each thread's stage-0 inputs v[16 b + r] = z[t + b NT + r M]
#pragma unroll
        for (int u = 0; u < 16 * NB; ++u)
        {
            const int e = B::pad(t + (u / 16) * NT + (u % 16) * M);
            if constexpr (!B::SPLIT) v[u] = reinterpret_cast<const cpx<T> *>(lds)[e];
            else if (h) v[u].y = lds[e];
            else v[u].x = lds[e];
        }
        __syncthreads();
    }
    radix16_stages<T, L, NB, 0>(v, lds, stw64, t);
    constexpr int NS = B::ns(B::NR16 - 1);
    constexpr int WS = NS == 1 ? 1 : NS + NS / 16;
    constexpr int RS = NSL + NSL / 16;
#pragma unroll
    for (int b = 0; b < NB; ++b)
    {
        const int k = (t + b * NT) & (NS - 1);
        if constexpr (NS > 1) twiddle16<T, NS>(v + 16 * b, stw64, k);
        dft16(v + 16 * b);
    }
    // the last radix-16 stage's outputs to LDS; the radix-RL stage reads d[j + r NSL]
    cpx<T> f[NF][RL];
#pragma unroll
    for (int h = 0; h < (B::SPLIT ? 2 : 1); ++h)
    {
#pragma unroll
        for (int b = 0; b < NB; ++b)
        {
            const int j = t + b * NT, k = j & (NS - 1);
            const int wb = B::pad((j / NS) * 16 * NS + k);
#pragma unroll
            for (int q = 0; q < 16; ++q)
            {
                if constexpr (!B::SPLIT) reinterpret_cast<cpx<T> *>(lds)[wb + q * WS] = v[16 * b + q];
                else lds[wb + q * WS] = h ? v[16 * b + q].y : v[16 * b + q].x;
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NF; ++i)
        {
            const int pa = B::pad(t + i * NT);
#pragma unroll
            for (int r = 0; r < RL; ++r)
            {
                if constexpr (!B::SPLIT) f[i][r] = reinterpret_cast<const cpx<T> *>(lds)[pa + r * RS];
                else if (h) f[i][r].y = lds[pa + r * RS];
                else f[i][r].x = lds[pa + r * RS];
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < NF; ++i)
    {
        const int j = t + i * NT;
        // twiddles w_L^{r j} = w_{2L}^{2 r j}
#pragma unroll
        for (int r = 1; r < RL; ++r) f[i][r] = cmul(f[i][r], tw[2 * r * j]);
        dftR<T, RL>(f[i]);
        // u = DFT(conj Z); the row is conj(u): out[2e] = Re u, out[2e+1] = -Im u
#pragma unroll
        for (int q = 0; q < RL; ++q)
            stor<S>::st(orow + 2 * (j + q * NSL), cpx<T>{f[i][q].x, -f[i][q].y});
    }
}

}   // namespace hpxfft_dev
