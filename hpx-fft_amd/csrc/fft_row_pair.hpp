// fft_row_pair.hpp -- row r2c kernel for rows of L = 2^13 or 2^14 complex
// points (Ny = 2^14, 2^15), with radix-32 stages split across lane pairs.
//
// One workgroup per row, L/16 threads with 16 points each.  Lanes 2p and
// 2p+1 form a pair that owns one radix-32 butterfly: lane h holds the inputs
// q = h + 2 q1 (q1 = 0..15), runs a 16-point DFT over q1 in registers, lane 1
// scales by w_32^{k1}, and the two lanes exchange their 16 values with a DPP
// quad permutation (a VALU operation, no LDS), after which lane h holds the
// outputs k1 + 16 h:
//     y[k1 + 16 k0] = sum_{q0} (-1)^{q0 k0} w_32^{q0 k1} sum_{q1} x[q0 + 2 q1] w_16^{q1 k1}.
// Stockham stages (butterfly j, k = j mod Ns, x_q = A[j + q n/R] w_{R Ns}^{q k},
// B[(j/Ns) R Ns + k + q Ns] = DFT_R(x)_q) of radices 32, 32 and RL = L/1024:
// two LDS exchanges per row instead of the three of 16, 16, 16, RL stages.
// The last stage is fused with the r2c split X[k] = f(Z[k], Z[L-k]): at
// L = 2^14 (RL = 16) lane h of pair p takes butterfly p or 1024 - p, and the
// pair swaps the eight values each needs from the other (DPP again); at
// L = 2^13 (RL = 8) each lane takes both butterflies j and 1024 - j.
// Butterflies 0 and 512 pair with themselves (lanes 0 and 1 of the row).
//
// Stage twiddles are formed in FP64 from four table values per butterfly and
// rounded once (as the radix-16 kernel); the r2c split reads its w_{2L}^k.  LDS exchanges split real and imaginary halves (8 bytes
// per point for FP64 and 4 for FP32), padded by one slot per 16.  16-bit
// storage: inputs converted (and unscaled by in_exp) on load, outputs held
// until the row maximum is known, stored as X * 2^e with e in out_exp.
#pragma once

#include "fft_row_big.hpp"

namespace hpxfft_dev {

// the value of the other lane of the pair (lanes 2p, 2p+1): DPP quad_perm [1,0,3,2]
__device__ __forceinline__ float pair_swap(float x)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, true));
}
__device__ __forceinline__ double pair_swap(double x)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)b, 0xB1, 0xF, 0xF, true);
    const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), 0xB1, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <typename T> __device__ __forceinline__ cpx<T> pair_swap(cpx<T> z) { return {pair_swap(z.x), pair_swap(z.y)}; }

template <typename T, int L> struct pairrow
{
    static constexpr int NB = 2;                  // radix-32 butterflies per lane pair and stage
    static constexpr int NT = L / (16 * NB);      // threads, 16 NB points each
    static constexpr int NP = NT / 2;             // lane pairs
    static constexpr int S = L / 32;              // radix-32 butterflies per stage (= NB NP)
    static constexpr int RL = L / 1024;           // radix of the final (fused) stage
    static constexpr int NPAIR = 1024 / (2 * NT); // final-stage butterfly pairs (j, 1024 - j) per thread
    static constexpr int LP = L + L / 16;         // padded LDS positions
    static constexpr int RS1 = 2 * S + (2 * S) / 16;   // stage-1 read stride (positions 2S apart)
    static constexpr int RS2 = 1024 + 1024 / 16;       // final-stage read stride (1024 apart)
    __device__ static __forceinline__ int pad(int e) { return e + (e >> 4); }
    static_assert((L == 8192 || L == 16384) && 2 * RL * NPAIR == 16 * NB, "L = 2^13 or 2^14");
};

// v[r] *= u^r, r = 1..R-1, u = w_{2L}^{ui} (FP64 table): u, u^2, u^4, u^8 are
// table values and each factor a product of at most four of them, formed in
// FP64 and rounded once (four gathers per butterfly, as twiddle16)
template <typename T, int R>
__device__ __forceinline__ void twiddle_pow(cpx<T> *v, const cpx<double> *__restrict__ tw64, int ui)
{
    const cpx<double> u1 = tw64[ui];
    cpx<double> u2 = {1.0, 0.0}, u4 = {1.0, 0.0}, u8 = {1.0, 0.0};
    if constexpr (R > 2) u2 = tw64[2 * ui];
    if constexpr (R > 4) u4 = tw64[4 * ui];
    if constexpr (R > 8) u8 = tw64[8 * ui];
    auto app = [&](int r, const cpx<double> &f) { v[r] = cmul(v[r], cpx<T>{(T)f.x, (T)f.y}); };
    app(1, u1);
    if constexpr (R > 2)
    {
        const cpx<double> u3 = cmul(u2, u1);
        app(2, u2);
        app(3, u3);
        if constexpr (R > 4)
        {
            app(4, u4);
            app(5, cmul(u4, u1));
            app(6, cmul(u4, u2));
            app(7, cmul(u4, u3));
            if constexpr (R > 8)
            {
                const cpx<double> u7 = cmul(u4, u3);
                app(8, u8);
                app(9, cmul(u8, u1));
                app(10, cmul(u8, u2));
                app(11, cmul(u8, u3));
                app(12, cmul(u8, u4));
                app(13, cmul(u8, cmul(u4, u1)));
                app(14, cmul(u8, cmul(u4, u2)));
                app(15, cmul(u8, u7));
            }
        }
    }
}

// cos and sin of pi k / 16, k = 0..8 (w_32^k = cos - i sin; k > 8 by symmetry)
__device__ constexpr double PAIR_C32[9] = {1.0, 0.98078528040323044913, 0.92387953251128675613,
                                           0.83146961230254523708, 0.70710678118654752440,
                                           0.55557023301960222474, 0.38268343236508977173,
                                           0.19509032201612826785, 0.0};

// the radix-32 butterfly of a lane pair: lane h holds x[h + 2 q1] in v[q1]
// and ends with y[k1 + 16 h] in v[k1].  Lane 1 scales the 16-point DFT by
// f(k1) (w_32^{k1} times a factor common to its inputs), lane 0 by 1
template <typename T, typename F>
__device__ __forceinline__ void pair_dft32(cpx<T> *v, int h, F f)
{
    dft16(v);
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1)
    {
        const cpx<T> w = f(k1);   // k1 = 0: the common factor alone (1 in stage 0)
        v[k1] = cmul(v[k1], cpx<T>{h ? w.x : T(1), h ? w.y : T(0)});
    }
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1)
    {
        const cpx<T> o = pair_swap(v[k1]);
        v[k1] = h ? o - v[k1] : v[k1] + o;
    }
}

// NW blocks of WN values out (block b of v at wb[b] + i*WS), then NR blocks of
// RN values in (block c at rb[c] + i*RS); real and imaginary halves in turn,
// barriers on both sides of each half
template <typename T, int NW, int WN, int WS, int NR, int RN, int RS>
__device__ __forceinline__ void pair_exchange(T *lds, cpx<T> *v, const int *wb, const int *rb)
{
    static_assert(NW * WN == NR * RN, "as many values in as out");
#pragma unroll
    for (int half = 0; half < 2; ++half)
    {
#pragma unroll
        for (int b = 0; b < NW; ++b)
#pragma unroll
            for (int i = 0; i < WN; ++i) lds[wb[b] + i * WS] = half ? v[b * WN + i].y : v[b * WN + i].x;
        __syncthreads();
#pragma unroll
        for (int c = 0; c < NR; ++c)
#pragma unroll
            for (int i = 0; i < RN; ++i)
            {
                if (half) v[c * RN + i].y = lds[rb[c] + i * RS];
                else v[c * RN + i].x = lds[rb[c] + i * RS];
            }
        __syncthreads();
    }
}

template <typename T, typename SO>
__device__ __forceinline__ void pair_store(SO *orow, int k, cpx<T> x)
{
    if constexpr (sizeof(SO) == 8) st_stream<SO>(orow + 2 * k, x);   // FP64: streaming, as the other row kernels
    else stor<SO>::st(orow + 2 * k, x);
}

// Forward r2c of rows of L = 2^LG complex points: workgroup g transforms row g.
// 16 NB = 32 points per thread, at most 128 VGPRs (four waves per SIMD): two
// workgroups per CU at L = 2^14, four at L = 2^13, as the radix-16 kernel.
template <typename T, typename SI, typename SO, int LG>
__global__ void __launch_bounds__((1 << LG) / 32, 4) row_r2c_pair_kernel(RowArgs a)
{
    constexpr int L = 1 << LG;
    using PR = pairrow<T, L>;
    constexpr int S = PR::S, NB = PR::NB, NP = PR::NP, NT = PR::NT, RL = PR::RL, NPAIR = PR::NPAIR;
    __shared__ T lds[PR::LP];
    __shared__ unsigned smax;
    const int t = threadIdx.x, p = t >> 1, h = t & 1;
    if (is16<SO>() && t == 0) smax = 0u;   // ordered before its use by the exchange barriers
    const int g = xcd_order((int)blockIdx.x, (int)gridDim.x);   // consecutive rows share an XCD
    const int b = g / a.nrows, row = g - b * a.nrows;
    const SI *irow = reinterpret_cast<const SI *>(a.in) + (long long)b * a.dist_in + (long long)row * a.ld_in;
    const cpx<T> *tw = reinterpret_cast<const cpx<T> *>(a.tw);
    const cpx<double> *tw64 = reinterpret_cast<const cpx<double> *>(a.tw64);
    // the table of length N sits at entries [N, 2N) of one array (w_N^m, m < N):
    // a.tw is the 2L-point table, so the 1024-point one starts 2L - 1024 before it
    const cpx<T> *w1024 = tw - (2 * L - 1024);

    // ---- stage 0 (Ns = 1): butterflies j = p + c NP; lane h loads z[j + (h + 2 q1) S]
    cpx<T> v[16 * NB];
#pragma unroll
    for (int c = 0; c < NB; ++c)
#pragma unroll
        for (int q1 = 0; q1 < 16; ++q1)
        {
            const SI *src = irow + 2 * (p + c * NP + (h + 2 * q1) * S);
            if constexpr (sizeof(SI) >= 4) v[16 * c + q1] = ld_stream<SI>(src);
            else v[16 * c + q1] = stor<SI>::ld(src);
        }
    if constexpr (is16<SI>())
    {
        if (a.in_exp)
        {
            const T sc = (T)ldexpf(1.0f, -a.in_exp[(long long)b * a.exp_dist + row]);
#pragma unroll
            for (int i = 0; i < 16 * NB; ++i) v[i] = scal(v[i], sc);
        }
    }
    int wb[NB], rb[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c)
    {
        const int j = p + c * NP;
        pair_dft32<T>(v + 16 * c, h, [](int k1) {   // w_32^{k1}, a constant rounded once
            const double cs = k1 <= 8 ? PAIR_C32[k1] : -PAIR_C32[16 - k1];
            const double sn = k1 <= 8 ? PAIR_C32[8 - k1] : PAIR_C32[k1 - 8];
            return cpx<T>{(T)cs, (T)-sn};
        });
        // outputs y[k1 + 16 h] -> B[32 j + 16 h + k1]; stage 1 reads A[j + (h + 2 q1) S]
        wb[c] = PR::pad(32 * j + 16 * h);
        rb[c] = PR::pad(j + h * S);
    }
    pair_exchange<T, NB, 16, 1, NB, 16, PR::RS1>(lds, v, wb, rb);

    // ---- stage 1 (Ns = 32): butterfly j, k = j mod 32: x_q w_1024^{q k}, q = h + 2 q1,
    // i.e. w_512^{q1 k} on the inputs of both lanes and w_1024^k, common to
    // lane 1's inputs, folded into its output factors
#pragma unroll
    for (int c = 0; c < NB; ++c)
    {
        const int j = p + c * NP, k = j & 31;
        if (k) twiddle_pow<T, 16>(v + 16 * c, tw64, k * (L / 256));
        // lane 1: w_1024^k w_32^{k1} = w_1024^{k + 32 k1}, consecutive entries of
        // the 1024-point table across the lanes of a wave
        pair_dft32<T>(v + 16 * c, h, [&](int k1) { return w1024[k + 32 * k1]; });
        // outputs y[k1 + 16 h] -> B[(j/32) 1024 + (j mod 32) + 32 (k1 + 16 h)]
        wb[c] = PR::pad((j >> 5) * 1024 + (j & 31) + 512 * h);
    }

    // ---- final stage (Ns = 1024, radix RL) fused with the r2c split: thread t
    // takes the butterfly pairs (j, 1024 - j), j = t + c NT; for t = 0, c = 0
    // the pair is butterflies 0 and 512, which each pair with themselves
    int rf[2 * NPAIR];
#pragma unroll
    for (int c = 0; c < NPAIR; ++c)
    {
        const int j = t + c * NT;
        rf[2 * c] = PR::pad(j);
        rf[2 * c + 1] = PR::pad(j ? 1024 - j : 512);
    }
    pair_exchange<T, NB, 16, 34, 2 * NPAIR, RL, PR::RS2>(lds, v, wb, rf);

    SO *orow = reinterpret_cast<SO *>(a.out) + (long long)b * a.dist_out + (long long)row * a.ld_out;
    int *oexp = (is16<SO>() && a.out_exp) ? a.out_exp + (long long)b * a.exp_dist + row : nullptr;
    // each output pair (X[k], X[L-k]) goes through put(): stored at once for
    // FP64 / FP32 storage; for 16-bit storage held (slot i) until the row
    // maximum is known.  X[L/2] is one output (k = L - k).
    constexpr int NSLOT = NPAIR * RL;
    cpx<T> xk[NSLOT], xl[NSLOT], x0 = {T(0), T(0)}, xL = {T(0), T(0)};
    int kk[NSLOT];
    auto put = [&](int i, int k, cpx<T> A, cpx<T> Bc) {
        cpx<T> u, w;
        r2c_val(A, Bc, tw[k], u, w);
        if constexpr (!is16<SO>())
        {
            pair_store<T, SO>(orow, k, u);
            pair_store<T, SO>(orow, L - k, w);
        }
        else
        {
            kk[i] = k;
            xk[i] = u;
            xl[i] = w;
        }
    };
#pragma unroll
    for (int c = 0; c < NPAIR; ++c)
    {
        const int j1 = t + c * NT, j2 = j1 ? 1024 - j1 : 512;
        cpx<T> *za = v + 2 * RL * c, *zb = za + RL;
        if (j1) twiddle_pow<T, RL>(za, tw64, 2 * j1);   // w_L^{q j}
        twiddle_pow<T, RL>(zb, tw64, 2 * j2);
        dftR<T, RL>(za);   // za[q] = Z[j1 + 1024 q]
        dftR<T, RL>(zb);   // zb[q] = Z[j2 + 1024 q]
        if (j1 != 0)
        {
            // Z[L - j1 - 1024 q] = Z[j2 + 1024 (RL - 1 - q)]
#pragma unroll
            for (int q = 0; q < RL; ++q) put(RL * c + q, j1 + 1024 * q, za[q], zb[RL - 1 - q]);
        }
        else
        {
            // butterfly 0: Z[1024 q] pairs with Z[1024 (RL - q)]; X[0], X[L] from Z[0], X[L/2] = conj Z[L/2]
#pragma unroll
            for (int q = 1; q < RL / 2; ++q) put(q - 1, 1024 * q, za[q], za[RL - q]);
            x0 = cpx<T>{za[0].x + za[0].y, T(0)};
            xL = cpx<T>{za[0].x - za[0].y, T(0)};
            const cpx<T> xh = conjc(za[RL / 2]);
            if constexpr (!is16<SO>())
            {
                pair_store<T, SO>(orow, 0, x0);
                pair_store<T, SO>(orow, L, xL);
                pair_store<T, SO>(orow, L / 2, xh);
            }
            else
            {
                kk[RL / 2 - 1] = L / 2;
                xk[RL / 2 - 1] = xl[RL / 2 - 1] = xh;
            }
            // butterfly 512: Z[512 + 1024 q] pairs with Z[512 + 1024 (RL - 1 - q)]
#pragma unroll
            for (int q = 0; q < RL / 2; ++q) put(RL / 2 + q, 512 + 1024 * q, zb[q], zb[RL - 1 - q]);
        }
    }
    if constexpr (is16<SO>())
    {
        const bool has0 = t == 0;   // the lane with X[0] and X[L]
        float m = has0 ? cmax(x0, xL) : 0.0f;
#pragma unroll
        for (int i = 0; i < NSLOT; ++i) m = fmaxf(m, cmax(xk[i], xl[i]));
        atomicMax(&smax, __float_as_uint(m));
        __syncthreads();
        const int ex = pick_exp(__uint_as_float(smax));
        const T sc = (T)ldexpf(1.0f, ex);
        if (t == 0 && oexp) *oexp = ex;
        if (has0)
        {
            stor<SO>::st(orow, scal(x0, sc));
            stor<SO>::st(orow + 2 * L, scal(xL, sc));
        }
#pragma unroll
        for (int i = 0; i < NSLOT; ++i)
        {
            stor<SO>::st(orow + 2 * kk[i], scal(xk[i], sc));
            if (kk[i] != L / 2) stor<SO>::st(orow + 2 * (L - kk[i]), scal(xl[i], sc));
        }
    }
}

}   // namespace hpxfft_dev
