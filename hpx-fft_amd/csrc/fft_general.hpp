// fft_general.hpp -- 2-D r2c passes for lengths that are not powers of two
// (FP64 / FP32, and 16-bit storage within one workgroup).  The reference
// accepts any even Ny and any Nx through FFTW (core/src/util/adapter_fftw.cpp:
// 9,28-29); these kernels cover the lengths whose prime factors are at most
// GEN_MAX_RADIX and whose data fit one workgroup's LDS, any other length
// whose Bluestein FFT (a power of two >= 2 len - 1) fits it, and longer
// lengths as two levels N = P Q (gen_strided_kernel, gen_rows_split_kernel),
// with the power-of-two kernels kept for power-of-two dimensions.
//
// One workgroup transforms NSEQ sequences of length n held in LDS (sequence
// s at [s*n, (s+1)*n), two ping-pong buffers in dynamic LDS sized to the
// plan), by Stockham autosort stages of runtime radices r (product n):
//   stage with sub-length Ns, butterfly j < n/r, k = j mod Ns:
//     x_q = A[j + q n/r] w_{r Ns}^{q k},  y = DFT_r(x),
//     B[(j / Ns) r Ns + k + q Ns] = y_q
// Radices 2, 3, 4, 5, 7 and 8 run as register butterflies (one LDS pass per
// stage); larger primes (up to 61) by a generic stage of two passes.
//
// Direct plans whose radices are all at most 8 can run in place instead, in
// one buffer (half the LDS, so twice the workgroups per CU): the loads scatter
// element e to position pos(e), the mixed-radix digit reversal of e, and
// decimation-in-time stages r_1, ..., r_S then leave X[k] at position k:
//   stage with sub-length Ls, butterfly (block c, j < Ls), base = c r Ls + j:
//     x_q = A[base + q Ls] w_{r Ls}^{q j},  A[base + q Ls] = DFT_r(x)_q;
//   e = d_S + r_S (d_{S-1} + r_{S-1} (... + r_2 d_1)), d_s < r_s:
//     pos(e) = sum_s d_s r_1 ... r_{s-1}.
// The plan's table holds pos() as ints after the twiddles.  Bluestein plans
// run in place too: decimation-in-frequency stages r_S, ..., r_1 take the
// natural order to X[k] at pos(k), the product with the kernel spectrum
// reads that spectrum stored in the same order, and the decimation-in-time
// stages bring the inverse transform back to natural order.
// Twiddles and the roots w_r^e = w_n^{e n/r} come from the plan's exact FP64
// table (rounded once for FP32); a register butterfly loads w^k and forms
// w^{qk} (q < r <= 8) as FP64 powers of it.
#pragma once

#include "fft_kernels.hpp"

namespace hpxfft_dev {

constexpr int GEN_MAX_STAGES = 24;
constexpr int GEN_MAX_RADIX = 61;
// threads per workgroup: 256, or up to GEN_NT_MAX for large LDS blocks (each
// kernel is built for both bounds, NTMAX = 256 and 1024; loops stride by the
// launch's blockDim)
constexpr int GEN_NT = 256;
constexpr int GEN_NT_MAX = 1024;
constexpr int GEN_LDS_BYTES = 128 * 1024;    // the largest pair of sequence buffers of a workgroup
template <typename T> constexpr int gen_max_points() { return GEN_LDS_BYTES / (2 * 2 * (int)sizeof(T)); }

// loop stride of a workgroup: the 256-thread build knows it at compile time,
// the larger build reads it from the launch (512 or 1024)
template <int NTMAX> __device__ __forceinline__ int gen_nth() { return NTMAX == GEN_NT ? GEN_NT : (int)blockDim.x; }

struct GenPlan
{
    int n;                          // FFT length run in LDS (the buffer stride of a sequence)
    int nstages;
    int radix[GEN_MAX_STAGES];      // stage radices, product n
    int len;                        // DFT length: n, or the length Bluestein's algorithm maps to n
    int off_c, off_b;               // Bluestein: chirp and kernel-spectrum offsets in the table
    int off_r;                      // >= 0: the pos() ints at this table offset (radices <= 7); -1: none
    int off_bp;                     // Bluestein: the kernel spectrum in pos() order at this offset
    int ip;                         // 1: in-place stages (needs off_r >= 0, or Bluestein); 0: ping-pong
};

// sequence buffers a plan's workgroup holds in LDS
__host__ __device__ inline int gen_nbuf(const GenPlan &g) { return g.ip ? 1 : 2; }

// LDS position of input element e of a sequence
__device__ __forceinline__ int gen_pos(const GenPlan &g, const cpx<double> *__restrict__ tab, int e)
{
    return g.ip && g.off_r >= 0 ? reinterpret_cast<const int *>(tab + g.off_r)[e] : e;
}

// x / d for 0 <= x < 2^20 and d >= 1, from inv = 1/d in FP32: the product is
// within 2^-21 * x of x/d, so the truncated quotient is off by at most one,
// which the remainder test corrects
__device__ __forceinline__ int gdiv(int x, int d, float inv)
{
    int q = (int)((float)x * inv);
    const int r = x - q * d;
    if (r < 0) --q;
    else if (r >= d) ++q;
    return q;
}

template <typename T> __device__ __forceinline__ cpx<T> tabv(const cpx<double> *__restrict__ tab, int m)
{
    const cpx<double> w = tab[m];
    return cpx<T>{(T)w.x, (T)w.y};
}

// DFT of R points in registers, forward sign; w[e] = w_R^e (e >= 1)
template <typename T, int R> __device__ __forceinline__ void dft_small(cpx<T> *x, const cpx<T> *w)
{
    if constexpr (R == 2)
    {
        const cpx<T> a = x[0], b = x[1];
        x[0] = a + b;
        x[1] = a - b;
    }
    else if constexpr (R == 4)
    {
        const cpx<T> s0 = x[0] + x[2], d0 = x[0] - x[2], s1 = x[1] + x[3], d1 = x[1] - x[3];
        x[0] = s0 + s1;
        x[2] = s0 - s1;
        x[1] = d0 + mul_mi(d1);
        x[3] = d0 - mul_mi(d1);
    }
    else if constexpr (R == 8)
    {
        // two 4-point DFTs of the even and odd inputs, X[k] = E[k] + w_8^k O[k],
        // X[k+4] = E[k] - w_8^k O[k] (w_8 = w[1], w_8^2 = -i, w_8^3 = w[3])
        cpx<T> e[4] = {x[0], x[2], x[4], x[6]}, o[4] = {x[1], x[3], x[5], x[7]};
        dft_small<T, 4>(e, w);
        dft_small<T, 4>(o, w);
        o[1] = cmul(o[1], w[1]);
        o[2] = mul_mi(o[2]);
        o[3] = cmul(o[3], w[3]);
#pragma unroll
        for (int k = 0; k < 4; ++k)
        {
            x[k] = e[k] + o[k];
            x[k + 4] = e[k] - o[k];
        }
    }
    else
    {
        cpx<T> y[R];
#pragma unroll
        for (int o = 0; o < R; ++o)
        {
            cpx<T> acc = x[0];
#pragma unroll
            for (int q = 1; q < R; ++q)
            {
                const int e = (q * o) % R;
                acc = e == 0 ? acc + x[q] : acc + cmul(x[q], w[e]);
            }
            y[o] = acc;
        }
#pragma unroll
        for (int o = 0; o < R; ++o) x[o] = y[o];
    }
}

// one Stockham stage of radix R (2, 3, 4, 5, 7, 8) over nseq sequences, A -> B
template <typename T, int R, int NTMAX>
__device__ void gen_stage_reg(const cpx<T> *A, cpx<T> *B, int nseq, int n, int Ns, const cpx<double> *__restrict__ tab)
{
    const int t = threadIdx.x, m = n / R, nb = nseq * m, tstep = n / (R * Ns);
    const float inv_m = 1.0f / (float)m, inv_ns = 1.0f / (float)Ns;
    cpx<T> w[R];
#pragma unroll
    for (int e = 1; e < R; ++e) w[e] = tabv<T>(tab, e * m);
    w[0] = cpx<T>{T(1), T(0)};
    // two butterflies per iteration (independent loads in flight together);
    // twiddles w^{qk} as powers of the one table value w^k, in FP64
    const int NTH = gen_nth<NTMAX>();
    for (int b0 = t; b0 < nb; b0 += 2 * NTH)
    {
        cpx<T> x[2][R];
        cpx<double> w1[2];
        int dst[2] = {-1, -1}, kk[2] = {0, 0};
#pragma unroll
        for (int u = 0; u < 2; ++u)
        {
            const int b = b0 + u * NTH;
            if (b < nb)
            {
                const int seq = gdiv(b, m, inv_m), j = b - seq * m;
                const int jb = gdiv(j, Ns, inv_ns), k = j - jb * Ns;
                const cpx<T> *src = A + seq * n + j;
#pragma unroll
                for (int q = 0; q < R; ++q) x[u][q] = src[q * m];
                w1[u] = tab[k * tstep];
                kk[u] = k;
                dst[u] = seq * n + jb * R * Ns + k;
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u)
        {
            if (dst[u] < 0) continue;
            if (kk[u] > 0)
            {
                cpx<double> wq = w1[u];
                x[u][1] = cmul(x[u][1], cpx<T>{(T)wq.x, (T)wq.y});
#pragma unroll
                for (int q = 2; q < R; ++q)
                {
                    wq = cmul(wq, w1[u]);
                    x[u][q] = cmul(x[u][q], cpx<T>{(T)wq.x, (T)wq.y});
                }
            }
            dft_small<T, R>(x[u], w);
            cpx<T> *dst_p = B + dst[u];
#pragma unroll
            for (int q = 0; q < R; ++q) dst_p[q * Ns] = x[u][q];
        }
    }
    __syncthreads();
}

// one in-place decimation-in-time stage of radix R (2, 3, 4, 5, 7, 8) over nseq
// sequences: butterfly (c, j) reads and writes the R positions
// c R Ls + j + q Ls, so the butterflies of a stage touch disjoint points
template <typename T, int R, int NTMAX, bool POST = false>
__device__ void gen_stage_ip(cpx<T> *A, int nseq, int n, int Ls, const cpx<double> *__restrict__ tab,
                             const cpx<double> *__restrict__ chirp = nullptr, int len = 0)
{
    const int t = threadIdx.x, m = n / R, nb = nseq * m, tstep = n / (R * Ls);
    const float inv_m = 1.0f / (float)m, inv_ls = 1.0f / (float)Ls;
    cpx<T> w[R];
#pragma unroll
    for (int e = 1; e < R; ++e) w[e] = tabv<T>(tab, e * m);
    w[0] = cpx<T>{T(1), T(0)};
    const int NTH = gen_nth<NTMAX>();
    for (int b0 = t; b0 < nb; b0 += 2 * NTH)
    {
        cpx<T> x[2][R];
        cpx<double> w1[2];
        int base[2] = {-1, -1}, jj[2] = {0, 0};
#pragma unroll
        for (int u = 0; u < 2; ++u)
        {
            const int b = b0 + u * NTH;
            if (b < nb)
            {
                const int seq = gdiv(b, m, inv_m), r = b - seq * m;
                const int c = gdiv(r, Ls, inv_ls), j = r - c * Ls;
                base[u] = seq * n + c * R * Ls + j;
#pragma unroll
                for (int q = 0; q < R; ++q) x[u][q] = A[base[u] + q * Ls];
                w1[u] = tab[j * tstep];
                jj[u] = j;
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u)
        {
            if (base[u] < 0) continue;
            if (jj[u] > 0)
            {
                cpx<double> wq = w1[u];
                x[u][1] = cmul(x[u][1], cpx<T>{(T)wq.x, (T)wq.y});
#pragma unroll
                for (int q = 2; q < R; ++q)
                {
                    wq = cmul(wq, w1[u]);
                    x[u][q] = cmul(x[u][q], cpx<T>{(T)wq.x, (T)wq.y});
                }
            }
            dft_small<T, R>(x[u], w);
            cpx<T> *p = A + base[u];
            if constexpr (POST)
            {
                // the last stage (R Ls = n, c = 0): output e = jj + q Ls, conjugated
                // and times the chirp below len (the rest is never read)
#pragma unroll
                for (int q = 0; q < R; ++q)
                {
                    const int e = jj[u] + q * Ls;
                    p[q * Ls] = e < len ? cmul(conjc(x[u][q]), tabv<T>(chirp, e)) : x[u][q];
                }
            }
            else
            {
#pragma unroll
                for (int q = 0; q < R; ++q) p[q * Ls] = x[u][q];
            }
        }
    }
    __syncthreads();
}

// Bluestein's middle pass, in place: the last decimation-in-frequency stage
// (radix R on blocks of R points, no twiddles), the product with the kernel
// spectrum in the stages' position order, conjugated, and the first
// decimation-in-time stage (radix R on the same blocks, no twiddles), in
// registers: two LDS passes and two barriers fewer than running them apart
template <typename T, int R, int NTMAX>
__device__ void gen_stage_blue_mid(cpx<T> *A, int nseq, int n, const cpx<double> *__restrict__ tab,
                                   const cpx<double> *__restrict__ bhp)
{
    const int t = threadIdx.x, per = n / R, nb = nseq * per;
    const float inv_per = 1.0f / (float)per;
    cpx<T> w[R];
#pragma unroll
    for (int e = 1; e < R; ++e) w[e] = tabv<T>(tab, e * per);
    w[0] = cpx<T>{T(1), T(0)};
    for (int b = t; b < nb; b += gen_nth<NTMAX>())
    {
        const int seq = gdiv(b, per, inv_per), c = b - seq * per;
        cpx<T> *p = A + seq * n + c * R;
        cpx<T> x[R];
#pragma unroll
        for (int q = 0; q < R; ++q) x[q] = p[q];
        dft_small<T, R>(x, w);
#pragma unroll
        for (int q = 0; q < R; ++q) x[q] = conjc(cmul(x[q], tabv<T>(bhp, c * R + q)));
        dft_small<T, R>(x, w);
#pragma unroll
        for (int q = 0; q < R; ++q) p[q] = x[q];
    }
    __syncthreads();
}

// one stage of a prime radix r > 7: twiddle A in place, then every output
// reads the r inputs of its butterfly from A and goes to B
template <typename T, int NTMAX>
__device__ void gen_stage_any(cpx<T> *A, cpx<T> *B, int nseq, int n, int r, int Ns, const cpx<double> *__restrict__ tab)
{
    const int t = threadIdx.x, m = n / r, total = nseq * n, tstep = n / (r * Ns);
    const float inv_n = 1.0f / (float)n, inv_m = 1.0f / (float)m, inv_ns = 1.0f / (float)Ns;
    for (int i = t; i < total; i += gen_nth<NTMAX>())
    {
        const int seq = gdiv(i, n, inv_n), idx = i - seq * n;
        const int q = gdiv(idx, m, inv_m), j = idx - q * m;
        const int k = j - gdiv(j, Ns, inv_ns) * Ns;
        if (q > 0 && k > 0) A[i] = cmul(A[i], tabv<T>(tab, q * k * tstep));
    }
    __syncthreads();
    for (int o = t; o < total; o += gen_nth<NTMAX>())
    {
        const int seq = gdiv(o, n, inv_n), oo = o - seq * n;
        const int qo = gdiv(oo, m, inv_m), j = oo - qo * m;
        const int jb = gdiv(j, Ns, inv_ns), k = j - jb * Ns;
        const cpx<T> *src = A + seq * n + j;
        cpx<T> acc = src[0];
        int e = 0;
        for (int q = 1; q < r; ++q)
        {
            e += qo;
            if (e >= r) e -= r;
            acc = acc + cmul(src[q * m], tabv<T>(tab, e * m));
        }
        B[seq * n + jb * r * Ns + k + qo * Ns] = acc;
    }
    __syncthreads();
}

// all stages of the plan on nseq sequences in A (B: the second buffer);
// returns the buffer holding the result
template <typename T, int NTMAX>
__device__ cpx<T> *gen_stages(cpx<T> *A, cpx<T> *B, int nseq, const GenPlan &g, const cpx<double> *__restrict__ tab)
{
    int Ns = 1;
    for (int s = 0; s < g.nstages; ++s)
    {
        const int r = g.radix[s];
        switch (r)
        {
        case 2: gen_stage_reg<T, 2, NTMAX>(A, B, nseq, g.n, Ns, tab); break;
        case 3: gen_stage_reg<T, 3, NTMAX>(A, B, nseq, g.n, Ns, tab); break;
        case 4: gen_stage_reg<T, 4, NTMAX>(A, B, nseq, g.n, Ns, tab); break;
        case 5: gen_stage_reg<T, 5, NTMAX>(A, B, nseq, g.n, Ns, tab); break;
        case 7: gen_stage_reg<T, 7, NTMAX>(A, B, nseq, g.n, Ns, tab); break;
        case 8: gen_stage_reg<T, 8, NTMAX>(A, B, nseq, g.n, Ns, tab); break;
        default: gen_stage_any<T, NTMAX>(A, B, nseq, g.n, r, Ns, tab); break;
        }
        cpx<T> *tmp = A;
        A = B;
        B = tmp;
        Ns *= r;
    }
    return A;
}

// one in-place decimation-in-frequency stage of radix R on blocks of Lb
// points: butterfly (c, j < Lb/R) reads the R positions c Lb + j + q Lb/R,
// transforms them and multiplies output p by w_{Lb}^{j p}
template <typename T, int R, int NTMAX, bool PRE = false>
__device__ void gen_stage_dif(cpx<T> *A, int nseq, int n, int Lb, const cpx<double> *__restrict__ tab,
                              const cpx<double> *__restrict__ chirp = nullptr, int len = 0)
{
    const int t = threadIdx.x, per = n / R, nb = nseq * per, m = Lb / R, tstep = n / Lb;
    const float inv_per = 1.0f / (float)per, inv_m = 1.0f / (float)m;
    cpx<T> w[R];
#pragma unroll
    for (int e = 1; e < R; ++e) w[e] = tabv<T>(tab, e * per);
    w[0] = cpx<T>{T(1), T(0)};
    const int NTH = gen_nth<NTMAX>();
    for (int b0 = t; b0 < nb; b0 += 2 * NTH)
    {
        cpx<T> x[2][R];
        cpx<double> w1[2];
        int base[2] = {-1, -1}, jj[2] = {0, 0};
#pragma unroll
        for (int u = 0; u < 2; ++u)
        {
            const int b = b0 + u * NTH;
            if (b < nb)
            {
                const int seq = gdiv(b, per, inv_per), r = b - seq * per;
                const int c = gdiv(r, m, inv_m), j = r - c * m;
                base[u] = seq * n + c * Lb + j;
                if constexpr (PRE)
                {
                    // the first stage (Lb = n, c = 0): element e = j + q m, times the
                    // chirp, zero beyond len (those points were never loaded)
#pragma unroll
                    for (int q = 0; q < R; ++q)
                    {
                        const int e = j + q * m;
                        x[u][q] = e < len ? cmul(A[base[u] + q * m], tabv<T>(chirp, e)) : cpx<T>{T(0), T(0)};
                    }
                }
                else
                {
#pragma unroll
                    for (int q = 0; q < R; ++q) x[u][q] = A[base[u] + q * m];
                }
                w1[u] = tab[j * tstep];
                jj[u] = j;
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u)
        {
            if (base[u] < 0) continue;
            dft_small<T, R>(x[u], w);
            if (jj[u] > 0)
            {
                cpx<double> wq = w1[u];
                x[u][1] = cmul(x[u][1], cpx<T>{(T)wq.x, (T)wq.y});
#pragma unroll
                for (int q = 2; q < R; ++q)
                {
                    wq = cmul(wq, w1[u]);
                    x[u][q] = cmul(x[u][q], cpx<T>{(T)wq.x, (T)wq.y});
                }
            }
            cpx<T> *p = A + base[u];
#pragma unroll
            for (int q = 0; q < R; ++q) p[q * m] = x[u][q];
        }
    }
    __syncthreads();
}

// all decimation-in-frequency stages (r_S first) of a plan whose radices are
// at most 8: natural order in A to X[k] at pos(k)
template <typename T, int NTMAX>
__device__ void gen_stages_dif(cpx<T> *A, int nseq, const GenPlan &g, const cpx<double> *__restrict__ tab)
{
    int Lb = g.n;
    for (int s = g.nstages - 1; s >= 0; --s)
    {
        const int r = g.radix[s];
        switch (r)
        {
        case 2: gen_stage_dif<T, 2, NTMAX>(A, nseq, g.n, Lb, tab); break;
        case 3: gen_stage_dif<T, 3, NTMAX>(A, nseq, g.n, Lb, tab); break;
        case 4: gen_stage_dif<T, 4, NTMAX>(A, nseq, g.n, Lb, tab); break;
        case 5: gen_stage_dif<T, 5, NTMAX>(A, nseq, g.n, Lb, tab); break;
        case 8: gen_stage_dif<T, 8, NTMAX>(A, nseq, g.n, Lb, tab); break;
        default: gen_stage_dif<T, 7, NTMAX>(A, nseq, g.n, Lb, tab); break;   // the host admits only 2..8 here
        }
        Lb /= r;
    }
}

// all stages of an in-place plan (radices <= 8) on the digit-reversed
// sequences in A; the result is in A in natural order
template <typename T, int NTMAX>
__device__ void gen_stages_ip(cpx<T> *A, int nseq, const GenPlan &g, const cpx<double> *__restrict__ tab,
                              int first = 0)
{
    int Ls = 1;
    for (int s = 0; s < first; ++s) Ls *= g.radix[s];
    for (int s = first; s < g.nstages; ++s)
    {
        const int r = g.radix[s];
        switch (r)
        {
        case 2: gen_stage_ip<T, 2, NTMAX>(A, nseq, g.n, Ls, tab); break;
        case 3: gen_stage_ip<T, 3, NTMAX>(A, nseq, g.n, Ls, tab); break;
        case 4: gen_stage_ip<T, 4, NTMAX>(A, nseq, g.n, Ls, tab); break;
        case 5: gen_stage_ip<T, 5, NTMAX>(A, nseq, g.n, Ls, tab); break;
        case 8: gen_stage_ip<T, 8, NTMAX>(A, nseq, g.n, Ls, tab); break;
        default: gen_stage_ip<T, 7, NTMAX>(A, nseq, g.n, Ls, tab); break;   // the host admits only 2..8 here
        }
        Ls *= r;
    }
}

// In-place Bluestein convolution of a power-of-two plan (radices 2, 4, 8),
// S >= 2 stages: the chirp product fused into the first DIF stage, the
// kernel-spectrum product between the last DIF and the first DIT stage
// (gen_stage_blue_mid), the conjugation and output chirp into the last DIT
// stage; natural order in and out
template <typename T, int NTMAX>
__device__ void gen_blue_ip(cpx<T> *A, int nseq, const GenPlan &g, const cpx<double> *__restrict__ tab)
{
    const cpx<double> *chirp = tab + g.off_c, *bhp = tab + g.off_bp;
    const int S = g.nstages, n = g.n, len = g.len;
    int Lb = n;
    for (int s = S - 1; s >= 1; --s)
    {
        const int r = g.radix[s];
        if (s == S - 1)
        {
            switch (r)
            {
            case 2: gen_stage_dif<T, 2, NTMAX, true>(A, nseq, n, Lb, tab, chirp, len); break;
            case 4: gen_stage_dif<T, 4, NTMAX, true>(A, nseq, n, Lb, tab, chirp, len); break;
            default: gen_stage_dif<T, 8, NTMAX, true>(A, nseq, n, Lb, tab, chirp, len); break;
            }
        }
        else
        {
            switch (r)
            {
            case 2: gen_stage_dif<T, 2, NTMAX>(A, nseq, n, Lb, tab); break;
            case 4: gen_stage_dif<T, 4, NTMAX>(A, nseq, n, Lb, tab); break;
            default: gen_stage_dif<T, 8, NTMAX>(A, nseq, n, Lb, tab); break;
            }
        }
        Lb /= r;
    }
    switch (g.radix[0])
    {
    case 2: gen_stage_blue_mid<T, 2, NTMAX>(A, nseq, n, tab, bhp); break;
    case 4: gen_stage_blue_mid<T, 4, NTMAX>(A, nseq, n, tab, bhp); break;
    default: gen_stage_blue_mid<T, 8, NTMAX>(A, nseq, n, tab, bhp); break;
    }
    int Ls = g.radix[0];
    for (int s = 1; s < S; ++s)
    {
        const int r = g.radix[s];
        if (s == S - 1)
        {
            switch (r)
            {
            case 2: gen_stage_ip<T, 2, NTMAX, true>(A, nseq, n, Ls, tab, chirp, len); break;
            case 4: gen_stage_ip<T, 4, NTMAX, true>(A, nseq, n, Ls, tab, chirp, len); break;
            default: gen_stage_ip<T, 8, NTMAX, true>(A, nseq, n, Ls, tab, chirp, len); break;
            }
        }
        else
        {
            switch (r)
            {
            case 2: gen_stage_ip<T, 2, NTMAX>(A, nseq, n, Ls, tab); break;
            case 4: gen_stage_ip<T, 4, NTMAX>(A, nseq, n, Ls, tab); break;
            default: gen_stage_ip<T, 8, NTMAX>(A, nseq, n, Ls, tab); break;
            }
        }
        Ls *= r;
    }
}

// The DFT of length g.len of nseq sequences held at A[s*g.n + e], e < g.len.
// Direct plans (g.len == g.n) run the Stockham stages.  Otherwise Bluestein's
// algorithm: with the chirp c[j] = w_{2 len}^{j^2} (tab + off_c),
//   X[k] = c[k] sum_j (x[j] c[j]) conj(c[k-j]),
// a circular convolution of length n >= 2 len - 1, computed as
// IDFT(DFT(a) * Bh) with Bh = DFT(b) / n (tab + off_b, b[m] = conj c[|m|]);
// the inverse DFT is conj(DFT(conj .)).  Returns the buffer with X at
// [s*g.n + k], k < g.len.  In-place plans (g.ip) take A loaded in the
// gen_pos() order and return A.
template <typename T, int NTMAX>
__device__ cpx<T> *gen_fft(cpx<T> *A, cpx<T> *B, int nseq, const GenPlan &g, const cpx<double> *__restrict__ tab,
                           int first = 0)
{
    if (g.ip && g.len == g.n)
    {
        gen_stages_ip<T, NTMAX>(A, nseq, g, tab, first);
        return A;
    }
    if (g.len == g.n) return gen_stages<T, NTMAX>(A, B, nseq, g, tab);
    if (g.ip && g.nstages >= 2)
    {
        gen_blue_ip<T, NTMAX>(A, nseq, g, tab);
        return A;
    }
    const int t = threadIdx.x, n = g.n, len = g.len, total = nseq * n;
    const cpx<double> *chirp = tab + g.off_c, *bh = tab + g.off_b;
    const float inv_n = 1.0f / (float)n;
    for (int i = t; i < total; i += gen_nth<NTMAX>())
    {
        const int e = i - gdiv(i, n, inv_n) * n;
        A[i] = e < len ? cmul(A[i], tabv<T>(chirp, e)) : cpx<T>{T(0), T(0)};
    }
    __syncthreads();
    if (g.ip)
    {
        // in place: DIF to the pos() order, the kernel spectrum in that order,
        // then DIT back to natural order
        gen_stages_dif<T, NTMAX>(A, nseq, g, tab);
        const cpx<double> *bhp = tab + g.off_bp;
        for (int i = t; i < total; i += gen_nth<NTMAX>())
        {
            const int p = i - gdiv(i, n, inv_n) * n;
            A[i] = conjc(cmul(A[i], tabv<T>(bhp, p)));
        }
        __syncthreads();
        gen_stages_ip<T, NTMAX>(A, nseq, g, tab);
        for (int i = t; i < total; i += gen_nth<NTMAX>())
        {
            const int e = i - gdiv(i, n, inv_n) * n;
            if (e < len) A[i] = cmul(conjc(A[i]), tabv<T>(chirp, e));
        }
        __syncthreads();
        return A;
    }
    cpx<T> *F = gen_stages<T, NTMAX>(A, B, nseq, g, tab);
    for (int i = t; i < total; i += gen_nth<NTMAX>())
    {
        const int e = i - gdiv(i, n, inv_n) * n;
        F[i] = conjc(cmul(F[i], tabv<T>(bh, e)));
    }
    __syncthreads();
    cpx<T> *G = gen_stages<T, NTMAX>(F, F == A ? B : A, nseq, g, tab);
    for (int i = t; i < total; i += gen_nth<NTMAX>())
    {
        const int e = i - gdiv(i, n, inv_n) * n;
        if (e < len) G[i] = cmul(conjc(G[i]), tabv<T>(chirp, e));
    }
    __syncthreads();
    return G;
}

__device__ __forceinline__ float absmax(cpx<float> z) { return fmaxf(fabsf(z.x), fabsf(z.y)); }

// Direct in-place plans: the loads fused with the first decimation-in-time
// stage.  The elements whose first digit d_1 is 0 are e0 < n/r_1, and
// e0 + q n/r_1 (q < r_1) sit at positions pos(e0) + q: one block of the
// first stage (radix r_1, no twiddles).  A thread loads the r_1 elements,
// transforms them and writes the block, one LDS pass and barrier fewer than
// scattering the loads and running the stage; gen_fft then starts at stage 1.
// ld(s, e) is element e of sequence s (zero outside the data); s_fast: the
// sequence index runs fastest across threads (adjacent columns), else the
// element index (a row's contiguous points).
template <typename T, int R, int NTMAX, typename LD>
__device__ void gen_load_first_r(cpx<T> *buf, int nseq, const GenPlan &g, const cpx<double> *__restrict__ tab,
                                 bool s_fast, LD ld)
{
    const int n = g.n, per = n / R, total = nseq * per;
    cpx<T> w[R];
#pragma unroll
    for (int e = 1; e < R; ++e) w[e] = tabv<T>(tab, e * per);
    w[0] = cpx<T>{T(1), T(0)};
    const float inv_per = 1.0f / (float)per, inv_ns = 1.0f / (float)nseq;
    for (int i = threadIdx.x; i < total; i += gen_nth<NTMAX>())
    {
        int s, e0;
        if (s_fast)
        {
            e0 = gdiv(i, nseq, inv_ns);
            s = i - e0 * nseq;
        }
        else
        {
            s = gdiv(i, per, inv_per);
            e0 = i - s * per;
        }
        cpx<T> x[R];
#pragma unroll
        for (int q = 0; q < R; ++q) x[q] = ld(s, e0 + q * per);
        dft_small<T, R>(x, w);
        cpx<T> *o = buf + s * n + gen_pos(g, tab, e0);
#pragma unroll
        for (int q = 0; q < R; ++q) o[q] = x[q];
    }
}

// whether a plan takes the fused loads (direct, in place)
__device__ __forceinline__ bool gen_fuses_first(const GenPlan &g)
{
    return g.ip && g.len == g.n && g.off_r >= 0 && g.nstages >= 1;
}

template <typename T, int NTMAX, typename LD>
__device__ void gen_load_first(cpx<T> *buf, int nseq, const GenPlan &g, const cpx<double> *__restrict__ tab,
                               bool s_fast, LD ld)
{
    switch (g.radix[0])
    {
    case 2: gen_load_first_r<T, 2, NTMAX>(buf, nseq, g, tab, s_fast, ld); break;
    case 3: gen_load_first_r<T, 3, NTMAX>(buf, nseq, g, tab, s_fast, ld); break;
    case 4: gen_load_first_r<T, 4, NTMAX>(buf, nseq, g, tab, s_fast, ld); break;
    case 5: gen_load_first_r<T, 5, NTMAX>(buf, nseq, g, tab, s_fast, ld); break;
    case 8: gen_load_first_r<T, 8, NTMAX>(buf, nseq, g, tab, s_fast, ld); break;
    default: gen_load_first_r<T, 7, NTMAX>(buf, nseq, g, tab, s_fast, ld); break;   // in-place radices are 2..8
    }
}

template <typename T> __device__ __forceinline__ cpx<T> *gen_lds()
{
    extern __shared__ __align__(16) unsigned char gen_smem[];
    return reinterpret_cast<cpx<T> *>(gen_smem);
}

// Row r2c of any even Ny: nseq rows per workgroup, L = Ny/2 complex points
// z[n] = x[2n] + i x[2n+1] each, then the r2c split
//   X[k] = Fe + w_{Ny}^k Fo, X[L-k] = conj(Fe - w_{Ny}^k Fo),
//   Fe = (Z[k] + conj Z[L-k]) / 2, Fo = (Z[k] - conj Z[L-k]) / (2i).
// tabL: the row plan's table; tab2: w_{Ny}^m (m < Ny).  blockIdx.x = row group
// (XCD order), .y = grid.  Dynamic LDS: gen_nbuf(g) * nseq * g.n complex, then nseq
// words.  16-bit storage S (T = float): the input row is unscaled by
// 2^-in_exp[row] when given, and the output row is stored times 2^e with
// e = pick_exp(its largest component), written to out_exp[row], as the
// power-of-two row kernels do.
template <typename T, typename S = T, int NTMAX = GEN_NT>
__global__ void __launch_bounds__(NTMAX) gen_row_r2c_kernel(RowArgs a, GenPlan g, const cpx<double> *tabL,
                                                            const cpx<double> *tab2, int nseq)
{
    constexpr bool S16 = is16<S>();
    cpx<T> *buf = gen_lds<T>();
    const int t = threadIdx.x, L = g.len, ST = g.n;
    unsigned *smax = reinterpret_cast<unsigned *>(buf + gen_nbuf(g) * nseq * ST);
    const long long b = blockIdx.y;
    const int row0 = xcd_order((int)blockIdx.x, (int)gridDim.x) * nseq;
    const int nrow = min(nseq, a.nrows - row0);
    const S *in = reinterpret_cast<const S *>(a.in) + b * a.dist_in + (long long)row0 * a.ld_in;
    S *out = reinterpret_cast<S *>(a.out) + b * a.dist_out + (long long)row0 * a.ld_out;
    if constexpr (S16)
    {
        for (int i = t; i < nseq; i += gen_nth<NTMAX>()) smax[i] = 0u;   // ordered before use by the barriers below
    }
    const float inv_l = 1.0f / (float)L;
    auto load = [&](int s, int e) {
        cpx<T> z = {T(0), T(0)};
        if (s < nrow)
        {
            z = stor<S>::ld(in + s * a.ld_in + 2 * e);
            if constexpr (S16)
            {
                if (a.in_exp) z = scal(z, (T)ldexpf(1.0f, -a.in_exp[b * a.exp_dist + row0 + s]));
            }
        }
        return z;
    };
    const bool fused = gen_fuses_first(g);
    if (fused) gen_load_first<T, NTMAX>(buf, nseq, g, tabL, false, load);
    else
    {
        for (int i = t; i < nseq * L; i += gen_nth<NTMAX>())
        {
            const int s = gdiv(i, L, inv_l), e = i - s * L;
            buf[s * ST + gen_pos(g, tabL, e)] = load(s, e);
        }
    }
    __syncthreads();
    const cpx<T> *Zs = gen_fft<T, NTMAX>(buf, buf + nseq * ST, nseq, g, tabL, fused ? 1 : 0);
    const int half = L / 2 + 1;   // k = 0 .. floor(L/2)
    const float inv_h = 1.0f / (float)half;
    // X[k] and X[L-k] of row s (X[0] and X[L] for k = 0; X[L/2] alone for 2k = L)
    auto split = [&](int s, int k, cpx<T> &x0, cpx<T> &x1) {
        const cpx<T> *Z = Zs + s * ST;
        if (k == 0)
        {
            const cpx<T> z0 = Z[0];
            x0 = cpx<T>{z0.x + z0.y, T(0)};
            x1 = cpx<T>{z0.x - z0.y, T(0)};
        }
        else if (2 * k == L)
        {
            x0 = conjc(Z[k]);
            x1 = x0;
        }
        else
        {
            const cpx<T> A = Z[k], Bc = conjc(Z[L - k]);
            const cpx<T> Fe = scal(A + Bc, T(0.5));
            const cpx<T> D = scal(A - Bc, T(0.5));
            const cpx<T> Fo = {D.y, -D.x};
            const cpx<T> tt = cmul(tabv<T>(tab2, k), Fo);
            x0 = Fe + tt;
            x1 = conjc(Fe - tt);
        }
    };
    if constexpr (S16)
    {
        for (int i = t; i < nrow * half; i += gen_nth<NTMAX>())
        {
            const int s = gdiv(i, half, inv_h), k = i - s * half;
            cpx<T> x0, x1;
            split(s, k, x0, x1);
            atomicMax(&smax[s], __float_as_uint(fmaxf(absmax(x0), absmax(x1))));
        }
        __syncthreads();
    }
    for (int i = t; i < nrow * half; i += gen_nth<NTMAX>())
    {
        const int s = gdiv(i, half, inv_h), k = i - s * half;
        cpx<T> x0, x1;
        split(s, k, x0, x1);
        S *o = out + s * a.ld_out;
        if constexpr (S16)
        {
            const int ex = pick_exp(__uint_as_float(smax[s]));
            const T sc = (T)ldexpf(1.0f, ex);
            x0 = scal(x0, sc);
            x1 = scal(x1, sc);
            if (k == 0 && a.out_exp) a.out_exp[b * a.exp_dist + row0 + s] = ex;
        }
        if (k == 0)
        {
            stor<S>::st(o, x0);
            stor<S>::st(o + 2 * L, x1);
        }
        else
        {
            stor<S>::st(o + 2 * k, x0);
            if (2 * k != L) stor<S>::st(o + 2 * (L - k), x1);
        }
    }
}

// Column c2c of any Nx: a workgroup takes W adjacent columns over all Nx rows
// (W * g.n <= gen_max_points), from a.in to a.out (the same spectrum for the
// forward transform, in place).  CONJ: the inverse transform, computed as
// conj(DFT(conj .)).  blockIdx.x = strip (XCD order: neighbouring strips,
// which share 128-byte lines, run behind one L2), .y = grid.
// tab: the column plan's table.  Dynamic LDS: gen_nbuf(g) * W * g.n complex, then W
// words.  16-bit storage S (T = float, forward only): rows are unscaled by
// 2^-row_exp[row] on load, and column c is stored times 2^e with e =
// pick_exp(its largest component), written to out_exp[c].
template <typename T, typename S, bool CONJ, int NTMAX = GEN_NT>
__global__ void __launch_bounds__(NTMAX) gen_col_kernel(ColArgs a, GenPlan g, const cpx<double> *tab, int W)
{
    constexpr bool S16 = is16<S>();
    cpx<T> *buf = gen_lds<T>();
    const int t = threadIdx.x, n = g.len, ST = g.n;
    unsigned *smax = reinterpret_cast<unsigned *>(buf + gen_nbuf(g) * W * ST);
    const long long b = blockIdx.y;
    const int col0 = xcd_order((int)blockIdx.x, (int)gridDim.x) * W;
    const int nvalid = min(W, a.ncols - col0);
    const S *src = reinterpret_cast<const S *>(a.in) + 2 * b * a.dist;
    S *dst = reinterpret_cast<S *>(a.out) + 2 * b * a.dist;
    if constexpr (S16)
    {
        if (t < W) smax[t] = 0u;   // ordered before use by the barriers below
    }
    const float inv_w = 1.0f / (float)W;
    // row segments: element (e, w) -> buf[w * ST + pos(e)]
    auto load = [&](int w, int e) {
        cpx<T> z = {T(0), T(0)};
        if (w < nvalid)
        {
            z = stor<S>::ld(src + 2 * ((long long)e * a.pitch + col0 + w));
            if constexpr (S16)
            {
                if (a.row_exp) z = scal(z, (T)ldexpf(1.0f, -a.row_exp[b * a.rexp_dist + e]));
            }
        }
        return CONJ ? conjc(z) : z;
    };
    const bool fused = gen_fuses_first(g);
    if (fused) gen_load_first<T, NTMAX>(buf, W, g, tab, true, load);
    else
    {
        for (int i = t; i < n * W; i += gen_nth<NTMAX>())
        {
            const int e = gdiv(i, W, inv_w), w = i - e * W;
            buf[w * ST + gen_pos(g, tab, e)] = load(w, e);
        }
    }
    __syncthreads();
    const cpx<T> *X = gen_fft<T, NTMAX>(buf, buf + ST * W, W, g, tab, fused ? 1 : 0);
    if constexpr (S16)
    {
        for (int i = t; i < n * W; i += gen_nth<NTMAX>())
        {
            const int e = gdiv(i, W, inv_w), w = i - e * W;
            if (w < nvalid) atomicMax(&smax[w], __float_as_uint(absmax(X[w * ST + e])));
        }
        __syncthreads();
    }
    for (int i = t; i < n * W; i += gen_nth<NTMAX>())
    {
        const int e = gdiv(i, W, inv_w), w = i - e * W;
        if (w >= nvalid) continue;
        cpx<T> z = X[w * ST + e];
        if constexpr (S16)
        {
            const int ex = pick_exp(__uint_as_float(smax[w]));
            z = scal(z, (T)ldexpf(1.0f, ex));
            if (e == 0 && a.out_exp) a.out_exp[b * a.cexp_dist + col0 + w] = ex;
        }
        stor<S>::st(dst + 2 * ((long long)e * a.pitch + col0 + w), CONJ ? conjc(z) : z);
    }
}

// Row c2r of any even Ny (the row half of the adjoint when HALVE, the
// unnormalised inverse c2r otherwise): half-spectrum row G[0..L] in, Ny reals
// out, nseq rows per workgroup (dynamic LDS: gen_nbuf(g) * nseq * g.n complex),
//   H = G on 0<k<L (halved when HALVE), H[0] = Re G[0], H[L] = Re G[L],
//   Z[k] = (H[k] + conj H[L-k]) + i (H[k] - conj H[L-k]) conj(w_{Ny}^k),
//   out[2n] + i out[2n+1] = IDFT_L(Z) = conj(DFT(conj Z)).
template <typename T, bool HALVE, int NTMAX = GEN_NT>
__global__ void __launch_bounds__(NTMAX) gen_row_c2r_kernel(RowArgs a, GenPlan g, const cpx<double> *tabL,
                                                            const cpx<double> *tab2, int nseq)
{
    cpx<T> *buf = gen_lds<T>();
    const int t = threadIdx.x, L = g.len, ST = g.n;
    const long long b = blockIdx.y;
    const int row0 = xcd_order((int)blockIdx.x, (int)gridDim.x) * nseq;
    const int nrow = min(nseq, a.nrows - row0);
    const T *in = reinterpret_cast<const T *>(a.in) + b * a.dist_in + (long long)row0 * a.ld_in;
    T *out = reinterpret_cast<T *>(a.out) + b * a.dist_out + (long long)row0 * a.ld_out;
    const T hs = HALVE ? T(0.5) : T(1);
    const float inv_l = 1.0f / (float)L;
    for (int i = t; i < nseq * L; i += gen_nth<NTMAX>())
    {
        const int s = gdiv(i, L, inv_l), k = i - s * L;
        cpx<T> z = {T(0), T(0)};
        if (s < nrow)
        {
            const T *g_in = in + s * a.ld_in;
            if (k == 0)
            {
                const T h0 = stor<T>::ld(g_in).x, hl = stor<T>::ld(g_in + 2 * L).x;
                z = cpx<T>{h0 + hl, h0 - hl};
            }
            else
            {
                const cpx<T> hk = scal(stor<T>::ld(g_in + 2 * k), hs), hl = scal(stor<T>::ld(g_in + 2 * (L - k)), hs);
                const cpx<T> w = tabv<T>(tab2, k);
                const cpx<T> wc = {w.x, -w.y};
                const cpx<T> s0 = hk + conjc(hl), d0 = cmul(hk - conjc(hl), wc);
                z = cpx<T>{s0.x - d0.y, s0.y + d0.x};
            }
        }
        buf[s * ST + gen_pos(g, tabL, k)] = conjc(z);
    }
    __syncthreads();
    const cpx<T> *U = gen_fft<T, NTMAX>(buf, buf + nseq * ST, nseq, g, tabL);
    for (int i = t; i < nrow * L; i += gen_nth<NTMAX>())
    {
        const int s = gdiv(i, L, inv_l), e = i - s * L;
        const cpx<T> u = U[s * ST + e];
        stor<T>::st(out + s * a.ld_out + 2 * e, cpx<T>{u.x, -u.y});
    }
}

}   // namespace hpxfft_dev

namespace hpxfft_dev {

// ---------------------------------------------------------------------------
// Two-level general transforms, for lengths N = P * Q beyond one workgroup:
// phase A takes the Q-point DFTs over elements m1 + P*e (one sequence per
// m1), times w_N^{m1 k2}; phase B the P-point DFTs over m1, to element
// Q*k1 + k2.  gen_strided_kernel runs either phase on W adjacent sequences
// ("columns") per workgroup, addressed by strides, so the same kernel serves
// columns of the spectrum and the complex points of a row.
//
// Lengths that no split into two one-workgroup lengths covers (a prime
// factor too large for in-LDS Bluestein) run Bluestein's algorithm with the
// convolution length M (a power of two >= 2N - 1) as two levels M = P * Q,
// in three passes over a scratch of M points per sequence:
//   1. a[m] = x[m] c[m] (0 for m >= N), m = m1 + P m2: Q-point DFTs over m2,
//      times w_M^{m1 k2}                                   (gen_strided_kernel)
//   2. for each k2: P-point DFT over m1, times Bh[k2 + Q k1] = DFT_M(b)/M,
//      inverse P-point DFT over k1, times w_M^{-n1 k2}       (gen_blue_mid_kernel)
//   3. for each n1: inverse Q-point DFT over k2 -> y[n1 + P n2];
//      X[n] = c[n] y[n] for n < N                           (gen_strided_kernel)
// with the chirp c[m] = w_{2N}^{m^2 mod 2N} and b[m] = conj c[|m|] (|m| < N,
// indices mod M).  Twiddle tables of more than 2^20 entries are split:
// w_n^m = lo[m mod 2^s] hi[m >> s] (gen_tw), formed in FP64 and rounded once.
struct GenStride
{
    const void *in;
    void *out;
    long long in_seq, in_col, in_sub, in_bat;      // complex-element strides: sequence point, column, blockIdx.y, .z
    long long out_seq, out_col, out_sub, out_bat;
    int ncols;      // valid columns
    int tw_n;       // > 0: output k of column c times w_{tw_n}^{(f k) mod tw_n}, f = c (tw_col) or blockIdx.y
    int tw_col;
    int conj_in, conj_out;
    int pre;        // 1: the input is a half-spectrum row G[0..L] turned into the c2r sequence z (2: interior halved)
    int prel;       // pre: L
    int tw_shift;   // split of the tw_n table (0: one table of tw_n entries)
    int tw2_shift;  // split of the w_Ny table of the c2r pre-processing and the r2c split
    // Bluestein passes 1 (blue_in) and 3 (blue_out): element e of sequence f
    // is index f + blue_p e of the length-blue_n sequence; chirp = w_{2 blue_n}
    // table split at chirp_shift; blue_conj: conjugate the pass-3 result
    int blue_in, blue_out, blue_n, blue_p, blue_conj, chirp_shift;
    const void *chirp;
};

// w_n^m from a table split at s (lo[i] = w_n^i, i < 2^s, then hi[j] = w_n^{j 2^s}),
// or from one table when s = 0; FP64 product, rounded once for FP32
template <typename T>
__device__ __forceinline__ cpx<T> gen_tw(const cpx<double> *__restrict__ t, int s, long long m)
{
    if (s == 0) return tabv<T>(t, (int)m);
    const cpx<double> w = cmul(t[(1LL << s) + (m >> s)], t[m & ((1LL << s) - 1)]);
    return cpx<T>{(T)w.x, (T)w.y};
}

// Bluestein chirp c[m] = w_{2N}^{m^2 mod 2N}
template <typename T>
__device__ __forceinline__ cpx<T> gen_chirp(const GenStride &a, long long m)
{
    const long long n2 = 2LL * a.blue_n;
    return gen_tw<T>(reinterpret_cast<const cpx<double> *>(a.chirp), a.chirp_shift, (m * m) % n2);
}

// element (e, w) of a workgroup's W x n block, for i < n W: sequence-fast
// order when the sequence stride is 1 (contiguous runs), else column-fast
__device__ __forceinline__ void gen_ew(int i, int n, int W, bool seq_fast, float inv_n, float inv_w, int &e, int &w)
{
    if (seq_fast)
    {
        w = gdiv(i, n, inv_n);
        e = i - w * n;
    }
    else
    {
        e = gdiv(i, W, inv_w);
        w = i - e * W;
    }
}

// GX: the general form (split twiddle tables, Bluestein passes 1 and 3);
// without it those paths are compiled out (the plain two-level phases)
template <typename T, bool GX = false, int NTMAX = GEN_NT>
__global__ void __launch_bounds__(NTMAX) gen_strided_kernel(GenStride a, GenPlan g, const cpx<double> *tab,
                                                            const cpx<double> *twt, const cpx<double> *tab2, int W)
{
    cpx<T> *buf = gen_lds<T>();
    const int t = threadIdx.x, n = g.len, ST = g.n;
    const int col0 = xcd_order((int)blockIdx.x, (int)gridDim.x) * W;
    const int nvalid = min(W, a.ncols - col0);
    const long long sub = blockIdx.y, bat = blockIdx.z;
    const T *src = reinterpret_cast<const T *>(a.in) + 2 * (bat * a.in_bat + sub * a.in_sub);
    T *dst = reinterpret_cast<T *>(a.out) + 2 * (bat * a.out_bat + sub * a.out_sub + (long long)col0 * a.out_col);
    const float inv_n = 1.0f / (float)n, inv_w = 1.0f / (float)W;
    const bool in_fast = a.in_seq == 1, out_fast = a.out_seq == 1;
    for (int i = t; i < n * W; i += gen_nth<NTMAX>())
    {
        int e, w;
        gen_ew(i, n, W, in_fast, inv_n, inv_w, e, w);
        cpx<T> z = {T(0), T(0)};
        // Bluestein pass 1: element e of sequence f is index m = f + P e, zero beyond N
        const bool blue_in = GX && a.blue_in;
        const long long mb = blue_in ? (long long)(a.tw_col ? col0 + w : (int)sub) + (long long)a.blue_p * e : 0;
        if (w < nvalid && (!blue_in || mb < a.blue_n))
        {
            const long long kk = (long long)e * a.in_seq + (long long)(col0 + w) * a.in_col;
            if (!a.pre) z = stor<T>::ld(src + 2 * kk);
            else
            {
                // c2r sequence from the half spectrum (as gen_row_c2r_kernel)
                const int L = a.prel, k = (int)kk;
                if (k == 0)
                {
                    const T h0 = stor<T>::ld(src).x, hl = stor<T>::ld(src + 2 * L).x;
                    z = cpx<T>{h0 + hl, h0 - hl};
                }
                else
                {
                    const T hs = a.pre == 2 ? T(0.5) : T(1);
                    const cpx<T> hk = scal(stor<T>::ld(src + 2 * k), hs), hl = scal(stor<T>::ld(src + 2 * (L - k)), hs);
                    const cpx<T> wv = gen_tw<T>(tab2, GX ? a.tw2_shift : 0, k);
                    const cpx<T> wc = {wv.x, -wv.y};
                    const cpx<T> s0 = hk + conjc(hl), d0 = cmul(hk - conjc(hl), wc);
                    z = cpx<T>{s0.x - d0.y, s0.y + d0.x};
                }
            }
            if (a.conj_in) z = conjc(z);
            if (blue_in) z = cmul(z, gen_chirp<T>(a, mb));
        }
        buf[w * ST + gen_pos(g, tab, e)] = z;
    }
    __syncthreads();
    const cpx<T> *X = gen_fft<T, NTMAX>(buf, buf + ST * W, W, g, tab);
    for (int i = t; i < n * W; i += gen_nth<NTMAX>())
    {
        int k, w;
        gen_ew(i, n, W, out_fast, inv_n, inv_w, k, w);
        if (w >= nvalid) continue;
        cpx<T> z = X[w * ST + k];
        const int f = a.tw_col ? col0 + w : (int)sub;
        if (a.tw_n > 0)
        {
            // f < P and k < Q in every use (m1 k2 < N <= 2^24), so the
            // product is already reduced and fits an int; the modulo only
            // guards other callers
            int m = f * k;
            if (m >= a.tw_n) m %= a.tw_n;
            z = cmul(z, gen_tw<T>(twt, GX ? a.tw_shift : 0, m));
        }
        if (a.conj_out) z = conjc(z);
        if (GX && a.blue_out)
        {
            // Bluestein pass 3: output k of sequence f is index n = f + P k, kept below N
            const long long nb = (long long)f + (long long)a.blue_p * k;
            if (nb >= a.blue_n) continue;
            z = cmul(z, gen_chirp<T>(a, nb));
            if (a.blue_conj) z = conjc(z);
        }
        stor<T>::st(dst + 2 * ((long long)k * a.out_seq + (long long)w * a.out_col), z);
    }
}

// Bluestein pass 2 (see GenStride): W sequences of P points (e = m1) per
// workgroup, sequence f = k2 (the column when tw_col, else blockIdx.y); the
// P-point DFT, the product with bh[k2 + Q k1], the inverse P-point DFT as
// conj(DFT(conj .)), then times w_M^{-n1 k2} (twt split at a.tw_shift),
// written back to the same positions (e = n1).  g: a ping-pong plan.
template <typename T, int NTMAX = GEN_NT>
__global__ void __launch_bounds__(NTMAX) gen_blue_mid_kernel(GenStride a, GenPlan g, const cpx<double> *tab,
                                                             const cpx<double> *bh, const cpx<double> *twt, int W,
                                                             int Q)
{
    cpx<T> *buf = gen_lds<T>();
    const int t = threadIdx.x, n = g.len, ST = g.n;
    const int col0 = xcd_order((int)blockIdx.x, (int)gridDim.x) * W;
    const int nvalid = min(W, a.ncols - col0);
    const long long sub = blockIdx.y, bat = blockIdx.z;
    const T *src = reinterpret_cast<const T *>(a.in) + 2 * (bat * a.in_bat + sub * a.in_sub);
    T *dst = reinterpret_cast<T *>(a.out) + 2 * (bat * a.out_bat + sub * a.out_sub + (long long)col0 * a.out_col);
    const float inv_n = 1.0f / (float)n, inv_w = 1.0f / (float)W;
    const bool in_fast = a.in_seq == 1, out_fast = a.out_seq == 1;
    for (int i = t; i < n * W; i += gen_nth<NTMAX>())
    {
        int e, w;
        gen_ew(i, n, W, in_fast, inv_n, inv_w, e, w);
        cpx<T> z = {T(0), T(0)};
        if (w < nvalid) z = stor<T>::ld(src + 2 * ((long long)e * a.in_seq + (long long)(col0 + w) * a.in_col));
        buf[w * ST + e] = z;
    }
    __syncthreads();
    cpx<T> *F = gen_stages<T, NTMAX>(buf, buf + ST * W, W, g, tab);
    for (int i = t; i < n * W; i += gen_nth<NTMAX>())
    {
        const int w = gdiv(i, n, inv_n), k1 = i - w * n;
        const long long k2 = a.tw_col ? col0 + w : sub;
        F[w * ST + k1] = conjc(cmul(F[w * ST + k1], tabv<T>(bh, (int)(k2 + (long long)Q * k1))));
    }
    __syncthreads();
    cpx<T> *G = gen_stages<T, NTMAX>(F, F == buf ? buf + ST * W : buf, W, g, tab);
    for (int i = t; i < n * W; i += gen_nth<NTMAX>())
    {
        int n1, w;
        gen_ew(i, n, W, out_fast, inv_n, inv_w, n1, w);
        if (w >= nvalid) continue;
        const long long k2 = a.tw_col ? col0 + w : sub;
        const cpx<T> z = cmul(conjc(G[w * ST + n1]), conjc(gen_tw<T>(twt, a.tw_shift, (long long)n1 * k2)));
        stor<T>::st(dst + 2 * ((long long)n1 * a.out_seq + (long long)w * a.out_col), z);
    }
}

// Phase B of two-level forward rows with the r2c split fused in (L = P Q,
// phase A's output in a.in as [row][m1][k2]): output k = Q k1 + k2 pairs
// with L - k = Q k1' + k2', where k2' = Q - k2 and k1' = P - 1 - k1 for
// k2 > 0, k2' = 0 and k1' = P - k1 for k2 = 0.  A workgroup takes Wh
// columns k2 <= Q/2 of one row (blockIdx.x = strip, XCD order; .y = row,
// .z = grid) and, in slots Wh.., their partners Q - k2, so it holds Z[k] and
// Z[L-k] for each of its pairs and writes X[k] and X[L-k] (the split as in
// gen_row_r2c_kernel; X[0], X[L] from Z[0]; X[L/2] = conj Z[L/2]).  A
// self-paired column (k2 = 0, or Q/2) writes the pairs with 2k <= L.
// tab: phase B's plan table; tab2: w_{Ny}^m.  Dynamic LDS: gen_nbuf(g) * 2 Wh
// * g.n complex.
template <typename T, bool GX = false, int NTMAX = GEN_NT>
__global__ void __launch_bounds__(NTMAX) gen_rows_split_kernel(GenStride a, GenPlan g, const cpx<double> *tab,
                                                               const cpx<double> *tab2, int Wh, int Q)
{
    cpx<T> *buf = gen_lds<T>();
    const int t = threadIdx.x, P = g.len, ST = g.n, W = 2 * Wh, L = P * Q, H2 = Q / 2 + 1;
    const int s0 = xcd_order((int)blockIdx.x, (int)gridDim.x) * Wh;
    const long long sub = blockIdx.y, bat = blockIdx.z;
    const T *src = reinterpret_cast<const T *>(a.in) + 2 * (bat * a.in_bat + sub * a.in_sub);
    T *dst = reinterpret_cast<T *>(a.out) + 2 * (bat * a.out_bat + sub * a.out_sub);
    const float inv_w = 1.0f / (float)W, inv_wh = 1.0f / (float)Wh;
    for (int i = t; i < P * W; i += gen_nth<NTMAX>())
    {
        const int e = gdiv(i, W, inv_w), w = i - e * W;
        const int kl = s0 + (w < Wh ? w : w - Wh);
        const int k2 = w < Wh ? kl : (kl == 0 ? 0 : Q - kl);
        cpx<T> z = {T(0), T(0)};
        if (kl < H2) z = stor<T>::ld(src + 2 * ((long long)e * a.in_seq + (long long)k2 * a.in_col));
        buf[w * ST + gen_pos(g, tab, e)] = z;
    }
    __syncthreads();
    const cpx<T> *X = gen_fft<T, NTMAX>(buf, buf + ST * W, W, g, tab);
    for (int i = t; i < P * Wh; i += gen_nth<NTMAX>())
    {
        const int k1 = gdiv(i, Wh, inv_wh), w = i - k1 * Wh;
        const int kl = s0 + w;
        if (kl >= H2) continue;
        const int k = Q * k1 + kl;
        const bool self = kl == 0 || 2 * kl == Q;
        if (self && 2 * k > L) continue;
        const cpx<T> Zk = X[w * ST + k1];
        if (k == 0)
        {
            stor<T>::st(dst, cpx<T>{Zk.x + Zk.y, T(0)});
            stor<T>::st(dst + 2 * L, cpx<T>{Zk.x - Zk.y, T(0)});
            continue;
        }
        if (2 * k == L)
        {
            stor<T>::st(dst + 2 * k, conjc(Zk));
            continue;
        }
        const int k1p = kl == 0 ? P - k1 : P - 1 - k1;
        const int wp = self ? w : Wh + w;
        const cpx<T> Bc = conjc(X[wp * ST + k1p]);
        const cpx<T> Fe = scal(Zk + Bc, T(0.5));
        const cpx<T> D = scal(Zk - Bc, T(0.5));
        const cpx<T> Fo = {D.y, -D.x};
        const cpx<T> tt = cmul(gen_tw<T>(tab2, GX ? a.tw2_shift : 0, k), Fo);
        stor<T>::st(dst + 2 * k, Fe + tt);
        stor<T>::st(dst + 2 * (L - k), conjc(Fe - tt));
    }
}

// the r2c split of a row's L-point complex transform Z, in place in the
// spectrum row (positions 0..L):
//   X[k] = Fe + w_{Ny}^k Fo, X[L-k] = conj(Fe - w_{Ny}^k Fo), X[0], X[L] from Z[0].
// blockIdx.y = row, .z = grid; one thread per pair (k, L-k), 2k <= L.
template <typename T>
__global__ void __launch_bounds__(256) gen_split_kernel(void *spec, long long row_stride, long long bat_stride, int L,
                                                         const cpx<double> *tab2, int tw2_shift)
{
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (2 * k > L) return;
    T *o = reinterpret_cast<T *>(spec) + 2 * ((long long)blockIdx.z * bat_stride + (long long)blockIdx.y * row_stride);
    if (k == 0)
    {
        const cpx<T> z0 = stor<T>::ld(o);
        stor<T>::st(o, cpx<T>{z0.x + z0.y, T(0)});
        stor<T>::st(o + 2 * L, cpx<T>{z0.x - z0.y, T(0)});
    }
    else if (2 * k == L)
    {
        stor<T>::st(o + 2 * k, conjc(stor<T>::ld(o + 2 * k)));
    }
    else
    {
        const cpx<T> A = stor<T>::ld(o + 2 * k), Bc = conjc(stor<T>::ld(o + 2 * (L - k)));
        const cpx<T> Fe = scal(A + Bc, T(0.5));
        const cpx<T> D = scal(A - Bc, T(0.5));
        const cpx<T> Fo = {D.y, -D.x};
        const cpx<T> tt = cmul(gen_tw<T>(tab2, tw2_shift, k), Fo);
        stor<T>::st(o + 2 * k, Fe + tt);
        stor<T>::st(o + 2 * (L - k), conjc(Fe - tt));
    }
}

// 16-bit storage with a length beyond one workgroup runs an FP32 plan on a
// workspace: these kernels convert into it and back.  blockIdx.y = row,
// .z = grid; one thread per complex pair.
// in: Ny/2 pairs of a 16-bit row (times 2^-in_exp[row] when given) -> FP32
template <typename S>
__global__ void __launch_bounds__(256) wide_in_kernel(const S *in, long long ld_in, long long dist_in,
                                                       const int *in_exp, long long exp_dist, float *ws,
                                                       long long ld_ws, long long dist_ws, int pairs)
{
    const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (j >= pairs) return;
    const long long b = blockIdx.z, row = blockIdx.y;
    cpx<float> z = stor<S>::ld(in + b * dist_in + row * ld_in + 2 * j);
    if (in_exp) z = scal(z, ldexpf(1.0f, -in_exp[b * exp_dist + row]));
    stor<float>::st(ws + b * dist_ws + row * ld_ws + 2 * j, z);
}

// the largest component of each spectrum column over `rows` rows per block
// (colmax cleared beforehand; float bits, so the unsigned maximum orders them)
__global__ void __launch_bounds__(256) wide_colmax_kernel(const float *ws, long long ld_ws, long long dist_ws,
                                                           int nrows, int ncols, int rows, unsigned *colmax)
{
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= ncols) return;
    const long long b = blockIdx.z;
    const int r0 = (int)blockIdx.y * rows, r1 = min(nrows, r0 + rows);
    float m = 0.0f;
    for (int r = r0; r < r1; ++r)
    {
        const cpx<float> z = stor<float>::ld(ws + b * dist_ws + (long long)r * ld_ws + 2 * c);
        m = fmaxf(m, fmaxf(fabsf(z.x), fabsf(z.y)));
    }
    atomicMax(&colmax[b * ncols + c], __float_as_uint(m));
}

// out: X * 2^e per column, e = pick_exp(column maximum), e -> col_exp
template <typename S>
__global__ void __launch_bounds__(256) wide_out_kernel(const float *ws, long long ld_ws, long long dist_ws, S *out,
                                                        long long ld_out, long long dist_out, int ncols,
                                                        const unsigned *colmax, int *col_exp, long long cexp_dist)
{
    const int c = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= ncols) return;
    const long long b = blockIdx.z, row = blockIdx.y;
    const int e = pick_exp(__uint_as_float(colmax[b * ncols + c]));
    const cpx<float> z = stor<float>::ld(ws + b * dist_ws + row * ld_ws + 2 * c);
    stor<S>::st(out + b * dist_out + row * ld_out + 2 * c, scal(z, ldexpf(1.0f, e)));
    if (row == 0 && col_exp) col_exp[b * cexp_dist + c] = e;
}

}   // namespace hpxfft_dev
