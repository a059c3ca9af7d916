// fft_general.hpp -- 2-D r2c passes for lengths that are not powers of two
// (FP64 / FP32).  The reference accepts any even Ny and any Nx through FFTW
// (core/src/util/adapter_fftw.cpp:9,28-29); these kernels cover the lengths
// whose prime factors are at most GEN_MAX_RADIX and whose data fit one
// workgroup's LDS, with the power-of-two kernels kept for power-of-two
// dimensions.
//
// One workgroup transforms NSEQ sequences of length n held in LDS, by
// Stockham autosort stages of runtime radices r (product n):
//   stage with sub-length Ns, butterfly j < n/r, k = j mod Ns:
//     x_q = A[j + q n/r] w_{r Ns}^{q k},  y = DFT_r(x),
//     B[(j / Ns) r Ns + k + q Ns] = y_q
// Stage twiddles come from the plan's exact table w_n^m (m < n); the radix-r
// DFT uses the r roots w_r^e = w_n^{e n / r} staged in LDS.  Sequences are
// stored sequence-major in LDS (seq * n + e).
#pragma once

#include "fft_kernels.hpp"

namespace hpxfft_dev {

constexpr int GEN_MAX_STAGES = 24;
constexpr int GEN_MAX_RADIX = 61;
constexpr int GEN_NT = 256;   // threads per workgroup
// LDS holds two buffers of up to GEN_LDS_BYTES / 2 each
constexpr int GEN_LDS_BYTES = 128 * 1024;
template <typename T> constexpr int gen_max_points() { return GEN_LDS_BYTES / (2 * 2 * (int)sizeof(T)); }

struct GenPlan
{
    int n;                          // transform length
    int nstages;
    int radix[GEN_MAX_STAGES];      // stage radices, product n
};

// Transform the NSEQ sequences in A (ping-pong with B); returns the buffer
// holding the result.  tab: w_n^m, m < n (forward sign).  Each stage runs as
// two LDS passes: twiddle the inputs in place, then every output reads the
// r inputs of its butterfly (no per-thread arrays of runtime size).
template <typename T>
__device__ cpx<T> *gen_fft(cpx<T> *A, cpx<T> *B, cpx<T> *roots, int nseq, const GenPlan &g,
                           const cpx<double> *__restrict__ tab)
{
    const int t = threadIdx.x, n = g.n, total = nseq * n;
    int Ns = 1;
    for (int s = 0; s < g.nstages; ++s)
    {
        const int r = g.radix[s], m = n / r;
        const long long tstep = n / (r * Ns);   // w_{r Ns}^{q k} = w_n^{q k tstep}, q k tstep < n
        // roots of this radix: w_r^e = w_n^{e n / r}
        for (int e = t; e < r; e += GEN_NT)
        {
            const cpx<double> w = tab[(long long)e * m];
            roots[e] = cpx<T>{(T)w.x, (T)w.y};
        }
        // twiddle: element j + q m of a sequence times w_{r Ns}^{q (j mod Ns)}
        for (int i = t; i < total; i += GEN_NT)
        {
            const int idx = i % n, q = idx / m, k = (idx - q * m) % Ns;
            if (q > 0 && k > 0)
            {
                const cpx<double> w = tab[(long long)q * k * tstep];
                A[i] = cmul(A[i], cpx<T>{(T)w.x, (T)w.y});
            }
        }
        __syncthreads();
        // radix-r DFTs: output qo of butterfly j -> B[(j / Ns) r Ns + (j mod Ns) + qo Ns]
        for (int i = t; i < total; i += GEN_NT)
        {
            const int seq = i / n, o = i - seq * n, qo = o / m, j = o - qo * m, k = j % Ns;
            const cpx<T> *src = A + (long long)seq * n + j;
            cpx<T> acc = src[0];
            int e = 0;
            for (int q = 1; q < r; ++q)
            {
                e += qo;
                if (e >= r) e -= r;
                acc = acc + cmul(src[(long long)q * m], roots[e]);
            }
            B[(long long)seq * n + (j / Ns) * r * Ns + k + qo * Ns] = acc;
        }
        __syncthreads();
        cpx<T> *tmp = A;
        A = B;
        B = tmp;
        Ns *= r;
    }
    return A;
}

// Row r2c of any even Ny: one row per workgroup, L = Ny/2 complex points
// z[n] = x[2n] + i x[2n+1], then the r2c split
//   X[k] = Fe + w_{Ny}^k Fo, X[L-k] = conj(Fe - w_{Ny}^k Fo),
//   Fe = (Z[k] + conj Z[L-k]) / 2, Fo = (Z[k] - conj Z[L-k]) / (2i).
// tabL: w_L^m (m < L); tab2: w_{Ny}^m (m < Ny).  blockIdx.x = row, .y = grid.
template <typename T>
__global__ void __launch_bounds__(GEN_NT) gen_row_r2c_kernel(RowArgs a, GenPlan g, const cpx<double> *tabL,
                                                            const cpx<double> *tab2)
{
    constexpr int CAP = gen_max_points<T>();
    __shared__ cpx<T> buf[2 * CAP];
    __shared__ cpx<T> roots[GEN_MAX_RADIX];
    const int t = threadIdx.x, L = g.n;
    const long long b = blockIdx.y;
    const int row = blockIdx.x;
    const T *in = reinterpret_cast<const T *>(a.in) + b * a.dist_in + (long long)row * a.ld_in;
    T *out = reinterpret_cast<T *>(a.out) + b * a.dist_out + (long long)row * a.ld_out;
    for (int e = t; e < L; e += GEN_NT) buf[e] = stor<T>::ld(in + 2 * e);
    __syncthreads();
    const cpx<T> *Z = gen_fft<T>(buf, buf + CAP, roots, 1, g, tabL);
    for (int k = t; 2 * k <= L; k += GEN_NT)
    {
        if (k == 0)
        {
            const cpx<T> z0 = Z[0];
            stor<T>::st(out, cpx<T>{z0.x + z0.y, T(0)});
            stor<T>::st(out + 2 * L, cpx<T>{z0.x - z0.y, T(0)});
        }
        else if (2 * k == L)
        {
            stor<T>::st(out + 2 * k, conjc(Z[k]));   // pairs with itself
        }
        else
        {
            const cpx<T> A = Z[k], Bc = conjc(Z[L - k]);
            const cpx<T> Fe = scal(A + Bc, T(0.5));
            const cpx<T> D = scal(A - Bc, T(0.5));
            const cpx<T> Fo = {D.y, -D.x};
            const cpx<double> w64 = tab2[k];
            const cpx<T> tt = cmul(cpx<T>{(T)w64.x, (T)w64.y}, Fo);
            stor<T>::st(out + 2 * k, Fe + tt);
            stor<T>::st(out + 2 * (L - k), conjc(Fe - tt));
        }
    }
}

// Column c2c of any Nx: a workgroup takes W adjacent columns over all Nx rows
// (W * Nx <= gen_max_points), from a.in to a.out (the same spectrum for the
// forward transform, in place).  CONJ: the inverse transform, computed as
// conj(DFT(conj .)).  blockIdx.x = strip, .y = grid.  tab: w_Nx^m (m < Nx).
template <typename T, bool CONJ>
__global__ void __launch_bounds__(GEN_NT) gen_col_kernel(ColArgs a, GenPlan g, const cpx<double> *tab, int W)
{
    constexpr int CAP = gen_max_points<T>();
    __shared__ cpx<T> buf[2 * CAP];
    __shared__ cpx<T> roots[GEN_MAX_RADIX];
    const int t = threadIdx.x, n = g.n;
    const long long b = blockIdx.y;
    const int col0 = blockIdx.x * W;
    const int nvalid = min(W, a.ncols - col0);
    const T *src = reinterpret_cast<const T *>(a.in) + 2 * b * a.dist;
    T *dst = reinterpret_cast<T *>(a.out) + 2 * b * a.dist;
    // load row segments: element (e, w) -> buf[w * n + e]
    for (int i = t; i < n * W; i += GEN_NT)
    {
        const int e = i / W, w = i - e * W;
        cpx<T> z = {T(0), T(0)};
        if (w < nvalid) z = stor<T>::ld(src + 2 * ((long long)e * a.pitch + col0 + w));
        buf[w * n + e] = CONJ ? conjc(z) : z;
    }
    __syncthreads();
    const cpx<T> *X = gen_fft<T>(buf, buf + CAP, roots, W, g, tab);
    for (int i = t; i < n * W; i += GEN_NT)
    {
        const int e = i / W, w = i - e * W;
        const cpx<T> z = X[w * n + e];
        if (w < nvalid) stor<T>::st(dst + 2 * ((long long)e * a.pitch + col0 + w), CONJ ? conjc(z) : z);
    }
}

// Row c2r of any even Ny (the row half of the adjoint when HALVE, the
// unnormalised inverse c2r otherwise): half-spectrum row G[0..L] in, Ny reals out,
//   H = G on 0<k<L (halved when HALVE), H[0] = Re G[0], H[L] = Re G[L],
//   Z[k] = (H[k] + conj H[L-k]) + i (H[k] - conj H[L-k]) conj(w_{Ny}^k),
//   out[2n] + i out[2n+1] = IDFT_L(Z) = conj(DFT(conj Z)).
template <typename T, bool HALVE>
__global__ void __launch_bounds__(GEN_NT) gen_row_c2r_kernel(RowArgs a, GenPlan g, const cpx<double> *tabL,
                                                            const cpx<double> *tab2)
{
    constexpr int CAP = gen_max_points<T>();
    __shared__ cpx<T> buf[2 * CAP];
    __shared__ cpx<T> roots[GEN_MAX_RADIX];
    const int t = threadIdx.x, L = g.n;
    const long long b = blockIdx.y;
    const int row = blockIdx.x;
    const T *in = reinterpret_cast<const T *>(a.in) + b * a.dist_in + (long long)row * a.ld_in;
    T *out = reinterpret_cast<T *>(a.out) + b * a.dist_out + (long long)row * a.ld_out;
    const T hs = HALVE ? T(0.5) : T(1);
    for (int k = t; k < L; k += GEN_NT)
    {
        cpx<T> z;
        if (k == 0)
        {
            const T h0 = stor<T>::ld(in).x, hl = stor<T>::ld(in + 2 * L).x;
            z = cpx<T>{h0 + hl, h0 - hl};
        }
        else
        {
            const cpx<T> hk = scal(stor<T>::ld(in + 2 * k), hs), hl = scal(stor<T>::ld(in + 2 * (L - k)), hs);
            const cpx<double> w64 = tab2[k];
            const cpx<T> wc = {(T)w64.x, -(T)w64.y};
            const cpx<T> s0 = hk + conjc(hl), d0 = cmul(hk - conjc(hl), wc);
            z = cpx<T>{s0.x - d0.y, s0.y + d0.x};
        }
        buf[k] = conjc(z);
    }
    __syncthreads();
    const cpx<T> *U = gen_fft<T>(buf, buf + CAP, roots, 1, g, tabL);
    for (int e = t; e < L; e += GEN_NT)
    {
        const cpx<T> u = U[e];
        stor<T>::st(out + 2 * e, cpx<T>{u.x, -u.y});
    }
}

}   // namespace hpxfft_dev
