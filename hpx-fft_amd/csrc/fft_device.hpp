// fft_device.hpp -- device building blocks for the gfx950 2-D r2c FFT.
//
// * complex arithmetic on float/double pairs,
// * in-register radix-2/4/8/16 forward DFT butterflies (natural-order out),
// * storage <-> compute conversions (f64, f32, f16, bf16 storage),
// * a workgroup Stockham engine: NSEQ independent L-point DFTs per
//   workgroup, E elements per thread held in VGPRs, one LDS exchange per
//   radix stage (optionally exchanged as separate real/imag halves so an
//   FP64 sub-FFT needs only 8 bytes of LDS per element).
//
// The 1-D transforms replace FFTW's r2c_1d / c2c_1d execute calls of
// core/src/util/adapter_fftw.cpp:12-15,32-35 (forward sign -1, unnormalised).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

namespace hpxfft_dev {

// ---------------------------------------------------------------------------
// XCD-aware block order.  gfx950 dispatches workgroup i (x fastest, then y,
// then z) to XCD i % 8, and each XCD has its own L2.  xcd_order(i, n) maps the
// dispatch index to a logical index such that XCD x works through one
// contiguous logical range, in order: neighbouring rows / column tiles then
// run at the same time behind the same L2, so the 128-byte lines they share
// (row pitches such as (Ny/2+1) complex are not multiples of 128 bytes) are
// fetched once and written back whole.  A bijection on [0, n) whatever the
// placement, so only locality depends on it.
constexpr int NUM_XCD = 8;

__device__ __forceinline__ int xcd_order(int i, int n)
{
    const int q = n / NUM_XCD, r = n % NUM_XCD;
    const int x = i % NUM_XCD, k = i / NUM_XCD;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

struct blk3 { int x, y, z; };

// the launch's (blockIdx.x, .y, .z) in XCD-contiguous logical order
__device__ __forceinline__ blk3 xcd_block()
{
    const int gx = (int)gridDim.x, gy = (int)gridDim.y;
    const int n = gx * gy * (int)gridDim.z;
    const int l = xcd_order((int)blockIdx.x + gx * ((int)blockIdx.y + gy * (int)blockIdx.z), n);
    const int yz = l / gx;
    return {l - yz * gx, yz % gy, yz / gy};
}

// ---------------------------------------------------------------------------
// complex helpers
template <typename T> struct cpx { T x, y; };

template <typename T> __device__ __forceinline__ cpx<T> mk(T a, T b) { return cpx<T>{a, b}; }
template <typename T> __device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T> __device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T> __device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T> __device__ __forceinline__ cpx<T> conjc(cpx<T> a) { return {a.x, -a.y}; }
template <typename T> __device__ __forceinline__ cpx<T> mul_mi(cpx<T> a) { return {a.y, -a.x}; }   // * (-i)
template <typename T> __device__ __forceinline__ cpx<T> scal(cpx<T> a, T s) { return {a.x * s, a.y * s}; }

template <typename T> struct kconst;
template <> struct kconst<double>
{
    static constexpr double r2 = 0.70710678118654752440;   // sqrt(2)/2
    static constexpr double c8 = 0.92387953251128675613;   // cos(pi/8)
    static constexpr double s8 = 0.38268343236508977173;   // sin(pi/8)
};
template <> struct kconst<float>
{
    static constexpr float r2 = 0.70710678118654752440f;
    static constexpr float c8 = 0.92387953251128675613f;
    static constexpr float s8 = 0.38268343236508977173f;
};

// ---------------------------------------------------------------------------
// forward DFT butterflies, in place, natural order output: a[q] <- sum_r a[r] w_R^{rq}
template <typename T> __device__ __forceinline__ void dft2(cpx<T> *a)
{
    cpx<T> t = a[0] - a[1];
    a[0] = a[0] + a[1];
    a[1] = t;
}

template <typename T> __device__ __forceinline__ void dft4v(cpx<T> &a0, cpx<T> &a1, cpx<T> &a2, cpx<T> &a3)
{
    cpx<T> t0 = a0 + a2, t1 = a0 - a2, t2 = a1 + a3, t3 = mul_mi(a1 - a3);
    a0 = t0 + t2;
    a1 = t1 + t3;
    a2 = t0 - t2;
    a3 = t1 - t3;
}

template <typename T> __device__ __forceinline__ void dft4(cpx<T> *a) { dft4v(a[0], a[1], a[2], a[3]); }

template <typename T> __device__ __forceinline__ cpx<T> mul_w8_1(cpx<T> z)   // * exp(-i pi/4)
{
    const T r = kconst<T>::r2;
    return {(z.x + z.y) * r, (z.y - z.x) * r};
}
template <typename T> __device__ __forceinline__ cpx<T> mul_w8_3(cpx<T> z)   // * exp(-3i pi/4)
{
    const T r = kconst<T>::r2;
    return {(z.y - z.x) * r, -(z.x + z.y) * r};
}

template <typename T> __device__ __forceinline__ void dft8(cpx<T> *a)
{
    cpx<T> e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
    cpx<T> o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
    dft4v(e0, e1, e2, e3);
    dft4v(o0, o1, o2, o3);
    o1 = mul_w8_1(o1);
    o2 = mul_mi(o2);
    o3 = mul_w8_3(o3);
    a[0] = e0 + o0;
    a[4] = e0 - o0;
    a[1] = e1 + o1;
    a[5] = e1 - o1;
    a[2] = e2 + o2;
    a[6] = e2 - o2;
    a[3] = e3 + o3;
    a[7] = e3 - o3;
}

// 16 = 4 x 4: n = n1 + 4 n2, k = 4 k1 + k2
template <typename T> __device__ __forceinline__ void dft16(cpx<T> *a)
{
    const T c = kconst<T>::c8, s = kconst<T>::s8;
    cpx<T> y[16];
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1)
    {
        cpx<T> b0 = a[n1], b1 = a[n1 + 4], b2 = a[n1 + 8], b3 = a[n1 + 12];
        dft4v(b0, b1, b2, b3);
        y[n1 * 4 + 0] = b0;
        y[n1 * 4 + 1] = b1;
        y[n1 * 4 + 2] = b2;
        y[n1 * 4 + 3] = b3;
    }
    // twiddles w16^{n1 k2}
    y[1 * 4 + 1] = cmul(y[5], cpx<T>{c, -s});     // w^1
    y[1 * 4 + 2] = mul_w8_1(y[6]);               // w^2
    y[1 * 4 + 3] = cmul(y[7], cpx<T>{s, -c});     // w^3
    y[2 * 4 + 1] = mul_w8_1(y[9]);               // w^2
    y[2 * 4 + 2] = mul_mi(y[10]);                // w^4
    y[2 * 4 + 3] = mul_w8_3(y[11]);              // w^6
    y[3 * 4 + 1] = cmul(y[13], cpx<T>{s, -c});    // w^3
    y[3 * 4 + 2] = mul_w8_3(y[14]);              // w^6
    y[3 * 4 + 3] = cmul(y[15], cpx<T>{-c, s});    // w^9
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2)
    {
        cpx<T> b0 = y[0 * 4 + k2], b1 = y[1 * 4 + k2], b2 = y[2 * 4 + k2], b3 = y[3 * 4 + k2];
        dft4v(b0, b1, b2, b3);
        a[0 + k2] = b0;
        a[4 + k2] = b1;
        a[8 + k2] = b2;
        a[12 + k2] = b3;
    }
}

template <typename T, int R> __device__ __forceinline__ void dftR(cpx<T> *a)
{
    if constexpr (R == 2) dft2(a);
    else if constexpr (R == 4) dft4(a);
    else if constexpr (R == 8) dft8(a);
    else if constexpr (R == 16) dft16(a);
}

// ---------------------------------------------------------------------------
// storage types.  f16/bf16 storage is always computed in f32.
struct bf16_t { uint16_t v; };

template <typename S> struct stor;
template <> struct stor<double>
{
    using pair = double2;
    __device__ static cpx<double> ld(const double *p)
    {
        double2 v = *reinterpret_cast<const double2 *>(p);
        return {v.x, v.y};
    }
    __device__ static void st(double *p, cpx<double> v) { *reinterpret_cast<double2 *>(p) = make_double2(v.x, v.y); }
};
template <> struct stor<float>
{
    __device__ static cpx<float> ld(const float *p)
    {
        float2 v = *reinterpret_cast<const float2 *>(p);
        return {v.x, v.y};
    }
    __device__ static void st(float *p, cpx<float> v) { *reinterpret_cast<float2 *>(p) = make_float2(v.x, v.y); }
};

// Streaming (nontemporal) access of one complex FP32/FP64 value: the load or
// store carries the nontemporal hint, so the data does not displace cache
// lines other kernels will re-read.  Used where measurement showed a gain
// (FP64 column-pass stores, FP32 row-pass input loads); other storage types
// take the plain access.
template <typename S> __device__ __forceinline__ cpx<S> ld_stream(const S *p)
{
    if constexpr (sizeof(S) >= 4)
    {
        typedef S v2 __attribute__((ext_vector_type(2)));
        const v2 v = __builtin_nontemporal_load(reinterpret_cast<const v2 *>(p));
        return cpx<S>{v.x, v.y};
    }
    else return stor<S>::ld(p);
}
template <typename S, typename T> __device__ __forceinline__ void st_stream(S *p, cpx<T> z)
{
    if constexpr (sizeof(S) >= 4 && sizeof(S) == sizeof(T))
    {
        typedef S v2 __attribute__((ext_vector_type(2)));
        v2 v;
        v.x = z.x;
        v.y = z.y;
        __builtin_nontemporal_store(v, reinterpret_cast<v2 *>(p));
    }
    else stor<S>::st(p, z);
}
// The same for one lane of a wave whose 64 lanes access 64 consecutive
// complex elements (DIR = +1: lane l at element e0 + l; DIR = -1: e0 - l).
// Rows start at any 16-byte offset of a 128-byte line, so such a wave access
// begins and ends inside lines it shares with the neighbouring access: those
// lanes take the plain access (the L2 merges the two halves of the line),
// the lanes of whole lines take the streaming one.  Only the cache hint
// depends on the test; the value loaded or stored is the same either way.
template <typename S, int DIR>
__device__ __forceinline__ bool in_whole_line(const S *p, int lane)
{
    constexpr long long ES = 2 * (long long)sizeof(S);   // bytes per complex element
    const long long a = (long long)reinterpret_cast<uintptr_t>(p);
    const long long lo = DIR > 0 ? a - lane * ES : a - (63 - lane) * ES;   // the wave's byte range [lo, lo + 64 ES)
    const long long line = a & ~127LL;
    return line >= lo && line + 128 <= lo + 64 * ES;
}
// streaming access where `whole` (in_whole_line of the lane), plain otherwise
template <typename S> __device__ __forceinline__ cpx<S> ld_stream_if(const S *p, bool whole)
{
    if (whole) return ld_stream<S>(p);
    return stor<S>::ld(p);
}
template <typename S, typename T> __device__ __forceinline__ void st_stream_if(S *p, cpx<T> z, bool whole)
{
    if (whole) st_stream<S>(p, z);
    else stor<S>::st(p, z);
}
template <> struct stor<__half>
{
    __device__ static cpx<float> ld(const __half *p)
    {
        __half2 v = *reinterpret_cast<const __half2 *>(p);
        float2 f = __half22float2(v);
        return {f.x, f.y};
    }
    __device__ static void st(__half *p, cpx<float> v)
    {
        *reinterpret_cast<__half2 *>(p) = __floats2half2_rn(v.x, v.y);
    }
};
__device__ __forceinline__ float bf16_to_f(uint32_t h) { return __uint_as_float(h << 16); }
// two floats -> two BF16 (round to nearest even) packed low/high in one word:
// one v_cvt_pk_bf16_f32 on gfx950
typedef float f32x2_v __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t f2_to_bf16x2(float lo, float hi)
{
    const f32x2_v f = {lo, hi};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f, bf16x2_v));
}
template <> struct stor<bf16_t>
{
    __device__ static cpx<float> ld(const bf16_t *p)
    {
        uint32_t v = *reinterpret_cast<const uint32_t *>(p);
        return {bf16_to_f(v & 0xffffu), bf16_to_f(v >> 16)};
    }
    __device__ static void st(bf16_t *p, cpx<float> v)
    {
        *reinterpret_cast<uint32_t *>(p) = f2_to_bf16x2(v.x, v.y);
    }
};

// single real element load/store (for c2r outputs and partial rows)
template <typename S, typename T> __device__ __forceinline__ T ld1(const S *p)
{
    if constexpr (sizeof(S) == 2)
    {
        if constexpr (__is_same(S, __half)) return __half2float(*p);
        else return bf16_to_f(reinterpret_cast<const bf16_t *>(p)->v);
    }
    else return (T)*p;
}

// ---------------------------------------------------------------------------
// Stockham engine.
//
// NSEQ sequences of length L per workgroup, NT = NSEQ*L/E threads, each
// holding E elements.  Stage radices: RM = min(E,16) repeated, then the
// remainder.  COL=false: LDS layout [seq][e] (row sub-FFTs: consecutive
// threads take consecutive elements of one sequence).  COL=true: LDS layout
// [e][seq] (column tiles: consecutive threads take consecutive sequences =
// adjacent columns, so global accesses are row segments).
//
// Stage s (radix R, Ns = product of earlier radices), butterfly j of a
// sequence: inputs e = j + r L/R, twiddled by w_{Ns R}^{r (j mod Ns)},
// DFT_R, outputs e = (j/Ns) Ns R + (j mod Ns) + q Ns   (Stockham autosort).
// ---------------------------------------------------------------------------
constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v >> 1); }

template <int L, int RM> struct stage_info
{
    static constexpr int nfull = ilog2c(L) / ilog2c(RM);
    static constexpr int rem = L >> (nfull * ilog2c(RM));   // 1 if none
    static constexpr int nstages = nfull + (rem > 1 ? 1 : 0);
    static constexpr int radix(int s) { return s < nfull ? RM : rem; }
    static constexpr int ns(int s) { return s == 0 ? 1 : ns(s - 1) * radix(s - 1); }
};

template <typename T, int L, int E, int NSEQ, bool COL, bool SPLIT> struct Engine
{
    static_assert((L & (L - 1)) == 0, "power of two");
    static constexpr int RM = E < 16 ? E : 16;
    static constexpr int NT = NSEQ * L / E;
    static_assert(NT >= 1 && NT * E == NSEQ * L, "E must divide NSEQ*L");
    static constexpr int RMX = (L >= RM ? RM : L) < 2 ? 2 : (L >= RM ? RM : L);
    using SI = stage_info<L, RMX>;
    static constexpr int NSTAGES = L > 1 ? SI::nstages : 0;

    // padding: one slot per 16 elements when a row of the layout is < 128 B
    static constexpr int ELEM_BYTES = SPLIT ? (int)sizeof(T) : 2 * (int)sizeof(T);
    static constexpr bool PAD = COL ? (NSEQ * ELEM_BYTES < 128 && L >= 16) : (L >= 16);
    static constexpr int LP = PAD ? L + L / 16 : L;
    static constexpr int LDS_ELEMS = NSEQ * LP;
    using lds_t = typename std::conditional<SPLIT, T, cpx<T>>::type;

    __device__ static __forceinline__ int pe(int e) { return PAD ? e + (e >> 4) : e; }
    __device__ static __forceinline__ int addr(int seq, int e) { return COL ? pe(e) * NSEQ + seq : seq * LP + pe(e); }

    template <int R> __device__ static __forceinline__ void bmap(int g, int &seq, int &j)
    {
        if constexpr (COL) { seq = g % NSEQ; j = g / NSEQ; }
        else { seq = g / (L / R); j = g % (L / R); }
    }

    template <int S> static constexpr int radix() { return SI::radix(S); }
    template <int S> static constexpr int nsv() { return SI::ns(S); }

    // twiddle + butterflies of stage S on registers v
    template <int S> __device__ static __forceinline__ void compute(cpx<T> *v, const cpx<T> *__restrict__ tw, int tws, int t)
    {
        constexpr int R = radix<S>();
        constexpr int Ns = nsv<S>();
#pragma unroll
        for (int i = 0; i < E / R; ++i)
        {
            if constexpr (Ns > 1)
            {
                int seq, j;
                bmap<R>(t + i * NT, seq, j);
                const int k = j % Ns;
                const int base = k * (L / (Ns * R)) * tws;
#pragma unroll
                for (int r = 1; r < R; ++r) v[i * R + r] = cmul(v[i * R + r], tw[base * r]);
            }
            dftR<T, R>(v + i * R);
        }
    }

    // write stage-S outputs to LDS, read stage S+1 inputs (Stockham exchange)
    template <int S> __device__ static __forceinline__ void exchange(cpx<T> *v, lds_t *lds, int t)
    {
        constexpr int R = radix<S>();
        constexpr int Ns = nsv<S>();
        constexpr int R2 = radix<S + 1>();
        int oa[E], ia[E];
#pragma unroll
        for (int i = 0; i < E / R; ++i)
        {
            int seq, j;
            bmap<R>(t + i * NT, seq, j);
            const int ob = (j / Ns) * Ns * R + (j % Ns);
#pragma unroll
            for (int q = 0; q < R; ++q) oa[i * R + q] = addr(seq, ob + q * Ns);
        }
#pragma unroll
        for (int i = 0; i < E / R2; ++i)
        {
            int seq, j;
            bmap<R2>(t + i * NT, seq, j);
#pragma unroll
            for (int r = 0; r < R2; ++r) ia[i * R2 + r] = addr(seq, j + r * (L / R2));
        }
        if constexpr (!SPLIT)
        {
#pragma unroll
            for (int e = 0; e < E; ++e) lds[oa[e]] = v[e];
            __syncthreads();
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = lds[ia[e]];
            __syncthreads();
        }
        else
        {
#pragma unroll
            for (int e = 0; e < E; ++e) lds[oa[e]] = v[e].x;
            __syncthreads();
#pragma unroll
            for (int e = 0; e < E; ++e) v[e].x = lds[ia[e]];
            __syncthreads();
#pragma unroll
            for (int e = 0; e < E; ++e) lds[oa[e]] = v[e].y;
            __syncthreads();
#pragma unroll
            for (int e = 0; e < E; ++e) v[e].y = lds[ia[e]];
            __syncthreads();
        }
    }

    template <int S> __device__ static __forceinline__ void run_from(cpx<T> *v, lds_t *lds, const cpx<T> *__restrict__ tw, int tws, int t)
    {
        if constexpr (S < NSTAGES)
        {
            compute<S>(v, tw, tws, t);
            if constexpr (S + 1 < NSTAGES)
            {
                exchange<S>(v, lds, t);
                run_from<S + 1>(v, lds, tw, tws, t);
            }
        }
    }

    // all stages; on entry v holds stage-0 inputs (see in_elem), on exit the
    // last stage's outputs (see out_elem).
    __device__ static __forceinline__ void run(cpx<T> *v, lds_t *lds, const cpx<T> *__restrict__ tw, int tws, int t)
    {
        run_from<0>(v, lds, tw, tws, t);
    }

    // (seq, element) that register slot `x` of thread t holds before stage 0
    __device__ static __forceinline__ void in_elem(int t, int x, int &seq, int &e)
    {
        if constexpr (NSTAGES == 0)
        {
            // L == 1: one element per sequence
            if constexpr (COL) { seq = (t + x * NT) % NSEQ; e = 0; }
            else { seq = t + x * NT; e = 0; }
        }
        else
        {
            constexpr int R = radix<0>();
            const int i = x / R, r = x % R;
            int j;
            bmap<R>(t + i * NT, seq, j);
            e = j + r * (L / R);
        }
    }
    // (seq, element) that register slot `x` holds after the last stage
    __device__ static __forceinline__ void out_elem(int t, int x, int &seq, int &e)
    {
        if constexpr (NSTAGES == 0) { in_elem(t, x, seq, e); }
        else
        {
            constexpr int S = NSTAGES - 1;
            constexpr int R = radix<S>();
            constexpr int Ns = nsv<S>();
            const int i = x / R, q = x % R;
            int j;
            bmap<R>(t + i * NT, seq, j);
            e = (j / Ns) * Ns * R + (j % Ns) + q * Ns;
        }
    }
};

}   // namespace hpxfft_dev
