// capi.hip -- plans, twiddle tables, launch sequencing and the C ABI of
// include/hpxfft_hip.h.
//
// A plan is the GPU image of hpxfft::shared::loop after initialize
// (core/src/shared/loop.cpp:158-189): dimensions, the chosen 1-D kernels
// (instead of FFTW plans), device scratch (instead of the transpose buffer
// trans_values_vec_, which the fused column pass does not need at full
// size), twiddle tables, timing events and the measurement map.
#include "launch.hpp"
#include "fft_general.hpp"
#include "batched.hpp"
#include "../../include/hpxfft_hip.h"

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

using namespace hpxfft_host;

// launch a general kernel built for the launch bound that covers `nt`
// threads (256, or 1024 for the larger workgroups); TARGS are its other
// template arguments in parentheses
#define GEN_UNPAREN(...) __VA_ARGS__
#define GEN_LAUNCH(KERN, TARGS, nt, grid, lds, stream, ...)                                                        \
    do                                                                                                             \
    {                                                                                                              \
        if ((nt) <= (unsigned)hpxfft_dev::GEN_NT)                                                                  \
            hipLaunchKernelGGL((KERN<GEN_UNPAREN TARGS, hpxfft_dev::GEN_NT>), grid, dim3(nt), lds, stream,         \
                               __VA_ARGS__);                                                                       \
        else                                                                                                       \
            hipLaunchKernelGGL((KERN<GEN_UNPAREN TARGS, hpxfft_dev::GEN_NT_MAX>), grid, dim3(nt), lds, stream,     \
                               __VA_ARGS__);                                                                       \
    } while (0)

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

// run an entry point's body; no exception crosses the C ABI
template <typename F> int guarded(F &&body)
{
    try
    {
        return body();
    }
    catch (const std::bad_alloc &)
    {
        return fail(HPXFFT_OUT_OF_MEMORY, "host allocation failed");
    }
    catch (const std::exception &e)
    {
        return fail(HPXFFT_HIP_ERROR, e.what());
    }
    catch (...)
    {
        return fail(HPXFFT_HIP_ERROR, "unexpected exception");
    }
}

#define HIPCHK(x)                                                                                                  \
    do                                                                                                             \
    {                                                                                                              \
        hipError_t e_ = (x);                                                                                       \
        if (e_ != hipSuccess)                                                                                      \
        {                                                                                                          \
            if (e_ == hipErrorOutOfMemory || e_ == hipErrorMemoryAllocation)                                       \
                return fail(HPXFFT_OUT_OF_MEMORY, std::string(#x) + ": " + hipGetErrorString(e_));                \
            return fail(HPXFFT_HIP_ERROR, std::string(#x) + ": " + hipGetErrorString(e_));                         \
        }                                                                                                          \
    } while (0)

bool is_pow2(size_t v) { return v && !(v & (v - 1)); }
int ilog2(size_t v)
{
    int r = 0;
    while ((size_t(1) << r) < v) ++r;
    return r;
}

// ---------------------------------------------------------------------------
// Twiddle tables: for every N = 2^k, k = 0..16, the entries
// w_N^m = exp(-2 pi i m / N), m < N, stored at [N, 2N) of one array.  Each
// value is evaluated directly (no recurrence) in long double after reducing
// 8m/N to the first octant, so axis and diagonal values are exact.  The FP32
// table is the FP64 table rounded (the only FP64 step of the FP32 path).
constexpr int TW_MAX_LG = 16;

void twiddle_ld(uint64_t m, uint64_t n, long double &c, long double &s)
{
    m %= n;
    const long double PI_L = 3.141592653589793238462643383279502884L;
    const uint64_t m8 = 8 * m, oct = m8 / n, rem = m8 - oct * n;
    const bool odd = oct & 1;
    const long double t = (long double)(odd ? n - rem : rem) / (long double)n * (PI_L / 4.0L);
    const long double ct = rem == 0 && !odd ? 1.0L : cosl(t);
    const long double st = rem == 0 && !odd ? 0.0L : sinl(t);
    switch (oct)
    {
    case 0: c = ct; s = st; break;
    case 1: c = st; s = ct; break;
    case 2: c = -st; s = ct; break;
    case 3: c = -ct; s = st; break;
    case 4: c = -ct; s = -st; break;
    case 5: c = -st; s = -ct; break;
    case 6: c = st; s = -ct; break;
    default: c = ct; s = -st; break;
    }
}

struct dev_tables
{
    double *tw64 = nullptr;   // complex double pairs
    float *tw32 = nullptr;    // complex float pairs
    double *stw64 = nullptr;  // radix-16 stage twiddles of the long-row kernels (stage_tw_off layout)
    float *stw32 = nullptr;
};

// The tables live in static device memory: one copy per device, filled from
// the host by the first get_tables call on that device, shared by every plan
// and by the allocation-free entry point, held until the process ends.
constexpr size_t TW_ENTRIES = size_t(2) << TW_MAX_LG;   // entries [N, 2N) for N = 1 .. 2^TW_MAX_LG
__device__ __align__(256) double tw64_table[2 * TW_ENTRIES];
__device__ __align__(256) float tw32_table[2 * TW_ENTRIES];
// the long-row kernels' stage tables: for NS = 16 and 256, the values w_{16 NS}^{m k},
// m = 1, 2, 4, 8, k < NS, as four arrays (entries of the tables above, regrouped)
__device__ __align__(256) double stw64_table[2 * hpxfft_dev::STAGE_TW_ENTRIES];
__device__ __align__(256) float stw32_table[2 * hpxfft_dev::STAGE_TW_ENTRIES];

std::mutex g_tab_mu;
std::map<int, dev_tables> g_tables;

// with device `dev` current
int get_tables(int dev, dev_tables &out)
{
    std::lock_guard<std::mutex> lk(g_tab_mu);
    auto it = g_tables.find(dev);
    if (it != g_tables.end())
    {
        out = it->second;
        return HPXFFT_OK;
    }
    const size_t n = TW_ENTRIES;
    std::vector<double> h64(2 * n, 0.0);
    std::vector<float> h32(2 * n, 0.0f);
    for (int k = 0; k <= TW_MAX_LG; ++k)
    {
        const uint64_t N = uint64_t(1) << k;
        for (uint64_t m = 0; m < N; ++m)
        {
            long double c, s;
            twiddle_ld(m, N, c, s);
            const size_t idx = N + m;
            h64[2 * idx] = (double)c;
            h64[2 * idx + 1] = -(double)s;
            h32[2 * idx] = (float)h64[2 * idx];
            h32[2 * idx + 1] = (float)h64[2 * idx + 1];
        }
    }
    constexpr int SE = hpxfft_dev::STAGE_TW_ENTRIES;
    std::vector<double> s64(2 * SE);
    std::vector<float> s32(2 * SE);
    for (const int NS : {16, 256})
    {
        const int off = NS == 16 ? hpxfft_dev::stage_tw_off<16>() : hpxfft_dev::stage_tw_off<256>();
        const uint64_t N = 16 * (uint64_t)NS;
        for (int q = 0; q < 4; ++q)
            for (int k = 0; k < NS; ++k)
            {
                const size_t src = N + ((uint64_t)(1 << q) * k) % N, dst = off + q * NS + k;   // w_N^{2^q k}
                s64[2 * dst] = h64[2 * src];
                s64[2 * dst + 1] = h64[2 * src + 1];
                s32[2 * dst] = h32[2 * src];
                s32[2 * dst + 1] = h32[2 * src + 1];
            }
    }
    dev_tables t;
    HIPCHK(hipGetSymbolAddress(reinterpret_cast<void **>(&t.tw64), HIP_SYMBOL(tw64_table)));
    HIPCHK(hipGetSymbolAddress(reinterpret_cast<void **>(&t.tw32), HIP_SYMBOL(tw32_table)));
    HIPCHK(hipGetSymbolAddress(reinterpret_cast<void **>(&t.stw64), HIP_SYMBOL(stw64_table)));
    HIPCHK(hipGetSymbolAddress(reinterpret_cast<void **>(&t.stw32), HIP_SYMBOL(stw32_table)));
    HIPCHK(hipMemcpy(t.tw64, h64.data(), sizeof(double) * 2 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(t.tw32, h32.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(t.stw64, s64.data(), sizeof(double) * 2 * SE, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(t.stw32, s32.data(), sizeof(float) * 2 * SE, hipMemcpyHostToDevice));
    g_tables[dev] = t;
    out = t;
    return HPXFFT_OK;
}

struct device_guard
{
    int prev = -1;
    explicit device_guard(int dev)
    {
        if (hipGetDevice(&prev) == hipSuccess && prev != dev) (void)hipSetDevice(dev);
        else prev = -1;
    }
    ~device_guard()
    {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

int check_device(int &dev)
{
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) return fail(HPXFFT_NO_DEVICE, "no HIP device available");
    HIPCHK(hipGetDevice(&dev));
    return HPXFFT_OK;
}

size_t storage_bytes(int prec) { return prec == P_F64 ? 8 : prec == P_F32 ? 4 : 2; }
size_t compute_bytes(int prec) { return prec == P_F64 ? 8 : 4; }
const char *prec_name(int prec)
{
    switch (prec)
    {
    case P_F64: return "f64";
    case P_F32: return "f32";
    case P_F16: return "f16-storage/f32-compute";
    default: return "bf16-storage/f32-compute";
    }
}

}   // namespace

// ---------------------------------------------------------------------------
// a dimension of length N = P * Q beyond one workgroup: phase A runs the
// Q-point plan A over elements m1 + P*e, phase B the P-point plan B
struct gen2_dim
{
    bool on = false;       // the dimension runs as two levels
    bool avail = false;    // tables and scratch in place (on, or a candidate the measure effort times)
    int P = 0, Q = 0;
    hpxfft_dev::GenPlan A{}, B{};
    int WA = 1, WB = 1;                    // sequences per workgroup in each phase
    size_t tabA = 0, tabB = 0, tabT = 0;   // table offsets (complex entries): plans A, B and w_{PQ}^m
    int tw_shift = 0;                      // w_{PQ} table split at 2^tw_shift (0: one table)
    // Bluestein over two levels: the length-N DFT as a chirp convolution of
    // length M = P*Q (a power of two >= 2N - 1), three passes
    bool blue = false;
    size_t N = 0;
    size_t tabC = 0, tabBh = 0;            // the chirp table w_{2N} and Bh = DFT_M(b) / M
    int chirp_shift = 0;
    size_t points() const { return (size_t)P * (size_t)Q; }   // scratch points per sequence
};

struct hpxfft_plan_s
{
    // problem
    size_t nx = 0, ny = 0, batch = 1, ld_in = 0, dist_in = 0, ld_out = 0, dist_out = 0;
    int prec = P_F64, effort = 0;
    hipStream_t stream = nullptr;
    int device = 0;
    int lgx = 0, lgL = 0;
    size_t ncols = 0;
    // row kernel for power-of-two rows of 2^13 / 2^14 complex points with FP32
    // arithmetic: 0 = radix-16 stages (row_r2c_big_kernel), 1 = lane-pair
    // radix-32 stages (row_r2c_pair_kernel)
    int row_kind = 0;
    double row_ms[2] = {};   // measure effort: row pass ms of each kind
    // column strategy
    bool four_step = false;
    int lgP = 0, lgQ = 0;
    int W = 8;
    size_t CW = 0;   // chunk width (columns) of the four-step scratch
    // lengths that are not powers of two: general kernels (F64 / F32 forward)
    bool gen_row = false, gen_col = false;
    // 16-bit storage with a length beyond one workgroup: the input converted
    // to an FP32 workspace (in-place padded layout), transformed by an FP32
    // plan of the same shape, converted back with per-column exponents
    bool wide16 = false;
    hpxfft_plan_s *wide = nullptr;
    float *wide_ws = nullptr;
    hpxfft_dev::GenPlan grow{}, gcol{};
    int gen_w = 1;                 // general column kernel: adjacent columns per workgroup
    int gen_rows = 1;              // general row kernels: rows per workgroup
    gen2_dim g2r, g2c;             // two-level general rows (Ny/2 = P*Q) / columns (Nx = P*Q)
    void *gtmp = nullptr;          // two-level rows: phase A output, Nx x Ny/2 per grid (compute type)
    void *gscr = nullptr;          // two-level columns: phase A output, Nx x (Ny/2+1) per grid
    double gcol_ms[2] = {};        // measure effort: general column pass ms, one level / two levels
    double gip_ms[2][2] = {};      // measure effort: general [row, column] pass ms, [ping-pong, in place]
    double *gen_tab = nullptr;     // exact FP64 tables w_n^m: rows L = Ny/2, split Ny, columns Nx
    size_t gtab_L = 0, gtab_2 = 0, gtab_N = 0;   // their offsets (complex entries)
    int tw2_shift = 0;                            // the w_Ny table split at 2^tw2_shift (0: whole)
    // device memory
    dev_tables tabs;
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    int *mid_exp = nullptr;       // 16-bit: per-row exponents after the row pass
    unsigned *colpart = nullptr;  // 16-bit four-step: per-workgroup column maxima, P x ncols
    unsigned *colmax = nullptr;   // 16-bit four-step: per-column maxima, ncols
    int *col_exp_own = nullptr;   // 16-bit: output exponents when the caller gives none
    void *adj_work = nullptr;     // adjoint: column-pass result
    size_t adj_bytes = 0;
    void *host_buf = nullptr;     // drop-in host path: device copy of the vector_2d
    size_t host_bytes = 0;
    // timing
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};   // start, rows done, columns done, column phase A done
    bool ran = false;
    bool timed = true;           // exec records the phase events (off for the allocation-free entry point)
    bool col_split = false;      // ev[3] valid: the last run's column pass was a four-step
    bool col_first = false;      // the last run was backward: column pass (ev0..ev1) before the rows
    double h2d = 0.0, d2h = 0.0;
    double plan_seconds = 0.0, plan_flops = 0.0;
    std::map<std::string, double> extra;
    // batched (uneven) plans
    bool batched = false;
    std::vector<size_t> bnx, bny;
    std::vector<hpxfft_plan> sub;   // unused (kept empty)
    int bcfg[2] = {0, 0};                                   // chosen configuration per phase
    void *bgrids = nullptr;                                 // device grid descriptors
    void *btasks[2][hpxfft_host::BATCH_NCFG] = {};          // device task tables
    int bntasks[2][hpxfft_host::BATCH_NCFG] = {};
    int bblocks[2][hpxfft_host::BATCH_NCFG] = {};
    double btune_ms[2][hpxfft_host::BATCH_NCFG] = {};
    double split_ms[COL_TILE_MAX_LG + 1] = {};   // four-step: column-pass ms per tried log2 P (effort >= measure)
    unsigned *bticket = nullptr;                            // work-queue counters: BATCH_QUEUES per phase
    void *bstage = nullptr;                                 // pinned staging of the grid descriptors, two slots
    hipEvent_t bstage_ev[2] = {nullptr, nullptr};           // each slot's last upload (the slot is free once it ran)
    int bstage_next = 0;                                    // slot of the next upload
    std::vector<void *> bptrs;                              // grid pointers the device descriptors hold now
};

namespace {

const void *tw_ptr(const hpxfft_plan_s *p, size_t N)
{
    if (p->prec == P_F64) return p->tabs.tw64 + 2 * N;
    return p->tabs.tw32 + 2 * N;
}
const void *tw64_ptr(const hpxfft_plan_s *p, size_t N) { return p->tabs.tw64 + 2 * N; }

// dynamic LDS of a general kernel: two buffers of `points` complex values in
// the compute type, then `words` 32-bit words (16-bit storage: row / column maxima)
size_t gen_lds_bytes(const hpxfft_plan_s *p, const hpxfft_dev::GenPlan &g, size_t points, size_t words = 0)
{
    return (size_t)hpxfft_dev::gen_nbuf(g) * points * 2 * compute_bytes(p->prec) + 4 * words;
}

// threads of a general workgroup whose dynamic LDS is `lds` bytes: 256 while
// three or more workgroups fit a CU's 160 KiB; a block that leaves room for
// only two (one) workgroups per CU takes 512 (1024) threads, so the CU still
// has 16-32 waves to hide latency with (measured: 3000 x 3000 FP64 columns,
// one column of 96 KiB per workgroup, 134 -> 116 us; rows of 48 KiB stay
// fastest at 256)
unsigned gen_threads(size_t lds, const hpxfft_dev::GenPlan &g)
{
    constexpr size_t CU_LDS = 160 * 1024;
    const size_t per_cu = CU_LDS / std::max<size_t>(lds, 1);
    // a prime radix above 7 runs the generic stage, whose outputs each loop
    // over r inputs: twice the threads hide that loop's latency (measured:
    // rows of 516 = 4*3*43 points, 46 -> 34 us at 512 threads)
    bool generic = false;
    for (int i = 0; i < g.nstages; ++i) generic = generic || g.radix[i] > 8;
    if (per_cu >= 3) return generic ? 512u : 256u;
    return per_cu == 2 ? 512u : 1024u;
}

// let every general kernel take up to GEN_LDS_BYTES of dynamic LDS (above the
// 64 KiB default); once per device
int gen_allow_lds(int dev)
{
    static std::mutex mu;
    static std::map<int, bool> done;
    std::lock_guard<std::mutex> lk(mu);
    if (done[dev]) return HPXFFT_OK;
    using namespace hpxfft_dev;
#define GEN_BOTH(KERN, ...) reinterpret_cast<const void *>(&KERN<__VA_ARGS__, GEN_NT>), \
                            reinterpret_cast<const void *>(&KERN<__VA_ARGS__, GEN_NT_MAX>)
    const void *fns[] = {
        GEN_BOTH(gen_row_r2c_kernel, double, double), GEN_BOTH(gen_row_r2c_kernel, float, float),
        GEN_BOTH(gen_row_r2c_kernel, float, __half), GEN_BOTH(gen_row_r2c_kernel, float, bf16_t),
        GEN_BOTH(gen_col_kernel, double, double, false), GEN_BOTH(gen_col_kernel, double, double, true),
        GEN_BOTH(gen_col_kernel, float, float, false), GEN_BOTH(gen_col_kernel, float, float, true),
        GEN_BOTH(gen_col_kernel, float, __half, false), GEN_BOTH(gen_col_kernel, float, bf16_t, false),
        GEN_BOTH(gen_row_c2r_kernel, double, false), GEN_BOTH(gen_row_c2r_kernel, double, true),
        GEN_BOTH(gen_row_c2r_kernel, float, false), GEN_BOTH(gen_row_c2r_kernel, float, true),
        GEN_BOTH(gen_strided_kernel, double, false), GEN_BOTH(gen_strided_kernel, float, false),
        GEN_BOTH(gen_strided_kernel, double, true), GEN_BOTH(gen_strided_kernel, float, true),
        GEN_BOTH(gen_rows_split_kernel, double, false), GEN_BOTH(gen_rows_split_kernel, float, false),
        GEN_BOTH(gen_rows_split_kernel, double, true), GEN_BOTH(gen_rows_split_kernel, float, true),
        GEN_BOTH(gen_blue_mid_kernel, double), GEN_BOTH(gen_blue_mid_kernel, float)};
#undef GEN_BOTH
    // the buffers plus up to 1024 words of per-row / per-column maxima
    for (const void *f : fns)
        HIPCHK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, GEN_LDS_BYTES + 4096));
    done[dev] = true;
    return HPXFFT_OK;
}

// Stockham radices for a length that is not a power of two: for its 2^e
// part 8s (then a 4 or a 2) where they take fewer stages than 4s (then a 2),
// else 4s; then odd primes ascending; false if a prime factor exceeds
// GEN_MAX_RADIX.  A radix-8 stage is heavier than a radix-4 one, so it pays
// only by removing a stage (measured: 4096-point Bluestein FFTs 6 -> 4
// stages, 2039 x 4078 FP64 0.315 -> 0.296 ms; 6000 = 8*2*3*5^3 in as many
// stages as 4*4*3*5^3 was 6% slower)
bool gen_factor(size_t n, hpxfft_dev::GenPlan &g)
{
    g = hpxfft_dev::GenPlan{};
    g.n = (int)n;
    size_t m = n;
    auto push = [&](int r) {
        if (g.nstages >= hpxfft_dev::GEN_MAX_STAGES) return false;
        g.radix[g.nstages++] = r;
        return true;
    };
    int e2 = 0;
    for (size_t r = m; r % 2 == 0; r /= 2) ++e2;
    if ((e2 + 2) / 3 < (e2 + 1) / 2)
        while (m % 8 == 0) { if (!push(8)) return false; m /= 8; }
    while (m % 4 == 0) { if (!push(4)) return false; m /= 4; }
    if (m % 2 == 0) { if (!push(2)) return false; m /= 2; }
    for (size_t f = 3; m > 1; f += 2)
    {
        while (m % f == 0)
        {
            if (f > (size_t)hpxfft_dev::GEN_MAX_RADIX || !push((int)f)) return false;
            m /= f;
        }
        if (f * f > m && m > 1)
        {
            if (m > (size_t)hpxfft_dev::GEN_MAX_RADIX || !push((int)m)) return false;
            m = 1;
        }
    }
    return true;
}

// a general plan for length n: Stockham stages when its prime factors are
// at most GEN_MAX_RADIX and it fits `cap` points (the ping-pong buffers), or
// in-place stages alone when its radices are at most 7 and it fits `ip_cap`
// points (one buffer of up to twice as many; measured against two levels:
// 6000 x 12000 FP64 2.58 -> 1.45 ms, 5000^2 FP64 0.59 -> 0.38 ms), else
// Bluestein's algorithm through the power-of-two FFT of length >= 2n - 1
// when that fits `cap`
bool gen_setup(size_t n, size_t cap, hpxfft_dev::GenPlan &g, size_t ip_cap = 0)
{
    bool long_ok = false;
    if (n > cap && n <= ip_cap && gen_factor(n, g))
    {
        long_ok = true;
        for (int i = 0; i < g.nstages; ++i) long_ok = long_ok && g.radix[i] <= 8;
    }
    if ((n <= cap || long_ok) && gen_factor(n, g))
    {
        g.len = (int)n;
        g.off_c = g.off_b = 0;
        // register radices only: in-place stages, their input positions
        // after the n twiddles.  In place is the estimate wherever the
        // radices allow it (one buffer instead of two; measure effort times
        // both forms, tune_gen_ip): 3000 x 3000 FP64 columns of 96 KiB per
        // workgroup 117 -> 79 us, rows of 1500 points 70 -> 58 us; rows of
        // 4 x 500 points 16 -> 17 us, the one case measured slower
        bool small = true;
        for (int i = 0; i < g.nstages; ++i) small = small && g.radix[i] <= 8;
        g.off_r = small ? g.n : -1;
        g.off_bp = -1;
        g.ip = small ? 1 : 0;
        return true;
    }
    size_t M = 1;
    while (M < 2 * n - 1) M *= 2;
    if (M > cap || !gen_factor(M, g)) return false;
    g.len = (int)n;
    g.off_c = (int)M;
    g.off_b = (int)(M + n);
    // the power-of-two FFT (radices 4 and 2) runs in place by default: the
    // kernel spectrum also in pos() order
    g.off_r = -1;
    g.off_bp = (int)(2 * M + n);
    g.ip = 1;
    return true;
}

// whether a plan can run its stages in place
bool gen_can_ip(const hpxfft_dev::GenPlan &g) { return g.off_r >= 0 || g.off_bp >= 0; }

// the last resort for a one-level length that no other plan covers (a prime
// factor above 61 that no split avoids): Bluestein's algorithm through an
// FFT of more than `cap` and at most 2 cap points, held in one LDS buffer
// and so run in place only
bool gen_setup_in_place_only(size_t n, size_t cap, hpxfft_dev::GenPlan &g)
{
    size_t M = 1;
    while (M < 2 * n - 1) M *= 2;
    if (M <= cap || M > 2 * cap || !gen_factor(M, g)) return false;
    g.len = (int)n;
    g.off_c = (int)M;
    g.off_b = (int)(M + n);
    g.off_r = -1;
    g.off_bp = (int)(2 * M + n);
    g.ip = 1;
    return true;
}


// tables of more than 2^20 entries are split in two of about sqrt(n) entries
constexpr size_t GEN_FULL_TABLE = size_t(1) << 20;
int split_shift(size_t n)
{
    if (n <= GEN_FULL_TABLE) return 0;
    int s = 0;
    while ((size_t(1) << (2 * s)) < n) ++s;
    return s;
}
// complex entries of w_n^m (m < n), whole or split at 2^s
size_t tw_table_size(size_t n, int s) { return s ? (size_t(1) << s) + ((n + (size_t(1) << s) - 1) >> s) : n; }

// longest two-level length, and longest length Bluestein's algorithm takes
// over two levels (its convolution length, a power of two < 2^24, is the
// scratch per sequence)
constexpr size_t GEN2_MAX = size_t(1) << 24;
constexpr size_t GEN2_BLUE_MAX = size_t(1) << 22;

// two-level split N = P * Q (N <= 2^24): both sub-lengths within one
// workgroup, direct plans preferred over Bluestein, then the most balanced
bool gen2_setup(size_t N, size_t cap, gen2_dim &d)
{
    if (N > GEN2_MAX) return false;
    int best = -1;
    double best_score = 0.0;
    for (size_t P = 2; P <= cap && P <= N / 2; ++P)
    {
        if (N % P) continue;
        const size_t Q = N / P;
        if (Q < 2 || Q > cap) continue;
        hpxfft_dev::GenPlan A, B;
        if (!gen_setup(Q, cap, A) || !gen_setup(P, cap, B)) continue;
        const double score = 10.0 * ((A.len != A.n) + (B.len != B.n)) + std::fabs(std::log2((double)P / (double)Q));
        if (best < 0 || score < best_score)
        {
            best = (int)P;
            best_score = score;
            d.A = A;
            d.B = B;
        }
    }
    if (best < 0) return false;
    // the phases hold several short sequences per workgroup, well within the
    // LDS: ping-pong stages (in place measured slower: 10000 x 10000 FP32
    // columns, 100 * 100, 0.69 -> 0.74 ms)
    d.A.ip = d.B.ip = 0;
    d.on = d.avail = true;
    d.P = best;
    d.Q = (int)(N / (size_t)best);
    d.tw_shift = split_shift(N);
    d.blue = false;
    return true;
}

// Bluestein's algorithm over two levels for a length N <= 2^22 that neither
// one workgroup nor a split into two one-workgroup lengths covers: the
// convolution length M (the power of two >= 2N - 1) as M = P*Q, P <= Q,
// both direct power-of-two plans (ping-pong stages)
bool gen2b_setup(size_t N, size_t cap, gen2_dim &d)
{
    if (N < 2 || N > GEN2_BLUE_MAX) return false;
    size_t M = 1;
    while (M < 2 * N - 1) M *= 2;
    size_t P = 1;
    while (P * P < M) P *= 2;
    if (P * P > M) P /= 2;
    const size_t Q = M / P;
    if (Q > cap || !gen_setup(Q, cap, d.A) || !gen_setup(P, cap, d.B)) return false;
    d.A.ip = d.B.ip = 0;
    d.on = d.avail = true;
    d.P = (int)P;
    d.Q = (int)Q;
    d.tw_shift = split_shift(M);
    d.blue = true;
    d.N = N;
    d.chirp_shift = split_shift(2 * N);
    return true;
}

// sequences per workgroup: up to `wmax` (a power of two) within `cap` points
int gen_width(const hpxfft_dev::GenPlan &g, size_t cap, size_t wmax)
{
    size_t w = 1;
    while (w * 2 <= wmax && w * 2 * (size_t)g.n <= cap) w *= 2;
    return (int)w;
}

// table entries (complex FP64) of a general plan: w_n^m (m < n); for
// Bluestein the chirp (len), the kernel spectrum (n) and the same spectrum in
// the in-place stages' pos() order (n); for other in-place plans the n int
// input positions, four per entry
size_t gen_table_size(const hpxfft_dev::GenPlan &g)
{
    if (g.len != g.n) return (size_t)(3 * g.n + g.len);
    return (size_t)g.n + (g.off_r >= 0 ? ((size_t)g.n + 3) / 4 : 0);
}

// input position of element e for in-place stages r_1 .. r_S: the digits of
// e = d_S + r_S (d_{S-1} + ... + r_2 d_1), weighted d_s -> r_1 ... r_{s-1}
int gen_digit_rev(const hpxfft_dev::GenPlan &g, int e)
{
    int pos = 0, w = g.n;
    for (int s = g.nstages - 1; s >= 0; --s)
    {
        const int r = g.radix[s];
        w /= r;
        pos += (e % r) * w;
        e /= r;
    }
    return pos;
}

// the chirp c[j] = w_{2len}^{j^2 mod 2len} (j < len) and DFT_n(b) with
// b[m] = conj c[m], b[n-m] = conj c[m] (0 < m < len), in long double (n a
// power of two >= 2 len - 1; an iterative radix-2 FFT)
void blue_spectrum(uint64_t len, uint64_t n, std::vector<long double> &cr, std::vector<long double> &ci,
                   std::vector<long double> &br, std::vector<long double> &bi)
{
    cr.assign(len, 0.0L);
    ci.assign(len, 0.0L);
    for (uint64_t j = 0; j < len; ++j)
    {
        long double c, sn;
        twiddle_ld((j * j) % (2 * len), 2 * len, c, sn);
        cr[j] = c;
        ci[j] = -sn;
    }
    br.assign(n, 0.0L);
    bi.assign(n, 0.0L);
    for (uint64_t m = 0; m < len; ++m)
    {
        br[m] = cr[m];
        bi[m] = -ci[m];
        if (m > 0)
        {
            br[n - m] = cr[m];
            bi[n - m] = -ci[m];
        }
    }
    // iterative radix-2 DIT FFT (n a power of two), forward sign
    for (uint64_t i = 1, j = 0; i < n; ++i)
    {
        uint64_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j)
        {
            std::swap(br[i], br[j]);
            std::swap(bi[i], bi[j]);
        }
    }
    for (uint64_t half = 1; half < n; half *= 2)
    {
        for (uint64_t k = 0; k < half; ++k)
        {
            long double c, sn;
            twiddle_ld(k, 2 * half, c, sn);
            const long double wr = c, wi = -sn;
            for (uint64_t s0 = 0; s0 < n; s0 += 2 * half)
            {
                const uint64_t a = s0 + k, b = a + half;
                const long double tr = br[b] * wr - bi[b] * wi, ti = br[b] * wi + bi[b] * wr;
                br[b] = br[a] - tr;
                bi[b] = bi[a] - ti;
                br[a] += tr;
                bi[a] += ti;
            }
        }
    }
}

// fill a general plan's table at h[2*off ...] (complex FP64 pairs), from long
// double values: w_n^m exact (octant-reduced), the chirp c[j] = w_{2len}^{j^2
// mod 2len} exact, and Bh = DFT_n(b) / n (blue_spectrum)
void gen_table(const hpxfft_dev::GenPlan &g, std::vector<double> &h, size_t off)
{
    const uint64_t n = (uint64_t)g.n;
    for (uint64_t m = 0; m < n; ++m)
    {
        long double c, sn;
        twiddle_ld(m, n, c, sn);
        h[2 * (off + m)] = (double)c;
        h[2 * (off + m) + 1] = -(double)sn;
    }
    if (g.off_r >= 0)
    {
        std::vector<int> pos(((size_t)n + 3) / 4 * 4, 0);
        for (uint64_t e = 0; e < n; ++e) pos[e] = gen_digit_rev(g, (int)e);
        std::memcpy(&h[2 * (off + (size_t)g.off_r)], pos.data(), sizeof(int) * pos.size());
    }
    if (g.len == g.n) return;
    const uint64_t len = (uint64_t)g.len;
    std::vector<long double> cr, ci, br, bi;
    blue_spectrum(len, n, cr, ci, br, bi);
    for (uint64_t j = 0; j < len; ++j)
    {
        h[2 * (off + n + j)] = (double)cr[j];
        h[2 * (off + n + j) + 1] = (double)ci[j];
    }
    for (uint64_t m = 0; m < n; ++m)
    {
        const size_t pm = (size_t)gen_digit_rev(g, (int)m);
        h[2 * (off + n + len + m)] = h[2 * (off + (size_t)g.off_bp + pm)] = (double)(br[m] / (long double)n);
        h[2 * (off + n + len + m) + 1] = h[2 * (off + (size_t)g.off_bp + pm) + 1] = (double)(bi[m] / (long double)n);
    }
}

std::string radices(const hpxfft_dev::GenPlan &g)
{
    std::string r;
    for (int i = 0; i < g.nstages; ++i) r += (i ? "x" : "") + std::to_string(g.radix[i]);
    return r.empty() ? "none" : r;
}

std::string gen_method(const hpxfft_dev::GenPlan &g)
{
    if (g.len == g.n)
        return std::string(g.ip ? "mixed-radix in-place decimation in time" : "mixed-radix Stockham") +
               " in LDS, radices " + radices(g);
    return "Bluestein (chirp convolution through a " + std::to_string(g.n) + "-point FFT in LDS, radices " +
           radices(g) + (g.ip ? ", in place: decimation in frequency, then in time)" : ", Stockham)");
}

int rows_split_width(const hpxfft_plan_s *p, hpxfft_dev::GenPlan &gB);

std::string describe(const hpxfft_plan_s *p)
{
    std::ostringstream os;
    if (p->batched)
    {
        os << "hpxfft-hip batched plan (gfx950): " << p->bnx.size() << " grids, precision " << prec_name(p->prec)
           << ", effort " << p->effort << "\n";
        for (int ph = 0; ph < 2; ++ph)
        {
            os << (ph == 0 ? "  row phase" : "  column phase") << ": persistent launch, configuration " << p->bcfg[ph]
               << " (" << batched_pts(ph, p->bcfg[ph]) << " points per task, " << p->bntasks[ph][p->bcfg[ph]]
               << " tasks, " << p->bblocks[ph][p->bcfg[ph]] << " workgroups)";
            if (p->effort > 0)
            {
                os << "; autotuned ms:";
                for (int c = 0; c < BATCH_NCFG; ++c) os << " cfg" << c << "=" << p->btune_ms[ph][c];
            }
            os << "\n";
        }
        size_t pts = 0;
        for (size_t i = 0; i < p->bnx.size(); ++i) pts += p->bnx[i] * p->bny[i];
        os << "  total real points " << pts << "\n";
        return os.str();
    }
    os << "hpxfft-hip plan (gfx950): Nx=" << p->nx << " Ny=" << p->ny << " batch=" << p->batch << " precision "
       << prec_name(p->prec) << " effort " << p->effort << "\n";
    if (p->wide16)
    {
        // the FP32 plan's lines, indented, after the conversion steps
        os << "  16-bit storage through an FP32 workspace (a length beyond one workgroup): rows converted in "
              "(times 2^-in_exp), the FP32 plan below in place, column maxima, converted out with per-column "
              "exponents\n";
        std::istringstream sub(describe(p->wide));
        for (std::string ln; std::getline(sub, ln);) os << "    " << ln << "\n";
        return os.str();
    }
    auto two_level = [&](const char *what, const gen2_dim &d) {
        if (d.blue)
        {
            os << "  " << what << " (general length " << d.N << ", Bluestein over two levels): chirp convolution of "
               << "length M = P*Q = " << d.P << " * " << d.Q << ", three passes through " << d.points()
               << " scratch points per sequence\n    pass 1: x c, " << d.Q << "-point " << gen_method(d.A)
               << " over elements m1+P*e, twiddle w_M^(m1 k2), " << d.WA << " sequence(s)/workgroup\n    pass 2: "
               << d.P << "-point " << gen_method(d.B) << ", times DFT_M(b)/M, inverse, twiddle w_M^(-n1 k2), " << d.WB
               << " sequence(s)/workgroup\n    pass 3: inverse " << d.Q << "-point, times c, indices n1+P*n2 < "
               << d.N << "\n" << (d.tw_shift ? "    twiddle tables split in two (FP64 product)\n" : "");
            return;
        }
        os << "  " << what << " (general length, two levels): " << (size_t)d.P * d.Q << " = P*Q = " << d.P << " * "
           << d.Q << "\n    phase A: " << d.Q << "-point " << gen_method(d.A) << " over elements m1+P*e, twiddle w^(m1 k2), "
           << d.WA << " sequence(s)/workgroup\n    phase B: " << d.P << "-point " << gen_method(d.B) << ", "
           << d.WB << " sequence(s)/workgroup\n" << (d.tw_shift ? "    twiddle table split in two (FP64 product)\n" : "");
    };
    if (p->g2r.on)
    {
        two_level("row r2c", p->g2r);
        hpxfft_dev::GenPlan gB;
        if (p->g2r.blue)
            os << "    then the r2c split in place (one thread per pair k, Ny/2-k)\n";
        else if (const int Wh = rows_split_width(p, gB))
            os << "    forward: phase B with the r2c split fused in, " << Wh << " columns k2 <= Q/2 and their "
               << "partners Q-k2 per workgroup, " << gen_method(gB) << "\n";
        else
            os << "    then the r2c split in place (one thread per pair k, Ny/2-k)\n";
    }
    else if (p->gen_row)
        os << "  row r2c kernel (general length): " << p->ny / 2 << "-point " << gen_method(p->grow) << ", "
           << p->gen_rows << " row(s)/workgroup, "
           << gen_threads(gen_lds_bytes(p, p->grow, (size_t)p->gen_rows * p->grow.n, (size_t)p->gen_rows), p->grow)
           << " threads + r2c split\n";
    else if (p->row_kind && row_pair_covers(p->prec, p->lgL))
        os << "  row r2c kernel: " << (size_t(1) << p->lgL) << "-point complex Stockham per row, radix-32 stages "
           << "split across lane pairs (DPP exchange), radix " << ((size_t(1) << p->lgL) / 1024)
           << " fused with the r2c split, " << ((size_t(1) << p->lgL) / 32) << " threads\n";
    else
    {
        const geom rg = row_geom(p->prec, p->lgL);
        os << "  row r2c kernel: " << (size_t(1) << p->lgL) << "-point complex Stockham per row + r2c split, "
           << rg.nseq << " row(s)/workgroup, " << rg.nt << " threads, " << rg.E << " elements/thread, radix "
           << (rg.E < 16 ? rg.E : 16) << (rg.split ? ", LDS exchange in re/im halves" : "") << "\n";
    }
    if (p->row_ms[0] > 0.0)
        os << "  row kernel autotuned (ms): radix-16 stages " << p->row_ms[0] << ", lane-pair radix-32 stages "
           << p->row_ms[1] << "\n";
    for (int d = 0; d < 2; ++d)
        if (p->gip_ms[d][0] > 0.0)
            os << "  general " << (d ? "column" : "row") << " pass autotuned (ms): ping-pong " << p->gip_ms[d][0]
               << ", in place " << p->gip_ms[d][1] << "\n";
    if (p->gcol_ms[0] > 0.0)
        os << "  general column pass autotuned (ms): one level " << p->gcol_ms[0] << ", two levels " << p->gcol_ms[1]
           << "\n";
    if (p->g2c.on) two_level("column c2c", p->g2c);
    else if (p->gen_col)
        os << "  column c2c kernel (general length): " << p->nx << "-point " << gen_method(p->gcol) << ", "
           << p->gen_w << " adjacent columns/workgroup\n";
    else if (!p->four_step)
    {
        const geom cg = col_geom(p->prec, p->lgx);
        os << "  column c2c kernel: " << p->nx << "-point, one tile of " << cg.nseq << " adjacent columns x all rows, "
           << cg.nt << " threads, transposes fused into row-segment loads/stores"
           << (cg.split ? ", LDS exchange in re/im halves" : "") << "\n";
    }
    else
    {
        const geom ga = tile_geom(p->prec, p->lgQ), gb = tile_geom(p->prec, p->lgP);
        os << "  column c2c four-step Nx = P*Q = " << (1 << p->lgP) << " * " << (1 << p->lgQ) << " (scratch "
           << p->scratch_bytes << " B, one block per strip)\n";
        os << "    phase A: " << (1 << p->lgQ) << "-point DFTs over rows m1+P*m2, twiddle w_Nx^(m1 k2), tiles of "
           << ga.nseq << " columns, " << ga.nt << " threads\n";
        os << "    phase B: " << (1 << p->lgP) << "-point DFTs, to rows Q*k1+k2, tiles of " << gb.nseq << " columns, "
           << gb.nt << " threads\n";
        bool tuned = false;
        for (int l = 0; l <= COL_TILE_MAX_LG; ++l) tuned = tuned || p->split_ms[l] > 0.0;
        if (tuned)
        {
            os << "    split autotuned (column pass ms):";
            for (int l = 0; l <= COL_TILE_MAX_LG; ++l)
                if (p->split_ms[l] > 0.0) os << " P=" << (1 << l) << ": " << p->split_ms[l];
            os << "\n";
        }
    }
    const double sr = (double)storage_bytes(p->prec);
    const double H = (double)p->nx * (double)p->ncols * 2.0 * sr;
    const double B = ((double)p->nx * p->ny * sr + H) + 2.0 * H;
    os << "  algorithmic bytes per grid (2-pass model): " << (long long)B << "\n";
    os << "  plan_flops (2.5 Ny log2 Ny per row + 5 Nx log2 Nx per column): " << p->plan_flops << "\n";
    return os.str();
}

// a HIP event destroyed when it leaves scope
struct scoped_event
{
    hipEvent_t ev = nullptr;
    scoped_event() = default;
    scoped_event(const scoped_event &) = delete;
    scoped_event &operator=(const scoped_event &) = delete;
    ~scoped_event()
    {
        if (ev) (void)hipEventDestroy(ev);
    }
};

double elapsed(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
    return (double)ms * 1e-3;
}

int alloc_events(hpxfft_plan_s *p)
{
    for (auto &e : p->ev)
        if (!e) HIPCHK(hipEventCreate(&e));
    return HPXFFT_OK;
}

// one phase of a two-level general dimension (gen_strided_kernel)
int launch_strided(hpxfft_plan_s *p, const hpxfft_dev::GenStride &a, const hpxfft_dev::GenPlan &g, size_t tab_off,
                   size_t twt_off, int W, unsigned gx, unsigned gy, unsigned gz)
{
    const auto *tab = reinterpret_cast<const hpxfft_dev::cpx<double> *>(p->gen_tab);
    const size_t lds = gen_lds_bytes(p, g, (size_t)W * g.n);
    const dim3 grid(gx, gy, gz);
    const unsigned nt = gen_threads(lds, g);
    // the general form only where a split table or a Bluestein pass needs it
    const bool general = a.tw_shift || a.tw2_shift || a.blue_in || a.blue_out;
    if (p->prec == P_F64 && general)
        GEN_LAUNCH(hpxfft_dev::gen_strided_kernel, (double, true), nt, grid, lds, p->stream, a, g, tab + tab_off,
                   tab + twt_off, tab + p->gtab_2, W);
    else if (p->prec == P_F64)
        GEN_LAUNCH(hpxfft_dev::gen_strided_kernel, (double, false), nt, grid, lds, p->stream, a, g, tab + tab_off,
                   tab + twt_off, tab + p->gtab_2, W);
    else if (general)
        GEN_LAUNCH(hpxfft_dev::gen_strided_kernel, (float, true), nt, grid, lds, p->stream, a, g, tab + tab_off,
                   tab + twt_off, tab + p->gtab_2, W);
    else
        GEN_LAUNCH(hpxfft_dev::gen_strided_kernel, (float, false), nt, grid, lds, p->stream, a, g, tab + tab_off,
                   tab + twt_off, tab + p->gtab_2, W);
    HIPCHK(hipGetLastError());
    return HPXFFT_OK;
}

unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// Bluestein pass 2 of a two-level dimension (gen_blue_mid_kernel), in place in `a`
int launch_blue_mid(hpxfft_plan_s *p, hpxfft_dev::GenStride a, const gen2_dim &d, int W, unsigned gx, unsigned gy,
                    unsigned gz)
{
    const auto *tab = reinterpret_cast<const hpxfft_dev::cpx<double> *>(p->gen_tab);
    const size_t lds = gen_lds_bytes(p, d.B, (size_t)W * d.B.n);
    const dim3 grid(gx, gy, gz);
    const unsigned nt = gen_threads(lds, d.B);
    a.tw_shift = d.tw_shift;
    if (p->prec == P_F64)
        GEN_LAUNCH(hpxfft_dev::gen_blue_mid_kernel, (double), nt, grid, lds, p->stream, a, d.B, tab + d.tabB,
                   tab + d.tabBh, tab + d.tabT, W, d.Q);
    else
        GEN_LAUNCH(hpxfft_dev::gen_blue_mid_kernel, (float), nt, grid, lds, p->stream, a, d.B, tab + d.tabB,
                   tab + d.tabBh, tab + d.tabT, W, d.Q);
    HIPCHK(hipGetLastError());
    return HPXFFT_OK;
}

// the Bluestein fields of a pass-1 / pass-3 GenStride
void blue_fields(const hpxfft_plan_s *p, const gen2_dim &d, hpxfft_dev::GenStride &a)
{
    a.blue_n = (int)d.N;
    a.blue_p = d.P;
    a.chirp_shift = d.chirp_shift;
    a.chirp = reinterpret_cast<const hpxfft_dev::cpx<double> *>(p->gen_tab) + d.tabC;
}

// two-level Bluestein column pass (Nx = N, M = P*Q) from spec_in to spec_out
// through the plan's scratch [m1][k2][col] of M rows; conj: the inverse
int run_columns_blue(hpxfft_plan_s *p, const void *spec_in, void *spec_out, bool conj)
{
    const gen2_dim &d = p->g2c;
    const long long P = d.P, Q = d.Q, M = (long long)d.points(), nc = (long long)p->ncols, ld = (long long)p->ld_out;
    const unsigned bz = (unsigned)p->batch;
    hpxfft_dev::GenStride a{};
    a.in = spec_in;
    a.in_seq = P * ld;
    a.in_col = 1;
    a.in_sub = ld;
    a.in_bat = (long long)p->dist_out;
    a.out = p->gscr;
    a.out_seq = nc;
    a.out_col = 1;
    a.out_sub = Q * nc;
    a.out_bat = M * nc;
    a.ncols = (int)nc;
    a.tw_n = (int)M;
    a.tw_shift = d.tw_shift;
    a.conj_in = conj;
    a.blue_in = 1;
    blue_fields(p, d, a);
    if (int rc = launch_strided(p, a, d.A, d.tabA, d.tabT, d.WA, cdiv(nc, d.WA), (unsigned)P, bz)) return rc;
    hpxfft_dev::GenStride m{};
    m.in = p->gscr;
    m.out = p->gscr;
    m.in_seq = m.out_seq = Q * nc;
    m.in_col = m.out_col = 1;
    m.in_sub = m.out_sub = nc;
    m.in_bat = m.out_bat = M * nc;
    m.ncols = (int)nc;
    if (int rc = launch_blue_mid(p, m, d, d.WB, cdiv(nc, d.WB), (unsigned)Q, bz)) return rc;
    hpxfft_dev::GenStride b{};
    b.in = p->gscr;
    b.in_seq = nc;
    b.in_col = 1;
    b.in_sub = Q * nc;
    b.in_bat = M * nc;
    b.out = spec_out;
    b.out_seq = P * ld;
    b.out_col = 1;
    b.out_sub = ld;
    b.out_bat = (long long)p->dist_out;
    b.ncols = (int)nc;
    b.conj_in = b.conj_out = 1;
    b.blue_out = 1;
    b.blue_conj = conj;
    blue_fields(p, d, b);
    return launch_strided(p, b, d.A, d.tabA, d.tabT, d.WA, cdiv(nc, d.WA), (unsigned)P, bz);
}

// two-level Bluestein rows (Ny/2 = N, M = P*Q) through the plan's scratch
// [row][m1][k2] of M points per row.  Forward: the real rows as complex
// points -> the spectrum rows, then the r2c split in place.  Backward (c2r,
// halve for the adjoint): the c2r sequence of each half-spectrum row -> `real`.
int run_rows_blue(hpxfft_plan_s *p, const void *real_in, void *real_out, const void *spec_in, void *spec_out,
                  bool backward, bool halve)
{
    const gen2_dim &d = p->g2r;
    const long long P = d.P, Q = d.Q, M = (long long)d.points(), L = (long long)d.N, nx = (long long)p->nx;
    const unsigned by = (unsigned)p->nx, bz = (unsigned)p->batch;
    hpxfft_dev::GenStride a{};
    a.in_seq = P;
    a.in_col = 1;
    if (!backward)
    {
        a.in = real_in;
        a.in_sub = (long long)p->ld_in / 2;
        a.in_bat = (long long)p->dist_in / 2;
    }
    else
    {
        a.in = spec_in;
        a.in_sub = (long long)p->ld_out;
        a.in_bat = (long long)p->dist_out;
        a.pre = halve ? 2 : 1;
        a.prel = (int)L;
        a.conj_in = 1;
        a.tw2_shift = p->tw2_shift;
    }
    a.out = p->gtmp;
    a.out_seq = 1;
    a.out_col = Q;
    a.out_sub = M;
    a.out_bat = nx * M;
    a.ncols = (int)P;
    a.tw_n = (int)M;
    a.tw_col = 1;
    a.tw_shift = d.tw_shift;
    a.blue_in = 1;
    blue_fields(p, d, a);
    if (int rc = launch_strided(p, a, d.A, d.tabA, d.tabT, d.WA, cdiv(P, d.WA), by, bz)) return rc;
    hpxfft_dev::GenStride m{};
    m.in = p->gtmp;
    m.out = p->gtmp;
    m.in_seq = m.out_seq = Q;
    m.in_col = m.out_col = 1;
    m.in_sub = m.out_sub = M;
    m.in_bat = m.out_bat = nx * M;
    m.ncols = (int)Q;
    m.tw_col = 1;
    if (int rc = launch_blue_mid(p, m, d, d.WB, cdiv(Q, d.WB), by, bz)) return rc;
    hpxfft_dev::GenStride b{};
    b.in = p->gtmp;
    b.in_seq = 1;
    b.in_col = Q;
    b.in_sub = M;
    b.in_bat = nx * M;
    b.out_seq = P;
    b.out_col = 1;
    b.ncols = (int)P;
    b.tw_col = 1;
    b.conj_in = b.conj_out = 1;
    b.blue_out = 1;
    blue_fields(p, d, b);
    if (!backward)
    {
        b.out = spec_out;
        b.out_sub = (long long)p->ld_out;
        b.out_bat = (long long)p->dist_out;
    }
    else
    {
        b.out = real_out;
        b.out_sub = (long long)p->ld_in / 2;
        b.out_bat = (long long)p->dist_in / 2;
        b.blue_conj = 1;
    }
    if (int rc = launch_strided(p, b, d.A, d.tabA, d.tabT, d.WA, cdiv(P, d.WA), by, bz)) return rc;
    if (!backward)
    {
        const auto *tab = reinterpret_cast<const hpxfft_dev::cpx<double> *>(p->gen_tab);
        const dim3 grid(cdiv((size_t)L / 2 + 1, 256), by, bz);
        if (p->prec == P_F64)
            hipLaunchKernelGGL(hpxfft_dev::gen_split_kernel<double>, grid, dim3(256), 0, p->stream, spec_out,
                               (long long)p->ld_out, (long long)p->dist_out, (int)L, tab + p->gtab_2, p->tw2_shift);
        else
            hipLaunchKernelGGL(hpxfft_dev::gen_split_kernel<float>, grid, dim3(256), 0, p->stream, spec_out,
                               (long long)p->ld_out, (long long)p->dist_out, (int)L, tab + p->gtab_2, p->tw2_shift);
        HIPCHK(hipGetLastError());
    }
    return HPXFFT_OK;
}

// two-level column pass (Nx = P*Q) from spec_in to spec_out through the
// plan's scratch [m1][k2][col]; conj: the inverse transform
int run_columns_2level(hpxfft_plan_s *p, const void *spec_in, void *spec_out, bool conj)
{
    if (p->g2c.blue) return run_columns_blue(p, spec_in, spec_out, conj);
    const gen2_dim &d = p->g2c;
    const long long P = d.P, Q = d.Q, nc = (long long)p->ncols, ld = (long long)p->ld_out;
    hpxfft_dev::GenStride a{};
    a.in = spec_in;
    a.in_seq = P * ld;
    a.in_col = 1;
    a.in_sub = ld;
    a.in_bat = (long long)p->dist_out;
    a.out = p->gscr;
    a.out_seq = nc;
    a.out_col = 1;
    a.out_sub = Q * nc;
    a.out_bat = (long long)p->nx * nc;
    a.ncols = (int)nc;
    a.tw_n = (int)p->nx;
    a.tw_shift = d.tw_shift;
    a.conj_in = conj;
    if (int rc = launch_strided(p, a, d.A, d.tabA, d.tabT, d.WA, cdiv(nc, d.WA), (unsigned)P, (unsigned)p->batch))
        return rc;
    hpxfft_dev::GenStride b{};
    b.in = p->gscr;
    b.in_seq = Q * nc;
    b.in_col = 1;
    b.in_sub = nc;
    b.in_bat = (long long)p->nx * nc;
    b.out = spec_out;
    b.out_seq = Q * ld;
    b.out_col = 1;
    b.out_sub = ld;
    b.out_bat = (long long)p->dist_out;
    b.ncols = (int)nc;
    b.conj_out = conj;
    return launch_strided(p, b, d.B, d.tabB, d.tabT, d.WB, cdiv(nc, d.WB), (unsigned)Q, (unsigned)p->batch);
}

// forward two-level rows: phase B with the r2c split fused in, each
// workgroup taking Wh columns k2 <= Q/2 and their partners Q - k2 (twice the
// sequences of the plain phase B), in place where the plan allows (gB), with
// Wh halved until the block fits the LDS; 0 if it never does (then the plain
// phase B and a split pass run)
int rows_split_width(const hpxfft_plan_s *p, hpxfft_dev::GenPlan &gB)
{
    gB = p->g2r.B;
    gB.ip = gen_can_ip(gB) ? 1 : 0;
    int Wh = p->g2r.WB;
    while (Wh > 1 && gen_lds_bytes(p, gB, (size_t)2 * Wh * gB.n) > (size_t)hpxfft_dev::GEN_LDS_BYTES) Wh /= 2;
    return gen_lds_bytes(p, gB, (size_t)2 * Wh * gB.n) <= (size_t)hpxfft_dev::GEN_LDS_BYTES ? Wh : 0;
}

// two-level rows (Ny/2 = L = P*Q).  Forward: the real rows of `real` as
// complex points -> plan scratch -> the spectrum rows of `spec`, then the
// r2c split in place.  Backward (c2r, halve for the adjoint): the c2r
// sequence of each half-spectrum row of `spec` -> scratch -> `real`.
int run_rows_2level(hpxfft_plan_s *p, const void *real_in, void *real_out, const void *spec_in, void *spec_out,
                    bool backward, bool halve)
{
    if (p->g2r.blue) return run_rows_blue(p, real_in, real_out, spec_in, spec_out, backward, halve);
    const gen2_dim &d = p->g2r;
    const long long P = d.P, Q = d.Q, L = P * Q;
    hpxfft_dev::GenStride a{};
    a.in_seq = P;
    a.in_col = 1;
    if (!backward)
    {
        a.in = real_in;
        a.in_sub = (long long)p->ld_in / 2;
        a.in_bat = (long long)p->dist_in / 2;
    }
    else
    {
        a.in = spec_in;
        a.in_sub = (long long)p->ld_out;
        a.in_bat = (long long)p->dist_out;
        a.pre = halve ? 2 : 1;
        a.prel = (int)L;
        a.conj_in = 1;
        a.tw2_shift = p->tw2_shift;
    }
    a.out = p->gtmp;
    a.out_seq = 1;
    a.out_col = Q;
    a.out_sub = L;
    a.out_bat = (long long)p->nx * L;
    a.ncols = (int)P;
    a.tw_n = (int)L;
    a.tw_col = 1;
    a.tw_shift = d.tw_shift;
    if (int rc = launch_strided(p, a, d.A, d.tabA, d.tabT, d.WA, cdiv(P, d.WA), (unsigned)p->nx, (unsigned)p->batch))
        return rc;
    hpxfft_dev::GenPlan gB;
    const int Wh = rows_split_width(p, gB);
    if (!backward && Wh > 0)
    {
        hpxfft_dev::GenStride b{};
        b.in = p->gtmp;
        b.in_seq = Q;
        b.in_col = 1;
        b.in_sub = L;
        b.in_bat = (long long)p->nx * L;
        b.out = spec_out;
        b.out_sub = (long long)p->ld_out;
        b.out_bat = (long long)p->dist_out;
        b.tw2_shift = p->tw2_shift;
        const size_t lds = gen_lds_bytes(p, gB, (size_t)2 * Wh * gB.n);
        const unsigned nt = gen_threads(lds, gB);
        const dim3 grid(cdiv((size_t)Q / 2 + 1, (size_t)Wh), (unsigned)p->nx, (unsigned)p->batch);
        const auto *tab = reinterpret_cast<const hpxfft_dev::cpx<double> *>(p->gen_tab);
        if (p->prec == P_F64 && b.tw2_shift)
            GEN_LAUNCH(hpxfft_dev::gen_rows_split_kernel, (double, true), nt, grid, lds, p->stream, b, gB,
                       tab + d.tabB, tab + p->gtab_2, Wh, (int)Q);
        else if (p->prec == P_F64)
            GEN_LAUNCH(hpxfft_dev::gen_rows_split_kernel, (double, false), nt, grid, lds, p->stream, b, gB,
                       tab + d.tabB, tab + p->gtab_2, Wh, (int)Q);
        else if (b.tw2_shift)
            GEN_LAUNCH(hpxfft_dev::gen_rows_split_kernel, (float, true), nt, grid, lds, p->stream, b, gB,
                       tab + d.tabB, tab + p->gtab_2, Wh, (int)Q);
        else
            GEN_LAUNCH(hpxfft_dev::gen_rows_split_kernel, (float, false), nt, grid, lds, p->stream, b, gB,
                       tab + d.tabB, tab + p->gtab_2, Wh, (int)Q);
        HIPCHK(hipGetLastError());
        return HPXFFT_OK;
    }
    hpxfft_dev::GenStride b{};
    b.in = p->gtmp;
    b.in_seq = Q;
    b.in_col = 1;
    b.in_sub = L;
    b.in_bat = (long long)p->nx * L;
    b.out_seq = Q;
    b.out_col = 1;
    b.ncols = (int)Q;
    if (!backward)
    {
        b.out = spec_out;
        b.out_sub = (long long)p->ld_out;
        b.out_bat = (long long)p->dist_out;
    }
    else
    {
        b.out = real_out;
        b.out_sub = (long long)p->ld_in / 2;
        b.out_bat = (long long)p->dist_in / 2;
        b.conj_out = 1;
    }
    if (int rc = launch_strided(p, b, d.B, d.tabB, d.tabT, d.WB, cdiv(Q, d.WB), (unsigned)p->nx, (unsigned)p->batch))
        return rc;
    if (!backward)
    {
        const auto *tab = reinterpret_cast<const hpxfft_dev::cpx<double> *>(p->gen_tab);
        const dim3 grid(cdiv((size_t)L / 2 + 1, 256), (unsigned)p->nx, (unsigned)p->batch);
        if (p->prec == P_F64)
            hipLaunchKernelGGL(hpxfft_dev::gen_split_kernel<double>, grid, dim3(256), 0, p->stream, spec_out,
                               (long long)p->ld_out, (long long)p->dist_out, (int)L, tab + p->gtab_2, p->tw2_shift);
        else
            hipLaunchKernelGGL(hpxfft_dev::gen_split_kernel<float>, grid, dim3(256), 0, p->stream, spec_out,
                               (long long)p->ld_out, (long long)p->dist_out, (int)L, tab + p->gtab_2, p->tw2_shift);
        HIPCHK(hipGetLastError());
    }
    return HPXFFT_OK;
}

// column pass on the spectrum `spec` (in place), or from `spec_in` to
// `spec_out` for the adjoint; `conj` = inverse transform.
int run_columns(hpxfft_plan_s *p, const void *spec_in, void *spec_out, bool conj, int *col_exp,
                hipEvent_t mid = nullptr)
{
    const size_t sb = storage_bytes(p->prec);
    const int W = p->four_step ? tile_geom(p->prec, p->lgQ).nseq : p->W;
    ColArgs a{};
    a.pitch = (long long)p->ld_out;
    a.dist = (long long)p->dist_out;
    a.ncols = (int)p->ncols;
    a.row_exp = p->mid_exp;
    a.out_exp = col_exp;
    a.rexp_dist = (long long)p->nx;
    a.cexp_dist = (long long)p->ncols;
    if (p->g2c.on) return run_columns_2level(p, spec_in, spec_out, conj);
    if (p->gen_col)
    {
        // general length (F64 / F32; in place for the forward transform)
        a.in = spec_in;
        a.out = spec_out;
        const dim3 grid((unsigned)((p->ncols + p->gen_w - 1) / p->gen_w), (unsigned)p->batch);
        const size_t lds = gen_lds_bytes(p, p->gcol, (size_t)p->gen_w * p->gcol.n, (size_t)p->gen_w);
        const unsigned nt = gen_threads(lds, p->gcol);
        const auto *tab = reinterpret_cast<const hpxfft_dev::cpx<double> *>(p->gen_tab) + p->gtab_N;
        using namespace hpxfft_dev;
        const GenPlan &g = p->gcol;
        const int W = p->gen_w;
        switch (p->prec)
        {
        case P_F64:
            if (conj) GEN_LAUNCH(gen_col_kernel, (double, double, true), nt, grid, lds, p->stream, a, g, tab, W);
            else GEN_LAUNCH(gen_col_kernel, (double, double, false), nt, grid, lds, p->stream, a, g, tab, W);
            break;
        case P_F32:
            if (conj) GEN_LAUNCH(gen_col_kernel, (float, float, true), nt, grid, lds, p->stream, a, g, tab, W);
            else GEN_LAUNCH(gen_col_kernel, (float, float, false), nt, grid, lds, p->stream, a, g, tab, W);
            break;
        case P_F16:
            if (conj) return fail(HPXFFT_INVALID_ARG, "16-bit storage plans are forward only");
            GEN_LAUNCH(gen_col_kernel, (float, __half, false), nt, grid, lds, p->stream, a, g, tab, W);
            break;
        default:
            if (conj) return fail(HPXFFT_INVALID_ARG, "16-bit storage plans are forward only");
            GEN_LAUNCH(gen_col_kernel, (float, bf16_t, false), nt, grid, lds, p->stream, a, g, tab, W);
            break;
        }
        HIPCHK(hipGetLastError());
        return HPXFFT_OK;
    }
    if (!p->four_step)
    {
        a.in = spec_in;
        a.out = spec_out;
        a.P = 1;
        a.Q = (int)p->nx;
        a.tw_L = tw_ptr(p, p->nx);
        a.tw_N = nullptr;
        const unsigned gx = (unsigned)((p->ncols + W - 1) / W);
        HIPCHK(col(p->prec, p->lgx, 0, conj, a, gx, 1, (unsigned)p->batch, p->stream));
        return HPXFFT_OK;
    }
    const int P = 1 << p->lgP, Q = 1 << p->lgQ;
    a.P = P;
    a.Q = Q;
    a.tw_N = tw_ptr(p, p->nx);
    a.tw64_N = tw64_ptr(p, p->nx);
    a.colpart = p->colpart;
    a.colmax = p->colmax;
    a.sdist = (long long)(p->nx * p->CW);
    const unsigned gx = (unsigned)((p->ncols + W - 1) / W), nb = (unsigned)p->batch;
    // phase A: spectrum -> scratch (one scratch slice per grid: all grids in one launch)
    a.in = spec_in;
    a.out = p->scratch;
    a.tw_L = tw_ptr(p, (size_t)Q);
    a.tw64_L = tw64_ptr(p, (size_t)Q);
    HIPCHK(col(p->prec, p->lgQ, 1, conj, a, gx, (unsigned)P, nb, p->stream));
    if (p->colpart)
    {
        // 16-bit: fold phase A's per-workgroup maxima into one per column (every entry is rewritten)
        hipLaunchKernelGGL(hpxfft_dev::colmax_reduce_kernel<0>, dim3((unsigned)((p->ncols + 63) / 64), nb), dim3(256),
                           0, p->stream, p->colpart, p->colmax, (int)p->ncols, P);
        HIPCHK(hipGetLastError());
    }
    if (mid) HIPCHK(hipEventRecord(mid, p->stream));
    // phase B: scratch -> spectrum; its tiles are shifted to whole 128-byte
    // lines of the output rows, which can take one more tile per row
    const int eline = (int)(128 / (2 * sb));
    a.in = p->scratch;
    a.out = spec_out;
    a.out_line_off = (int)((reinterpret_cast<uintptr_t>(spec_out) / (2 * sb)) % (uintptr_t)eline);
    a.tw_L = tw_ptr(p, (size_t)P);
    a.tw64_L = tw64_ptr(p, (size_t)P);
    const unsigned gxb = (unsigned)((p->ncols + eline - 2) / W + 1);
    HIPCHK(col(p->prec, p->lgP, 2, conj, a, gxb, (unsigned)Q, nb, p->stream));
    return HPXFFT_OK;
}

// the kernels move whole complex elements (real rows as complex pairs), so
// both device pointers must be aligned to one complex element of the storage
int check_aligned(const hpxfft_plan_s *p, const void *a, const void *b)
{
    const size_t al = 2 * storage_bytes(p->prec);
    if ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) % al)
        return fail(HPXFFT_INVALID_ARG, "device pointers must be aligned to one complex element (" +
                                            std::to_string(al) + " bytes) of the plan's storage");
    return HPXFFT_OK;
}

// the row pass of power-of-two rows with row kernel `kind` (0: radix-16 stages;
// 1: lane-pair radix-32 stages, where the plan's rows allow it)
hipError_t launch_rows_pow2(const hpxfft_plan_s *p, const RowArgs &r, int kind)
{
    if (kind && row_pair_covers(p->prec, p->lgL))
        return row_r2c_pair(p->prec, p->lgL, r, (unsigned)(p->nx * p->batch), p->stream);
    const geom rg = row_geom(p->prec, p->lgL);
    const unsigned gx = (unsigned)((p->nx + rg.nseq - 1) / rg.nseq);
    return row_r2c(p->prec, p->lgL, r, gx, (unsigned)p->batch, p->stream);
}

int exec_wide16(hpxfft_plan_s *p, const void *d_in, void *d_out, const int *in_exp, int *col_exp);

int exec_forward(hpxfft_plan_s *p, const void *d_in, void *d_out, const int *in_exp, int *col_exp)
{
    if (!d_in || !d_out) return fail(HPXFFT_INVALID_ARG, "null device pointer");
    if (int rc = check_aligned(p, d_in, d_out)) return rc;
    if (d_in == d_out && (p->ld_in != 2 * p->ld_out || p->dist_in != 2 * p->dist_out))
        return fail(HPXFFT_INVALID_ARG, "in-place exec needs ld_in == 2*ld_out and dist_in == 2*dist_out");
    const bool s16 = p->prec == P_F16 || p->prec == P_BF16;
    if (s16 && !col_exp) col_exp = p->col_exp_own;
    if (p->wide16) return exec_wide16(p, d_in, d_out, in_exp, col_exp);
    device_guard g(p->device);
    if (p->timed)
    {
        if (int rc = alloc_events(p)) return rc;
        HIPCHK(hipEventRecord(p->ev[0], p->stream));
    }
    RowArgs r{};
    r.in = d_in;
    r.out = d_out;
    r.ld_in = (long long)p->ld_in;
    r.ld_out = (long long)(2 * p->ld_out);
    r.dist_in = (long long)p->dist_in;
    r.dist_out = (long long)(2 * p->dist_out);
    r.nrows = (int)p->nx;
    r.tw = tw_ptr(p, p->ny);
    r.tw64 = tw64_ptr(p, p->ny);
    r.stw64 = p->tabs.stw64;
    r.stw32 = p->tabs.stw32;
    r.in_exp = s16 ? in_exp : nullptr;
    r.out_exp = s16 ? p->mid_exp : nullptr;
    r.exp_dist = (long long)p->nx;
    if (p->g2r.on)
    {
        if (int rc = run_rows_2level(p, d_in, nullptr, nullptr, d_out, false, false)) return rc;
    }
    else if (p->gen_row)
    {
        const int ns = p->gen_rows;
        const size_t lds = gen_lds_bytes(p, p->grow, (size_t)ns * p->grow.n, (size_t)ns);
        const dim3 grid((unsigned)((p->nx + ns - 1) / ns), (unsigned)p->batch);
        const unsigned nt = gen_threads(lds, p->grow);
        const auto *tab = reinterpret_cast<const hpxfft_dev::cpx<double> *>(p->gen_tab);
        const auto *tL = tab + p->gtab_L, *t2 = tab + p->gtab_2;
        using namespace hpxfft_dev;
        switch (p->prec)
        {
        case P_F64: GEN_LAUNCH(gen_row_r2c_kernel, (double, double), nt, grid, lds, p->stream, r, p->grow, tL, t2, ns); break;
        case P_F32: GEN_LAUNCH(gen_row_r2c_kernel, (float, float), nt, grid, lds, p->stream, r, p->grow, tL, t2, ns); break;
        case P_F16: GEN_LAUNCH(gen_row_r2c_kernel, (float, __half), nt, grid, lds, p->stream, r, p->grow, tL, t2, ns); break;
        default: GEN_LAUNCH(gen_row_r2c_kernel, (float, bf16_t), nt, grid, lds, p->stream, r, p->grow, tL, t2, ns); break;
        }
        HIPCHK(hipGetLastError());
    }
    else
    {
        HIPCHK(launch_rows_pow2(p, r, p->row_kind));
    }
    if (p->timed) HIPCHK(hipEventRecord(p->ev[1], p->stream));
    if (int rc = run_columns(p, d_out, d_out, false, s16 ? col_exp : nullptr, p->timed ? p->ev[3] : nullptr))
        return rc;
    if (p->timed) HIPCHK(hipEventRecord(p->ev[2], p->stream));
    p->ran = p->timed;
    p->col_split = p->four_step;
    p->col_first = false;
    p->h2d = p->d2h = 0.0;
    return HPXFFT_OK;
}

// 16-bit storage through the FP32 workspace: convert in (unscaling rows by
// in_exp), the FP32 plan in place, the column maxima, convert out (scaling
// columns, exponents to col_exp).  Timing: the FP32 plan records the row /
// column boundaries into this plan's events, so first_fftw covers the
// conversion in and the rows, second_fftw the columns and the conversion out.
int exec_wide16(hpxfft_plan_s *p, const void *d_in, void *d_out, const int *in_exp, int *col_exp)
{
    hpxfft_plan_s *w = p->wide;
    const long long nc = (long long)p->ncols, ldw = 2 * nc, distw = (long long)p->nx * ldw;
    const dim3 gin(cdiv(p->ny / 2, 256), (unsigned)p->nx, (unsigned)p->batch);
    const dim3 gcol(cdiv(p->ncols, 256), (unsigned)p->nx, (unsigned)p->batch);
    if (p->timed)
    {
        if (int rc = alloc_events(p)) return rc;
        HIPCHK(hipEventRecord(p->ev[0], p->stream));
    }
    using namespace hpxfft_dev;
    if (p->prec == P_F16)
        hipLaunchKernelGGL(wide_in_kernel<__half>, gin, dim3(256), 0, p->stream, static_cast<const __half *>(d_in),
                           (long long)p->ld_in, (long long)p->dist_in, in_exp, (long long)p->nx, p->wide_ws, ldw,
                           distw, (int)(p->ny / 2));
    else
        hipLaunchKernelGGL(wide_in_kernel<bf16_t>, gin, dim3(256), 0, p->stream, static_cast<const bf16_t *>(d_in),
                           (long long)p->ld_in, (long long)p->dist_in, in_exp, (long long)p->nx, p->wide_ws, ldw,
                           distw, (int)(p->ny / 2));
    HIPCHK(hipGetLastError());
    // the FP32 plan records its phase boundaries into this plan's events
    hipEvent_t own[3] = {w->ev[1], w->ev[2], w->ev[3]};
    const bool wt = w->timed;
    if (p->timed)
    {
        w->ev[1] = p->ev[1];
        w->ev[2] = p->ev[2];
        w->ev[3] = p->ev[3];
    }
    w->timed = p->timed;
    const int rc = exec_forward(w, p->wide_ws, p->wide_ws, nullptr, nullptr);
    w->ev[1] = own[0];
    w->ev[2] = own[1];
    w->ev[3] = own[2];
    w->timed = wt;
    if (rc) return rc;
    const long long ncm = (long long)p->ncols * (long long)p->batch;
    hipLaunchKernelGGL(zero_u32_kernel<0>, dim3(cdiv((size_t)ncm, 256)), dim3(256), 0, p->stream, p->colmax, ncm);
    constexpr int ROWS = 64;
    const dim3 gmax(cdiv(p->ncols, 256), cdiv(p->nx, ROWS), (unsigned)p->batch);
    hipLaunchKernelGGL(wide_colmax_kernel, gmax, dim3(256), 0, p->stream, p->wide_ws, ldw, distw, (int)p->nx,
                       (int)p->ncols, ROWS, p->colmax);
    if (p->prec == P_F16)
        hipLaunchKernelGGL(wide_out_kernel<__half>, gcol, dim3(256), 0, p->stream, p->wide_ws, ldw, distw,
                           static_cast<__half *>(d_out), (long long)(2 * p->ld_out), (long long)(2 * p->dist_out),
                           (int)p->ncols, p->colmax, col_exp, (long long)p->ncols);
    else
        hipLaunchKernelGGL(wide_out_kernel<bf16_t>, gcol, dim3(256), 0, p->stream, p->wide_ws, ldw, distw,
                           static_cast<bf16_t *>(d_out), (long long)(2 * p->ld_out), (long long)(2 * p->dist_out),
                           (int)p->ncols, p->colmax, col_exp, (long long)p->ncols);
    HIPCHK(hipGetLastError());
    if (p->timed) HIPCHK(hipEventRecord(p->ev[2], p->stream));
    p->ran = p->timed;
    p->col_split = w->four_step;
    p->col_first = false;
    p->h2d = p->d2h = 0.0;
    return HPXFFT_OK;
}

// timed repetitions per candidate of the planner's searches: measure 3,
// patient 5, exhaustive 10 (more repetitions, less timing noise in the choice)
int tune_reps(int effort) { return effort >= HPXFFT_EXHAUSTIVE ? 10 : effort >= HPXFFT_PATIENT ? 5 : 3; }

// four-step splits Nx = P*Q with both sub-lengths in [2^COL_TILE_MIN_LG, 2^COL_TILE_MAX_LG]
void split_range(int lgx, int &lo, int &hi)
{
    lo = std::max(COL_TILE_MIN_LG, lgx - COL_TILE_MAX_LG);
    hi = std::min(COL_TILE_MAX_LG, lgx - COL_TILE_MIN_LG);
}

// effort >= measure: time the column pass of every valid split on a zeroed
// spectrum of the plan's shape (forward, in place) and keep the fastest.
// If the temporary spectrum cannot be allocated the default split stays.
int tune_split(hpxfft_plan_s *p)
{
    int lo, hi;
    split_range(p->lgx, lo, hi);
    if (lo >= hi) return HPXFFT_OK;
    const size_t bytes = p->batch * p->dist_out * 2 * storage_bytes(p->prec);
    void *tmp = nullptr;
    if (hipMalloc(&tmp, bytes) != hipSuccess)
    {
        (void)hipGetLastError();
        return HPXFFT_OK;
    }
    const bool s16 = p->prec == P_F16 || p->prec == P_BF16;
    int rc = HPXFFT_OK;
    int best = p->lgP;
    double best_ms = 0.0;
    hipError_t e = hipMemsetAsync(tmp, 0, bytes, p->stream);
    if (e == hipSuccess && s16) e = hipMemsetAsync(p->mid_exp, 0, sizeof(int) * p->nx * p->batch, p->stream);
    for (int l = lo; l <= hi && e == hipSuccess && rc == HPXFFT_OK; ++l)
    {
        p->lgP = l;
        p->lgQ = p->lgx - l;
        const int REPS = tune_reps(p->effort);
        rc = run_columns(p, tmp, tmp, false, s16 ? p->col_exp_own : nullptr);   // warm-up
        if (rc == HPXFFT_OK) e = hipEventRecord(p->ev[0], p->stream);
        for (int r = 0; r < REPS && rc == HPXFFT_OK && e == hipSuccess; ++r)
            rc = run_columns(p, tmp, tmp, false, s16 ? p->col_exp_own : nullptr);
        if (rc == HPXFFT_OK && e == hipSuccess) e = hipEventRecord(p->ev[2], p->stream);
        if (rc == HPXFFT_OK && e == hipSuccess) e = hipEventSynchronize(p->ev[2]);
        if (rc == HPXFFT_OK && e == hipSuccess)
        {
            const double ms = elapsed(p->ev[0], p->ev[2]) * 1e3 / REPS;
            p->split_ms[l] = ms;
            if (best_ms == 0.0 || ms < best_ms) { best_ms = ms; best = l; }
        }
    }
    (void)hipStreamSynchronize(p->stream);
    (void)hipFree(tmp);
    p->lgP = best;
    p->lgQ = p->lgx - best;
    if (rc != HPXFFT_OK) return rc;
    HIPCHK(e);
    return HPXFFT_OK;
}

// measure effort, one-level general rows / columns whose radices allow it:
// time the forward transform with the ping-pong and the in-place stages on
// zeroed buffers of the plan's layout (the row and column passes are timed
// apart by the exec's own events) and keep the faster form of each pass.
// If the buffers cannot be allocated the estimate stays.
int tune_gen_ip(hpxfft_plan_s *p)
{
    // both forms must fit: the ping-pong one holds twice the points
    const size_t cap = p->prec == P_F64 ? hpxfft_dev::gen_max_points<double>() : hpxfft_dev::gen_max_points<float>();
    const bool row_ok = p->gen_row && !p->g2r.on && gen_can_ip(p->grow) && (size_t)p->gen_rows * p->grow.n <= cap;
    const bool col_ok = p->gen_col && !p->g2c.on && gen_can_ip(p->gcol) && (size_t)p->gen_w * p->gcol.n <= cap;
    if (!row_ok && !col_ok) return HPXFFT_OK;
    const size_t sb = storage_bytes(p->prec);
    const size_t bin = p->batch * p->dist_in * sb, bout = p->batch * p->dist_out * 2 * sb;
    void *in = nullptr, *out = nullptr;
    if (hipMalloc(&in, bin) != hipSuccess || hipMalloc(&out, bout) != hipSuccess)
    {
        (void)hipGetLastError();
        (void)hipFree(in);
        return HPXFFT_OK;
    }
    const int ip_row = p->grow.ip, ip_col = p->gcol.ip;
    int rc = HPXFFT_OK;
    hipError_t e = hipMemsetAsync(in, 0, bin, p->stream);
    for (int ip = 0; ip < 2 && e == hipSuccess && rc == HPXFFT_OK; ++ip)
    {
        if (row_ok) p->grow.ip = ip;
        if (col_ok) p->gcol.ip = ip;
        const int REPS = tune_reps(p->effort);
        double row_s = 0.0, col_s = 0.0;
        for (int r = 0; r <= REPS && rc == HPXFFT_OK && e == hipSuccess; ++r)   // r = 0: warm-up
        {
            rc = exec_forward(p, in, out, nullptr, nullptr);
            if (rc == HPXFFT_OK) e = hipEventSynchronize(p->ev[2]);
            if (rc == HPXFFT_OK && e == hipSuccess && r > 0)
            {
                row_s += elapsed(p->ev[0], p->ev[1]);
                col_s += elapsed(p->ev[1], p->ev[2]);
            }
        }
        p->gip_ms[0][ip] = row_ok ? row_s * 1e3 / REPS : 0.0;
        p->gip_ms[1][ip] = col_ok ? col_s * 1e3 / REPS : 0.0;
    }
    (void)hipStreamSynchronize(p->stream);
    (void)hipFree(in);
    (void)hipFree(out);
    p->ran = false;
    p->grow.ip = ip_row;
    p->gcol.ip = ip_col;
    if (rc != HPXFFT_OK) return rc;
    HIPCHK(e);
    if (row_ok) p->grow.ip = p->gip_ms[0][1] < p->gip_ms[0][0] ? 1 : 0;
    if (col_ok) p->gcol.ip = p->gip_ms[1][1] < p->gip_ms[1][0] ? 1 : 0;
    return HPXFFT_OK;
}

// measure effort, general columns with narrow strips: time the one-level
// and the two-level column pass on a zeroed spectrum of the plan's shape and
// keep the faster (the scratch of the two-level form is freed if unused)
int tune_gen_columns(hpxfft_plan_s *p)
{
    const size_t bytes = p->batch * p->dist_out * 2 * storage_bytes(p->prec);
    void *tmp = nullptr;
    if (hipMalloc(&tmp, bytes) != hipSuccess)
    {
        (void)hipGetLastError();
        p->g2c.on = false;
        return HPXFFT_OK;
    }
    int rc = HPXFFT_OK;
    hipError_t e = hipMemsetAsync(tmp, 0, bytes, p->stream);
    for (int two = 0; two < 2 && e == hipSuccess && rc == HPXFFT_OK; ++two)
    {
        p->g2c.on = two == 1;
        const int REPS = tune_reps(p->effort);
        rc = run_columns(p, tmp, tmp, false, nullptr);   // warm-up
        if (rc == HPXFFT_OK) e = hipEventRecord(p->ev[0], p->stream);
        for (int r = 0; r < REPS && rc == HPXFFT_OK && e == hipSuccess; ++r) rc = run_columns(p, tmp, tmp, false, nullptr);
        if (rc == HPXFFT_OK && e == hipSuccess) e = hipEventRecord(p->ev[2], p->stream);
        if (rc == HPXFFT_OK && e == hipSuccess) e = hipEventSynchronize(p->ev[2]);
        if (rc == HPXFFT_OK && e == hipSuccess) p->gcol_ms[two] = elapsed(p->ev[0], p->ev[2]) * 1e3 / REPS;
    }
    (void)hipStreamSynchronize(p->stream);
    (void)hipFree(tmp);
    p->g2c.on = p->gcol_ms[1] > 0.0 && p->gcol_ms[1] < p->gcol_ms[0];
    if (!p->g2c.on)
    {
        (void)hipFree(p->gscr);
        p->gscr = nullptr;
        p->g2c.avail = false;
    }
    if (rc != HPXFFT_OK) return rc;
    HIPCHK(e);
    return HPXFFT_OK;
}

// patient effort and above: time both row kernels of a plan whose rows take
// either (2^13 / 2^14 complex points, FP32 arithmetic) on a zeroed grid of the
// plan's shape and keep the faster (measured: the radix-16 kernel, 15-45%
// ahead on MI355X).  If the buffers cannot be allocated the default stays.
int tune_rows(hpxfft_plan_s *p)
{
    if (p->gen_row || !row_pair_covers(p->prec, p->lgL)) return HPXFFT_OK;
    const size_t sb = storage_bytes(p->prec);
    const size_t bin = p->batch * p->dist_in * sb, bout = p->batch * p->dist_out * 2 * sb;
    void *tin = nullptr, *tout = nullptr;
    if (hipMalloc(&tin, bin) != hipSuccess || hipMalloc(&tout, bout) != hipSuccess)
    {
        (void)hipGetLastError();
        (void)hipFree(tin);
        return HPXFFT_OK;
    }
    const bool s16 = p->prec == P_F16 || p->prec == P_BF16;
    RowArgs r{};
    r.in = tin;
    r.out = tout;
    r.ld_in = (long long)p->ld_in;
    r.ld_out = (long long)(2 * p->ld_out);
    r.dist_in = (long long)p->dist_in;
    r.dist_out = (long long)(2 * p->dist_out);
    r.nrows = (int)p->nx;
    r.tw = tw_ptr(p, p->ny);
    r.tw64 = tw64_ptr(p, p->ny);
    r.stw64 = p->tabs.stw64;
    r.stw32 = p->tabs.stw32;
    r.out_exp = s16 ? p->mid_exp : nullptr;
    r.exp_dist = (long long)p->nx;
    const int REPS = tune_reps(p->effort);
    hipError_t e = hipMemsetAsync(tin, 0, bin, p->stream);
    double best = 0.0;
    for (int kind = 0; kind < 2 && e == hipSuccess; ++kind)
    {
        e = launch_rows_pow2(p, r, kind);   // warm-up
        if (e == hipSuccess) e = hipEventRecord(p->ev[0], p->stream);
        for (int i = 0; i < REPS && e == hipSuccess; ++i) e = launch_rows_pow2(p, r, kind);
        if (e == hipSuccess) e = hipEventRecord(p->ev[2], p->stream);
        if (e == hipSuccess) e = hipEventSynchronize(p->ev[2]);
        if (e == hipSuccess)
        {
            p->row_ms[kind] = elapsed(p->ev[0], p->ev[2]) * 1e3 / REPS;
            if (kind == 0 || p->row_ms[kind] < best)
            {
                best = p->row_ms[kind];
                p->row_kind = kind;
            }
        }
    }
    (void)hipStreamSynchronize(p->stream);
    (void)hipFree(tin);
    (void)hipFree(tout);
    HIPCHK(e);
    return HPXFFT_OK;
}

// the power-of-two column pass: one tile of W adjacent columns over all Nx
// rows, or (Nx above the direct kernels' reach) the four-step Nx = P*Q with
// the balanced split; sets the four-step scratch size (one P x Q x TW block
// per strip of TW columns, one slice per grid), allocates nothing
int choose_columns(hpxfft_plan_s *p)
{
    p->W = col_geom(p->prec, p->lgx).nseq;
    const int dmax = col_direct_max_lg(p->prec);
    p->four_step = !p->gen_col && p->lgx > dmax;
    if (!p->four_step) return HPXFFT_OK;
    p->lgP = p->lgx / 2;
    p->lgQ = p->lgx - p->lgP;
    if (p->lgP < COL_TILE_MIN_LG) { p->lgP = COL_TILE_MIN_LG; p->lgQ = p->lgx - p->lgP; }
    if (p->lgQ > COL_TILE_MAX_LG || p->lgP > COL_TILE_MAX_LG) return fail(HPXFFT_UNSUPPORTED_SIZE, "Nx too large");
    const size_t cb = compute_bytes(p->prec) * 2;
    const size_t tw_cols = (size_t)tile_geom(p->prec, p->lgQ).nseq;
    p->CW = (p->ncols + tw_cols - 1) / tw_cols * tw_cols;
    p->scratch_bytes = p->nx * p->CW * cb * p->batch;
    return HPXFFT_OK;
}

int plan_common(hpxfft_plan_s *p);
void free_plan(hpxfft_plan_s *p);

// 16-bit storage with a length beyond one workgroup: an FP32 plan of the same
// shape on an in-place padded FP32 workspace (Nx x 2(Ny/2+1) floats per
// grid), the column maxima and the caller-less column exponents
int plan_wide16(hpxfft_plan_s *p)
{
    auto *w = new hpxfft_plan_s();
    w->nx = p->nx;
    w->ny = p->ny;
    w->batch = p->batch;
    w->ld_in = 2 * p->ncols;
    w->ld_out = p->ncols;
    w->dist_in = p->nx * w->ld_in;
    w->dist_out = p->nx * w->ld_out;
    w->prec = P_F32;
    w->effort = p->effort;
    w->stream = p->stream;
    w->device = p->device;
    if (int rc = plan_common(w))
    {
        free_plan(w);
        return rc;
    }
    p->wide = w;
    HIPCHK(hipMalloc(&p->wide_ws, sizeof(float) * p->batch * w->dist_in));
    HIPCHK(hipMalloc(&p->colmax, sizeof(unsigned) * p->ncols * p->batch));
    HIPCHK(hipMalloc(&p->col_exp_own, sizeof(int) * p->ncols * p->batch));
    if (int rc = alloc_events(p)) return rc;
    p->plan_flops = w->plan_flops;
    return HPXFFT_OK;
}

int plan_common(hpxfft_plan_s *p)
{
    const std::string dims = std::to_string(p->nx) + " x " + std::to_string(p->ny);
    if (p->nx == 0 || p->ny < 2 || p->ny % 2)
        return fail(HPXFFT_UNSUPPORTED_SIZE, "need Nx >= 1 and an even Ny >= 2 (got " + dims + ")");
    // grids run along the second or third launch dimension
    if (p->batch > 65535)
        return fail(HPXFFT_UNSUPPORTED_SIZE, "at most 65535 grids per plan (got " + std::to_string(p->batch) + ")");
    {
        // every element offset the kernels form fits a signed 64-bit index
        const unsigned __int128 lim = (unsigned __int128)1 << 62;
        const unsigned __int128 ext_in = (unsigned __int128)(p->batch - 1) * p->dist_in +
                                         (unsigned __int128)p->nx * p->ld_in;
        const unsigned __int128 ext_out =
            2 * ((unsigned __int128)(p->batch - 1) * p->dist_out + (unsigned __int128)p->nx * p->ld_out);
        if (p->ld_in >= lim || p->ld_out >= lim || ext_in >= lim || ext_out >= lim)
            return fail(HPXFFT_INVALID_ARG, "layout: the strides address more than 2^62 elements");
    }
    // the general kernels take lengths that are not powers of two, and
    // powers of two beyond the power-of-two kernels (Nx > 2^16, Ny > 2^15)
    p->gen_row = !is_pow2(p->ny) || p->ny / 2 > (size_t(1) << ROW_MAX_LG);
    p->gen_col = !is_pow2(p->nx) || p->nx > (size_t(1) << (2 * COL_TILE_MAX_LG));
    if (p->gen_row || p->gen_col)
    {
        // lengths that are not powers of two: general kernels (FP64 / FP32;
        // 16-bit storage within one workgroup, forward only; 16-bit lengths
        // beyond one workgroup run through an FP32 workspace, plan_wide16)
        const bool s16 = p->prec == P_F16 || p->prec == P_BF16;
        const size_t cap = p->prec == P_F64 ? hpxfft_dev::gen_max_points<double>()
                                            : hpxfft_dev::gen_max_points<float>();
        hpxfft_dev::GenPlan one{};
        auto one_level = [&](size_t n) {
            return gen_setup(n, cap, one, 2 * cap) || gen_setup_in_place_only(n, cap, one);
        };
        p->wide16 = s16 && ((p->gen_row && !one_level(p->ny / 2)) || (p->gen_col && !one_level(p->nx)));
        if (p->wide16)
        {
            // the FP32 plan of plan_wide16 must cover both lengths
            const size_t c32 = hpxfft_dev::gen_max_points<float>();
            auto fp32_ok = [&](size_t n) {
                gen2_dim d;
                return gen_setup(n, c32, one, 2 * c32) || gen2_setup(n, c32, d) ||
                       gen_setup_in_place_only(n, c32, one) || gen2b_setup(n, c32, d);
            };
            for (int dim = 0; dim < 2; ++dim)
            {
                const bool gen = dim ? p->gen_col : p->gen_row;
                const size_t n = dim ? p->nx : p->ny / 2;
                if (gen && !fp32_ok(n))
                    return fail(HPXFFT_UNSUPPORTED_SIZE, std::string(dim ? "column length Nx" : "row length Ny/2") +
                                                             " = " + std::to_string(n) + " has no FP32 plan (prime "
                                                             "factors, beyond 2^22 points without a split; got " +
                                                             dims + ")");
            }
        }
        if (!p->wide16)
        {
            auto why = [&](const char *what, size_t n) {
                return std::string(what) + " = " + std::to_string(n) + " needs at most " + std::to_string(cap) +
                       " points with prime factors <= " + std::to_string(hpxfft_dev::GEN_MAX_RADIX) + " (" +
                       std::to_string(2 * cap) + " with prime factors <= 7), or a Bluestein FFT (the power of two >= 2n-1) of at most " + std::to_string(2 * cap) +
                       " points, or (up to 2^24 points) a split n = P*Q into two lengths of at most " +
                       std::to_string(cap) + " points, each with such prime factors or a Bluestein FFT of at most " +
                       std::to_string(cap) + " points, or (up to 2^22 points) Bluestein's algorithm over two levels, "
                       "for this precision (got " + dims + ")";
            };
            const size_t wmax = 128 / (2 * storage_bytes(p->prec));   // one 128-byte segment
            if (p->gen_row && !gen_setup(p->ny / 2, cap, p->grow, 2 * cap))
            {
                if (!s16 && gen2_setup(p->ny / 2, cap, p->g2r))
                {
                    if (p->nx > 65535 || p->batch > 65535)
                        return fail(HPXFFT_UNSUPPORTED_SIZE,
                                    "two-level rows take at most 65535 rows and grids (got " + dims + ")");
                    p->g2r.WA = std::min(gen_width(p->g2r.A, cap, wmax), p->g2r.P);
                    p->g2r.WB = std::min(gen_width(p->g2r.B, cap, wmax), p->g2r.Q);
                }
                else if (!gen_setup_in_place_only(p->ny / 2, cap, p->grow))
                {
                    if (s16)
                        return fail(HPXFFT_UNSUPPORTED_SIZE, "16-bit storage takes lengths that are not powers of two "
                                                             "only within one workgroup: F64 or F32 for row length "
                                                             "Ny/2 = " + std::to_string(p->ny / 2) + " (got " + dims + ")");
                    if (!gen2b_setup(p->ny / 2, cap, p->g2r))
                        return fail(HPXFFT_UNSUPPORTED_SIZE, why("row length Ny/2", p->ny / 2));
                    if (p->nx > 65535 || p->batch > 65535)
                        return fail(HPXFFT_UNSUPPORTED_SIZE,
                                    "two-level rows take at most 65535 rows and grids (got " + dims + ")");
                    p->g2r.WA = std::min(gen_width(p->g2r.A, cap, wmax), p->g2r.P);
                    p->g2r.WB = std::min(gen_width(p->g2r.B, cap, wmax), p->g2r.Q);
                }
            }
            if (p->gen_col && !gen_setup(p->nx, cap, p->gcol, 2 * cap))
            {
                if (!s16 && gen2_setup(p->nx, cap, p->g2c))
                {
                    p->g2c.WA = gen_width(p->g2c.A, cap, wmax);
                    p->g2c.WB = gen_width(p->g2c.B, cap, wmax);
                }
                else if (!gen_setup_in_place_only(p->nx, cap, p->gcol))
                {
                    if (s16)
                        return fail(HPXFFT_UNSUPPORTED_SIZE, "16-bit storage takes lengths that are not powers of two "
                                                             "only within one workgroup: F64 or F32 for column length "
                                                             "Nx = " + std::to_string(p->nx) + " (got " + dims + ")");
                    if (!gen2b_setup(p->nx, cap, p->g2c))
                        return fail(HPXFFT_UNSUPPORTED_SIZE, why("column length Nx", p->nx));
                    p->g2c.WA = gen_width(p->g2c.A, cap, wmax);
                    p->g2c.WB = gen_width(p->g2c.B, cap, wmax);
                }
            }
            if (p->gen_row && !p->g2r.on)
            {
                // short rows share a workgroup: about 2048 points per workgroup
                const size_t ST = (size_t)p->grow.n;
                p->gen_rows = (int)std::max<size_t>(1, std::min(cap / ST, std::min<size_t>(p->nx, 2048 / ST)));
            }
            if (p->gen_col && !p->g2c.on)
            {
                // one 128-byte row segment per workgroup where the LDS allows
                p->gen_w = gen_width(p->gcol, cap, wmax);
                // narrow strips (16- or 32-byte row segments): at measure effort the
                // two-level form, with whole segments, is timed against it
                if (!s16 && p->effort >= HPXFFT_MEASURE && p->gen_w <= 2 && gen2_setup(p->nx, cap, p->g2c))
                {
                    p->g2c.on = false;
                    p->g2c.WA = gen_width(p->g2c.A, cap, wmax);
                    p->g2c.WB = gen_width(p->g2c.B, cap, wmax);
                }
            }
        }
    }
    p->lgx = p->gen_col ? 0 : ilog2(p->nx);
    p->lgL = p->gen_row ? 0 : ilog2(p->ny / 2);
    if (p->lgL > ROW_MAX_LG || p->lgx > 2 * COL_TILE_MAX_LG)
        return fail(HPXFFT_UNSUPPORTED_SIZE, "grid too large: Nx <= 65536 and Ny <= 32768 are supported");
    p->ncols = p->ny / 2 + 1;
    if (p->ld_in < p->ny || p->ld_out < p->ncols || (p->batch > 1 && (p->dist_in < p->nx * p->ld_in ||
                                                                      p->dist_out < p->nx * p->ld_out)))
        return fail(HPXFFT_INVALID_ARG, "layout: need ld_in >= Ny, ld_out >= Ny/2+1 and non-overlapping grids");
    if (p->ld_in % 2 != 0)
        return fail(HPXFFT_INVALID_ARG, "ld_in must be even (rows are read as complex pairs)");
    if (p->dist_in % 2 != 0) return fail(HPXFFT_INVALID_ARG, "dist_in must be even");
    if (int rc = check_device(p->device)) return rc;
    if (int rc = get_tables(p->device, p->tabs)) return rc;
    if (p->wide16) return plan_wide16(p);
    if (p->gen_row || p->gen_col)
        if (int rc = gen_allow_lds(p->device)) return rc;
    if (p->gen_row || p->gen_col)
    {
        // exact FP64 tables (the long-double evaluation of the power-of-two tables)
        const bool row1 = p->gen_row && !p->g2r.on, col1 = p->gen_col && !p->g2c.on;
        p->tw2_shift = p->gen_row ? split_shift(p->ny) : 0;   // w_Ny, split only for two-level rows
        const size_t L = row1 ? gen_table_size(p->grow) : 0;
        const size_t N2 = p->gen_row ? tw_table_size(p->ny, p->tw2_shift) : 0;
        const size_t NX = col1 ? gen_table_size(p->gcol) : 0;
        p->gtab_L = 0;
        p->gtab_2 = L;
        p->gtab_N = L + N2;
        size_t total = L + N2 + NX;
        // two-level dimensions: plans A and B, then w_N^m (m < N) for the phase-A twiddle
        auto place2 = [&](gen2_dim &d) {
            if (!d.avail) return;
            d.tabA = total;
            d.tabB = d.tabA + gen_table_size(d.A);
            d.tabT = d.tabB + gen_table_size(d.B);
            total = d.tabT + tw_table_size(d.points(), d.tw_shift);
            if (d.blue)
            {
                d.tabC = total;
                total += tw_table_size(2 * d.N, d.chirp_shift);
                d.tabBh = total;
                total += d.points();
            }
        };
        place2(p->g2r);
        place2(p->g2c);
        std::vector<double> h(2 * total);
        auto fill_w = [&](size_t off, size_t n) {   // w_n^m, m < n
            for (size_t m = 0; m < n; ++m)
            {
                long double c, sn;
                twiddle_ld(m, n, c, sn);
                h[2 * (off + m)] = (double)c;
                h[2 * (off + m) + 1] = -(double)sn;
            }
        };
        auto fill_tw = [&](size_t off, size_t n, int sh) {   // w_n^m, whole or split at 2^sh
            if (!sh)
            {
                fill_w(off, n);
                return;
            }
            const size_t lo = size_t(1) << sh, nhi = (n + lo - 1) >> sh;
            // lo[i] = w_n^i (i < 2^sh < n), then hi[j] = w_n^{j 2^sh}
            for (size_t i = 0; i < lo + nhi; ++i)
            {
                long double c, sn;
                twiddle_ld((uint64_t)(i < lo ? i : (i - lo) << sh), n, c, sn);
                h[2 * (off + i)] = (double)c;
                h[2 * (off + i) + 1] = -(double)sn;
            }
        };
        if (row1) gen_table(p->grow, h, p->gtab_L);
        if (p->gen_row) fill_tw(p->gtab_2, p->ny, p->tw2_shift);   // w_Ny^m for the r2c split
        if (col1) gen_table(p->gcol, h, p->gtab_N);
        for (gen2_dim *d : {&p->g2r, &p->g2c})
        {
            if (!d->avail) continue;
            gen_table(d->A, h, d->tabA);
            gen_table(d->B, h, d->tabB);
            fill_tw(d->tabT, d->points(), d->tw_shift);
            if (d->blue)
            {
                // the chirp's w_{2N} table and Bh = DFT_M(b) / M
                fill_tw(d->tabC, 2 * d->N, d->chirp_shift);
                std::vector<long double> cr, ci, br, bi;
                const uint64_t M = d->points();
                blue_spectrum(d->N, M, cr, ci, br, bi);
                for (uint64_t k = 0; k < M; ++k)
                {
                    h[2 * (d->tabBh + k)] = (double)(br[k] / (long double)M);
                    h[2 * (d->tabBh + k) + 1] = (double)(bi[k] / (long double)M);
                }
            }
        }
        HIPCHK(hipMalloc(&p->gen_tab, sizeof(double) * h.size()));
        HIPCHK(hipMemcpy(p->gen_tab, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
        const size_t cb = 2 * compute_bytes(p->prec);
        if (p->g2r.on) HIPCHK(hipMalloc(&p->gtmp, p->batch * p->nx * p->g2r.points() * cb));
        if (p->g2c.avail) HIPCHK(hipMalloc(&p->gscr, p->batch * p->g2c.points() * p->ncols * cb));
    }
    if (int rc = choose_columns(p)) return rc;
    if (p->four_step) HIPCHK(hipMalloc(&p->scratch, p->scratch_bytes));
    if (p->prec == P_F16 || p->prec == P_BF16)
    {
        HIPCHK(hipMalloc(&p->mid_exp, sizeof(int) * p->nx * p->batch));
        HIPCHK(hipMalloc(&p->col_exp_own, sizeof(int) * p->ncols * p->batch));
        if (p->four_step)
        {
            int lo, hi;   // room for the largest P any split may use
            split_range(p->lgx, lo, hi);
            HIPCHK(hipMalloc(&p->colpart, sizeof(unsigned) * p->ncols * ((size_t)1 << std::max(hi, p->lgP)) * p->batch));
            HIPCHK(hipMalloc(&p->colmax, sizeof(unsigned) * p->ncols * p->batch));
        }
    }
    if (int rc = alloc_events(p)) return rc;
    if (p->four_step && p->effort >= HPXFFT_MEASURE)
        if (int rc = tune_split(p)) return rc;
    if (p->effort >= HPXFFT_PATIENT)
        if (int rc = tune_rows(p)) return rc;
    if (p->effort >= HPXFFT_MEASURE && ((p->gen_row && !p->g2r.on && gen_can_ip(p->grow)) ||
                                        (p->gen_col && !p->g2c.on && gen_can_ip(p->gcol))))
        if (int rc = tune_gen_ip(p)) return rc;
    if (p->g2c.avail && !p->g2c.on)
        if (int rc = tune_gen_columns(p)) return rc;
    const double lx = std::log2((double)p->nx), ly = std::log2((double)p->ny);
    p->plan_flops = (double)p->batch * ((double)p->nx * 2.5 * (double)p->ny * ly +
                                        (double)p->ncols * 5.0 * (double)p->nx * lx);
    return HPXFFT_OK;
}

void free_plan(hpxfft_plan_s *p)
{
    if (!p) return;
    device_guard g(p->device);
    for (auto s : p->sub) free_plan(s);
    free_plan(p->wide);
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    else (void)hipDeviceSynchronize();
    for (auto &e : p->ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipFree(p->scratch);
    (void)hipFree(p->mid_exp);
    (void)hipFree(p->colpart);
    (void)hipFree(p->colmax);
    (void)hipFree(p->gen_tab);
    (void)hipFree(p->gtmp);
    (void)hipFree(p->gscr);
    (void)hipFree(p->col_exp_own);
    (void)hipFree(p->adj_work);
    (void)hipFree(p->host_buf);
    (void)hipFree(p->wide_ws);
    (void)hipFree(p->bgrids);
    for (auto &ph : p->btasks)
        for (auto &t : ph) (void)hipFree(t);
    (void)hipFree(p->bticket);
    for (auto &e : p->bstage_ev)
        if (e) (void)hipEventDestroy(e);
    if (p->bstage) (void)hipHostFree(p->bstage);
    delete p;
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free entry point (hpxfft_exec_r2c_2d_noalloc): the
// four-step column pass's intermediate, one array per device, held until the
// process ends.  Host code never passes this symbol itself to a kernel (on the
// host it names a stand-in, not device memory): the device address comes from
// hipGetSymbolAddress, once per device, in noalloc_setup.  Calls that overlap
// on one device share the array and race; each call rewrites every scratch
// entry it reads, so nothing needs clearing between calls.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_noalloc_mu;
std::map<int, void *> g_noalloc_scratch;   // device -> address of its scratch array

// first call on device `dev` (current): the scratch address and the twiddle
// tables (uploaded once per device, shared with the plans); later calls only
// look them up
int noalloc_setup(int dev, hipStream_t stream, void *&scr, dev_tables &tabs)
{
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    auto it = g_noalloc_scratch.find(dev);
    if (it == g_noalloc_scratch.end())
    {
        bool have_tables;   // a plan on this device may have uploaded them already
        {
            std::lock_guard<std::mutex> lt(g_tab_mu);
            have_tables = g_tables.count(dev) != 0;
        }
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        HIPCHK(hipStreamIsCapturing(stream, &cs));
        if (!have_tables && cs != hipStreamCaptureStatusNone)
            return fail(HPXFFT_INVALID_ARG, "the first allocation-free call on a device uploads its twiddle tables "
                                            "and cannot be captured: make one call before capturing");
        void *ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        it = g_noalloc_scratch.emplace(dev, ptr).first;
    }
    scr = it->second;
    return get_tables(dev, tabs);
}

}   // namespace

// ---------------------------------------------------------------------------
// C ABI
extern "C" {

const char *hpxfft_last_error(void) { return g_last_error.c_str(); }

const char *hpxfft_version(void) { return "hpxfft-hip 0.1 (gfx950)"; }

int hpxfft_plan_flag_from_string(const char *flag, int *effort)
{
    if (!flag || !effort) return fail(HPXFFT_INVALID_ARG, "null argument");
    const std::string f(flag);
    if (f == "estimate") *effort = HPXFFT_ESTIMATE;
    else if (f == "measure") *effort = HPXFFT_MEASURE;
    else if (f == "patient") *effort = HPXFFT_PATIENT;
    else if (f == "exhaustive") *effort = HPXFFT_EXHAUSTIVE;
    else return fail(HPXFFT_INVALID_ARG, "Invalid FFTW plan flag string");
    return HPXFFT_OK;
}

int hpxfft_plan_r2c_2d_many(hpxfft_plan *plan, size_t nx, size_t ny, size_t batch, size_t ld_in, size_t dist_in,
                            size_t ld_out, size_t dist_out, int precision, int effort, void *stream)
{
    try
    {
        if (!plan) return fail(HPXFFT_INVALID_ARG, "null plan pointer");
        *plan = nullptr;
        if (precision < HPXFFT_F64 || precision > HPXFFT_BF16_STORAGE)
            return fail(HPXFFT_INVALID_ARG, "unknown precision");
        if (effort < HPXFFT_ESTIMATE || effort > HPXFFT_EXHAUSTIVE) return fail(HPXFFT_INVALID_ARG, "unknown effort");
        if (batch == 0) return fail(HPXFFT_INVALID_ARG, "batch must be >= 1");
        const auto t0 = std::chrono::steady_clock::now();
        auto *p = new hpxfft_plan_s();
        p->nx = nx;
        p->ny = ny;
        p->batch = batch;
        p->ld_in = ld_in;
        p->dist_in = batch > 1 ? dist_in : nx * ld_in;
        p->ld_out = ld_out;
        p->dist_out = batch > 1 ? dist_out : nx * ld_out;
        p->prec = precision;
        p->effort = effort;
        p->stream = reinterpret_cast<hipStream_t>(stream);
        if (int rc = plan_common(p))
        {
            std::string msg = g_last_error;
            free_plan(p);
            g_last_error = msg;
            return rc;
        }
        p->plan_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        *plan = p;
        return HPXFFT_OK;
    }
    catch (const std::bad_alloc &)
    {
        return fail(HPXFFT_OUT_OF_MEMORY, "host allocation failed");
    }
    catch (const std::exception &e)
    {
        return fail(HPXFFT_HIP_ERROR, e.what());
    }
}

int hpxfft_plan_r2c_2d(hpxfft_plan *plan, size_t nx, size_t ny, size_t ld_in, int precision, int effort, void *stream)
{
    const size_t ncy = ny / 2 + 1;
    return hpxfft_plan_r2c_2d_many(plan, nx, ny, 1, ld_in, nx * ld_in, ncy, nx * ncy, precision, effort, stream);
}

int hpxfft_exec_r2c_2d(hpxfft_plan plan, const void *d_in, void *d_out)
{
    if (!plan || plan->batched) return fail(HPXFFT_INVALID_ARG, "invalid plan");
    try
    {
        return exec_forward(plan, d_in, d_out, nullptr, nullptr);
    }
    catch (const std::exception &e)
    {
        return fail(HPXFFT_HIP_ERROR, e.what());
    }
}

int hpxfft_exec_r2c_2d_noalloc(size_t nx, size_t ny, size_t ld_in, size_t ld_out, int precision, const void *d_in,
                               void *d_out, void *stream)
{
    return guarded([&]() -> int {
        if (precision != HPXFFT_F64 && precision != HPXFFT_F32)
            return fail(HPXFFT_INVALID_ARG, "the allocation-free entry point takes F64 or F32");
        const std::string dims = std::to_string(nx) + " x " + std::to_string(ny);
        if (!is_pow2(nx) || !is_pow2(ny) || ny < 2 || nx > (size_t(1) << (2 * COL_TILE_MAX_LG)) ||
            ny / 2 > (size_t(1) << ROW_MAX_LG))
            return fail(HPXFFT_UNSUPPORTED_SIZE, "the allocation-free entry point takes powers of two, Nx <= 65536 "
                                                 "and 2 <= Ny <= 32768 (got " + dims + ")");
        if (ld_in < ny || ld_in % 2 || ld_out < ny / 2 + 1)
            return fail(HPXFFT_INVALID_ARG, "layout: need an even ld_in >= Ny and ld_out >= Ny/2+1");
        hpxfft_plan_s p;   // the plan's fields only: no events, no device memory of its own
        p.nx = nx;
        p.ny = ny;
        p.ld_in = ld_in;
        p.dist_in = nx * ld_in;
        p.ld_out = ld_out;
        p.dist_out = nx * ld_out;
        p.prec = precision;
        p.stream = static_cast<hipStream_t>(stream);
        p.timed = false;
        p.lgx = ilog2(nx);
        p.lgL = ilog2(ny / 2);
        p.ncols = ny / 2 + 1;
        if (int rc = check_device(p.device)) return rc;
        void *scr = nullptr;
        if (int rc = noalloc_setup(p.device, p.stream, scr, p.tabs)) return rc;
        if (int rc = choose_columns(&p)) return rc;
        if (p.four_step && p.scratch_bytes > NOALLOC_SCRATCH_BYTES)
            return fail(HPXFFT_UNSUPPORTED_SIZE,
                        "the allocation-free entry point's scratch holds " + std::to_string(NOALLOC_SCRATCH_BYTES) +
                            " bytes; " + dims + " needs " + std::to_string(p.scratch_bytes) +
                            " (Nx x (Ny/2+1) complex values, rounded to whole column strips)");
        p.scratch = p.four_step ? scr : nullptr;
        const int rc = exec_forward(&p, d_in, d_out, nullptr, nullptr);
        p.scratch = nullptr;
        return rc;
    });
}

int hpxfft_exec_r2c_2d_scaled(hpxfft_plan plan, const void *d_in, void *d_out, const int *d_in_row_exp,
                              int *d_out_col_exp)
{
    if (!plan || plan->batched) return fail(HPXFFT_INVALID_ARG, "invalid plan");
    if (plan->prec != P_F16 && plan->prec != P_BF16)
        return fail(HPXFFT_INVALID_ARG, "scaled exec is for 16-bit storage plans");
    if (!d_out_col_exp) return fail(HPXFFT_INVALID_ARG, "d_out_col_exp is required");
    try
    {
        return exec_forward(plan, d_in, d_out, d_in_row_exp, d_out_col_exp);
    }
    catch (const std::exception &e)
    {
        return fail(HPXFFT_HIP_ERROR, e.what());
    }
}

int hpxfft_exec_r2c_2d_host(hpxfft_plan plan, double *inout)
{
    return guarded([&]() -> int {
        if (!plan || plan->batched) return fail(HPXFFT_INVALID_ARG, "invalid plan");
        if (!inout) return fail(HPXFFT_INVALID_ARG, "null host pointer");
        if (plan->prec != P_F64) return fail(HPXFFT_INVALID_ARG, "host exec is for F64 plans");
        if (plan->ld_in != 2 * plan->ld_out || plan->batch != 1)
            return fail(HPXFFT_INVALID_ARG, "host exec needs the in-place padded layout (ld_in == 2*(Ny/2+1))");
        device_guard g(plan->device);
        const size_t bytes = plan->nx * plan->ld_in * sizeof(double);
        if (!plan->host_buf || plan->host_bytes < bytes)
        {
            (void)hipFree(plan->host_buf);
            plan->host_buf = nullptr;
            plan->host_bytes = 0;
            HIPCHK(hipMalloc(&plan->host_buf, bytes));
            plan->host_bytes = bytes;
        }
        scoped_event e0, e1;   // destroyed on every return path
        HIPCHK(hipEventCreate(&e0.ev));
        HIPCHK(hipEventCreate(&e1.ev));
        HIPCHK(hipEventRecord(e0.ev, plan->stream));
        HIPCHK(hipMemcpyAsync(plan->host_buf, inout, bytes, hipMemcpyHostToDevice, plan->stream));
        if (int rc = exec_forward(plan, plan->host_buf, plan->host_buf, nullptr, nullptr)) return rc;
        HIPCHK(hipMemcpyAsync(inout, plan->host_buf, bytes, hipMemcpyDeviceToHost, plan->stream));
        HIPCHK(hipEventRecord(e1.ev, plan->stream));
        HIPCHK(hipEventSynchronize(e1.ev));
        plan->h2d = elapsed(e0.ev, plan->ev[0]);
        plan->d2h = elapsed(plan->ev[2], e1.ev);
        return HPXFFT_OK;
    });
}

}   // extern "C"

namespace {

// spectrum -> real through an inverse column c2c into the plan's work buffer,
// then a row c2r: the r2c adjoint (halve) or the unnormalised inverse c2r
int exec_backward(hpxfft_plan_s *plan, const void *d_spec, void *d_real, bool halve)
{
    const char *what = halve ? "adjoint" : "c2r";
    if (!plan || plan->batched) return fail(HPXFFT_INVALID_ARG, "invalid plan");
    if (!d_spec || !d_real) return fail(HPXFFT_INVALID_ARG, "null device pointer");
    if (int rc = check_aligned(plan, d_spec, d_real)) return rc;
    if (plan->prec != P_F64 && plan->prec != P_F32)
        return fail(HPXFFT_INVALID_ARG, std::string(what) + " is implemented for F64 and F32 plans");
    device_guard g(plan->device);
    const size_t sb = storage_bytes(plan->prec);
    const size_t bytes = plan->batch * plan->dist_out * 2 * sb;
    if (!plan->adj_work)
    {
        HIPCHK(hipMalloc(&plan->adj_work, bytes));
        plan->adj_bytes = bytes;
    }
    if (int rc = alloc_events(plan)) return rc;
    HIPCHK(hipEventRecord(plan->ev[0], plan->stream));
    if (int rc = run_columns(plan, d_spec, plan->adj_work, true, nullptr, plan->ev[3])) return rc;
    HIPCHK(hipEventRecord(plan->ev[1], plan->stream));
    RowArgs r{};
    r.in = plan->adj_work;
    r.out = d_real;
    r.ld_in = (long long)(2 * plan->ld_out);
    r.ld_out = (long long)plan->ld_in;
    r.dist_in = (long long)(2 * plan->dist_out);
    r.dist_out = (long long)plan->dist_in;
    r.nrows = (int)plan->nx;
    r.tw = tw_ptr(plan, plan->ny);
    r.tw64 = tw64_ptr(plan, plan->ny);
    r.stw64 = plan->tabs.stw64;
    r.stw32 = plan->tabs.stw32;
    if (plan->g2r.on)
    {
        if (int rc = run_rows_2level(plan, nullptr, d_real, plan->adj_work, nullptr, true, halve)) return rc;
    }
    else if (plan->gen_row)
    {
        const int ns = plan->gen_rows;
        const size_t lds = gen_lds_bytes(plan, plan->grow, (size_t)ns * plan->grow.n);
        const dim3 grid((unsigned)((plan->nx + ns - 1) / ns), (unsigned)plan->batch);
        const unsigned nt = gen_threads(lds, plan->grow);
        const auto *tab = reinterpret_cast<const hpxfft_dev::cpx<double> *>(plan->gen_tab);
        const auto *tL = tab + plan->gtab_L, *t2 = tab + plan->gtab_2;
        using namespace hpxfft_dev;
        const hipStream_t st = plan->stream;
        if (plan->prec == P_F64)
        {
            if (halve) GEN_LAUNCH(gen_row_c2r_kernel, (double, true), nt, grid, lds, st, r, plan->grow, tL, t2, ns);
            else GEN_LAUNCH(gen_row_c2r_kernel, (double, false), nt, grid, lds, st, r, plan->grow, tL, t2, ns);
        }
        else
        {
            if (halve) GEN_LAUNCH(gen_row_c2r_kernel, (float, true), nt, grid, lds, st, r, plan->grow, tL, t2, ns);
            else GEN_LAUNCH(gen_row_c2r_kernel, (float, false), nt, grid, lds, st, r, plan->grow, tL, t2, ns);
        }
        HIPCHK(hipGetLastError());
    }
    else
    {
        const geom rg = row_geom(plan->prec, plan->lgL);
        const unsigned gx = (unsigned)((plan->nx + rg.nseq - 1) / rg.nseq);
        HIPCHK(row_c2r(plan->prec, plan->lgL, halve, r, gx, (unsigned)plan->batch, plan->stream));
    }
    HIPCHK(hipEventRecord(plan->ev[2], plan->stream));
    plan->ran = true;
    plan->col_split = plan->four_step;
    plan->col_first = true;
    return HPXFFT_OK;
}

}   // namespace

extern "C" {

int hpxfft_exec_r2c_2d_adjoint(hpxfft_plan plan, const void *d_spec, void *d_real)
{
    return guarded([&]() -> int {
        return exec_backward(plan, d_spec, d_real, true);
    });
}

int hpxfft_exec_c2r_2d(hpxfft_plan plan, const void *d_spec, void *d_real)
{
    return guarded([&]() -> int {
        return exec_backward(plan, d_spec, d_real, false);
    });
}

int hpxfft_get_measurement(hpxfft_plan plan, const char *key, double *value)
{
    return guarded([&]() -> int {
        if (!plan || !key || !value) return fail(HPXFFT_INVALID_ARG, "null argument");
        const std::string k(key);
        double v = 0.0;
        if (k == "plan") v = plan->plan_seconds;
        else if (k == "plan_flops") v = plan->plan_flops;
        else if (k == "first_trans" || k == "second_trans") v = 0.0;
        else if (k == "h2d") v = plan->h2d;
        else if (k == "d2h") v = plan->d2h;
        else if (k == "total" || k == "first_fftw" || k == "second_fftw")
        {
            if (plan->ran)
            {
                device_guard g(plan->device);
                HIPCHK(hipEventSynchronize(plan->ev[2]));
                if (k == "total") v = elapsed(plan->ev[0], plan->ev[2]);
                else if (k == "first_fftw") v = elapsed(plan->ev[0], plan->ev[1]);
                else v = elapsed(plan->ev[1], plan->ev[2]);
            }
        }
        else if (k == "column_a" || k == "column_b")
        {
            // the two launches of a four-step column pass (0 when the last run had no split)
            if (plan->ran && plan->col_split)
            {
                device_guard g(plan->device);
                HIPCHK(hipEventSynchronize(plan->ev[2]));
                // forward: rows ev0..ev1, columns ev1..ev3..ev2; backward: columns ev0..ev3..ev1, rows ev1..ev2
                if (plan->col_first)
                    v = k == "column_a" ? elapsed(plan->ev[0], plan->ev[3]) : elapsed(plan->ev[3], plan->ev[1]);
                else
                    v = k == "column_a" ? elapsed(plan->ev[1], plan->ev[3]) : elapsed(plan->ev[3], plan->ev[2]);
            }
        }
        else
        {
            auto it = plan->extra.find(k);
            if (it != plan->extra.end()) v = it->second;
            else plan->extra[k] = 0.0;
        }
        *value = v;
        return HPXFFT_OK;
    });
}

int hpxfft_plan_describe(hpxfft_plan plan, char *buf, size_t len, size_t *needed)
{
    return guarded([&]() -> int {
        if (!plan) return fail(HPXFFT_INVALID_ARG, "null plan");
        const std::string s = describe(plan);
        if (needed) *needed = s.size();
        if (buf && len)
        {
            const size_t n = std::min(len - 1, s.size());
            std::memcpy(buf, s.data(), n);
            buf[n] = '\0';
        }
        return HPXFFT_OK;
    });
}

int hpxfft_write_plan(hpxfft_plan plan, const char *path)
{
    return guarded([&]() -> int {
        if (!plan || !path) return fail(HPXFFT_INVALID_ARG, "null argument");
        FILE *f = std::fopen(path, "a");
        if (!f) return fail(HPXFFT_IO_ERROR, std::string("Failed to open file: ") + path);
        const std::string s = describe(plan);
        std::fprintf(f, "GPU r2c 2D plan:\n%s\n", s.c_str());
        std::fclose(f);
        return HPXFFT_OK;
    });
}

int hpxfft_plan_destroy(hpxfft_plan plan)
{
    return guarded([&]() -> int {
        if (!plan) return HPXFFT_OK;
        free_plan(plan);
        return HPXFFT_OK;
    });
}

}   // extern "C"

namespace {

int batched_run(hpxfft_plan_s *p, int cfg_row, int cfg_col)
{
    const int cfg[2] = {cfg_row, cfg_col};
    for (int ph = 0; ph < 2; ++ph)
    {
        unsigned *tk = p->bticket + ph * BATCH_QUEUES;
        hipLaunchKernelGGL(hpxfft_dev::zero_u32_kernel<0>, dim3(1), dim3(64), 0, p->stream, tk,
                           (long long)BATCH_QUEUES);
        HIPCHK(hipGetLastError());
        const int c = cfg[ph];
        const int nb = std::min(p->bblocks[ph][c], std::max(1, p->bntasks[ph][c]));
        HIPCHK(batched_launch(ph, c, p->bgrids, p->btasks[ph][c], p->bntasks[ph][c], tk, p->tabs.tw32, nb,
                              p->stream));
        if (ph == 0) HIPCHK(hipEventRecord(p->ev[1], p->stream));
    }
    return HPXFFT_OK;
}

// upload the grid descriptors for d_grids unless the device already holds
// them (the same pointers as the last upload): then nothing is enqueued.  The
// pinned staging has two slots used in turn; a slot is rewritten only after
// the copy that last read it has run, so no call waits for the stream.
int batched_upload_grids(hpxfft_plan_s *p, void *const *d_grids)
{
    const size_t n = p->bnx.size(), gb = batched_grid_desc_bytes();
    for (size_t i = 0; i < n; ++i)
    {
        if (!d_grids[i]) return fail(HPXFFT_INVALID_ARG, "null grid pointer " + std::to_string(i));
        if (int rc = check_aligned(p, d_grids[i], d_grids[i])) return rc;
    }
    if (p->bptrs.size() == n && std::equal(p->bptrs.begin(), p->bptrs.end(), d_grids)) return HPXFFT_OK;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIPCHK(hipStreamIsCapturing(p->stream, &cs));
    if (cs != hipStreamCaptureStatusNone)
        return fail(HPXFFT_INVALID_ARG, "a captured batched exec must use the grid pointers of the plan's previous "
                                        "exec (a new pointer set is uploaded from host memory, which a graph "
                                        "cannot replay)");
    const int slot = p->bstage_next;
    HIPCHK(hipEventSynchronize(p->bstage_ev[slot]));
    char *st = reinterpret_cast<char *>(p->bstage) + (size_t)slot * gb * n;
    for (size_t i = 0; i < n; ++i)
        batched_fill_grid(st + i * gb, reinterpret_cast<float *>(d_grids[i]), (int)p->bnx[i], (int)p->bny[i]);
    // forget the old set first: if the copy fails, the next call uploads again
    p->bptrs.clear();
    HIPCHK(hipMemcpyAsync(p->bgrids, st, gb * n, hipMemcpyHostToDevice, p->stream));
    HIPCHK(hipEventRecord(p->bstage_ev[slot], p->stream));
    p->bstage_next = 1 - slot;
    p->bptrs.assign(d_grids, d_grids + n);
    return HPXFFT_OK;
}

// time every candidate configuration of both phases on scratch grids of the
// plan's sizes and keep the fastest (the data values do not affect timing)
int batched_autotune(hpxfft_plan_s *p)
{
    size_t total = 0;
    std::vector<size_t> off(p->bnx.size());
    for (size_t i = 0; i < p->bnx.size(); ++i)
    {
        off[i] = total;
        total += p->bnx[i] * 2 * (p->bny[i] / 2 + 1);
    }
    float *buf = nullptr;
    HIPCHK(hipMalloc(&buf, total * sizeof(float)));
    std::vector<void *> ptrs(p->bnx.size());
    for (size_t i = 0; i < ptrs.size(); ++i) ptrs[i] = buf + off[i];
    int rc = batched_upload_grids(p, ptrs.data());
    hipEvent_t a = nullptr, b = nullptr;
    if (!rc && hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess)
    {
        for (int ph = 0; ph < 2 && !rc; ++ph)
        {
            double best = 1e30;
            for (int c = 0; c < BATCH_NCFG && !rc; ++c)
            {
                const int nb = std::min(p->bblocks[ph][c], std::max(1, p->bntasks[ph][c]));
                float ms = 0.0f, acc = 0.0f;
                const int reps = 2 + tune_reps(p->effort);   // two warm-up runs, then timed
                for (int rep = 0; rep < reps && !rc; ++rep)
                {
                    unsigned *tk = p->bticket + ph * BATCH_QUEUES;
                    hipLaunchKernelGGL(hpxfft_dev::zero_u32_kernel<0>, dim3(1), dim3(64), 0, p->stream, tk,
                                       (long long)BATCH_QUEUES);
                    (void)hipEventRecord(a, p->stream);
                    if (batched_launch(ph, c, p->bgrids, p->btasks[ph][c], p->bntasks[ph][c], tk, p->tabs.tw32, nb,
                                       p->stream) != hipSuccess)
                        rc = fail(HPXFFT_HIP_ERROR, "batched autotune launch failed");
                    (void)hipEventRecord(b, p->stream);
                    (void)hipEventSynchronize(b);
                    (void)hipEventElapsedTime(&ms, a, b);
                    if (rep >= 2) acc += ms;   // two warm-up runs
                }
                p->btune_ms[ph][c] = acc / (reps - 2);
                if (p->btune_ms[ph][c] < best)
                {
                    best = p->btune_ms[ph][c];
                    p->bcfg[ph] = c;
                }
            }
        }
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    (void)hipStreamSynchronize(p->stream);
    (void)hipFree(buf);
    p->bptrs.clear();   // the descriptors point into the freed buffer: the next exec uploads
    return rc;
}

int plan_batched(hpxfft_plan_s *p)
{
    for (size_t i = 0; i < p->bnx.size(); ++i)
    {
        const size_t x = p->bnx[i], y = p->bny[i];
        if (!is_pow2(x) || !is_pow2(y) || x < (size_t)BATCH_MIN_SIDE || y < (size_t)BATCH_MIN_SIDE ||
            x > (size_t)BATCH_MAX_SIDE || y > (size_t)BATCH_MAX_SIDE)
            return fail(HPXFFT_UNSUPPORTED_SIZE, "batched grids need power-of-two sides in [64, 2048] (grid " +
                                                     std::to_string(i) + ": " + std::to_string(x) + " x " +
                                                     std::to_string(y) + ")");
    }
    if (int rc = check_device(p->device)) return rc;
    if (int rc = get_tables(p->device, p->tabs)) return rc;
    if (int rc = alloc_events(p)) return rc;
    const size_t n = p->bnx.size();
    HIPCHK(hipMalloc(&p->bgrids, batched_grid_desc_bytes() * n));
    HIPCHK(hipHostMalloc(&p->bstage, 2 * batched_grid_desc_bytes() * n, hipHostMallocDefault));
    for (auto &e : p->bstage_ev)
    {
        HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        HIPCHK(hipEventRecord(e, p->stream));   // both slots start free
    }
    HIPCHK(hipMalloc(&p->bticket, 2 * BATCH_QUEUES * sizeof(unsigned)));
    const size_t tb = batched_task_desc_bytes();
    for (int ph = 0; ph < 2; ++ph)
    {
        for (int c = 0; c < BATCH_NCFG; ++c)
        {
            const int pts = batched_pts(ph, c);
            std::vector<char> tasks;
            int nt = 0;
            for (size_t i = 0; i < n; ++i)
            {
                const int nx = (int)p->bnx[i], ny = (int)p->bny[i];
                if (ph == 0)
                {
                    const int rows = std::max(1, pts / (ny / 2));
                    for (int r = 0; r < nx; r += rows, ++nt)
                    {
                        tasks.resize((size_t)(nt + 1) * tb);
                        batched_fill_task(tasks.data() + (size_t)nt * tb, (int)i, r);
                    }
                }
                else
                {
                    const int w = std::max(1, pts / nx), ncols = ny / 2 + 1;
                    for (int st = 0; st * w < ncols; ++st, ++nt)
                    {
                        tasks.resize((size_t)(nt + 1) * tb);
                        batched_fill_task(tasks.data() + (size_t)nt * tb, (int)i, st);
                    }
                }
            }
            p->bntasks[ph][c] = nt;
            HIPCHK(hipMalloc(&p->btasks[ph][c], tasks.size()));
            HIPCHK(hipMemcpy(p->btasks[ph][c], tasks.data(), tasks.size(), hipMemcpyHostToDevice));
            p->bblocks[ph][c] = batched_resident_blocks(ph, c);
        }
    }
    double fl = 0.0;
    for (size_t i = 0; i < n; ++i)
    {
        const double x = (double)p->bnx[i], y = (double)p->bny[i];
        fl += x * 2.5 * y * std::log2(y) + (y / 2 + 1) * 5.0 * x * std::log2(x);
    }
    p->plan_flops = fl;
    if (p->effort >= HPXFFT_MEASURE)
    {
        if (int rc = batched_autotune(p)) return rc;
    }
    return HPXFFT_OK;
}

}   // namespace

extern "C" {

int hpxfft_plan_r2c_2d_batched(hpxfft_plan *plan, size_t count, const size_t *nx, const size_t *ny, int precision,
                               int effort, void *stream)
{
    try
    {
        if (!plan) return fail(HPXFFT_INVALID_ARG, "null plan pointer");
        *plan = nullptr;
        if (count == 0 || !nx || !ny) return fail(HPXFFT_INVALID_ARG, "need count >= 1 and size arrays");
        if (precision != HPXFFT_F32) return fail(HPXFFT_INVALID_ARG, "batched plans are FP32");
        if (effort < HPXFFT_ESTIMATE || effort > HPXFFT_EXHAUSTIVE) return fail(HPXFFT_INVALID_ARG, "unknown effort");
        const auto t0 = std::chrono::steady_clock::now();
        auto *p = new hpxfft_plan_s();
        p->batched = true;
        p->prec = precision;
        p->effort = effort;
        p->stream = reinterpret_cast<hipStream_t>(stream);
        p->bnx.assign(nx, nx + count);
        p->bny.assign(ny, ny + count);
        if (int rc = plan_batched(p))
        {
            std::string msg = g_last_error;
            free_plan(p);
            g_last_error = msg;
            return rc;
        }
        p->plan_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        *plan = p;
        return HPXFFT_OK;
    }
    catch (const std::bad_alloc &)
    {
        return fail(HPXFFT_OUT_OF_MEMORY, "host allocation failed");
    }
    catch (const std::exception &e)
    {
        return fail(HPXFFT_HIP_ERROR, e.what());
    }
}

int hpxfft_plan_batched_set_config(hpxfft_plan plan, int row_cfg, int col_cfg)
{
    if (!plan || !plan->batched) return fail(HPXFFT_INVALID_ARG, "not a batched plan");
    if (row_cfg < 0 || row_cfg >= BATCH_NCFG || col_cfg < 0 || col_cfg >= BATCH_NCFG)
        return fail(HPXFFT_INVALID_ARG, "configuration out of range");
    plan->bcfg[0] = row_cfg;
    plan->bcfg[1] = col_cfg;
    return HPXFFT_OK;
}

int hpxfft_plan_set_column_split(hpxfft_plan plan, int log2_p)
{
    if (!plan || plan->batched || !plan->four_step)
        return fail(HPXFFT_INVALID_ARG, "the plan's column pass is not a four-step");
    int lo, hi;
    split_range(plan->lgx, lo, hi);
    if (log2_p < lo || log2_p > hi)
        return fail(HPXFFT_INVALID_ARG, "column split out of range: log2 P in [" + std::to_string(lo) + ", " +
                                            std::to_string(hi) + "] for Nx = " + std::to_string(plan->nx));
    // the scratch holds TW-column strips of Nx rows whatever the split, and the
    // 16-bit column maxima were sized for the largest P of the range
    plan->lgP = log2_p;
    plan->lgQ = plan->lgx - log2_p;
    return HPXFFT_OK;
}

int hpxfft_plan_set_row_kernel(hpxfft_plan plan, int kind)
{
    if (!plan || plan->batched) return fail(HPXFFT_INVALID_ARG, "invalid plan");
    if (kind < 0 || kind > 1) return fail(HPXFFT_INVALID_ARG, "row kernel kind out of range (0..1)");
    if (kind != 0 && (plan->gen_row || !row_pair_covers(plan->prec, plan->lgL)))
        return fail(HPXFFT_INVALID_ARG, "the lane-pair row kernel covers rows of 2^13 and 2^14 complex points "
                                        "(Ny = 16384, 32768) with FP32 arithmetic; this plan has Ny = " +
                                            std::to_string(plan->ny) + ", precision " + prec_name(plan->prec));
    plan->row_kind = kind;
    return HPXFFT_OK;
}

int hpxfft_exec_r2c_2d_batched(hpxfft_plan plan, void *const *d_grids)
{
    return guarded([&]() -> int {
        if (!plan || !plan->batched) return fail(HPXFFT_INVALID_ARG, "not a batched plan");
        if (!d_grids) return fail(HPXFFT_INVALID_ARG, "null grid pointer array");
        device_guard g(plan->device);
        if (int rc = batched_upload_grids(plan, d_grids)) return rc;
        HIPCHK(hipEventRecord(plan->ev[0], plan->stream));
        if (int rc = batched_run(plan, plan->bcfg[0], plan->bcfg[1])) return rc;
        HIPCHK(hipEventRecord(plan->ev[2], plan->stream));
        plan->ran = true;
        plan->col_split = false;
        return HPXFFT_OK;
    });
}

}   // extern "C"
