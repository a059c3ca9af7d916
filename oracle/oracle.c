/*
 * oracle.c -- CPU FP64 oracle for the 2-D real-to-complex FFT path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * C++ drop-in class, the Python binding) links, loads or calls this file.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * it, and only as the checker / the timed CPU comparator.
 *
 * It restates, in plain C with OpenMP, what HPX-FFT's shared-memory loop
 * computes (core/src/shared/loop.cpp:56-113) with the arithmetic that the
 * reference delegates to FFTW 3.3.10 (core/src/util/adapter_fftw.cpp:9,14,
 * 28-29,34: 1-D r2c of length Ny, 1-D c2c FFTW_FORWARD of length Nx, in
 * place, unnormalised).  FFTW is not vendored in the reference and is not
 * installed here, so the 1-D transforms are written from the textbook
 * definitions:
 *
 *   X[k] = sum_n x[n] exp(-2 pi i k n / N)                 (FFTW_FORWARD)
 *
 * Lengths: any Nx >= 1 and any even Ny >= 2 (powers of two take a radix-2
 * path, other lengths a mixed-radix one).
 *
 * Pins (all exercised by tests/test_oracle.py):
 *   - the reference's only known-answer test, the 4x6 case with exact
 *     equality (test/src/test_shared_loop.cpp:15-34,53);
 *   - a long-double brute-force O(N^2) DFT (oracle_dft_r2c_2d_brute);
 *   - closed forms (ramp input of examples/hpxfft/shared_loop_2d.cpp:33-40,
 *     deltas, constants), Parseval, linearity, the shift theorem;
 *   - numpy.fft.rfft2 (pocketfft) fixtures committed under tests/golden/.
 *
 * Layout (identical to FFTW's in-place padded r2c, and to the reference's
 * vector_2d<double>(Nx, 2*(Ny/2+1)), core/include/hpxfft/util/vector_2d.hpp):
 *   real input row i at values[i*ld + 0 .. Ny-1], ld >= 2*(Ny/2+1);
 *   complex output row i at values[i*ld + 2*k, 2*k+1], k = 0..Ny/2.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Seeded inputs: SplitMix64 (Steele, Lea & Flood, OOPSLA 2014).  The k-th
 * output of a generator seeded with `seed` is mix(seed + (k+1)*gamma), so any
 * sample can be generated independently (counter-based).  Sample k is the
 * k-th real sample in row-major order over the Nx x Ny real grid (padding
 * excluded and set to zero). */
static const uint64_t SM64_GAMMA = 0x9e3779b97f4a7c15ULL;

static inline uint64_t sm64_mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

static inline uint64_t sm64_at(uint64_t seed, uint64_t k) { return sm64_mix(seed + (k + 1) * SM64_GAMMA); }

/* u in [0,1) with 53 random bits */
static inline double sm64_u01(uint64_t seed, uint64_t k) { return (double)(sm64_at(seed, k) >> 11) * 0x1.0p-53; }

/* dist: 0 = U[0,1), 1 = U[-1,1), 2 = N(0,1) (Box-Muller on consecutive pairs),
 *       3 = ramp x[i][j] = j (examples/hpxfft/shared_loop_2d.cpp:33-40) */
void oracle_fill(double *x, size_t nx, size_t ny, size_t ld, uint64_t seed, int dist)
{
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)nx; ++i)
    {
        double *row = x + (size_t)i * ld;
        for (size_t j = 0; j < ny; ++j)
        {
            const uint64_t k = (uint64_t)i * ny + j;
            double v;
            if (dist == 0)
                v = sm64_u01(seed, k);
            else if (dist == 1)
                v = 2.0 * sm64_u01(seed, k) - 1.0;
            else if (dist == 2)
            {
                const uint64_t p = k & ~(uint64_t)1;
                const double u1 = sm64_u01(seed, p), u2 = sm64_u01(seed, p + 1);
                const double r = sqrt(-2.0 * log(1.0 - u1));
                const double th = 2.0 * 3.14159265358979323846 * u2;
                v = (k & 1) ? r * sin(th) : r * cos(th);
            }
            else
                v = (double)j;
            row[j] = v;
        }
        for (size_t j = ny; j < ld; ++j) row[j] = 0.0;
    }
}

/* ------------------------------------------------------------------------ */
/* Twiddles w_N^m = exp(-2 pi i m / N), evaluated directly (no recurrence) in
 * long double after reducing m/N to the first octant, so that the values at
 * multiples of pi/4 are exact (cos(pi/2) is exactly 0, etc.).  This is what
 * makes the 4x6 known-answer test bit-exact. */
void oracle_twiddle(uint64_t m, uint64_t n, double *re, double *im)
{
    m %= n;
    /* angle a = 2*pi*m/n ; reduce with 8*m vs n in integer arithmetic */
    const uint64_t m8 = 8 * m; /* octant index = floor(8m/n) */
    const uint64_t oct = m8 / n;
    const long double PI_L = 3.141592653589793238462643383279502884L;
    /* residual within the octant, as a fraction of pi/4 */
    const uint64_t rem = m8 - oct * n; /* in [0, n) */
    long double c, s;                  /* cos, sin of angle in [0, 2pi) */
    long double t, ct, st;
    /* angle = oct*pi/4 + rem/n * pi/4 */
    if (oct % 2 == 0)
    {
        t = (long double)rem / (long double)n * (PI_L / 4.0L);
        ct = cosl(t);
        st = sinl(t);
    }
    else
    {
        /* use the complementary angle so the reduced argument is <= pi/4 */
        t = (long double)(n - rem) / (long double)n * (PI_L / 4.0L);
        ct = cosl(t);
        st = sinl(t);
    }
    /* base angle b = oct*pi/4 (even) or (oct+1)*pi/4 (odd), then +/- t */
    switch (oct)
    {
    case 0: c = ct; s = st; break;                 /* t */
    case 1: c = st; s = ct; break;                 /* pi/2 - t */
    case 2: c = -st; s = ct; break;                /* pi/2 + t */
    case 3: c = -ct; s = st; break;                /* pi - t */
    case 4: c = -ct; s = -st; break;               /* pi + t */
    case 5: c = -st; s = -ct; break;               /* 3pi/2 - t */
    case 6: c = st; s = -ct; break;                /* 3pi/2 + t */
    default: c = ct; s = -st; break;               /* 2pi - t */
    }
    if (rem == 0 && (oct % 2 == 0))
    {
        /* exact axis values */
        c = (oct == 0) ? 1.0L : (oct == 2) ? 0.0L : (oct == 4) ? -1.0L : 0.0L;
        s = (oct == 0) ? 0.0L : (oct == 2) ? 1.0L : (oct == 4) ? 0.0L : -1.0L;
    }
    *re = (double)c;
    *im = -(double)s; /* forward sign */
}

/* ------------------------------------------------------------------------ */
/* 1-D complex FFT, sign = -1 forward (FFTW_FORWARD), +1 backward,
 * unnormalised, any length n >= 1.  Powers of two: iterative radix-2
 * decimation in time with a table of n/2 twiddles.  Other lengths:
 * recursive decimation in time by the smallest prime factor p of n,
 *   X[k + q m] = sum_{r<p} w_n^{r (k + q m)} S_r[k],  n = p m,
 * S_r the length-m transform of x[r + p j], with a table of all n twiddles
 * (O(n sum of prime factors)). */
typedef struct
{
    size_t n;
    int pow2;
    double *tw; /* powers of two: w_n^m, m < n/2; otherwise w_n^m, m < n (forward sign) */
} fft_tab;

static int fft_tab_init(fft_tab *t, size_t n)
{
    t->n = n;
    t->tw = NULL;
    t->pow2 = (n & (n - 1)) == 0;
    if (n < 2) return 0;
    const size_t cnt = t->pow2 ? n / 2 : n;
    t->tw = (double *)malloc(sizeof(double) * 2 * cnt);
    if (!t->tw) return -1;
    for (size_t m = 0; m < cnt; ++m) oracle_twiddle(m, n, &t->tw[2 * m], &t->tw[2 * m + 1]);
    return 0;
}

static void fft_tab_free(fft_tab *t) { free(t->tw); t->tw = NULL; }

static size_t smallest_factor(size_t n)
{
    if (n % 2 == 0) return 2;
    for (size_t f = 3; f * f <= n; f += 2)
        if (n % f == 0) return f;
    return n;
}

/* length-n transform of in[0], in[s], ..., in[(n-1)s] into out[0..n);
 * tw: the table of the full length N (n divides N); work: N complex */
static void fft_rec(const double *tw, size_t N, const double *in, size_t s, double *out, size_t n, int sign,
                    double *work)
{
    if (n == 1)
    {
        out[0] = in[0];
        out[1] = in[1];
        return;
    }
    const size_t p = smallest_factor(n), m = n / p, step = N / n;
    for (size_t r = 0; r < p; ++r) fft_rec(tw, N, in + 2 * r * s, s * p, out + 2 * r * m, m, sign, work);
    for (size_t q = 0; q < p; ++q)
        for (size_t k = 0; k < m; ++k)
        {
            double ar = 0.0, ai = 0.0;
            for (size_t r = 0; r < p; ++r)
            {
                const size_t e = (r * (k + q * m)) % n;
                const double wr = tw[2 * e * step], wi = sign < 0 ? tw[2 * e * step + 1] : -tw[2 * e * step + 1];
                const double xr = out[2 * (r * m + k)], xi = out[2 * (r * m + k) + 1];
                ar += xr * wr - xi * wi;
                ai += xr * wi + xi * wr;
            }
            work[2 * (k + q * m)] = ar;
            work[2 * (k + q * m) + 1] = ai;
        }
    memcpy(out, work, sizeof(double) * 2 * n);
}

static void fft_c2c_tab(const fft_tab *t, double *a, int sign)
{
    const size_t n = t->n;
    if (n < 2) return;
    if (!t->pow2)
    {
        double *buf = (double *)malloc(sizeof(double) * 4 * n);
        if (!buf) abort();
        memcpy(buf, a, sizeof(double) * 2 * n);
        fft_rec(t->tw, n, buf, 1, a, n, sign, buf + 2 * n);
        free(buf);
        return;
    }
    /* bit reversal permutation */
    for (size_t i = 1, j = 0; i < n; ++i)
    {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j)
        {
            double r = a[2 * i], m = a[2 * i + 1];
            a[2 * i] = a[2 * j];
            a[2 * i + 1] = a[2 * j + 1];
            a[2 * j] = r;
            a[2 * j + 1] = m;
        }
    }
    for (size_t len = 2; len <= n; len <<= 1)
    {
        const size_t half = len >> 1, step = n / len;
        for (size_t i = 0; i < n; i += len)
        {
            for (size_t k = 0; k < half; ++k)
            {
                const double wr = t->tw[2 * k * step];
                const double wi = sign < 0 ? t->tw[2 * k * step + 1] : -t->tw[2 * k * step + 1];
                double *u = a + 2 * (i + k), *v = a + 2 * (i + k + half);
                const double vr = v[0] * wr - v[1] * wi;
                const double vi = v[0] * wi + v[1] * wr;
                v[0] = u[0] - vr;
                v[1] = u[1] - vi;
                u[0] += vr;
                u[1] += vi;
            }
        }
    }
}

int oracle_fft_c2c_1d(double *a, size_t n, int sign)
{
    fft_tab t;
    if (fft_tab_init(&t, n)) return -1;
    fft_c2c_tab(&t, a, sign);
    fft_tab_free(&t);
    return 0;
}

/* 1-D r2c of length N (even), N/2+1 outputs, via one N/2-point complex FFT
 * of z[n] = x[2n] + i x[2n+1] and the split
 *   X[k] = Fe[k] + w_N^k Fo[k],  Fe = (Z[k]+conj Z[L-k])/2,
 *                                Fo = (Z[k]-conj Z[L-k])/(2i),  L = N/2.
 * `inout` holds N reals on entry and N/2+1 complex on exit (in place, as
 * FFTW's in-place r2c; needs N+2 doubles).  tab is for length N/2, wN holds
 * w_N^k for k < N/2. */
static void fft_r2c_inplace(const fft_tab *tab, const double *wN, double *inout, size_t N)
{
    const size_t L = N / 2;
    if (L == 0) return;
    fft_c2c_tab(tab, inout, -1); /* z -> Z, first L complex slots */
    const double z0r = inout[0], z0i = inout[1];
    for (size_t k = 1; k < L - k; ++k)
    {
        double *A = inout + 2 * k, *B = inout + 2 * (L - k);
        const double ar = A[0], ai = A[1], br = B[0], bi = -B[1]; /* B := conj Z[L-k] */
        const double fer = 0.5 * (ar + br), fei = 0.5 * (ai + bi);
        /* Fo = (A - B)/(2i) = -i (A-B)/2 */
        const double dr = 0.5 * (ar - br), di = 0.5 * (ai - bi);
        const double for_ = di, foi = -dr;
        const double wr = wN[2 * k], wi = wN[2 * k + 1];
        const double tr = wr * for_ - wi * foi, ti = wr * foi + wi * for_;
        /* X[k] = Fe + w Fo ; X[L-k] = conj(Fe - w Fo) */
        A[0] = fer + tr;
        A[1] = fei + ti;
        B[0] = fer - tr;
        B[1] = -(fei - ti);
    }
    if (L % 2 == 0 && L >= 2)
    {
        /* k = L/2 pairs with itself: X[L/2] = conj Z[L/2] */
        inout[2 * (L / 2) + 1] = -inout[2 * (L / 2) + 1];
    }
    inout[0] = z0r + z0i;
    inout[1] = 0.0;
    inout[2 * L] = z0r - z0i;
    inout[2 * L + 1] = 0.0;
}

int oracle_fft_r2c_1d(const double *in, double *out, size_t n)
{
    if (n < 2 || n % 2) return -1;
    fft_tab t;
    if (fft_tab_init(&t, n / 2)) return -1;
    double *wN = (double *)malloc(sizeof(double) * n);
    if (!wN) { fft_tab_free(&t); return -1; }
    for (size_t k = 0; k < n / 2; ++k) oracle_twiddle(k, n, &wN[2 * k], &wN[2 * k + 1]);
    memmove(out, in, sizeof(double) * n);
    fft_r2c_inplace(&t, wN, out, n);
    free(wN);
    fft_tab_free(&t);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Restatement of hpxfft::shared::loop (core/src/shared/loop.cpp).
 *   initialize (:158-189): dims, zero-filled transpose buffer
 *     (Ny/2+1) x 2Nx (:167), r2c plan of length Ny and forward c2c plan of
 *     length Nx (:172-180);
 *   fft_2d_r2c_par (:56-113): four phases, each a parallel loop with a
 *     barrier after it, timed into total / first_fftw / first_trans /
 *     second_fftw / second_trans (:106-110).
 * hpx::experimental::for_loop(par) becomes an OpenMP parallel for. */
typedef struct
{
    size_t nx, ny, ld, ncy; /* ncy = Ny/2+1 */
    fft_tab row_tab;        /* length Ny/2 */
    fft_tab col_tab;        /* length Nx */
    double *wN;             /* w_Ny^k, k < Ny/2 */
    double *trans;          /* ncy x 2*nx */
} oracle_loop;

void *oracle_loop_init(size_t nx, size_t ny, size_t ld)
{
    if (nx == 0 || ny < 2 || ny % 2 || ld < 2 * (ny / 2 + 1)) return NULL;
    oracle_loop *p = (oracle_loop *)calloc(1, sizeof(oracle_loop));
    if (!p) return NULL;
    p->nx = nx;
    p->ny = ny;
    p->ld = ld;
    p->ncy = ny / 2 + 1;
    if (fft_tab_init(&p->row_tab, ny / 2) || fft_tab_init(&p->col_tab, nx)) goto fail;
    p->wN = (double *)malloc(sizeof(double) * ny);
    p->trans = (double *)malloc(sizeof(double) * p->ncy * 2 * nx);
    if (!p->wN || !p->trans) goto fail;
    for (size_t k = 0; k < ny / 2; ++k) oracle_twiddle(k, ny, &p->wN[2 * k], &p->wN[2 * k + 1]);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)(p->ncy * 2 * nx); ++i) p->trans[i] = 0.0;
    return p;
fail:
    fft_tab_free(&p->row_tab);
    fft_tab_free(&p->col_tab);
    free(p->wN);
    free(p->trans);
    free(p);
    return NULL;
}

void oracle_loop_free(void *h)
{
    oracle_loop *p = (oracle_loop *)h;
    if (!p) return;
    fft_tab_free(&p->row_tab);
    fft_tab_free(&p->col_tab);
    free(p->wN);
    free(p->trans);
    free(p);
}

/* timings[5] = total, first_fftw, first_trans, second_fftw, second_trans (s).
 * nthreads <= 0 means "all available". */
int oracle_loop_exec(void *h, double *values, int nthreads, double *timings)
{
    oracle_loop *p = (oracle_loop *)h;
    if (!p || !values) return -1;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    const size_t nx = p->nx, ny = p->ny, ld = p->ld, ncy = p->ncy;
    double *trans = p->trans;
    const double t0 = omp_get_wtime();
    /* phase 1: 1-D r2c along y for every row (loop.cpp:61-69, :6-9) */
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)nx; ++i) fft_r2c_inplace(&p->row_tab, p->wN, values + (size_t)i * ld, ny);
    const double t1 = omp_get_wtime();
    /* phase 2: trans[index][j] = values[j][index] (loop.cpp:72-80, :18-25) */
#pragma omp parallel for schedule(static)
    for (long long index = 0; index < (long long)ncy; ++index)
    {
        double *dst = trans + (size_t)index * 2 * nx;
        for (size_t j = 0; j < nx; ++j)
        {
            dst[2 * j] = values[j * ld + 2 * index];
            dst[2 * j + 1] = values[j * ld + 2 * index + 1];
        }
    }
    const double t2 = omp_get_wtime();
    /* phase 3: 1-D c2c forward along x on each transposed row (loop.cpp:83-91) */
#pragma omp parallel for schedule(static)
    for (long long index = 0; index < (long long)ncy; ++index) fft_c2c_tab(&p->col_tab, trans + (size_t)index * 2 * nx, -1);
    const double t3 = omp_get_wtime();
    /* phase 4: values[j][index] = trans[index][j] (loop.cpp:94-101, :46-53) */
#pragma omp parallel for schedule(static)
    for (long long index = 0; index < (long long)ncy; ++index)
    {
        const double *src = trans + (size_t)index * 2 * nx;
        for (size_t j = 0; j < nx; ++j)
        {
            values[j * ld + 2 * index] = src[2 * j];
            values[j * ld + 2 * index + 1] = src[2 * j + 1];
        }
    }
    const double t4 = omp_get_wtime();
    if (timings)
    {
        timings[0] = t4 - t0;
        timings[1] = t1 - t0;
        timings[2] = t2 - t1;
        timings[3] = t3 - t2;
        timings[4] = t4 - t3;
    }
    return 0;
}

/* One-shot convenience wrapper. */
int oracle_r2c_2d(double *values, size_t nx, size_t ny, size_t ld, int nthreads)
{
    void *h = oracle_loop_init(nx, ny, ld);
    if (!h) return -1;
    int rc = oracle_loop_exec(h, values, nthreads, NULL);
    oracle_loop_free(h);
    return rc;
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------------------------------ */
/* Brute-force 2-D r2c DFT in long double, O((NxNy)^2) -- for tiny grids.
 * in: Nx x Ny reals with row pitch ld_in; out: Nx x (Ny/2+1) complex
 * interleaved, row pitch 2*(Ny/2+1) doubles.  Exponent index reduced mod
 * N exactly in integers before evaluating cos/sin. */
void oracle_dft_r2c_2d_brute(const double *in, size_t nx, size_t ny, size_t ld_in, double *out)
{
    const size_t ncy = ny / 2 + 1;
    const long double TWO_PI = 6.283185307179586476925286766559005768L;
#pragma omp parallel for schedule(dynamic)
    for (long long kx = 0; kx < (long long)nx; ++kx)
    {
        for (size_t ky = 0; ky < ncy; ++ky)
        {
            long double sr = 0.0L, si = 0.0L;
            for (size_t m = 0; m < nx; ++m)
            {
                for (size_t n = 0; n < ny; ++n)
                {
                    /* phase = kx*m/nx + ky*n/ny  (mod 1), exact rational */
                    const uint64_t num = ((uint64_t)kx * m % nx) * ny + ((uint64_t)ky * n % ny) * nx;
                    const uint64_t den = (uint64_t)nx * ny;
                    const long double ph = TWO_PI * (long double)(num % den) / (long double)den;
                    const long double v = in[m * ld_in + n];
                    sr += v * cosl(ph);
                    si -= v * sinl(ph);
                }
            }
            out[(size_t)kx * 2 * ncy + 2 * ky] = (double)sr;
            out[(size_t)kx * 2 * ncy + 2 * ky + 1] = (double)si;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Adjoint of the 2-D r2c map R : R^{Nx x Ny} -> C^{Nx x (Ny/2+1)} under the
 * real inner product <A,B> = Re sum conj(A) B over the half spectrum:
 *   (R* Y)[m][n] = Re sum_kx sum_{ky=0}^{Ny/2} Y[kx][ky] e^{+2 pi i (kx m/Nx + ky n/Ny)}.
 * This function has no counterpart in the reference (which only computes the
 * forward transform).  Brute force in long double for tiny grids.
 * Y: Nx x (Ny/2+1) complex, pitch 2*(Ny/2+1) doubles; out: Nx x Ny reals,
 * pitch ld_out. */
void oracle_adjoint_r2c_2d_brute(const double *Y, size_t nx, size_t ny, double *out, size_t ld_out)
{
    const size_t ncy = ny / 2 + 1;
    const long double TWO_PI = 6.283185307179586476925286766559005768L;
#pragma omp parallel for schedule(dynamic)
    for (long long m = 0; m < (long long)nx; ++m)
    {
        for (size_t n = 0; n < ny; ++n)
        {
            long double s = 0.0L;
            for (size_t kx = 0; kx < nx; ++kx)
            {
                for (size_t ky = 0; ky < ncy; ++ky)
                {
                    const uint64_t num = ((uint64_t)kx * m % nx) * ny + ((uint64_t)ky * n % ny) * nx;
                    const uint64_t den = (uint64_t)nx * ny;
                    const long double ph = TWO_PI * (long double)(num % den) / (long double)den;
                    const long double yr = Y[kx * 2 * ncy + 2 * ky], yi = Y[kx * 2 * ncy + 2 * ky + 1];
                    /* Re[(yr + i yi)(cos ph + i sin ph)] */
                    s += yr * cosl(ph) - yi * sinl(ph);
                }
            }
            out[(size_t)m * ld_out + n] = (double)s;
        }
    }
}

/* FP64 FFT-based adjoint: inverse (+i) c2c along x on every column, then,
 * per row, halve the interior bins 0<ky<Ny/2, keep only the real parts of
 * the DC and Nyquist bins, and run an unnormalised c2r (Hermitian
 * extension) of length Ny.  Y is not modified.  out pitch ld_out. */
int oracle_adjoint_r2c_2d(const double *Y, size_t nx, size_t ny, double *out, size_t ld_out, int nthreads)
{
    if (nx == 0 || ny < 2 || ny % 2) return -1;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    const size_t ncy = ny / 2 + 1;
    double *G = (double *)malloc(sizeof(double) * 2 * nx * ncy);
    if (!G) return -1;
    fft_tab ct, rt;
    if (fft_tab_init(&ct, nx)) { free(G); return -1; }
    if (fft_tab_init(&rt, ny)) { fft_tab_free(&ct); free(G); return -1; }
    int err = 0;
    /* column pass: G[:,ky] = IDFT_x(Y[:,ky]) (unnormalised, +i) */
#pragma omp parallel
    {
        double *col = (double *)malloc(sizeof(double) * 2 * nx);
        if (!col)
        {
#pragma omp atomic write
            err = 1;
        }
        else
        {
#pragma omp for schedule(static)
            for (long long ky = 0; ky < (long long)ncy; ++ky)
            {
                for (size_t m = 0; m < nx; ++m)
                {
                    col[2 * m] = Y[m * 2 * ncy + 2 * ky];
                    col[2 * m + 1] = Y[m * 2 * ncy + 2 * ky + 1];
                }
                fft_c2c_tab(&ct, col, +1);
                for (size_t m = 0; m < nx; ++m)
                {
                    G[m * 2 * ncy + 2 * ky] = col[2 * m];
                    G[m * 2 * ncy + 2 * ky + 1] = col[2 * m + 1];
                }
            }
            free(col);
        }
    }
    /* row pass: full Hermitian extension + complex inverse FFT of length Ny */
#pragma omp parallel
    {
        double *row = (double *)malloc(sizeof(double) * 2 * ny);
        if (!row)
        {
#pragma omp atomic write
            err = 1;
        }
        else
        {
#pragma omp for schedule(static)
            for (long long m = 0; m < (long long)nx; ++m)
            {
                const double *g = G + (size_t)m * 2 * ncy;
                for (size_t ky = 0; ky < ncy; ++ky)
                {
                    double hr = g[2 * ky], hi = g[2 * ky + 1];
                    if (ky == 0 || ky == ny / 2)
                        hi = 0.0;
                    else
                    {
                        hr *= 0.5;
                        hi *= 0.5;
                    }
                    row[2 * ky] = hr;
                    row[2 * ky + 1] = hi;
                    if (ky != 0 && ky != ny / 2)
                    {
                        row[2 * (ny - ky)] = hr;
                        row[2 * (ny - ky) + 1] = -hi;
                    }
                }
                fft_c2c_tab(&rt, row, +1);
                for (size_t n = 0; n < ny; ++n) out[(size_t)m * ld_out + n] = row[2 * n];
            }
            free(row);
        }
    }
    fft_tab_free(&ct);
    fft_tab_free(&rt);
    free(G);
    return err ? -1 : 0;
}
