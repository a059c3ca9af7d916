/*
 * hpxfft_hip.h -- C ABI of the MI355X (gfx950) 2-D real-to-complex FFT.
 *
 * Plain C: pointers, sizes and status codes; no C++ or torch types and no
 * exception ever crosses this boundary.  Every entry point below replaces a
 * member of HPX-FFT's shared-memory class surface (hpxfft::shared::loop,
 * core/include/hpxfft/shared/loop.hpp:15-57) or of its FFTW adapter
 * (core/include/hpxfft/util/adapter_fftw.hpp:9-79); the citation is on each
 * declaration.  The C++ drop-in class include/hpxfft/gpu/loop.hpp is a thin
 * header-only layer over these functions.
 *
 * Transform (FFTW_FORWARD sign, unnormalised, as FFTW's r2c used by the
 * reference, core/src/util/adapter_fftw.cpp:9,28-29):
 *     X[kx][ky] = sum_m sum_n x[m][n] exp(-2 pi i (kx m / Nx + ky n / Ny)),
 *     kx in [0,Nx), ky in [0, Ny/2].
 * Layouts: input row m of grid g at in[g*dist_in + m*ld_in + n] (reals);
 * output X[kx][ky] at out[g*dist_out + kx*ld_out + ky] (complex = 2 reals,
 * interleaved re,im).  In place (out == in) needs ld_in == 2*ld_out and
 * dist_in == 2*dist_out (the reference's padded vector_2d(Nx, 2(Ny/2+1))).
 *
 * Threading: a plan is bound to one HIP stream; calls on one plan must be
 * serialised by the caller; distinct plans are independent.
 */
#ifndef HPXFFT_HIP_H_INCLUDED
#define HPXFFT_HIP_H_INCLUDED

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The reference reports errors only by exceptions
 * (std::invalid_argument, adapter_fftw.hpp:42; std::runtime_error,
 * loop.cpp:200); the C++ layer maps these codes back to those types. */
typedef enum
{
    HPXFFT_OK = 0,
    HPXFFT_INVALID_ARG = 1,      /* bad pointer, flag, layout, or call order */
    HPXFFT_UNSUPPORTED_SIZE = 2, /* size outside what the kernels cover: powers of two (Nx <= 2^16,
                                    Ny <= 2^15), or for F64/F32 plans any Nx and even Ny whose
                                    lengths (Nx, Ny/2) have prime factors <= 61 within 4096 (F64) /
                                    8192 (F32) points (prime factors <= 7: within 8192 / 16384
                                    points), or otherwise up to 4096 (F64) / 8192 (F32)
                                    points by Bluestein's algorithm, or up to 2^20 points as a
                                    product of two lengths each within 4096 / 8192 points with
                                    such prime factors or within 2048 / 4096 points by Bluestein
                                    (F16/BF16 storage: the one-level lengths, forward) */
    HPXFFT_NO_DEVICE = 3,        /* no HIP device visible */
    HPXFFT_OUT_OF_MEMORY = 4,
    HPXFFT_HIP_ERROR = 5,
    HPXFFT_IO_ERROR = 6
} hpxfft_status;

/* Arithmetic / storage of a plan. */
typedef enum
{
    HPXFFT_F64 = 0,          /* FP64 storage and arithmetic (the reference's `real` = double, loop.hpp:9) */
    HPXFFT_F32 = 1,          /* FP32 storage and arithmetic, twiddles generated in FP64 then rounded */
    HPXFFT_F16_STORAGE = 2,  /* IEEE half storage, FP32 arithmetic, power-of-two scale per row of each pass */
    HPXFFT_BF16_STORAGE = 3  /* bfloat16 storage, FP32 arithmetic, same scaling */
} hpxfft_precision;

/* Planning effort, the image of the FFTW plan-flag strings (wider searches,
 * as FFTW_MEASURE < FFTW_PATIENT < FFTW_EXHAUSTIVE, adapter_fftw.hpp:13-18). */
typedef enum
{
    HPXFFT_ESTIMATE = 0,   /* no timing: balanced four-step split, radix-16 row kernel, in-place general stages */
    HPXFFT_MEASURE = 1,    /* time the four-step splits, the general kernels' stage forms and one/two levels,
                              the batched task shapes; keep the fastest (3 timed runs per candidate) */
    HPXFFT_PATIENT = 2,    /* measure's search, plus the row kernels of 2^13- / 2^14-point rows (FP32
                              arithmetic); 5 timed runs per candidate */
    HPXFFT_EXHAUSTIVE = 3  /* patient's search with 10 timed runs per candidate */
} hpxfft_effort;

typedef struct hpxfft_plan_s *hpxfft_plan;

/* Replaces hpxfft::util::fftw_adapter::string_to_fftw_plan_flag
 * (core/include/hpxfft/util/adapter_fftw.hpp:22-44): "estimate", "measure",
 * "patient", "exhaustive" -> effort; anything else -> HPXFFT_INVALID_ARG with
 * last_error "Invalid FFTW plan flag string". */
int hpxfft_plan_flag_from_string(const char *flag, int *effort);

/* Replaces hpxfft::shared::loop::initialize (core/src/shared/loop.cpp:158-189)
 * for one in-place grid: Nx rows, Ny reals per row, row pitch ld_in reals
 * (the reference derives Nx = n_row, Ny = n_col - 2 from the vector_2d,
 * loop.cpp:163-165).  Output pitch is Ny/2+1 complex; exec may be in place
 * when ld_in == 2*(Ny/2+1).  `stream` is a hipStream_t (NULL = default). */
int hpxfft_plan_r2c_2d(hpxfft_plan *plan, size_t nx, size_t ny, size_t ld_in, int precision, int effort,
                       void *stream);

/* Same as hpxfft_plan_r2c_2d for `batch` grids with explicit layouts (no
 * reference counterpart for the batch; used by the torch binding). */
int hpxfft_plan_r2c_2d_many(hpxfft_plan *plan, size_t nx, size_t ny, size_t batch, size_t ld_in, size_t dist_in,
                            size_t ld_out, size_t dist_out, int precision, int effort, void *stream);

/* Replaces the body of hpxfft::shared::loop::fft_2d_r2c_par / _seq
 * (core/src/shared/loop.cpp:56-155): row r2c, transpose, column c2c,
 * transpose back, on device memory, enqueued on the plan's stream (no host
 * synchronisation).  d_in/d_out are device pointers of the plan's storage
 * type (double, float, _Float16 or bfloat16 bit patterns), each aligned to
 * one complex element (16, 8 or 4 bytes; HPXFFT_INVALID_ARG otherwise, as
 * for every exec function below). */
int hpxfft_exec_r2c_2d(hpxfft_plan plan, const void *d_in, void *d_out);

/* Allocation-free form of hpxfft_exec_r2c_2d.  It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * One grid, Nx x Ny reals at row pitch ld_in in, Nx x (Ny/2+1) complex at
 * row pitch ld_out out (in place when d_in == d_out and ld_in == 2*ld_out),
 * F64 or F32, both sides powers of two (Nx <= 65536, 2 <= Ny <= 32768),
 * planned as hpxfft_plan_r2c_2d at "estimate" on every call and enqueued on
 * `stream` (a hipStream_t, NULL = default) with no plan object: the results
 * equal that plan's bit for bit.  No hipMalloc, hipFree or other allocation
 * happens.  A column pass that runs as a four-step (Nx > 1024) keeps its
 * intermediate in a static device array of 128 MiB (134217728 bytes) per
 * device, held until the process ends; a grid whose intermediate,
 * Nx x (Ny/2+1) complex values of the precision (rounded up to whole strips
 * of 32 FP64 or 64 FP32 columns), exceeds it is refused with
 * HPXFFT_UNSUPPORTED_SIZE.  Every call on a device shares that device's array:
 * calls that overlap on one device (other streams, other threads) race.  The
 * first call on a device finds the array's address and, unless a plan has
 * done it already, uploads the twiddle tables: it must not be captured
 * (HPXFFT_INVALID_ARG if it is); later calls may be. */
int hpxfft_exec_r2c_2d_noalloc(size_t nx, size_t ny, size_t ld_in, size_t ld_out, int precision, const void *d_in,
                               void *d_out, void *stream);

/* 16-bit storage plans: value = stored * 2^-exp.  d_in_row_exp (optional,
 * Nx ints per grid) scales the input rows; d_out_col_exp (required, Ny/2+1
 * ints per grid) receives the per-column exponent of the output (the second
 * pass works on columns, i.e. on the rows of the reference's transposed
 * buffer, loop.cpp:167). */
int hpxfft_exec_r2c_2d_scaled(hpxfft_plan plan, const void *d_in, void *d_out, const int *d_in_row_exp,
                              int *d_out_col_exp);

/* Drop-in host path for fft_2d_r2c_par (loop.cpp:56-113) on the reference's
 * in-place padded buffer: copies `inout` (Nx x ld_in doubles) to the device,
 * transforms, copies back and synchronises.  F64 plans only.  Adds the
 * measurement keys "h2d" and "d2h".  The first call allocates the plan's
 * device copy of the buffer. */
int hpxfft_exec_r2c_2d_host(hpxfft_plan plan, double *inout);

/* Adjoint of the r2c map under Re<A,B> on the half spectrum (no reference
 * counterpart):  (R* Y)[m][n] = Re sum_kx sum_{ky<=Ny/2} Y[kx][ky]
 * exp(+2 pi i (kx m/Nx + ky n/Ny)).  d_spec uses the plan's output layout
 * and is not modified; d_real uses the plan's input layout.  F64/F32 plans.
 * The first backward call on a plan (this or hpxfft_exec_c2r_2d) allocates
 * its work buffer, one spectrum; later calls allocate nothing.
 *
 * Every exec function except the host path and the batched one enqueues
 * only kernels and event records on the plan's stream, so a caller may
 * capture it into a HIP graph once that first allocation is done (make one
 * backward call before capturing a backward one). */
int hpxfft_exec_r2c_2d_adjoint(hpxfft_plan plan, const void *d_spec, void *d_real);

/* Inverse 2-D c2r, unnormalised, in FFTW's FFTW_BACKWARD convention: the
 * reference declares the backward direction (core/include/hpxfft/util/
 * adapter_fftw.hpp:11) but never plans it.  x[m][n] = sum_kx sum_ky
 * X[kx][ky] exp(+2 pi i (kx m/Nx + ky n/Ny)) over the Hermitian extension of
 * the half spectrum, the imaginary parts of the ky = 0 and ky = Ny/2 bins
 * ignored after the column transform (as numpy.fft.irfft2 times Nx*Ny), so
 * c2r(r2c(x)) = Nx*Ny*x.  d_spec uses the plan's output layout and is not
 * modified; d_real uses the plan's input layout.  F64/F32 plans. */
int hpxfft_exec_c2r_2d(hpxfft_plan plan, const void *d_spec, void *d_real);

/* Replaces hpxfft::shared::loop::get_measurement (loop.cpp:192): seconds of
 * the last exec for "total", "first_fftw", "first_trans", "second_fftw",
 * "second_trans" (the transposes are fused into the column pass and read
 * 0), "h2d", "d2h"; "plan" (planning seconds) and "plan_flops"; and, when
 * the column pass runs as two launches (Nx > 1024, or > 512 for 16-bit
 * storage), "column_a" and "column_b" (seconds of each launch of the last
 * exec; 0 otherwise).  The backward execs (adjoint, c2r) run the column
 * pass first: for them "first_fftw" is the column pass and "second_fftw"
 * the row pass; batched plans give their row and column launches.  Unknown keys
 * give 0.0, as std::map::operator[] does in the reference.  Waits for the
 * last exec to finish. */
int hpxfft_get_measurement(hpxfft_plan plan, const char *key, double *value);

/* Replaces hpxfft::shared::loop::write_plans_to_file (loop.cpp:194-212):
 * appends a description of the chosen kernels to `path`; HPXFFT_IO_ERROR
 * with last_error "Failed to open file: <path>" if it cannot be opened. */
int hpxfft_write_plan(hpxfft_plan plan, const char *path);

/* Plan text (same content as hpxfft_write_plan) into buf (NUL-terminated,
 * truncated to len).  Returns the full length in *needed if non-NULL. */
int hpxfft_plan_describe(hpxfft_plan plan, char *buf, size_t len, size_t *needed);

/* Replaces ~loop (loop.hpp:30): frees device buffers and the stream state. */
int hpxfft_plan_destroy(hpxfft_plan plan);

/* Batched uneven grids (no reference counterpart): `count` independent
 * in-place grids, grid i of nx[i] x ny[i] in the padded layout
 * (ld = 2*(ny[i]/2+1) reals), all transformed by one persistent launch.
 * F32 only. */
int hpxfft_plan_r2c_2d_batched(hpxfft_plan *plan, size_t count, const size_t *nx, const size_t *ny, int precision,
                               int effort, void *stream);

/* d_grids: host array of `count` device pointers, one per grid (in place).
 * When they differ from the previous exec's, their descriptors are uploaded
 * by a host-to-device copy on the plan's stream (two pinned staging slots
 * used in turn; the host waits only if the copy made two calls before has
 * not run yet).  With the same pointers as the previous exec, nothing but
 * kernels and event records is enqueued, so such a call may be captured
 * into a HIP graph; capturing a call with a new pointer set gives
 * HPXFFT_INVALID_ARG. */
int hpxfft_exec_r2c_2d_batched(hpxfft_plan plan, void *const *d_grids);

/* No reference counterpart: overrides the task shape a batched plan runs in
 * each phase (row_cfg, col_cfg in 0..3: the candidates its autotuner times;
 * the plan text lists them).  HPXFFT_INVALID_ARG for another value or a plan
 * that is not batched. */
int hpxfft_plan_batched_set_config(hpxfft_plan plan, int row_cfg, int col_cfg);

/* No reference counterpart (FFTW's planner choice is internal to
 * fftw_plan_dft_1d, core/src/util/adapter_fftw.cpp:28-29): runs the plan's
 * four-step column pass with the split Nx = P * Q, P = 2^log2_p, instead of
 * the one it chose (estimate: the balanced split; measure and above: the
 * fastest one timed).  log2_p must lie in the range of splits the kernels
 * cover (sub-lengths 2^5 .. 2^8).  HPXFFT_INVALID_ARG for another value or
 * for a plan whose column pass is not a four-step (the plan text says which
 * it is).  Lets a test check every split the measure effort may keep. */
int hpxfft_plan_set_column_split(hpxfft_plan plan, int log2_p);

/* No reference counterpart: selects the row kernel of a plan whose rows are
 * 2^13 or 2^14 complex points (Ny = 16384 or 32768, powers of two) with FP32
 * arithmetic (FP32, FP16 or BF16 storage): 0 = radix-16 Stockham stages,
 * three LDS exchanges per row; 1 = radix-32 stages whose butterflies span a
 * lane pair (the pair exchanges values by DPP) and a radix-8/16 stage fused
 * with the r2c split, two LDS exchanges per row.  HPXFFT_INVALID_ARG for
 * another kind, or kind 1 on a plan with other rows or FP64 arithmetic. */
int hpxfft_plan_set_row_kernel(hpxfft_plan plan, int kind);

/* Message of the last failing call on this thread ("" if none). */
const char *hpxfft_last_error(void);

/* Library version string. */
const char *hpxfft_version(void);

#ifdef __cplusplus
}
#endif

#endif /* HPXFFT_HIP_H_INCLUDED */
