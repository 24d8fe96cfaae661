/*
 * nrfft.h -- C ABI of the MI355X (gfx950) FFT library for the numrs FFT path.
 *
 * Conventions of the numrs routines this library replaces:
 *   - complex data are interleaved re,im; lengths are powers of two;
 *   - isign = +1 is the kernel exp(+2 pi i jk/N), isign = -1 exp(-2 pi i jk/N);
 *   - inverses are unnormalised: four1 forward then inverse gives N*x,
 *     realft forward then inverse gives (n/2)*x.
 * Every entry point is extern "C", takes plain pointers and sizes, returns an
 * nrfft_status (0 = success, negative = error) and does not throw.
 *
 * Two groups of entry points:
 *   1. Host-pointer calls with the reference signatures (drop-in): data live
 *      in host memory; the call stages them through HBM, runs the HIP kernels
 *      and returns when the result is back in host memory.
 *   2. Device-pointer plans: data already resident in HBM, execution enqueued
 *      on the caller's HIP stream (passed as void*, a hipStream_t; NULL = the
 *      default stream).  A plan is not to be executed concurrently on two
 *      streams (it owns one workspace).
 * All calls are thread-safe.
 */
#ifndef NRFFT_H
#define NRFFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum nrfft_status {
    NRFFT_OK = 0,
    NRFFT_E_ISIGN = -1,          /* isign not +1/-1 (ConvlvError::InvalidIsign, Fourn.rs:371-373) */
    NRFFT_E_DIMS = -2,           /* ndim == 0, ndim > len(nn), or a dimension <= 1 (Fourn.rs:368-376) */
    NRFFT_E_NOT_POW2 = -3,       /* a length that is not a power of two */
    NRFFT_E_LEN = -4,            /* length mismatch (CorrelError::LengthMismatch, Correlation.rs:14-16) */
    NRFFT_E_EMPTY = -5,          /* empty input (ConvlvError/CorrelError::EmptyInput) */
    NRFFT_E_RESP_TOO_LONG = -6,  /* m > n (ConvlvError::ResponseTooLong, Convolve.rs:16-18) */
    NRFFT_E_HIP = -7,            /* a HIP runtime call failed (see nrfft_last_error) */
    NRFFT_E_NOMEM = -8,          /* device or host allocation failed */
    NRFFT_E_UNSUPPORTED = -9,    /* size beyond what this build plans for */
    NRFFT_E_INVALID = -10        /* null pointer / bad plan / bad argument */
} nrfft_status;

typedef enum nrfft_dtype { NRFFT_F32 = 0, NRFFT_F64 = 1 } nrfft_dtype;

/* ---- library ------------------------------------------------------------ */
const char* nrfft_version(void);
const char* nrfft_status_string(int status);
/* Message of the last failing call on this thread ("" if none). */
const char* nrfft_last_error(void);
/* Select the HIP device used by this thread and build its twiddle tables
 * (FP64, generated on the GPU).  Optional: every call initialises lazily. */
int nrfft_init(int device);

/* ---- 1. host-pointer drop-in calls ------------------------------------- */

/* four1: FFT_1.rs:5 `pub fn four1(data: &mut [f64], nn: usize, isign: i32)`.
 * In place on 2*nn values.  nn must be a power of two (1 allowed). */
int nrfft_four1_f64(double* data, size_t nn, int isign);
int nrfft_four1_f32(float* data, size_t nn, int isign);

/* FFTProcessor::fft_batch (FFT_1.rs:185-189) for `batch` transforms of equal
 * length nn stored back to back (2*nn*batch values). */
int nrfft_four1_batch_f64(double* data, size_t nn, size_t batch, int isign);
int nrfft_four1_batch_f32(float* data, size_t nn, size_t batch, int isign);

/* realft: Real_FT.rs:4 `pub fn realft(data: &mut [f64], n: usize, isign: i32)`.
 * In place on n values.  isign=+1: n reals -> packed spectrum
 * [F_0, F_{n/2}, Re F_1, Im F_1, ...], F_k = sum_j x_j e^{+2 pi i jk/n};
 * isign=-1: packed spectrum -> (n/2) x.  n must be a power of two >= 2. */
int nrfft_realft_f64(double* data, size_t n, int isign);
int nrfft_realft_f32(float* data, size_t n, int isign);

/* RealFTProcessor::process_batch (Real_FT.rs:365-369) for `count` host
 * buffers of n reals each (equal lengths, one isign), transformed in place by
 * one batched execution.  Same semantics and checks as nrfft_realft_*. */
int nrfft_realft_batch_f64(double* const* data, size_t n, size_t count, int isign);
int nrfft_realft_batch_f32(float* const* data, size_t n, size_t count, int isign);

/* fourn: the in-memory NR fourn that Real_FT3.rs:35 calls as
 * Fourn(&mut data, &nn, ndim, isign) (no such function exists in the
 * reference; Fourn.rs:56 is a data-less stub).  C order: nn[0] slowest.
 * Validation follows Fourn.rs:367-378; dimensions must be powers of two. */
int nrfft_fourn_f64(double* data, const size_t* nn, size_t ndim, int isign);
int nrfft_fourn_f32(float* data, const size_t* nn, size_t ndim, int isign);

/* convlv: Convolve.rs:8 `pub fn convlv(data: &[f64], respns: &[f64], isign: i32)
 * -> Result<Array1<f64>, ConvlvError>`.  Writes n values to ans.
 * isign=+1 convolution, -1 deconvolution (bins with |R|^2 < 1e-12 -> 0).
 * The response is placed as Convolve.rs:41-63 places it. */
int nrfft_convlv_f64(const double* data, size_t n, const double* respns, size_t m, int isign, double* ans);
int nrfft_convlv_f32(const float* data, size_t n, const float* respns, size_t m, int isign, float* ans);

/* convlv_batch (Convolve.rs:241-250), extended to one response per signal:
 * signal b is data[b*n .. b*n+n), its response respns[b*m .. b*m+m). */
int nrfft_convlv_batch_f64(const double* data, size_t n, const double* respns, size_t m, size_t batch,
                           int isign, double* ans);
int nrfft_convlv_batch_f32(const float* data, size_t n, const float* respns, size_t m, size_t batch,
                           int isign, float* ans);

/* correl: Correlation.rs:8 `pub fn correl(data1: &[f64], data2: &[f64])
 * -> Result<Array1<f64>, CorrelError>`.  n <= 32: linear correlation at lags
 * 0..n-1 (Correlation.rs:37-50); n > 32 (power of two): circular correlation,
 * negative lags wrapped to the top (NR correl). */
int nrfft_correl_f64(const double* data1, const double* data2, size_t n, double* ans);
int nrfft_correl_f32(const float* data1, const float* data2, size_t n, float* ans);

/* ---- 2. device-pointer plans ------------------------------------------ */
typedef struct nrfft_plan_s* nrfft_plan;

/* C2C: `batch` transforms of shape nn[0..ndim) (C order) stored back to back.
 * ndim == 1 is a batched four1, ndim >= 2 a batched fourn. */
int nrfft_plan_c2c(nrfft_plan* plan, int dtype, const size_t* nn, size_t ndim, size_t batch);
/* Out of place or in place (d_out == d_in). */
int nrfft_execute_c2c(nrfft_plan plan, const void* d_in, void* d_out, int isign, void* stream);

/* Split-precision form of an FP32 C2C plan (no counterpart in the reference):
 * radix-16 stages as 16x16 complex DFT-matrix products on MFMA with FP16
 * hi/lo (NRFFT_SPLIT_F16, 3 products) or BF16 hi/mid/lo (NRFFT_SPLIT_BF16,
 * 6 products) operands and FP32 accumulation; twiddles in FP32 from FP64
 * tables.  Passes whose length has no MFMA kernel run the plain FP32 kernels.
 * Executed with nrfft_execute_c2c. */
enum { NRFFT_SPLIT_F16 = 1, NRFFT_SPLIT_BF16 = 2 };
int nrfft_plan_c2c_split(nrfft_plan* plan, const size_t* nn, size_t ndim, size_t batch, int prec);

/* realft on `batch` signals of n reals, in place. */
int nrfft_plan_realft(nrfft_plan* plan, int dtype, size_t n, size_t batch);
int nrfft_execute_realft(nrfft_plan plan, void* d_data, int isign, void* stream);

/* convlv over `batch` signals of length n, one response of length m per
 * signal (d_respns holds batch*m values).  Fused R2C -> spectrum
 * multiply/divide -> C2R in one launch when n/2 fits one work-group. */
int nrfft_plan_convlv(nrfft_plan* plan, int dtype, size_t n, size_t m, size_t batch);
int nrfft_execute_convlv(nrfft_plan plan, const void* d_data, const void* d_respns, void* d_ans, int isign,
                         void* stream);

/* correl over `batch` pairs of length n. */
int nrfft_plan_correl(nrfft_plan* plan, int dtype, size_t n, size_t batch);
int nrfft_execute_correl(nrfft_plan plan, const void* d_data1, const void* d_data2, void* d_ans, void* stream);

/* Allocation-free batched four1 on device pointers.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same transform, kernels and default tunings as nrfft_plan_c2c(dtype, {nn},
 * 1, batch) + nrfft_execute_c2c, but no plan: the four-step intermediate of
 * nn > 8192 lives in a 64 MiB device array built into the library (one per
 * GPU, held until the process ends), so nn * sizeof(complex) must not exceed
 * 64 MiB (FP32 nn <= 2^23, FP64 nn <= 2^22; larger -> NRFFT_E_UNSUPPORTED).
 * nrfft_init(device) must have run on the current device (else
 * NRFFT_E_INVALID); the call then calls no allocating routine.  Calls that
 * overlap in time on one GPU share that array and race; calls ordered on one
 * stream do not.  In place allowed (d_out == d_in). */
int nrfft_execute_four1_batch_noalloc(int dtype, const void* d_in, void* d_out, size_t nn, size_t batch, int isign,
                                      void* stream);

/* Profiling aid (no counterpart in the reference): enqueue only launches
 * [first, first + count) of a C2C plan, e.g. one pass of a four-step plan. */
int nrfft_execute_c2c_launches(nrfft_plan plan, const void* d_in, void* d_out, int isign, void* stream, int first,
                               int count);

/* ---- twofft / cosine transforms / 3-D real FFT ------------------------------
 * Numerical Recipes semantics with 0-based storage (the Rust sources are 1-based
 * NR translations; the divergences are listed in the oracle).  Host-pointer
 * calls stage through HBM like the ones above. */
/* FFT_2.rs:3 twofft(data1, data2, fft1, fft2): the isign = +1 spectra of two
 * real sequences of n points from one complex FFT.  fft1 / fft2 hold 2n + 2
 * values (FFT_2.rs:6-7): bins 0..n-1 interleaved in the first 2n, then 0, 0. */
int nrfft_twofft_f64(const double* data1, const double* data2, size_t n, double* fft1, double* fft2);
int nrfft_twofft_f32(const float* data1, const float* data2, size_t n, float* fft1, float* fft2);
/* Cos_FT.rs:7 cosft1(y, n): DCT-I in place on y[0..n] (n + 1 values):
 * Y_k = (y_0 + (-1)^k y_n)/2 + sum_{j=1}^{n-1} y_j cos(pi j k / n); its own
 * inverse up to n/2. */
int nrfft_cosft1_f64(double* y, size_t n);
int nrfft_cosft1_f32(float* y, size_t n);
/* Cos_FT2.rs:7 cosft2(y, n, isign): in place on y[0..n-1]; isign = 1:
 * Y_k = sum_j y_j cos(pi k (j + 1/2) / n); isign = -1: the inverse
 * y_j = Y_0/2 + sum_{k>=1} Y_k cos(pi k (j + 1/2) / n) (= (n/2) x the input of
 * the forward call). */
int nrfft_cosft2_f64(double* y, size_t n, int isign);
int nrfft_cosft2_f32(float* y, size_t n, int isign);
/* Real_FT3.rs:8 rlft3(data, speq, nn1, nn2, nn3, isign): 3-D real FFT, C order.
 * isign = 1: data (nn1 x nn2 x nn3 reals) becomes the complex half spectrum
 * F[k1][k2][k3], k3 < nn3/2, and speq (nn1 x 2 nn2 reals) F[k1][k2][nn3/2];
 * isign = -1 inverts (data <- (nn1 nn2 nn3 / 2) x).  nn1 = 1 gives the 2-D
 * transform. */
int nrfft_rlft3_f64(double* data, double* speq, size_t nn1, size_t nn2, size_t nn3, int isign);
int nrfft_rlft3_f32(float* data, float* speq, size_t nn1, size_t nn2, size_t nn3, int isign);
/* Four_FS.rs:55 FourFS::fourfs(nn, ndim, isign) -- the out-of-core n-D FFT:
 * in place on a host array of prod(nn) interleaved complex values (C order;
 * an mmap-ed file works), streamed through HBM with at most device_bytes of
 * device memory (0: 1 GiB).  Results equal fourn's.  An array that fits the
 * budget makes one round trip; a larger one streams axis by axis in chunks
 * through pinned host staging.  Strided axes up to 8192 points; a 1-D
 * transform larger than half the budget runs a four-step over
 * the host array and needs `scratch` (same size as data), else scratch may
 * be NULL.  Errors follow Four_FS.rs:328-339 (NRFFT_E_DIMS "Invalid
 * dimensions" / "Invalid dimension size", NRFFT_E_ISIGN). */
int nrfft_fourfs_f64(double* data, double* scratch, const size_t* nn, size_t ndim, int isign, size_t device_bytes);
int nrfft_fourfs_f32(float* data, float* scratch, const size_t* nn, size_t ndim, int isign, size_t device_bytes);
/* Device plans of the same transforms (batch back to back; twofft outputs
 * are 2n values per transform, cosft1 n + 1, cosft2 n, rlft3 speq
 * 2 nn1 nn2).  Outputs must not overlap inputs except in/out of cosft1/2
 * (in place allowed) and rlft3 (in place by definition). */
int nrfft_plan_twofft(nrfft_plan* plan, int dtype, size_t n, size_t batch);
int nrfft_execute_twofft(nrfft_plan plan, const void* d_data1, const void* d_data2, void* d_fft1, void* d_fft2,
                         void* stream);
int nrfft_plan_cosft1(nrfft_plan* plan, int dtype, size_t n, size_t batch);
int nrfft_execute_cosft1(nrfft_plan plan, const void* d_in, void* d_out, void* stream);
int nrfft_plan_cosft2(nrfft_plan* plan, int dtype, size_t n, size_t batch);
int nrfft_execute_cosft2(nrfft_plan plan, const void* d_in, void* d_out, int isign, void* stream);
int nrfft_plan_rlft3(nrfft_plan* plan, int dtype, size_t nn1, size_t nn2, size_t nn3, size_t batch);
int nrfft_execute_rlft3(nrfft_plan plan, void* d_data, void* d_speq, int isign, void* stream);

/* ---- uneven batches -------------------------------------------------------
 * Many transforms of different lengths and dimensionality in one persistent
 * work-queue kernel (two dispatches, one per pass phase, replayed as a
 * captured hipGraph).  Item i has ndim[i] (1 or 2) and dims dims[2 i]
 * (, dims[2 i + 1]) in C order, powers of two >= 64 (2-D sides up to 8192 in
 * FP32 / 4096 in FP64; 1-D up to 2^24 in FP32 / 2^22 in FP64), and lives at
 * complex-element offset offs[i] of both d_in and d_out (in place allowed).
 * FFTProcessor::fft_batch (FFT_1.rs:185-189) is the reference counterpart
 * for 1-D items. */
typedef struct nrfft_uneven_s* nrfft_uneven;
int nrfft_plan_uneven(nrfft_uneven* plan, int dtype, size_t nitems, const int* ndim, const size_t* dims,
                      const size_t* offs);
int nrfft_execute_uneven(nrfft_uneven plan, const void* d_in, void* d_out, int isign, void* stream);
/* Same work as plain launches (reset, phase 1, phase 2) instead of a graph replay. */
int nrfft_execute_uneven_nograph(nrfft_uneven plan, const void* d_in, void* d_out, int isign, void* stream);
/* Waits for the stream and reports NRFFT_E_HIP if the last execution left tasks untaken. */
int nrfft_uneven_check(nrfft_uneven plan, void* stream);
int nrfft_uneven_info(nrfft_uneven plan, size_t* tasks_phase1, size_t* tasks_phase2, size_t* workspace_bytes,
                      unsigned* groups);
int nrfft_uneven_destroy(nrfft_uneven plan);
/* Autotuner of the uneven-batch size classes: for every 1-D length 2^k,
 * log2n_min <= k <= log2n_max, that takes the four-step path, times each
 * legal split N = 2^l1 x 2^(k - l1) on a batch of that class (iters
 * executions) and caches the fastest per (dtype, N) with the C2C tunings
 * (nrfft_tune_cache_save/load persist both).  Plans created afterwards use it.
 * best_l1[k - log2n_min] / best_ms[...] receive the choice and its time (0 for
 * single-pass classes); either may be NULL. */
int nrfft_autotune_uneven(int dtype, int log2n_min, int log2n_max, int iters, int* best_l1, float* best_ms);

/* FFTProcessor::fft_batch (FFT_1.rs:185-189): `count` host buffers, buffer i
 * holding 2*nn[i] interleaved values, transformed in place in one launch. */
int nrfft_fft_batch_f64(double* const* data, const size_t* nn, size_t count, int isign);
int nrfft_fft_batch_f32(float* const* data, const size_t* nn, size_t count, int isign);

/* Seeded synthetic inputs generated in HBM (bench and test utility): n draws
 * of uniform[-1,1) from SplitMix64(seed), starting at draw `start`, the same
 * stream as numrs_amd/synth.py; FP32 values are the FP64 draws rounded. */
int nrfft_gen_uniform_f32(void* d, size_t n, uint64_t seed, uint64_t start, void* stream);
int nrfft_gen_uniform_f64(void* d, size_t n, uint64_t seed, uint64_t start, void* stream);

/* ---- autotuner (no counterpart in the reference) -------------------------
 * A C2C plan's tunables -- tune[0] log2 of the four-step column length,
 * tune[1] / tune[2] log2 tile widths of the column / transpose passes (which
 * set threads, LDS and so work-groups per CU), tune[3] transforms per batch
 * chunk (0 = whole batch), tune[4] resident work-groups per CU of the
 * persistent passes, tune[5] XCD-contiguous tile order (0/1) -- are cached per
 * (shape, batch, dtype).  best_tune receives 6 values.
 * nrfft_autotune_c2c times every candidate on the current device with HIP
 * events (iters executions each) and caches the fastest; plans created
 * afterwards for that key use it.  -1 leaves an entry at the planner default.
 * nrfft_plan_c2c_tuned builds one plan with the given tunables (ntune <= 7;
 * tune[6], if given, is 0 for the plain form or NRFFT_SPLIT_F16 /
 * NRFFT_SPLIT_BF16 for the split-precision form of an FP32 plan); it neither
 * reads nor changes the cache.  Split-precision plans never use cached
 * tunings (their split must leave a pass of 2^8..2^10 points). */
int nrfft_plan_c2c_tuned(nrfft_plan* plan, int dtype, const size_t* nn, size_t ndim, size_t batch, const int* tune,
                         size_t ntune);
int nrfft_autotune_c2c(int dtype, const size_t* nn, size_t ndim, size_t batch, int iters, int* best_tune,
                       double* best_ms);
/* The tunables a plan was built with (up to 6 values, order as above; -1 =
 * planner default). */
int nrfft_plan_tune(nrfft_plan plan, int* tune, size_t ntune);
/* Text cache file; also read at the first plan creation from $NRFFT_TUNE_CACHE. */
int nrfft_tune_cache_save(const char* path);
int nrfft_tune_cache_load(const char* path);

int nrfft_plan_destroy(nrfft_plan plan);
/* Bytes of HBM workspace the plan owns. */
size_t nrfft_plan_workspace_bytes(nrfft_plan plan);
/* Number of kernel launches one execution enqueues. */
int nrfft_plan_num_launches(nrfft_plan plan);

#ifdef __cplusplus
}
#endif

#endif /* NRFFT_H */
