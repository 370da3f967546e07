/*
 * fcwt_hip.h -- C ABI of the MI355X (gfx950) implementation of fCWT's
 * FCWT::cwt hot path (Morlet wavelet, real float32 input).
 *
 * Plain C types only: pointers, sizes, status codes.  Every entry point
 * returns FCWT_HIP_OK (0) or a negative status; fcwt_hip_last_error() then
 * holds a message (thread-local).  Nothing throws across this boundary.
 *
 * Each declaration names the reference interface it replaces
 * (fastlib/fCWT, src/fcwt/...).  A plan is not safe for concurrent use from
 * several threads (the reference's FCWT object is not either: fcwt.cpp:450
 * mutates this->size and the wavelet table during cwt).
 */
#ifndef FCWT_HIP_H
#define FCWT_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define FCWT_HIP_OK 0
#define FCWT_HIP_EINVAL (-1)      /* invalid argument */
#define FCWT_HIP_ENOMEM (-2)      /* host or device allocation failed */
#define FCWT_HIP_EDEVICE (-3)     /* HIP runtime error, or no usable gfx950 device */
#define FCWT_HIP_EUNSUPPORTED (-4) /* size outside what this build handles */
#define FCWT_HIP_EIO (-5)          /* plan-cache file could not be read or written */

/* Scale types; same values as enum SCALETYPE, src/fcwt/fcwt.h:65. */
#define FCWT_HIP_LINSCALES 0
#define FCWT_HIP_LOGSCALES 1
#define FCWT_HIP_LINFREQS 2

/* Plan flags. */
/* Bins covered per scale follow the reference's scalar branch
 * (src/fcwt/fcwt.cpp:276-280) instead of its AVX branch (fcwt.cpp:203-214,
 * the branch every shipped build compiles). */
#define FCWT_HIP_COVERAGE_SCALAR 0x1u

/* Message for the last failing call on this thread ("" if none). */
const char* fcwt_hip_last_error(void);

/* Library version string. */
const char* fcwt_hip_version(void);

/* ---------------------------------------------------------------------- */
/* Host arithmetic of the reference, bit-exact (no device involved).       */
/* ---------------------------------------------------------------------- */

/* find2power, src/fcwt/fcwt.h:154-165: smallest m with 2^m >= n. */
int fcwt_hip_find2power(int n);

/* Scales::Scales(wavelet, st, fs, f0, f1, fn), src/fcwt/fcwt.cpp:111-190;
 * writes fn float32 scales.  (The wavelet argument only feeds four_wavelen,
 * which the scale formulas never use.) */
int fcwt_hip_scales(int scaletype, int fs, float f0, float f1, int fn, float* scales);

/* Scales::getFrequencies, src/fcwt/fcwt.cpp:133-137: freqs[i] = fs / scales[i]. */
int fcwt_hip_frequencies(const float* scales, int fn, int fs, float* freqs);

/* Morlet::generate(int size), src/fcwt/fcwt.cpp:51-67: frequency-domain
 * mother table of `size` float32 values (host copy). */
int fcwt_hip_morlet_generate(float fb, int size, float* mother);

/* Morlet::getSupport, src/fcwt/fcwt.h:88: (int)(fb*scale*3.0f). */
int fcwt_hip_morlet_support(float fb, float scale);

/* Morlet::generate(real, imag, size, scale), src/fcwt/fcwt.cpp:68-84:
 * writes the 2*support+1 time-domain taps (support = fcwt_hip_morlet_support)
 * to real[] and imag[], normalised with `size` as the reference does. */
int fcwt_hip_morlet_generate_time(float fb, float scale, int size, float* real, float* imag);

/* Morlet::getWavelet(scale, pwav, pn), src/fcwt/fcwt.cpp:86-105: time-domain
 * wavelet, pn interleaved complex64 values (first pn of the 2*support+1 taps). */
int fcwt_hip_morlet_wavelet(float fb, float scale, float* wavelet, int pn);

/* The spectrum each output row is the inverse DFT of, as planned (host only,
 * no device): for scale i, bins [0, covered[i]) hold X[q]*mother[idx_i(q)]
 * (fcwt.cpp:197-311) and bins up to band[i] may also hold the product an
 * earlier scale left in the reference's reused O1 buffer (fcwt.cpp:449,
 * 466-467, 514-518); segments[i] counts the distinct windows.  Products whose
 * window is below 1e-12 of the window peak are left out of band[i].
 * Any output pointer may be NULL. */
int fcwt_hip_describe_rows(int n, const float* scales, int fn, float fb, int threads, unsigned flags,
                           int* covered, int* band, int* segments);

/* ---------------------------------------------------------------------- */
/* Plans and execution on the GPU.                                         */
/* ---------------------------------------------------------------------- */

typedef struct fcwt_hip_plan_s* fcwt_hip_plan;

/* Everything FCWT::cwt sets up per call before its scale loop
 * (src/fcwt/fcwt.cpp:447-511: padded size N = 2^find2power(n), buffers,
 * mother table wavelet->generate(N)), done once on the current HIP device:
 * per-scale band metadata, mother table (float64 maths rounded to float32),
 * twiddle tables, launch schedule.
 *   n          signal length (>= 1)
 *   scales     fn float32 scales (host pointer; copied)
 *   fb         Morlet bandwidth (Morlet::Morlet(fb), fcwt.cpp:41-49)
 *   threads    the FCWT constructor's thread count (fcwt.h:116); it only
 *              selects how many bins per scale the reference covers
 *              (fcwt.cpp:213-214 / 277-278), which the plan reproduces
 *   normalize  use_normalization (fcwt.h:116): divide rows by N (fcwt.cpp:431-445)
 *   flags      FCWT_HIP_* plan flags */
int fcwt_hip_plan_create(fcwt_hip_plan* plan, int n, const float* scales, int fn, float fb, int threads,
                         int normalize, unsigned flags);

/* Frees the plan and its device memory. */
int fcwt_hip_plan_destroy(fcwt_hip_plan plan);

/* Plan facts: signal length, padded FFT size N, number of scales. */
int fcwt_hip_plan_shape(fcwt_hip_plan plan, int* n, int* N, int* fn);

/* Waits for the device and reports whether the last executes of this plan
 * completed their in-kernel synchronisation (FCWT_HIP_EDEVICE if a
 * wide-band team wait failed; the output is then invalid).  fcwt_hip_cwt
 * performs this check itself; callers of fcwt_hip_cwt_device call this after
 * their stream work completes.  No reference counterpart (the reference has
 * no asynchronous entry point). */
int fcwt_hip_plan_check(fcwt_hip_plan plan);

/* Kernel timing for benchmarks and tuning (no reference counterpart; the
 * reference times whole calls with std::chrono, benchmark.cpp:225-230).
 * enable != 0 makes every later execute record HIP events around each kernel
 * on its stream; the call also clears the totals. */
int fcwt_hip_plan_profile(fcwt_hip_plan plan, int enable);

/* Waits for the recorded events and reports, per kernel kind that ran
 * ("forward_fft", "row_writer_direct", "row_spreader",
 * "row_writer_two_pass", "row_team", "adjoint", "row_columns",
 * "column_transpose"), the summed milliseconds and launch
 * count since the last fcwt_hip_plan_profile.  Writes at most `max` entries;
 * *count gets the number of kinds that ran. */
int fcwt_hip_plan_profile_read(fcwt_hip_plan plan, int max, const char** names, double* ms, long long* launches,
                               int* count);

/* How the plan splits its fn rows over the row kernels: rows written by the
 * single-work-group kernel ("row_writer_direct"), by the XCD-team kernel
 * ("row_team") and by two-pass kernels: the column path ("row_columns") or
 * the generic pair ("row_spreader" + "row_writer_two_pass").  Any pointer may be NULL.  Benchmarks use it to
 * price each kernel's output bytes; no reference counterpart. */
int fcwt_hip_plan_row_split(fcwt_hip_plan plan, int* direct_rows, int* team_rows, int* two_pass_rows);

/* Bytes of device memory the plan holds (excluding caller buffers). */
int fcwt_hip_plan_device_bytes(fcwt_hip_plan plan, size_t* bytes);

/* FCWT::cwt(float* in, int n, complex<float>* out, Scales*),
 * src/fcwt/fcwt.cpp:532-534 (core 447-529), host pointers: input holds n
 * float32 samples, output receives fn*n interleaved complex64 values, row i
 * (scale i) at output + 2*i*n.  Copies in, runs, copies out; returns when the
 * output is complete. */
int fcwt_hip_cwt(fcwt_hip_plan plan, const float* input, float* output);

/* The same with device pointers, enqueued on `stream` (a hipStream_t; NULL =
 * the default stream).  Returns once enqueued.  d_output holds fn*n complex64. */
int fcwt_hip_cwt_device(fcwt_hip_plan plan, const float* d_input, float* d_output, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): fcwt_hip_cwt_device without a plan object, allocating
 * nothing.  The plan's tables, spectrum, schedule and scratch live in a
 * static device array of 64 MiB that the library holds on each GPU it runs
 * on until the process ends; the arguments are those of fcwt_hip_plan_create
 * followed by those of fcwt_hip_cwt_device.  Limits: n <= 131072 samples
 * (padded size 2^17), and the plan must fit the array (FCWT_HIP_EUNSUPPORTED
 * otherwise; 3000 scales at n = 100,000 need about 52 MiB).  The array keeps
 * the last call's plan: a call with the same n, scales and options reuses
 * it, another rebuilds it (host work, table kernels and one stream
 * synchronisation).  Calls that overlap on one GPU race on the array. */
int fcwt_hip_cwt_device_noalloc(const float* d_input, int n, const float* scales, int fn, float fb, int threads,
                                int normalize, unsigned flags, float* d_output, void* stream);

/* FCWT::cwt(complex<float>* in, int n, complex<float>* out, Scales*),
 * src/fcwt/fcwt.cpp:536-538 (core 447-529; complex branch 486-490: the
 * zero-padded complex signal goes through a forward c2c FFT instead of the
 * r2c one), host pointers: input holds n interleaved complex64 samples
 * (2n floats), output as fcwt_hip_cwt.  Only the positive-frequency half
 * of the spectrum enters any row (the Morlet window is one-sided). */
int fcwt_hip_ccwt(fcwt_hip_plan plan, const float* input, float* output);

/* The same with device pointers, enqueued on `stream`. */
int fcwt_hip_ccwt_device(fcwt_hip_plan plan, const float* d_input, float* d_output, void* stream);

/* The forward transform alone: X[k] = sum_{t<n} x[t] e^{-2 pi i k t/N} for
 * k < N/2 of the zero-padded input (real, or complex with complex_input),
 * written to d_spec as N/2 complex64 -- what the reference's FFTW r2c / c2c
 * plans write to Ihat (src/fcwt/fcwt.cpp:486-500), the lower half of which
 * is all the rows read.  Device pointers, enqueued on `stream`. */
int fcwt_hip_spectrum_device(fcwt_hip_plan plan, const float* d_input, float* d_spec, int complex_input, void* stream);

/* Adjoint of the plan's transform: the input gradient from an output
 * gradient (no reference counterpart -- the reference has no derivative; the
 * torch.autograd.Function of fcwt_amd uses this).  d_grad holds fn*n complex64,
 * G_i[t] = dL/dRe W_i[t] + i dL/dIm W_i[t]; d_xbar receives n float32
 * (complex_input = 0: dL/dx for real input) or n interleaved complex64
 * (complex_input != 0: dL/dRe x + i dL/dIm x):
 *   xbar[t] = c sum_{q<N/2} A[q] e^{2 pi i q t/N}   (real part for real input)
 *   A[q]    = sum_i w_i(q) sum_{t'<n} G_i[t'] e^{-2 pi i q t'/N}
 * with w_i the window the forward applied to row i (stale bins included) and
 * c = 1/N when normalising.  Enqueued on `stream`; the first call allocates
 * the plan's adjoint workspace (about 256 MiB at most). */
int fcwt_hip_adjoint_device(fcwt_hip_plan plan, const float* d_grad, float* d_xbar, int complex_input, void* stream);

/* The same with host pointers; returns when xbar is complete. */
int fcwt_hip_adjoint(fcwt_hip_plan plan, const float* grad, float* xbar, int complex_input);

/* The adjoint for `signals` signals of the plan's length and scales at once
 * (the backward pass of a batch): d_grad holds signals*fn*n complex64
 * (signal s's rows at s*fn*n), d_xbar receives signals*n values (float32,
 * or complex64 with complex_input), signal s at s*n.  All row transforms of
 * the batch go through shared launches.  signals <= 65535. */
int fcwt_hip_adjoint_batch_device(fcwt_hip_plan plan, const float* d_grad, float* d_xbar, int signals,
                                  int complex_input, void* stream);

/* ---------------------------------------------------------------------- */
/* Tuning: the analogue of FCWT::create_FFT_optimization_plan and
 * load_FFT_optimization_plan (src/fcwt/fcwt.cpp:313-393), which make and
 * read FFTW wisdom files n{N}_t{T}.wis.  fcwt_hip_tune times the schedule
 * choices of padded size N (power of two) on this GPU -- the XCD-team
 * kernel's scratch buffer count, the column path's scratch per batch, and
 * which of the two takes the wide rows where both can -- and writes the best
 * to `path` (a short text file) and to this process's table.  effort: 0 keeps the defaults
 * (FFTW_ESTIMATE), 1 takes the best of three runs of each candidate
 * (FFTW_MEASURE), 2 the best of seven over a wider candidate set
 * (FFTW_PATIENT, FFTW_EXHAUSTIVE).  fcwt_hip_tune_load reads
 * such a file; plans created afterwards for that N use it. */
int fcwt_hip_tune(int N, int threads, int effort, const char* path);
int fcwt_hip_tune_load(const char* path);

/* ---------------------------------------------------------------------- */
/* Batches of signals (no reference counterpart: the reference transforms one
 * signal per FCWT::cwt call).  Signal c has lengths[c] samples and its own
 * fns[c] scales at scales[c]; the Morlet bandwidth, threads, normalisation
 * and flags are common.  Signals that pad to the same N share one plan, so an
 * execute makes one forward-FFT launch and one launch of each row kernel per
 * distinct N for all of their rows together. */
typedef struct fcwt_hip_batch_s* fcwt_hip_batch;

int fcwt_hip_batch_create(fcwt_hip_batch* batch, int count, const int* lengths, const float* const* scales,
                          const int* fns, float fb, int threads, int normalize, unsigned flags);

int fcwt_hip_batch_destroy(fcwt_hip_batch batch);

/* d_inputs[c]: lengths[c] float32 samples, d_outputs[c]: fns[c]*lengths[c]
 * complex64 (host arrays of `count` device pointers), enqueued on `stream`.
 * When the same pointers come twice in a row, the launches are captured into
 * a HIP graph, and later executes with those pointers replay it (one graph
 * launch); other pointers drop it.  FCWT_HIP_NO_GRAPHS=1 keeps every execute
 * eager. */
int fcwt_hip_batch_execute_device(fcwt_hip_batch batch, const float* const* d_inputs, float* const* d_outputs,
                                  void* stream);

/* Executes of the batch that were served by a graph launch so far. */
int fcwt_hip_batch_graph_launches(fcwt_hip_batch batch, long long* launches);

/* As fcwt_hip_plan_check, for every plan of the batch. */
int fcwt_hip_batch_check(fcwt_hip_batch batch);

/* Plans (distinct padded sizes) in the batch and rows on the XCD-team kernel. */
int fcwt_hip_batch_shape(fcwt_hip_batch batch, int* groups, int* team_rows);

/* Launches of row kernels (direct, XCD-team, column and two-pass kernels;
   not the forward transforms or the fills that clear counters) that one
   execute of the batch makes.  When two or more padded sizes have rows on
   the XCD-team kernel, one launch takes the team rows and the direct rows of
   all of them.  Introspection of this build's schedule; no counterpart in the
   reference, which makes one FCWT::cwt call per signal (fcwt.cpp:447-529). */
int fcwt_hip_batch_row_launches(fcwt_hip_batch batch, int* launches);

#ifdef __cplusplus
}
#endif

#endif /* FCWT_HIP_H */
