/*
 * genfft_amd.h -- C ABI of the MI355X (gfx950) FFT engine.
 *
 * This is the drop-in boundary under the genfft-compatible C++ surface in
 * <genfft_amd/fft.h> (and any other binding: ctypes, cgo, JNI ...).  Plain C,
 * plain pointers and sizes; every entry point returns an int status
 * (GFFT_OK = 0) and records a message readable with gfft_last_error().
 *
 * Reference interfaces replaced (paths relative to the genFFT tree):
 *   genfft::FFT<T>(n)                 include/genFFT/fft.h:60-64
 *        -> gfft_plan_create(kind = GFFT_C2C)
 *   FFT<T>::transform<inv>(out, in)   include/genFFT/fft.h:80-85
 *        -> gfft_execute(plan, in, out, inv, stream)
 *   FFT<T>::transform_real(out, in)   include/genFFT/fft.h:90-94
 *        -> GFFT_R2C plan without GFFT_FLAG_HALF
 *   FFT<T>::transform_interleave      include/genFFT/fft.h:100-105
 *        -> gfft_interleave_real() + GFFT_C2C execute
 *   RealFFT<T>::forward(out,in,half)  include/genFFT/FFTReal.h:204-213
 *        -> GFFT_R2C plan (GFFT_FLAG_HALF selects n/2+1 bins)
 *   DIT<T>::apply(out, in, half)      include/genFFT/fft.h:180-188,
 *                                     generic/fft_dit_impl_generic.inl:27-61
 *        -> GFFT_DIT plan
 *   separate_2x_real_FFT              include/genFFT/FFTReal.h:36-66
 *        -> gfft_separate_2x_real()
 *   FFTVert<T>::transform<inv>        include/genFFT/fft.h:146-163
 *        -> GFFT_COLS plan
 *   FFT2D<T>::transform<inv>          include/genFFT/fft.h:214-218
 *        -> GFFT_C2C_2D plan
 *   RealFFT2D<T>::forward             include/genFFT/FFTReal.h:83-104
 *        -> GFFT_R2C_2D plan
 *   RealFFT2D<T>::inverse             include/genFFT/FFTReal.h:120-130 (declared,
 *                                     body is assert(!"TODO") in the reference)
 *        -> GFFT_C2R_2D plan
 *   backend::GetImpl / GetVertImpl / GetDITImpl (the plan factory seam,
 *   include/genFFT/FFTBackend.h:40-70) -> the plan cache of the C++ layer.
 *
 * Conventions (same as the reference): forward kernel exp(-2*pi*i*j*k/N),
 * inverse unscaled (inverse(forward(x)) = N*x), natural order in and out,
 * interleaved complex (re, im), power-of-two sizes 2^0 .. 2^23.
 */
#ifndef GENFFT_AMD_H
#define GENFFT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef GFFT_API
#define GFFT_API __attribute__((visibility("default")))
#endif

#define GFFT_VERSION 1

/* status codes */
#define GFFT_OK 0
#define GFFT_ERR_INVALID_ARG 1     /* null pointer, bad kind/precision/stride */
#define GFFT_ERR_UNSUPPORTED 2     /* size or layout outside what the engine supports */
#define GFFT_ERR_HIP 3             /* a HIP runtime call failed */
#define GFFT_ERR_ALLOC 4           /* device or host allocation failed */
#define GFFT_ERR_NO_DEVICE 5       /* no usable gfx950 device */

/* transform kinds */
#define GFFT_C2C 0      /* batched 1D complex -> complex                          */
#define GFFT_R2C 1      /* batched 1D real -> complex (RealFFT::forward)          */
#define GFFT_C2C_2D 2   /* 2D complex (FFT2D), n = width (cols), n2 = height      */
#define GFFT_R2C_2D 3   /* 2D real-input forward (RealFFT2D::forward)             */
#define GFFT_COLS 4     /* column FFTs (FFTVert): n = column length, n2 = columns */
#define GFFT_DIT 5      /* real-FFT unpack only (DIT::apply), n = real length     */
#define GFFT_C2R 6      /* batched half-spectrum -> real, unscaled inverse         */
#define GFFT_C2R_2D 7   /* 2D inverse of GFFT_R2C_2D (RealFFT2D::inverse): n = width */
                        /* (reals per row), n2 = height; reads columns 0..n/2 of   */
                        /* the complex input (pitch in complex), writes reals      */
                        /* (pitch in scalars); unscaled                            */

/* precisions */
#define GFFT_F32 0      /* complex64  (interleaved float)  */
#define GFFT_F64 1      /* complex128 (interleaved double) */
/* 16-bit storage with FP32 arithmetic: a complex element is two 16-bit
 * halves (real part first), 4 bytes; C2C, C2C_2D and COLS plans.  No
 * reference counterpart (genFFT is float/double only). */
#define GFFT_F16 2      /* IEEE binary16 pairs */
#define GFFT_BF16 3     /* bfloat16 pairs      */

/* gfft_execute_scaled scale modes */
#define GFFT_SCALE_GIVEN 0 /* scales[b] supplied by the caller                      */
#define GFFT_SCALE_AUTO 1  /* library writes scales[b] (power of two <= 1) first    */

/* flags */
#define GFFT_FLAG_HALF 1 /* R2C / DIT: write n/2+1 bins (else all n, Hermitian-filled) */
/* plan_create allocates no workspace: the plan runs only through
 * gfft_execute_noalloc, with the library's static workspace (batches are cut
 * into chunks that fit it); plans needing more than
 * GFFT_NOALLOC_SCRATCH_BYTES are rejected with GFFT_ERR_UNSUPPORTED */
#define GFFT_FLAG_NO_WORKSPACE 2

/* Size of the static device workspace gfft_execute_noalloc uses, per GPU. */
#define GFFT_NOALLOC_SCRATCH_BYTES (64ll << 20)

/* directions */
#define GFFT_FORWARD 0
#define GFFT_INVERSE 1

typedef struct gfft_plan_s *gfft_plan;

typedef struct gfft_desc {
    int32_t kind;       /* GFFT_C2C ... GFFT_C2R_2D                                */
    int32_t precision;  /* GFFT_F32 / GFFT_F64 / GFFT_F16 / GFFT_BF16                */
    int64_t n;          /* 1D length; 2D width (columns); COLS column length (rows) */
    int64_t n2;         /* 2D height (rows); COLS number of columns; else ignored   */
    int64_t batch;      /* independent transforms (>= 1)                            */
    int64_t in_stride;  /* 2D / COLS row pitch of the input: complex elements, but   */
                        /* scalars for R2C_2D (as RealFFT2D::forward); 0 = dense     */
    int64_t out_stride; /* 2D / COLS row pitch of the output, complex elements; 0 = dense */
    int64_t in_dist;    /* distance between batch members, elements of the input   */
                        /* type (complex, or real for R2C); 0 = dense               */
    int64_t out_dist;   /* same for the output (real for C2R); 0 = dense.  A real   */
                        /* side may have any pitch and distance: when one is odd   */
                        /* the kernels (which move reals as complex pairs) run on  */
                        /* a dense copy in the plan's workspace, one extra pass    */
    int32_t flags;      /* GFFT_FLAG_*                                              */
    int32_t reserved;   /* must be 0                                                */
} gfft_desc;

/* Creates a plan: validates the descriptor, selects kernels, uploads the
 * FP64-generated twiddle tables and allocates the plan's device workspace.
 * The plan belongs to the HIP device current at creation; every call that
 * executes it (and the batch and filter plans alike) must be made with that
 * device current, else it returns GFFT_ERR_INVALID_ARG. */
GFFT_API int gfft_plan_create(gfft_plan *plan, const gfft_desc *desc);

/* Releases a plan and its workspace. NULL is accepted. */
GFFT_API int gfft_plan_destroy(gfft_plan plan);

/* Executes on device-resident buffers, asynchronously on `stream`
 * (a hipStream_t; NULL = the null stream).  Input and output may overlap in
 * any way.  Where a plan's launches may write part of the output before all
 * of the input is read (every plan whose tile pass reads the input and
 * writes the output, a 1D batch beyond the 1 GiB workspace run in chunks)
 * and the buffers overlap other than exactly in place (in == out with equal
 * input and output pitches and distances in bytes), the input is first
 * copied to a plan-owned buffer (allocated on first use) and the plan runs
 * from there.
 * direction: GFFT_FORWARD or GFFT_INVERSE (R2C / R2C_2D / DIT are forward
 * only, C2R inverse only).  Calls may share a plan from any threads and
 * streams, as with genFFT's shared plans (FFTLevel.h:76-92): a call on a
 * stream other than the plan's previous one first waits (hipStreamWaitEvent)
 * for the plan's previous work, because they share its workspace. */
GFFT_API int gfft_execute(gfft_plan plan, const void *in, void *out, int direction, void *stream);

/* gfft_execute that allocates nothing (no device memory, no events): for
 * callers that cannot allocate at call time -- code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  No reference
 * counterpart (genFFT allocates nothing on the device at all).  The plan's
 * workspace is replaced by one static device array of
 * GFFT_NOALLOC_SCRATCH_BYTES per GPU (held from the library's first use on
 * that GPU until the process ends); a plan needing more, or a call whose
 * buffers overlap other than exactly in place where the plan would stage its
 * input, returns GFFT_ERR_UNSUPPORTED.  Create the plan with
 * GFFT_FLAG_NO_WORKSPACE to keep its batches chunked to that size and to skip
 * its own workspace (any plan that fits may be run).  Every call on one GPU
 * shares that array: calls that overlap in time on one GPU (other streams,
 * other threads) race and must be ordered by the caller.  Nothing in the
 * array needs clearing between calls (each launch writes what it later
 * reads). */
GFFT_API int gfft_execute_noalloc(gfft_plan plan, const void *in, void *out, int direction, void *stream);

/* gfft_execute for 16-bit storage plans with a per-transform scale:
 * transform b (one batch member: a 1D transform, a 2D image, a COLS matrix)
 * is stored as scales[b] * FFT(x_b), the factor applied in FP32 before the
 * first rounding to 16 bits.  `scales` is a device array of `batch` floats.
 * GFFT_SCALE_GIVEN reads it.  GFFT_SCALE_AUTO first writes it, on `stream`:
 * scales[b] = the largest power of two <= 1 with
 *     len_b * max_j(|Re x_bj| + |Im x_bj|) * scales[b] <= 2^15   (FP16)
 * (len_b = the transform's point count, which bounds every partial and final
 * DFT value), so no stored value overflows FP16's 65504; for BF16 the bound
 * is 2^120 (FP32's exponent range), i.e. 1 for any finite input.  Powers of
 * two are exact: divide the output by scales[b] to undo it.  No reference
 * counterpart. */
GFFT_API int gfft_execute_scaled(gfft_plan plan, const void *in, void *out, int direction, float *scales,
                                 int scale_mode, void *stream);

/* gfft_execute on input in bit-reversed order: element j of each transform
 * (C2C) or row j of each matrix (COLS) holds what the natural order puts at
 * bit-reverse(j).  Replaces FFT<T>::transform_no_scramble
 * (include/genFFT/fft.h:69-73) and FFTVert<T>::transform_no_scramble
 * (fft.h:132-136) on device data: the rows are gathered into natural order in
 * a plan-owned buffer (allocated on first use), then the plan runs; in == out
 * is allowed.  complex64 / complex128 C2C and COLS plans. */
GFFT_API int gfft_execute_scrambled(gfft_plan plan, const void *in, void *out, int direction, void *stream);

/* Same as gfft_execute for host buffers: stages through plan-owned device
 * memory and returns after the result is back in `out` (synchronous). */
GFFT_API int gfft_execute_host(gfft_plan plan, const void *in, void *out, int direction);

/* Runs the plan once like gfft_execute, recording a HIP event pair on
 * `stream` around each kernel launch, waits for it, and writes the duration
 * of launch i (milliseconds) to ms[i] for i < max_launches.  *count receives
 * the number of launches.  For benchmarks and profiles; no reference
 * counterpart (genFFT has no timing hooks). */
GFFT_API int gfft_plan_profile(gfft_plan plan, const void *in, void *out, int direction, void *stream,
                               float *ms, int max_launches, int *count);

/* Runs launch `launch` of the plan `reps` times back to back on `stream`
 * between one HIP event pair, waits, and writes the mean duration of one
 * launch (milliseconds) to *ms: a kernel's time without the cost of an event
 * pair around each launch.  The buffers are those of a gfft_execute call
 * (no staging copy is made); repeating a launch of an in-place plan changes
 * the data, not the timing.  GFFT_ERR_UNSUPPORTED for the launches of an
 * uneven batch's work queue.  No reference counterpart. */
GFFT_API int gfft_plan_profile_launch(gfft_plan plan, const void *in, void *out, int direction, void *stream,
                                      int launch, int reps, float *ms);

/* Device workspace the plan holds, in bytes. */
GFFT_API int gfft_plan_workspace(gfft_plan plan, size_t *bytes);

/* Bytes a call reads from `in` and writes to `out` (from each pointer to the
 * last byte touched, pitch gaps included): what a binding checks its buffers
 * against before it calls gfft_execute.  No reference counterpart (genFFT's
 * callers size their own arrays). */
GFFT_API int gfft_plan_extents(gfft_plan plan, size_t *in_bytes, size_t *out_bytes);

/* Human-readable list of the plan's kernel launches (for logs and tests). */
GFFT_API int gfft_plan_describe(gfft_plan plan, char *buf, size_t len);

/* Plans `desc` without touching a device (no allocation, no upload), proves
 * every launch's index range lies inside its buffer, and writes the same
 * description as gfft_plan_describe.  Usable on machines without a GPU. */
GFFT_API int gfft_plan_explain(const gfft_desc *desc, char *buf, size_t len);

/* separate_2x_real_FFT on device data: `in` holds the n-point FFT of
 * x + i*y; out1 / out2 receive the n-point spectra of x and y.  in may alias
 * out1.  batch transforms at distance n (complex elements). */
GFFT_API int gfft_separate_2x_real(int precision, void *out1, void *out2, const void *in, int64_t n,
                          int64_t batch, void *stream);

/* transform_interleave's packing step: out[k] = in1[k] + i*in2[k], n reals
 * each, batch rows at distance n. */
GFFT_API int gfft_interleave_real(int precision, void *out, const void *in1, const void *in2, int64_t n,
                         int64_t batch, void *stream);

/* RealFFT2D<T>::forward_2x's packing step (include/genFFT/FFTReal.h:106-118,
 * rows scrambled at :167-180) in one launch: row r of out (pitch out_stride,
 * complex elements) = row r of in1 + i * row r of in2 (pitches in_stride1 /
 * in_stride2, real elements), width elements per row, height rows.  Each
 * input keeps its own pitch (genFFT offsets in2's second half by in_stride1,
 * FFTReal.h:178, which agrees only when the two pitches are equal). */
GFFT_API int gfft_interleave_real_2d(int precision, void *out, int64_t out_stride, const void *in1,
                                     int64_t in_stride1, const void *in2, int64_t in_stride2, int64_t width,
                                     int64_t height, void *stream);

/* ---- uneven batches (no reference counterpart: genFFT runs one transform
 * per call) ------------------------------------------------------------------
 * `count` transforms of power-of-two lengths sizes[j] (1..2^23), job j reading
 * in + in_offsets[j] and writing out + out_offsets[j] (complex elements),
 * forward or (unscaled) inverse per directions[j].  Lengths up to 4096 (c64)
 * / 2048 (c128) run in one persistent launch that drains a device-side task
 * queue; longer lengths run as per-length batched launches.  Jobs must not
 * write where another job reads or writes (a job may be in place). */
typedef struct gfft_batch_plan_s *gfft_batch_plan;

#define GFFT_BATCH_GRAPH 0    /* capture the launches once into a HIP graph and replay it */
#define GFFT_BATCH_EAGER 1    /* enqueue the launches directly */
#define GFFT_BATCH_PER_JOB 2  /* the plain path: one single-transform plan launch per job */

GFFT_API int gfft_batch_plan_create(gfft_batch_plan *plan, int precision, int64_t count, const int64_t *sizes,
                                    const int64_t *in_offsets, const int64_t *out_offsets,
                                    const int32_t *directions);
/* Device buffers; asynchronous on `stream`.  The graph is re-captured when in
 * or out change (after the previous graph has finished).  Calls on different
 * streams are ordered like gfft_execute's: they share the plan's scratch and
 * work-queue counter. */
GFFT_API int gfft_batch_execute(gfft_batch_plan plan, const void *in, void *out, int mode, void *stream);
GFFT_API int gfft_batch_plan_describe(gfft_batch_plan plan, char *buf, size_t len);
GFFT_API int gfft_batch_plan_destroy(gfft_batch_plan plan);

/* ---- fused spectral filter and its adjoint (no reference counterpart) ------
 * On `batch` real rows of length n (64 .. 16384 complex64 path / 8192
 * complex128 path), H holding the n/2+1 bins of a Hermitian filter:
 *   forward   y  = irfft(H * rfft(x))      (conj(H) when conj_h != 0)
 *   backward  gx = irfft(conj(H) * rfft(gy)),
 *             gH = sum_rows w_k/n * rfft(gy)_k * conj(rfft(x)_k), w = 1 at k = 0, n/2, else 2
 * irfft is normalised (1/n) and ignores the imaginary parts of bins 0 and n/2.
 * Each row is one workgroup doing FFT -> filter -> inverse FFT in LDS: one
 * HBM read and one write per row.  gH is the gradient of a real loss with
 * respect to H (d/dRe + i d/dIm).  Rows are contiguous (distance n reals). */
typedef struct gfft_filter_plan_s *gfft_filter_plan;
GFFT_API int gfft_filter_plan_create(gfft_filter_plan *plan, int precision, int64_t n, int64_t batch);
GFFT_API int gfft_filter_forward(gfft_filter_plan plan, const void *x, const void *H, void *y, int conj_h,
                                 void *stream);
GFFT_API int gfft_filter_backward(gfft_filter_plan plan, const void *x, const void *H, const void *gy, void *gx,
                                  void *gH, void *stream);
GFFT_API int gfft_filter_plan_destroy(gfft_filter_plan plan);

/* ---- autotuner (no reference counterpart) ---------------------------------
 * Times every candidate kernel choice for a batched C2C of length n (four-step
 * split, tile widths, pipelined kernels, radix order / threads of the
 * single-pass kernel) on `batch` transforms, `reps` runs each, keeps the
 * fastest in the process-wide tuning table that later gfft_plan_create calls
 * consult, and writes a JSON report of all candidates to `report`.  The table
 * is saved / loaded as JSON; GFFT_TUNING_FILE names a file loaded at the first
 * plan creation. */
GFFT_API int gfft_autotune(int precision, int64_t n, int64_t batch, int reps, char *report, size_t len);
GFFT_API int gfft_tuning_save(const char *path);
GFFT_API int gfft_tuning_load(const char *path);
GFFT_API int gfft_tuning_clear(void);

/* 1 if p is device (or managed) memory visible to the current device. */
GFFT_API int gfft_is_device_pointer(const void *p);

/* Static text for a status code; the last error message of this thread. */
GFFT_API const char *gfft_error_string(int status);
GFFT_API const char *gfft_last_error(void);

GFFT_API int gfft_version(void);

#ifdef __cplusplus
}
#endif

#endif /* GENFFT_AMD_H */
