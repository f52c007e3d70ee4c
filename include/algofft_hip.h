/*
 * algofft_hip — C ABI of the MI355X (gfx950) FFT backend for algo-fft.
 *
 * This is the boundary a cgo binding of algo-fft's `gpu` package binds to.
 * Each entry point names the Go interface method it backs
 * (gpu/backend.go, gpu/plan.go, gpu/types.go, gpu/errors.go in
 * github.com/cwbudde/algo-fft).  Plain pointers, sizes and integer status
 * codes only: no C++ or framework types cross this line.
 *
 * Data layout: complex values are interleaved (re, im) exactly like Go's
 * complex64/complex128 slices; real values are float/double arrays.
 * Element counts are in complex elements unless stated otherwise.
 *
 * Threading: a plan's executions are serialised by a lock inside the plan,
 * so one plan may be shared by several host threads (plan.go:14: plans are
 * safe for concurrent use during transforms); other objects are used from one
 * thread at a time.  Distinct plans (and clones) run concurrently on distinct
 * streams.  The last error message is thread-local.
 */
#ifndef ALGOFFT_HIP_H
#define ALGOFFT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: 1:1 with the Go sentinel errors ---------------------- */
typedef enum {
    AFFT_OK = 0,
    AFFT_ERR_NO_BACKEND = 1,          /* gpu/errors.go:7   ErrNoBackend */
    AFFT_ERR_BACKEND_UNAVAILABLE = 2, /* gpu/errors.go:11  ErrBackendUnavailable */
    AFFT_ERR_NOT_IMPLEMENTED = 3,     /* gpu/errors.go:14  ErrNotImplemented (also wrong dynamic type) */
    AFFT_ERR_INVALID_LENGTH = 4,      /* gpu/errors.go:17  ErrInvalidLength */
    AFFT_ERR_NIL_SLICE = 5,           /* gpu/errors.go:20  ErrNilSlice */
    AFFT_ERR_LENGTH_MISMATCH = 6,     /* gpu/errors.go:23  ErrLengthMismatch */
    AFFT_ERR_INVALID_STRIDE = 7,      /* errors.go:21      ErrInvalidStride */
    AFFT_ERR_INVALID_SPECTRUM = 8,    /* errors.go:25      ErrInvalidSpectrum */
    AFFT_ERR_DEVICE = 9,              /* HIP runtime failure (allocation, launch, copy) */
    AFFT_ERR_INVALID_ARGUMENT = 10    /* bad handle or descriptor field */
} afft_status;

/* gpu/types.go:10-13 PrecisionKind */
typedef enum { AFFT_COMPLEX64 = 0, AFFT_COMPLEX128 = 1 } afft_precision;

typedef enum { AFFT_FORWARD = -1, AFFT_INVERSE = 1 } afft_direction;

/* Transform kinds of the extended plan descriptor. */
typedef enum { AFFT_C2C = 0, AFFT_R2C = 1, AFFT_C2R = 2 } afft_kind;

/* Execution forms selectable per plan (afft_plan_desc.flags). */
enum {
    AFFT_FLAG_INPLACE = 1 << 0,   /* gpu/types.go:40 PlanOptions.InPlace */
    AFFT_FLAG_MFMA = 1 << 1,      /* radix-16 stages on matrix cores (hi/lo split) */
    AFFT_FLAG_NORMALIZE = 1 << 2, /* forward scaled by 1/N (PlanRealT.ForwardNormalized) */
    AFFT_FLAG_UNITARY = 1 << 3,   /* forward scaled by 1/sqrt(N) (ForwardUnitary) */
    AFFT_FLAG_REF_CHIRP = 1 << 4  /* Bluestein chirp without the k^2 mod 2N reduction */
};

/* gpu/types.go:16-22 DeviceInfo (strings are NUL-terminated, truncated). */
typedef struct {
    char name[128];
    char vendor[64];
    char driver[64];
    int64_t memory_mb;
    char compute_cap[64];
} afft_device_info;

/* gpu/types.go:25-29 BackendInfo */
typedef struct {
    char name[64];
    char version[64];
    char description[256];
} afft_backend_info;

/* gpu/types.go:33-42 PlanOptions */
typedef struct {
    int device_index;
    int stream_count;
    int in_place;
} afft_plan_options;

/* Extended plan descriptor for the surface the Go `gpu` package lacks today
 * but the CPU Plan API has (plan_options.go:46-76: Batch, Stride; plus the
 * real and N-D plans).  dims[0..rank-1] are row-major, last fastest.
 * For C2C, strides/distances are in complex elements; 0 means contiguous
 * defaults (stride 1, distance = product of dims).  For R2C/C2R the real side
 * uses real elements and the complex side complex elements; defaults are
 * the packed layouts N and N/2+1 on the last dimension. */
typedef struct {
    int kind;            /* afft_kind */
    int precision;       /* afft_precision */
    int rank;            /* 1..8 */
    int64_t dims[8];
    int64_t batch;       /* >= 1 */
    int64_t istride, idist, ostride, odist;
    int flags;
    int strategy;        /* afft_strategy for 1-D plans, 0 = planner's choice */
    int64_t split;       /* four-step first factor n1 (0 = balanced) */
} afft_plan_desc;

/* Forced kernel strategy (cf. PlanOptions.Strategy, plan_options.go:51-53,
 * and the KernelStrategy values of internal/fftypes/strategy.go:4-14). */
typedef enum {
    AFFT_STRATEGY_AUTO = 0,      /* KernelAuto: wisdom, then cost model, then heuristic */
    AFFT_STRATEGY_STOCKHAM = 2,  /* KernelStockham: one compile-time LDS pass set */
    AFFT_STRATEGY_SIXSTEP = 3,   /* KernelSixStep: four-step with first factor `split` */
    AFFT_STRATEGY_BLUESTEIN = 5, /* KernelBluestein */
    AFFT_STRATEGY_MIXED = 16     /* runtime-size mixed-radix LDS kernel */
} afft_strategy;

typedef struct afft_context_s* afft_context;
typedef struct afft_buffer_s* afft_buffer;
typedef struct afft_stream_s* afft_stream;
typedef struct afft_plan_s* afft_plan;

/* ---- errors -------------------------------------------------------------- */
/* Message of the last failing call on this thread ("" if none). */
const char* afft_last_error(void);
/* Fixed text of a status code. */
const char* afft_status_string(int status);

/* ---- Backend (gpu/backend.go:7-12) --------------------------------------- */
int afft_backend_get_info(afft_backend_info* out);          /* Backend.Info */
int afft_backend_available(void);                            /* Backend.Available: 1 or 0 */
int afft_device_count(int* count);                           /* Backend.Devices (count) */
int afft_device_get_info(int index, afft_device_info* out); /* Backend.Devices (entry) */

/* ---- Context (gpu/backend.go:15-24) -------------------------------------- */
int afft_context_create(int device_index, afft_context* out); /* Backend.NewContext */
int afft_context_device(afft_context ctx, afft_device_info* out); /* Context.Device */
int afft_context_destroy(afft_context ctx);                   /* Context.Close */

/* ---- Buffer (gpu/backend.go:27-35; Context.NewBuffer :18) ---------------- */
int afft_buffer_create(afft_context ctx, int64_t elem_count, int precision, afft_buffer* out);
int64_t afft_buffer_len(afft_buffer buf);                       /* Buffer.Len */
int afft_buffer_precision(afft_buffer buf);                     /* Buffer.Precision */
/* Buffer.Upload / Buffer.Download: host_len is the host slice length in
 * complex elements; it must be >= Len (ErrLengthMismatch otherwise). */
int afft_buffer_upload(afft_buffer buf, const void* host, int64_t host_len);
int afft_buffer_download(afft_buffer buf, void* host, int64_t host_len);
void* afft_buffer_device_ptr(afft_buffer buf);
int afft_buffer_destroy(afft_buffer buf);                       /* Buffer.Close */

/* Page-locked host memory (no counterpart in the reference: its gpu package
 * only takes Go slices).  Host slices in such memory are transferred by DMA
 * directly, without staging, by afft_plan_forward / afft_plan_inverse. */
int afft_host_alloc(afft_context ctx, int64_t bytes, void** out);
int afft_host_free(void* ptr);

/* ---- Stream (gpu/backend.go:38-41; Context.NewStream :20) ---------------- */
int afft_stream_create(afft_context ctx, afft_stream* out);
int afft_stream_synchronize(afft_stream s);                     /* Stream.Synchronize */
int afft_stream_destroy(afft_stream s);                         /* Stream.Close */
/* Underlying hipStream_t, as an opaque pointer. */
void* afft_stream_native(afft_stream s);

/* ---- PlanImpl (gpu/backend.go:45-51; Context.NewFFTPlan :22) ------------- */
/* Context.NewFFTPlan(n, precision, opts): one 1-D C2C transform of n points. */
int afft_plan_create(afft_context ctx, int64_t n, int precision, const afft_plan_options* opts,
                     afft_plan* out);
/* Extended constructor: batch/stride (Plan.ForwardBatch, plan.go:504;
 * Plan.ForwardStrided, plan_strided.go:16), real (NewPlanReal32/64,
 * plan_real_constructors.go:8-43) and N-D (NewPlan2D/3D/ND, plan_2d.go:44,
 * plan_3d.go:42, plan_nd.go:43). */
int afft_plan_create_desc(afft_context ctx, const afft_plan_desc* desc, afft_plan* out);
int64_t afft_plan_len(afft_plan p);           /* PlanImpl.Len: points of one transform */
int afft_plan_precision(afft_plan p);         /* PlanImpl.Precision */
/* Required host/device lengths of the input and output side, in elements
 * of that side's type (complex, or real for the real side of R2C/C2R). */
int afft_plan_io_len(afft_plan p, int direction, int64_t* in_len, int64_t* out_len);
/* PlanImpl.Forward / Inverse with host slices (the Go []complex64/128 case):
 * stages through device memory on the plan's stream and synchronises.
 * dst may equal src.  Lengths are host slice lengths (elements). */
int afft_plan_forward(afft_plan p, void* dst, int64_t dst_len, const void* src, int64_t src_len);
int afft_plan_inverse(afft_plan p, void* dst, int64_t dst_len, const void* src, int64_t src_len);
/* PlanImpl.Forward / Inverse with device-resident data (the Buffer case):
 * enqueues on `stream` (hipStream_t as void*, NULL = the plan's stream)
 * and returns without synchronising.  A plan with workspace that is enqueued
 * on a different stream than its previous execution makes the new stream
 * wait for that execution first (outside stream capture), so executions of
 * one plan never overlap on its workspace; clones overlap freely. */
int afft_plan_exec_device(afft_plan p, int direction, void* dst_dev, const void* src_dev, void* stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory while they execute (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): afft_plan_exec_device with the
 * plan's scratch taken from one static device array per GPU of
 * afft_noalloc_scratch_bytes() bytes (64 MiB) instead of memory the plan
 * allocates on its first execution.  Same kernels, same results.  A plan
 * whose workspace exceeds the array fails with ErrInvalidLength before
 * anything is enqueued.  All calls on one GPU share the array: calls that
 * overlap in time on one GPU race on it (order them on one stream). */
int afft_plan_exec_device_noalloc(afft_plan p, int direction, void* dst_dev, const void* src_dev, void* stream);
int64_t afft_noalloc_scratch_bytes(void);
/* PlanReal2D/3D.ForwardFull and InverseFull (plan_real_2d.go:231-276,
 * :370-389; plan_real_3d.go:241-285, :392-419) for a real plan of one
 * transform in the contiguous layout.  Forward: n reals in, the full
 * D x ... x W spectrum of n complex bins out (bins past W/2 filled as
 * conj(X[-k mod dims])).  Inverse: n complex bins in, of which the first
 * W/2+1 of every row are used, n reals out.  Host slices (lengths in
 * elements, exactly n each, else ErrLengthMismatch) or device pointers
 * enqueued on `stream` like afft_plan_exec_device. */
int afft_plan_forward_full(afft_plan p, void* dst, int64_t dst_len, const void* src, int64_t src_len);
int afft_plan_inverse_full(afft_plan p, void* dst, int64_t dst_len, const void* src, int64_t src_len);
int afft_plan_exec_full_device(afft_plan p, int direction, void* dst_dev, const void* src_dev, void* stream);
int afft_plan_exec_buffers(afft_plan p, int direction, afft_buffer dst, afft_buffer src, afft_stream s);
/* Waits for the work enqueued on the plan's own stream (exec calls made with
 * a NULL stream). */
int afft_plan_synchronize(afft_plan p);
/* Algorithm string recorded in wisdom (cf. Plan.Algorithm, plan.go:157). */
int afft_plan_describe(afft_plan p, char* buf, int cap);
/* Bytes of device workspace the plan holds. */
int64_t afft_plan_workspace_bytes(afft_plan p);
int afft_plan_destroy(afft_plan p);           /* PlanImpl.Close */

/* Plan.Clone / Plan.NewExecutor (plan.go:1300, executor.go:6-43): a plan of
 * the same description with its own stream and workspace, so host threads
 * that each use their own clone never contend for one plan. */
int afft_plan_clone(afft_plan p, afft_plan* out);

/* ---- FastPlan (plan_fast.go:21-133) ---------------------------------------
 * Power-of-two plans that run as one kernel launch, with the launch resolved
 * at creation.  NewFastPlan's errors: ErrInvalidLength unless n >= 1 is a
 * power of two, ErrNotImplemented when no single-launch kernel covers n
 * (n > 16384 complex64 / 8192 complex128).  batch transforms of n contiguous
 * points.  Forward / Inverse take device pointers, validate nothing and take
 * no lock: the caller guarantees valid buffers of n*batch elements, the
 * plan's device current, and one host thread per fast plan.  In place is
 * allowed.  Launches go to `stream` (NULL: the fast plan's own). */
typedef struct afft_fast_plan_s* afft_fast_plan;
int afft_fast_plan_create(afft_context ctx, int64_t n, int64_t batch, int precision, void* stream,
                          afft_fast_plan* out);
int64_t afft_fast_plan_len(afft_fast_plan fp);
void afft_fast_forward(afft_fast_plan fp, void* dst, const void* src);
void afft_fast_inverse(afft_fast_plan fp, void* dst, const void* src);
int afft_fast_plan_destroy(afft_fast_plan fp);

/* ---- ragged batches ------------------------------------------------------
 * No single counterpart in the reference: a caller there loops Plan.Forward
 * (plan.go:332) over plans of different sizes, as ForwardBatch loops one size
 * (plan.go:504-530).  Here `count` complex transforms of lengths sizes[i]
 * are packed back to back (job i starts at the sum of the earlier lengths)
 * and run as one persistent work-queue launch. */
typedef struct afft_ragged_s* afft_ragged;
int afft_ragged_create(afft_context ctx, int precision, const int64_t* sizes, int64_t count, int flags,
                       afft_ragged* out);
int64_t afft_ragged_total_len(afft_ragged r);                    /* elements of the packed layout */
int afft_ragged_offsets(afft_ragged r, int64_t* offsets, int64_t cap);
/* One launch of the persistent kernel (out of place, device pointers). */
int afft_ragged_exec(afft_ragged r, int direction, void* dst, const void* src, void* stream);
/* The same launch captured once per (direction, dst, src, stream) in a
 * hipGraph and replayed. */
int afft_ragged_exec_graph(afft_ragged r, int direction, void* dst, const void* src, void* stream);
/* The plain per-size path: one batched plan execution (Plan.ForwardBatch,
 * plan.go:504-530) per run of consecutive jobs of equal length.  With the
 * jobs sorted by length that is one batched plan per distinct size. */
int afft_ragged_exec_per_size(afft_ragged r, int direction, void* dst, const void* src, void* stream);
/* Every job through a batch-1 plan of its own size (Plan.Forward,
 * plan.go:332-369): one plan execution per job. */
int afft_ragged_exec_per_job(afft_ragged r, int direction, void* dst, const void* src, void* stream);
int afft_ragged_destroy(afft_ragged r);

/* ---- planning: candidates, autotuner, wisdom -------------------------------
 * The autotuner replaces MeasureAndSelect (internal/fft/measure.go:89-228);
 * wisdom keeps the reference's text format "size:precision:features:
 * algorithm:timestamp" (internal/planner/wisdom.go:96-205), with the GPU
 * architecture number (0x950 for gfx950) in the features field. */
/* Candidate strategies of a 1-D C2C shape; returns their count (arrays may
 * be NULL to query it). */
int afft_plan_candidates(const afft_plan_desc* desc, int* strategies, int64_t* splits, int cap);
/* The 32 cost-model features of (shape, candidate). */
int afft_plan_features(const afft_plan_desc* desc, int strategy, int64_t split, float* out32);
/* Milliseconds of one forward execution (device buffers): after `warmup`
 * untimed executions, the mean of `iters` timed ones, best of three rounds. */
int afft_tune_time(afft_context ctx, const afft_plan_desc* desc, int warmup, int iters, double* ms);
/* Times every candidate (times[i] = ms of candidate i of
 * afft_plan_candidates, -1 if it could not run) and stores the winner as
 * wisdom in the context. */
int afft_tune_plan(afft_context ctx, const afft_plan_desc* desc, int warmup, int iters, int* best_strategy,
                   int64_t* best_split, double* best_ms, double* times, int cap);
/* Features word of the wisdom key for a plan description: the device's
 * architecture number (0x950) in bits 0-15 and the shape class in bits 16-19
 * (batch bucket 0: 1, 1: 2-15, 2: 16-255, 3: 256-4095, 4: 4096 and more;
 * +8 for strided layouts), so that a winner is applied only to plans of the
 * class it was timed for (the field of planner/wisdom.go:17-30). */
uint64_t afft_wisdom_features(afft_context ctx, const afft_plan_desc* desc);
int afft_wisdom_export(afft_context ctx, char* buf, int64_t cap, int64_t* needed);  /* Wisdom.Export */
int afft_wisdom_import(afft_context ctx, const char* text);                          /* Wisdom.Import */
int afft_wisdom_clear(afft_context ctx);

/* ---- learned cost model -----------------------------------------------------
 * A 32->64->64->1 MLP predicting log runtime of (shape, candidate), trained
 * and evaluated by the library's own HIP kernels.  No counterpart in the
 * reference; it stands in front of the wisdom/timing planner. */
typedef struct afft_mlp_s* afft_mlp;
int afft_mlp_param_count(void);
int afft_mlp_create(afft_context ctx, uint64_t seed, afft_mlp* out);
int afft_mlp_set_normalization(afft_mlp m, const float* mean32, const float* std32, float ymean, float ystd);
int afft_mlp_get_params(afft_mlp m, float* out);
int afft_mlp_set_params(afft_mlp m, const float* in);
int afft_mlp_loss_grad(afft_mlp m, const float* x, const float* t, int64_t n, float* loss, float* grad);
int afft_mlp_train(afft_mlp m, const float* x, const float* t, int64_t n, int epochs, int batch, float lr,
                   uint64_t seed, float* final_loss);
int afft_mlp_predict(afft_mlp m, const float* x, int64_t n, float* y);
int afft_mlp_save(afft_mlp m, const char* path);   /* weights + normalisation + FNV-1a checksum */
int afft_mlp_load(afft_mlp m, const char* path);
int afft_mlp_destroy(afft_mlp m);
/* AUTO plans on this context ask the model when no wisdom entry matches. */
int afft_context_set_cost_model(afft_context ctx, afft_mlp m);

/* ---- convolution and correlation -------------------------------------------
 * Convolve / Convolve128 (convolve.go:9-115), ConvolveReal
 * (convolve_real.go:9-80) and CrossCorrelate / AutoCorrelate
 * (correlate.go:6-79) on the device.  Output length na+nb-1.
 *   AFFT_CONV_LINEAR     complex linear convolution
 *   AFFT_CONV_CORRELATE  complex cross-correlation (b reversed and
 *                        conjugated; output index k is lag k-(nb-1))
 *   AFFT_CONV_REAL       real linear convolution: a, b and dst are float
 *                        (AFFT_COMPLEX64) or double (AFFT_COMPLEX128) */
enum { AFFT_CONV_LINEAR = 0, AFFT_CONV_CORRELATE = 1, AFFT_CONV_REAL = 2 };
typedef struct afft_conv_s* afft_conv;
int afft_conv_create(afft_context ctx, int precision, int64_t na, int64_t nb, int64_t batch, int mode,
                     afft_conv* out);
/* Device pointers: a = batch rows of na, b = batch rows of nb, dst = batch
 * rows of na+nb-1 (complex elements, real for AFFT_CONV_REAL); enqueued on
 * stream (NULL = own). */
int afft_conv_exec(afft_conv c, void* dst, const void* a, const void* b, void* stream);
int afft_conv_destroy(afft_conv c);
/* Host slices, one convolution, synchronised; the checks and errors of
 * Convolve and ConvolveReal (nil -> ErrNilSlice, empty -> ErrInvalidLength,
 * dst length != na+nb-1 -> ErrLengthMismatch). */
int afft_convolve_host(afft_context ctx, int precision, int mode, void* dst, int64_t dst_len, const void* a,
                       int64_t na, const void* b, int64_t nb);

#ifdef __cplusplus
}
#endif

#endif /* ALGOFFT_HIP_H */
