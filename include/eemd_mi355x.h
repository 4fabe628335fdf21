/*
 * eemd_mi355x.h -- C ABI of the MI355X EEMD engine (libeemd_mi355x.so).
 *
 * Plain C: pointers, sizes, integer status codes; no exceptions or C++/torch
 * types cross this boundary.  Every entry point returns EEMD_OK (0) or an
 * EEMD_ERR_* code; eemd_last_error() gives the message of the context's last
 * failure.
 *
 * What each entry point replaces in the reference (paths relative to the
 * reference root):
 *
 *   eemd_run_batch      gpu_eemd_full(x[E,N], E, tau, max_sifts, max_imf, ...)
 *                       -> (imfs[E,maxm,N], nsift[E], nmode[E])
 *                       (emd_gpu_v5_dsbt.py:354-384, emd_gpu.py:301-334),
 *                       together with the host-side level-wise averaging of
 *                       emd_par._gpu_decompose (emd_par.py:225-250) and
 *                       emd_ref.eemd (emd_ref.py:199-217).  One call handles
 *                       many signals of different lengths (the reference
 *                       launches once per signal).
 *   eemd_run_batch_dev  the same on device-resident buffers (no H2D / D2H).
 *   eemd_noise          the noise matrix of emd_par.py:233 / emd_ref.py:204,
 *                       here the counter-based normals the kernels generate.
 *   eemd_ctx_create /   the kernel factory + buffer allocation of
 *   eemd_ctx_destroy    emd_gpu_v5_dsbt.py:137-168, 343-347, 367-376: device
 *                       scratch is owned by the context and reused.
 *
 * Error conventions (emd_gpu.py:307-315, emd_par.py:281-286): invalid shapes
 * or parameters -> EEMD_ERR_BAD_ARG (the Python layer raises ValueError);
 * device failures -> EEMD_ERR_DEVICE; allocation failures -> EEMD_ERR_OOM
 * (RuntimeError).
 *
 * Threading: one context per host thread.  Calls are synchronous on the
 * context's stream unless the _dev variant is given a stream, in which case
 * work is enqueued on that stream and the call returns after enqueueing.
 */
#ifndef EEMD_MI355X_H
#define EEMD_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EEMD_OK 0
#define EEMD_ERR_BAD_ARG 1
#define EEMD_ERR_DEVICE 2
#define EEMD_ERR_OOM 3

#define EEMD_END_NATURAL 0 /* emd_ref.py:77-103 end='natural' */
#define EEMD_END_MIRROR 1  /* emd_ref.py:92-95 end='mirror' */
#define EEMD_END_LEARNED 2 /* learned boundary extension, mirror fallback */

#define EEMD_FLAG_MEMBER_IMFS 1 /* also return every member's IMFs */
#define EEMD_FLAG_TRACE 2       /* per-sift extrema counts + index digests */
#define EEMD_FLAG_PHASES 4      /* per-workgroup phase clocks (eemd_last_phases) */
#define EEMD_FLAG_GRAPH 8       /* capture the launch sequence in a hipGraph and replay it on
                                   later calls with the same plan, parameters and buffers */
#define EEMD_FLAG_LEVEL_SUMS 16 /* imf_mean holds per-level sums over the members (not
                                   divided by the member counts) and residual is x minus
                                   their sum: for callers that combine batches of one
                                   ensemble, as emd_par.py:234-250 does */
#define EEMD_FLAG_TUNE_BLOCK 32 /* eemd_autotune may change the block-Thomas block size, and
                                   a later call with this flag and automatic launch
                                   parameters uses the tuned block (results then move by
                                   the seam truncation, at most 8 rho^S relative) */

typedef struct eemd_ctx eemd_ctx;

typedef struct eemd_params {
  int32_t fp64;             /* 1: FP64 arithmetic, 0: FP32 (spline solve + ensemble sum in FP64) */
  int32_t end_mode;         /* EEMD_END_* */
  int32_t max_imf;          /* 1..64 IMFs per member (emd_ref.py:125) */
  int32_t max_sifts;        /* sifts per IMF cap (emd_ref.py:130) */
  int32_t min_extrema;      /* stop when the residual has fewer extrema (emd_ref.py:123) */
  int32_t block;            /* decay-seamed block-Thomas block B; <= 0: exact Thomas */
  int32_t halo;             /* seam halo S in knots */
  int32_t threads;          /* workgroup size 64 / 256 / 512 / 1024; 0: automatic */
  int32_t members_per_task; /* ensemble members per queue entry; 0: automatic (one count for
                               the batch: the most members whose task on the longest signal
                               is at most one resident workgroup's share of the work) */
  int32_t flags;            /* EEMD_FLAG_* */
  double tau;               /* SD stopping tolerance (emd_ref.py:140-146); <= 0: fixed sifts */
  uint64_t seed;            /* Philox key when no explicit noise is given */
  int32_t wave_windows;     /* FP32: solve windows a wave at a time when a workgroup has at most
                               this many (0: automatic, < 0: never) */
  int32_t member0;          /* Philox member index of member 0 (batches of one ensemble) */
  int32_t signal0;          /* Philox signal index of signal 0 (a batch split over calls) */
} eemd_params;

/* Defaults: FP64, mirror, max_imf 10, max_sifts 10, min_extrema 4, B 1024, S 30,
 * automatic workgroup / task sizes, tau 0, seed 0, member0 0, signal0 0. */
void eemd_params_default(eemd_params* p);

int eemd_ctx_create(int device, eemd_ctx** out);
int eemd_ctx_destroy(eemd_ctx* ctx);
const char* eemd_last_error(const eemd_ctx* ctx);

/*
 * EEMD of nsig signals (concatenated in x; signal s is x[offsets[s] ..
 * offsets[s+1]), length >= 3) with E members each.  Member e of signal s
 * decomposes x_s + sigma_s * z_{s,e}, where z is the explicit standard-normal
 * matrix `noise` (per signal an [E][n_s] block, blocks concatenated in signal
 * order) or, when noise is NULL, the Philox normals of eemd_noise() for signal
 * index signal0 + s and members member0 ...
 *
 * Outputs (host memory, caller-owned):
 *   imf_mean   [sum_s max_imf * n_s]  per signal a [max_imf][n_s] block; levels
 *              >= n_modes[s] are zero
 *   residual   [sum_s n_s]            x - sum of the mean IMFs (emd_ref.py:217)
 *   n_modes    [nsig]                 levels produced by at least one member
 *   nmode      [nsig*E]               IMFs of each member
 *   nsift      [nsig*E]               total sifts of each member
 *   per_mode_sifts [nsig*E*max_imf]   optional (may be NULL)
 *   member_imfs    optional (may be NULL; requires EEMD_FLAG_MEMBER_IMFS):
 *              per signal an [E][max_imf][n_s] block of double (fp64) or
 *              float (fp32); levels >= nmode are zero
 *   trace      optional (may be NULL; requires EEMD_FLAG_TRACE):
 *              [nsig*E][max_imf][max_sifts][4] uint64 = {maxima count, minima
 *              count, digest of maxima indices, digest of minima indices} of
 *              the knots each executed sift used (digest: sum of splitmix64)
 */
int eemd_run_batch(eemd_ctx* ctx, const double* x, const int64_t* offsets, int32_t nsig,
                   int32_t E, const double* sigmas, const double* noise, const eemd_params* prm,
                   double* imf_mean, double* residual, int32_t* n_modes, int32_t* nmode,
                   int32_t* nsift, int32_t* per_mode_sifts, void* member_imfs, uint64_t* trace);

/* Same, with x, noise, imf_mean, residual, n_modes, nmode, nsift,
 * per_mode_sifts, member_imfs and trace in device memory (offsets and sigmas
 * stay on the host).  stream: hipStream_t, or NULL for the context's stream
 * (then the call waits for completion). */
int eemd_run_batch_dev(eemd_ctx* ctx, const double* x, const int64_t* offsets, int32_t nsig,
                       int32_t E, const double* sigmas, const double* noise,
                       const eemd_params* prm, double* imf_mean, double* residual,
                       int32_t* n_modes, int32_t* nmode, int32_t* nsift, int32_t* per_mode_sifts,
                       void* member_imfs, uint64_t* trace, void* stream);

/* Size of the allocation-free scratch: a static device array each GPU the
 * library runs on holds until the process ends. */
#define EEMD_NOALLOC_SCRATCH_BYTES (256ull << 20)

/* eemd_run_batch_dev without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as eemd_run_batch_dev; the
 * kernels' scratch (per-workgroup slabs, ensemble accumulators, the task
 * queue and descriptors) comes from the static EEMD_NOALLOC_SCRATCH_BYTES
 * array of the context's device instead of context buffers, so the launch
 * uses as many workgroups as fit in it.  EEMD_ERR_BAD_ARG when the batch
 * does not fit (one workgroup slab plus one accumulator block per signal)
 * or with EEMD_FLAG_GRAPH.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU from different contexts race (calls of one
 * context are ordered). */
int eemd_run_batch_dev_noalloc(eemd_ctx* ctx, const double* x, const int64_t* offsets,
                               int32_t nsig, int32_t E, const double* sigmas, const double* noise,
                               const eemd_params* prm, double* imf_mean, double* residual,
                               int32_t* n_modes, int32_t* nmode, int32_t* nsift,
                               int32_t* per_mode_sifts, void* member_imfs, uint64_t* trace,
                               void* stream);

/* Times candidate launch configurations on the given device-resident batch
 * (x and noise as for eemd_run_batch_dev; the timed runs write to context
 * buffers) and returns the fastest: *best = *prm with threads,
 * members_per_task, (FP32) wave_windows and, with EEMD_FLAG_TUNE_BLOCK in
 * prm->flags, block set; *best_ms its kernel time.  The context remembers the
 * choice for this batch shape (lengths, E, FP mode, IMF/sift counts, block,
 * halo, end mode) and applies it to later calls whose threads,
 * members_per_task and wave_windows are all 0 (the block only when such a
 * call also sets EEMD_FLAG_TUNE_BLOCK).  Coordinate search: workgroup size
 * (64 threads: one wave, one member at a time), then the block size, then
 * members per task, then the FP32 wave-solve threshold; each candidate runs
 * `reps` times and keeps its fastest run.  The reference picks its launch
 * shape and block size by hand (emd_gpu_v5_dsbt.py:343, K / nthreads / B
 * arguments of gpu_eemd_full). */
int eemd_autotune(eemd_ctx* ctx, const double* x, const int64_t* offsets, int32_t nsig, int32_t E,
                  const double* sigmas, const double* noise, const eemd_params* prm, int32_t reps,
                  eemd_params* best, double* best_ms);

/* The standard normals member e of signal `signal` receives when no explicit
 * noise is given: out[e][i] for members member0 .. member0+E-1 (host memory). */
int eemd_noise(eemd_ctx* ctx, uint64_t seed, int32_t signal, int32_t member0, int32_t E,
               int64_t n, double* out);

/* Device time (ms, HIP events on the launch stream) of the last call's
 * persistent sift kernel and mean kernel, and its launch configuration.
 * After an asynchronous eemd_run_batch_dev call this waits for its events.
 * A call replayed from a graph (EEMD_FLAG_GRAPH) reports the whole graph in
 * sift_ms and 0 in mean_ms. */
int eemd_last_timing(eemd_ctx* ctx, double* sift_ms, double* mean_ms, int32_t* threads,
                     int32_t* grid, int32_t* ntasks);

/* Free and total device memory (hipMemGetInfo) of the context's device: the
 * reference sizes its batches from the free memory (emd_par.py:208-222). */
int eemd_mem_info(eemd_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes);

/* Graph captures and graph launches (EEMD_FLAG_GRAPH calls) since the context
 * was created: a replayed call adds a launch without a capture. */
int eemd_graph_stats(eemd_ctx* ctx, int64_t* captures, int64_t* launches);

/* Mean per-workgroup time (ms) of the last call's phases when it ran with
 * EEMD_FLAG_PHASES: [0] block-Thomas solve, [1] fused envelope / update /
 * extrema passes, [2] segment seam fix-up and prefix, [3] end extension
 * setup, noise, finalisation and queue overhead.  No
 * counterpart in the reference (its phase split used an in-kernel clock,
 * dbg_serial_split_v0.py:25-42). */
int eemd_last_phases(eemd_ctx* ctx, double* ms4);

/* Instantaneous frequency (Hz) and amplitude of each row of x [rows][n]
 * (host memory), as audio_eemd/spectra.py:32-48 computes them per IMF:
 * z = x + i H(x) (scipy.signal.hilbert: FFT, one-sided spectrum, inverse
 * FFT; hipFFT here), inst_freq = np.gradient(np.unwrap(np.angle(z)), 1/sr)
 * / (2 pi), inst_amp = |z|.  n in [2, 2^27]. */
int eemd_hilbert_if(eemd_ctx* ctx, const double* x, int64_t rows, int64_t n, double sr,
                    double* inst_freq, double* inst_amp);

/* STFT of each row of x [rows][n] (host memory) as audio_eemd/spectra.py:18-29
 * asks scipy.signal.stft for it: periodic Hann window of nperseg, noverlap
 * overlap, nperseg/2 zeros added at both ends, no padding to whole frames,
 * one-sided spectrum scaled by 1 / sum(window).  zxx: [rows][nperseg/2+1]
 * [nseg] complex as interleaved (re, im) doubles, nseg = (n + 2*(nperseg/2)
 * - noverlap) / (nperseg - noverlap); frame k is centred on sample
 * k*(nperseg-noverlap). */
int eemd_stft(eemd_ctx* ctx, const double* x, int64_t rows, int64_t n, int32_t nperseg,
              int32_t noverlap, double* zxx);

/* ------------------------------------------------------------------------
 * Learned boundary extension.  No counterpart in the reference: its GPU
 * kernels use natural ends and its CPU oracle the mirror rule
 * (emd_ref.py:92-95), which EEMD_END_LEARNED falls back to when an envelope
 * has fewer than 16 extrema.
 *
 * Model: MLP 64 -> 128 -> 128 -> 128 -> 16, GELU (tanh form), BF16
 * parameters.  Input per end: the 16 extrema nearest to the end on both
 * envelopes (gaps scaled by their mean, values standardised); output: 4
 * virtual extrema per envelope beyond the end.  Parameters are FP32, packed
 * per layer as W[in][out] then bias[out] (EEMD_MLP_NPARAM floats).
 * ------------------------------------------------------------------------ */
#define EEMD_MLP_NPARAM 43408

typedef struct eemd_mlp eemd_mlp;

/* Install the boundary model used by EEMD_END_LEARNED calls on this context
 * (parameters rounded to BF16). */
int eemd_set_boundary_model(eemd_ctx* ctx, const float* params);

/* Training state on one device: FP32 master parameters, Adam moments, BF16
 * copy; LeCun-uniform initial weights from `seed`, zero biases. */
int eemd_mlp_create(int device, uint64_t seed, eemd_mlp** out);
int eemd_mlp_destroy(eemd_mlp* m);
const char* eemd_mlp_last_error(const eemd_mlp* m);
/* Host FP32 parameters in / out (set also resets the Adam state). */
int eemd_mlp_set(eemd_mlp* m, const float* params);
int eemd_mlp_get(eemd_mlp* m, float* params);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on the mean squared error of a
 * batch (multiple of 16): x [batch][64], y [batch][16] device memory;
 * loss (host, optional) is the batch loss before the step. */
int eemd_mlp_step(eemd_mlp* m, const float* x, const float* y, int32_t batch, float lr,
                  float* loss);
/* Gradient of the batch loss w.r.t. the parameters (host, EEMD_MLP_NPARAM)
 * without updating them. */
int eemd_mlp_grad(eemd_mlp* m, const float* x, const float* y, int32_t batch, float* grad,
                  float* loss);
/* Inference with the arithmetic of the in-kernel extension: x [batch][64],
 * y [batch][16], device memory. */
int eemd_mlp_infer(eemd_mlp* m, const float* x, int32_t batch, float* y);

#ifdef __cplusplus
}
#endif

#endif /* EEMD_MI355X_H */
