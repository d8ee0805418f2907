/*
 * gmmhip.h -- C ABI of the MI355X (gfx950) diagonal-GMM statistics
 * accumulator.  Plain C: opaque handle, plain pointers and sizes, int status
 * codes.  One handle drives one GPU through one HIP stream and is not
 * thread-safe.
 *
 * Each entry point names the reference interface it replaces.  Reference
 * paths are relative to GMM/lib_bmssr/ of GMMTeam/GMM.
 */
#ifndef GMMHIP_H
#define GMMHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gmmhip_ctx *gmmhip_t;

/* Status codes.  The C++ adapter rethrows every non-zero code as
 * std::runtime_error (or std::logic_error for GMMHIP_ERR_ARG/STATE), which is
 * what the reference's checkError / NaN checks / insertModel throw
 * (trainer/CU_GMMStatsEstimator_CUDA.cu:1036-1053,
 * trainer/OL_GMMStatsEstimator_CUDA.cpp:383-387, :455-456,
 * trainer/OL_GMMStatsEstimator.cpp:279-285). */
enum {
    GMMHIP_OK = 0,
    GMMHIP_ERR_HIP = 1,      /* a HIP runtime call failed */
    GMMHIP_ERR_OOM = 2,      /* device allocation failed / over budget */
    GMMHIP_ERR_NAN = 3,      /* total log-likelihood is NaN */
    GMMHIP_ERR_ARG = 4,      /* bad argument (dims, null pointer, sizes) */
    GMMHIP_ERR_VAR = 5,      /* model variance below min_var */
    GMMHIP_ERR_STATE = 6,    /* no model / no frames inserted */
    GMMHIP_ERR_MLLR = 7,     /* MLLR under-determined (CheckInvertibility) */
    GMMHIP_ERR_NODEV = 8,    /* no HIP device */
    GMMHIP_ERR_TIMEOUT = 9,  /* a wait ran out of time */
    GMMHIP_ERR_CLOSED = 10   /* the server is shutting down */
};

/* Arithmetic of the E-step contractions. */
enum {
    GMMHIP_PREC_FP32 = 0,    /* FP32 MFMA throughout (plain form) */
    GMMHIP_PREC_BF16X3 = 1,  /* split bf16 hi+lo operands, FP32 accumulate */
    GMMHIP_PREC_FP16X3 = 2   /* split fp16 hi+lo operands, FP32 accumulate */
};

/* Accumulation flags, setAccFlags(mixProbsOnly, meanStats, varStats, ...)
 * (trainer/OL_GMMStatsEstimator.h:118-123).  n is always produced. */
enum {
    GMMHIP_ACC_MEAN = 1,     /* export F = sum g*x   */
    GMMHIP_ACC_VAR = 2,      /* export S = sum g*x*x */
    GMMHIP_ACC_FULLVAR = 4   /* with ACC_VAR on a diagonal model: S = packed
                                upper triangle of sum g x x^T (setAccFlags'
                                varFull, used by fMLLR); FP32 */
};

/* Options, the reference's GMMStatsEstimator_GPU::OPTIONS
 * (trainer/CU_GMMStatsEstimator_CUDA.h:56-63) plus the estimator's public
 * knobs _minLogLike/_minGamma/_min_var (trainer/OL_GMMStatsEstimator.h:61-66).
 * Zero-initialised fields take the defaults noted. */
typedef struct gmmhip_opts {
    float min_loglike;   /* default -1e20f */
    float min_gamma;     /* default 1e-4f  */
    float min_var;       /* default 1e-6f  */
    int precision;       /* GMMHIP_PREC_*, default FP32 */
    int chunk_frames;    /* frames per FP32 partial accumulator, 0 = auto */
    int verbosity;       /* 0 silent */
    int profile;         /* 1: time every kernel with HIP events */
    int ll_frames;       /* FP32 log-likelihood frame tile: 64, 128, 256;
                            0 = tuning table, else 128 */
    int comp_tile;       /* mixed-form statistics component tile: 128 or 256
                            components per block (256 at mixed depths 64
                            and 128 only); 0 = tuning table, else 128 */
} gmmhip_opts;

/* Fill *o with the defaults above. */
void gmmhip_default_opts(gmmhip_opts *o);

/* Number of HIP devices; replaces GMMStatsEstimator_GPU::deviceCount
 * (trainer/CU_GMMStatsEstimator_CUDA.h:199). */
int gmmhip_device_count(int *count);

/* Create a handle on `device`; replaces the GMMStatsEstimator_GPU(int GPU_id)
 * constructor (trainer/CU_GMMStatsEstimator_CUDA.h:65) and
 * updateOptionsOnGPU (trainer/OL_GMMStatsEstimator_CUDA.cpp:114-123).
 * opts may be NULL (defaults). */
int gmmhip_create(int device, const gmmhip_opts *opts, gmmhip_t *out);

/* Destroy a handle and free its device memory; replaces
 * ~GMMStatsEstimator_GPU (trainer/CU_GMMStatsEstimator_CUDA.h:66). */
int gmmhip_destroy(gmmhip_t h);

/* Replace the options; replaces updateOptionsOnGPU
 * (trainer/OL_GMMStatsEstimator_CUDA.cpp:114-123). */
int gmmhip_set_opts(gmmhip_t h, const gmmhip_opts *opts);

/* Text of the last error on this handle ("" if none). */
const char *gmmhip_last_error(gmmhip_t h);

/* Insert a diagonal model of M active components (all weights > 0, already
 * in the order produced by GMModel::RearrangeMixtures), D dims, row-major
 * mu/var [M*D].  Computes log-weights, inverse variances and Gaussian
 * constants (insertModel, trainer/OL_GMMStatsEstimator.cpp:234-313), packs
 * them for the MFMA contraction and uploads them (the CUDA insertModel +
 * uploadModel, trainer/OL_GMMStatsEstimator_CUDA.cpp:127-276,
 * trainer/CU_GMMStatsEstimator_CUDA.cu:284-318).  Returns GMMHIP_ERR_VAR if a
 * variance is below min_var. */
int gmmhip_set_model(gmmhip_t h, int M, int D, const float *w,
                     const float *mu, const float *var);

/* Insert a full-covariance model: M active components (weights > 0, in
 * RearrangeMixtures order), D <= 127 dims, row-major mu [M*D] and
 * symmetric cov [M*D*D].  Replaces insertModel's full-covariance branch
 * (trainer/OL_GMMStatsEstimator.cpp:243-299: getCovInverse / getLogDet by
 * Cholesky, :156-229) + uploadModel with alignedDimVar = D*D
 * (trainer/CU_GMMStatsEstimator_CUDA.cu:284-318).  Afterwards
 * gmmhip_accumulate exports S as the packed upper triangle of sum g x x^T,
 * D(D+1)/2 per component in the reference's GMMStats order (row k holds
 * columns k..D-1, trainer/OL_GMMStatsEstimator.cpp:1358-1364), and
 * gmmhip_loglikes works as for a diagonal model.  FP32 only (opts.precision
 * must be FP32); not usable with gmmhip_batch_adapt.  A diagonal model
 * inserted later replaces it.  Returns GMMHIP_ERR_VAR when a covariance is
 * not positive definite ("cholesky factorization failed"). */
int gmmhip_set_model_full(gmmhip_t h, int M, int D, const float *w,
                          const float *mu, const float *cov);

/* Upload N frames of D floats given as row pointers (the reference's
 * float** vectors; rows may be non-contiguous) into device memory, where they
 * stay resident until replaced.  Replaces alignVectors + uploadParam
 * (trainer/OL_GMMStatsEstimator_CUDA.cpp:292-328,
 * trainer/CU_GMMStatsEstimator_CUDA.cu:431-470). */
int gmmhip_set_frames(gmmhip_t h, const float *const *rows, uint64_t N, int D);

/* Same, from one contiguous host buffer with row stride ld (floats). */
int gmmhip_set_frames_contig(gmmhip_t h, const float *x, uint64_t N, int D,
                             uint64_t ld);

/* Use N frames already resident in device memory at dx (row stride ld
 * floats).  The memory is borrowed: it must stay valid while the handle uses
 * it.  No reference counterpart (the reference always uploads from host); it
 * is the same call with the upload removed. */
int gmmhip_set_frames_device(gmmhip_t h, const float *dx, uint64_t N, int D,
                             uint64_t ld);

/* Accumulate statistics of the resident frames under the inserted model:
 * n[m] += sum_t g_tm, F[m*D+d] += sum_t g_tm x_td, S[m*D+d] += sum_t g_tm x_td^2
 * over posteriors g > min_gamma, *tot_ll += sum_t LL_t, *nframes += N.
 * Accumulate-into semantics like GMMStats (caller resets).  F/S may be NULL
 * when the matching flag is off.  Replaces accumulateStatsMT
 * (trainer/OL_GMMStatsEstimator_CUDA.cpp:396-511) = compAccStats +
 * getAccStats + export2GMMStats (trainer/CU_GMMStatsEstimator_CUDA.cu:721-865,
 * trainer/OL_GMMStatsEstimator_CUDA.cpp:593-648).  Returns GMMHIP_ERR_NAN if
 * the total log-likelihood is NaN. */
int gmmhip_accumulate(gmmhip_t h, int flags, double *n, double *F, double *S,
                      double *tot_ll, uint64_t *nframes);

/* No counterpart in the reference: for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The plain FP32 E-step of gmmhip_accumulate on N frames
 * [N][D] in DEVICE memory (row-major, borrowed for the call), adding into
 * DEVICE outputs n[M], F[M*D], S[M*D] and tot_ll[1] (FP64) and into the host
 * counter *nframes, with no device or host allocation, no host
 * synchronisation and no memset: the work is enqueued on the handle's
 * stream (gmmhip_stream / gmmhip_sync), its scratch is a file-scope device
 * array of GMMHIP_NOALLOC_SCRATCH_BYTES per device, held until the process
 * ends.  Results: n, F, S bit-identical to gmmhip_accumulate in FP32 on the
 * same frames; tot_ll equal to FP64 rounding.  Runs the FP32 form whatever
 * opts.precision says.  Limits: Mp x KP <= 524288 (Mp = M rounded up to 16,
 * KP = 2D+1 rounded up to 16; e.g. M <= 4096 at D <= 63, M <= 2048 at
 * D <= 127), GMMHIP_ERR_ARG beyond; any N.  Needs a diagonal model and no
 * Gaussian-selection mask.  Every call on one device shares the scratch
 * array: calls that overlap on one device (two streams, two threads) race.
 * The host cannot see a NaN total here: check *tot_ll after syncing. */
#define GMMHIP_NOALLOC_SCRATCH_BYTES 25297152u /* 24.1 MiB */
int gmmhip_accumulate_noalloc(gmmhip_t h, const float *frames, uint64_t N, int D, int flags,
                              double *n, double *F, double *S, double *tot_ll,
                              uint64_t *nframes);

/* Robust auxiliary statistics of the resident frames under the inserted
 * model (diagonal or full), added into aux[M*(D+2)]: per component
 * [aux_d (D) | aux2 | aux3] with, over every consecutive frame pair
 * (t-1, t) of the frame set and posteriors g before the cut,
 *   p_m = g_t,m g_t-1,m,  s = sum_m [p_m > min_gamma] p_m,  h_m = p_m / s,
 *   aux2_m += sqrt(g_t,m) (g_t,m > min_gamma),
 *   aux_m[d] += h_m |x_td - x_t-1,d|, aux3_m += h_m  (h_m > min_gamma).
 * Replaces accumulateAuxStats (trainer/OL_GMMStatsEstimator.cpp:1226-1280,
 * as accumulateStats drives it, :1174-1189) and the GPU's
 * normAuxGammasKernel + accAuxStatsKernel + the aux part of getAccStats
 * (trainer/CU_GMMStatsEstimator_CUDA.cu:922-940, :861-862; same
 * [M][D+2] layout).  Pairs run over the whole frame set: they do not break
 * at thread or upload-chunk boundaries as the reference's do, and the first
 * frame of the set starts the first pair.  Computed in FP32 with FP64
 * sums whatever opts.precision says.  The statistics n/F/S come from
 * gmmhip_accumulate, as before. */
int gmmhip_accumulate_aux(gmmhip_t h, double *aux);

/* Gaussian selection: restrict every frame's log-likelihood, posteriors and
 * statistics to a list of components (the reference's logLikeMask in "use"
 * mode, mask2use + maskIdxs, trainer/OL_GMMStatsEstimator.h:73-114,
 * trainer/OL_GMMStatsEstimator.cpp:444-512 with mixIdxs; its CUDA estimator
 * ignores the mask, OL_GMMStatsEstimator_CUDA.cpp:352, :421).  Frame t's
 * components are idx[offsets[t] .. offsets[t+1]) (model order, < M); a
 * frame with none is clamped (LL = min_loglike, no statistics).  The mask
 * belongs to the inserted model and N resident frames: accumulate,
 * loglikes and posteriors fail with GMMHIP_ERR_ARG while it does not fit
 * them; offsets = NULL removes it.  While set, those calls run the GEMM
 * form in FP32 (whatever opts.precision says); gmmhip_batch_adapt and the
 * server do not use it. */
int gmmhip_set_mask(gmmhip_t h, uint64_t N, const uint64_t *offsets, const int *idx);

/* Per-frame log-likelihoods ll[N] and posteriors before the cut
 * post[N*M] (host, row-major) of the resident frames: the reference's
 * logLikeMask in "fill" mode (mask2fill, trainer/OL_GMMStatsEstimator.cpp:
 * 1168-1171; unsupported by its CUDA estimator, OL_GMMStatsEstimator_CUDA.cpp:
 * 352-353).  A clamped frame gets min_gamma in every component (1 when
 * M = 1), as compGammas leaves it.  FP32 GEMM form; a set mask applies. */
int gmmhip_posteriors(gmmhip_t h, float *ll, float *post);

/* Per-frame log-likelihoods of the resident frames into out[N] (host).
 * Replaces compVecsLogLikeMT (trainer/OL_GMMStatsEstimator_CUDA.cpp:333-391)
 * = compLogLikes + getLogLikes (trainer/CU_GMMStatsEstimator_CUDA.cu:606-643),
 * without the reference's truncation beyond 262,136 frames per upload. */
int gmmhip_loglikes(gmmhip_t h, float *out);

/* Device-side variant: per-frame log-likelihoods into device memory d_out[N]
 * on the handle's stream (no host copy). */
int gmmhip_loglikes_device(gmmhip_t h, float *d_out);

/* Kernel timers of a handle created with opts.profile = 1, like the
 * reference's GPU_CUDA_TIMER accumulators
 * (trainer/CU_GMMStatsEstimator_CUDA.cu:50-55, :338-352): ms[i] receives the
 * summed milliseconds and count[i] the number of launches of kernel class i
 * (GMMHIP_KT_*) since the handle was created or the timers were reset.
 * n is the capacity of both arrays. */
enum {
    GMMHIP_KT_LOGLIK = 0,    /* per-frame log-sum-exp kernel */
    GMMHIP_KT_STATS = 1,     /* posteriors + statistics kernel */
    GMMHIP_KT_REDUCE = 2,    /* FP64 partial reduction kernels */
    GMMHIP_KT_COUNT = 3
};
int gmmhip_kernel_times(gmmhip_t h, double *ms, uint64_t *count, int n);
int gmmhip_reset_kernel_times(gmmhip_t h);

/* ---- Batched per-utterance adaptation --------------------------------
 * Many utterances of uneven length against the inserted model (the UBM):
 * per-utterance occupancies and first-order statistics, MAP-adapted means
 * (UpdateModelEM/UpdateModelMAP with --mean, trainer/OL_ModelAdapt.cpp:
 * 993-1106) and global MLLR transforms (ComputeW_MLLR with ComputeKiGi_MLLR,
 * prepareKiGiInitMatrices and CheckInvertibility, OL_ModelAdapt.cpp:
 * 1417-1581, :1851-1878).  The reference adapts one model per process
 * (projects/Model_ADAPT/Model_ADAPT.cpp:173-240); this is its batched form. */
enum {
    GMMHIP_BATCH_QUEUE = 0,          /* one persistent work-queue launch */
    GMMHIP_BATCH_PER_UTTERANCE = 1   /* one launch set per utterance */
};

typedef struct gmmhip_batch_cfg {
    int tau;                /* MAP relevance factor (an int, as --tau) */
    int mllr_prior;         /* 1: UBM prior (--mllr-init) */
    int mode;               /* GMMHIP_BATCH_* */
    int chunk_frames;       /* frames per work item, 0 = auto */
    int graph;              /* 1: capture the device pipeline in a hipGraph and
                               replay it for batches of the same shape and
                               buffers */
    int outputs_on_device;  /* 1: the output pointers are device memory */
} gmmhip_batch_cfg;

/* MLLR outputs need M <= 4096 (GMMHIP_ERR_ARG otherwise). */
typedef struct gmmhip_batch_out {
    float *means;           /* [U][M][D] MAP-adapted means, or NULL */
    double *occ;            /* [U][M] occupancies n, or NULL */
    double *first;          /* [U][M][D] first-order statistics F, or NULL */
    double *mllr_W;         /* [U][D][D+1] MLLR transforms [A | b], or NULL */
    int *mllr_status;       /* [U] 0 or GMMHIP_ERR_MLLR; required with mllr_W */
} gmmhip_batch_out;

/* frames: the U utterances concatenated, [offsets[U]][D] row-major, in host
 * memory or (frames_on_device = 1) device memory; offsets[U+1] (host) with
 * offsets[0] = 0.  The frames become the handle's current frames. */
int gmmhip_batch_adapt(gmmhip_t h, const float *frames, int frames_on_device,
                       const uint64_t *offsets, int U, int D, const gmmhip_batch_cfg *cfg,
                       gmmhip_batch_out *out);
/* Number of times the handle has captured the batch pipeline into a
 * hipGraph (cfg.graph = 1).  A graph is replayed only while the batch shape,
 * the model and every device buffer it captured are unchanged; any device
 * reallocation in between (a larger loglikes / auxiliary / full-covariance
 * call on the same handle, for example) makes the next call capture anew.
 * Diagnostic; no counterpart in the reference. */
int gmmhip_batch_graph_captures(gmmhip_t h, uint64_t *captures);

/* ---- Tuning table -----------------------------------------------------
 * Tile choices per (path, precision, D, M) picked by a timed sweep
 * (python -m gmm_amd.autotune) and stored as text in
 * gmm_amd/tuning/gfx950.tune (or $GMMHIP_TUNE_FILE): the log-likelihood
 * frame tile, the most frames per statistics work item and (mixed forms)
 * the statistics component tile.  Used when opts.ll_frames /
 * opts.chunk_frames (cfg->chunk_frames for batches) / opts.comp_tile are 0.
 * No counterpart in the reference, whose launch shapes are compile-time
 * constants (trainer/CU_GMMStatsEstimator_CUDA.cu:606-643). */
enum { GMMHIP_TUNE_ESTEP = 0, GMMHIP_TUNE_BATCH = 1 };
/* Reload the table from path (NULL: the default); returns the number of
 * entries, or -1 if the file cannot be read (the table is then empty). */
int gmmhip_tune_load(const char *path);
/* The table's choice for (path, precision, D, M), nearest M (then D) when
 * there is no exact entry; GMMHIP_ERR_STATE when it has none for the path
 * and precision. */
int gmmhip_tune_lookup(int path, int precision, int D, int M, int *ll_frames,
                       int *chunk_frames, int *comp_tile);
/* Tiles the handle's last accumulate / loglikes / batch call used
 * (comp_tile: components per mixed statistics block, 0 when no mixed
 * statistics kernel ran). */
int gmmhip_tiles_used(gmmhip_t h, int *ll_frames, int *chunk_frames, int *comp_tile);

/* ---- Serving loop -------------------------------------------------------
 * A server keeps one UBM resident in HBM and answers requests of three kinds
 * and any length, submitted from any thread:
 *   SCORE  the utterance's log-likelihood (sum over frames, natural log) and
 *          optionally its per-frame values -- the reference's getLogLike,
 *          projects/GetLogLikeGMM/main.cpp:150-188;
 *   MAP    MAP-adapted means (relevance factor tau) -- Model_ADAPT with
 *          --type 1 --mean, projects/Model_ADAPT/Model_ADAPT.cpp:173-240 and
 *          trainer/OL_ModelAdapt.cpp:993-1106;
 *   MLLR   the global MLLR transform [A | b] -- ComputeW_MLLR,
 *          trainer/OL_ModelAdapt.cpp:1417-1581.
 * A dispatcher thread admits waiting requests (first come, first served)
 * into a batch while the batch's device memory fits the slot budget, and
 * launches it on the slot's stream; up to `slots` batches run at once, so
 * requests arriving while others are on the GPU are admitted as soon as a
 * slot frees (continuous batching).  A batch runs the persistent work-queue
 * pipeline of gmmhip_batch_adapt.  Work items never straddle requests and
 * every per-request reduction has a fixed order, so a request's results are
 * bit-identical whichever requests share its batch (and equal to
 * gmmhip_batch_adapt with the same chunk_frames).  The reference has no
 * serving mode. */
typedef struct gmmhip_server *gmmhip_server_t;

enum {
    GMMHIP_REQ_SCORE = 0,
    GMMHIP_REQ_MAP = 1,
    GMMHIP_REQ_MLLR = 2
};
#define GMMHIP_REQ_PENDING (-1)

typedef struct gmmhip_server_cfg {
    gmmhip_opts opts;        /* thresholds and precision (FP32 or mixed; the
                                mixed form centres and scales frames by the
                                UBM's global mean and deviation) */
    int slots;               /* batches in flight, 0 = 2 */
    uint64_t slot_bytes;     /* device budget of one batch: frames, per-frame
                                values, partial sums and outputs; 0 = 1 GiB.
                                Each slot also pins slot_bytes of host
                                memory for its staging (a batch's input and
                                output regions are disjoint parts of it) */
    int max_batch;           /* requests per batch, 0 = 1024 */
    int chunk_frames;        /* frames per statistics work item, 0 = 512 */
    int tau;                 /* MAP relevance factor */
    int mllr_prior;          /* 1: UBM prior on the MLLR statistics */
} gmmhip_server_cfg;

/* One request.  The caller owns it and the buffers it points to until the
 * server has finished it (status != GMMHIP_REQ_PENDING). */
typedef struct gmmhip_request {
    int kind;                /* GMMHIP_REQ_* */
    const float *frames;     /* [n_frames][D] host memory */
    uint64_t n_frames;
    float *frame_ll;         /* optional [n_frames] per-frame log-likelihood */
    float *means;            /* MAP: [M][D], required */
    double *occ;             /* MAP / MLLR: optional [M] occupancies */
    double *mllr_W;          /* MLLR: [D][D+1], required */
    /* written by the server */
    double ll_sum;           /* sum over frames of the frame log-likelihood */
    double latency_ms;       /* submission to completion */
    volatile int status;     /* GMMHIP_REQ_PENDING, then 0 or GMMHIP_ERR_*
                                (GMMHIP_ERR_MLLR: transform under-determined,
                                mllr_W holds zeros) */
} gmmhip_request;

typedef struct gmmhip_server_stats {
    uint64_t submitted, completed, failed, batches, frames;
    uint64_t fp32_fallbacks;   /* FP16 form: requests outside the FP16 range
                                  after centring and scaling, run in FP32 */
    double p50_ms, p90_ms, p99_ms, p999_ms, max_ms, mean_ms;  /* latency */
    double mean_batch_requests;
    double mean_batch_gpu_ms;  /* device time of one batch (event pair) */
    double wall_s;             /* since creation or the last reset */
    int precision;             /* GMMHIP_PREC_* in use (FP32 if the model's
                                  split coefficients overflow FP16) */
} gmmhip_server_stats;

/* M active components (weights w[M], means mu[M][D], variances var[M][D]),
 * M <= 4096. */
int gmmhip_server_create(int device, const gmmhip_server_cfg *cfg, int M, int D, const float *w,
                         const float *mu, const float *var, gmmhip_server_t *out);
/* Queue a request (non-blocking).  Sets r->status = GMMHIP_REQ_PENDING. */
int gmmhip_server_submit(gmmhip_server_t s, gmmhip_request *r);
/* Block until r is finished or timeout_ms (< 0: no limit) passes; returns
 * r->status, or GMMHIP_ERR_TIMEOUT. */
int gmmhip_server_wait(gmmhip_server_t s, gmmhip_request *r, double timeout_ms);
/* Block until every submitted request is finished. */
int gmmhip_server_drain(gmmhip_server_t s);
/* Latency percentiles over the requests completed since creation or the
 * last reset (reset = 1 clears after reading). */
int gmmhip_server_stats_get(gmmhip_server_t s, gmmhip_server_stats *st, int reset);
const char *gmmhip_server_last_error(gmmhip_server_t s);
/* Finishes queued and running requests, then frees the server. */
void gmmhip_server_destroy(gmmhip_server_t s);

/* The handle's HIP stream (hipStream_t as void*), for callers that order
 * their own device work with the handle's. */
void *gmmhip_stream(gmmhip_t h);

/* Synchronise the handle's stream. */
int gmmhip_sync(gmmhip_t h);

#ifdef __cplusplus
}
#endif

#endif /* GMMHIP_H */
