/*
 * mirorr_gpu.h -- C ABI of the MI355X block-matching registration library.
 *
 * Plain pointers and sizes only (no C++ or framework types), so that the
 * reference's host code (C++/ITK) or any FFI can bind it. Every entry point
 * returns a mirorr_status; nothing throws across the boundary. Status codes
 * mirror the reference's exception classes:
 *   MIRORR_ERR_ALGO      <- BlockMatcherGPUExceptionAlgo / std::runtime_error
 *                           "No block matches" (itkBlockMatcherGPU.h:57-71,
 *                           itkBlockMatcher.txx:936-943)
 *   MIRORR_ERR_PLATFORM  <- BlockMatcherGPUExceptionPlatform, on which the
 *                           reference falls back to its CPU matcher
 *                           (itkMirorrRegistrationMethod.txx:363-385)
 *   MIRORR_ERR_ARG       <- std::out_of_range / invalid parameters
 *                           (itkAbstractBlockMatcher.txx:83-99)
 *
 * Volumes are float32 with x fastest, then y, then z (ITK raster order);
 * masks are uint8 (non-zero = inside), NULL meaning all 255 as the reference
 * fills default masks (itkBlockMatcher.txx:507-529).
 *
 * Threading: one handle per host thread. All work of a context runs on the
 * context's own HIP stream; calls that return host data synchronise it.
 */
#ifndef MIRORR_GPU_H
#define MIRORR_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MIRORR_ABI_VERSION 1

typedef enum mirorr_status {
  MIRORR_OK = 0,
  MIRORR_ERR_ARG = 1,
  MIRORR_ERR_PLATFORM = 2,
  MIRORR_ERR_ALGO = 3,
  MIRORR_ERR_NOMEM = 4,
  MIRORR_ERR_STATE = 5,
  MIRORR_PENDING = 6 /* mirorr_poll: the request has not finished yet */
} mirorr_status;

typedef struct mirorr_ctx mirorr_ctx;
typedef struct mirorr_bm mirorr_bm;

/* Image geometry: size, spacing (mm), origin (mm) and row-major direction
 * cosines, as itk::Image carries them. */
typedef struct mirorr_geom {
  int dims[3];
  double spacing[3];
  double origin[3];
  double direction[9];
} mirorr_geom;

/* Block-matcher parameters (itkAbstractBlockMatcher.h:101-127, defaults at
 * itkAbstractBlockMatcher.txx:24-43). */
typedef struct mirorr_bm_params {
  int block_width;       /* BW, voxels per block edge          (default 4)   */
  int block_gap;         /* BG, stride between candidates      (default 1)   */
  int nhood_width;       /* NW, candidates per axis            (default 5)   */
  int nhood_gap;         /* NG, stride between blocks          (default 2)   */
  double portion_kept;   /* PortionMatchesKept in (0,1]        (default 0.5) */
  int variance_select;   /* 0: reference behaviour (no culling, the culling
                            code is disabled at itkBlockMatcher.txx:664-719);
                            1: keep the top portion_kept blocks by variance */
  int verbosity;         /* 0..3                                             */
  int block_metric;      /* MIRORR_METRIC_NC (default), _SSD or _CR          */
  int exact_rescore;     /* 1 (default): blocks of the float NCC search
                            (BW 4, BG 1, NW 5-13) whose float best or
                            best-second gap lies next to the score filter's
                            thresholds (1e-7, 1e-4), and exact ties, are
                            searched again in double with the reference's
                            operation order, so their records and the kept
                            set are the reference's; 0: float scores as
                            searched (for measuring their error)          */
  int storage;           /* MIRORR_STORAGE_F32 (default) or _F16: how
                            mirorr_bm_set_volume keeps the images on the
                            device (see mirorr_volume_create_storage)      */
} mirorr_bm_params;

/* Block metrics (--blockmetric nc | sd | cr; BlockMatcher's metric map,
 * itkBlockMatcher.h:316-323). The search keeps the largest score under every
 * metric, as the reference does. */
#define MIRORR_METRIC_NC 0  /* normalizedCorrelation, itkBlockMatcher.txx:46-101  */
#define MIRORR_METRIC_SSD 1 /* sumOfSquareDifference, itkBlockMatcher.txx:195-227 */
#define MIRORR_METRIC_CR 2  /* correlationRatio, itkBlockMatcher.txx:105-191      */

enum { MIRORR_ROLE_BASE = 0, MIRORR_ROLE_SEARCH = 1 };

/* ---- context --------------------------------------------------------- */

/* Replaces the OpenCL context/queue set-up of BlockMatcherGPU::Initialize
 * (itkBlockMatcherGPU.txx:597-723). hbm_budget = 0 means no limit. */
mirorr_status mirorr_ctx_create(int device, size_t hbm_budget, mirorr_ctx** out);
void mirorr_ctx_destroy(mirorr_ctx* ctx);
/* Message of the last failed call on this context (never NULL). */
const char* mirorr_last_error(const mirorr_ctx* ctx);
/* Waits for all work queued on the context's stream. */
mirorr_status mirorr_ctx_synchronize(mirorr_ctx* ctx);
/* The context's HIP stream (hipStream_t) for callers that interleave work. */
void* mirorr_ctx_stream(mirorr_ctx* ctx);

/* Iteration graphs (no counterpart in the reference, whose outer loop
 * itkSymMirorrRegistrationMethod.txx:456-692 runs on the host): with `on`,
 * mirorr_register captures an outer iteration of each level (the
 * resampling, both block-matching passes, the LTS start, its first four
 * inner iterations and the finish) as a HIP graph at the level's second
 * iteration, keeps it on the context and replays it for the later
 * iterations, also of later registrations, while every host value its
 * launches read is unchanged; further inner iterations, when the LTS needs
 * them, are queued as usual. Results are identical either way. Off by
 * default, or as the environment variable MIRORR_ITERATION_GRAPH says at
 * context creation; ignored while per-kernel timing is on. */
mirorr_status mirorr_ctx_set_iteration_graphs(mirorr_ctx* ctx, int on);
int mirorr_ctx_iteration_graphs(const mirorr_ctx* ctx);

/* Per-kernel timing (no counterpart in the reference; the OpenCL host timed
 * its kernel with CL_QUEUE_PROFILING_ENABLE, itkBlockMatcherGPU.txx:380-391).
 * `on` selects the instrumented kernels (0 switches profiling off):
 * MIRORR_PROF_FAST ("ncc_tile_fast") and MIRORR_PROF_MASKED ("ncc_tile_masked");
 * with MIRORR_PROF_FINEST_ONLY a registration instruments only the launches
 * of its last (full-resolution) level. Each selected launch is bracketed by
 * HIP events on the context stream and reports its work in voxel-evals;
 * mirorr_ctx_profile_get sums launches, milliseconds and work. The events
 * add a few microseconds per launch. */
#define MIRORR_PROF_FAST 1
#define MIRORR_PROF_MASKED 2
#define MIRORR_PROF_FINEST_ONLY 4
mirorr_status mirorr_ctx_set_profiling(mirorr_ctx* ctx, int on);
mirorr_status mirorr_ctx_profile_get(mirorr_ctx* ctx, const char* kernel, int64_t* launches,
                                     double* total_ms, double* work);
mirorr_status mirorr_ctx_profile_reset(mirorr_ctx* ctx);

/* ---- block matcher (itk::AbstractBlockMatcher, itkAbstractBlockMatcher.h:41-313) */

void mirorr_bm_default_params(mirorr_bm_params* p);
mirorr_status mirorr_bm_create(mirorr_ctx* ctx, const mirorr_bm_params* p,
                               mirorr_bm** out);
void mirorr_bm_destroy(mirorr_bm* bm);
/* Set{BlockWidth,BlockGap,NhoodWidth,NhoodGap,PortionMatchesKept,Verbosity}
 * (itkAbstractBlockMatcher.h:101-127). */
mirorr_status mirorr_bm_set_params(mirorr_bm* bm, const mirorr_bm_params* p);

/* SetBaseImage/SetSearchImage + Set{Base,Search}Mask
 * (itkAbstractBlockMatcher.h:85-95); uploads host arrays (the reference's
 * updateBuffers, itkBlockMatcherGPU.txx:725-830). Base and search must share
 * one grid (the reference bounds the search by the base size,
 * itkBlockMatcher.txx:736-740). */
mirorr_status mirorr_bm_set_volume(mirorr_bm* bm, int role, const float* data,
                                   const uint8_t* mask, const mirorr_geom* geom);
/* Same with device pointers that stay owned by the caller and must outlive
 * the next match; nothing is copied. */
mirorr_status mirorr_bm_set_volume_device(mirorr_bm* bm, int role,
                                          const float* d_data,
                                          const uint8_t* d_mask,
                                          const mirorr_geom* geom);

/* SwapImageOrder (itkAbstractBlockMatcher.txx:101-112): base <-> search
 * without re-upload (itkBlockMatcherGPU.txx:186-202). */
mirorr_status mirorr_bm_swap(mirorr_bm* bm);

/* GetOffsets (itkAbstractBlockMatcher.h:139-146, itkBlockMatcherGPU.txx:204-522):
 * runs the exhaustive NCC search and the score filter and returns the kept
 * block centres (base) and matched centres (search) in mm, n x 3 doubles, in
 * block raster order. best/second/block_index (block raster id) are optional
 * per kept block. Returns MIRORR_ERR_ALGO when no block is kept, like the
 * reference GPU matcher. cap is the capacity in points. */
mirorr_status mirorr_bm_get_offsets(mirorr_bm* bm, double* base_mm,
                                    double* match_mm, size_t cap, size_t* n,
                                    float* best, float* second,
                                    int32_t* block_index);

/* Number of blocks of the current grid (itkBlockMatcher.txx:531-600). */
mirorr_status mirorr_bm_block_count(mirorr_bm* bm, size_t* nblocks,
                                    int grid[3]);

/* Unfiltered per-block results of the last search, block raster order:
 * block_idx / match_idx are n x 3 voxel indices of the block corners, best /
 * second the float score buffer (itkBlockMatcher.txx:650-656). Any pointer
 * may be NULL. Runs the search if the inputs changed since the last one. */
mirorr_status mirorr_bm_get_raw(mirorr_bm* bm, int32_t* block_idx,
                                int32_t* match_idx, float* best, float* second,
                                size_t cap, size_t* nblocks);

/* Enqueues the search on the context stream without waiting (for timing). */
mirorr_status mirorr_bm_match_async(mirorr_bm* bm);

/* Voxel-evals of one search over the current grid: sum over blocks of the
 * valid candidates times BW^3 (the unit of the throughput metric). */
mirorr_status mirorr_bm_voxel_evals(mirorr_bm* bm, int64_t* evals);

/* Tile shapes of the NCC search (BW 4, BG 1). Each neighbourhood width has a
 * few shapes (blocks per workgroup, candidate rows per lane); every shape
 * gives the same records. The first search of a (width, gap, storage,
 * grid) in the process times each shape that fits and keeps the fastest
 * (MIRORR_TILE_AUTOTUNE=0 in the environment keeps the first shape).
 * No counterpart in the reference, whose work-group shape is fixed
 * (itkBlockMatcherGPU.txx:330-338).
 * mirorr_tile_force pins width nhood_width to its index-th shape (0 = the
 * default; a negative index returns it to the autotuner).
 * mirorr_tile_autotune_report writes the choices so far (one line per key,
 * with each trial's time) and returns the report's length. */
mirorr_status mirorr_tile_force(int nhood_width, int index);
int mirorr_tile_autotune_report(char* buf, size_t cap);

/* ---- device-resident registration ------------------------------------ */

typedef struct mirorr_volume mirorr_volume;

/* Registration parameters: the mirorr CLI options of the hot path
 * (mirorr.cxx:99-281; MirorrPyramidImplement / MirorrRegistrationMethod
 * setters, itkMirorrPyramidImplement.h:104-139, itkMirorrRegistrationMethod.h:135-276). */
typedef struct mirorr_reg_params {
  int symmetric;      /* --reg-mode symmetric (1, default) | classic (0)        */
  int affine;         /* -t affine (1) | rigid Euler3D (0, default)             */
  int block_width;    /* --blockwidth (4)                                        */
  int block_gap;      /* --blockgap (1)                                          */
  int nhood_width;    /* --nhoodwidth (7)                                        */
  int nhood_gap;      /* --nhoodgap (3)                                          */
  int iterations;     /* -n: 0 keeps 5 per level, negative halves per level      */
  int pyr_start;      /* -a (1)                                                  */
  int pyr_end;        /* -c: 0 runs every level                                  */
  int pyr_num;        /* -d (1024)                                               */
  int pyr_min_size;   /* -e (32)                                                 */
  double portion_kept;/* --portion-kept (0.5); used by variance selection only   */
  int variance_select;/* 0: reference behaviour                                  */
  int verbosity;
  int centre_init;    /* centred initial transform (1) unless --no-centering-init */
  /* Early-stop thresholds, 1e-3 in the reference: the outer loop stops when
   * the squared parameter change falls below stop_tolerance
   * (itkSymMirorrRegistrationMethod.txx:456-692); the LTS inner loop runs
   * while the change and its change exceed lts_tolerance
   * (itkMirorrRegistrationMethod.txx:785-786). 0 disables early stopping. */
  double stop_tolerance;
  double lts_tolerance;
  int block_metric;   /* --blockmetric: MIRORR_METRIC_NC (default), _SSD, _CR    */
  int resampling;     /* --resampling-mode: MIRORR_RESAMPLE_MIDDLE (default),    */
                      /* _BASIC, _MOVING or _FIXED                               */
} mirorr_reg_params;

/* Resampling modes (MirorrRegistrationMethod's EResampling*): the grid each
 * level's images are resampled onto. Symmetric mode: middle and basic use
 * the half-way grid, moving and fixed that image's grid, all centred.
 * Classic mode: middle the half-way grid about the moving centre, moving and
 * basic the moving grid, fixed the fixed grid mapped back by the level's
 * start transform (itkMirorrRegistrationMethod.txx:450-531). */
#define MIRORR_RESAMPLE_MIDDLE 0
#define MIRORR_RESAMPLE_BASIC 1
#define MIRORR_RESAMPLE_MOVING 2
#define MIRORR_RESAMPLE_FIXED 3

typedef struct mirorr_reg_stats {
  int levels;
  int iterations[16];      /* outer iterations run per level                 */
  int64_t points[16];      /* matched pairs in the last iteration per level  */
  int inner_iterations;    /* LTS inner iterations, all levels               */
  int64_t voxel_evals;     /* NCC voxel-evals, all levels and passes         */
  int level_dims[16][3];   /* half-way grid per level                        */
  double seconds;          /* wall time of mirorr_register                   */
  int64_t rescored_blocks; /* NCC blocks re-searched in double because their
                              float best or best-second gap lay next to the
                              score filter's thresholds, all levels and passes */
} mirorr_reg_stats;

void mirorr_reg_default_params(mirorr_reg_params* p);

/* An image (and optional uint8 mask) uploaded once and kept in HBM; the
 * reference's NIfTI images after ReadAndResampleImages
 * (itkMirorrPyramidWrapper.txx:175-310). */
mirorr_status mirorr_volume_create(mirorr_ctx* ctx, const float* data, const uint8_t* mask,
                                   const mirorr_geom* geom, mirorr_volume** out);
void mirorr_volume_destroy(mirorr_volume* v);

/* Volume storage: float (the reference's float image3d uploads,
 * itkBlockMatcherGPU.txx:808-817), or binary16 holding (v - offset) * scale
 * with offset the nearest integer to the middle of the finite range and
 * scale a power of two (integer intensities of up to 12 bits are exact).
 * A registration of two binary16 volumes keeps its pyramid levels and
 * half-way images in binary16 and widens them to float where the search
 * stages them (NCC^2 is unchanged by the affine encoding; only the
 * rounding of interpolated values to binary16 changes the result). It
 * needs the NCC metric and blocks of width 4, block gap 1, nhood width 5
 * to 13. No counterpart in the reference: the image precision is its
 * float throughout. */
#define MIRORR_STORAGE_F32 0
#define MIRORR_STORAGE_F16 1
mirorr_status mirorr_volume_create_storage(mirorr_ctx* ctx, const float* data,
                                           const uint8_t* mask, const mirorr_geom* geom,
                                           int storage, mirorr_volume** out);
/* The binary16 encoding mirorr_volume_create_storage applies (host only):
 * out[i] holds the IEEE binary16 bits of (data[i] - offset) * scale,
 * computed in float and rounded to nearest even. */
mirorr_status mirorr_half_encode(const float* data, size_t n, uint16_t* out, double* offset,
                                 double* scale);
/* A volume's storage and, for binary16, its encoding. */
mirorr_status mirorr_volume_info(const mirorr_volume* v, int* storage, double* offset,
                                 double* scale);

/* MirorrPyramidImplement::RunMirorr (itkMirorrPyramidImplement.txx:71-158):
 * pyramid build, level loop, resample / block-match / LTS iterations.
 * init_params/init_center give the start transform (NULL: the centred
 * initialisation of itkMirorrPyramidWrapper.txx:611-623). out_params holds
 * 6 (rigid) or 12 (affine) ITK parameters; out_center their centre. */
mirorr_status mirorr_register(mirorr_ctx* ctx, const mirorr_reg_params* p,
                              const mirorr_volume* fixed, const mirorr_volume* moving,
                              const double* init_params, const double* init_center,
                              double* out_params, double* out_center, mirorr_reg_stats* stats);

/* The registrations of n pairs (n <= 256) with one set of parameters, run
 * in lockstep. No counterpart in the reference, which registers one pair
 * per process (mirorr.cxx:649). Each outer iteration queues every pair's
 * resampling, window statistics and block preparation, then per matching
 * pass ONE launch of each search kernel over a flat work-list of all the
 * pairs' tiles (pairs of any sizes), and one launch per LTS phase for all
 * pairs, with one host read of all the pairs' flags per chunk of inner
 * iterations. Pairs whose schedules have fewer levels join at their first
 * level (the finest levels align); a pair whose stop rule fires leaves its
 * level while the others go on. Pair i's result and statistics equal those
 * of mirorr_register on that pair alone. Needs the NCC metric with the
 * tiled search (blocks of width 4, gap 1, nhood width 5 to 13) and volumes
 * of one storage; per-kernel timing must be off. Arrays: fixed[n],
 * moving[n]; init_params n x 12 and init_center n x 3 (or NULL: the
 * centred initialisation of each pair); out_params n x 12 (the first 6 of
 * each row for rigid transforms); out_center n x 3 or NULL; stats[n] or NULL. */
mirorr_status mirorr_register_batch(mirorr_ctx* ctx, const mirorr_reg_params* p, int n,
                                    const mirorr_volume* const* fixed,
                                    const mirorr_volume* const* moving, const double* init_params,
                                    const double* init_center, double* out_params,
                                    double* out_center, mirorr_reg_stats* stats);

/* ---- serving loop ---------------------------------------------------- */

/* A registration server on one GPU. It has no counterpart in the reference,
 * whose command line runs one pair per process (mirorr.cxx:309-619); it is
 * the request loop a service puts in front of mirorr_register. Requests are
 * admitted onto `slots` workers, each with its own context and HIP stream,
 * whenever a worker is free and the request's device memory bound (volumes,
 * pyramids and workspace) fits in what the running requests and the workers'
 * kept workspaces leave of `hbm_budget`; workspaces of idle workers are
 * released when a request would not fit otherwise. Admission follows arrival
 * order, except that later requests that fit may overtake one waiting for
 * memory, at most 8 times. The work of admitted requests overlaps on the
 * device. A request that cannot fit even alone completes with
 * MIRORR_ERR_NOMEM. Every worker's context charges the
 * server's budget, so the limit also holds if a bound were exceeded.
 * hbm_budget = 0: the device's free memory at creation minus a reserve of
 * min(8 GiB, 10%). slots <= 0: 4; at most 64. */
typedef struct mirorr_server mirorr_server;

/* One registration: mirorr_volume_create of both images and mirorr_register
 * with these parameters. Masks and init_params/init_center may be NULL (as
 * for mirorr_register). borrow = 0: mirorr_submit copies the host arrays;
 * borrow = 1: the caller keeps the volume and mask arrays unchanged until
 * mirorr_poll has returned the request's result, and the worker uploads
 * from them directly. */
typedef struct mirorr_request {
  mirorr_reg_params params;
  const float* fixed;
  const uint8_t* fixed_mask;
  mirorr_geom fixed_geom;
  const float* moving;
  const uint8_t* moving_mask;
  mirorr_geom moving_geom;
  const double* init_params;
  const double* init_center;
  int borrow;
  int storage;            /* MIRORR_STORAGE_F32 (0) or _F16 for both volumes  */
} mirorr_request;

typedef struct mirorr_request_result {
  mirorr_status status;   /* mirorr_register's status                       */
  double params[12];      /* as mirorr_register's out_params / out_center    */
  double center[3];
  mirorr_reg_stats stats;
  double t_submit;        /* seconds since the server's creation             */
  double t_start;         /* admitted onto a worker                          */
  double t_end;           /* result ready                                    */
  int slot;               /* worker that ran it (-1: never admitted)         */
  size_t bytes_bound;     /* the device memory bound used for admission      */
  char error[256];        /* message of a failed request                     */
} mirorr_request_result;

typedef struct mirorr_server_info {
  int slots;
  int pending;            /* submitted, not admitted                         */
  int running;
  int64_t completed;      /* results produced (collected or not)             */
  size_t hbm_budget;
  size_t hbm_used;        /* bytes held by the workers' contexts now         */
  size_t hbm_peak;
  int max_running;        /* most requests running at once so far            */
  size_t reserve_peak;    /* largest sum, at an admission, of what the slots
                             were allowed to hold (running: kept workspace
                             plus the request's bound; idle: kept workspace);
                             admission keeps it within hbm_budget            */
  int batch_max;          /* requests a slot may take at once               */
  int max_batch;          /* largest batch run so far                        */
} mirorr_server_info;

mirorr_status mirorr_server_create(int device, size_t hbm_budget, int slots,
                                   mirorr_server** out);
/* Batching: a slot that takes a request also takes up to max_pairs - 1
 * later pending requests with the same parameters, storage and kind of
 * start that fit the HBM budget, and runs them together through
 * mirorr_register_batch (requests that cannot be batched, e.g. another
 * block metric, run alone). Each result equals the request's alone; a
 * batch's requests start and end together. 1 (the default) runs every
 * request alone. No counterpart in the reference (mirorr.cxx:649). */
mirorr_status mirorr_server_set_batch(mirorr_server* server, int max_pairs);
/* Queues a request; *ticket identifies it to mirorr_poll. */
mirorr_status mirorr_submit(mirorr_server* s, const mirorr_request* req, int64_t* ticket);
/* The result of a ticket. timeout_s = 0: returns MIRORR_PENDING at once if
 * it has not finished; < 0: waits for it; > 0: waits at most that long.
 * MIRORR_OK means *out holds the result (its own status in out->status); a
 * result is handed out once, after which the ticket is unknown
 * (MIRORR_ERR_ARG). */
mirorr_status mirorr_poll(mirorr_server* s, int64_t ticket, double timeout_s,
                          mirorr_request_result* out);
mirorr_status mirorr_server_info_get(mirorr_server* s, mirorr_server_info* out);
/* Message of the last failed mirorr_server_create / mirorr_submit /
 * mirorr_poll call on this host thread. */
const char* mirorr_server_last_error(void);
/* Completes the requests not yet admitted with MIRORR_ERR_STATE ("server
 * shut down"), waits for the running ones and for threads inside
 * mirorr_poll to return, then frees the workers. No call on the server may
 * start once this has been called. */
void mirorr_server_destroy(mirorr_server* s);

/* OptimizeTransform on given pairs (x_i block centres, y_i matches, n x 3 mm):
 * classic (itkMirorrRegistrationMethod.txx:758-1003) composes the fit into
 * params/center (in-out); symmetric (itkSymMirorrRegistrationMethod.txx:729-1442)
 * starts from identity half-space transforms and returns them as 3x4
 * [M | offset] in fwd34/bwd34 with params = t12F o t12B^-1. */
mirorr_status mirorr_lts_fit(mirorr_ctx* ctx, int symmetric, int affine, const double* x,
                             const double* y, size_t n, double* params, double* center,
                             double* fwd34, double* bwd34, int* iterations);

/* mirorr_lts_fit without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, semantics and results as mirorr_lts_fit; every device
 * buffer lives in one file-scope device array that each GPU holds until the
 * process ends (about 66 MB), sized for at most 1048576 pairs: larger n is
 * MIRORR_ERR_ARG. All calls on one GPU share that GPU's array, so calls that
 * overlap on one GPU (from several contexts or threads) race and must be
 * serialised by the caller. */
mirorr_status mirorr_lts_fit_noalloc(mirorr_ctx* ctx, int symmetric, int affine, const double* x,
                                     const double* y, size_t n, double* params, double* center,
                                     double* fwd34, double* bwd34, int* iterations);

/* ResampleImageFilter with LinearInterpolateImageFunction as the registration
 * uses it: out(p) = in(matrix * p + offset), 0 outside; out_mask (nullable)
 * is 255 where the point falls inside the input grid. Host arrays. */
mirorr_status mirorr_resample(mirorr_ctx* ctx, const float* data, const mirorr_geom* gin,
                              const double matrix[9], const double offset[3],
                              const mirorr_geom* gout, float* out, uint8_t* out_mask);

/* Interpolators of the final output resampling (--final-interpolator). */
#define MIRORR_INTERP_NN 0      /* "nn": NearestNeighborInterpolateImageFunction        */
#define MIRORR_INTERP_LINEAR 1  /* "linear": LinearInterpolateImageFunction             */
#define MIRORR_INTERP_BSPLINE 2 /* "bspline" (default): cubic BSplineInterpolateImage-  */
                                /* Function, mirror boundaries                          */
#define MIRORR_INTERP_SINC 3    /* "sinc": WindowedSincInterpolateImageFunction, radius */
                                /* 5, Welch window, zero-flux Neumann edges             */

/* Final output resampling of the command line's --save-resampled-fixed/-moving,
 * --save-fixed and --save-moving (itkMirorrPyramidImplement.txx:364-428
 * GetInterpolatorFromString, 436-470 GetResampledImage): ResampleImageFilter
 * with default value 0, out(p) = in(matrix * p + offset) for the physical
 * points p of gout, evaluated only inside [-0.5, n-0.5) of the input grid.
 * Host arrays; dims of gout at most 65535 per axis. */
mirorr_status mirorr_resample_final(mirorr_ctx* ctx, const float* data, const mirorr_geom* gin,
                                    const double matrix[9], const double offset[3],
                                    const mirorr_geom* gout, int interpolator, float* out);

/* Level `level` (0 = coarsest) of the recursive Gaussian pyramid of `data`
 * for a schedule of rows x 3 shrink factors (coarsest first). out holds the
 * level image; out_geom its geometry. */
mirorr_status mirorr_pyramid_level(mirorr_ctx* ctx, const float* data, const mirorr_geom* geom,
                                   const int* schedule, int rows, int level, float* out,
                                   mirorr_geom* out_geom);

#ifdef __cplusplus
}
#endif

#endif /* MIRORR_GPU_H */
