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
  MIRORR_ERR_STATE = 5
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
} mirorr_bm_params;

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

#ifdef __cplusplus
}
#endif

#endif /* MIRORR_GPU_H */
