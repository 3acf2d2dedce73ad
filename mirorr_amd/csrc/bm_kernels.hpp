// Launchers of the block-matching kernels (implemented in bm_kernels.hip).
#pragma once

#include <string>
#include <vector>

#include "common.hpp"
#include "profiler.hpp"

namespace mirorr {

// Per-block preparation of a base image (centred block, sums, mask bits).
struct BlockPrep {
  float* xt = nullptr;        // nblocks x BW^3 centred voxels x - mean (float)
  float* sxx = nullptr;       // nblocks: sum (x - mean)^2 over all voxels
  float* sumxt = nullptr;     // nblocks: sum of the float centred voxels
  uint32_t* flags = nullptr;  // nblocks: bit0 base mask all on, bit1 all off,
                              // bit2 dropped by the variance selection
};

constexpr uint32_t kBlockSkip = 4u;  // flags bit: block not searched

// Block-variance selection (device state and scratch).
struct VarSelState {
  unsigned long long k, k2, k3;  // ascending rank of the cut-off key, then within bins
  unsigned int bin1, bin2, bin3;
  unsigned long long kept;       // blocks kept by the last selection
};
struct VarSelScratch {
  VarSelState* st = nullptr;
  unsigned int* hist = nullptr;        // 4096 + 4096 + 256 bins
  unsigned long long* evals = nullptr; // optional: += voxel-evals of the kept blocks
};
constexpr int kVarSelHistBins = 4096 + 4096 + 256;

// Per-position window statistics of a search image (window = BW^3 voxels
// starting at the position): x = mean over all voxels, y = 1 / (sum over
// search-mask-on voxels of (v - mean)^2), or -1 when that sum is zero (the
// score is then -2). Computed in double, stored float.
struct SearchStats {
  float2* ms = nullptr;  // dims[0]*dims[1]*dims[2]
};

// Raw per-block results in block raster order.
struct BMResult {
  float* best = nullptr;
  float* second = nullptr;
  int32_t* match = nullptr;  // nblocks x 3 voxel index of the matched corner
};

// Working memory of one search (tile queue for the masked path).
//
// Exact re-scoring (rlist set): the tiled kernels accumulate in float, so
// their NCC^2 values differ from the reference's double sums
// (itkBlockMatcher.txx:46-101) by up to a few 1e-5. The score filter keeps a
// block when best > 1e-7 and |best - second| > 1e-4 (itkBlockMatcher.txx:
// 886-916); the tiled kernels list every block whose float best or gap lies
// within 1e-4 of a threshold (and, with rescore_ties, every exact tie),
// and those blocks are searched again in double with the reference's
// operation order and float score buffer, so the filter decides every block
// as the reference does.
struct BMScratch {
  int32_t* queue = nullptr;  // ntiles entries
  int32_t* count = nullptr;  // 1 counter
  unsigned long long* qbits = nullptr;  // ntiles entries: blocks per queued tile
  int64_t ntiles_cap = 0;
  Profiler* prof = nullptr;  // optional: event timing + voxel-evals per launch
  int32_t* rlist = nullptr;  // nblocks entries: blocks to re-score (NULL: none)
  int32_t* rcount = nullptr; // 1 counter
  int rescore_ties = 0;
  // 1: the listed blocks get the reference's record (best, second, match);
  // 0: only the filter's decision (and a kept block's match) is the
  // reference's, which is all a registration uses.
  int exact_records = 1;
  unsigned long long* rescored = nullptr;  // optional: += blocks re-scored
};

// One pair's search in a batched launch: the arguments of launch_ncc_search.
struct SearchJob {
  const void* x = nullptr;
  const uint8_t* bmask = nullptr;
  const void* y = nullptr;
  const uint8_t* smask = nullptr;
  BMGrid g{};
  BlockPrep bp;
  SearchStats st;
  BMScratch scr;
  BMResult res;
};

// The NCC searches of several pairs (one block and neighbourhood geometry,
// volumes of one storage, any image sizes) in one launch of each tiled
// kernel over a flat list of every pair's tiles, then one re-scoring launch
// over every pair's list. Each pair's records equal those of
// launch_ncc_search on that pair. `dev` is device memory of at least
// search_batch_bytes(n) bytes for the per-pair arguments (rewritten by
// every call; calls on one buffer are ordered by the stream).
size_t search_batch_bytes(int njobs);
void launch_ncc_search_batch(const std::vector<SearchJob>& jobs, void* dev, size_t dev_bytes,
                             hipStream_t s, int storage = kStoreF32);

// Image pointers of the search and block preparation hold elements of
// `storage` (kStoreF32 float, kStoreF16 half_t; half needs 4-voxel blocks).
void launch_search_stats(const void* y, const uint8_t* smask, const BMGrid& g,
                         SearchStats st, hipStream_t s, int storage = kStoreF32);
void launch_block_prep(const void* x, const uint8_t* bmask, const BMGrid& g,
                       BlockPrep bp, hipStream_t s, int storage = kStoreF32);
// Keeps the top `portion` of the blocks by centred sum of squares (bp.sxx):
// K = ceil(portion * nblocks) clamped to [1, nblocks], the cut-off is the
// K-th largest value and every block at or above it is kept (ties kept);
// the others get kBlockSkip in bp.flags. The reference has this culling
// compiled out (itkBlockMatcher.txx:664-719, keeping the top 1 - portion
// instead). Returns false (and touches nothing) when every block is kept.
bool launch_variance_select(const BMGrid& g, const BlockPrep& bp, double portion,
                            VarSelScratch& scr, hipStream_t s);

// Exhaustive NCC search. Chooses the tiled fast kernels for BW=4, BG=1 and a
// supported NW, else the generic kernel (float images only).
void launch_ncc_search(const void* x, const uint8_t* bmask, const void* y,
                       const uint8_t* smask, const BMGrid& g, const BlockPrep& bp,
                       const SearchStats& st, BMScratch& scratch, BMResult res,
                       hipStream_t s, int storage = kStoreF32);
// Generic kernel only (any BW/BG/NW), double accumulation; used for
// unsupported shapes and as an independent device cross-check.
void launch_ncc_generic(const float* x, const uint8_t* bmask, const float* y,
                        const uint8_t* smask, const BMGrid& g, BMResult res,
                        hipStream_t s, const uint32_t* flags = nullptr);
// Exhaustive search under another block metric (MIRORR_METRIC_SSD or
// MIRORR_METRIC_CR; MIRORR_METRIC_NC runs launch_ncc_generic), any BW up to
// kMaxMetricBW, BG and NW.
constexpr int kMaxMetricBW = 16;
void launch_metric_search(int metric, const float* x, const uint8_t* bmask, const float* y,
                          const uint8_t* smask, const BMGrid& g, BMResult res, hipStream_t s,
                          const uint32_t* flags = nullptr);
// Tiles of the largest tile shape count for this grid (queue sizing), and
// whether the tiled kernels serve it.
int64_t tile_count(const BMGrid& g);
// The tile shape a search over `g` with volumes of `storage` runs with now,
// as an opaque code that changes whenever that choice does: a pinned shape,
// the tuned one, or -1 when the next search would tune.
int tile_choice(const BMGrid& g, int storage);
bool tiled_supported(const BMGrid& g);
// The autotuner's choices so far, one line per (width, gap, storage, grid):
// the chosen shape and each trial's time.
std::string tile_autotune_report();
// Pins the tile shape of width nw to its index-th variant (0 = the default),
// or returns it to the autotuner (index < 0). False for an unknown index.
bool tile_force(int nw, int index);

// Score filter + mm conversion (itkBlockMatcher.txx:886-916,
// itkAbstractBlockMatcher.h:271-287). kept[i] in {0,1}; pts are nblocks x 3.
void launch_postprocess(const BMGrid& g, const Geom& geom, const BMResult& res,
                        uint8_t* kept, double* base_mm, double* match_mm,
                        hipStream_t s);

// Score filter + mm conversion written as LTS pairs: the base point goes to
// X (Y when swap), the matched point to Y (X when swap); count += kept.
void launch_postprocess_pairs(const BMGrid& g, const Geom& geom, const BMResult& res, int swap,
                              uint8_t* kept, double* X, double* Y, unsigned long long* count,
                              hipStream_t s);

// Ordered compaction of nblocks records by a 0/1 flag array.
struct Compactor {
  int32_t* chunk_counts = nullptr;  // nchunks
  int32_t* chunk_offsets = nullptr; // nchunks + 1
  int64_t cap_chunks = 0;
};
int64_t compact_chunks(int64_t n);
void launch_compact_index(const uint8_t* flags, int64_t n, Compactor& c,
                          int32_t* out_index, hipStream_t s);

}  // namespace mirorr
