// CPU oracle for the Mirorr block-matching registration hot path.
//
// TEST INFRASTRUCTURE ONLY. This library is a plain C++ restatement of the
// reference algorithms (aehrc/Mirorr, src/mirorr/*) used as the checker for
// the HIP implementation in mirorr_amd/. Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg may load it. The product never links it.
//
// Every function cites the reference file:line it restates. Arithmetic types
// follow the reference: float voxels, double accumulation, float score buffer.
#pragma once
#include <cstdint>
#include <cstddef>

#ifdef __cplusplus
extern "C" {
#endif

// Geometry of a 3-D volume: x fastest, then y, then z (ITK raster order).
typedef struct {
  int dims[3];
  double spacing[3];
  double origin[3];
  double direction[9];  // row-major 3x3 direction cosines
} orc_geom;

// ---- Block matcher (itkBlockMatcher.txx, itkBlockMatcherThreaded.txx) -----

// NCC^2 of one block pair, masked (itkBlockMatcher.txx:46-101).
double orc_ncc_pair(const float* base, const uint8_t* base_mask,
                    const float* search, const uint8_t* search_mask,
                    const int dims[3], const int base_idx[3],
                    const int search_idx[3], int bw);

// Unmasked variant (itkAbstractBlockMatcher.h:167-204).
double orc_ncc_pair_unmasked(const float* base, const float* search,
                             const int dims[3], const int base_idx[3],
                             const int search_idx[3], int bw);

// Block grid (itkBlockMatcher.txx:531-600): per-axis count and start.
void orc_block_grid(const int dims[3], int bw, int ng, int nblocks[3],
                    int start[3]);

// Full GetOffsets. Raw per-block outputs are written in block raster order
// (x fastest) whatever the thread count: block_idx/match_idx are n x 3 voxel
// indices, best/second the float score buffer. Kept points (mm) come out in
// the reference's thread-concatenation order (itkBlockMatcherThreaded.txx:
// 73-89). nthreads <= 0 means single-threaded BlockMatcher::GetOffsets.
// mask pointers may be NULL (reference default: all 255,
// itkBlockMatcher.txx:507-529). Returns 0 on success, 1 if the region has no
// block (the reference throws, itkBlockMatcher.txx:936-943).
int orc_bm_get_offsets(const float* base, const uint8_t* base_mask,
                       const float* search, const uint8_t* search_mask,
                       const orc_geom* geom, int bw, int bg, int nw, int ng,
                       int nthreads, int32_t* block_idx, int32_t* match_idx,
                       float* best, float* second, double* base_mm,
                       double* match_mm, int64_t* n_kept);

// GetOffsets restricted to the blocks with block_keep[b] != 0 (NULL: all).
// Skipped blocks get best = second = -3 and their own position as match.
// SSD and correlation ratio of one block pair (BlockMatcher's other metrics).
double orc_ssd_pair(const float* base, const uint8_t* base_mask, const float* search,
                    const uint8_t* search_mask, const int dims[3], const int base_idx[3],
                    const int search_idx[3], int bw);
double orc_cr_pair(const float* base, const uint8_t* base_mask, const float* search,
                   const uint8_t* search_mask, const int dims[3], const int base_idx[3],
                   const int search_idx[3], int bw);

// orc_bm_get_offsets_sel under block metric `metric` (0 NCC, 1 SSD, 2 CR).
int orc_bm_get_offsets_metric(const float* base, const uint8_t* base_mask,
                              const float* search, const uint8_t* search_mask,
                              const orc_geom* geom, int bw, int bg, int nw, int ng,
                              int nthreads, const uint8_t* block_keep, int metric,
                              int32_t* block_idx, int32_t* match_idx, float* best,
                              float* second, double* base_mm, double* match_mm,
                              int64_t* n_kept);

int orc_bm_get_offsets_sel(const float* base, const uint8_t* base_mask,
                           const float* search, const uint8_t* search_mask,
                           const orc_geom* geom, int bw, int bg, int nw, int ng,
                           int nthreads, const uint8_t* block_keep, int32_t* block_idx,
                           int32_t* match_idx, float* best, float* second, double* base_mm,
                           double* match_mm, int64_t* n_kept);

// Block-variance selection (a live version of the culling compiled out at
// itkBlockMatcher.txx:664-719, keeping the top `portion` of the blocks by
// centred sum of squares; ties with the cut-off are kept). Writes keep[b]
// and (nullable) keys[b] in block raster order; returns the number kept.
int64_t orc_variance_select(const float* base, const int dims[3], int bw, int ng,
                            double portion, uint8_t* keep, float* keys);

// Candidates of the blocks with keep[b] != 0 (NULL: all blocks).
int64_t orc_bm_candidate_count_sel(const int dims[3], int bw, int bg, int nw, int ng,
                                   const uint8_t* keep);

// Number of candidates evaluated for the whole grid (voxel-evals / bw^3).
int64_t orc_bm_candidate_count(const int dims[3], int bw, int bg, int nw,
                               int ng);

// ---- Resampling and pyramid (imaging.cpp) -------------------------------
void orc_resample(const float* in, const uint8_t* in_mask, const orc_geom* gin,
                  const double M[9], const double off[3], const orc_geom* gout,
                  float* out, uint8_t* out_mask);
int orc_gaussian_kernel(double variance, double max_error, int max_width, double* out,
                        int cap);
void orc_gaussian_smooth(const float* in, const int dims[3], const double var[3],
                         double max_error, float* out);
void orc_pyramid_level_geom(const orc_geom* in, const int factors[3], orc_geom* out);
void orc_pyramid_step(const float* prev, const orc_geom* prev_geom, const int ratio[3],
                      const orc_geom* out_geom, float* out);
int orc_schedule_symmetric(const int msize[3], const double mspacing[3], const int fsize[3],
                           const double fspacing[3], unsigned min_length, unsigned max_levels,
                           int* msched, int* fsched, int cap);
int orc_cropped_level(int level, int default_level, int total_levels);
void orc_cropped_min_max(int level_min, int level_max, int total, int* min_out, int* max_out);

#ifdef __cplusplus
}
#endif
