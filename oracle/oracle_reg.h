// Oracle registration API (test infrastructure only; see oracle.h).
#pragma once
#include <cstdint>

#include "oracle.h"

#ifdef __cplusplus
#include <vector>

#include "xform.hpp"
namespace orc {
extern int g_verbosity;
// tol: the inner loop's thresholds on the parameter change (1e-3 in the
// reference, itkMirorrRegistrationMethod.txx:785-786).
int lts_classic(Xf& T, const std::vector<double>& X, const std::vector<double>& Y,
                double tol = 1e-3);
int lts_symmetric(Xf& T, Xf& t12F, Xf& t12B, const std::vector<double>& X,
                  const std::vector<double>& Y, double tol = 1e-3);
}  // namespace orc
extern "C" {
#endif

typedef struct orc_reg_params {
  int symmetric;     // --reg-mode symmetric (1, default) or classic (0)
  int affine;        // -t affine (1) or rigid Euler3D (0)
  int bw, bg, nw, ng;
  int iterations;    // -n: 0 keeps 5 per level, negative halves per level
  int pyr_start, pyr_end, pyr_num, pyr_min_size;  // -a -c -d -e
  double portion_kept;
  int nthreads;      // block-matcher threads; 0 = single-threaded matcher
  int verbosity;
  int centre_init;   // CenteredTransformInitializer GeometryOn (default 1)
  int variance_select;  // block-variance selection at levels >= 2 (default 0)
  // Early-stop thresholds, 1e-3 in the reference: outer loop on the squared
  // parameter change (itkSymMirorrRegistrationMethod.txx:456-692), inner LTS
  // loop on the change and its change (itkMirorrRegistrationMethod.txx:785-786).
  // 0 disables early stopping (fixed iteration counts), for comparisons.
  double stop_tolerance;
  double lts_tolerance;
  int metric;        // --blockmetric: 0 nc (default), 1 sd, 2 cr
  int resampling;    // --resampling-mode: 0 middle (default), 1 basic, 2 moving, 3 fixed
} orc_reg_params;

typedef struct orc_reg_stats {
  int levels;
  int iterations[16];        // outer iterations run per level
  int64_t points[16];        // matched pairs in the last iteration per level
  int inner_iterations;      // LTS inner iterations, all levels
  int64_t voxel_evals;       // NCC voxel-evals, all levels and passes
  int level_dims[16][3];     // half-space grid per level
} orc_reg_stats;

void orc_reg_default_params(orc_reg_params* p);
int orc_register(const float* fixed, const uint8_t* fmask, const orc_geom* fg,
                 const float* moving, const uint8_t* mmask, const orc_geom* mg,
                 const orc_reg_params* p, const double* init_params,
                 const double* init_center, double* out_params, double* out_center,
                 orc_reg_stats* stats);
// The level loop with exact correspondences from a truth transform (moving
// -> fixed physical points) in place of the block search.
int orc_register_synthetic(const orc_geom* fg, const orc_geom* mg, const orc_reg_params* p,
                           const double truth_M[9], const double truth_t[3],
                           const double* init_params, const double* init_center,
                           double* out_params, double* out_center, orc_reg_stats* stats);

#ifdef __cplusplus
}
#endif
