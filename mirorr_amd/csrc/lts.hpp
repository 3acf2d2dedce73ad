// Least-trimmed-squares regression on the device (implemented in lts.hip).
#pragma once

#include "common.hpp"
#include "imaging.hpp"
#include "xform.hpp"

namespace mirorr {

// Device state of one LTS run (one OptimizeTransform call).
struct LtsState {
  xf::Xf F, B;       // current fits: final_transform(F) and final_transformB
  double rc[3];      // rotation centre (classic)
  double last[12];   // last_parameters
  double fraction, delta, last_delta;
  double tol;        // inner-loop threshold (1e-3 in the reference)
  int iter, active;
  int finished;      // the transform update of this run is done (finish_kernel)
  int kind, symmetric;
  int status;        // bit0: a rigid square root needed re-orthogonalising
  unsigned long long n_kept;  // kept pairs (points with a finite residual)
  unsigned long long k;       // order statistic of the trimmed threshold
  unsigned long long k2, k3;  // remaining ranks after the 1st and 2nd digit
  unsigned int bin1, bin2;
  unsigned int ncand;
  double thr;
};

// Device state of the outer iteration loop of one level.
struct OuterState {
  xf::Xf T;              // m_Transform (centred frame in symmetric mode)
  xf::Xf t12F, t12B;     // half-space transforms (symmetric)
  xf::Xf fixT, movT;     // translations by the image centres
  double lastp[12];      // m_LastTransformParameters
  double sqdiff;         // squared parameter change of the last iteration
  int stop;              // early breakout (sqdiff < stop_tol)
  double stop_tol;       // outer-loop threshold (1e-3 in the reference)
  int inner_total;       // LTS inner iterations so far
  // resampling index maps for the next iteration
  IndexMap mapF, mapM, mmapF, mmapM;
  Geom gF, gM;           // level grids of the fixed and moving images
  Geom gFfull, gMfull;   // full-resolution grids (masks)
  Geom gHF, gHM;         // half-space grids (fixed / moving direction)
};

struct LtsBuffers {
  double* r = nullptr;                // npts residuals
  unsigned int* hist1 = nullptr;      // 4096, zero between iterations
  unsigned int* hist2 = nullptr;      // 4096, zero between iterations
  unsigned long long* cand = nullptr; // npts candidate keys
  double* part = nullptr;             // nparts x 32 moment partials
  int64_t nparts_cap = 0;
};

int64_t lts_parts(int64_t npts);

// One LTS run over pairs (X[i], Y[i]) with kept[i] != 0; st->n_kept must
// hold the number of kept pairs. lts_start resets the state from the outer
// transform; the host then queues iterations (kernels are predicated on
// st->active, so a chunk can be queued before reading the flag back).
void lts_start(LtsState* st, const OuterState* outer, int kind, int symmetric, double tol,
               LtsBuffers& b, hipStream_t s);
void lts_enqueue_iteration(const double* X, const double* Y, const uint8_t* kept, int64_t npts,
                           LtsState* st, OuterState* outer, LtsBuffers& buf, int iteration,
                           hipStream_t s);
// Composes the fits into the outer transforms (symmetric: t12F/t12B and
// T = t12F o t12B^-1; classic: T o F), updates the stop flag and the
// resampling maps of the next iteration.
// Queued after each chunk of inner iterations: once the LTS is inactive it
// composes the fit into the outer state (exactly once per run). Every launch
// then writes flags[0] = active, flags[1] = the outer stop flag and
// flags[2] = the inner iterations run so far (flags holds 3 ints).
void lts_finish(LtsState* st, OuterState* outer, int* flags, hipStream_t s);
// Computes the resampling maps of the current outer state (level start).
void outer_prepare_maps(OuterState* outer, int symmetric, hipStream_t s);
// Zeroes the kept-pair counter (before the block-matching passes).
void lts_reset_count(LtsState* st, hipStream_t s);

// The LTS of several pairs at once: one launch per phase for all pairs,
// each pair with its own state, pairs and buffers (an array of LtsItem in
// device memory). Every pair's results are those of the one-pair calls.
struct LtsItem {
  const double* X;
  const double* Y;
  const uint8_t* kept;
  int64_t npts;          // as the one-pair calls' npts (at least 1)
  LtsState* st;
  OuterState* outer;
  double* r;
  unsigned int* hist1;
  unsigned int* hist2;
  unsigned long long* cand;
  double* part;          // lts_parts(npts) partials
  int* flags;            // {active, stop, iterations}, written by lts_finish_multi
};
void lts_start_multi(const LtsItem* items, int n, int kind, int symmetric, double tol,
                     hipStream_t s);
// max_npts, min_npts: the largest and the smallest npts of the items.
void lts_enqueue_iteration_multi(const LtsItem* items, int n, int64_t max_npts,
                                 int64_t min_npts, int iteration, hipStream_t s);
void lts_finish_multi(const LtsItem* items, int n, hipStream_t s);
void lts_reset_count_multi(const LtsItem* items, int n, hipStream_t s);

}  // namespace mirorr
