// Oracle: multi-resolution registration driver (test infrastructure only).
//
// Restates the level loop of MirorrPyramidImplement::RunMirorr
// (itkMirorrPyramidImplement.txx:71-158), the pyramid construction
// (:575-609), the classic and symmetric StartOptimization
// (itkMirorrRegistrationMethod.txx:422-659, itkSymMirorrRegistrationMethod.txx:238-720)
// with middle-space resampling, ComputeHalfSpaceTransforms (:139-201) and the
// centred initial transform of itkMirorrPyramidWrapper.txx:611-623.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "oracle_reg.h"

using namespace orc;

namespace {

struct Img {
  std::vector<float> v;
  orc_geom g;
};

void to_phys(const orc_geom& g, const double idx[3], double p[3]) {
  for (int r = 0; r < 3; ++r) {
    double s = g.origin[r];
    for (int c = 0; c < 3; ++c) s += g.direction[3 * r + c] * g.spacing[c] * idx[c];
    p[r] = s;
  }
}

// ComputeImageCenter: physical point of continuous index (size - 1) / 2.
void image_centre(const orc_geom& g, double c[3]) {
  const double idx[3] = {(g.dims[0] - 1) / 2.0, (g.dims[1] - 1) / 2.0, (g.dims[2] - 1) / 2.0};
  to_phys(g, idx, c);
}

size_t nvox(const orc_geom& g) { return (size_t)g.dims[0] * g.dims[1] * g.dims[2]; }

// Recursive pyramid: outputs for every schedule row (coarsest first).
std::vector<Img> build_pyramid(const float* data, const orc_geom& g, const std::vector<int>& sched,
                               int rows) {
  std::vector<Img> out(rows);
  for (int l = rows - 1; l >= 0; --l) {
    const int f[3] = {sched[3 * l], sched[3 * l + 1], sched[3 * l + 2]};
    orc_pyramid_level_geom(&g, f, &out[l].g);
    out[l].v.resize(nvox(out[l].g));
    int ratio[3];
    const float* prev;
    const orc_geom* pg;
    if (l == rows - 1) {
      for (int a = 0; a < 3; ++a) ratio[a] = f[a];
      prev = data;
      pg = &g;
    } else {
      for (int a = 0; a < 3; ++a) ratio[a] = f[a] / sched[3 * (l + 1) + a];
      prev = out[l + 1].v.data();
      pg = &out[l + 1].g;
    }
    orc_pyramid_step(prev, pg, ratio, &out[l].g, out[l].v.data());
  }
  return out;
}

// Resample src (and its mask) under transform T (output -> input physical).
void resample(const float* src, const uint8_t* smask, const orc_geom& sg, const Xf& T,
              const orc_geom& og, std::vector<float>* out, std::vector<uint8_t>* omask,
              bool default_mask_src) {
  double M[9];
  for (int i = 0; i < 9; ++i) M[i] = T.M.a[i / 3][i % 3];
  if (out) {
    out->resize(nvox(og));
    orc_resample(src, nullptr, &sg, M, T.off, &og, out->data(), nullptr);
  }
  if (omask) {
    omask->resize(nvox(og));
    // A missing mask is the reference's default all-255 mask image.
    orc_resample(src, default_mask_src ? nullptr : smask, &sg, M, T.off, &og, nullptr,
                 omask->data());
  }
}

struct Matches {
  std::vector<double> x, y;
  int64_t evals = 0;
};

// Exact-correspondence mode (a check of the level loop's transform algebra,
// used only by orc_register_synthetic): with g_truth set, a block-matching
// pass returns, for every block of the base grid, its centre x (the
// blockPositionToPoint convention, itkAbstractBlockMatcher.h:271-287) and
// the exact point y = to_search(to_phys(x)) where its content lies in the
// search image, instead of searching the images. g_truth maps moving
// physical points to fixed ones, as m_Transform does.
thread_local const Xf* g_truth = nullptr;

void exact_pairs(const orc_geom& gb, const orc_reg_params& p, const Xf& to_phys,
                 const Xf& to_search, std::vector<double>& out_b, std::vector<double>& out_m) {
  int nb[3], st[3];
  orc_block_grid(gb.dims, p.bw, p.ng, nb, st);
  out_b.clear();
  out_m.clear();
  for (int z = 0; z < nb[2]; ++z)
    for (int y = 0; y < nb[1]; ++y)
      for (int x = 0; x < nb[0]; ++x) {
        const int idx[3] = {st[0] + p.ng * x, st[1] + p.ng * y, st[2] + p.ng * z};
        double c[3], q[3], r[3];
        for (int k = 0; k < 3; ++k) {
          double v = gb.origin[k];
          for (int j = 0; j < 3; ++j) v += gb.direction[3 * k + j] * gb.spacing[j] * idx[j];
          c[k] = v + gb.spacing[k] * (0.5 * p.bw);
        }
        to_phys.apply(c, q);
        to_search.apply(q, r);
        out_b.insert(out_b.end(), c, c + 3);
        out_m.insert(out_m.end(), r, r + 3);
      }
}

// Composition a o b (b applied first).
Xf after(const Xf& a, const Xf& b) {
  Xf r = a;
  r.compose(b, true);
  return r;
}

// One block-matching pass. With portion < 1 (variance selection on, at a
// level whose portion is not forced to 1.0) only the selected base blocks
// are searched.
void match(const std::vector<float>& base, const std::vector<uint8_t>& bmask,
           const std::vector<float>& search, const std::vector<uint8_t>& smask,
           const orc_geom& bg, const orc_reg_params& p, double portion,
           std::vector<double>& out_b, std::vector<double>& out_m, int64_t* evals) {
  int nb[3], st[3];
  orc_block_grid(bg.dims, p.bw, p.ng, nb, st);
  const int64_t n = (int64_t)nb[0] * nb[1] * nb[2];
  std::vector<double> b(3 * std::max<int64_t>(n, 1)), m(3 * std::max<int64_t>(n, 1));
  std::vector<uint8_t> keep;
  if (portion < 1.0 && n > 0) {
    keep.resize(n);
    orc_variance_select(base.data(), bg.dims, p.bw, p.ng, portion, keep.data(), nullptr);
  }
  const uint8_t* kp = keep.empty() ? nullptr : keep.data();
  int64_t k = 0;
  if (orc_bm_get_offsets_metric(base.data(), bmask.data(), search.data(), smask.data(), &bg,
                                p.bw, p.bg, p.nw, p.ng, p.nthreads, kp, p.metric, nullptr,
                                nullptr, nullptr, nullptr, b.data(), m.data(), &k) != 0) {
    std::fprintf(stderr, "oracle: no block matches occurred\n");
    k = 0;
  }
  *evals += orc_bm_candidate_count_sel(bg.dims, p.bw, p.bg, p.nw, p.ng, kp) * p.bw * p.bw * p.bw;
  out_b.assign(b.begin(), b.begin() + 3 * k);
  out_m.assign(m.begin(), m.begin() + 3 * k);
}

// Middle-space grid of the symmetric method (itkSymMirorrRegistrationMethod.txx:255-283).
orc_geom middle_grid(const orc_geom& F, const orc_geom& M, const double cF[3],
                     const double cM[3], bool add_moving_centre, const double* direction) {
  orc_geom h = F;
  for (int a = 0; a < 3; ++a) {
    const double dF = F.origin[a] - cF[a], dM = M.origin[a] - cM[a];
    h.origin[a] = (dF + dM) / 2.0 + (add_moving_centre ? cM[a] : 0.0);
    h.spacing[a] = (F.spacing[a] + M.spacing[a]) / 2.0;
    h.dims[a] = (int)(((double)(F.dims[a] + M.dims[a]) + 0.5) / 2.0);
  }
  for (int i = 0; i < 9; ++i) h.direction[i] = direction[i];
  return h;
}

// Level grid under a resampling mode (itkSymMirorrRegistrationMethod.txx:
// 264-301; itkMirorrRegistrationMethod.txx:450-531). Symmetric grids are in
// centred coordinates and basic equals middle there; classic fixed maps the
// fixed origin back through T, classic basic uses the moving grid and
// direction for both images.
orc_geom mode_grid(int mode, bool classic, const orc_geom& F, const orc_geom& M,
                   const double cF[3], const double cM[3], const Xf& T, const double* direction) {
  if (mode == 0 || (mode == 1 && !classic)) return middle_grid(F, M, cF, cM, classic, direction);
  const bool fixed = mode == 3;
  orc_geom h = fixed ? F : M;
  const double* c = fixed ? cF : cM;
  if (!classic) {
    for (int a = 0; a < 3; ++a) h.origin[a] -= c[a];
  } else if (fixed) {
    double o[3];
    T.inverse().apply(F.origin, o);
    for (int a = 0; a < 3; ++a) h.origin[a] = o[a];
  }
  const double* d = (classic && mode == 1) ? M.direction : direction;
  for (int i = 0; i < 9; ++i) h.direction[i] = d[i];
  return h;
}

Xf translation_xf(const double t[3]) {
  Xf x;
  x.kind = kEuler3D;
  x.set_identity();
  const double p[6] = {0, 0, 0, t[0], t[1], t[2]};
  x.set_params(p);
  return x;
}

// ComputeHalfSpaceTransforms: t12F = sqrt(T), t12B = sqrt(T^-1).
void half_space(const Xf& T, Xf& t12F, Xf& t12B) {
  M4 m{};
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) m.a[i][j] = T.M.a[i][j];
    m.a[i][3] = T.t[i];
  }
  m.a[3][3] = 1.0;
  const M4 m12 = m4_sqrt_db(m, 1e-14, 50);
  const M4 im12 = m4_sqrt_db(m4_inverse(m), 1e-14, 50);
  M3 K, L;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      K.a[i][j] = m12.a[i][j];
      L.a[i][j] = im12.a[i][j];
    }
  t12F.set_matrix(K);
  t12B.set_matrix(L);
  const double a[3] = {m12.a[0][3], m12.a[1][3], m12.a[2][3]};
  const double b[3] = {im12.a[0][3], im12.a[1][3], im12.a[2][3]};
  t12F.set_translation(a);
  t12B.set_translation(b);
}

struct LevelIn {
  const Img* F;
  const Img* M;
  const float* Ffull;  // full-resolution data (mask default uses its grid)
  const uint8_t* Fmask;
  const orc_geom* Fg;
  const float* Mfull;
  const uint8_t* Mmask;
  const orc_geom* Mg;
};

int params_sq_delta(const double* a, const double* b, int n, double* out) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += (a[i] - b[i]) * (a[i] - b[i]);
  *out = s;
  return 0;
}

// Portion of blocks kept by the variance selection at a level: levels 0 and
// 1 (the two coarsest) use every block (itkMirorrPyramidImplement.txx:116-119);
// 1.0 when the selection is off.
double level_portion(const orc_reg_params& p, int level) {
  if (!p.variance_select || level <= 1) return 1.0;
  return p.portion_kept == 0.0 ? 0.5 : p.portion_kept;
}

// Symmetric StartOptimization for one level; T is updated in place.
void start_opt_symmetric(Xf& T, const LevelIn& L, int max_iter, const orc_reg_params& p,
                         orc_reg_stats* st, int level) {
  const double portion = level_portion(p, level);
  double cF[3], cM[3];
  image_centre(L.F->g, cF);
  image_centre(L.M->g, cM);
  const orc_geom gF = mode_grid(p.resampling, false, L.F->g, L.M->g, cF, cM, T, L.F->g.direction);
  const orc_geom gM = mode_grid(p.resampling, false, L.F->g, L.M->g, cF, cM, T, L.M->g.direction);
  for (int a = 0; a < 3; ++a) st->level_dims[level][a] = gF.dims[a];
  const double c0[3] = {T.c[0], T.c[1], T.c[2]};
  set_center_to_origin(T);
  const Xf fT = translation_xf(cF), mT = translation_xf(cM);
  T.compose(fT.inverse(), false);
  T.compose(mT, true);
  Xf t12F, t12B;
  t12F.kind = t12B.kind = T.kind;
  half_space(T, t12F, t12B);
  const int np = T.nparams();
  double lastp[12], cur[12];
  T.get_params(lastp);
  std::vector<float> Fh, Mh;
  std::vector<uint8_t> Fmh, Mmh;
  int it = 0;
  const int maxi = std::abs(max_iter);
  for (; it < maxi; ++it) {
    double par[12];
    Xf cF12, cM12;
    cF12.kind = cM12.kind = T.kind;
    t12F.get_params(par);
    cF12.set_params(par);
    cF12.compose(fT, false);
    t12B.get_params(par);
    cM12.set_params(par);
    cM12.compose(mT, false);
    resample(L.F->v.data(), nullptr, L.F->g, cF12, gF, &Fh, nullptr, true);
    resample(L.M->v.data(), nullptr, L.M->g, cM12, gM, &Mh, nullptr, true);
    resample(L.Ffull, L.Fmask, *L.Fg, cF12, gF, nullptr, &Fmh, L.Fmask == nullptr);
    resample(L.Mfull, L.Mmask, *L.Mg, cM12, gM, nullptr, &Mmh, L.Mmask == nullptr);
    std::vector<double> xb, ym, fb, fm;
    if (g_truth) {
      // moving half -> moving -> fixed -> fixed half, and the swapped pass
      exact_pairs(gM, p, after(*g_truth, cM12), cF12.inverse(), xb, ym);
      exact_pairs(gF, p, after(g_truth->inverse(), cF12), cM12.inverse(), fb, fm);
    } else {
      match(Mh, Mmh, Fh, Fmh, gM, p, portion, xb, ym, &st->voxel_evals);  // base = moving
      match(Fh, Fmh, Mh, Mmh, gF, p, portion, fb, fm, &st->voxel_evals);  // swapped
    }
    xb.insert(xb.end(), fm.begin(), fm.end());
    ym.insert(ym.end(), fb.begin(), fb.end());
    st->points[level] = (int64_t)xb.size() / 3;
    st->inner_iterations += lts_symmetric(T, t12F, t12B, xb, ym, p.lts_tolerance);
    T.get_params(cur);
    double d;
    params_sq_delta(lastp, cur, np, &d);
    if (p.verbosity >= 1)
      std::printf("oracle level %d iteration %d: %lld blocks, delta %.6g\n", level, it + 1,
                  (long long)st->points[level], d), std::fflush(stdout);
    std::memcpy(lastp, cur, sizeof(double) * np);
    if (d < p.stop_tolerance) {
      ++it;
      break;
    }
  }
  st->iterations[level] = it;
  double par[12];
  Xf cF12, cM12;
  cF12.kind = cM12.kind = T.kind;
  t12F.get_params(par);
  cF12.set_params(par);
  cF12.compose(fT, false);
  t12B.get_params(par);
  cM12.set_params(par);
  cM12.compose(mT, false);
  const Xf inv = cM12.inverse();
  T.set_identity();
  T.compose(inv, true);
  T.compose(cF12, false);
  update_center(T, c0);
}

// Classic StartOptimization with middle resampling.
void start_opt_classic(Xf& T, const LevelIn& L, int max_iter, const orc_reg_params& p,
                       orc_reg_stats* st, int level) {
  const double portion = level_portion(p, level);
  double cF[3], cM[3];
  image_centre(L.F->g, cF);
  image_centre(L.M->g, cM);
  const orc_geom gF = mode_grid(p.resampling, true, L.F->g, L.M->g, cF, cM, T, L.F->g.direction);
  const orc_geom gM = mode_grid(p.resampling, true, L.F->g, L.M->g, cF, cM, T, L.M->g.direction);
  for (int a = 0; a < 3; ++a) st->level_dims[level][a] = gF.dims[a];
  Xf I;
  I.kind = T.kind;
  I.set_identity();
  std::vector<float> Fh, Mh;
  std::vector<uint8_t> Fmh, Mmh;
  resample(L.M->v.data(), nullptr, L.M->g, I, gM, &Mh, nullptr, true);
  resample(L.Mfull, L.Mmask, *L.Mg, I, gM, nullptr, &Mmh, L.Mmask == nullptr);
  const int np = T.nparams();
  double lastp[12], cur[12];
  T.get_params(lastp);
  int it = 0;
  const int maxi = std::abs(max_iter);
  for (; it < maxi; ++it) {
    resample(L.F->v.data(), nullptr, L.F->g, T, gF, &Fh, nullptr, true);
    resample(L.Ffull, L.Fmask, *L.Fg, T, gF, nullptr, &Fmh, L.Fmask == nullptr);
    std::vector<double> xb, ym;
    if (g_truth)  // the moving grid holds moving points; Fh(h) = F(T(h))
      exact_pairs(gM, p, *g_truth, T.inverse(), xb, ym);
    else
      match(Mh, Mmh, Fh, Fmh, gM, p, portion, xb, ym, &st->voxel_evals);
    st->points[level] = (int64_t)xb.size() / 3;
    st->inner_iterations += lts_classic(T, xb, ym, p.lts_tolerance);
    T.get_params(cur);
    double d;
    params_sq_delta(lastp, cur, np, &d);
    if (p.verbosity >= 1)
      std::printf("oracle level %d iteration %d: %lld blocks, delta %.6g\n", level, it + 1,
                  (long long)st->points[level], d), std::fflush(stdout);
    std::memcpy(lastp, cur, sizeof(double) * np);
    if (d < p.stop_tolerance) {
      ++it;
      break;
    }
  }
  st->iterations[level] = it;
}

}  // namespace

extern "C" {

void orc_reg_default_params(orc_reg_params* p) {
  // mirorr.cxx:99-281 defaults, except pyr-end 0 (all levels).
  p->symmetric = 1;
  p->affine = 0;
  p->bw = 4;
  p->bg = 1;
  p->nw = 7;
  p->ng = 3;
  p->iterations = 0;
  p->pyr_start = 1;
  p->pyr_end = 0;
  p->pyr_num = 1024;
  p->pyr_min_size = 32;
  p->portion_kept = 0.5;
  p->variance_select = 0;
  p->stop_tolerance = 1e-3;
  p->lts_tolerance = 1e-3;
  p->metric = 0;
  p->resampling = 0;
  p->nthreads = 0;
  p->verbosity = 0;
  p->centre_init = 1;
}

static int register_impl(const float* fixed, const uint8_t* fmask, const orc_geom* fg,
                         const float* moving, const uint8_t* mmask, const orc_geom* mg,
                         const orc_reg_params* p, const double* init_params,
                         const double* init_center, double* out_params, double* out_center,
                         orc_reg_stats* stats) {
  orc_reg_stats st{};
  g_verbosity = p->verbosity;
  // Schedules (PyramidScheduleTuner::UpdateSymmetric) and cropping.
  int msched[3 * 32], fsched[3 * 32];
  const int rows = orc_schedule_symmetric(mg->dims, mg->spacing, fg->dims, fg->spacing,
                                          (unsigned)p->pyr_min_size, (unsigned)p->pyr_num,
                                          msched, fsched, 32);
  if (rows <= 0) return 1;
  int level_min = p->pyr_start == 0 ? 1 : p->pyr_start;
  int level_max = p->pyr_end == 0 ? rows : p->pyr_end;
  int mn, mx;
  orc_cropped_min_max(level_min, level_max, rows, &mn, &mx);
  std::vector<int> ms(msched + 3 * (mn - 1), msched + 3 * mx);
  std::vector<int> fs(fsched + 3 * (mn - 1), fsched + 3 * mx);
  const int nlev = mx - mn + 1;
  const std::vector<Img> Fp = build_pyramid(fixed, *fg, fs, nlev);
  const std::vector<Img> Mp = build_pyramid(moving, *mg, ms, nlev);
  int t_max = orc_cropped_level(level_max, nlev, rows);
  t_max = std::min(t_max, nlev);  // the reference over-runs the cropped pyramid here
  // Initial transform.
  Xf T;
  T.kind = p->affine ? kAffine : kEuler3D;
  T.set_identity();
  if (init_params) {
    if (init_center) T.set_center(init_center);
    T.set_params(init_params);
  } else if (p->centre_init) {
    double cM[3], cF[3];
    image_centre(*mg, cM);
    image_centre(*fg, cF);
    const double t[3] = {cF[0] - cM[0], cF[1] - cM[1], cF[2] - cM[2]};
    T.set_center(cM);
    T.set_translation(t);
    // Composed into an identity transform with centre 0 before RunMirorr
    // (itkMirorrPyramidWrapper.txx:626-640).
    set_center_to_origin(T);
  }
  double cur[12];
  T.get_params(cur);
  int max_it = p->iterations != 0 ? std::abs(p->iterations) : 5;
  const bool halve = p->iterations < 0;
  st.levels = t_max;
  for (int level = 0; level < t_max && level < 16; ++level) {
    T.set_params(cur);
    LevelIn L{&Fp[level], &Mp[level], fixed, fmask, fg, moving, mmask, mg};
    if (p->symmetric)
      start_opt_symmetric(T, L, max_it, *p, &st, level);
    else
      start_opt_classic(T, L, max_it, *p, &st, level);
    T.get_params(cur);
    if (halve) max_it = std::max(1, max_it / 2);
  }
  T.get_params(out_params);
  if (out_center)
    for (int k = 0; k < 3; ++k) out_center[k] = T.c[k];
  if (stats) *stats = st;
  return 0;
}

int orc_register(const float* fixed, const uint8_t* fmask, const orc_geom* fg,
                 const float* moving, const uint8_t* mmask, const orc_geom* mg,
                 const orc_reg_params* p, const double* init_params,
                 const double* init_center, double* out_params, double* out_center,
                 orc_reg_stats* stats) {
  try {
    return register_impl(fixed, fmask, fg, moving, mmask, mg, p, init_params, init_center,
                         out_params, out_center, stats);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "oracle registration failed: %s\n", e.what());
    return 2;
  }
}

// The level loop with exact correspondences from a truth transform
// (y = M x + t, moving -> fixed physical points) in place of the block
// search: images of the given geometries are all zero and only set the
// grids. A check of the loop's transform algebra (centring, half-space
// roots, composition, re-centring, level hand-over).
int orc_register_synthetic(const orc_geom* fg, const orc_geom* mg, const orc_reg_params* p,
                           const double truth_M[9], const double truth_t[3],
                           const double* init_params, const double* init_center,
                           double* out_params, double* out_center, orc_reg_stats* stats) {
  Xf truth;
  truth.kind = kAffine;
  truth.set_identity();
  M3 m;
  for (int i = 0; i < 9; ++i) m.a[i / 3][i % 3] = truth_M[i];
  truth.set_matrix(m, false);
  truth.set_translation(truth_t);
  std::vector<float> F(nvox(*fg), 0.0f), M(nvox(*mg), 0.0f);
  g_truth = &truth;
  const int rc = orc_register(F.data(), nullptr, fg, M.data(), nullptr, mg, p, init_params,
                              init_center, out_params, out_center, stats);
  g_truth = nullptr;
  return rc;
}

}  // extern "C"
