// Device-resident multi-resolution registration (host driver + C ABI).
//
// Mirrors MirorrPyramidImplement::RunMirorr (itkMirorrPyramidImplement.txx:71-158)
// with the classic (itkMirorrRegistrationMethod.txx:422-659) and symmetric
// (itkSymMirorrRegistrationMethod.txx:238-720) StartOptimization in
// middle-space resampling mode. All per-voxel and per-block work runs on the
// GPU and the transforms of the iteration live in device memory (OuterState,
// LtsState); the host only reads a few flags per chunk of LTS inner
// iterations (the last chunk's carry the outer stop flag). Level set-up (grids, centring, half-space
// square roots) is FP64 host code on 4x4 matrices.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "bm_kernels.hpp"
#include "context.hpp"
#include "imaging.hpp"
#include "lts.hpp"
#include "schedule.hpp"
#include "storage.hpp"

using namespace mirorr;

struct mirorr_volume {
  mirorr_ctx* ctx = nullptr;
  DeviceBuffer data, mask;
  Geom geom{};
  bool has_mask = false;
  int storage = mirorr::kStoreF32;  // element type of `data`
  double enc_offset = 0.0, enc_scale = 1.0;  // half storage: (v - offset) * scale
  float enc_zero = 0.0f;            // the encoding of intensity 0 (half storage)
};

namespace {

size_t nvox(const Geom& g) { return (size_t)g.dims[0] * g.dims[1] * g.dims[2]; }

void to_phys(const Geom& g, const double idx[3], double p[3]) {
  for (int r = 0; r < 3; ++r) {
    double s = g.origin[r];
    for (int c = 0; c < 3; ++c) s += g.dir[3 * r + c] * g.spacing[c] * idx[c];
    p[r] = s;
  }
}
// ComputeImageCenter: physical point of continuous index (size - 1) / 2.
void centre_of(const Geom& g, double c[3]) {
  const double idx[3] = {(g.dims[0] - 1) / 2.0, (g.dims[1] - 1) / 2.0, (g.dims[2] - 1) / 2.0};
  to_phys(g, idx, c);
}

// Output geometry of a pyramid level for total shrink factors f
// (MultiResolutionPyramidImageFilter output information).
Geom level_geom(const Geom& in, const int f[3]) {
  Geom o = in;
  double ds[3];
  for (int a = 0; a < 3; ++a) {
    o.spacing[a] = in.spacing[a] * f[a];
    o.dims[a] = std::max(1, (int)std::floor((double)in.dims[a] / f[a]));
    ds[a] = o.spacing[a] - in.spacing[a];
  }
  for (int r = 0; r < 3; ++r) {
    double v = 0.0;
    for (int c = 0; c < 3; ++c) v += in.dir[3 * r + c] * ds[c];
    o.origin[r] = in.origin[r] + 0.5 * v;
  }
  return o;
}

// Middle (half-way) grid (itkSymMirorrRegistrationMethod.txx:255-283; the
// classic method adds the moving image centre, itkMirorrRegistrationMethod.txx:495-514).
Geom middle_grid(const Geom& F, const Geom& M, const double cF[3], const double cM[3],
                 bool add_moving_centre, const double* dir) {
  Geom h = F;
  for (int a = 0; a < 3; ++a) {
    h.origin[a] = ((F.origin[a] - cF[a]) + (M.origin[a] - cM[a])) / 2.0 +
                  (add_moving_centre ? cM[a] : 0.0);
    h.spacing[a] = (F.spacing[a] + M.spacing[a]) / 2.0;
    h.dims[a] = (int)(((double)(F.dims[a] + M.dims[a]) + 0.5) / 2.0);
  }
  for (int i = 0; i < 9; ++i) h.dir[i] = dir[i];
  return h;
}

// Grid the images are resampled onto for one level under a resampling mode
// (itkSymMirorrRegistrationMethod.txx:264-301, itkMirorrRegistrationMethod.txx:
// 450-531). Symmetric: the grid is in centred coordinates (basic = middle).
// Classic: middle adds the moving centre; fixed takes the fixed grid with its
// origin mapped by the inverse of the level's start transform T; moving and
// basic take the moving grid (basic with the moving direction for both
// images, as its resampler uses the moving image as reference).
Geom level_grid(int mode, bool classic, const Geom& F, const Geom& M, const double cF[3],
                const double cM[3], const xf::Xf& T, const double* dir) {
  if (mode == MIRORR_RESAMPLE_MIDDLE || (mode == MIRORR_RESAMPLE_BASIC && !classic))
    return middle_grid(F, M, cF, cM, classic, dir);
  const bool fixed = mode == MIRORR_RESAMPLE_FIXED;
  Geom h = fixed ? F : M;
  const double* c = fixed ? cF : cM;
  if (!classic) {
    for (int a = 0; a < 3; ++a) h.origin[a] -= c[a];
  } else if (fixed) {
    double o[3];
    xf::apply(xf::inverse(T), F.origin, o);
    for (int a = 0; a < 3; ++a) h.origin[a] = o[a];
  }
  const double* d = (classic && mode == MIRORR_RESAMPLE_BASIC) ? M.dir : dir;
  for (int i = 0; i < 9; ++i) h.dir[i] = d[i];
  return h;
}

struct Level {
  DeviceBuffer data;
  Geom geom;
};

// Recursive pyramid of one image (coarsest first) into `out`, whose level
// buffers are reused when large enough. Levels hold elements of `storage`.
void build_pyramid(const void* d_in, const Geom& g, const std::vector<int>& sched, int rows,
                   std::vector<Level>& out, DeviceBuffer& tmp, DeviceBuffer& map, Budget* bud,
                   hipStream_t s, int storage = kStoreF32, float enc_zero = 0.0f) {
  if ((int)out.size() < rows) out.resize(rows);
  map.ensure(sizeof(IndexMap), bud);
  for (int l = rows - 1; l >= 0; --l) {
    const int f[3] = {sched[3 * l], sched[3 * l + 1], sched[3 * l + 2]};
    out[l].geom = level_geom(g, f);
    out[l].data.ensure(storage_bytes(storage) * nvox(out[l].geom), bud);
    int ratio[3];
    const void* prev;
    const Geom* pg;
    if (l == rows - 1) {
      for (int a = 0; a < 3; ++a) ratio[a] = f[a];
      prev = d_in;
      pg = &g;
    } else {
      for (int a = 0; a < 3; ++a) ratio[a] = f[a] / sched[3 * (l + 1) + a];
      prev = out[l + 1].data.as<void>();
      pg = &out[l + 1].geom;
    }
    tmp.ensure(sizeof(double) * (2 * nvox(*pg) + 3 * 64), bud);
    launch_pyramid_step(prev, *pg, ratio, out[l].geom, out[l].data.as<void>(), tmp.as<double>(),
                        map.as<IndexMap>(), s, storage, enc_zero);
  }
}

// Shrink factors of the levels a registration runs (coarsest first): the
// schedule cropped to [pyr_start, pyr_end] (itkPyramidScheduleTuner.h:148-161).
struct LevelPlan {
  int nlev = 0, t_max = 0;
  std::vector<int> fs, ms;
};

LevelPlan plan_levels(const mirorr_reg_params* p, const Geom& FG, const Geom& MG) {
  const Schedule sch = schedule_symmetric(MG.dims, MG.spacing, FG.dims, FG.spacing,
                                          (unsigned)p->pyr_min_size, (unsigned)p->pyr_num);
  if (sch.rows <= 0) throw Error(kErrArg, "empty pyramid schedule");
  const int rows = sch.rows;
  const int level_min = p->pyr_start == 0 ? 1 : p->pyr_start;
  const int level_max = p->pyr_end == 0 ? rows : p->pyr_end;
  int mn = cropped_level(level_min, 1, rows), mx = cropped_level(level_max, rows, rows);
  if (mn > mx) mx = mn;
  LevelPlan pl;
  pl.nlev = mx - mn + 1;
  pl.ms.assign(sch.moving.begin() + 3 * (mn - 1), sch.moving.begin() + 3 * mx);
  pl.fs.assign(sch.fixed.begin() + 3 * (mn - 1), sch.fixed.begin() + 3 * mx);
  pl.t_max = std::min(cropped_level(level_max, pl.nlev, rows), pl.nlev);
  return pl;
}

// Per-registration device workspace, grown to the largest level.
struct Workspace {
  DeviceBuffer Fh, Mh, Fmh, Mmh;              // resampled images and masks
  DeviceBuffer statsF, statsM;                // window statistics
  DeviceBuffer xtF, xtM, sxxF, sxxM, sumF, sumM, flF, flM;  // block preparation
  DeviceBuffer best, second, match;           // raw BM results (one pass)
  DeviceBuffer queue, qbits, qcount;
  DeviceBuffer best2, second2, match2, queue2, qbits2, qcount2;  // the second pass's
  DeviceBuffer X, Y, kept;                    // LTS pairs
  DeviceBuffer r, hist1, hist2, cand, part;
  DeviceBuffer lts, outer, mapI, flagsD;  // flagsD: Flags for one read-back
  DeviceBuffer pyr_tmp, pyr_map;
  DeviceBuffer vsel_st, vsel_hist, vsel_evals;  // block-variance selection
  DeviceBuffer rlist, rcount, rlist2, rcount2;  // blocks listed for the exact re-scoring
  DeviceBuffer rescored, rescored2;             // blocks re-scored in double, per pass
};

// What lts_finish writes for the host: the LTS's active flag, the outer
// stop flag and the inner iterations run so far.
struct Flags {
  int active;
  int stop;
  int iters;
  int pad;
};

// Bytes of every host value an outer iteration's launches depend on: two
// iterations with equal keys queue the same launches with the same
// arguments. Fields are appended one by one (no struct padding).
struct KeyBytes {
  std::vector<unsigned char> b;
  template <class T>
  void add(const T& v) {
    static_assert(std::is_trivially_copyable<T>::value, "plain values only");
    const unsigned char* p = reinterpret_cast<const unsigned char*>(&v);
    b.insert(b.end(), p, p + sizeof(T));
  }
  void geom(const Geom& g) {
    for (int a = 0; a < 3; ++a) add(g.dims[a]);
    for (int a = 0; a < 3; ++a) add(g.spacing[a]);
    for (int a = 0; a < 3; ++a) add(g.origin[a]);
    for (int i = 0; i < 9; ++i) add(g.dir[i]);
  }
  void buffers(const Workspace& w) {
    for (const DeviceBuffer* d :
         {&w.Fh, &w.Mh, &w.Fmh, &w.Mmh, &w.statsF, &w.statsM, &w.xtF, &w.xtM, &w.sxxF, &w.sxxM,
          &w.sumF, &w.sumM, &w.flF, &w.flM, &w.best, &w.second, &w.match, &w.queue, &w.qbits,
          &w.qcount, &w.best2, &w.second2, &w.match2, &w.queue2, &w.qbits2, &w.qcount2, &w.X,
          &w.Y, &w.kept, &w.r, &w.hist1, &w.hist2, &w.cand, &w.part, &w.lts, &w.outer, &w.mapI,
          &w.flagsD, &w.vsel_st, &w.vsel_hist, &w.vsel_evals, &w.rlist, &w.rcount, &w.rlist2,
          &w.rcount2, &w.rescored, &w.rescored2})
      add(d->as<void>());
  }
};

// An instantiated graph of one outer iteration of a level, captured from the
// launches an enqueue function makes on a stream (and on the streams it
// forks from it), with the key of the host values it was captured under.
struct IterationGraph {
  hipGraphExec_t exec = nullptr;
  std::vector<unsigned char> key;
  int64_t evals = 0;  // voxel-evals the captured enqueue counted on the host
  IterationGraph() = default;
  IterationGraph(const IterationGraph&) = delete;
  IterationGraph& operator=(const IterationGraph&) = delete;
  ~IterationGraph() { reset(); }
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    key.clear();
    evals = 0;
  }
  template <class F>
  void capture(hipStream_t s, F&& enqueue) {
    reset();
    MIRORR_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipGraph_t g = nullptr;
    try {
      enqueue();
    } catch (...) {
      // leave the stream out of capture mode; the partial graph is dropped
      if (hipStreamEndCapture(s, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
      throw;
    }
    MIRORR_HIP(hipStreamEndCapture(s, &g));
    const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      exec = nullptr;
      throw Error(kErrPlatform, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
  }
};

}  // namespace

namespace mirorr {

struct RegCache {
  Workspace w;
  std::vector<Level> F, M;  // pyramid levels of the fixed and moving image
  Flags* flags = nullptr;   // pinned host flags read once per LTS chunk
  // Per level, the last captured outer iteration (mirorr_ctx_set_iteration_graphs);
  // replayed while the level's key matches.
  IterationGraph graphs[16];
  RegCache() { MIRORR_HIP(hipHostMalloc((void**)&flags, sizeof(Flags), hipHostMallocDefault)); }
  ~RegCache() {
    if (flags) (void)hipHostFree(flags);
  }
};

// The context's registration cache, created on first use (caller holds c->mu).
RegCache& reg_cache(mirorr_ctx* c) {
  if (!c->reg) c->reg = std::make_shared<RegCache>();
  return *c->reg;
}

// Batched registrations: one registration cache per pair, the per-pair
// arguments of the batched searches and LTS, and the pairs' Flags (device and
// pinned host).
struct BatchCache {
  std::vector<std::unique_ptr<RegCache>> regs;
  DeviceBuffer args[2], items, flagsD;
  Flags* flagsH = nullptr;
  int cap = 0;
  BatchCache() = default;
  BatchCache(const BatchCache&) = delete;
  BatchCache& operator=(const BatchCache&) = delete;
  ~BatchCache() {
    if (flagsH) (void)hipHostFree(flagsH);
  }
  void ensure(int n, Budget* bud) {
    while ((int)regs.size() < n) regs.emplace_back(new RegCache());
    for (DeviceBuffer& a : args) a.ensure(search_batch_bytes(n), bud);
    items.ensure(sizeof(LtsItem) * n, bud);
    flagsD.ensure(sizeof(Flags) * n, bud);
    if (cap < n) {
      if (flagsH) (void)hipHostFree(flagsH);
      flagsH = nullptr;
      cap = 0;
      MIRORR_HIP(hipHostMalloc((void**)&flagsH, sizeof(Flags) * n, hipHostMallocDefault));
      cap = n;
    }
  }
};

BatchCache& batch_cache(mirorr_ctx* c) {
  if (!c->batch) c->batch = std::make_shared<BatchCache>();
  return *c->batch;
}

struct RegRun {
  mirorr_ctx* ctx;
  const mirorr_reg_params* p;
  Workspace& w;
  mirorr_reg_stats* stats;
  int kind;
  bool symmetric;
  hipStream_t s;
  Budget* bud;
  bool finest = false;  // the current level is the last (full-resolution) one
  int storage = kStoreF32;  // element type of the levels and half-way images

  // The context's second stream (and fork/join events), created on first use.
  hipStream_t second_stream() {
    if (!ctx->stream2) {
      MIRORR_HIP(hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking));
      MIRORR_HIP(hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
      MIRORR_HIP(hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
    }
    return ctx->stream2;
  }

  // Block preparation of a base image and, with portion < 1, the variance
  // selection over its blocks. Returns whether blocks were dropped.
  bool prepare_base(const void* base, const uint8_t* bmask, const BMGrid& g, DeviceBuffer& xt,
                    DeviceBuffer& sxx, DeviceBuffer& sum, DeviceBuffer& fl, double portion,
                    hipStream_t s) {
    const BlockPrep bp{xt.as<float>(), sxx.as<float>(), sum.as<float>(), fl.as<uint32_t>()};
    launch_block_prep(base, bmask, g, bp, s, storage);
    if (!(portion < 1.0)) return false;
    VarSelScratch vs{w.vsel_st.as<VarSelState>(), w.vsel_hist.as<unsigned int>(),
                     w.vsel_evals.as<unsigned long long>()};
    return launch_variance_select(g, bp, portion, vs, s);
  }

  // One NCC pass base -> search over the half grid as a tiled-search job,
  // into the first or (second_set) the second set of result buffers.
  SearchJob ncc_job(const void* base, const uint8_t* bmask, const void* search,
                    const uint8_t* smask, const Geom& g_base, DeviceBuffer& xt, DeviceBuffer& sxx,
                    DeviceBuffer& sum, DeviceBuffer& fl, DeviceBuffer& stats_search,
                    bool second_set) {
    SearchJob j;
    j.x = base;
    j.bmask = bmask;
    j.y = search;
    j.smask = smask;
    j.g = make_grid(g_base.dims, p->block_width, p->block_gap, p->nhood_width, p->nhood_gap);
    j.bp = BlockPrep{xt.as<float>(), sxx.as<float>(), sum.as<float>(), fl.as<uint32_t>()};
    j.st = SearchStats{stats_search.as<float2>()};
    j.res = second_set
                ? BMResult{w.best2.as<float>(), w.second2.as<float>(), w.match2.as<int32_t>()}
                : BMResult{w.best.as<float>(), w.second.as<float>(), w.match.as<int32_t>()};
    DeviceBuffer& qu = second_set ? w.queue2 : w.queue;
    DeviceBuffer& qc = second_set ? w.qcount2 : w.qcount;
    DeviceBuffer& qb = second_set ? w.qbits2 : w.qbits;
    j.scr = BMScratch{qu.as<int32_t>(), qc.as<int32_t>(), qb.as<unsigned long long>(),
                      tile_count(j.g),
                      (ctx->prof.finest_only() && !finest) ? nullptr : &ctx->prof};
    j.scr.rlist = (second_set ? w.rlist2 : w.rlist).as<int32_t>();
    j.scr.rcount = (second_set ? w.rcount2 : w.rcount).as<int32_t>();
    j.scr.rescored = (second_set ? w.rescored2 : w.rescored).as<unsigned long long>();
    j.scr.exact_records = 0;  // the pairs need the filter's decision and the match only
    return j;
  }

  // The LTS pairs of a pass's records (the second pass's after the first's).
  void post_pass(const BMGrid& g, const Geom& g_base, const BMResult& res, int swap,
                 bool selected, LtsState* st, hipStream_t s) {
    const int64_t off = swap ? g.nblocks : 0;
    launch_postprocess_pairs(g, g_base, res, swap, w.kept.as<uint8_t>() + off,
                             w.X.as<double>() + 3 * off, w.Y.as<double>() + 3 * off,
                             &st->n_kept, s);
    if (!selected) stats->voxel_evals += voxel_evals(g);
  }

  // One BM pass base -> search over the half grid, writing LTS pairs. With
  // `selected`, the prepared flags drop blocks and the pass's voxel-evals are
  // counted on the device.
  void match_pass(const void* base, const uint8_t* bmask, const void* search,
                  const uint8_t* smask, const Geom& g_base, DeviceBuffer& xt, DeviceBuffer& sxx,
                  DeviceBuffer& sum, DeviceBuffer& fl, DeviceBuffer& stats_search, int swap,
                  bool selected, LtsState* st, hipStream_t s, bool second_set) {
    const BMGrid g = make_grid(g_base.dims, p->block_width, p->block_gap, p->nhood_width,
                               p->nhood_gap);
    if (g.nblocks <= 0) return;
    if (p->block_metric == MIRORR_METRIC_NC && tiled_supported(g)) {
      SearchJob j = ncc_job(base, bmask, search, smask, g_base, xt, sxx, sum, fl, stats_search,
                            second_set);
      launch_ncc_search(j.x, j.bmask, j.y, j.smask, j.g, j.bp, j.st, j.scr, j.res, s, storage);
      post_pass(g, g_base, j.res, swap, selected, st, s);
      return;
    }
    if (storage != kStoreF32)
      throw Error(kErrArg, "half-precision images need the NCC metric and the tiled search "
                           "(blocks of width 4, block gap 1, nhood width 5 to 13)");
    const BMResult res =
        second_set ? BMResult{w.best2.as<float>(), w.second2.as<float>(), w.match2.as<int32_t>()}
                   : BMResult{w.best.as<float>(), w.second.as<float>(), w.match.as<int32_t>()};
    launch_metric_search(p->block_metric, static_cast<const float*>(base), bmask,
                         static_cast<const float*>(search), smask, g, res, s,
                         selected ? fl.as<uint32_t>() : nullptr);
    post_pass(g, g_base, res, swap, selected, st, s);
  }

  static int64_t voxel_evals(const BMGrid& g) {
    int64_t per[3] = {0, 0, 0};
    for (int a = 0; a < 3; ++a)
      for (int i = 0; i < g.nb[a]; ++i) {
        int f, c;
        cand_range(g.start[a] + i * g.ng, g.dims[a], g.bw, g.bg, g.nw, &f, &c);
        per[a] += c;
      }
    return per[0] * per[1] * per[2] * (int64_t)g.bw * g.bw * g.bw;
  }

  void ensure_level_buffers(const Geom& gh) {
    const size_t n = nvox(gh);
    const BMGrid g = make_grid(gh.dims, p->block_width, p->block_gap, p->nhood_width,
                               p->nhood_gap);
    const int64_t nb = std::max<int64_t>(g.nblocks, 1);
    const int64_t n3 = (int64_t)g.bw * g.bw * g.bw;
    for (DeviceBuffer* b : {&w.Fh, &w.Mh}) b->ensure(storage_bytes(storage) * n, bud);
    for (DeviceBuffer* b : {&w.Fmh, &w.Mmh}) b->ensure(n, bud);
    for (DeviceBuffer* b : {&w.statsF, &w.statsM}) b->ensure(sizeof(float2) * n, bud);
    for (DeviceBuffer* b : {&w.xtF, &w.xtM}) b->ensure(sizeof(float) * nb * n3, bud);
    for (DeviceBuffer* b : {&w.sxxF, &w.sxxM, &w.sumF, &w.sumM}) b->ensure(sizeof(float) * nb, bud);
    for (DeviceBuffer* b : {&w.flF, &w.flM}) b->ensure(sizeof(uint32_t) * nb, bud);
    for (DeviceBuffer* b : {&w.best, &w.second, &w.best2, &w.second2})
      b->ensure(sizeof(float) * nb, bud);
    for (DeviceBuffer* b : {&w.match, &w.match2}) b->ensure(sizeof(int32_t) * 3 * nb, bud);
    for (DeviceBuffer* b : {&w.queue, &w.queue2})
      b->ensure(sizeof(int32_t) * std::max<int64_t>(tile_count(g), 1), bud);
    for (DeviceBuffer* b : {&w.qbits, &w.qbits2})
      b->ensure(sizeof(unsigned long long) * std::max<int64_t>(tile_count(g), 1), bud);
    for (DeviceBuffer* b : {&w.qcount, &w.qcount2, &w.rcount, &w.rcount2})
      b->ensure(sizeof(int32_t), bud);
    for (DeviceBuffer* b : {&w.rlist, &w.rlist2}) b->ensure(sizeof(int32_t) * nb, bud);
    const int64_t np = 2 * nb;
    w.X.ensure(sizeof(double) * 3 * np, bud);
    w.Y.ensure(sizeof(double) * 3 * np, bud);
    w.kept.ensure(np, bud);
    w.r.ensure(sizeof(double) * np, bud);
    w.cand.ensure(sizeof(unsigned long long) * np, bud);
    const int64_t parts = lts_parts(np);
    w.part.ensure(sizeof(double) * 32 * parts, bud);
    w.hist1.ensure(sizeof(unsigned int) * 4096, bud);
    w.hist2.ensure(sizeof(unsigned int) * 4096, bud);
    w.vsel_st.ensure(sizeof(VarSelState), bud);
    w.vsel_hist.ensure(sizeof(unsigned int) * kVarSelHistBins, bud);
  }

  LtsItem lts_item(LtsState* st, OuterState* outer, int64_t npts, int* flags) {
    return LtsItem{w.X.as<double>(), w.Y.as<double>(), w.kept.as<uint8_t>(), npts, st, outer,
                   w.r.as<double>(), w.hist1.as<unsigned int>(), w.hist2.as<unsigned int>(),
                   w.cand.as<unsigned long long>(), w.part.as<double>(), flags};
  }

  LtsBuffers lts_buffers(int64_t npts) {
    LtsBuffers lb;
    lb.r = w.r.as<double>();
    lb.hist1 = w.hist1.as<unsigned int>();
    lb.hist2 = w.hist2.as<unsigned int>();
    lb.cand = w.cand.as<unsigned long long>();
    lb.part = w.part.as<double>();
    lb.nparts_cap = lts_parts(npts);
    return lb;
  }

  // Inner iterations [it0, it1) (predicated on the LTS being active), the
  // finish and one read-back of its flags: when the run has converged, the
  // outer update is already done and its stop flag known.
  void enqueue_lts_chunk(LtsState* st, OuterState* outer, int64_t npts, Flags* h_flags, int it0,
                         int it1) {
    static_assert(sizeof(Flags) == 4 * sizeof(int), "lts_finish writes the first 3 ints");
    LtsBuffers lb = lts_buffers(npts);
    for (int j = it0; j < it1; ++j)
      lts_enqueue_iteration(w.X.as<double>(), w.Y.as<double>(), w.kept.as<uint8_t>(), npts, st,
                            outer, lb, j, s);
    lts_finish(st, outer, w.flagsD.as<int>(), s);
    MIRORR_HIP(hipMemcpyAsync(h_flags, w.flagsD.as<int>(), sizeof(Flags), hipMemcpyDeviceToHost,
                              s));
  }

  // Chunk sizes of the predicated LTS: the first chunk (`first`, given to
  // lts_begin and the same to lts_complete), then one of kLtsNext, then
  // kLtsChunk each. An iteration queued after the LTS has converged is a
  // no-op that still costs its launches; a chunk too short costs a host
  // round trip. lts_first is the first-chunk size of this pair's next LTS.
  static constexpr int kLtsChunk = 4, kLtsNext = 2, kLtsFirstMax = 8;
  int lts_first = kLtsChunk;

  // The inner iterations of the LTS that has just completed, as a first-chunk
  // size for the next one: successive LTS runs of a registration take
  // similar counts, since their inputs change little between iterations.
  static int first_chunk_after(const Flags& f) {
    return std::min(std::max(f.iters, 1), kLtsFirstMax);
  }

  // The LTS start and its first chunk, queued without a host wait (the
  // flags buffer is allocated beforehand: this may run under graph capture).
  void lts_begin(LtsState* st, OuterState* outer, int64_t npts, Flags* h_flags, int first) {
    if (w.flagsD.bytes() < sizeof(Flags)) throw Error(kErrState, "LTS flags buffer missing");
    if (first < 1 || first > 101) throw Error(kErrArg, "LTS first chunk out of range");
    LtsBuffers lb = lts_buffers(npts);
    lts_start(st, outer, kind, symmetric ? 1 : 0, p->lts_tolerance, lb, s);
    enqueue_lts_chunk(st, outer, npts, h_flags, 0, first);
  }

  // Waits for the first chunk's flags and runs further chunks while the LTS
  // is active (at most 101 inner iterations).
  void lts_complete(LtsState* st, OuterState* outer, int64_t npts, Flags* h_flags, int first) {
    MIRORR_HIP(hipStreamSynchronize(s));
    int it = first, step = kLtsNext;
    while (h_flags->active && it < 101) {
      const int it1 = std::min(it + step, 101);
      enqueue_lts_chunk(st, outer, npts, h_flags, it, it1);
      MIRORR_HIP(hipStreamSynchronize(s));
      it = it1;
      step = kLtsChunk;
    }
  }

  // Runs the LTS to convergence with predicated chunks; returns when done.
  void run_lts(LtsState* st, OuterState* outer, int64_t npts, Flags* h_flags) {
    w.flagsD.ensure(sizeof(Flags), bud);
    lts_begin(st, outer, npts, h_flags, kLtsChunk);
    lts_complete(st, outer, npts, h_flags, kLtsChunk);
  }
};

size_t registration_bytes_bound(const mirorr_reg_params* p, const Geom& FG, bool fmask,
                                const Geom& MG, bool mmask, int storage) {
  const size_t nF = nvox(FG), nM = nvox(MG);
  const size_t eb = storage_bytes(storage);
  // The two uploaded volumes and masks.
  size_t total = eb * (nF + nM) + (fmask ? nF : 0) + (mmask ? nM : 0);
  // Pyramid levels (build_pyramid), its double scratch, sized by the finest
  // input, and index map.
  const LevelPlan plan = plan_levels(p, FG, MG);
  for (int l = 0; l < plan.nlev; ++l) {
    total += eb * nvox(level_geom(FG, &plan.fs[3 * l]));
    total += eb * nvox(level_geom(MG, &plan.ms[3 * l]));
  }
  total += sizeof(double) * (2 * std::max(nF, nM) + 3 * 64) + sizeof(IndexMap);
  // Level buffers (RegRun::ensure_level_buffers) at a grid no larger than
  // the per-axis maxima of the two images: every resampling mode's grid is.
  Geom gh = FG;
  for (int a = 0; a < 3; ++a) gh.dims[a] = std::max(FG.dims[a], MG.dims[a]);
  const size_t n = nvox(gh);
  const BMGrid g = make_grid(gh.dims, p->block_width, p->block_gap, p->nhood_width,
                             p->nhood_gap);
  const size_t nb = (size_t)std::max<int64_t>(g.nblocks, 1);
  const size_t n3 = (size_t)g.bw * g.bw * g.bw;
  const size_t tiles = (size_t)std::max<int64_t>(tile_count(g), 1);
  const size_t np = 2 * nb;
  total += 2 * eb * n + 2 * n + 2 * sizeof(float2) * n;  // Fh, Mh, masks, stats
  total += 2 * sizeof(float) * nb * n3 + 4 * sizeof(float) * nb + 2 * sizeof(uint32_t) * nb;
  // best, second, match and the masked-tile queue, one set per pass
  total += 2 * (2 * sizeof(float) * nb + 3 * sizeof(int32_t) * nb);
  total += 2 * (tiles * (sizeof(int32_t) + sizeof(unsigned long long)) + sizeof(int32_t));
  total += 2 * 3 * sizeof(double) * np + np + sizeof(double) * np +
           sizeof(unsigned long long) * np;  // X, Y, kept, r, cand
  total += sizeof(double) * 32 * (size_t)lts_parts((int64_t)np);
  total += 2 * sizeof(unsigned int) * 4096 + sizeof(VarSelState) +
           sizeof(unsigned int) * kVarSelHistBins + 3 * sizeof(unsigned long long);
  total += 2 * (sizeof(int32_t) * nb + sizeof(int32_t));  // re-scoring lists, one per pass
  total += sizeof(LtsState) + sizeof(OuterState) + 2 * sizeof(IndexMap) + sizeof(Flags);
  return total;
}

}  // namespace mirorr

namespace {

constexpr int kMaxBatchPairs = 256;

void check_reg_params(const mirorr_reg_params* p) {
  if (p->block_width < 1 || p->block_gap < 1 || p->nhood_width < 1 || p->nhood_gap < 1)
    throw Error(kErrArg, "block-matcher parameters must be positive");
  // SetPortionMatchesKept (itkAbstractBlockMatcher.txx:83-99): 0 means 0.5.
  if (p->portion_kept < 0.0 || p->portion_kept > 1.0)
    throw Error(kErrArg, "portion_kept should be between 0 and 1.0");
  if (p->block_metric < MIRORR_METRIC_NC || p->block_metric > MIRORR_METRIC_CR)
    throw Error(kErrArg, "unknown block metric");
  if (p->resampling < MIRORR_RESAMPLE_MIDDLE || p->resampling > MIRORR_RESAMPLE_FIXED)
    throw Error(kErrArg, "Unknown resampling type");
  if (p->block_metric != MIRORR_METRIC_NC && p->block_width > kMaxMetricBW)
    throw Error(kErrArg, "block width above 16 for this block metric");
  if (!(p->stop_tolerance >= 0.0) || !(p->lts_tolerance >= 0.0))
    throw Error(kErrArg, "stop tolerances must be non-negative");
}

// One registration (MirorrPyramidImplement::RunMirorr, itkMirorrPyramidImplement.txx:71-158)
// as host state between the launches of its levels and iterations: the
// pyramids, the transform between levels, the current level's outer state
// and grids. The one-pair driver and the batched driver below queue its
// launches.
struct PairReg {
  mirorr_ctx* ctx;
  const mirorr_reg_params* p;
  const mirorr_volume* fixed;
  const mirorr_volume* moving;
  RegCache& cache;
  RegRun run;
  mirorr_reg_stats* stats;
  std::chrono::steady_clock::time_point t0;
  LevelPlan plan;
  xf::Xf T;
  double cur[12];
  int max_it = 5;
  bool halve = false;
  // the current level
  OuterState o{};
  double c0[3] = {0.0, 0.0, 0.0};
  BMGrid gh{};
  bool tiled = false, prep = false;
  double portion = 1.0;
  int64_t npts = 0;
  const uint8_t* fmask = nullptr;
  const uint8_t* mmask = nullptr;
  int fmode = kMaskInside, mmode = kMaskInside;
  bool selM = false, selF = false;
  int level = -1;

  PairReg(mirorr_ctx* c, const mirorr_reg_params* pp, const mirorr_volume* f,
          const mirorr_volume* m, const double* init_params, const double* init_center,
          mirorr_reg_stats* st, RegCache& rc, hipStream_t s)
      : ctx(c), p(pp), fixed(f), moving(m), cache(rc),
        run{c, pp, rc.w, st, pp->affine ? xf::kAffine : xf::kRigid, pp->symmetric != 0, s,
            &c->budget},
        stats(st), t0(std::chrono::steady_clock::now()) {
    if (fixed->storage != moving->storage)
      throw Error(kErrArg, "the fixed and moving volumes have different storage");
    run.storage = fixed->storage;
    Workspace& w = run.w;
    Budget* bud = run.bud;
    const Geom& FG = fixed->geom;
    const Geom& MG = moving->geom;
    plan = plan_levels(p, FG, MG);
    build_pyramid(fixed->data.as<void>(), FG, plan.fs, plan.nlev, cache.F, w.pyr_tmp, w.pyr_map,
                  bud, s, fixed->storage, fixed->enc_zero);
    build_pyramid(moving->data.as<void>(), MG, plan.ms, plan.nlev, cache.M, w.pyr_tmp, w.pyr_map,
                  bud, s, moving->storage, moving->enc_zero);
    // Initial transform: given, or centred (itkMirorrPyramidWrapper.txx:611-623:
    // centre of the moving image, translation between the image centres), then
    // composed into an identity transform with centre 0 as the command line
    // does before RunMirorr (itkMirorrPyramidWrapper.txx:626-640).
    T = xf::identity(run.kind);
    if (init_params) {
      if (init_center) xf::set_center(T, init_center);
      xf::set_params(T, init_params);
    } else if (p->centre_init) {
      double cM[3], cF[3], t[3];
      centre_of(MG, cM);
      centre_of(FG, cF);
      for (int k = 0; k < 3; ++k) t[k] = cF[k] - cM[k];
      xf::set_center(T, cM);
      xf::set_translation(T, t);
      xf::set_center_to_origin(T);
    }
    xf::get_params(T, cur);
    max_it = p->iterations != 0 ? std::abs(p->iterations) : 5;
    halve = p->iterations < 0;
    w.lts.ensure(sizeof(LtsState), bud);
    w.outer.ensure(sizeof(OuterState), bud);
    w.mapI.ensure(sizeof(IndexMap), bud);
    w.flagsD.ensure(sizeof(Flags), bud);
    std::memset(stats, 0, sizeof(*stats));
    w.vsel_evals.ensure(sizeof(unsigned long long), bud);
    MIRORR_HIP(hipMemsetAsync(w.vsel_evals.as<void>(), 0, sizeof(unsigned long long), s));
    for (DeviceBuffer* b : {&w.rescored, &w.rescored2}) {
      b->ensure(sizeof(unsigned long long), bud);
      MIRORR_HIP(hipMemsetAsync(b->as<void>(), 0, sizeof(unsigned long long), s));
    }
    stats->levels = plan.t_max;
  }

  int levels() const { return std::min(plan.t_max, 16); }
  LtsState* st() { return run.w.lts.as<LtsState>(); }
  OuterState* outer() { return run.w.outer.as<OuterState>(); }

  // Level start: the outer state, the resampling maps and the level's
  // buffers on the device; in classic mode the moving image and mask,
  // resampled once onto the grid.
  void begin_level(int lv) {
    level = lv;
    Workspace& w = run.w;
    Budget* bud = run.bud;
    hipStream_t s = run.s;
    const Geom& FG = fixed->geom;
    const Geom& MG = moving->geom;
    std::vector<Level>& Fp = cache.F;
    std::vector<Level>& Mp = cache.M;
    run.finest = level + 1 == levels();
    xf::set_params(T, cur);
    const Geom& gF = Fp[level].geom;
    const Geom& gM = Mp[level].geom;
    double cF[3], cM[3];
    centre_of(gF, cF);
    centre_of(gM, cM);
    o = OuterState{};
    o.gF = gF;
    o.gM = gM;
    o.gFfull = FG;
    o.gMfull = MG;
    for (int k = 0; k < 3; ++k) c0[k] = T.c[k];
    if (run.symmetric) {
      o.gHF = level_grid(p->resampling, false, gF, gM, cF, cM, T, gF.dir);
      o.gHM = level_grid(p->resampling, false, gF, gM, cF, cM, T, gM.dir);
      // Centre m_Transform and express it between centred images.
      xf::set_center_to_origin(T);
      o.fixT = xf::translation(cF);
      o.movT = xf::translation(cM);
      xf::compose(T, xf::inverse(o.fixT), false);
      xf::compose(T, o.movT, true);
      // ComputeHalfSpaceTransforms (itkSymMirorrRegistrationMethod.txx:139-201).
      xf::M4 m;
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) m.a[i][j] = (i == 3 && j == 3) ? 1.0 : 0.0;
      for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) m.a[i][j] = T.M.a[i][j];
        m.a[i][3] = T.t[i];
      }
      const xf::M4 m12 = xf::sqrt_db(m, 1e-14, 50);
      const xf::M4 im12 = xf::sqrt_db(xf::inverse(m), 1e-14, 50);
      o.t12F = xf::identity(run.kind);
      o.t12B = xf::identity(run.kind);
      xf::M3 K, L;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          K.a[i][j] = m12.a[i][j];
          L.a[i][j] = im12.a[i][j];
        }
      if (!xf::set_matrix(o.t12F, K) || !xf::set_matrix(o.t12B, L))
        throw Error(kErrAlgo, "Attempting to set a non-orthogonal rotation matrix");
      const double a[3] = {m12.a[0][3], m12.a[1][3], m12.a[2][3]};
      const double b[3] = {im12.a[0][3], im12.a[1][3], im12.a[2][3]};
      xf::set_translation(o.t12F, a);
      xf::set_translation(o.t12B, b);
    } else {
      o.gHF = level_grid(p->resampling, true, gF, gM, cF, cM, T, gF.dir);
      o.gHM = level_grid(p->resampling, true, gF, gM, cF, cM, T, gM.dir);
    }
    o.T = T;
    o.stop_tol = p->stop_tolerance;
    xf::get_params(T, o.lastp);
    for (int a = 0; a < 3; ++a) stats->level_dims[level][a] = o.gHF.dims[a];
    run.ensure_level_buffers(o.gHF);
    MIRORR_HIP(hipMemcpyAsync(outer(), &o, sizeof(o), hipMemcpyHostToDevice, s));
    outer_prepare_maps(outer(), run.symmetric ? 1 : 0, s);
    const int64_t nb = make_grid(o.gHF.dims, p->block_width, p->block_gap, p->nhood_width,
                                 p->nhood_gap).nblocks;
    npts = run.symmetric ? 2 * nb : nb;
    fmask = fixed->has_mask ? fixed->mask.as<uint8_t>() : nullptr;
    mmask = moving->has_mask ? moving->mask.as<uint8_t>() : nullptr;
    fmode = fixed->has_mask ? kMaskLinear : kMaskInside;
    mmode = moving->has_mask ? kMaskLinear : kMaskInside;
    gh = make_grid(o.gHF.dims, p->block_width, p->block_gap, p->nhood_width, p->nhood_gap);
    tiled = p->block_metric == MIRORR_METRIC_NC && tiled_supported(gh);
    // Portion of blocks searched: levels 0 and 1 use every block
    // (itkMirorrPyramidImplement.txx:116-119); the selection is opt-in.
    portion = (p->variance_select && level >= 2)
                  ? (p->portion_kept == 0.0 ? 0.5 : p->portion_kept)
                  : 1.0;
    prep = tiled || portion < 1.0;
    selM = selF = false;
    if (!run.symmetric) {
      // The moving image and mask are resampled once onto the grid.
      const xf::M3 I = xf::eye3();
      const double z[3] = {0.0, 0.0, 0.0};
      IndexMap mi[2] = {make_index_map(I, z, gM, o.gHM), make_index_map(I, z, MG, o.gHM)};
      w.mapI.ensure(sizeof(mi), bud);
      MIRORR_HIP(hipMemcpyAsync(w.mapI.as<IndexMap>(), mi, sizeof(mi), hipMemcpyHostToDevice, s));
      launch_resample(Mp[level].data.as<void>(), gM.dims, w.mapI.as<IndexMap>(), mmask, MG.dims,
                      w.mapI.as<IndexMap>() + 1, mmode, o.gHM.dims, w.Mh.as<void>(),
                      w.Mmh.as<uint8_t>(), s, run.storage, moving->enc_zero);
      if (prep)
        selM = run.prepare_base(w.Mh.as<void>(), w.Mmh.as<uint8_t>(), gh, w.xtM, w.sxxM, w.sumM,
                                w.flM, portion, s);
      // (the half-grid stays fixed through the level; hipMemcpy is ordered)
      MIRORR_HIP(hipStreamSynchronize(s));
    }
  }

  // 1. resampling onto the half-way grid under the current transforms and
  // the kept-pair counter, on s.
  void enqueue_resample(hipStream_t s) {
    Workspace& w = run.w;
    const Geom& FG = fixed->geom;
    const Geom& MG = moving->geom;
    launch_resample(cache.F[level].data.as<void>(), cache.F[level].geom.dims, &outer()->mapF,
                    fmask, FG.dims, &outer()->mmapF, fmode, o.gHF.dims, w.Fh.as<void>(),
                    w.Fmh.as<uint8_t>(), s, run.storage, fixed->enc_zero);
    if (run.symmetric)
      launch_resample(cache.M[level].data.as<void>(), cache.M[level].geom.dims, &outer()->mapM,
                      mmask, MG.dims, &outer()->mmapM, mmode, o.gHM.dims, w.Mh.as<void>(),
                      w.Mmh.as<uint8_t>(), s, run.storage, moving->enc_zero);
    lts_reset_count(st(), s);
  }

  // For the tiled search, the window statistics of the search images and
  // (symmetric) the block preparation of the base images: the first pass's
  // on s, the second pass's on s_b (which waits for the resampling).
  void enqueue_prep(hipStream_t s, hipStream_t s_b) {
    Workspace& w = run.w;
    if (tiled) {
      launch_search_stats(w.Fh.as<void>(), w.Fmh.as<uint8_t>(), gh,
                          SearchStats{w.statsF.as<float2>()}, s, run.storage);
      if (run.symmetric)
        launch_search_stats(w.Mh.as<void>(), w.Mmh.as<uint8_t>(), gh,
                            SearchStats{w.statsM.as<float2>()}, s_b, run.storage);
    }
    if (run.symmetric && prep) {
      selM = run.prepare_base(w.Mh.as<void>(), w.Mmh.as<uint8_t>(), gh, w.xtM, w.sxxM, w.sumM,
                              w.flM, portion, s);
      selF = run.prepare_base(w.Fh.as<void>(), w.Fmh.as<uint8_t>(), gh, w.xtF, w.sxxF, w.sumF,
                              w.flF, portion, s_b);  // the second pass's base
    }
  }

  // 2. block matching: pass 0 takes the moving half as base, pass 1
  // (symmetric) the fixed half.
  void match(int pass, hipStream_t s, bool second_set) {
    Workspace& w = run.w;
    if (pass == 0)
      run.match_pass(w.Mh.as<void>(), w.Mmh.as<uint8_t>(), w.Fh.as<void>(), w.Fmh.as<uint8_t>(),
                     o.gHM, w.xtM, w.sxxM, w.sumM, w.flM, w.statsF, 0, selM, st(), s, second_set);
    else
      run.match_pass(w.Fh.as<void>(), w.Fmh.as<uint8_t>(), w.Mh.as<void>(), w.Mmh.as<uint8_t>(),
                     o.gHF, w.xtF, w.sxxF, w.sumF, w.flF, w.statsM, 1, selF, st(), s, second_set);
  }
  // The same pass as one pair's job of a batched search, and its pairs.
  SearchJob job(int pass) {
    Workspace& w = run.w;
    if (pass == 0)
      return run.ncc_job(w.Mh.as<void>(), w.Mmh.as<uint8_t>(), w.Fh.as<void>(),
                         w.Fmh.as<uint8_t>(), o.gHM, w.xtM, w.sxxM, w.sumM, w.flM, w.statsF, false);
    return run.ncc_job(w.Fh.as<void>(), w.Fmh.as<uint8_t>(), w.Mh.as<void>(), w.Mmh.as<uint8_t>(),
                       o.gHF, w.xtF, w.sxxF, w.sumF, w.flF, w.statsM, true);
  }
  void pairs(int pass, hipStream_t s) {
    const SearchJob j = job(pass);
    run.post_pass(j.g, pass == 0 ? o.gHM : o.gHF, j.res, pass, pass == 0 ? selM : selF, st(), s);
  }

  void report(int it) {
    unsigned long long nk = 0;
    double d = 0.0;
    hipStream_t s = run.s;
    MIRORR_HIP(hipMemcpyAsync(&nk, &st()->n_kept, sizeof(nk), hipMemcpyDeviceToHost, s));
    MIRORR_HIP(hipMemcpyAsync(&d, &outer()->sqdiff, sizeof(d), hipMemcpyDeviceToHost, s));
    MIRORR_HIP(hipStreamSynchronize(s));
    std::printf("level %d iteration %d: %llu blocks, delta %.6g\n", level, it + 1, nk, d);
    std::fflush(stdout);
  }

  // Level end after `it` outer iterations: the outer state read back and the
  // transform for the next level.
  void end_level(int it) {
    hipStream_t s = run.s;
    stats->iterations[level] = it;
    MIRORR_HIP(hipMemcpyAsync(&o, outer(), sizeof(o), hipMemcpyDeviceToHost, s));
    unsigned long long nk = 0;
    MIRORR_HIP(hipMemcpyAsync(&nk, &st()->n_kept, sizeof(nk), hipMemcpyDeviceToHost, s));
    MIRORR_HIP(hipStreamSynchronize(s));
    stats->points[level] = (int64_t)nk;
    stats->inner_iterations += o.inner_total;
    if (run.symmetric) {
      // Final composition and re-centring (itkSymMirorrRegistrationMethod.txx:661-692).
      auto centered = [&](const xf::Xf& t12, const xf::Xf& shift) {
        double par[12];
        xf::get_params(t12, par);
        xf::Xf c = xf::identity(run.kind);
        xf::set_params(c, par);
        xf::compose(c, shift, false);
        return c;
      };
      const xf::Xf cF12 = centered(o.t12F, o.fixT), cM12 = centered(o.t12B, o.movT);
      T = xf::identity(run.kind);
      xf::compose(T, xf::inverse(cM12), true);
      xf::compose(T, cF12, false);
      xf::update_center(T, c0);
    } else {
      T = o.T;
    }
    xf::get_params(T, cur);
    if (halve) max_it = std::max(1, max_it / 2);
  }

  void finish(double* out_params, double* out_center) {
    hipStream_t s = run.s;
    Workspace& w = run.w;
    unsigned long long sel_evals = 0, rescored[2] = {0, 0};
    MIRORR_HIP(hipMemcpyAsync(&sel_evals, w.vsel_evals.as<void>(), sizeof(sel_evals),
                              hipMemcpyDeviceToHost, s));
    MIRORR_HIP(hipMemcpyAsync(&rescored[0], w.rescored.as<void>(), sizeof(rescored[0]),
                              hipMemcpyDeviceToHost, s));
    MIRORR_HIP(hipMemcpyAsync(&rescored[1], w.rescored2.as<void>(), sizeof(rescored[1]),
                              hipMemcpyDeviceToHost, s));
    MIRORR_HIP(hipStreamSynchronize(s));
    stats->voxel_evals += (int64_t)sel_evals;
    stats->rescored_blocks = (int64_t)(rescored[0] + rescored[1]);
    const int np = xf::nparams(T);
    for (int i = 0; i < np; ++i) out_params[i] = cur[i];
    if (out_center)
      for (int k = 0; k < 3; ++k) out_center[k] = T.c[k];
    stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

void run_registration(mirorr_ctx* ctx, const mirorr_reg_params* p, const mirorr_volume* fixed,
                      const mirorr_volume* moving, const double* init_params,
                      const double* init_center, double* out_params, double* out_center,
                      mirorr_reg_stats* stats) {
  hipStream_t s = ctx->stream;
  RegCache& cache = reg_cache(ctx);
  PairReg R(ctx, p, fixed, moving, init_params, init_center, stats, cache, s);
  RegRun& run = R.run;
  Workspace& w = run.w;
  Flags* flags = cache.flags;
  for (int level = 0; level < R.levels(); ++level) {
    R.begin_level(level);
    LtsState* st = R.st();
    OuterState* outer = R.outer();
    // One outer iteration up to the LTS's first chunk, queued on `s`.
    auto enqueue_iteration = [&] {
      // Without variance selection (whose scratch the passes share) the two
      // symmetric passes are independent until the LTS: the second runs on
      // the context's second stream, forked after the resampling and joined
      // before the LTS. While the search kernels are being timed the passes
      // run one after the other, so that each kernel's time is its own.
      const bool fork =
          run.symmetric && !(R.portion < 1.0) && !ctx->prof.on() && ctx->overlap_passes;
      hipStream_t s_b = fork ? run.second_stream() : s;
      // If anything throws between the fork and the join, the second pass's
      // kernels may still be in flight on the second stream: wait for them
      // before the buffers they use can be freed or rewritten.
      struct JoinOnUnwind {
        hipStream_t s = nullptr;
        ~JoinOnUnwind() {
          if (s) (void)hipStreamSynchronize(s);
        }
      } join_guard;
      if (fork) join_guard.s = s_b;
      R.enqueue_resample(s);
      if (fork) {
        MIRORR_HIP(hipEventRecord(ctx->ev_fork, s));
        MIRORR_HIP(hipStreamWaitEvent(s_b, ctx->ev_fork, 0));
      }
      R.enqueue_prep(s, s_b);
      R.match(0, s, false);
      if (run.symmetric) R.match(1, s_b, fork);
      if (fork) {
        MIRORR_HIP(hipEventRecord(ctx->ev_join, s_b));
        MIRORR_HIP(hipStreamWaitEvent(s, ctx->ev_join, 0));
        join_guard.s = nullptr;  // the first stream now orders everything after the join
      }
      // 3. LTS start and its first chunk
      run.lts_begin(st, outer, R.npts, flags, run.lts_first);
    };
    // Iteration graph of the level: every launch of an outer iteration up to
    // the LTS's first chunk takes the same arguments within a level (the
    // transforms live on the device). The graph cached for this level is
    // replayed when its key (every host value those launches read) matches;
    // otherwise the first iteration runs as it is (its searches tune their
    // tile shapes outside any capture) and the second is captured. Host-side
    // counts the captured enqueue adds (the voxel-evals) are added per replay.
    const bool graphs = ctx->iteration_graphs && !ctx->prof.on();
    IterationGraph& graph = cache.graphs[level];
    auto iteration_key = [&] {
      KeyBytes k;
      k.buffers(w);
      for (int v : {p->symmetric, p->affine, p->block_width, p->block_gap, p->nhood_width,
                    p->nhood_gap, p->variance_select, p->block_metric, p->resampling})
        k.add(v);
      k.add(p->portion_kept);
      k.add(p->lts_tolerance);
      k.add(run.storage);
      k.add(fixed->enc_zero);
      k.add(moving->enc_zero);
      k.add(cache.F[level].data.as<void>());
      k.add(cache.M[level].data.as<void>());
      k.add(R.fmask);
      k.add(R.mmask);
      for (const Geom* g : {&fixed->geom, &moving->geom, (const Geom*)&cache.F[level].geom,
                            (const Geom*)&cache.M[level].geom, (const Geom*)&R.o.gHF,
                            (const Geom*)&R.o.gHM})
        k.geom(*g);
      k.add(R.npts);
      k.add(R.portion);
      k.add(R.tiled);
      k.add((const void*)flags);
      k.add(ctx->overlap_passes);
      k.add((const void*)ctx->stream2);
      if (R.tiled)
        for (const Geom* g : {(const Geom*)&R.o.gHM, (const Geom*)&R.o.gHF})
          k.add(tile_choice(make_grid(g->dims, p->block_width, p->block_gap, p->nhood_width,
                                      p->nhood_gap),
                            run.storage));
      return std::move(k.b);
    };
    bool replay = graphs && graph.exec && graph.key == iteration_key();
    int it = 0;
    for (; it < R.max_it; ++it) {
      if (replay) {
        MIRORR_HIP(hipGraphLaunch(graph.exec, s));
        stats->voxel_evals += graph.evals;
      } else if (graphs && it >= 1) {
        std::vector<unsigned char> key = iteration_key();
        const int64_t ev0 = stats->voxel_evals;
        graph.capture(s, enqueue_iteration);
        graph.evals = stats->voxel_evals - ev0;
        graph.key = std::move(key);
        MIRORR_HIP(hipGraphLaunch(graph.exec, s));
        replay = true;
      } else {
        enqueue_iteration();
      }
      // 3. further LTS chunks while it is active; the transform update
      // (a graph replayed here may come from an earlier registration: with
      // graphs on, every first chunk is kLtsChunk, as captured)
      run.lts_complete(st, outer, R.npts, flags, run.lts_first);  // leaves flags->stop read back
      // otherwise the next LTS's first chunk is sized by this one's count
      if (!graphs) run.lts_first = RegRun::first_chunk_after(*flags);
      if (p->verbosity >= 1) R.report(it);  // as the reference's iteration report
      if (flags->stop) {
        ++it;
        break;
      }
    }
    R.end_level(it);
  }
  R.finish(out_params, out_center);
}

// Registrations of several pairs in lockstep, with one set of launches per
// step for all of them where the step is batched: per outer iteration, each
// pair's resampling, window statistics and block preparation, then one
// launch of each search kernel over every pair's tiles per pass (a flat
// work-list of uneven pairs), each pair's LTS pairs, and one launch per LTS
// phase for all pairs, with one host read of every pair's flags per chunk
// of inner iterations. Pairs with fewer pyramid levels join at the level
// where their schedule starts (the finest levels align); a pair whose stop
// rule fires leaves its level's iterations while the others go on. Every
// pair's result and statistics are those of mirorr_register on that pair
// alone. No counterpart in the reference, which registers one pair per
// process (mirorr.cxx:649).
void run_registration_batch(mirorr_ctx* ctx, const mirorr_reg_params* p, int n,
                            const mirorr_volume* const* fixed, const mirorr_volume* const* moving,
                            const double* init_params, const double* init_center,
                            double* out_params, double* out_center, mirorr_reg_stats* stats) {
  if (ctx->prof.on()) throw Error(kErrArg, "batched registrations do not time their kernels");
  if (p->block_metric != MIRORR_METRIC_NC)
    throw Error(kErrArg, "batched registrations need the NCC metric");
  {
    const int dims[3] = {64, 64, 64};
    if (!tiled_supported(make_grid(dims, p->block_width, p->block_gap, p->nhood_width,
                                   p->nhood_gap)))
      throw Error(kErrArg, "batched registrations need the tiled search (blocks of width 4, "
                           "block gap 1, nhood width 5 to 13)");
  }
  for (int i = 0; i < n; ++i)
    if (!fixed[i] || !moving[i]) throw Error(kErrArg, "null volume in the batch");
  const int storage = fixed[0]->storage;
  for (int i = 0; i < n; ++i)
    if (fixed[i]->storage != storage || moving[i]->storage != storage)
      throw Error(kErrArg, "the volumes of a batch have different storage");
  hipStream_t s = ctx->stream;
  BatchCache& bc = batch_cache(ctx);
  bc.ensure(n, &ctx->budget);
  std::vector<std::unique_ptr<PairReg>> R;
  R.reserve(n);
  for (int i = 0; i < n; ++i)
    R.emplace_back(new PairReg(ctx, p, fixed[i], moving[i],
                               init_params ? init_params + 12 * i : nullptr,
                               init_center ? init_center + 3 * i : nullptr, &stats[i],
                               *bc.regs[i], s));
  const int kind = p->affine ? xf::kAffine : xf::kRigid;
  const bool symmetric = p->symmetric != 0;
  int L = 0;
  for (const auto& r : R) L = std::max(L, r->levels());
  std::vector<int> its(n, 0);
  std::vector<char> going(n, 0);
  for (int gl = 0; gl < L; ++gl) {
    std::vector<int> A;  // pairs at a level of their schedule now
    for (int i = 0; i < n; ++i) {
      const int lv = gl - (L - R[i]->levels());
      if (lv < 0) continue;
      R[i]->begin_level(lv);
      A.push_back(i);
      its[i] = 0;
      going[i] = 1;
    }
    for (int it = 0;; ++it) {
      std::vector<int> I;  // pairs iterating now
      for (int i : A)
        if (going[i] && it < R[i]->max_it) I.push_back(i);
      if (I.empty()) break;
      const int nI = (int)I.size();
      // 1. resampling, statistics and block preparation, per pair; without
      // variance selection (whose scratch a pair's passes share) the second
      // pass runs on the context's second stream, as in mirorr_register
      bool fork = symmetric && ctx->overlap_passes;
      for (int i : I) fork = fork && !(R[i]->portion < 1.0);
      hipStream_t s_b = fork ? R[I[0]]->run.second_stream() : s;
      struct JoinOnUnwind {
        hipStream_t s = nullptr;
        ~JoinOnUnwind() {
          if (s) (void)hipStreamSynchronize(s);
        }
      } join_guard;
      for (int i : I) R[i]->enqueue_resample(s);
      if (fork) {
        join_guard.s = s_b;
        MIRORR_HIP(hipEventRecord(ctx->ev_fork, s));
        MIRORR_HIP(hipStreamWaitEvent(s_b, ctx->ev_fork, 0));
      }
      for (int i : I) R[i]->enqueue_prep(s, s_b);
      // 2. the passes, each one launch per search kernel for every pair
      for (int pass = 0; pass < (symmetric ? 2 : 1); ++pass) {
        hipStream_t sp = pass == 0 ? s : s_b;
        std::vector<SearchJob> jobs;
        for (int i : I) {
          SearchJob j = R[i]->job(pass);
          if (j.g.nblocks > 0) jobs.push_back(j);
        }
        launch_ncc_search_batch(jobs, bc.args[pass].as<void>(), bc.args[pass].bytes(), sp,
                                storage);
        for (int i : I) R[i]->pairs(pass, sp);
      }
      if (fork) {
        MIRORR_HIP(hipEventRecord(ctx->ev_join, s_b));
        MIRORR_HIP(hipStreamWaitEvent(s, ctx->ev_join, 0));
        join_guard.s = nullptr;
      }
      // 3. the LTS of every pair, one launch per phase
      std::vector<LtsItem> items(nI);
      int64_t max_npts = 1, min_npts = INT64_MAX;
      for (int k = 0; k < nI; ++k) {
        PairReg& r = *R[I[k]];
        items[k] = r.run.lts_item(r.st(), r.outer(), r.npts, &bc.flagsD.as<Flags>()[k].active);
        max_npts = std::max<int64_t>(max_npts, r.npts);
        min_npts = std::min<int64_t>(min_npts, r.npts);
      }
      // pageable source: read when the call returns
      MIRORR_HIP(hipMemcpyAsync(bc.items.as<void>(), items.data(), sizeof(LtsItem) * nI,
                                hipMemcpyHostToDevice, s));
      const LtsItem* di = bc.items.as<LtsItem>();
      lts_start_multi(di, nI, kind, symmetric ? 1 : 0, p->lts_tolerance, s);
      // chunks as in the single registration; the first is the largest of
      // the pairs' own first-chunk sizes
      int step = 1;
      for (int k = 0; k < nI; ++k) step = std::max(step, R[I[k]]->run.lts_first);
      for (int c = 0, chunk = 0; c < 101; ++chunk) {
        const int c1 = std::min(c + step, 101);
        for (int j = c; j < c1; ++j)
          lts_enqueue_iteration_multi(di, nI, max_npts, min_npts, j, s);
        lts_finish_multi(di, nI, s);
        MIRORR_HIP(hipMemcpyAsync(bc.flagsH, bc.flagsD.as<Flags>(), sizeof(Flags) * nI,
                                  hipMemcpyDeviceToHost, s));
        MIRORR_HIP(hipStreamSynchronize(s));
        bool any = false;
        for (int k = 0; k < nI; ++k) any = any || bc.flagsH[k].active != 0;
        if (!any) break;
        c = c1;
        step = chunk == 0 ? RegRun::kLtsNext : RegRun::kLtsChunk;
      }
      std::vector<char> stop(nI);
      for (int k = 0; k < nI; ++k) {
        stop[k] = bc.flagsH[k].stop != 0;
        R[I[k]]->run.lts_first = RegRun::first_chunk_after(bc.flagsH[k]);
      }
      for (int k = 0; k < nI; ++k) {
        const int i = I[k];
        its[i] = it + 1;
        if (p->verbosity >= 1) R[i]->report(it);
        if (stop[k]) going[i] = 0;
      }
    }
    for (int i : A) R[i]->end_level(its[i]);
  }
  for (int i = 0; i < n; ++i)
    R[i]->finish(out_params + 12 * i, out_center ? out_center + 3 * i : nullptr);
}

}  // namespace

extern "C" {

void mirorr_reg_default_params(mirorr_reg_params* p) {
  if (!p) return;
  // mirorr.cxx:99-281 (pyr-end 0 = all levels, the usual choice; the CLI
  // default of 1 runs only the coarsest level).
  p->symmetric = 1;
  p->affine = 0;
  p->block_width = 4;
  p->block_gap = 1;
  p->nhood_width = 7;
  p->nhood_gap = 3;
  p->iterations = 0;
  p->pyr_start = 1;
  p->pyr_end = 0;
  p->pyr_num = 1024;
  p->pyr_min_size = 32;
  p->portion_kept = 0.5;
  p->variance_select = 0;
  p->verbosity = 0;
  p->centre_init = 1;
  p->stop_tolerance = 1e-3;
  p->lts_tolerance = 1e-3;
  p->block_metric = MIRORR_METRIC_NC;
  p->resampling = MIRORR_RESAMPLE_MIDDLE;
}

mirorr_status mirorr_volume_create(mirorr_ctx* c, const float* data, const uint8_t* mask,
                                   const mirorr_geom* geom, mirorr_volume** out) {
  return mirorr_volume_create_storage(c, data, mask, geom, MIRORR_STORAGE_F32, out);
}

}  // extern "C"

namespace {

constexpr size_t kStageChunk = size_t(8) << 20;  // bytes per pinned staging chunk

// Copies `bytes` bytes to device memory `dst` through the context's two
// pinned staging chunks: fill(offset, count, chunk) writes bytes
// [offset, offset + count) of the stream into a chunk, and the chunk goes to
// the device while the host fills the other one. Queued on the context's
// stream; the caller synchronises it before the source may change.
template <class Fill>
void staged_upload(mirorr_ctx* c, unsigned char* dst, size_t bytes, Fill fill) {
  for (int i = 0; i < 2; ++i) {
    if (!c->stage[i])
      MIRORR_HIP(hipHostMalloc(&c->stage[i], kStageChunk, hipHostMallocDefault));
    if (!c->stage_ev[i])
      MIRORR_HIP(hipEventCreateWithFlags(&c->stage_ev[i], hipEventDisableTiming));
  }
  for (size_t off = 0, k = 0; off < bytes; off += kStageChunk, ++k) {
    const int b = (int)(k & 1);
    const size_t cnt = std::min(kStageChunk, bytes - off);
    // The chunk's last copy, of this upload or an earlier one, has been
    // read (an event never recorded completes at once).
    MIRORR_HIP(hipEventSynchronize(c->stage_ev[b]));
    fill(off, cnt, static_cast<unsigned char*>(c->stage[b]));
    MIRORR_HIP(hipMemcpyAsync(dst + off, c->stage[b], cnt, hipMemcpyHostToDevice, c->stream));
    MIRORR_HIP(hipEventRecord(c->stage_ev[b], c->stream));
  }
}

}  // namespace

extern "C" {

mirorr_status mirorr_volume_create_storage(mirorr_ctx* c, const float* data, const uint8_t* mask,
                                           const mirorr_geom* geom, int storage,
                                           mirorr_volume** out) {
  if (!c || !out) return MIRORR_ERR_ARG;
  *out = nullptr;
  mirorr_volume* v = nullptr;
  std::lock_guard<std::mutex> lk(c->mu);  // the staging chunks are the context's
  mirorr_status st = guarded(c, [&] {
    if (!data || !geom) throw Error(kErrArg, "null volume data or geometry");
    if (storage != MIRORR_STORAGE_F32 && storage != MIRORR_STORAGE_F16)
      throw Error(kErrArg, "unknown volume storage");
    for (int i = 0; i < 3; ++i)
      if (geom->dims[i] < 1 || !(geom->spacing[i] > 0.0)) throw Error(kErrArg, "bad geometry");
    v = new mirorr_volume();
    v->ctx = c;
    v->geom = to_geom(*geom);
    v->storage = storage;
    const size_t n = (size_t)geom->dims[0] * geom->dims[1] * geom->dims[2];
    v->data.ensure(n * storage_bytes(storage), &c->budget);
    // Uploads go through pinned staging chunks (the host fills one while
    // the other crosses the bus), twice the rate of a pageable copy.
    unsigned char* dst = v->data.as<unsigned char>();
    if (storage == MIRORR_STORAGE_F16) {
      // Encoded on the host, chunk by chunk into the staging memory: half
      // the bytes cross the bus.
      const HalfEncoding e = half_encoding(data, n);
      v->enc_offset = e.offset;
      v->enc_scale = e.scale;
      v->enc_zero = e.zero();
      staged_upload(c, dst, n * sizeof(half_t), [&](size_t off, size_t cnt, unsigned char* h) {
        encode_half(data + off / sizeof(half_t), cnt / sizeof(half_t), e,
                    reinterpret_cast<half_t*>(h));
      });
    } else {
      const unsigned char* src = reinterpret_cast<const unsigned char*>(data);
      staged_upload(c, dst, n * sizeof(float), [&](size_t off, size_t cnt, unsigned char* h) {
        std::memcpy(h, src + off, cnt);
      });
    }
    if (mask) {
      v->mask.ensure(n, &c->budget);
      staged_upload(c, v->mask.as<unsigned char>(), n,
                    [&](size_t off, size_t cnt, unsigned char* h) {
                      std::memcpy(h, mask + off, cnt);
                    });
      v->has_mask = true;
    }
    MIRORR_HIP(hipStreamSynchronize(c->stream));
  });
  if (st != MIRORR_OK) {
    delete v;
    return st;
  }
  *out = v;
  return MIRORR_OK;
}

mirorr_status mirorr_volume_info(const mirorr_volume* v, int* storage, double* offset,
                                 double* scale) {
  if (!v) return MIRORR_ERR_ARG;
  if (storage) *storage = v->storage;
  if (offset) *offset = v->enc_offset;
  if (scale) *scale = v->enc_scale;
  return MIRORR_OK;
}

void mirorr_volume_destroy(mirorr_volume* v) {
  if (!v) return;
  (void)hipSetDevice(v->ctx->device);
  (void)hipStreamSynchronize(v->ctx->stream);
  delete v;
}

mirorr_status mirorr_register(mirorr_ctx* c, const mirorr_reg_params* p,
                              const mirorr_volume* fixed, const mirorr_volume* moving,
                              const double* init_params, const double* init_center,
                              double* out_params, double* out_center, mirorr_reg_stats* stats) {
  if (!c || !p || !fixed || !moving || !out_params) return MIRORR_ERR_ARG;
  mirorr_reg_stats local{};
  std::lock_guard<std::mutex> lk(c->mu);
  return guarded(c, [&] {
    check_reg_params(p);
    run_registration(c, p, fixed, moving, init_params, init_center, out_params, out_center,
                     stats ? stats : &local);
  });
}

mirorr_status mirorr_register_batch(mirorr_ctx* c, const mirorr_reg_params* p, int n,
                                    const mirorr_volume* const* fixed,
                                    const mirorr_volume* const* moving, const double* init_params,
                                    const double* init_center, double* out_params,
                                    double* out_center, mirorr_reg_stats* stats) {
  if (!c || !p || n < 0 || (n > 0 && (!fixed || !moving || !out_params))) return MIRORR_ERR_ARG;
  if (n == 0) return MIRORR_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  return guarded(c, [&] {
    check_reg_params(p);
    if (n > kMaxBatchPairs) throw Error(kErrArg, "more than 256 pairs in a batch");
    std::vector<mirorr_reg_stats> local;
    if (!stats) {
      local.resize(n);
      stats = local.data();
    }
    run_registration_batch(c, p, n, fixed, moving, init_params, init_center, out_params,
                           out_center, stats);
  });
}

mirorr_status mirorr_lts_fit(mirorr_ctx* c, int symmetric, int affine, const double* x,
                             const double* y, size_t n, double* params, double* center,
                             double* fwd34, double* bwd34, int* iterations) {
  if (!c || !x || !y || !params) return MIRORR_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  return guarded(c, [&] {
    hipStream_t s = c->stream;
    Budget* bud = &c->budget;
    const int kind = affine ? xf::kAffine : xf::kRigid;
    mirorr_reg_params rp;
    mirorr_reg_default_params(&rp);
    mirorr_reg_stats stats{};
    RegCache& cache = reg_cache(c);
    RegRun run{c, &rp, cache.w, &stats, kind, symmetric != 0, s, bud};
    Workspace& w = run.w;
    const int64_t np = (int64_t)n;
    w.X.ensure(sizeof(double) * 3 * std::max<int64_t>(np, 1), bud);
    w.Y.ensure(sizeof(double) * 3 * std::max<int64_t>(np, 1), bud);
    w.kept.ensure(std::max<int64_t>(np, 1), bud);
    w.r.ensure(sizeof(double) * std::max<int64_t>(np, 1), bud);
    w.cand.ensure(sizeof(unsigned long long) * std::max<int64_t>(np, 1), bud);
    w.part.ensure(sizeof(double) * 32 * std::max<int64_t>(lts_parts(np), 1), bud);
    w.hist1.ensure(sizeof(unsigned int) * 4096, bud);
    w.hist2.ensure(sizeof(unsigned int) * 4096, bud);
    w.lts.ensure(sizeof(LtsState), bud);
    w.outer.ensure(sizeof(OuterState), bud);
    std::vector<uint8_t> ones(std::max<int64_t>(np, 1), 1);
    if (np > 0) {
      MIRORR_HIP(hipMemcpyAsync(w.X.as<double>(), x, sizeof(double) * 3 * np,
                                hipMemcpyHostToDevice, s));
      MIRORR_HIP(hipMemcpyAsync(w.Y.as<double>(), y, sizeof(double) * 3 * np,
                                hipMemcpyHostToDevice, s));
    }
    MIRORR_HIP(hipMemcpyAsync(w.kept.as<uint8_t>(), ones.data(), ones.size(),
                              hipMemcpyHostToDevice, s));
    OuterState o{};
    const Geom unit{{1, 1, 1}, {1, 1, 1}, {0, 0, 0}, {1, 0, 0, 0, 1, 0, 0, 0, 1}};
    o.gF = o.gM = o.gFfull = o.gMfull = o.gHF = o.gHM = unit;
    o.fixT = o.movT = xf::identity(xf::kRigid);
    o.t12F = o.t12B = xf::identity(kind);
    o.T = xf::identity(kind);
    if (!symmetric) {
      if (center) xf::set_center(o.T, center);
      xf::set_params(o.T, params);
    }
    xf::get_params(o.T, o.lastp);
    LtsState* st = w.lts.as<LtsState>();
    OuterState* outer = w.outer.as<OuterState>();
    MIRORR_HIP(hipMemcpyAsync(outer, &o, sizeof(o), hipMemcpyHostToDevice, s));
    const unsigned long long nk = (unsigned long long)np;
    MIRORR_HIP(hipMemcpyAsync(&st->n_kept, &nk, sizeof(nk), hipMemcpyHostToDevice, s));
    run.run_lts(st, outer, std::max<int64_t>(np, 1), cache.flags);
    MIRORR_HIP(hipMemcpyAsync(&o, outer, sizeof(o), hipMemcpyDeviceToHost, s));
    MIRORR_HIP(hipStreamSynchronize(s));
    xf::get_params(o.T, params);
    if (center)
      for (int k = 0; k < 3; ++k) center[k] = o.T.c[k];
    for (int r = 0; r < 3; ++r) {
      for (int k = 0; k < 3; ++k) {
        if (fwd34) fwd34[4 * r + k] = o.t12F.M.a[r][k];
        if (bwd34) bwd34[4 * r + k] = o.t12B.M.a[r][k];
      }
      if (fwd34) fwd34[4 * r + 3] = o.t12F.off[r];
      if (bwd34) bwd34[4 * r + 3] = o.t12B.off[r];
    }
    if (iterations) *iterations = o.inner_total;
  });
}

mirorr_status mirorr_resample(mirorr_ctx* c, const float* data, const mirorr_geom* gin,
                              const double matrix[9], const double offset[3],
                              const mirorr_geom* gout, float* out, uint8_t* out_mask) {
  if (!c || !data || !gin || !matrix || !offset || !gout || !out) return MIRORR_ERR_ARG;
  return guarded(c, [&] {
    hipStream_t s = c->stream;
    const Geom gi = to_geom(*gin), go = to_geom(*gout);
    DeviceBuffer din, dout, dmask, dmap;
    din.ensure(sizeof(float) * nvox(gi), &c->budget);
    dout.ensure(sizeof(float) * nvox(go), &c->budget);
    if (out_mask) dmask.ensure(nvox(go), &c->budget);
    xf::M3 M;
    for (int i = 0; i < 9; ++i) M.a[i / 3][i % 3] = matrix[i];
    const IndexMap map = make_index_map(M, offset, gi, go);
    dmap.ensure(sizeof(IndexMap), &c->budget);
    MIRORR_HIP(hipMemcpyAsync(din.as<float>(), data, sizeof(float) * nvox(gi),
                              hipMemcpyHostToDevice, s));
    MIRORR_HIP(hipMemcpyAsync(dmap.as<IndexMap>(), &map, sizeof(map), hipMemcpyHostToDevice, s));
    launch_resample(din.as<float>(), gi.dims, dmap.as<IndexMap>(), nullptr, gi.dims,
                    dmap.as<IndexMap>(), out_mask ? kMaskInside : kMaskNone, go.dims,
                    dout.as<float>(), out_mask ? dmask.as<uint8_t>() : nullptr, s);
    MIRORR_HIP(hipMemcpyAsync(out, dout.as<float>(), sizeof(float) * nvox(go),
                              hipMemcpyDeviceToHost, s));
    if (out_mask)
      MIRORR_HIP(hipMemcpyAsync(out_mask, dmask.as<uint8_t>(), nvox(go), hipMemcpyDeviceToHost, s));
    MIRORR_HIP(hipStreamSynchronize(s));
  });
}

mirorr_status mirorr_pyramid_level(mirorr_ctx* c, const float* data, const mirorr_geom* geom,
                                   const int* schedule, int rows, int level, float* out,
                                   mirorr_geom* out_geom) {
  if (!c || !data || !geom || !schedule || rows < 1 || level < 0 || level >= rows)
    return MIRORR_ERR_ARG;
  return guarded(c, [&] {
    hipStream_t s = c->stream;
    const Geom g = to_geom(*geom);
    DeviceBuffer din, tmp, map;
    din.ensure(sizeof(float) * nvox(g), &c->budget);
    MIRORR_HIP(hipMemcpyAsync(din.as<float>(), data, sizeof(float) * nvox(g),
                              hipMemcpyHostToDevice, s));
    const std::vector<int> sched(schedule, schedule + 3 * rows);
    std::vector<Level> lv;
    build_pyramid(din.as<float>(), g, sched, rows, lv, tmp, map, &c->budget, s);
    const Geom& lg = lv[level].geom;
    MIRORR_HIP(hipMemcpyAsync(out, lv[level].data.as<float>(), sizeof(float) * nvox(lg),
                              hipMemcpyDeviceToHost, s));
    MIRORR_HIP(hipStreamSynchronize(s));
    if (out_geom) {
      for (int i = 0; i < 3; ++i) {
        out_geom->dims[i] = lg.dims[i];
        out_geom->spacing[i] = lg.spacing[i];
        out_geom->origin[i] = lg.origin[i];
      }
      for (int i = 0; i < 9; ++i) out_geom->direction[i] = lg.dir[i];
    }
  });
}

}  // extern "C"
