// C-ABI: context and block matcher (replaces BlockMatcherGPU,
// itkBlockMatcherGPU.txx:186-522,597-830, behind the AbstractBlockMatcher
// contract of itkAbstractBlockMatcher.h:41-313).
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <vector>

#include "bm_kernels.hpp"
#include "context.hpp"
#include "storage.hpp"

using namespace mirorr;

struct VolumeSlot {
  const void* data = nullptr;     // float, or half_t when storage is kStoreF16
  int storage = kStoreF32;
  const uint8_t* mask = nullptr;
  mirorr_geom geom{};
  uint64_t version = 0;  // bumps whenever data/mask/geometry change
  DeviceBuffer own_data, own_mask;
  bool set = false;
};

struct mirorr_bm {
  mirorr_ctx* ctx = nullptr;
  mirorr_bm_params p{};
  VolumeSlot slot[2];  // [MIRORR_ROLE_BASE], [MIRORR_ROLE_SEARCH]
  // Derived data cached per slot and volume version.
  DeviceBuffer stats[2];
  uint64_t stats_ver[2] = {0, 0};
  int stats_bw[2] = {0, 0};
  DeviceBuffer prep_xt[2], prep_sxx[2], prep_sum[2], prep_flags[2];
  uint64_t prep_ver[2] = {0, 0};
  int prep_key[2][2] = {{0, 0}, {0, 0}};  // (bw, ng) the prep was made for
  double prep_portion[2] = {1.0, 1.0};    // variance selection applied to the flags
  bool prep_selected[2] = {false, false};
  DeviceBuffer vsel_st, vsel_hist, vsel_evals;
  // Results of the last search.
  DeviceBuffer best, second, match, kept, base_mm, match_mm, index;
  DeviceBuffer queue, qbits, qcount, chunk_counts, chunk_offsets;
  DeviceBuffer rlist, rcount;  // blocks listed for the exact re-scoring
  bool results_valid = false;
  uint64_t results_key[2] = {0, 0};
  BMGrid grid{};
};

static std::atomic<uint64_t> g_version{1};

static void check_params(const mirorr_bm_params* p) {
  if (!p) throw Error(kErrArg, "null parameters");
  if (p->block_width < 1 || p->block_width > 16) throw Error(kErrArg, "block_width out of range");
  if (p->block_gap < 1) throw Error(kErrArg, "block_gap must be >= 1");
  if (p->nhood_width < 1 || p->nhood_width > 64) throw Error(kErrArg, "nhood_width out of range");
  if (p->nhood_gap < 1) throw Error(kErrArg, "nhood_gap must be >= 1");
  // SetPortionMatchesKept (itkAbstractBlockMatcher.txx:83-99): 0 means 0.5.
  if (p->portion_kept < 0.0 || p->portion_kept > 1.0)
    throw Error(kErrArg, "portion_kept should be between 0 and 1.0");
  if (p->block_metric < MIRORR_METRIC_NC || p->block_metric > MIRORR_METRIC_CR)
    throw Error(kErrArg, "unknown block metric");
  if (p->storage != MIRORR_STORAGE_F32 && p->storage != MIRORR_STORAGE_F16)
    throw Error(kErrArg, "unknown image storage");
}

static BMGrid current_grid(const mirorr_bm* bm) {
  const mirorr_geom& g = bm->slot[MIRORR_ROLE_BASE].geom;
  return make_grid(g.dims, bm->p.block_width, bm->p.block_gap, bm->p.nhood_width,
                   bm->p.nhood_gap);
}

static void require_inputs(const mirorr_bm* bm) {
  if (!bm->slot[0].set || !bm->slot[1].set) throw Error(kErrState, "base and search images must be set");
  const mirorr_geom& a = bm->slot[0].geom;
  const mirorr_geom& b = bm->slot[1].geom;
  for (int i = 0; i < 3; ++i)
    if (a.dims[i] != b.dims[i]) throw Error(kErrArg, "base and search images must share one grid");
}

// Runs the search (stats, prep, NCC kernels) when inputs or parameters
// changed since the last one. Enqueued on the context stream.
static void run_search(mirorr_bm* bm) {
  require_inputs(bm);
  const BMGrid g = current_grid(bm);
  if (bm->slot[0].storage != bm->slot[1].storage)
    throw Error(kErrArg, "base and search images have different storage");
  const int storage = bm->slot[0].storage;
  const uint64_t key0 = bm->slot[0].version ^ ((uint64_t)g.bw << 40) ^ ((uint64_t)g.nw << 48) ^
                        ((uint64_t)g.ng << 56) ^ ((uint64_t)g.bg << 32) ^
                        ((uint64_t)bm->p.block_metric << 24) ^ ((uint64_t)bm->p.exact_rescore << 20);
  const uint64_t key1 = bm->slot[1].version;
  if (bm->results_valid && bm->results_key[0] == key0 && bm->results_key[1] == key1 &&
      std::memcmp(&bm->grid, &g, sizeof(g)) == 0)
    return;
  hipStream_t s = bm->ctx->stream;
  Budget* bud = &bm->ctx->budget;
  const VolumeSlot& B = bm->slot[MIRORR_ROLE_BASE];
  const VolumeSlot& S = bm->slot[MIRORR_ROLE_SEARCH];
  bm->grid = g;
  const int64_t nb = g.nblocks;
  if (nb <= 0) throw Error(kErrAlgo, "No block matches occurred: the image is smaller than a block");
  const int64_t nvox = (int64_t)g.dims[0] * g.dims[1] * g.dims[2];
  bm->best.ensure(sizeof(float) * nb, bud);
  bm->second.ensure(sizeof(float) * nb, bud);
  bm->match.ensure(sizeof(int32_t) * 3 * nb, bud);
  BMResult res{bm->best.as<float>(), bm->second.as<float>(), bm->match.as<int32_t>()};
  const int si = MIRORR_ROLE_SEARCH, bi = MIRORR_ROLE_BASE;
  const int metric = bm->p.block_metric;
  const bool tiled = metric == MIRORR_METRIC_NC && tiled_supported(g);
  const double portion = bm->p.variance_select ? bm->p.portion_kept : 1.0;
  // Base-image preparation (tiled search, or sums for the variance selection).
  if (tiled || portion < 1.0) {
    if (bm->prep_ver[bi] != B.version || bm->prep_key[bi][0] != g.bw ||
        bm->prep_key[bi][1] != g.ng || bm->prep_portion[bi] != portion) {
      const int64_t n3 = (int64_t)g.bw * g.bw * g.bw;
      bm->prep_xt[bi].ensure(sizeof(float) * nb * n3, bud);
      bm->prep_sxx[bi].ensure(sizeof(float) * nb, bud);
      bm->prep_sum[bi].ensure(sizeof(float) * nb, bud);
      bm->prep_flags[bi].ensure(sizeof(uint32_t) * nb, bud);
      const BlockPrep bp{bm->prep_xt[bi].as<float>(), bm->prep_sxx[bi].as<float>(),
                         bm->prep_sum[bi].as<float>(), bm->prep_flags[bi].as<uint32_t>()};
      launch_block_prep(B.data, B.mask, g, bp, s, storage);
      bm->prep_selected[bi] = false;
      if (portion < 1.0) {
        bm->vsel_st.ensure(sizeof(VarSelState), bud);
        bm->vsel_hist.ensure(sizeof(unsigned int) * kVarSelHistBins, bud);
        bm->vsel_evals.ensure(sizeof(unsigned long long), bud);
        MIRORR_HIP(hipMemsetAsync(bm->vsel_evals.as<void>(), 0, sizeof(unsigned long long), s));
        VarSelScratch vs{bm->vsel_st.as<VarSelState>(), bm->vsel_hist.as<unsigned int>(),
                         bm->vsel_evals.as<unsigned long long>()};
        bm->prep_selected[bi] = launch_variance_select(g, bp, portion, vs, s);
      }
      bm->prep_ver[bi] = B.version;
      bm->prep_key[bi][0] = g.bw;
      bm->prep_key[bi][1] = g.ng;
      bm->prep_portion[bi] = portion;
    }
  }
  const BlockPrep bp{bm->prep_xt[bi].as<float>(), bm->prep_sxx[bi].as<float>(),
                     bm->prep_sum[bi].as<float>(), bm->prep_flags[bi].as<uint32_t>()};
  if (tiled) {
    // Search-image statistics (slot SEARCH).
    if (bm->stats_ver[si] != S.version || bm->stats_bw[si] != g.bw) {
      bm->stats[si].ensure(sizeof(float2) * nvox, bud);
      launch_search_stats(S.data, S.mask, g, SearchStats{bm->stats[si].as<float2>()}, s, storage);
      bm->stats_ver[si] = S.version;
      bm->stats_bw[si] = g.bw;
    }
    const int64_t nt = tile_count(g);
    bm->queue.ensure(sizeof(int32_t) * nt, bud);
    bm->qbits.ensure(sizeof(unsigned long long) * nt, bud);
    bm->qcount.ensure(sizeof(int32_t), bud);
    BMScratch scr{bm->queue.as<int32_t>(), bm->qcount.as<int32_t>(),
                  bm->qbits.as<unsigned long long>(), nt, &bm->ctx->prof};
    if (bm->p.exact_rescore) {
      bm->rlist.ensure(sizeof(int32_t) * nb, bud);
      bm->rcount.ensure(sizeof(int32_t), bud);
      scr.rlist = bm->rlist.as<int32_t>();
      scr.rcount = bm->rcount.as<int32_t>();
      scr.rescore_ties = 1;  // raw records of exact ties are the reference's too
    }
    launch_ncc_search(B.data, B.mask, S.data, S.mask, g, bp,
                      SearchStats{bm->stats[si].as<float2>()}, scr, res, s, storage);
  } else {
    if (storage != kStoreF32)
      throw Error(kErrArg, "half-precision images need the NCC metric and the tiled search "
                           "(blocks of width 4, block gap 1, nhood width 5 to 13)");
    launch_metric_search(metric, static_cast<const float*>(B.data), B.mask,
                         static_cast<const float*>(S.data), S.mask, g, res, s,
                         bm->prep_selected[bi] ? bp.flags : nullptr);
  }
  bm->results_valid = true;
  bm->results_key[0] = key0;
  bm->results_key[1] = key1;
}

extern "C" {

mirorr_status mirorr_ctx_create(int device, size_t hbm_budget, mirorr_ctx** out) {
  if (!out) return MIRORR_ERR_ARG;
  *out = nullptr;
  mirorr_ctx* c = new (std::nothrow) mirorr_ctx();
  if (!c) return MIRORR_ERR_NOMEM;
  c->device = device;
  c->budget.limit = hbm_budget;
  mirorr_status st = guarded(c, [&] {
    int n = 0;
    MIRORR_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) throw Error(kErrPlatform, "no such HIP device");
    MIRORR_HIP(hipSetDevice(device));
    MIRORR_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    const char* e = std::getenv("MIRORR_ITERATION_GRAPH");
    if (e && e[0]) c->iteration_graphs = e[0] != '0';
  });
  if (st != MIRORR_OK) {
    delete c;
    return st;
  }
  *out = c;
  return MIRORR_OK;
}

void mirorr_ctx_destroy(mirorr_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamDestroy(c->stream);
  }
  if (c->stream2) {
    (void)hipStreamSynchronize(c->stream2);
    (void)hipStreamDestroy(c->stream2);
  }
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  for (int i = 0; i < 2; ++i) {
    if (c->stage[i]) (void)hipHostFree(c->stage[i]);
    if (c->stage_ev[i]) (void)hipEventDestroy(c->stage_ev[i]);
  }
  delete c;
}

const char* mirorr_last_error(const mirorr_ctx* c) {
  return c ? c->last_error.c_str() : "null context";
}

mirorr_status mirorr_ctx_synchronize(mirorr_ctx* c) {
  if (!c) return MIRORR_ERR_ARG;
  return guarded(c, [&] { MIRORR_HIP(hipStreamSynchronize(c->stream)); });
}

void* mirorr_ctx_stream(mirorr_ctx* c) { return c ? (void*)c->stream : nullptr; }

mirorr_status mirorr_ctx_set_profiling(mirorr_ctx* c, int on) {
  if (!c) return MIRORR_ERR_ARG;
  return guarded(c, [&] {
    c->prof.collect(c->stream);
    c->prof.enable((on & (kProfFast | kProfMasked)) ? on & (kProfFast | kProfMasked | kProfFinest) : 0,
                   &c->budget);
  });
}

mirorr_status mirorr_ctx_set_iteration_graphs(mirorr_ctx* c, int on) {
  if (!c) return MIRORR_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  c->iteration_graphs = on != 0;
  return MIRORR_OK;
}

int mirorr_ctx_iteration_graphs(const mirorr_ctx* c) { return c && c->iteration_graphs ? 1 : 0; }

mirorr_status mirorr_ctx_profile_get(mirorr_ctx* c, const char* kernel, int64_t* launches,
                                     double* total_ms, double* work) {
  if (!c || !kernel) return MIRORR_ERR_ARG;
  return guarded(c, [&] {
    c->prof.collect(c->stream);
    const auto& a = c->prof.aggregates();
    const auto it = a.find(kernel);
    const ProfAgg z{};
    const ProfAgg& g = it == a.end() ? z : it->second;
    if (launches) *launches = g.launches;
    if (total_ms) *total_ms = g.ms;
    if (work) *work = g.units;
  });
}

mirorr_status mirorr_ctx_profile_reset(mirorr_ctx* c) {
  if (!c) return MIRORR_ERR_ARG;
  return guarded(c, [&] {
    c->prof.collect(c->stream);
    c->prof.reset();
  });
}

void mirorr_bm_default_params(mirorr_bm_params* p) {
  if (!p) return;
  // itkAbstractBlockMatcher.txx:24-43
  p->block_width = 4;
  p->block_gap = 1;
  p->nhood_width = 5;
  p->nhood_gap = 2;
  p->portion_kept = 0.5;
  p->variance_select = 0;
  p->verbosity = 0;
  p->block_metric = MIRORR_METRIC_NC;
  p->exact_rescore = 1;
  p->storage = MIRORR_STORAGE_F32;
}

mirorr_status mirorr_bm_create(mirorr_ctx* c, const mirorr_bm_params* p, mirorr_bm** out) {
  if (!c || !out) return MIRORR_ERR_ARG;
  *out = nullptr;
  mirorr_bm* bm = nullptr;
  mirorr_status st = guarded(c, [&] {
    mirorr_bm_params q;
    mirorr_bm_default_params(&q);
    if (p) q = *p;
    if (q.portion_kept == 0.0) q.portion_kept = 0.5;
    check_params(&q);
    bm = new mirorr_bm();
    bm->ctx = c;
    bm->p = q;
  });
  if (st == MIRORR_OK) *out = bm;
  return st;
}

void mirorr_bm_destroy(mirorr_bm* bm) {
  if (!bm) return;
  (void)hipSetDevice(bm->ctx->device);
  (void)hipStreamSynchronize(bm->ctx->stream);
  delete bm;
}

mirorr_status mirorr_bm_set_params(mirorr_bm* bm, const mirorr_bm_params* p) {
  if (!bm) return MIRORR_ERR_ARG;
  return guarded(bm->ctx, [&] {
    mirorr_bm_params q = *p;
    if (q.portion_kept == 0.0) q.portion_kept = 0.5;
    check_params(&q);
    bm->p = q;
    bm->results_valid = false;
  });
}

static void check_geom(const mirorr_geom* g) {
  if (!g) throw Error(kErrArg, "null geometry");
  for (int i = 0; i < 3; ++i) {
    if (g->dims[i] < 1) throw Error(kErrArg, "image dimensions must be positive");
    if (!(g->spacing[i] > 0.0)) throw Error(kErrArg, "spacing must be positive");
  }
}

mirorr_status mirorr_bm_set_volume(mirorr_bm* bm, int role, const float* data,
                                   const uint8_t* mask, const mirorr_geom* geom) {
  if (!bm) return MIRORR_ERR_ARG;
  return guarded(bm->ctx, [&] {
    if (role != MIRORR_ROLE_BASE && role != MIRORR_ROLE_SEARCH) throw Error(kErrArg, "bad role");
    if (!data) throw Error(kErrArg, "null image data");
    check_geom(geom);
    VolumeSlot& s = bm->slot[role];
    const size_t n = (size_t)geom->dims[0] * geom->dims[1] * geom->dims[2];
    s.storage = bm->p.storage;
    s.own_data.ensure(n * storage_bytes(s.storage), &bm->ctx->budget);
    std::vector<half_t> h;
    if (s.storage == kStoreF16) {
      h.resize(n);
      encode_half(data, n, half_encoding(data, n), h.data());
      MIRORR_HIP(hipMemcpyAsync(s.own_data.as<void>(), h.data(), n * sizeof(half_t),
                                hipMemcpyHostToDevice, bm->ctx->stream));
    } else {
      MIRORR_HIP(hipMemcpyAsync(s.own_data.as<float>(), data, n * sizeof(float),
                                hipMemcpyHostToDevice, bm->ctx->stream));
    }
    s.data = s.own_data.as<void>();
    if (mask) {
      s.own_mask.ensure(n, &bm->ctx->budget);
      MIRORR_HIP(hipMemcpyAsync(s.own_mask.as<uint8_t>(), mask, n, hipMemcpyHostToDevice,
                                bm->ctx->stream));
      s.mask = s.own_mask.as<uint8_t>();
    } else {
      s.mask = nullptr;
    }
    // Host arrays may be reused by the caller right after return.
    MIRORR_HIP(hipStreamSynchronize(bm->ctx->stream));
    s.geom = *geom;
    s.version = ++g_version;
    s.set = true;
  });
}

mirorr_status mirorr_bm_set_volume_device(mirorr_bm* bm, int role, const float* d_data,
                                          const uint8_t* d_mask, const mirorr_geom* geom) {
  if (!bm) return MIRORR_ERR_ARG;
  return guarded(bm->ctx, [&] {
    if (role != MIRORR_ROLE_BASE && role != MIRORR_ROLE_SEARCH) throw Error(kErrArg, "bad role");
    if (!d_data) throw Error(kErrArg, "null image data");
    check_geom(geom);
    VolumeSlot& s = bm->slot[role];
    s.own_data.reset();
    s.own_mask.reset();
    s.storage = kStoreF32;
    s.data = d_data;
    s.mask = d_mask;
    s.geom = *geom;
    s.version = ++g_version;
    s.set = true;
  });
}

int mirorr_tile_autotune_report(char* buf, size_t cap) {
  const std::string r = tile_autotune_report();
  if (buf && cap > 0) std::snprintf(buf, cap, "%s", r.c_str());
  return (int)r.size();
}

mirorr_status mirorr_tile_force(int nhood_width, int index) {
  return tile_force(nhood_width, index) ? MIRORR_OK : MIRORR_ERR_ARG;
}

mirorr_status mirorr_bm_swap(mirorr_bm* bm) {
  if (!bm) return MIRORR_ERR_ARG;
  return guarded(bm->ctx, [&] {
    std::swap(bm->slot[0], bm->slot[1]);
    std::swap(bm->stats[0], bm->stats[1]);
    std::swap(bm->stats_ver[0], bm->stats_ver[1]);
    std::swap(bm->stats_bw[0], bm->stats_bw[1]);
    std::swap(bm->prep_xt[0], bm->prep_xt[1]);
    std::swap(bm->prep_sxx[0], bm->prep_sxx[1]);
    std::swap(bm->prep_sum[0], bm->prep_sum[1]);
    std::swap(bm->prep_flags[0], bm->prep_flags[1]);
    std::swap(bm->prep_ver[0], bm->prep_ver[1]);
    std::swap(bm->prep_key[0], bm->prep_key[1]);
    std::swap(bm->prep_portion[0], bm->prep_portion[1]);
    std::swap(bm->prep_selected[0], bm->prep_selected[1]);
    bm->results_valid = false;
  });
}

mirorr_status mirorr_bm_match_async(mirorr_bm* bm) {
  if (!bm) return MIRORR_ERR_ARG;
  return guarded(bm->ctx, [&] {
    // Time the whole search: window statistics and block preparation too.
    bm->results_valid = false;
    bm->stats_ver[0] = bm->stats_ver[1] = 0;
    bm->prep_ver[0] = bm->prep_ver[1] = 0;
    run_search(bm);
  });
}

mirorr_status mirorr_bm_block_count(mirorr_bm* bm, size_t* nblocks, int grid[3]) {
  if (!bm) return MIRORR_ERR_ARG;
  return guarded(bm->ctx, [&] {
    if (!bm->slot[0].set) throw Error(kErrState, "base image not set");
    const BMGrid g = current_grid(bm);
    if (nblocks) *nblocks = (size_t)g.nblocks;
    if (grid)
      for (int i = 0; i < 3; ++i) grid[i] = g.nb[i];
  });
}

mirorr_status mirorr_bm_voxel_evals(mirorr_bm* bm, int64_t* evals) {
  if (!bm || !evals) return MIRORR_ERR_ARG;
  return guarded(bm->ctx, [&] {
    if (!bm->slot[0].set) throw Error(kErrState, "base image not set");
    const BMGrid g = current_grid(bm);
    if (bm->p.variance_select && bm->p.portion_kept < 1.0) {
      // Only the selected blocks are searched: counted on the device.
      run_search(bm);
      if (bm->prep_selected[MIRORR_ROLE_BASE]) {
        unsigned long long e = 0;
        MIRORR_HIP(hipMemcpyAsync(&e, bm->vsel_evals.as<void>(), sizeof(e),
                                  hipMemcpyDeviceToHost, bm->ctx->stream));
        MIRORR_HIP(hipStreamSynchronize(bm->ctx->stream));
        *evals = (int64_t)e;
        return;
      }
    }
    int64_t per[3] = {0, 0, 0};
    for (int a = 0; a < 3; ++a)
      for (int i = 0; i < g.nb[a]; ++i) {
        int f, c;
        cand_range(g.start[a] + i * g.ng, g.dims[a], g.bw, g.bg, g.nw, &f, &c);
        per[a] += c;
      }
    *evals = per[0] * per[1] * per[2] * (int64_t)g.bw * g.bw * g.bw;
  });
}

mirorr_status mirorr_bm_get_raw(mirorr_bm* bm, int32_t* block_idx, int32_t* match_idx,
                                float* best, float* second, size_t cap, size_t* nblocks) {
  if (!bm) return MIRORR_ERR_ARG;
  return guarded(bm->ctx, [&] {
    run_search(bm);
    const BMGrid& g = bm->grid;
    const size_t n = (size_t)g.nblocks;
    if (nblocks) *nblocks = n;
    if (cap < n) throw Error(kErrArg, "capacity too small for the block grid");
    hipStream_t s = bm->ctx->stream;
    if (match_idx)
      MIRORR_HIP(hipMemcpyAsync(match_idx, bm->match.as<int32_t>(), 3 * n * sizeof(int32_t),
                                hipMemcpyDeviceToHost, s));
    if (best)
      MIRORR_HIP(hipMemcpyAsync(best, bm->best.as<float>(), n * sizeof(float),
                                hipMemcpyDeviceToHost, s));
    if (second)
      MIRORR_HIP(hipMemcpyAsync(second, bm->second.as<float>(), n * sizeof(float),
                                hipMemcpyDeviceToHost, s));
    MIRORR_HIP(hipStreamSynchronize(s));
    if (block_idx) {
      size_t i = 0;
      for (int z = 0; z < g.nb[2]; ++z)
        for (int y = 0; y < g.nb[1]; ++y)
          for (int x = 0; x < g.nb[0]; ++x, ++i) {
            block_idx[3 * i + 0] = g.start[0] + g.ng * x;
            block_idx[3 * i + 1] = g.start[1] + g.ng * y;
            block_idx[3 * i + 2] = g.start[2] + g.ng * z;
          }
    }
  });
}

mirorr_status mirorr_bm_get_offsets(mirorr_bm* bm, double* base_mm, double* match_mm,
                                    size_t cap, size_t* n, float* best, float* second,
                                    int32_t* block_index) {
  if (!bm) return MIRORR_ERR_ARG;
  return guarded(bm->ctx, [&] {
    run_search(bm);
    const BMGrid& g = bm->grid;
    const int64_t nb = g.nblocks;
    hipStream_t s = bm->ctx->stream;
    Budget* bud = &bm->ctx->budget;
    bm->kept.ensure(nb, bud);
    bm->base_mm.ensure(sizeof(double) * 3 * nb, bud);
    bm->match_mm.ensure(sizeof(double) * 3 * nb, bud);
    bm->index.ensure(sizeof(int32_t) * nb, bud);
    const int64_t nc = compact_chunks(nb);
    bm->chunk_counts.ensure(sizeof(int32_t) * nc, bud);
    bm->chunk_offsets.ensure(sizeof(int32_t) * (nc + 1), bud);
    BMResult res{bm->best.as<float>(), bm->second.as<float>(), bm->match.as<int32_t>()};
    const Geom geom = to_geom(bm->slot[MIRORR_ROLE_BASE].geom);
    launch_postprocess(g, geom, res, bm->kept.as<uint8_t>(), bm->base_mm.as<double>(),
                       bm->match_mm.as<double>(), s);
    Compactor cp{bm->chunk_counts.as<int32_t>(), bm->chunk_offsets.as<int32_t>(), nc};
    launch_compact_index(bm->kept.as<uint8_t>(), nb, cp, bm->index.as<int32_t>(), s);
    int32_t nk = 0;
    MIRORR_HIP(hipMemcpyAsync(&nk, bm->chunk_offsets.as<int32_t>() + nc, sizeof(int32_t),
                              hipMemcpyDeviceToHost, s));
    MIRORR_HIP(hipStreamSynchronize(s));
    if (n) *n = (size_t)nk;
    if (nk == 0)
      throw Error(kErrAlgo,
                  "WARNING: No block matches occurred. Images probably have zero overlap. "
                  "Try supplying a better initial transform.");
    if ((size_t)nk > cap) throw Error(kErrArg, "capacity too small for the kept points");
    // Gather on the host from full per-block copies (sizes are small).
    std::vector<int32_t> idx(nk);
    std::vector<double> bmm(3 * nb), mmm(3 * nb);
    std::vector<float> vb, vs;
    MIRORR_HIP(hipMemcpyAsync(idx.data(), bm->index.as<int32_t>(), nk * sizeof(int32_t),
                              hipMemcpyDeviceToHost, s));
    MIRORR_HIP(hipMemcpyAsync(bmm.data(), bm->base_mm.as<double>(), 3 * nb * sizeof(double),
                              hipMemcpyDeviceToHost, s));
    MIRORR_HIP(hipMemcpyAsync(mmm.data(), bm->match_mm.as<double>(), 3 * nb * sizeof(double),
                              hipMemcpyDeviceToHost, s));
    if (best) {
      vb.resize(nb);
      MIRORR_HIP(hipMemcpyAsync(vb.data(), bm->best.as<float>(), nb * sizeof(float),
                                hipMemcpyDeviceToHost, s));
    }
    if (second) {
      vs.resize(nb);
      MIRORR_HIP(hipMemcpyAsync(vs.data(), bm->second.as<float>(), nb * sizeof(float),
                                hipMemcpyDeviceToHost, s));
    }
    MIRORR_HIP(hipStreamSynchronize(s));
    for (int32_t k = 0; k < nk; ++k) {
      const int32_t i = idx[k];
      for (int r = 0; r < 3; ++r) {
        if (base_mm) base_mm[3 * k + r] = bmm[3 * i + r];
        if (match_mm) match_mm[3 * k + r] = mmm[3 * i + r];
      }
      if (best) best[k] = vb[i];
      if (second) second[k] = vs[i];
      if (block_index) block_index[k] = i;
    }
  });
}

}  // extern "C"
