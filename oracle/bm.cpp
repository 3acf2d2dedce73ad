// Oracle: exhaustive NCC block matcher (test infrastructure only; see oracle.h).
//
// Restates, for float images and uint8 masks:
//   BlockMatcher::normalizedCorrelation      itkBlockMatcher.txx:46-101
//   BlockMatcher::ComputeRegionForMatching   itkBlockMatcher.txx:531-600
//   BlockMatcher::GetOffsets(region, ...)    itkBlockMatcher.txx:631-848
//   BlockMatcher::PostProcessOffsets         itkBlockMatcher.txx:850-944
//   AbstractBlockMatcher::blockPositionToPoint  itkAbstractBlockMatcher.h:271-287
//   SparseImageRegionConstIterator scan order   itkSparseImageRegionConstIterator.h:196-243
//   BlockMatcherThreaded split and reduction    itkBlockMatcherThreaded.txx:22-205
#include "oracle.h"

#include <algorithm>
#include <utility>
#include <cmath>
#include <thread>
#include <vector>

namespace {

inline int64_t lin(const int dims[3], int x, int y, int z) {
  return (int64_t)x + (int64_t)dims[0] * ((int64_t)y + (int64_t)dims[1] * z);
}

// Number of samples a SparseImageRegionConstIterator with gap `g` visits on
// an axis of `size` voxels (itkSparseImageRegionConstIterator.h:230-243 sets
// the end at start + ((size-1)/g)*g with C truncating division; a size of 0
// with g >= 2 still visits the start once).
inline int sparse_count(long size, int g) {
  if (size <= 0) return 1;
  return (int)((size - 1) / g) + 1;
}

// Candidate start and count on one axis for a block whose first voxel is p
// (itkBlockMatcher.txx:730-756). searchEnd = L - BW + 1 uses the BASE image
// size, as the reference does.
inline void cand_range(int p, int L, int bw, int bg, int nw, int* first,
                       int* count) {
  long start = (long)p - (long)(nw / 2) * bg;
  long end = start + (long)nw * bg;
  const long search_end = (long)L - bw + 1;
  while (start < 0) start += bg;
  while (end > search_end) end -= bg;
  *first = (int)start;
  *count = sparse_count(end - start, bg);
}

struct Region {  // block-start region (index + size) as the sparse iterator sees it
  int index[3];
  long size[3];
};

}  // namespace

extern "C" double orc_ncc_pair(const float* base, const uint8_t* bmask,
                               const float* search, const uint8_t* smask,
                               const int dims[3], const int bi[3],
                               const int si[3], int bw) {
  // itkBlockMatcher.txx:57-68: means over ALL voxels of both blocks.
  double count = 0, mean_s = 0, mean_b = 0;
  for (int z = 0; z < bw; ++z)
    for (int y = 0; y < bw; ++y)
      for (int x = 0; x < bw; ++x) {
        mean_b += base[lin(dims, bi[0] + x, bi[1] + y, bi[2] + z)];
        mean_s += search[lin(dims, si[0] + x, si[1] + y, si[2] + z)];
        ++count;
      }
  mean_b /= count;
  mean_s /= count;
  // itkBlockMatcher.txx:70-95: sums only where both masks are non-zero
  // (USE_MASK_IN_METRIC, itkAbstractBlockMatcher.h:36).
  double prod = 0.0, bss = 0.0, sss = 0.0;
  for (int z = 0; z < bw; ++z)
    for (int y = 0; y < bw; ++y)
      for (int x = 0; x < bw; ++x) {
        const int64_t ob = lin(dims, bi[0] + x, bi[1] + y, bi[2] + z);
        const int64_t os = lin(dims, si[0] + x, si[1] + y, si[2] + z);
        const bool on = (bmask ? bmask[ob] != 0 : true) &&
                        (smask ? smask[os] != 0 : true);
        if (!on) continue;
        const double b = base[ob] - mean_b;
        const double s = search[os] - mean_s;
        prod += b * s;
        bss += b * b;
        sss += s * s;
      }
  // itkBlockMatcher.txx:97-100
  if (bss <= 0.0 || sss <= 0.0) return -2.0;
  return prod * prod / bss / sss;
}

// BlockMatcher::sumOfSquareDifference (itkBlockMatcher.txx:195-227): the
// mean of (base - search)^2 over voxels inside both masks; -2 when none is.
extern "C" double orc_ssd_pair(const float* base, const uint8_t* bmask, const float* search,
                               const uint8_t* smask, const int dims[3], const int bi[3],
                               const int si[3], int bw) {
  double acc = 0.0, n = 0.0;
  for (int z = 0; z < bw; ++z)
    for (int y = 0; y < bw; ++y)
      for (int x = 0; x < bw; ++x) {
        const int64_t ob = lin(dims, bi[0] + x, bi[1] + y, bi[2] + z);
        const int64_t os = lin(dims, si[0] + x, si[1] + y, si[2] + z);
        if ((bmask && !bmask[ob]) || (smask && !smask[os])) continue;
        const double d = (double)base[ob] - (double)search[os];
        acc += d * d;
        n += 1.0;
      }
  return n == 0.0 ? -2.0 : acc / n;
}

// BlockMatcher::correlationRatio (itkBlockMatcher.txx:105-191) with the
// matcher's scale_factor 0.1 (itkBlockMatcher.h:314). The (bin, base value)
// pairs of the voxels inside both masks, bin = floor(search * 0.1), are
// sorted; per bin, base values above 10 give a count, sum and sum of
// squares; bins with more than 2 such values add their variance
// (E[x^2] - E[x]^2) to a sum of variances; every such value adds to the
// totals. Score 1 - sum_var / var_total when the total count exceeds 2 and
// sum_var > 0, kept when above 0.4, else -2. With no voxel inside both masks
// the reference reads the front of an empty list and then returns -2 through
// the same test; -2 is returned here directly.
extern "C" double orc_cr_pair(const float* base, const uint8_t* bmask, const float* search,
                              const uint8_t* smask, const int dims[3], const int bi[3],
                              const int si[3], int bw) {
  std::vector<std::pair<int, double>> v;
  v.reserve((size_t)bw * bw * bw);
  for (int z = 0; z < bw; ++z)
    for (int y = 0; y < bw; ++y)
      for (int x = 0; x < bw; ++x) {
        const int64_t ob = lin(dims, bi[0] + x, bi[1] + y, bi[2] + z);
        const int64_t os = lin(dims, si[0] + x, si[1] + y, si[2] + z);
        if ((bmask && !bmask[ob]) || (smask && !smask[os])) continue;
        v.emplace_back((int)std::floor((double)search[os] * 0.1), (double)base[ob]);
      }
  if (v.empty()) return -2.0;
  std::sort(v.begin(), v.end());
  double tot_s = 0.0, tot_s2 = 0.0, tot_n = 0.0, sum_var = 0.0;
  size_t i = 0;
  while (i < v.size()) {
    const int bin = v[i].first;
    double s = 0.0, s2 = 0.0, n = 0.0;
    for (; i < v.size() && v[i].first == bin; ++i) {
      const double xv = v[i].second;
      if (xv > 10.0) {
        s += xv;
        s2 += xv * xv;
        n += 1.0;
      }
    }
    if (n > 2.0) {
      const double m = s / n;
      sum_var += s2 / n - m * m;
    }
    tot_s += s;
    tot_s2 += s2;
    tot_n += n;
  }
  if (!(tot_n > 2.0 && sum_var > 0.0)) return -2.0;
  const double m = tot_s / tot_n;
  const double cr = 1.0 - sum_var / (tot_s2 / tot_n - m * m);
  return cr > 0.4 ? cr : -2.0;
}

extern "C" double orc_ncc_pair_unmasked(const float* base, const float* search,
                                        const int dims[3], const int bi[3],
                                        const int si[3], int bw) {
  return orc_ncc_pair(base, nullptr, search, nullptr, dims, bi, si, bw);
}

extern "C" void orc_block_grid(const int dims[3], int bw, int ng,
                               int nblocks[3], int start[3]) {
  // itkBlockMatcher.txx:543-566
  for (int a = 0; a < 3; ++a) {
    const int n = (dims[a] - bw) / ng + 1;
    const int region = ng * (n - 1) + bw;
    const int gap_at_end = dims[a] - region;
    nblocks[a] = n;
    start[a] = gap_at_end / 2;
  }
}

namespace {

struct BMArgs {
  const float* base;
  const uint8_t* bmask;
  const float* search;
  const uint8_t* smask;
  const orc_geom* g;
  int bw, bg, nw, ng;
  int nb[3], bstart[3];
  int32_t* block_idx;
  int32_t* match_idx;
  float* best;
  float* second;
  const uint8_t* keep;  // per block (raster order): 0 = not selected, skipped
  int metric;           // 0 NCC, 1 SSD, 2 CR (BlockMatcher's metric map)
};

using PairMetric = double (*)(const float*, const uint8_t*, const float*, const uint8_t*,
                              const int*, const int*, const int*, int);

PairMetric metric_fn(int metric) {
  return metric == 1 ? orc_ssd_pair : metric == 2 ? orc_cr_pair : orc_ncc_pair;
}

// blockPositionToPoint (itkAbstractBlockMatcher.h:271-287): the physical
// point of the block corner plus spacing*BW/2, added WITHOUT the direction.
void block_to_point(const orc_geom* g, const int idx[3], int bw, double out[3]) {
  for (int r = 0; r < 3; ++r) {
    double p = g->origin[r];
    for (int c = 0; c < 3; ++c)
      p += g->direction[3 * r + c] * g->spacing[c] * (double)idx[c];
    out[r] = p + g->spacing[r] * (0.5 * bw);
  }
}

// BlockMatcher::GetOffsets(region, ...) for one block-start region
// (itkBlockMatcher.txx:631-848) followed by PostProcessOffsets
// (itkBlockMatcher.txx:850-944). Kept points are appended to out_b / out_m.
void match_region(const BMArgs& a, const Region& r, std::vector<double>& out_b,
                  std::vector<double>& out_m) {
  const int* dims = a.g->dims;
  // Block starts visited by the sparse iterator with gap NG.
  int cnt[3];
  for (int k = 0; k < 3; ++k) cnt[k] = sparse_count(r.size[k], a.ng);
  for (int kz = 0; kz < cnt[2]; ++kz)
    for (int ky = 0; ky < cnt[1]; ++ky)
      for (int kx = 0; kx < cnt[0]; ++kx) {
        const int p[3] = {r.index[0] + kx * a.ng, r.index[1] + ky * a.ng,
                          r.index[2] + kz * a.ng};
        const int gb[3] = {(p[0] - a.bstart[0]) / a.ng, (p[1] - a.bstart[1]) / a.ng,
                           (p[2] - a.bstart[2]) / a.ng};
        const int64_t id = gb[0] + (int64_t)a.nb[0] * (gb[1] + (int64_t)a.nb[1] * gb[2]);
        if (a.keep && !a.keep[id]) {
          // A block dropped by the variance selection is not searched: its
          // record keeps the initial score -3 and the block's own position.
          if (a.block_idx) for (int k = 0; k < 3; ++k) a.block_idx[3 * id + k] = p[k];
          if (a.match_idx) for (int k = 0; k < 3; ++k) a.match_idx[3 * id + k] = p[k];
          if (a.best) a.best[id] = -3.0f;
          if (a.second) a.second[id] = -3.0f;
          continue;
        }
        int first[3], ncand[3];
        for (int k = 0; k < 3; ++k)
          cand_range(p[k], dims[k], a.bw, a.bg, a.nw, &first[k], &ncand[k]);
        // Float score buffer (USE_FLOAT_SCORE_BUFFER, itkAbstractBlockMatcher.h:33-35),
        // initialised to -3 with the first candidate as match (:764-774).
        float best = -3.0f, second = -3.0f;
        int bidx[3] = {first[0], first[1], first[2]};
        const PairMetric score = metric_fn(a.metric);
        for (int cz = 0; cz < ncand[2]; ++cz)
          for (int cy = 0; cy < ncand[1]; ++cy)
            for (int cx = 0; cx < ncand[0]; ++cx) {
              const int s[3] = {first[0] + cx * a.bg, first[1] + cy * a.bg,
                                first[2] + cz * a.bg};
              const double v = score(a.base, a.bmask, a.search, a.smask, dims, p, s, a.bw);
              // itkBlockMatcher.txx:823-833: float buffer compared as double.
              if ((double)best < v) {
                second = best;
                best = (float)v;
                bidx[0] = s[0]; bidx[1] = s[1]; bidx[2] = s[2];
              } else if ((double)second < v) {
                second = (float)v;
              }
            }
        // Raw per-block record in global raster order.
        if (a.block_idx) for (int k = 0; k < 3; ++k) a.block_idx[3 * id + k] = p[k];
        if (a.match_idx) for (int k = 0; k < 3; ++k) a.match_idx[3 * id + k] = bidx[k];
        if (a.best) a.best[id] = best;
        if (a.second) a.second[id] = second;
        // PostProcessOffsets filter (itkBlockMatcher.txx:886-916).
        const double cb = best, c2 = second;
        if (cb > 1e-7 && std::fabs(cb - c2) > 1e-4) {
          double pb[3], pm[3];
          block_to_point(a.g, p, a.bw, pb);
          block_to_point(a.g, bidx, a.bw, pm);
          out_b.insert(out_b.end(), pb, pb + 3);
          out_m.insert(out_m.end(), pm, pm + 3);
        }
      }
}

// BlockMatcherThreaded::SplitRequestedRegion (itkBlockMatcherThreaded.txx:97-155).
// Returns the number of pieces the region splits into.
unsigned split_region(const Region& full, int bw, int ng, unsigned tid,
                      unsigned tcount, Region& out) {
  out = full;
  const long* s = full.size;
  const int axis = (s[0] > s[1]) ? ((s[0] > s[2]) ? 0 : 2) : ((s[1] > s[2]) ? 1 : 2);
  // (splitSize - BW) is unsigned in the reference: a negative difference wraps.
  const unsigned long diff = (unsigned long)s[axis] - (unsigned long)bw;
  if (diff < (unsigned long)ng) return 1;
  const unsigned n = (unsigned)((s[axis] - 1) / ng + 1);
  const unsigned per = n / tcount, res = n % tcount;
  const unsigned max_used = per > 0 ? tcount : res - 1;
  unsigned this_n = per + (tid < res ? 1u : 0u);
  unsigned first = tid * per + (tid < res ? tid : res);
  out.index[axis] += ng * (int)first;
  out.size[axis] = (long)(1u + (unsigned)ng * (this_n - 1u));
  return max_used + 1;
}

}  // namespace

extern "C" int64_t orc_bm_candidate_count_sel(const int dims[3], int bw, int bg, int nw, int ng,
                                              const uint8_t* keep) {
  int nb[3], st[3];
  orc_block_grid(dims, bw, ng, nb, st);
  int64_t tot = 0;
  for (int z = 0; z < nb[2]; ++z)
    for (int y = 0; y < nb[1]; ++y)
      for (int x = 0; x < nb[0]; ++x) {
        const int64_t b = x + (int64_t)nb[0] * (y + (int64_t)nb[1] * z);
        if (keep && !keep[b]) continue;
        int f, cx, cy, cz;
        cand_range(st[0] + x * ng, dims[0], bw, bg, nw, &f, &cx);
        cand_range(st[1] + y * ng, dims[1], bw, bg, nw, &f, &cy);
        cand_range(st[2] + z * ng, dims[2], bw, bg, nw, &f, &cz);
        tot += (int64_t)cx * cy * cz;
      }
  return tot;
}

extern "C" int64_t orc_bm_candidate_count(const int dims[3], int bw, int bg,
                                          int nw, int ng) {
  int nb[3], st[3];
  orc_block_grid(dims, bw, ng, nb, st);
  int64_t per_axis[3] = {0, 0, 0};
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < nb[a]; ++i) {
      int f, c;
      cand_range(st[a] + i * ng, dims[a], bw, bg, nw, &f, &c);
      per_axis[a] += c;
    }
  return per_axis[0] * per_axis[1] * per_axis[2];
}

extern "C" int orc_bm_get_offsets(const float* base, const uint8_t* bmask,
                                  const float* search, const uint8_t* smask,
                                  const orc_geom* g, int bw, int bg, int nw,
                                  int ng, int nthreads, int32_t* block_idx,
                                  int32_t* match_idx, float* best,
                                  float* second, double* base_mm,
                                  double* match_mm, int64_t* n_kept) {
  return orc_bm_get_offsets_sel(base, bmask, search, smask, g, bw, bg, nw, ng, nthreads, nullptr,
                                block_idx, match_idx, best, second, base_mm, match_mm, n_kept);
}

// Block-variance selection. The reference's culling is compiled out
// (itkBlockMatcher.txx:664-719: variance sum(x^2)/n - mean^2, cut-off by
// nth_element at floor(portion * N), blocks at or below it skipped, which
// keeps the top 1 - portion and drops everything at portion 1.0). The
// selection restated here keeps the top `portion` instead:
//   key_b  = float(sum over the block of (x - mean)^2), two-pass in double
//            over all BW^3 voxels (masks ignored, as in the dead code);
//   K      = ceil(portion * N) clamped to [1, N];
//   cutoff = the K-th largest key; keep every block with key >= cutoff
//            (blocks tied with the cut-off are all kept).
extern "C" int64_t orc_variance_select(const float* base, const int dims[3], int bw, int ng,
                                       double portion, uint8_t* keep, float* keys) {
  int nb[3], st[3];
  orc_block_grid(dims, bw, ng, nb, st);
  if (nb[0] <= 0 || nb[1] <= 0 || nb[2] <= 0) return 0;
  const int64_t N = (int64_t)nb[0] * nb[1] * nb[2];
  std::vector<float> key(N);
  const int n3 = bw * bw * bw;
  for (int64_t b = 0; b < N; ++b) {
    const int p[3] = {st[0] + ng * (int)(b % nb[0]), st[1] + ng * (int)((b / nb[0]) % nb[1]),
                      st[2] + ng * (int)(b / ((int64_t)nb[0] * nb[1]))};
    double s = 0.0;
    for (int z = 0; z < bw; ++z)
      for (int y = 0; y < bw; ++y)
        for (int x = 0; x < bw; ++x)
          s += base[(p[0] + x) + (int64_t)dims[0] * ((p[1] + y) + (int64_t)dims[1] * (p[2] + z))];
    const double mu = s / n3;
    double ss = 0.0;
    for (int z = 0; z < bw; ++z)
      for (int y = 0; y < bw; ++y)
        for (int x = 0; x < bw; ++x) {
          const double d =
              base[(p[0] + x) + (int64_t)dims[0] * ((p[1] + y) + (int64_t)dims[1] * (p[2] + z))] -
              mu;
          ss += d * d;
        }
    key[b] = (float)ss;
    if (keys) keys[b] = key[b];
  }
  int64_t K = (int64_t)std::ceil(portion * (double)N);
  K = std::min<int64_t>(std::max<int64_t>(K, 1), N);
  std::vector<float> sorted(key);
  std::nth_element(sorted.begin(), sorted.begin() + (N - K), sorted.end());
  const float cutoff = sorted[N - K];
  int64_t kept = 0;
  for (int64_t b = 0; b < N; ++b) {
    keep[b] = key[b] >= cutoff ? 1 : 0;
    kept += keep[b];
  }
  return kept;
}

extern "C" int orc_bm_get_offsets_sel(const float* base, const uint8_t* bmask,
                                      const float* search, const uint8_t* smask,
                                      const orc_geom* g, int bw, int bg, int nw, int ng,
                                      int nthreads, const uint8_t* block_keep,
                                      int32_t* block_idx, int32_t* match_idx, float* best,
                                      float* second, double* base_mm, double* match_mm,
                                      int64_t* n_kept) {
  return orc_bm_get_offsets_metric(base, bmask, search, smask, g, bw, bg, nw, ng, nthreads,
                                   block_keep, 0, block_idx, match_idx, best, second, base_mm,
                                   match_mm, n_kept);
}

extern "C" int orc_bm_get_offsets_metric(const float* base, const uint8_t* bmask,
                                         const float* search, const uint8_t* smask,
                                         const orc_geom* g, int bw, int bg, int nw, int ng,
                                         int nthreads, const uint8_t* block_keep, int metric,
                                         int32_t* block_idx, int32_t* match_idx, float* best,
                                         float* second, double* base_mm, double* match_mm,
                                         int64_t* n_kept) {
  BMArgs a{base, bmask, search, smask, g, bw, bg, nw, ng, {0, 0, 0}, {0, 0, 0},
           block_idx, match_idx, best, second, block_keep, metric};
  orc_block_grid(g->dims, bw, ng, a.nb, a.bstart);
  // ComputeRegionForMatching: region of block starts (itkBlockMatcher.txx:566-574).
  Region full;
  for (int k = 0; k < 3; ++k) {
    full.index[k] = a.bstart[k];
    full.size[k] = (long)ng * (a.nb[k] - 1) + 1;
    if (a.nb[k] <= 0) {
      if (n_kept) *n_kept = 0;
      return 1;
    }
  }
  std::vector<std::vector<double>> lb, lm;
  if (nthreads <= 0) {
    lb.resize(1);
    lm.resize(1);
    match_region(a, full, lb[0], lm[0]);
  } else {
    const unsigned T = (unsigned)nthreads;
    lb.resize(T);
    lm.resize(T);
    std::vector<Region> regs(T);
    std::vector<unsigned> totals(T);
    for (unsigned t = 0; t < T; ++t) totals[t] = split_region(full, bw, ng, t, T, regs[t]);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; ++t) {
      if (t >= totals[t]) continue;  // ThreaderCallback: idle thread
      th.emplace_back([&, t]() { match_region(a, regs[t], lb[t], lm[t]); });
    }
    for (auto& x : th) x.join();
  }
  // ReduceLists: concatenate in thread order (itkBlockMatcherThreaded.txx:73-89).
  int64_t n = 0;
  for (size_t t = 0; t < lb.size(); ++t) {
    const size_t m = lb[t].size() / 3;
    for (size_t i = 0; i < m; ++i, ++n) {
      for (int k = 0; k < 3; ++k) {
        if (base_mm) base_mm[3 * n + k] = lb[t][3 * i + k];
        if (match_mm) match_mm[3 * n + k] = lm[t][3 * i + k];
      }
    }
  }
  if (n_kept) *n_kept = n;
  return 0;
}
