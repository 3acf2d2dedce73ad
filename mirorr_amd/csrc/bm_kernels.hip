// Exhaustive NCC block search for gfx950 (MI355X).
//
// Replaces the reference OpenCL kernel optKmatch (kernels/optKmatchMask.cl:67-332)
// and the CPU matcher it mirrors (itkBlockMatcher.txx:46-101,631-848). The
// score of block b at candidate c is r^2 = Sxy^2 / (Sxx * Syy) with means over
// all BW^3 voxels and sums over voxels where both masks are on; -2 when a sum of
// squares is zero (itkBlockMatcher.txx:97-100).
//
// Design (not a translation of the OpenCL work-item-per-voxel scheme):
//  * A search image is summarised once per match by window statistics
//    (mean and masked centred sum of squares of every BW^3 window, in double).
//  * A base image is summarised once per match by centred blocks x - mean.
//  * Fast path (BW=4, BG=1): one workgroup owns a tile of 1 x TY x TZ blocks
//    stacked in y/z, whose union search window is staged in LDS once, shifted
//    by a tile reference value. A lane owns KY x NW candidates (one z offset,
//    KY y offsets, all NW x offsets) and accumulates Sum x~ * y' with one FMA
//    per voxel-eval, the centred block held in 64 VGPRs and rows of the window
//    read as 16-byte LDS loads shared by 4*KY*NW FMAs. Sxy, Syy then follow
//    from the window statistics (1/Syy stored). Blocks whose base mask or own search window
//    is not all on are queued (tile + block bitmask) for the masked kernel
//    (masked sums, 4-5 ops/voxel-eval); the rest of the tile stays here.
//  * Top-2 per block uses the float score buffer semantics of the reference
//    with ties on the best broken by the earliest candidate in scan order.
#include "bm_kernels.hpp"
#include "../../include/mirorr_gpu.h"
#include "select.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

namespace mirorr {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxWindow = 64;  // window planes (and rows) a tile may span
// Masked row tasks are grouped by their skip pattern: 4 bits of live planes,
// 4 bits of live rows and one bit for a partly masked base block. The
// bucket counters are 16-bit halves of 32-bit words (at most 64 x 169 tasks).
constexpr int kPatBuckets = 512;

__device__ __forceinline__ int64_t lin3(int nx, int ny, int x, int y, int z) {
  return (int64_t)x + (int64_t)nx * ((int64_t)y + (int64_t)ny * z);
}

// ---------------------------------------------------------------------------
// Window statistics of a search image.
template <int BWT>
__global__ __launch_bounds__(kThreads) void search_stats_kernel(
    const float* __restrict__ y, const uint8_t* __restrict__ sm, int nx, int ny,
    int nz, int bw_rt, float2* __restrict__ out) {
  const int bw = BWT > 0 ? BWT : bw_rt;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int yy = blockIdx.y;
  const int z = blockIdx.z;
  if (x > nx - bw || yy > ny - bw || z > nz - bw) return;
  double sum = 0.0;
  for (int dz = 0; dz < bw; ++dz)
    for (int dy = 0; dy < bw; ++dy) {
      const float* r = y + lin3(nx, ny, x, yy + dy, z + dz);
#pragma unroll
      for (int dx = 0; dx < (BWT > 0 ? BWT : 1); ++dx) sum += r[dx];
      if (BWT == 0)
        for (int dx = 1; dx < bw; ++dx) sum += r[dx];
    }
  const double mu = sum / (double)(bw * bw * bw);
  double ss = 0.0;
  for (int dz = 0; dz < bw; ++dz)
    for (int dy = 0; dy < bw; ++dy) {
      const int64_t o = lin3(nx, ny, x, yy + dy, z + dz);
      for (int dx = 0; dx < bw; ++dx) {
        if (sm && sm[o + dx] == 0) continue;
        const double d = (double)y[o + dx] - mu;
        ss += d * d;
      }
    }
  out[lin3(nx, ny, x, yy, z)] = make_float2((float)mu, ss > 0.0 ? (float)(1.0 / ss) : -__builtin_huge_valf());
}

// BW = 4: a thread owns one (x, y) column of window positions and slides
// along z over a chunk. Each of the four 4x4 planes of the current window is
// summarised once, when it enters, by FP64 sums: all 16 values (for the mean)
// and, over the voxels whose search mask is on, the count and the first and
// second moments of y - c0 (c0 = the column's first value, which keeps the
// terms small). Then
//   Syy = sum_on (y - mu)^2 = S2 - 2 m S1 + m^2 n,   m = mu - c0,
// where S1, S2, n are the on-voxel sums over the window. A window whose 64
// values are all equal has Syy = 0 exactly, as the reference's two-pass sum
// gives, so the per-plane minimum and maximum detect that case.
constexpr int kStatsZ = 16;  // z positions per thread

template <bool MASK, class E>
__global__ __launch_bounds__(kThreads) void search_stats4_kernel(
    const E* __restrict__ y, const uint8_t* __restrict__ sm, int nx, int ny, int nz,
    float2* __restrict__ out) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int yy = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int z0 = blockIdx.z * kStatsZ;
  if (x > nx - 4 || yy > ny - 4) return;
  const int zend = min(z0 + kStatsZ, nz - 3);  // exclusive end of output z
  if (z0 >= zend) return;
  const int64_t nxy = (int64_t)nx * ny;
  const E* ycol = y + lin3(nx, ny, x, yy, 0);  // the column's (x, y); + z * nxy
  const uint8_t* mcol = MASK ? sm + lin3(nx, ny, x, yy, 0) : nullptr;
  const double c0 = (double)(float)ycol[(int64_t)z0 * nxy];
  double ps[4], p1[4], p2[4];
  int pn[4];
  float pmin[4], pmax[4];
  auto load = [&](int slot, int z) {
    const E* yp = ycol + (int64_t)z * nxy;
    float v[16];
    uint32_t mw[4] = {0x01010101u, 0x01010101u, 0x01010101u, 0x01010101u};
#pragma unroll
    for (int dy = 0; dy < 4; ++dy) {
#pragma unroll
      for (int dx = 0; dx < 4; ++dx) v[4 * dy + dx] = (float)yp[dy * nx + dx];
      if (MASK) {
        const uint8_t* mr = mcol + (int64_t)z * nxy + dy * nx;
        mw[dy] = (uint32_t)mr[0] | ((uint32_t)mr[1] << 8) | ((uint32_t)mr[2] << 16) |
                 ((uint32_t)mr[3] << 24);
      }
    }
    double s = 0.0, s1 = 0.0, s2 = 0.0;
    int cnt = 0;
    float lo = __builtin_huge_valf(), hi = -__builtin_huge_valf();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      lo = fminf(lo, v[k]);
      hi = fmaxf(hi, v[k]);
      s += (double)v[k];
      // An off voxel adds exact zeros, so the sums are those over the on
      // voxels, in the same order, without a branch.
      const double d = (double)v[k] - c0;
      const bool on = !MASK || ((mw[k >> 2] >> (8 * (k & 3))) & 0xffu) != 0u;
      const double dw = on ? d : 0.0;
      s1 += dw;
      s2 = fma(dw, d, s2);
      cnt += on ? 1 : 0;
    }
    ps[slot] = s;
    p1[slot] = s1;
    p2[slot] = s2;
    pn[slot] = cnt;
    pmin[slot] = lo;
    pmax[slot] = hi;
  };
  auto emit = [&](int z) {
    const double mu = (ps[0] + ps[1] + ps[2] + ps[3]) * (1.0 / 64.0);
    const double m = mu - c0;
    const double s1 = p1[0] + p1[1] + p1[2] + p1[3];
    const double s2 = p2[0] + p2[1] + p2[2] + p2[3];
    const double n = (double)(pn[0] + pn[1] + pn[2] + pn[3]);
    const float lo = fminf(fminf(pmin[0], pmin[1]), fminf(pmin[2], pmin[3]));
    const float hi = fmaxf(fmaxf(pmax[0], pmax[1]), fmaxf(pmax[2], pmax[3]));
    double ss = fma(m * n - 2.0 * s1, m, s2);
    if (lo == hi || n == 0.0) ss = 0.0;
    out[lin3(nx, ny, x, yy, z)] = make_float2((float)mu, ss > 0.0 ? (float)(1.0 / ss) : -__builtin_huge_valf());
  };
  // Slots 0..3 hold the planes z .. z + 3 of the current window: one plane
  // enters per step and the slots shift down.
  load(0, z0);
  load(1, z0 + 1);
  load(2, z0 + 2);
#pragma unroll 1
  for (int z = z0; z < zend; ++z) {
    load(3, z + 3);
    emit(z);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      ps[k] = ps[k + 1];
      p1[k] = p1[k + 1];
      p2[k] = p2[k + 1];
      pn[k] = pn[k + 1];
      pmin[k] = pmin[k + 1];
      pmax[k] = pmax[k + 1];
    }
  }
}

// ---------------------------------------------------------------------------
// Per-block preparation of a base image: one wavefront per block, a lane per
// voxel (BW^3 <= 64 handled in one pass; larger blocks loop).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <int BWT>
__global__ __launch_bounds__(kThreads) void block_prep_kernel(
    const float* __restrict__ x, const uint8_t* __restrict__ bm, BMGrid g,
    float* __restrict__ xt, float* __restrict__ sxx, float* __restrict__ sumxt,
    uint32_t* __restrict__ flags) {
  const int lane = threadIdx.x & 63;
  const int b = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);  // nblocks < 2^31
  if (b >= g.nblocks) return;  // wave-uniform
  const int bw = BWT > 0 ? BWT : g.bw, n = bw * bw * bw;
  const int bxi = b % g.nb[0];
  const int byi = (b / g.nb[0]) % g.nb[1];
  const int bzi = b / (g.nb[0] * g.nb[1]);
  const int px = g.start[0] + g.ng * bxi;
  const int py = g.start[1] + g.ng * byi;
  const int pz = g.start[2] + g.ng * bzi;
  auto voxel = [&](int k) {
    const int dx = k % bw, dy = (k / bw) % bw, dz = k / (bw * bw);
    return lin3(g.dims[0], g.dims[1], px + dx, py + dy, pz + dz);
  };
  double s = 0.0;
  for (int k = lane; k < n; k += 64) s += x[voxel(k)];
  const double mx = wave_sum(s) / (double)n;
  double s2 = 0.0, st = 0.0;
  int on = 0;
  for (int k = lane; k < n; k += 64) {
    const int64_t o = voxel(k);
    const double d = (double)x[o] - mx;
    const float f = (float)d;
    xt[(int64_t)b * n + k] = f;
    s2 += d * d;
    st += (double)f;
    on += (bm == nullptr || bm[o] != 0) ? 1 : 0;
  }
  s2 = wave_sum(s2);
  st = wave_sum(st);
  on = wave_sum_i(on);
  if (lane == 0) {
    sxx[b] = (float)s2;
    sumxt[b] = (float)st;
    flags[b] = (on == n ? 1u : 0u) | (on == 0 ? 2u : 0u);
  }
}

// BW = 4: four lanes per block, lane q owning the block's plane dz = q (four
// rows of four voxels). Sums are reduced within the quad; the centred voxels
// of a wave's 16 blocks form one contiguous 4 KB run of xt.
__device__ __forceinline__ double quad_sum(double v) {
  v += __shfl_xor(v, 1, 4);
  v += __shfl_xor(v, 2, 4);
  return v;
}
__device__ __forceinline__ int quad_sum_i(int v) {
  v += __shfl_xor(v, 1, 4);
  v += __shfl_xor(v, 2, 4);
  return v;
}

template <class E>
__global__ __launch_bounds__(kThreads) void block_prep4_kernel(
    const E* __restrict__ x, const uint8_t* __restrict__ bm, BMGrid g,
    float* __restrict__ xt, float* __restrict__ sxx, float* __restrict__ sumxt,
    uint32_t* __restrict__ flags) {
  const int t = blockIdx.x * kThreads + threadIdx.x;  // nblocks * 4 < 2^31
  const int b = t >> 2, dz = t & 3;
  if (b >= g.nblocks) return;  // the four lanes of a block leave together
  const int bxi = b % g.nb[0];
  const int byi = (b / g.nb[0]) % g.nb[1];
  const int bzi = b / (g.nb[0] * g.nb[1]);
  const int px = g.start[0] + g.ng * bxi;
  const int py = g.start[1] + g.ng * byi;
  const int pz = g.start[2] + g.ng * bzi + dz;
  float v[16];
  double s = 0.0;
#pragma unroll
  for (int dy = 0; dy < 4; ++dy) {
    const E* row = x + lin3(g.dims[0], g.dims[1], px, py + dy, pz);
#pragma unroll
    for (int dx = 0; dx < 4; ++dx) {
      v[dy * 4 + dx] = (float)row[dx];
      s += (double)v[dy * 4 + dx];
    }
  }
  const double mx = quad_sum(s) * (1.0 / 64.0);
  double s2 = 0.0, st = 0.0;
  float4* out = reinterpret_cast<float4*>(xt + (int64_t)b * 64 + dz * 16);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    float f[4];
#pragma unroll
    for (int dx = 0; dx < 4; ++dx) {
      const double d = (double)v[r * 4 + dx] - mx;
      f[dx] = (float)d;
      s2 = fma(d, d, s2);
      st += (double)f[dx];
    }
    out[r] = make_float4(f[0], f[1], f[2], f[3]);
  }
  int on = 16;
  if (bm != nullptr) {
    on = 0;
#pragma unroll
    for (int dy = 0; dy < 4; ++dy) {
      const uint8_t* row = bm + lin3(g.dims[0], g.dims[1], px, py + dy, pz);
#pragma unroll
      for (int dx = 0; dx < 4; ++dx) on += row[dx] != 0 ? 1 : 0;
    }
  }
  s2 = quad_sum(s2);
  st = quad_sum(st);
  on = quad_sum_i(on);
  if (dz == 0) {
    sxx[b] = (float)s2;
    sumxt[b] = (float)st;
    flags[b] = (on == 64 ? 1u : 0u) | (on == 0 ? 2u : 0u);
  }
}

// ---------------------------------------------------------------------------
// Top-2 bookkeeping shared by the kernels.
struct Top2 {
  float best;
  int scan;  // scan-order index of the best candidate
  float second;
};

__device__ __forceinline__ void top2_init(Top2& t) {
  t.best = -3.0f;
  t.scan = INT_MAX;
  t.second = -3.0f;
}

// Candidates reach a lane in increasing scan order: strict '>' keeps the
// earliest (itkBlockMatcher.txx:823-833 on the float buffer).
__device__ __forceinline__ void top2_push(Top2& t, float s, int scan) {
  if (s > t.best) {
    t.second = t.best;
    t.best = s;
    t.scan = scan;
  } else {
    t.second = fmaxf(t.second, s);
  }
}

// The same update as top2_push in three operations: the new second is the
// median of (best, second, s) since second <= best.
__device__ __forceinline__ void top2_step(Top2& t, float s, int scan) {
  t.scan = s > t.best ? scan : t.scan;
  t.second = __builtin_amdgcn_fmed3f(t.best, t.second, s);
  t.best = fmaxf(t.best, s);
}

// The record of a candidate sequence followed by a later one in scan order
// (the same result as pushing b's candidates after a's).
__device__ __forceinline__ Top2 top2_append(const Top2& a, const Top2& b) {
  Top2 r;
  if (b.best > a.best) {
    r.best = b.best;
    r.scan = b.scan;
    r.second = fmaxf(b.second, a.best);
  } else {
    r.best = a.best;
    r.scan = a.scan;
    r.second = fmaxf(a.second, b.best);
  }
  return r;
}

// Order-independent merge of per-lane records of one block: best = max
// score, earliest scan index on ties; second = max of all seconds and of the
// bests that did not win (equal bests make best == second, as in the
// reference's sequential update).
__device__ inline Top2 top2_merge(const float* tb, const int* ti, const float* ts,
                                  int first, int n) {
  Top2 r;
  top2_init(r);
  int win = -1;
  for (int j = first; j < first + n; ++j) {
    if (tb[j] > r.best || (tb[j] == r.best && ti[j] < r.scan)) {
      r.best = tb[j];
      r.scan = ti[j];
      win = j;
    }
  }
  float sec = -3.0f;
  for (int j = first; j < first + n; ++j) {
    sec = fmaxf(sec, ts[j]);
    if (j != win) sec = fmaxf(sec, tb[j]);
  }
  r.second = sec;
  return r;
}

__device__ __forceinline__ float ncc_score(float sxy, float sxx, float syy) {
  // -2 when either sum of squares is not positive (itkBlockMatcher.txx:97-98).
  if (!(sxx > 0.0f) || !(syy > 0.0f)) return -2.0f;
  return sxy * sxy * __builtin_amdgcn_rcpf(sxx * syy);
}

// ---------------------------------------------------------------------------
// Tiled kernels, BW = 4, BG = 1.
template <int NW_, int KY_, int TY_, int TZ_>
struct TileCfg {
  static constexpr int BW = 4;
  static constexpr int NW = NW_;
  static constexpr int KY = KY_;
  static constexpr int TY = TY_;
  static constexpr int TZ = TZ_;
  static constexpr int H = NW / 2;
  static constexpr int NCY = (NW + KY - 1) / KY;
  static constexpr int TPB = NW * NCY;  // fast-path tasks per block
  static constexpr int NB = TY * TZ;    // blocks per tile
  static constexpr int NT = TPB * NB;   // fast-path tasks per tile
  static constexpr int WX = NW + BW - 1;
  static constexpr int PX = (WX + 3) & ~3;  // LDS row pitch (floats)
  static constexpr int NV = PX / 4;         // float4 per row
  static constexpr int YPAD = NCY * KY - NW;
  static constexpr int GT = NB * NW * NW;  // masked-path tasks per tile
  // Rows per staged plane of the fast kernel's window: with 12-float rows an
  // odd count makes the plane stride an odd number of 16-byte words, so the
  // lanes of a wave (consecutive cz) read rows in different LDS banks.
  __host__ __device__ static constexpr int fast_pitch(int wy) {
    return PX % 8 == 4 ? (wy | 1) : wy;
  }
  // Workgroups per CU the masked kernel's registers are budgeted for: four
  // (128 VGPRs) up to NW 9, three for the wider rows of NW 11 and 13.
  static constexpr int MWG = NW <= 9 ? 4 : 3;
  static_assert(NT <= kThreads, "tile too large");
  static_assert(NB <= 64, "block bitmasks are 64-bit");
};

// Static LDS of the masked kernel (bytes, rounded up).
template <class T>
constexpr size_t masked_static_lds() {
  return T::NB * (16 * 16 + 8 + 4 + 4 + 8) + T::NB * T::NW * T::NW * (4 + 4 + 4 + 2) +
         2 * 8 * kMaxWindow + 2 * kPatBuckets + 256;
}

struct TileArgs {
  const void* y;  // float or half_t elements (the kernels' E)
  const uint8_t* sm;
  const uint8_t* bm;
  BMGrid g;
  const float* xt;
  const float* sxx;
  const float* sumxt;
  const uint32_t* bflags;
  const float2* stats;
  int ntx, nty, ntz;
  int32_t* queue;
  unsigned long long* qbits;  // per queued tile: the blocks left to the masked kernel
  int32_t* qcount;
  int msplit;                 // masked kernel: workgroups per queued tile
  float* out_best;
  float* out_second;
  int32_t* out_match;
  unsigned long long* work;  // optional: voxel-evals done by this launch
  int32_t* rlist;            // optional: blocks to re-score in double
  int32_t* rcount;
  int rties;                 // list exact ties too
};

// Blocks of the tiled search listed for the exact re-scoring (below): the
// float best or best-second gap lies within kRescoreBand of a score-filter
// threshold (best > 1e-7, |best - second| > 1e-4), or (with ties) the best is
// an exact tie, whose float-buffer record depends on how the double score
// rounds although the filter culls it either way. The band is twice the
// largest float error measured on the phantom, resampled-with-masks and
// CT/MR workloads (2.3e-5, at high-contrast blocks whose best window lies
// across an intensity step from the block's own position); a GPU test
// checks that margin on those workloads.
constexpr double kRescoreBand = 5e-5;
constexpr int kRescoreMaxCand = 13 * 13 * 13;
constexpr int kRescoreMaxWin = 16 * 16 * 16;  // (NW + BW - 1)^3 at NW 13

__device__ __forceinline__ bool near_filter(float best, float second, int ties) {
  const double b = best, s = second;
  return fabs(b - 1e-7) < kRescoreBand || fabs(fabs(b - s) - 1e-4) < kRescoreBand ||
         (ties && best == second && b > 0.0);
}

// Voxel-evals of the valid blocks of a tile (profiling only).
template <class T>
__device__ unsigned long long tile_evals(const TileArgs& a, int px, int py0, int pz0,
                                         const int* bvalid) {
  const BMGrid& g = a.g;
  auto cnt = [&](int p, int L) {
    return min(p - T::H + T::NW - 1, L - T::BW) - max(p - T::H, 0) + 1;
  };
  unsigned long long tot = 0;
  const int cx = cnt(px, g.dims[0]);
  for (int bl = 0; bl < T::NB; ++bl) {
    if (bvalid[bl] != 1) continue;
    const int py = py0 + g.ng * (bl % T::TY), pz = pz0 + g.ng * (bl / T::TY);
    tot += (unsigned long long)cx * cnt(py, g.dims[1]) * cnt(pz, g.dims[2]);
  }
  return tot * 64ull;
}

// Workgroups are dealt to the 8 XCDs round-robin (workgroup i runs on XCD
// i % 8, each XCD with its own L2). Renumbering them so that each XCD walks a
// contiguous run of tiles keeps neighbouring tiles, which re-read the same
// window rows and statistics, in one L2. A bijection for any tile count.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int bid, int nt) {
  const int q = nt / kXcds, r = nt % kXcds;
  const int x = bid % kXcds;
  return x * q + min(x, r) + bid / kXcds;
}

template <class T>
__device__ __forceinline__ void tile_origin(const TileArgs& a, int tile, int& bx,
                                            int& by0, int& bz0) {
  bx = tile % a.ntx;
  by0 = ((tile / a.ntx) % a.nty) * T::TY;
  bz0 = (tile / (a.ntx * a.nty)) * T::TZ;
}

template <class T>
__device__ __forceinline__ void write_block_result(const TileArgs& a, int64_t b,
                                                   int px, int py, int pz,
                                                   const Top2& r) {
  const int cx = r.scan % T::NW;
  const int cy = (r.scan / T::NW) % T::NW;
  const int cz = r.scan / (T::NW * T::NW);
  a.out_best[b] = r.best;
  a.out_second[b] = r.second;
  a.out_match[3 * b + 0] = px - T::H + cx;
  a.out_match[3 * b + 1] = py - T::H + cy;
  a.out_match[3 * b + 2] = pz - T::H + cz;
  if (a.rlist && near_filter(r.best, r.second, a.rties)) a.rlist[atomicAdd(a.rcount, 1)] = (int32_t)b;
}

// Record of a block dropped by the variance selection: the initial score
// buffer (-3) with the block's own position as match.
__device__ __forceinline__ void write_skipped(const TileArgs& a, int64_t b, int px, int py,
                                              int pz) {
  a.out_best[b] = -3.0f;
  a.out_second[b] = -3.0f;
  a.out_match[3 * b + 0] = px;
  a.out_match[3 * b + 1] = py;
  a.out_match[3 * b + 2] = pz;
}

// Stages a tile's search window into LDS: rows of PX floats (wx fastest,
// then wy, then wz), y - ref inside the image and 0 outside. One item is a
// 4-wide column group of one row, so the row index needs one division by
// the runtime height WY, done with a float reciprocal (exact here: rows are
// small integers and the quotient's fractional part is at least 0.5 / WY
// away from an integer).
//  * MASKED: voxels whose search mask is off stage as 0, the mask bits go
//    to `wm` (one byte per voxel, same layout), a voxel whose search mask is
//    on sets bit wy of zrow[wz], and a window voxel that is outside the image
//    or whose search mask is off sets bit wy of zbad[wz].
//  * otherwise: a voxel whose search mask is off sets bit wy of zrow[wz].
template <class T, bool MASKED, class E>
__device__ __forceinline__ void stage_window(const TileArgs& a, float* win, uint8_t* wm,
                                             unsigned long long* zrow, int WY, int WZ, int wx0,
                                             int wy0, int wz0, float ref,
                                             unsigned long long* zbad = nullptr) {
  const int nx = a.g.dims[0], ny = a.g.dims[1], nz = a.g.dims[2];
  const E* Y = static_cast<const E*>(a.y);
  const int items = T::NV * WY * WZ;
  const float inv_wy = 1.0f / (float)WY;
  for (int it = threadIdx.x; it < items; it += kThreads) {
    const int v = it % T::NV;
    const int row = it / T::NV;
    const int wz = (int)(((float)row + 0.5f) * inv_wy);
    const int wy = row - wz * WY;
    const int gy = wy0 + wy, gz = wz0 + wz;
    float q[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    uint32_t mbits = 0u;
    if (gy >= 0 && gy < ny && gz >= 0 && gz < nz) {
      const int64_t rb = lin3(nx, ny, 0, gy, gz);
      const int gx0 = wx0 + 4 * v;
      float y[4];
      uint32_t inside, on;  // bit k: voxel k inside image and window / search mask on
      if (gx0 >= 0 && gx0 + 3 < nx && 4 * v + 3 < T::WX) {
        // interior: the four loads issue together
        inside = 0xfu;
#pragma unroll
        for (int k = 0; k < 4; ++k) y[k] = (float)Y[rb + gx0 + k];
        on = inside;
        if (a.sm) {
          uint8_t m[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) m[k] = a.sm[rb + gx0 + k];
#pragma unroll
          for (int k = 0; k < 4; ++k)
            if (m[k] == 0) on &= ~(1u << k);
        }
      } else {
        inside = 0u;
        on = 0u;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int gx = gx0 + k;
          y[k] = 0.0f;
          if (4 * v + k < T::WX && gx >= 0 && gx < nx) {
            inside |= 1u << k;
            y[k] = (float)Y[rb + gx];
            if (a.sm == nullptr || a.sm[rb + gx] != 0) on |= 1u << k;
          }
        }
      }
      const uint32_t use = MASKED ? on : inside;
#pragma unroll
      for (int k = 0; k < 4; ++k) q[k] = ((use >> k) & 1u) ? y[k] - ref : 0.0f;
      if (MASKED) {
#pragma unroll
        for (int k = 0; k < 4; ++k) mbits |= ((on >> k) & 1u) << (8 * k);
        if (on) atomicOr(&zrow[wz], 1ull << wy);
        const uint32_t cols = 4 * v + 3 < T::WX ? 0xfu : (1u << (T::WX - 4 * v)) - 1u;
        if (cols & ~on) atomicOr(&zbad[wz], 1ull << wy);
      } else if (inside & ~on) {
        atomicOr(&zrow[wz], 1ull << wy);
      }
    } else if (MASKED && 4 * v < T::WX) {
      atomicOr(&zbad[wz], 1ull << wy);  // a row outside the image
    }
    reinterpret_cast<float4*>(win)[row * T::NV + v] = make_float4(q[0], q[1], q[2], q[3]);
    if (MASKED) reinterpret_cast<uint32_t*>(wm)[row * T::NV + v] = mbits;
  }
}

// Blocks of a tile whose own search window (WX rows from the block's offset
// in y and z) holds a search-mask-off voxel, from the per-plane row bits.
template <class T>
__device__ __forceinline__ bool window_has_off(const unsigned long long* zrow, int bl, int ng) {
  const int bly = bl % T::TY, blz = bl / T::TY;
  const unsigned long long ym = ((1ull << T::WX) - 1ull) << (bly * ng);
  unsigned long long any = 0ull;
#pragma unroll
  for (int k = 0; k < T::WX; ++k) any |= zrow[blz * ng + k];
  return (any & ym) != 0ull;
}

// Record of a block whose base mask is off at every voxel: each candidate
// inside the image scores -2 (the base sum of squares over both masks is 0,
// itkBlockMatcher.txx:96-97), so the first candidate in scan order wins and
// the second is -2 when there are at least two candidates, else -3.
template <class T>
__device__ __forceinline__ void write_all_off(const TileArgs& a, int64_t b, int px, int py,
                                              int pz) {
  const int p[3] = {px, py, pz};
  int lo[3], n = 1;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    lo[d] = max(p[d] - T::H, 0);
    n *= min(p[d] - T::H + T::NW - 1, a.g.dims[d] - T::BW) - lo[d] + 1;
  }
  a.out_best[b] = -2.0f;
  a.out_second[b] = n >= 2 ? -2.0f : -3.0f;
  a.out_match[3 * b + 0] = lo[0];
  a.out_match[3 * b + 1] = lo[1];
  a.out_match[3 * b + 2] = lo[2];
}

// Block states in a tile: 0 outside the grid, 1 searched, 2 skipped,
// 3 left to the masked kernel, 4 base mask off everywhere (no search).
__device__ __forceinline__ int block_state(const TileArgs& a, int gy, int gz, int64_t b) {
  if (gy >= a.g.nb[1] || gz >= a.g.nb[2]) return 0;
  return (a.bflags[b] & kBlockSkip) ? 2 : 1;
}

// Orderable encodings for LDS atomics: max over the key is max over the
// score, then min over the scan index.
__device__ __forceinline__ uint32_t ord_f32(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unord_f32(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}
__device__ __forceinline__ unsigned long long best_key(float s, int scan) {
  return ((unsigned long long)ord_f32(s) << 32) | (unsigned long long)(0xffffffffu - (uint32_t)scan);
}

// Per-block merge state in LDS.
struct MergeSlot {
  unsigned long long best;  // best_key of the winner
  uint32_t second;          // ord_f32 of the second best
  uint32_t pad;
};

__device__ __forceinline__ void merge_init(MergeSlot* m, int nb) {
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    m[i].best = 0ull;
    m[i].second = ord_f32(-3.0f);
  }
}
// Phase 1: the winner. Phase 2: the second = max of all seconds and of the
// bests that did not win (equal bests give best == second, as the
// reference's sequential update does).
__device__ __forceinline__ void merge_phase1(MergeSlot* m, int bl, const Top2& t) {
  if (t.scan != INT_MAX) atomicMax(&m[bl].best, best_key(t.best, t.scan));
}
__device__ __forceinline__ void merge_phase2(MergeSlot* m, int bl, const Top2& t) {
  if (t.scan == INT_MAX) return;
  uint32_t v = ord_f32(t.second);
  if (best_key(t.best, t.scan) != m[bl].best) v = max(v, ord_f32(t.best));
  atomicMax(&m[bl].second, v);
}
__device__ __forceinline__ Top2 merge_result(const MergeSlot& m) {
  Top2 r;
  r.best = unord_f32((uint32_t)(m.best >> 32));
  r.scan = (int)(0xffffffffu - (uint32_t)(m.best & 0xffffffffull));
  r.second = unord_f32(m.second);
  return r;
}

// Fast path: all base and search masks of the tile are on. Tile `tile` of
// the search `a`, run by workgroup `wg` of its launch.
template <class T, class E>
__device__ __forceinline__ void tile_fast(const TileArgs& a, int tile, int wg) {
  extern __shared__ float4 dyn4[];
  float* win = reinterpret_cast<float*>(dyn4);
  __shared__ float4 xs4[T::NB * 16];
  __shared__ float bsxx[T::NB], bsum[T::NB];
  __shared__ int bvalid[T::NB];
  __shared__ MergeSlot mrg[T::NB];
  __shared__ int s_any;
  __shared__ unsigned long long s_bad, s_valid;
  __shared__ unsigned long long zrow[kMaxWindow];  // off-mask rows per window plane
  __shared__ float bref[T::NB];  // per-block shift: the window mean at the block's position

  const BMGrid& g = a.g;
  const int nx = g.dims[0], ny = g.dims[1], nz = g.dims[2];
  const int tid = threadIdx.x;
  int bx, by0, bz0;
  tile_origin<T>(a, tile, bx, by0, bz0);
  const int NG = g.ng;
  const int WY = T::fast_pitch((T::TY - 1) * NG + T::WX + T::YPAD);
  const int WZ = (T::TZ - 1) * NG + T::WX;
  const int px = g.start[0] + NG * bx;
  const int py0 = g.start[1] + NG * by0;
  const int pz0 = g.start[2] + NG * bz0;
  const int wx0 = px - T::H, wy0 = py0 - T::H, wz0 = pz0 - T::H;

  if (tid == 0) {
    s_any = 0;
    s_bad = 0ull;
    s_valid = 0ull;
  }
  merge_init(mrg, T::NB);
  for (int i = tid; i < WZ; i += kThreads) zrow[i] = 0ull;
  __syncthreads();
  if (tid < T::NB) {
    const int gy = by0 + tid % T::TY, gz = bz0 + tid / T::TY;
    const int64_t b = bx + (int64_t)g.nb[0] * (gy + (int64_t)g.nb[1] * gz);
    int st = block_state(a, gy, gz, b);
    if (st == 1 && (a.bflags[b] & 2u)) {  // base mask off everywhere: no search
      write_all_off<T>(a, b, px, py0 + NG * (tid % T::TY), pz0 + NG * (tid / T::TY));
      st = 4;
    }
    bvalid[tid] = st;
    if (st == 1) {
      s_any = 1;
      atomicOr(&s_valid, 1ull << tid);
      bsxx[tid] = a.sxx[b];
      bsum[tid] = a.sumxt[b];
      bref[tid] = a.stats[lin3(nx, ny, px, py0 + NG * (tid % T::TY), pz0 + NG * (tid / T::TY))].x;
      if ((a.bflags[b] & 1u) == 0u) atomicOr(&s_bad, 1ull << tid);  // base partly masked
    } else if (st == 2) {
      write_skipped(a, b, px, py0 + NG * (tid % T::TY), pz0 + NG * (tid / T::TY));
    }
  }
  __syncthreads();
  if (!s_any) return;  // every block is skipped or outside the grid (workgroup-uniform)
  // Stage the union search window as it is (zeros outside the image); each
  // lane shifts the rows it loads by its block's own reference, so that the
  // accumulated products stay of the size of the contrast around the block
  // (a shift shared by the tile leaves products of the size of the
  // intensity step between a block's region and the shift, whose float
  // rounding is large next to the Sxy of a low-contrast block). A
  // search-mask-off voxel sends the blocks whose own window holds it to the
  // masked kernel; the other blocks stay here.
  stage_window<T, false, E>(a, win, nullptr, zrow, WY, WZ, wx0, wy0, wz0, 0.0f);
  for (int i = tid; i < T::NB * 16; i += kThreads) {
    const int bl = i / 16;
    const int gy = by0 + bl % T::TY, gz = bz0 + bl / T::TY;
    if (gy < g.nb[1] && gz < g.nb[2]) {
      const int64_t b = bx + (int64_t)g.nb[0] * (gy + (int64_t)g.nb[1] * gz);
      xs4[i] = reinterpret_cast<const float4*>(a.xt)[b * 16 + (i % 16)];
    }
  }
  __syncthreads();
  if (tid < T::NB && ((s_valid >> tid) & 1ull) && window_has_off<T>(zrow, tid, NG))
    atomicOr(&s_bad, 1ull << tid);

  __syncthreads();
  const unsigned long long todo = s_bad & s_valid;
  if (todo) {
    if (tid == 0) {
      const int q = atomicAdd(a.qcount, 1);
      a.queue[q] = tile;
      a.qbits[q] = todo;
    }
    if (todo == s_valid) return;  // workgroup-uniform: nothing left here
    if (tid < T::NB && ((todo >> tid) & 1ull)) bvalid[tid] = 3;
    __syncthreads();
  }
  if (a.work && tid == 0)
    atomicAdd(a.work + wg % kWorkStripe, tile_evals<T>(a, px, py0, pz0, bvalid));

  // Workgroup-uniform: every candidate window of the tile's blocks (those
  // inside the grid or not) lies inside the image.
  const bool interior = px - T::H >= 0 && px - T::H + T::NW - 1 <= nx - T::BW &&
                        py0 - T::H >= 0 &&
                        py0 + NG * (T::TY - 1) - T::H + T::NW - 1 <= ny - T::BW &&
                        pz0 - T::H >= 0 && pz0 + NG * (T::TZ - 1) - T::H + T::NW - 1 <= nz - T::BW;
  Top2 t2;
  top2_init(t2);
  const int bl = tid / T::TPB;
  if (tid < T::NT && bvalid[bl] == 1) {
    const int r = tid % T::TPB;  // consecutive lanes: consecutive cz
    const int cz = r % T::NW;
    const int cyg = r / T::NW;
    const int bly = bl % T::TY, blz = bl / T::TY;
    float acc[T::KY][T::NW];
#pragma unroll
    for (int ky = 0; ky < T::KY; ++ky)
#pragma unroll
      for (int cx = 0; cx < T::NW; ++cx) acc[ky][cx] = 0.0f;
    const float* rowbase = win + ((blz * NG + cz) * WY + (bly * NG + cyg * T::KY)) * T::PX;
    const float rb = bref[bl];
#pragma unroll 1
    for (int iz = 0; iz < 4; ++iz) {
      // One z-plane of the centred block (16 values) stays in VGPRs.
      float xp[16];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const float4 q = xs4[bl * 16 + iz * 4 + k];
        xp[4 * k + 0] = q.x;
        xp[4 * k + 1] = q.y;
        xp[4 * k + 2] = q.z;
        xp[4 * k + 3] = q.w;
      }
      const float* plane = rowbase + iz * WY * T::PX;
#pragma unroll
      for (int rr = 0; rr < T::KY + 3; ++rr) {
        float row[T::PX];
        const float4* rp = reinterpret_cast<const float4*>(plane + rr * T::PX);
#pragma unroll
        for (int v = 0; v < T::NV; ++v) {
          const float4 q = rp[v];
          row[4 * v + 0] = q.x - rb;
          row[4 * v + 1] = q.y - rb;
          row[4 * v + 2] = q.z - rb;
          row[4 * v + 3] = q.w - rb;
        }
#pragma unroll
        for (int ky = 0; ky < T::KY; ++ky) {
          const int iy = rr - ky;
          if (iy < 0 || iy > 3) continue;
#pragma unroll
          for (int cx = 0; cx < T::NW; ++cx)
#pragma unroll
            for (int ix = 0; ix < 4; ++ix)
              acc[ky][cx] = fmaf(xp[iy * 4 + ix], row[cx + ix], acc[ky][cx]);
        }
      }
    }
    // Epilogue: scores of the lane's candidates, in scan order. With
    // t = Sxy^2 / Sxx, the score is max(t * iyy, -2): iyy = 1 / Syy, or -inf
    // for a window with Syy = 0, whose product is -inf or NaN, both taken to
    // -2 by fmaxf. A block with Sxx = 0 scores -2 everywhere: 1 / Sxx is
    // then NaN.
    const int py = py0 + NG * bly, pz = pz0 + NG * blz;
    const float sxx = bsxx[bl], sumx = bsum[bl];
    const float isxx = sxx > 0.0f ? __builtin_amdgcn_rcpf(sxx) : __builtin_nanf("");
    if (interior) {
      // Every candidate of every block of the tile lies inside the image:
      // no range checks, and the statistics of a row are read at constant
      // offsets from one address.
#pragma unroll
      for (int ky = 0; ky < T::KY; ++ky) {
        const int cy = cyg * T::KY + ky;
        const float2* srow =
            a.stats + lin3(nx, ny, px - T::H, min(py - T::H + cy, ny - T::BW), pz - T::H + cz);
#pragma unroll
        for (int cx = 0; cx < T::NW; ++cx) {
          const float2 st = srow[cx];
          const float sxy = fmaf(-(st.x - rb), sumx, acc[ky][cx]);
          float sc = fmaxf(sxy * sxy * isxx * st.y, -2.0f);
          if (T::YPAD > 0 && cy >= T::NW) sc = -__builtin_huge_valf();
          top2_step(t2, sc, (cz * T::NW + cy) * T::NW + cx);
        }
      }
    } else {
      const int lox = max(px - T::H, 0), hix = min(px - T::H + T::NW - 1, nx - T::BW);
      const int loy = max(py - T::H, 0), hiy = min(py - T::H + T::NW - 1, ny - T::BW);
      const int loz = max(pz - T::H, 0), hiz = min(pz - T::H + T::NW - 1, nz - T::BW);
      // Candidates outside the image score -inf, which changes neither the
      // best nor the second (the scores start at -3); their statistics are
      // read at a clamped position.
      const int czg = pz - T::H + cz;
      const bool zok = czg >= loz && czg <= hiz;
      const int czc = min(max(czg, 0), nz - T::BW);
#pragma unroll
      for (int ky = 0; ky < T::KY; ++ky) {
        const int cy = cyg * T::KY + ky;
        const int cyv = py - T::H + cy;
        const bool yok = zok && cy < T::NW && cyv >= loy && cyv <= hiy;
        const float2* srow = a.stats + lin3(nx, ny, 0, min(max(cyv, 0), ny - T::BW), czc);
#pragma unroll
        for (int cx = 0; cx < T::NW; ++cx) {
          const int cxg = px - T::H + cx;
          const float2 st = srow[min(max(cxg, 0), nx - T::BW)];
          const float sxy = fmaf(-(st.x - rb), sumx, acc[ky][cx]);
          float sc = fmaxf(sxy * sxy * isxx * st.y, -2.0f);
          sc = (yok && cxg >= lox && cxg <= hix) ? sc : -__builtin_huge_valf();
          top2_step(t2, sc, (cz * T::NW + cy) * T::NW + cx);
        }
      }
    }
  }
  if (tid < T::NT) merge_phase1(mrg, bl, t2);
  __syncthreads();
  if (tid < T::NT) merge_phase2(mrg, bl, t2);
  __syncthreads();
  if (tid < T::NB && bvalid[tid] == 1) {
    const int gy = by0 + tid % T::TY, gz = bz0 + tid / T::TY;
    const int64_t b = bx + (int64_t)g.nb[0] * (gy + (int64_t)g.nb[1] * gz);
    write_block_result<T>(a, b, px, py0 + NG * (tid % T::TY), pz0 + NG * (tid / T::TY),
                          merge_result(mrg[tid]));
  }
}

// Search mask values 4v .. 4v+3 of the staged row at float offset ro.
template <class T>
__device__ __forceinline__ void load_mask(const uint8_t* wv, int ro, int v, float* vr) {
  const uint32_t m = reinterpret_cast<const uint32_t*>(wv + ro)[v];
  vr[4 * v + 0] = (float)(m & 0xffu);
  vr[4 * v + 1] = (float)((m >> 8) & 0xffu);
  vr[4 * v + 2] = (float)((m >> 16) & 0xffu);
  vr[4 * v + 3] = (float)(m >> 24);
}

// Masked path: sums over voxels where both masks are on. The window is
// staged as v * y (v = search mask) plus the mask bytes.
// With t = v (y - mu), one FMA per voxel (v * y - v * mu, rounded once):
//  * Base block fully on (the usual case at image borders): Sxy = Sum x~ t,
//    Syy = Sum t^2, Sxx = Sum v x~^2 (4 FMAs per voxel). Syy is summed from
//    the same float t as Sxy rather than taken from the window statistic:
//    with few voxels on, r^2 is a ratio of nearly equal terms and both must
//    carry the same centring (one voxel on gives exactly 1).
//  * Base block partly masked (base mask b): Sxy = Sum b x~ t,
//    Syy = Sum b t^2, Sxx = Sum b x~^2 v (5 ops per voxel).
template <class T, class E>
__device__ __forceinline__ void tile_masked(const TileArgs& a, int wg) {
  constexpr int kRows = T::NW * T::NW;  // row tasks per block
  extern __shared__ float4 dyn4[];
  __shared__ float rbest[T::NB * kRows], rsecond[T::NB * kRows];  // per-task records
  __shared__ int rscan[T::NB * kRows];
  __shared__ uint16_t live_list[T::NB * kRows];  // tasks that need arithmetic
  __shared__ unsigned int pat_cnt[kPatBuckets / 2];  // counts, then next list slots
  __shared__ int s_nlive;
  __shared__ float4 xh4[T::NB * 16];  // base-masked centred blocks
  __shared__ unsigned long long bbits[T::NB];  // base mask bit per voxel
  __shared__ int todo_list[T::NB];
  __shared__ int s_ntodo;
  __shared__ unsigned long long s_fullbase;  // queued blocks whose base mask is fully on
  __shared__ unsigned long long zon[kMaxWindow];   // rows holding a search-mask-on voxel
  __shared__ unsigned long long zbad[kMaxWindow];  // rows holding an off or outside voxel
  __shared__ float bsxx[T::NB], bsum[T::NB];       // block sums of fully-on bases
  __shared__ int s_nfull;
  __shared__ int bvalid[T::NB];
  __shared__ float bref[T::NB];  // per-block shift of the full rows

  // Workgroup (q, part) searches every msplit-th queued block of tile q
  // (more, shorter workgroups where the level has few tiles).
  const int q = wg / a.msplit, part = wg % a.msplit;
  if (q >= *a.qcount) return;
  const int tile = a.queue[q];
  const unsigned long long queued = a.qbits[q];
  unsigned long long todo = 0ull;  // blocks to search here
  for (int bl = 0, c = 0; bl < T::NB; ++bl) {
    if (((queued >> bl) & 1ull) == 0ull) continue;
    if (c % a.msplit == part) todo |= 1ull << bl;
    ++c;
  }
  if (todo == 0ull) return;
  const BMGrid& g = a.g;
  const int nx = g.dims[0], ny = g.dims[1], nz = g.dims[2];
  const int tid = threadIdx.x;
  int bx, by0, bz0;
  tile_origin<T>(a, tile, bx, by0, bz0);
  const int NG = g.ng;
  const int WY = (T::TY - 1) * NG + T::WX;
  const int WZ = (T::TZ - 1) * NG + T::WX;
  const int px = g.start[0] + NG * bx;
  const int py0 = g.start[1] + NG * by0;
  const int pz0 = g.start[2] + NG * bz0;
  const int wx0 = px - T::H, wy0 = py0 - T::H, wz0 = pz0 - T::H;
  const int total = T::PX * WY * WZ;
  float* wy_ = reinterpret_cast<float*>(dyn4);
  uint8_t* wv_ = reinterpret_cast<uint8_t*>(wy_ + total);  // search mask bytes

  if (tid < T::NB) {
    bvalid[tid] = ((todo >> tid) & 1ull) ? 1 : 0;
    bbits[tid] = 0ull;
  }
  for (int i = tid; i < WZ; i += kThreads) {
    zon[i] = 0ull;
    zbad[i] = 0ull;
  }
  for (int i = tid; i < kPatBuckets / 2; i += kThreads) pat_cnt[i] = 0u;
  if (tid == 0) {
    s_fullbase = 0ull;
    s_nlive = 0;
    s_nfull = 0;
  }
  __syncthreads();
  if (tid < T::NB && ((todo >> tid) & 1ull)) {
    const int gy = by0 + tid % T::TY, gz = bz0 + tid / T::TY;
    const int64_t b = bx + (int64_t)g.nb[0] * (gy + (int64_t)g.nb[1] * gz);
    if (a.bm == nullptr || (a.bflags[b] & 1u) != 0u) {
      atomicOr(&s_fullbase, 1ull << tid);
      bsxx[tid] = a.sxx[b];
      bsum[tid] = a.sumxt[b];
      bref[tid] = a.stats[lin3(nx, ny, px, py0 + NG * (tid % T::TY), pz0 + NG * (tid / T::TY))].x;
    }
  }
  __syncthreads();
  if (tid == 0) {
    // The queued blocks, densely numbered: fully-on bases first, then the
    // partly masked ones, so that few waves run both task loops.
    const unsigned long long fb = s_fullbase;
    int n = 0;
    for (int bl = 0; bl < T::NB; ++bl)
      if (((todo & fb) >> bl) & 1ull) todo_list[n++] = bl;
    for (int bl = 0; bl < T::NB; ++bl)
      if (((todo & ~fb) >> bl) & 1ull) todo_list[n++] = bl;
    s_ntodo = n;
  }
  __syncthreads();
  // Staged as v * y; full rows shift their loads by the block's reference
  // as the fast kernel does, masked rows centre each voxel on the
  // candidate's own mean with one rounding.
  stage_window<T, true, E>(a, wy_, wv_, zon, WY, WZ, wx0, wy0, wz0, 0.0f, zbad);
  for (int i = tid; i < T::NB * 64; i += kThreads) {
    const int bl = i / 64, k = i % 64;
    const int gy = by0 + bl % T::TY, gz = bz0 + bl / T::TY;
    float xh = 0.0f;
    if ((todo >> bl) & 1ull) {
      const int64_t b = bx + (int64_t)g.nb[0] * (gy + (int64_t)g.nb[1] * gz);
      bool on = true;
      if (a.bm != nullptr && (a.bflags[b] & 1u) == 0u) {
        const int qy = py0 + NG * (bl % T::TY), qz = pz0 + NG * (bl / T::TY);
        on = a.bm[lin3(nx, ny, px + (k & 3), qy + ((k >> 2) & 3), qz + (k >> 4))] != 0;
      }
      if (on) {
        xh = a.xt[b * 64 + k];
        atomicOr(&bbits[bl], 1ull << k);
      }
    }
    reinterpret_cast<float*>(xh4)[i] = xh;
  }
  __syncthreads();

  if (a.work && tid == 0)
    atomicAdd(a.work + wg % kWorkStripe, tile_evals<T>(a, px, py0, pz0, bvalid));
  // Row tasks (block k, cz, cy) of the queued blocks, numbered
  // k * NW^2 + cz * NW + cy. Each task's top-2 record goes to LDS: empty for
  // rows outside the image, -2 at every candidate for rows whose window holds
  // no search-mask-on voxel (Syy = 0), and otherwise computed by a thread.
  // Rows of a fully-on base block whose whole window is inside the image and
  // search-mask on ("full" rows) are scored as the fast kernel scores them
  // (one FMA per voxel-eval, block and window statistics); the others take
  // the masked sums. The rows to compute are compacted, full rows from the
  // front of the list and masked rows from the back, and dealt to the
  // threads strided, so nearly every lane works and few waves run both
  // paths; then one wave per block appends the block's records in scan
  // order.
  const int ntodo = s_ntodo;
  const int ntasks = ntodo * kRows;
  const int lox = max(px - T::H, 0), hix = min(px - T::H + T::NW - 1, nx - T::BW);
  for (int base = 0; base < ntasks; base += kThreads) {  // workgroup-uniform
    const int i = base + tid;
    bool live = false, full = false;
    unsigned int pm = 0u, rm = 0u;  // planes and rows of the task's window with an on voxel
    if (i < ntasks) {
      const int k = i / kRows, r = i % kRows;
      const int bl = todo_list[k];
      const int cz = r / T::NW, cy = r % T::NW;
      const int bly = bl % T::TY, blz = bl / T::TY;
      const int py = py0 + NG * bly, pz = pz0 + NG * blz;
      const int czg = pz - T::H + cz, cyg = py - T::H + cy;
      Top2 tr;
      top2_init(tr);
      if (czg >= max(pz - T::H, 0) && czg <= min(pz - T::H + T::NW - 1, nz - T::BW) &&
          cyg >= max(py - T::H, 0) && cyg <= min(py - T::H + T::NW - 1, ny - T::BW)) {
#pragma unroll
        for (int iz = 0; iz < 4; ++iz) {
          const unsigned int rp = (unsigned int)(zon[blz * NG + cz + iz] >> (bly * NG + cy)) & 0xfu;
          rm |= rp;
          pm |= (rp != 0u ? 1u : 0u) << iz;
        }
        if (rm != 0u) {
          live = true;
          if (((s_fullbase >> bl) & 1ull) && bbits[bl] == ~0ull) {
            unsigned long long rows_bad = 0ull;
#pragma unroll
            for (int iz = 0; iz < 4; ++iz) rows_bad |= zbad[blz * NG + cz + iz];
            full = ((rows_bad >> (bly * NG + cy)) & 0xfull) == 0ull;
          }
        } else {
#pragma unroll
          for (int cx = 0; cx < T::NW; ++cx) {
            const int cxg = px - T::H + cx;
            if (cxg < lox || cxg > hix) continue;
            top2_step(tr, -2.0f, (cz * T::NW + cy) * T::NW + cx);
          }
        }
      }
      if (!live) {
        rbest[i] = tr.best;
        rsecond[i] = tr.second;
        rscan[i] = tr.scan;
      }
      if (live && !full) {
        // A window plane or row without an on voxel, or a base plane or row
        // without one, adds exact zeros to every sum: the task skips it. Until
        // the task is computed its record's scan slot holds INT_MIN + pattern
        // (records hold scans >= 0 or INT_MAX; full tasks hold 0 until then).
        const unsigned long long bits = bbits[bl];
        unsigned int bpm = 0u, brm = 0u;
#pragma unroll
        for (int iz = 0; iz < 4; ++iz) {
          const unsigned int pl = (unsigned int)(bits >> (16 * iz)) & 0xffffu;
          bpm |= (pl != 0u ? 1u : 0u) << iz;
#pragma unroll
          for (int iy = 0; iy < 4; ++iy) brm |= (((pl >> (4 * iy)) & 0xfu) != 0u ? 1u : 0u) << iy;
        }
        const unsigned int pat = (pm & bpm) | ((rm & brm) << 4) | (bits == ~0ull ? 0u : 256u);
        rscan[i] = (int)(0x80000000u | pat);
        atomicAdd(&pat_cnt[pat >> 1], 1u << (16 * (pat & 1u)));
      } else if (full) {
        rscan[i] = 0;
      }
    }
    const unsigned long long balf = __ballot(full);
    const unsigned long long balm = __ballot(live && !full);
    int fbase = 0, mbase = 0;
    if ((tid & 63) == 0) {
      if (balf != 0ull) fbase = atomicAdd(&s_nfull, (int)__popcll(balf));
      if (balm != 0ull) mbase = atomicAdd(&s_nlive, (int)__popcll(balm));
    }
    fbase = __shfl(fbase, 0, 64);
    mbase = __shfl(mbase, 0, 64);
    const unsigned long long below = (1ull << (tid & 63)) - 1ull;
    if (full) live_list[fbase + __popcll(balf & below)] = (uint16_t)i;
    (void)mbase;
  }
  __syncthreads();
  const int nfull = s_nfull, nmasked = s_nlive;
  const int nlive = nfull + nmasked;
  // The masked tasks go to the back of the list grouped by pattern (counting
  // sort: fully-on bases first), so that the lanes of a wave mostly skip the
  // same planes and rows. Their order does not change any record.
  if (nmasked > 0) {
    if (tid < 64) {  // exclusive prefix of the bucket counts, 8 buckets a lane
      constexpr int kPer = kPatBuckets / 64;
      unsigned int c[kPer], s = 0u;
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        c[u] = (pat_cnt[(tid * kPer + u) >> 1] >> (16 * (u & 1))) & 0xffffu;
        s += c[u];
      }
      unsigned int inc = s;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const unsigned int t = __shfl_up(inc, o, 64);
        if (tid >= o) inc += t;
      }
      unsigned int at = (unsigned int)(T::NB * kRows - nmasked) + (inc - s);
#pragma unroll
      for (int u = 0; u < kPer; u += 2) {
        pat_cnt[(tid * kPer + u) >> 1] = at | ((at + c[u]) << 16);
        at += c[u] + c[u + 1];
      }
    }
    __syncthreads();
    for (int i = tid; i < ntasks; i += kThreads) {
      const int m = rscan[i];
      if (m < 0) {
        const unsigned int pat = (unsigned int)m & 0x1ffu;
        const unsigned int sh = 16 * (pat & 1u);
        const unsigned int old = atomicAdd(&pat_cnt[pat >> 1], 1u << sh);
        live_list[(old >> sh) & 0xffffu] = (uint16_t)i;
      }
    }
    __syncthreads();
  }
#pragma unroll 1
  for (int j = tid; j < nlive; j += kThreads) {
    const bool full = j < nfull;
    const int i = live_list[full ? j : T::NB * kRows - nmasked + (j - nfull)];
    const int k = i / kRows, r = i % kRows;
    const int bl = todo_list[k];
    const int cz = r / T::NW, cy = r % T::NW;
    const int bly = bl % T::TY, blz = bl / T::TY;
    const int py = py0 + NG * bly, pz = pz0 + NG * blz;
    const int czg = pz - T::H + cz, cyg = py - T::H + cy;
    Top2 tr;
    top2_init(tr);
    const float2* srow = a.stats + lin3(nx, ny, 0, cyg, czg);
    const int rowoff = ((blz * NG + cz) * WY + (bly * NG + cy)) * T::PX;
    if (full) {
      // Every voxel of every candidate window of this row is on and the
      // base block is fully on: Sxy = Sum x~ (y - rb) - (mu - rb) Sum x~,
      // Sxx and Syy from the block and window statistics.
      const float rb = bref[bl];
      float acc[T::NW];
#pragma unroll
      for (int cx = 0; cx < T::NW; ++cx) acc[cx] = 0.0f;
#pragma unroll 1
      for (int iz = 0; iz < 4; ++iz) {
#pragma unroll
        for (int iy = 0; iy < 4; ++iy) {
          const int ro = rowoff + (iz * WY + iy) * T::PX;
          float yr[T::PX];
          const float4* yp = reinterpret_cast<const float4*>(wy_ + ro);
#pragma unroll
          for (int v = 0; v < T::NV; ++v) {
            const float4 q = yp[v];
            yr[4 * v + 0] = q.x - rb; yr[4 * v + 1] = q.y - rb;
            yr[4 * v + 2] = q.z - rb; yr[4 * v + 3] = q.w - rb;
          }
          const float4 xq4 = xh4[bl * 16 + iz * 4 + iy];
          const float xh[4] = {xq4.x, xq4.y, xq4.z, xq4.w};
#pragma unroll
          for (int cx = 0; cx < T::NW; ++cx)
#pragma unroll
            for (int ix = 0; ix < 4; ++ix) acc[cx] = fmaf(xh[ix], yr[cx + ix], acc[cx]);
        }
      }
      const float sxx = bsxx[bl], sumx = bsum[bl];
      const float isxx = sxx > 0.0f ? __builtin_amdgcn_rcpf(sxx) : __builtin_nanf("");
#pragma unroll
      for (int cx = 0; cx < T::NW; ++cx) {
        const int cxg = px - T::H + cx;
        if (cxg < lox || cxg > hix) continue;
        const float2 st = srow[cxg];
        const float sxy = fmaf(-(st.x - rb), sumx, acc[cx]);
        top2_push(tr, fmaxf(sxy * sxy * isxx * st.y, -2.0f), (cz * T::NW + cy) * T::NW + cx);
      }
      rbest[i] = tr.best;
      rsecond[i] = tr.second;
      rscan[i] = tr.scan;
      continue;
    }
    float dmu[T::NW];
#pragma unroll
    for (int cx = 0; cx < T::NW; ++cx) {
      const int cxg = min(max(px - T::H + cx, lox), hix);
      dmu[cx] = srow[cxg].x;
    }
    const unsigned long long bits = bbits[bl];
    const unsigned int pat = (unsigned int)rscan[i] & 0x1ffu;  // set above
    float sxy[T::NW], sxx[T::NW], s3[T::NW];
#pragma unroll
    for (int cx = 0; cx < T::NW; ++cx) sxy[cx] = sxx[cx] = s3[cx] = 0.0f;
    if (bits == ~0ull) {
#pragma unroll 1
      for (int iz = 0; iz < 4; ++iz) {
        if (((pat >> iz) & 1u) == 0u) continue;
#pragma unroll
        for (int iy = 0; iy < 4; ++iy) {
          if (((pat >> (4 + iy)) & 1u) == 0u) continue;
          const int ro = rowoff + (iz * WY + iy) * T::PX;
          float yr[T::PX], vr[T::PX];
          const float4* yp = reinterpret_cast<const float4*>(wy_ + ro);
#pragma unroll
          for (int v = 0; v < T::NV; ++v) {
            const float4 q = yp[v];
            yr[4 * v + 0] = q.x; yr[4 * v + 1] = q.y; yr[4 * v + 2] = q.z; yr[4 * v + 3] = q.w;
            load_mask<T>(wv_, ro, v, vr);
          }
          const float4 xq4 = xh4[bl * 16 + iz * 4 + iy];
          const float xh[4] = {xq4.x, xq4.y, xq4.z, xq4.w};
          float xs[4];
#pragma unroll
          for (int ix = 0; ix < 4; ++ix) xs[ix] = xh[ix] * xh[ix];
#pragma unroll
          for (int cx = 0; cx < T::NW; ++cx)
#pragma unroll
            for (int ix = 0; ix < 4; ++ix) {
              const float tm = fmaf(-vr[cx + ix], dmu[cx], yr[cx + ix]);
              sxy[cx] = fmaf(xh[ix], tm, sxy[cx]);
              s3[cx] = fmaf(tm, tm, s3[cx]);
              sxx[cx] = fmaf(xs[ix], vr[cx + ix], sxx[cx]);
            }
        }
      }
#pragma unroll
      for (int cx = 0; cx < T::NW; ++cx) {
        const int cxg = px - T::H + cx;
        if (cxg < lox || cxg > hix) continue;
        top2_push(tr, ncc_score(sxy[cx], sxx[cx], s3[cx]), (cz * T::NW + cy) * T::NW + cx);
      }
    } else {
      // s3 = Sum b t^2 (Syy over both masks)
#pragma unroll 1
      for (int iz = 0; iz < 4; ++iz) {
        if (((pat >> iz) & 1u) == 0u) continue;
#pragma unroll
        for (int iy = 0; iy < 4; ++iy) {
          if (((pat >> (4 + iy)) & 1u) == 0u) continue;
          const int ro = rowoff + (iz * WY + iy) * T::PX;
          float yr[T::PX], vr[T::PX];
          const float4* yp = reinterpret_cast<const float4*>(wy_ + ro);
#pragma unroll
          for (int v = 0; v < T::NV; ++v) {
            const float4 q = yp[v];
            yr[4 * v + 0] = q.x; yr[4 * v + 1] = q.y; yr[4 * v + 2] = q.z; yr[4 * v + 3] = q.w;
            load_mask<T>(wv_, ro, v, vr);
          }
          const float4 xq4 = xh4[bl * 16 + iz * 4 + iy];
          const float xh[4] = {xq4.x, xq4.y, xq4.z, xq4.w};
          const uint32_t nib = (uint32_t)(bits >> (iz * 16 + iy * 4)) & 0xfu;
          float bb[4], xs[4];
#pragma unroll
          for (int ix = 0; ix < 4; ++ix) {
            bb[ix] = (float)((nib >> ix) & 1u);
            xs[ix] = xh[ix] * xh[ix];
          }
#pragma unroll
          for (int cx = 0; cx < T::NW; ++cx)
#pragma unroll
            for (int ix = 0; ix < 4; ++ix) {
              const float tm = fmaf(-vr[cx + ix], dmu[cx], yr[cx + ix]);
              sxy[cx] = fmaf(xh[ix], tm, sxy[cx]);
              s3[cx] = fmaf(tm * bb[ix], tm, s3[cx]);
              sxx[cx] = fmaf(xs[ix], vr[cx + ix], sxx[cx]);
            }
        }
      }
#pragma unroll
      for (int cx = 0; cx < T::NW; ++cx) {
        const int cxg = px - T::H + cx;
        if (cxg < lox || cxg > hix) continue;
        top2_push(tr, ncc_score(sxy[cx], sxx[cx], s3[cx]), (cz * T::NW + cy) * T::NW + cx);
      }
    }
    rbest[i] = tr.best;
    rsecond[i] = tr.second;
    rscan[i] = tr.scan;
  }
  __syncthreads();
  // Wave w merges blocks w, w + 4, ...: lane L appends rows
  // [L*C, L*C + C) in order, then lanes combine pairwise in order.
  constexpr int C = (kRows + 63) / 64;
  const int lane = tid & 63;
  for (int k = tid >> 6; k < ntodo; k += kThreads / 64) {  // wave-uniform
    Top2 acc;
    top2_init(acc);
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int r = lane * C + c;
      if (r < kRows) {
        Top2 t;
        t.best = rbest[k * kRows + r];
        t.second = rsecond[k * kRows + r];
        t.scan = rscan[k * kRows + r];
        acc = top2_append(acc, t);
      }
    }
#pragma unroll
    for (int sh = 1; sh < 64; sh <<= 1) {
      Top2 o;
      o.best = __shfl_down(acc.best, sh, 64);
      o.second = __shfl_down(acc.second, sh, 64);
      o.scan = __shfl_down(acc.scan, sh, 64);
      if ((lane & (2 * sh - 1)) == 0 && lane + sh < 64) acc = top2_append(acc, o);
    }
    if (lane == 0) {
      const int bl = todo_list[k];
      const int gy = by0 + bl % T::TY, gz = bz0 + bl / T::TY;
      const int64_t b = bx + (int64_t)g.nb[0] * (gy + (int64_t)g.nb[1] * gz);
      write_block_result<T>(a, b, px, py0 + NG * (bl % T::TY), pz0 + NG * (bl / T::TY), acc);
    }
  }
}

template <class T, class E>
__global__ __launch_bounds__(kThreads, 4) void ncc_tile_fast(TileArgs a) {
  tile_fast<T, E>(a, xcd_tile(blockIdx.x, gridDim.x), blockIdx.x);
}

template <class T, class E>
__global__ __launch_bounds__(kThreads, T::MWG) void ncc_tile_masked(TileArgs a) {
  tile_masked<T, E>(a, blockIdx.x);
}

// The search of several pairs in one launch: workgroups are numbered over a
// flat list of every pair's work items (tiles, or queued-tile parts), pair p
// owning items [start[p], start[p + 1]). The pairs share the tile shape and
// hence the LDS layout; each has its own arguments (images, grid, queues).
__device__ __forceinline__ int pair_of(const int* __restrict__ start, int n, int item) {
  int lo = 0, hi = n - 1;  // the last p with start[p] <= item
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= item) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <class T, class E>
__global__ __launch_bounds__(kThreads, 4) void ncc_tile_fast_multi(
    const TileArgs* __restrict__ args, const int* __restrict__ start, int n) {
  const int item = xcd_tile(blockIdx.x, gridDim.x);
  const int p = pair_of(start, n, item);
  const TileArgs a = args[p];
  tile_fast<T, E>(a, item - start[p], blockIdx.x);
}

template <class T, class E>
__global__ __launch_bounds__(kThreads, T::MWG) void ncc_tile_masked_multi(
    const TileArgs* __restrict__ args, const int* __restrict__ start, int n) {
  const int p = pair_of(start, n, blockIdx.x);
  const TileArgs a = args[p];
  tile_masked<T, E>(a, blockIdx.x - start[p]);
}

__global__ void reset_counters(int32_t* c, int32_t* d) {
  *c = 0;
  if (d) *d = 0;
}

// ---------------------------------------------------------------------------
// Generic kernel: any BW, BG, NW. One workgroup per block; candidates in
// scan order strided over the threads; double two-pass sums as the reference.
__global__ __launch_bounds__(kThreads) void ncc_generic_kernel(
    const float* __restrict__ x, const uint8_t* __restrict__ bm,
    const float* __restrict__ y, const uint8_t* __restrict__ sm, BMGrid g,
    float* out_best, float* out_second, int32_t* out_match, const uint32_t* flags) {
  __shared__ float tb[kThreads], ts[kThreads];
  __shared__ int ti[kThreads];
  const int64_t b = blockIdx.x;
  const int nx = g.dims[0], ny = g.dims[1];
  const int bxi = (int)(b % g.nb[0]);
  const int byi = (int)((b / g.nb[0]) % g.nb[1]);
  const int bzi = (int)(b / ((int64_t)g.nb[0] * g.nb[1]));
  const int p[3] = {g.start[0] + g.ng * bxi, g.start[1] + g.ng * byi, g.start[2] + g.ng * bzi};
  if (flags && (flags[b] & kBlockSkip)) {  // dropped by the variance selection
    if (threadIdx.x == 0) {
      out_best[b] = -3.0f;
      out_second[b] = -3.0f;
      for (int k = 0; k < 3; ++k) out_match[3 * b + k] = p[k];
    }
    return;
  }
  int first[3], cnt[3];
  for (int k = 0; k < 3; ++k) cand_range(p[k], g.dims[k], g.bw, g.bg, g.nw, &first[k], &cnt[k]);
  const int bw = g.bw;
  const double n = (double)bw * bw * bw;
  double mb = 0.0;
  for (int dz = 0; dz < bw; ++dz)
    for (int dy = 0; dy < bw; ++dy)
      for (int dx = 0; dx < bw; ++dx) mb += x[lin3(nx, ny, p[0] + dx, p[1] + dy, p[2] + dz)];
  mb /= n;
  Top2 t2;
  top2_init(t2);
  const int total = cnt[0] * cnt[1] * cnt[2];
  for (int k = threadIdx.x; k < total; k += kThreads) {
    const int kx = k % cnt[0], ky = (k / cnt[0]) % cnt[1], kz = k / (cnt[0] * cnt[1]);
    const int s[3] = {first[0] + kx * g.bg, first[1] + ky * g.bg, first[2] + kz * g.bg};
    double ms = 0.0;
    for (int dz = 0; dz < bw; ++dz)
      for (int dy = 0; dy < bw; ++dy)
        for (int dx = 0; dx < bw; ++dx) ms += y[lin3(nx, ny, s[0] + dx, s[1] + dy, s[2] + dz)];
    ms /= n;
    double sp = 0.0, sb = 0.0, ss = 0.0;
    for (int dz = 0; dz < bw; ++dz)
      for (int dy = 0; dy < bw; ++dy)
        for (int dx = 0; dx < bw; ++dx) {
          const int64_t ob = lin3(nx, ny, p[0] + dx, p[1] + dy, p[2] + dz);
          const int64_t os = lin3(nx, ny, s[0] + dx, s[1] + dy, s[2] + dz);
          if ((bm && bm[ob] == 0) || (sm && sm[os] == 0)) continue;
          const double u = (double)x[ob] - mb, v = (double)y[os] - ms;
          sp += u * v;
          sb += u * u;
          ss += v * v;
        }
    const float sc = (sb <= 0.0 || ss <= 0.0) ? -2.0f : (float)(sp * sp / sb / ss);
    top2_push(t2, sc, k);
  }
  tb[threadIdx.x] = t2.best;
  ti[threadIdx.x] = t2.scan;
  ts[threadIdx.x] = t2.second;
  __syncthreads();
  if (threadIdx.x == 0) {
    const Top2 r = top2_merge(tb, ti, ts, 0, kThreads);
    const int k = r.scan;
    const int kx = k % cnt[0], ky = (k / cnt[0]) % cnt[1], kz = k / (cnt[0] * cnt[1]);
    out_best[b] = r.best;
    out_second[b] = r.second;
    out_match[3 * b + 0] = first[0] + kx * g.bg;
    out_match[3 * b + 1] = first[1] + ky * g.bg;
    out_match[3 * b + 2] = first[2] + kz * g.bg;
  }
}

// ---------------------------------------------------------------------------
// The other block metrics (itkBlockMatcher.txx:105-227), searched like
// ncc_generic_kernel: one workgroup per block, scan order, the same top-2
// (the search keeps the LARGEST score whatever the metric).

// Base block of workgroup `b` staged in LDS: values and the base-mask bit.
__device__ __forceinline__ void stage_base_block(const float* __restrict__ x,
                                                 const uint8_t* __restrict__ bm, const BMGrid& g,
                                                 const int p[3], float* lx, uint8_t* lm) {
  const int bw = g.bw, n3 = bw * bw * bw;
  for (int i = threadIdx.x; i < n3; i += kThreads) {
    const int dx = i % bw, dy = (i / bw) % bw, dz = i / (bw * bw);
    const int64_t o = lin3(g.dims[0], g.dims[1], p[0] + dx, p[1] + dy, p[2] + dz);
    lx[i] = x[o];
    lm[i] = (!bm || bm[o] != 0) ? 1 : 0;
  }
  __syncthreads();
}

__device__ __forceinline__ void block_corner(const BMGrid& g, int64_t b, int p[3]) {
  p[0] = g.start[0] + g.ng * (int)(b % g.nb[0]);
  p[1] = g.start[1] + g.ng * (int)((b / g.nb[0]) % g.nb[1]);
  p[2] = g.start[2] + g.ng * (int)(b / ((int64_t)g.nb[0] * g.nb[1]));
}

__device__ __forceinline__ void write_generic_result(const BMGrid& g, int64_t b,
                                                     const int first[3], const int cnt[3],
                                                     const Top2& r, float* out_best,
                                                     float* out_second, int32_t* out_match) {
  const int k = r.scan;
  const int kx = k % cnt[0], ky = (k / cnt[0]) % cnt[1], kz = k / (cnt[0] * cnt[1]);
  out_best[b] = r.best;
  out_second[b] = r.second;
  out_match[3 * b + 0] = first[0] + kx * g.bg;
  out_match[3 * b + 1] = first[1] + ky * g.bg;
  out_match[3 * b + 2] = first[2] + kz * g.bg;
}

// Sum of squared differences (itkBlockMatcher.txx:195-227): the mean of
// (base - search)^2 over the voxels inside both masks, -2 when there are
// none. One thread per candidate; the differences of two floats and their
// squares are exact in double, so the sum matches the reference's to the
// last rounding of the additions.
__global__ __launch_bounds__(kThreads) void ssd_generic_kernel(
    const float* __restrict__ x, const uint8_t* __restrict__ bm,
    const float* __restrict__ y, const uint8_t* __restrict__ sm, BMGrid g,
    float* out_best, float* out_second, int32_t* out_match, const uint32_t* flags) {
  extern __shared__ float dyn[];
  __shared__ float tb[kThreads], ts[kThreads];
  __shared__ int ti[kThreads];
  const int64_t b = blockIdx.x;
  int p[3];
  block_corner(g, b, p);
  if (flags && (flags[b] & kBlockSkip)) {
    if (threadIdx.x == 0) {
      out_best[b] = -3.0f;
      out_second[b] = -3.0f;
      for (int k = 0; k < 3; ++k) out_match[3 * b + k] = p[k];
    }
    return;
  }
  const int bw = g.bw, n3 = bw * bw * bw, nx = g.dims[0], ny = g.dims[1];
  float* lx = dyn;
  uint8_t* lm = reinterpret_cast<uint8_t*>(dyn + n3);
  stage_base_block(x, bm, g, p, lx, lm);
  int first[3], cnt[3];
  for (int k = 0; k < 3; ++k) cand_range(p[k], g.dims[k], g.bw, g.bg, g.nw, &first[k], &cnt[k]);
  Top2 t2;
  top2_init(t2);
  const int total = cnt[0] * cnt[1] * cnt[2];
  for (int k = threadIdx.x; k < total; k += kThreads) {
    const int kx = k % cnt[0], ky = (k / cnt[0]) % cnt[1], kz = k / (cnt[0] * cnt[1]);
    const int s0 = first[0] + kx * g.bg, s1 = first[1] + ky * g.bg, s2 = first[2] + kz * g.bg;
    double sum = 0.0;
    int count = 0;
    int i = 0;
    for (int dz = 0; dz < bw; ++dz)
      for (int dy = 0; dy < bw; ++dy) {
        const int64_t row = lin3(nx, ny, s0, s1 + dy, s2 + dz);
        for (int dx = 0; dx < bw; ++dx, ++i) {
          if (!lm[i] || (sm && sm[row + dx] == 0)) continue;
          const double d = (double)lx[i] - (double)y[row + dx];
          sum += d * d;
          ++count;
        }
      }
    top2_push(t2, count == 0 ? -2.0f : (float)(sum / count), k);
  }
  tb[threadIdx.x] = t2.best;
  ti[threadIdx.x] = t2.scan;
  ts[threadIdx.x] = t2.second;
  __syncthreads();
  if (threadIdx.x == 0)
    write_generic_result(g, b, first, cnt, top2_merge(tb, ti, ts, 0, kThreads), out_best,
                         out_second, out_match);
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}

// Correlation ratio (itkBlockMatcher.txx:105-191). The voxels inside both
// masks are grouped by the search intensity's bin floor(y * 0.1); a bin's
// statistics use only base values above 10. The score is
// 1 - (sum over bins with more than 2 such values of their variance) /
// (variance of every such value), kept when above 0.4, else -2; -2 also
// when fewer than 3 values count or the bin variances sum to 0 (and when no
// voxel is inside both masks, where the reference reads an empty list).
// One wave per candidate, lane l holding voxels l, l + 64, ...; the bins are
// visited in increasing order (the order of the reference's sort), each by a
// wave minimum and three wave sums in double. Sums within a bin are formed
// in another order than the reference's sorted one, so scores agree to
// rounding.
__global__ __launch_bounds__(kThreads) void cr_generic_kernel(
    const float* __restrict__ x, const uint8_t* __restrict__ bm,
    const float* __restrict__ y, const uint8_t* __restrict__ sm, BMGrid g,
    float* out_best, float* out_second, int32_t* out_match, const uint32_t* flags) {
  extern __shared__ float dyn[];
  constexpr int kWaves = kThreads / 64;
  __shared__ float tb[kWaves], ts[kWaves];
  __shared__ int ti[kWaves];
  const int64_t b = blockIdx.x;
  int p[3];
  block_corner(g, b, p);
  if (flags && (flags[b] & kBlockSkip)) {
    if (threadIdx.x == 0) {
      out_best[b] = -3.0f;
      out_second[b] = -3.0f;
      for (int k = 0; k < 3; ++k) out_match[3 * b + k] = p[k];
    }
    return;
  }
  const int bw = g.bw, n3 = bw * bw * bw, nx = g.dims[0], ny = g.dims[1];
  float* lx = dyn;
  uint8_t* lm = reinterpret_cast<uint8_t*>(dyn + n3);
  stage_base_block(x, bm, g, p, lx, lm);
  int first[3], cnt[3];
  for (int k = 0; k < 3; ++k) cand_range(p[k], g.dims[k], g.bw, g.bg, g.nw, &first[k], &cnt[k]);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  Top2 t2;
  top2_init(t2);
  const int total = cnt[0] * cnt[1] * cnt[2];
  for (int k = wave; k < total; k += kWaves) {  // wave-uniform
    const int kx = k % cnt[0], ky = (k / cnt[0]) % cnt[1], kz = k / (cnt[0] * cnt[1]);
    const int s0 = first[0] + kx * g.bg, s1 = first[1] + ky * g.bg, s2 = first[2] + kz * g.bg;
    // Voxels of this lane: bin key (INT_MAX when outside a mask) and value.
    uint64_t pending = 0;  // bit j: voxel lane + 64 j still to be binned
    for (int i = lane, j = 0; i < n3; i += 64, ++j) {
      const int dx = i % bw, dy = (i / bw) % bw, dz = i / (bw * bw);
      const int64_t o = lin3(nx, ny, s0 + dx, s1 + dy, s2 + dz);
      if (lm[i] && (!sm || sm[o] != 0)) pending |= 1ull << j;
    }
    double sum_var = 0.0, sx = 0.0, sx2 = 0.0, cx = 0.0;
    while (true) {
      int lane_min = INT_MAX;
      for (int i = lane, j = 0; i < n3; i += 64, ++j) {
        if (!((pending >> j) & 1)) continue;
        const int dx = i % bw, dy = (i / bw) % bw, dz = i / (bw * bw);
        const float yv = y[lin3(nx, ny, s0 + dx, s1 + dy, s2 + dz)];
        lane_min = min(lane_min, (int)floor((double)yv * 0.1));
      }
      const int key = wave_min(lane_min);
      if (key == INT_MAX) break;
      double s = 0.0, s2v = 0.0, c = 0.0;
      for (int i = lane, j = 0; i < n3; i += 64, ++j) {
        if (!((pending >> j) & 1)) continue;
        const int dx = i % bw, dy = (i / bw) % bw, dz = i / (bw * bw);
        const float yv = y[lin3(nx, ny, s0 + dx, s1 + dy, s2 + dz)];
        if ((int)floor((double)yv * 0.1) != key) continue;
        pending &= ~(1ull << j);
        const double xv = lx[i];
        if (xv > 10.0) {
          s += xv;
          s2v += xv * xv;
          c += 1.0;
        }
      }
      s = wave_sum(s);
      s2v = wave_sum(s2v);
      c = wave_sum(c);
      if (c > 2.0) {
        const double m = s / c;
        sum_var += s2v / c - m * m;
      }
      sx += s;
      sx2 += s2v;
      cx += c;
    }
    float score = -2.0f;
    if (cx > 2.0 && sum_var > 0.0) {
      const double m = sx / cx;
      const double cr = 1.0 - sum_var / (sx2 / cx - m * m);
      if (cr > 0.4) score = (float)cr;
    }
    top2_push(t2, score, k);  // every lane, identical up to rounding; lane 0 reports
  }
  if (lane == 0) {
    tb[wave] = t2.best;
    ti[wave] = t2.scan;
    ts[wave] = t2.second;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    write_generic_result(g, b, first, cnt, top2_merge(tb, ti, ts, 0, kWaves), out_best,
                         out_second, out_match);
}

// ---------------------------------------------------------------------------
// Exact re-scoring of the blocks next to the score filter's thresholds.
//
// The tiled kernels (BW 4, BG 1) accumulate in float, so their NCC^2 values
// differ from the reference's double sums (itkBlockMatcher.txx:46-101) by up
// to a few 1e-5. The filter keeps a block when best > 1e-7 and |best -
// second| > 1e-4 (itkBlockMatcher.txx:886-916), so a block whose float best
// or gap lies within kRescoreBand of a threshold could land on the other
// side of it. The tiled kernels list such blocks (and on request exact
// ties), and one workgroup per listed block rebuilds its record as the
// reference computes it:
//  1. every candidate is scored in float with its own window mean, together
//     with a rigorous interval [lo, hi] holding the reference's double
//     score (first-order rounding-error bounds of the float sums, padded);
//  2. the candidates whose hi reaches the second largest lo (less 2^-20)
//     are scored again in double with the reference's operation order (the
//     means over all voxels summed in raster order, the masked centred sums
//     in raster order, then prod * prod / bss / sss), without fused
//     multiply-adds; the others lie below the best and the second and
//     cannot change the record;
//  3. the record follows the reference's float score buffer
//     (itkBlockMatcher.txx:764-774,823-833): for the double scores v_1..v_n
//     in scan order it ends with
//       best   = float(max_k v_k),
//       match  = the last k with v_k > double(float(max(-3, v_1..v_{k-1}))),
//       second = float(max over the other candidates), -3 with one candidate,
//     computed over the candidates of step 2 (the others never replace the
//     best after a candidate within 2^-20 of the maximum, and never attain
//     the second).
__device__ __forceinline__ int cand_total(const BMGrid& g, int64_t b, int p[3], int first[3],
                                          int cnt[3]) {
  block_corner(g, b, p);
  for (int k = 0; k < 3; ++k) cand_range(p[k], g.dims[k], g.bw, g.bg, g.nw, &first[k], &cnt[k]);
  return cnt[0] * cnt[1] * cnt[2];
}

// Images of a re-scoring (base and search, masks NULL = all on) and the
// search image's window statistics (mean over all voxels in .x).
struct RescoreArgs {
  const void* x;  // float or half_t elements (the kernel's E)
  const uint8_t* bm;
  const void* y;
  const uint8_t* sm;
  const float2* stats;
  unsigned long long* count;  // optional: += blocks re-scored
  int exact_records;          // 1: the reference's record; 0: its filter decision only
};

struct RescoreLds {
  union {
    float2 iv[kRescoreMaxCand];  // phase 1: (lo, hi) of the double score
    double v[kRescoreMaxCand];   // phase 2: the double score, -inf off the exact set
  };
  float win[kRescoreMaxWin];    // the union of the candidate windows
  uint8_t won[kRescoreMaxWin];  // its search mask (1 = on)
  float base[64];               // the block
  float xm[64];                 // float(x - mean), 0 where the base mask is off
  uint8_t bon[64];
  int16_t elist[kRescoreMaxCand];  // phase 2: the candidates scored in double
  double wave[kThreads / 64];
  int iwave[kThreads / 64];
  int anyoff;                   // a base or window voxel is masked off
  int ne;
};

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Maximum of v over the workgroup (all threads get it).
__device__ __forceinline__ double group_max(double v, RescoreLds& L) {
  v = wave_max(v);
  if ((threadIdx.x & 63) == 0) L.wave[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = L.wave[0];
  for (int w = 1; w < kThreads / 64; ++w) r = fmax(r, L.wave[w]);
  __syncthreads();
  return r;
}

// The two largest of v over the workgroup (m1 >= m2; -inf when absent).
__device__ __forceinline__ void group_top2(float a, float b, RescoreLds& L, float* m1,
                                           float* m2) {
  // (a, b) is a thread's own top-2; merge by waves then across waves.
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const float c = __shfl_xor(a, o, 64), d = __shfl_xor(b, o, 64);
    const float hi = fmaxf(a, c);
    b = fmaxf(fminf(a, c), fmaxf(b, d));
    a = hi;
  }
  if ((threadIdx.x & 63) == 0) {
    L.wave[threadIdx.x >> 6] = __hiloint2double(__float_as_int(a), __float_as_int(b));
  }
  __syncthreads();
  float x = -__builtin_huge_valf(), y = -__builtin_huge_valf();
  for (int w = 0; w < kThreads / 64; ++w) {
    const double q = L.wave[w];
    const float c = __int_as_float(__double2hiint(q)), d = __int_as_float(__double2loint(q));
    const float hi = fmaxf(x, c);
    y = fmaxf(fminf(x, c), fmaxf(y, d));
    x = hi;
  }
  __syncthreads();
  *m1 = x;
  *m2 = y;
}

// Candidate k's voxel v (raster order in the block) in the staged window.
__device__ __forceinline__ int win_index(int w0, int v, int ex, int ey) {
  return w0 + ((v >> 4) * ey + ((v >> 2) & 3)) * ex + (v & 3);
}

constexpr float kU = 5.9604645e-8f;          // 2^-24
constexpr float kGamma = 64.0f * kU * 1.01f;  // error factor of a 64-term float sum

// Phase 1 when every base and window voxel is on: one task per run of up to
// kLeanRun candidates of a candidate row (cz, cy), scored from shared row
// loads, with Sxy = Sum x~ t and Syy = Sum t^2 in float (t = y - mu, mu the
// window mean of the statistics) and Sxx = the block's double sum (the same
// for every candidate). Each candidate's sums are taken in the same order
// whichever task holds it.
constexpr int kLeanRun = 3;
template <int NW>
__device__ void lean_phase1(const RescoreArgs& ra, const BMGrid& g, const int first[3],
                            const int cnt[3], int ex, int ey, double bss, float bx, float sx,
                            RescoreLds& L, float* l1, float* l2, float* h1, float* h2) {
  constexpr int C = kLeanRun;
  constexpr int NR = (NW + C - 1) / C;  // tasks per candidate row
  const int nx = g.dims[0], ny = g.dims[1];
  const int rows = cnt[1] * cnt[2];
  const float isxx = (float)(1.0 / bss);
  for (int task = threadIdx.x; task < rows * NR; task += kThreads) {
    const int r = task / NR, c0 = (task % NR) * C;
    if (c0 >= cnt[0]) continue;
    const int ky = r % cnt[1], kz = r / cnt[1];
    float mu[C], syy[C], sxy[C], A[C];
    const float2* srow = ra.stats + lin3(nx, ny, first[0] + c0, first[1] + ky, first[2] + kz);
#pragma unroll
    for (int cx = 0; cx < C; ++cx) {
      mu[cx] = c0 + cx < cnt[0] ? srow[cx].x : 0.0f;
      syy[cx] = 0.0f;
      sxy[cx] = 0.0f;
      A[cx] = 0.0f;
    }
#pragma unroll 4
    for (int q = 0; q < 16; ++q) {  // block row (iz, iy) = (q / 4, q % 4)
      const float* wr = L.win + ((kz + (q >> 2)) * ey + ky + (q & 3)) * ex + c0;
      float w[C + 3];
#pragma unroll
      for (int i = 0; i < C + 3; ++i) w[i] = c0 + i < ex ? wr[i] : 0.0f;
      float xr[4];
#pragma unroll
      for (int ix = 0; ix < 4; ++ix) xr[ix] = L.xm[4 * q + ix];
#pragma unroll
      for (int cx = 0; cx < C; ++cx)
#pragma unroll
        for (int ix = 0; ix < 4; ++ix) {
          const float t = w[cx + ix] - mu[cx];
          sxy[cx] = fmaf(xr[ix], t, sxy[cx]);
          syy[cx] = fmaf(t, t, syy[cx]);
          A[cx] = fmaf(fabsf(xr[ix]), fabsf(t), A[cx]);
        }
    }
#pragma unroll
    for (int cx = 0; cx < C; ++cx) {
      if (c0 + cx >= cnt[0]) continue;
      // Error bounds as in the general phase 1; with every voxel on, the
      // exact centred values sum to zero over the block and over the
      // window, so the error e of the float mean (|e| <= 2u |mu|) enters
      // Sxy only through the float block sum (|Sum x~| <= |sx| + gamma bx)
      // and Syy only as 64 e^2.
      const float amu = fabsf(mu[cx]);
      const float d = kGamma * A[cx] + 3.0f * kU * A[cx] + 2.0f * kU * amu * (fabsf(sx) + kGamma * bx);
      const float dyy = kGamma * syy[cx] + 5.0f * kU * syy[cx] +
                        256.0f * (kU * amu) * (kU * amu);
      float lo, hi;
      if (syy[cx] - dyy > 0.0f) {
        const float smax = fabsf(sxy[cx]) + d, smin = fmaxf(fabsf(sxy[cx]) - d, 0.0f);
        hi = smax * smax * isxx / (syy[cx] - dyy) * (1.0f + 16.0f * kU) + 1e-9f;
        lo = smin * smin * isxx / (syy[cx] + dyy) * (1.0f - 16.0f * kU) - 1e-9f;
      } else {
        lo = -2.0f;  // the window sum may be zero: scored in double
        hi = __builtin_huge_valf();
      }
      L.iv[r * cnt[0] + c0 + cx] = make_float2(lo, hi);
      const float nl1 = fmaxf(*l1, lo);
      *l2 = fmaxf(fminf(*l1, lo), *l2);
      *l1 = nl1;
      const float nh1 = fmaxf(*h1, hi);
      *h2 = fmaxf(fminf(*h1, hi), *h2);
      *h1 = nh1;
    }
  }
}

// Re-scores block b (BW 4, BG 1) and rewrites its record. Called by every
// thread of the workgroup with the same b.
template <class E>
__device__ void rescore_block(const RescoreArgs& ra, const BMGrid& g, int64_t b, float* best,
                              float* second, int32_t* match, RescoreLds& L) {
#pragma clang fp contract(off)
  const int nx = g.dims[0], ny = g.dims[1];
  const E* X = static_cast<const E*>(ra.x);
  const E* Y = static_cast<const E*>(ra.y);
  int p[3], first[3], cnt[3];
  const int total = cand_total(g, b, p, first, cnt);
  const int ex = cnt[0] + 3, ey = cnt[1] + 3, ez = cnt[2] + 3;
  if (threadIdx.x == 0) L.anyoff = 0;
  __syncthreads();
  // The union of the candidate windows, kStageBatch voxels per thread with
  // their loads in flight together.
  constexpr int kStageBatch = 8;
  const int nwin = ex * ey * ez;
  for (int i0 = threadIdx.x; i0 < nwin; i0 += kStageBatch * kThreads) {
    float yv[kStageBatch];
    uint8_t mv[kStageBatch];
#pragma unroll
    for (int u = 0; u < kStageBatch; ++u) {
      const int i = min(i0 + u * kThreads, nwin - 1);
      const int wx = i % ex, wy = (i / ex) % ey, wz = i / (ex * ey);
      const int64_t o = lin3(nx, ny, first[0] + wx, first[1] + wy, first[2] + wz);
      yv[u] = (float)Y[o];
      mv[u] = ra.sm == nullptr ? (uint8_t)1 : ra.sm[o];
    }
#pragma unroll
    for (int u = 0; u < kStageBatch; ++u) {
      const int i = i0 + u * kThreads;
      if (i >= nwin) break;
      L.win[i] = yv[u];
      const bool on = mv[u] != 0;
      L.won[i] = on ? 1 : 0;
      if (!on) L.anyoff = 1;
    }
  }
  if (threadIdx.x < 64) {
    const int k = threadIdx.x;
    const int64_t o = lin3(nx, ny, p[0] + (k & 3), p[1] + ((k >> 2) & 3), p[2] + (k >> 4));
    L.base[k] = (float)X[o];
    const bool on = ra.bm == nullptr || ra.bm[o] != 0;
    L.bon[k] = on ? 1 : 0;
    if (!on) L.anyoff = 1;
  }
  __syncthreads();
  // The block mean over all voxels, summed in raster order as the
  // reference does (itkBlockMatcher.txx:57-68).
  double count = 0.0, mb = 0.0;
  for (int k = 0; k < 64; ++k) {
    mb += L.base[k];
    count += 1.0;
  }
  mb /= count;
  if (threadIdx.x < 64)
    L.xm[threadIdx.x] = L.bon[threadIdx.x] ? (float)(L.base[threadIdx.x] - mb) : 0.0f;
  __syncthreads();

  // Phase 1: float scores with intervals.
  constexpr float u = kU;
  constexpr float gam = kGamma;
  // thread's top-2 of lo and of hi
  float l1 = -__builtin_huge_valf(), l2 = -__builtin_huge_valf();
  float h1 = -__builtin_huge_valf(), h2 = -__builtin_huge_valf();
  const bool lean = !L.anyoff && cnt[1] * cnt[2] <= kThreads;
  if (lean) {
    double bss = 0.0;  // the reference's base sum of squares (every voxel on)
    float bx = 0.0f, sx = 0.0f;
    for (int k = 0; k < 64; ++k) {
      const double d = L.base[k] - mb;
      bss += d * d;
      bx += fabsf(L.xm[k]);
      sx += L.xm[k];
    }
    if (bss > 0.0) {
#define MIRORR_LEAN(NWV) \
  lean_phase1<NWV>(ra, g, first, cnt, ex, ey, bss, bx, sx, L, &l1, &l2, &h1, &h2)
      switch (g.nw) {
        case 5: MIRORR_LEAN(5); break;
        case 7: MIRORR_LEAN(7); break;
        case 9: MIRORR_LEAN(9); break;
        case 11: MIRORR_LEAN(11); break;
        default: MIRORR_LEAN(13); break;
      }
#undef MIRORR_LEAN
    } else {
      for (int k = threadIdx.x; k < total; k += kThreads) L.iv[k] = make_float2(-2.0f, -2.0f);
      l1 = h1 = threadIdx.x < total ? -2.0f : l1;
      l2 = h2 = threadIdx.x + kThreads < total ? -2.0f : l2;
    }
  }
  for (int k = lean ? total : threadIdx.x; k < total; k += kThreads) {
    const int kx = k % cnt[0], ky = (k / cnt[0]) % cnt[1], kz = k / (cnt[0] * cnt[1]);
    const int w0 = (kz * ey + ky) * ex + kx;
    const float mu = ra.stats[lin3(nx, ny, first[0] + kx, first[1] + ky, first[2] + kz)].x;
    float sxy = 0.0f, syy = 0.0f, sxx = 0.0f, A = 0.0f, Bx = 0.0f, C = 0.0f;
#pragma unroll 8
    for (int v = 0; v < 64; ++v) {
      const int wi = win_index(w0, v, ex, ey);
      const float m = (L.won[wi] && L.bon[v]) ? 1.0f : 0.0f;
      const float t = m * (L.win[wi] - mu);
      const float xv = L.xm[v] * m;
      sxy = fmaf(xv, t, sxy);
      syy = fmaf(t, t, syy);
      sxx = fmaf(xv, xv, sxx);
      A = fmaf(fabsf(xv), fabsf(t), A);
      Bx += fabsf(xv);
      C += fabsf(t);
    }
    const float amu = fabsf(mu);
    const float dxy = gam * A + 3.0f * u * (A + amu * Bx);
    const float dyy = gam * syy + 5.0f * u * (syy + amu * C) + 64.0f * (u * amu) * (u * amu);
    const float dxx = gam * sxx + 3.0f * u * sxx;
    float lo, hi;
    if (sxx == 0.0f) {
      // No voxel inside both masks, or the block's centred values are all
      // exactly zero there (float(x - mean) is 0 only when x - mean is):
      // the reference's -2.
      lo = hi = -2.0f;
    } else if (sxx - dxx > 0.0f && syy - dyy > 0.0f) {
      const float smax = fabsf(sxy) + dxy, smin = fmaxf(fabsf(sxy) - dxy, 0.0f);
      hi = smax * smax / ((sxx - dxx) * (syy - dyy)) * (1.0f + 16.0f * u) + 1e-9f;
      lo = smin * smin / ((sxx + dxx) * (syy + dyy)) * (1.0f - 16.0f * u) - 1e-9f;
    } else {
      lo = -2.0f;  // the window sum may be zero: scored in double
      hi = __builtin_huge_valf();
    }
    L.iv[k] = make_float2(lo, hi);
    const float nl1 = fmaxf(l1, lo);
    l2 = fmaxf(fminf(l1, lo), l2);
    l1 = nl1;
    const float nh1 = fmaxf(h1, hi);
    h2 = fmaxf(fminf(h1, hi), h2);
    h1 = nh1;
  }
  float L1, L2, U1, U2;
  group_top2(l1, l2, L, &L1, &L2);  // (its syncs also publish L.iv)
  group_top2(h1, h2, L, &U1, &U2);
  if (!ra.exact_records) {
    // The filter's decision alone (the registration): the true best M and
    // second S satisfy L1 <= M <= U1, S >= L2 and, when one candidate's lo
    // exceeds every other hi, S <= U2. A decision certain by more than 1e-6
    // (above the float rounding of the filter's comparisons) is written as a
    // record the filter decides the same way; the rest is scored in double.
    constexpr float mg = 1e-6f;
    const bool cull = U1 < 1e-7f - mg || (total >= 2 && U1 - L2 < 1e-4f - mg);
    const bool keep = L1 > 1e-7f + mg && (total < 2 || L1 - U2 > 1e-4f + mg);
    if (cull || keep) {
      for (int k = threadIdx.x; k < total && keep; k += kThreads)
        if (L.iv[k].y == U1) {  // the unique best candidate
          match[3 * b + 0] = first[0] + k % cnt[0];
          match[3 * b + 1] = first[1] + (k / cnt[0]) % cnt[1];
          match[3 * b + 2] = first[2] + k / (cnt[0] * cnt[1]);
        }
      if (threadIdx.x == 0) {
        best[b] = 0.5f * (L1 + U1);
        second[b] = cull ? best[b] : (total >= 2 ? 0.5f * (L2 + U2) : -3.0f);
      }
      __syncthreads();
      return;
    }
  }
  const float cut = (total >= 2 ? L2 : L1) - 9.5367432e-7f;  // less 2^-20

  // Phase 2: the candidates that can reach the top two, in double, one wave
  // per candidate with a lane per voxel; each sum is accumulated in raster
  // order from the lanes' terms, as the reference's loops add them (a voxel
  // outside either mask adds an exact zero where the reference skips it).
  if (threadIdx.x == 0) L.ne = 0;
  __syncthreads();
  for (int k = threadIdx.x; k < total; k += kThreads) {
    if (L.iv[k].y >= cut) L.elist[atomicAdd(&L.ne, 1)] = (int16_t)k;
    L.v[k] = -__builtin_huge_val();  // the thread's own slot: its interval was read above
  }
  __syncthreads();
  const int ne = L.ne;
  {
    const int ln = threadIdx.x & 63;
    const double uu = L.base[ln] - mb;
    for (int e = threadIdx.x >> 6; e < ne; e += kThreads / 64) {  // wave-uniform
      const int k = L.elist[e];
      const int kx = k % cnt[0], ky = (k / cnt[0]) % cnt[1], kz = k / (cnt[0] * cnt[1]);
      const int wi = win_index((kz * ey + ky) * ex + kx, ln, ex, ey);
      const double yv = L.win[wi];
      const bool on = L.bon[ln] && L.won[wi];
      double ms = 0.0;
      for (int q = 0; q < 64; ++q) ms += __shfl(yv, q, 64);
      ms /= count;
      // Sums over the voxels inside both masks (itkBlockMatcher.txx:70-95).
      const double tt = yv - ms;
      const double pv = on ? uu * tt : 0.0, bv = on ? uu * uu : 0.0, sv = on ? tt * tt : 0.0;
      double prod = 0.0, bss = 0.0, sss = 0.0;
      for (int q = 0; q < 64; ++q) {
        prod += __shfl(pv, q, 64);
        bss += __shfl(bv, q, 64);
        sss += __shfl(sv, q, 64);
      }
      if (ln == 0) L.v[k] = (bss <= 0.0 || sss <= 0.0) ? -2.0 : prod * prod / bss / sss;
    }
  }
  __syncthreads();
  // Each thread replays a contiguous run of candidates.
  const int per = (total + kThreads - 1) / kThreads;
  const int k0 = min((int)threadIdx.x * per, total), k1 = min(k0 + per, total);
  double run_max = -3.0;
  for (int k = k0; k < k1; ++k) run_max = fmax(run_max, L.v[k]);
  // Exclusive prefix maximum of the runs (the buffer starts at -3).
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double inc = run_max;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double t = __shfl_up(inc, o, 64);
    if (lane >= o) inc = fmax(inc, t);
  }
  double pre = __shfl_up(inc, 1, 64);
  if (lane == 0) pre = -3.0;
  if (lane == 63) L.wave[wv] = inc;
  __syncthreads();
  for (int w = 0; w < wv; ++w) pre = fmax(pre, L.wave[w]);
  __syncthreads();
  int last = -1;
  for (int k = k0; k < k1; ++k) {
    const double v = L.v[k];
    if ((double)(float)pre < v) last = k;
    pre = fmax(pre, v);
  }
  // The last replacement over the whole sequence.
  int m = last;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
  if (lane == 0) L.iwave[wv] = m;
  __syncthreads();
  int I = L.iwave[0];
  for (int w = 1; w < kThreads / 64; ++w) I = max(I, L.iwave[w]);
  double other = -3.0;
  for (int k = k0; k < k1; ++k)
    if (k != I) other = fmax(other, L.v[k]);
  const double top = group_max(run_max, L);
  const double sec = group_max(other, L);
  if (threadIdx.x == 0) {
    best[b] = (float)top;
    second[b] = (float)sec;
    match[3 * b + 0] = first[0] + I % cnt[0];
    match[3 * b + 1] = first[1] + (I / cnt[0]) % cnt[1];
    match[3 * b + 2] = first[2] + I / (cnt[0] * cnt[1]);
  }
  __syncthreads();
}

// Re-scores the listed blocks, one workgroup per block; the grid strides
// over the list.
template <class E>
__global__ __launch_bounds__(kThreads) void rescore_list_kernel(BMGrid g, float* best,
                                                                float* second, int32_t* match,
                                                                RescoreArgs ra,
                                                                const int32_t* __restrict__ list,
                                                                const int32_t* __restrict__ count) {
  __shared__ RescoreLds L;
  const int n = *count;
  if (blockIdx.x == 0 && threadIdx.x == 0 && ra.count && n > 0)
    atomicAdd(ra.count, (unsigned long long)n);
  for (int i = blockIdx.x; i < n; i += gridDim.x)
    rescore_block<E>(ra, g, list[i], best, second, match, L);
}

// The re-scoring of several pairs' lists in one launch: workgroup w takes
// entries w, w + gridDim.x, ... of every pair's list.
struct RescoreItem {
  BMGrid g;
  float* best;
  float* second;
  int32_t* match;
  RescoreArgs ra;
  const int32_t* list;   // NULL: no re-scoring for this pair
  const int32_t* count;
};

template <class E>
__global__ __launch_bounds__(kThreads) void rescore_multi_kernel(
    const RescoreItem* __restrict__ items, int npairs) {
  __shared__ RescoreLds L;
  for (int p = 0; p < npairs; ++p) {
    const RescoreItem it = items[p];
    if (!it.list) continue;
    const int n = *it.count;
    if (blockIdx.x == 0 && threadIdx.x == 0 && it.ra.count && n > 0)
      atomicAdd(it.ra.count, (unsigned long long)n);
    for (int i = blockIdx.x; i < n; i += gridDim.x)
      rescore_block<E>(it.ra, it.g, it.list[i], it.best, it.second, it.match, L);
  }
}

// Zeroes the masked-tile and re-score counters of several pairs.
__global__ void reset_counters_multi(int32_t* const* __restrict__ ctr, int n) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n && ctr[i]) *ctr[i] = 0;
}

__global__ void postprocess_kernel(BMGrid g, Geom geom, const float* __restrict__ best,
                                   const float* __restrict__ second,
                                   const int32_t* __restrict__ match, uint8_t* kept,
                                   double* base_mm, double* match_mm) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= g.nblocks) return;
  const double cb = best[b], c2 = second[b];
  kept[b] = (cb > 1e-7 && fabs(cb - c2) > 1e-4) ? 1 : 0;
  const int bxi = (int)(b % g.nb[0]);
  const int byi = (int)((b / g.nb[0]) % g.nb[1]);
  const int bzi = (int)(b / ((int64_t)g.nb[0] * g.nb[1]));
  const int p[3] = {g.start[0] + g.ng * bxi, g.start[1] + g.ng * byi, g.start[2] + g.ng * bzi};
  const int m[3] = {match[3 * b], match[3 * b + 1], match[3 * b + 2]};
  const double half = 0.5 * g.bw;
  for (int r = 0; r < 3; ++r) {
    double pb = geom.origin[r], pm = geom.origin[r];
    for (int c = 0; c < 3; ++c) {
      const double dsc = geom.dir[3 * r + c] * geom.spacing[c];
      pb += dsc * (double)p[c];
      pm += dsc * (double)m[c];
    }
    base_mm[3 * b + r] = pb + geom.spacing[r] * half;
    match_mm[3 * b + r] = pm + geom.spacing[r] * half;
  }
}

// Pair output for the LTS: X gets the moving-side point and Y the fixed-side
// point of every block; `swap` marks the second symmetric pass, whose base
// image is the fixed one (itkSymMirorrRegistrationMethod.txx:595-602).
__global__ __launch_bounds__(kThreads) void postprocess_pairs_kernel(
    BMGrid g, Geom geom, const float* __restrict__ best, const float* __restrict__ second,
    const int32_t* __restrict__ match, int swap, uint8_t* kept, double* X, double* Y,
    unsigned long long* count) {
  // Grid-stride over the blocks with one count update per workgroup: the
  // counter is a single device-wide address, so updates are kept few.
  __shared__ unsigned int wsum[kThreads / 64];
  const double half = 0.5 * g.bw;
  double* pb = swap ? Y : X;
  double* pm = swap ? X : Y;
  unsigned int nk = 0;
  for (int64_t b = (int64_t)blockIdx.x * kThreads + threadIdx.x; b < g.nblocks;
       b += (int64_t)gridDim.x * kThreads) {
    const double cb = (double)best[b];
    const double c2 = (double)second[b];
    const bool k = cb > 1e-7 && fabs(cb - c2) > 1e-4;
    nk += k ? 1u : 0u;
    kept[b] = k ? 1 : 0;
    const int bxi = (int)(b % g.nb[0]);
    const int byi = (int)((b / g.nb[0]) % g.nb[1]);
    const int bzi = (int)(b / ((int64_t)g.nb[0] * g.nb[1]));
    const int p[3] = {g.start[0] + g.ng * bxi, g.start[1] + g.ng * byi,
                      g.start[2] + g.ng * bzi};
    const int m[3] = {match[3 * b], match[3 * b + 1], match[3 * b + 2]};
    for (int r = 0; r < 3; ++r) {
      double vb = geom.origin[r], vm = geom.origin[r];
      for (int c = 0; c < 3; ++c) {
        const double dsc = geom.dir[3 * r + c] * geom.spacing[c];
        vb += dsc * (double)p[c];
        vm += dsc * (double)m[c];
      }
      pb[3 * b + r] = vb + geom.spacing[r] * half;
      pm[3 * b + r] = vm + geom.spacing[r] * half;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) nk += __shfl_xor(nk, o, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = nk;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned int tot = 0;
    for (int w = 0; w < kThreads / 64; ++w) tot += wsum[w];
    if (tot) atomicAdd(count, (unsigned long long)tot);
  }
}

// ---------------------------------------------------------------------------
constexpr int kChunk = 1024;  // elements per compaction workgroup

__global__ __launch_bounds__(kThreads) void compact_count(const uint8_t* f, int64_t n,
                                                         int32_t* counts) {
  __shared__ int s;
  if (threadIdx.x == 0) s = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kChunk;
  int c = 0;
  for (int k = threadIdx.x; k < kChunk; k += kThreads) {
    const int64_t i = base + k;
    if (i < n && f[i]) ++c;
  }
  atomicAdd(&s, c);
  __syncthreads();
  if (threadIdx.x == 0) counts[blockIdx.x] = s;
}

__global__ __launch_bounds__(1024) void compact_scan(const int32_t* counts, int nchunks,
                                                     int32_t* offsets) {
  __shared__ int32_t sh[1024];
  __shared__ int32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nchunks; base += 1024) {
    const int i = base + threadIdx.x;
    const int32_t v = i < nchunks ? counts[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      const int32_t t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nchunks) offsets[i] = carry + sh[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += sh[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) offsets[nchunks] = carry;
}

__global__ __launch_bounds__(kThreads) void compact_scatter(const uint8_t* f, int64_t n,
                                                           const int32_t* offsets,
                                                           int32_t* out) {
  // Each thread owns 4 consecutive elements of the chunk.
  __shared__ int32_t sh[kThreads];
  const int64_t base = (int64_t)blockIdx.x * kChunk + 4 * threadIdx.x;
  int flags[4], c = 0;
  for (int k = 0; k < 4; ++k) {
    const int64_t i = base + k;
    flags[k] = (i < n && f[i]) ? 1 : 0;
    c += flags[k];
  }
  sh[threadIdx.x] = c;
  __syncthreads();
  for (int o = 1; o < kThreads; o <<= 1) {
    const int32_t t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  int pos = offsets[blockIdx.x] + sh[threadIdx.x] - c;
  for (int k = 0; k < 4; ++k)
    if (flags[k]) out[pos++] = (int32_t)(base + k);
}

// ---------------------------------------------------------------------------
// Block-variance selection: exact K-th largest block key (the float bits of
// the centred sum of squares, non-negative, so ordered as unsigned ints) by
// three radix digits (bits 31..20, 19..8, 7..0): a histogram pass over the
// blocks, then a one-workgroup bin search, per digit.
__global__ __launch_bounds__(kThreads) void vsel_init(VarSelState* st, unsigned int* hist,
                                                      unsigned long long k) {
  for (int i = threadIdx.x; i < kVarSelHistBins; i += kThreads) hist[i] = 0u;
  if (threadIdx.x == 0) {
    st->k = k;
    st->k2 = st->k3 = 0ull;
    st->bin1 = st->bin2 = st->bin3 = 0u;
    st->kept = 0ull;
  }
}

template <int PASS>
__global__ __launch_bounds__(kThreads) void vsel_hist(const float* __restrict__ key, int64_t n,
                                                      VarSelState* st, unsigned int* hist) {
  constexpr int NB = PASS == 2 ? 256 : 4096;
  __shared__ unsigned int h[NB];
  for (int i = threadIdx.x; i < NB; i += kThreads) h[i] = 0u;
  __syncthreads();
  const unsigned int b1 = st->bin1, b2 = st->bin2;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kThreads) {
    const uint32_t u = __float_as_uint(key[i]);
    if (PASS == 0) {
      atomicAdd(&h[u >> 20], 1u);
    } else if (PASS == 1) {
      if ((u >> 20) == b1) atomicAdd(&h[(u >> 8) & 0xfffu], 1u);
    } else {
      if ((u >> 8) == ((b1 << 12) | b2)) atomicAdd(&h[u & 0xffu], 1u);
    }
  }
  __syncthreads();
  unsigned int* gh = hist + (PASS == 0 ? 0 : PASS == 1 ? 4096 : 8192);
  for (int i = threadIdx.x; i < NB; i += kThreads)
    if (h[i]) atomicAdd(&gh[i], h[i]);
}

// Bin search of one digit (one workgroup).
template <int PASS>
__global__ __launch_bounds__(kThreads) void vsel_scan(VarSelState* st, const unsigned int* hist) {
  if (PASS == 0) sel::find_bin<4096 / kThreads>(hist, st->k, &st->bin1, &st->k2);
  if (PASS == 1) sel::find_bin<4096 / kThreads>(hist + 4096, st->k2, &st->bin2, &st->k3);
  if (PASS == 2) sel::find_bin<256 / kThreads>(hist + 8192, st->k3, &st->bin3, &st->k3);
}

// Marks the blocks below the cut-off and accounts the kept ones.
__global__ __launch_bounds__(kThreads) void vsel_mark(const float* __restrict__ key,
                                                      uint32_t* __restrict__ flags, BMGrid g,
                                                      VarSelState* st,
                                                      unsigned long long* evals) {
  const uint32_t cut = (st->bin1 << 20) | (st->bin2 << 8) | st->bin3;
  unsigned long long kept = 0, ev = 0;
  for (int64_t b = (int64_t)blockIdx.x * kThreads + threadIdx.x; b < g.nblocks;
       b += (int64_t)gridDim.x * kThreads) {
    const bool keep = __float_as_uint(key[b]) >= cut;
    flags[b] = (flags[b] & ~kBlockSkip) | (keep ? 0u : kBlockSkip);
    if (keep) {
      const int bxi = (int)(b % g.nb[0]);
      const int byi = (int)((b / g.nb[0]) % g.nb[1]);
      const int bzi = (int)(b / ((int64_t)g.nb[0] * g.nb[1]));
      int f, cx, cy, cz;
      cand_range(g.start[0] + g.ng * bxi, g.dims[0], g.bw, g.bg, g.nw, &f, &cx);
      cand_range(g.start[1] + g.ng * byi, g.dims[1], g.bw, g.bg, g.nw, &f, &cy);
      cand_range(g.start[2] + g.ng * bzi, g.dims[2], g.bw, g.bg, g.nw, &f, &cz);
      kept += 1;
      ev += (unsigned long long)cx * cy * cz;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    kept += __shfl_xor(kept, o, 64);
    ev += __shfl_xor(ev, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    if (kept) atomicAdd(&st->kept, kept);
    if (evals && ev) atomicAdd(evals, ev * (unsigned long long)(g.bw * g.bw * g.bw));
  }
}

// ---------------------------------------------------------------------------
// Tile configurations (BW = 4, BG = 1): see kTileVariants below.

template <class T>
int64_t tiles_of(const BMGrid& g) {
  const int64_t nty = (g.nb[1] + T::TY - 1) / T::TY;
  const int64_t ntz = (g.nb[2] + T::TZ - 1) / T::TZ;
  return (int64_t)g.nb[0] * nty * ntz;
}

template <class T, class E>
void run_tiled(const BMGrid& g, TileArgs a, BMScratch& scr, hipStream_t s) {
  a.ntx = g.nb[0];
  a.nty = (g.nb[1] + T::TY - 1) / T::TY;
  a.ntz = (g.nb[2] + T::TZ - 1) / T::TZ;
  const int64_t nt = (int64_t)a.ntx * a.nty * a.ntz;
  if (nt > scr.ntiles_cap) throw Error(kErrState, "tile queue too small");
  a.queue = scr.queue;
  a.qbits = scr.qbits;
  a.qcount = scr.count;
  a.rlist = scr.rlist;
  a.rcount = scr.rcount;
  a.rties = scr.rescore_ties;
  hipLaunchKernelGGL(reset_counters, dim3(1), dim3(1), 0, s, scr.count, scr.rcount);
  const int WYf = T::fast_pitch((T::TY - 1) * g.ng + T::WX + T::YPAD);
  const int WZ = (T::TZ - 1) * g.ng + T::WX;
  const size_t lds_fast = sizeof(float) * (size_t)T::PX * WYf * WZ;
  if (scr.prof) a.work = scr.prof->begin("ncc_tile_fast", kProfFast, s);
  hipLaunchKernelGGL((ncc_tile_fast<T, E>), dim3((unsigned)nt), dim3(kThreads), lds_fast, s, a);
  MIRORR_CHECK_LAUNCH();
  if (scr.prof) scr.prof->end(kProfFast, s);
  const int WYm = (T::TY - 1) * g.ng + T::WX;
  const size_t lds_masked = (sizeof(float) + 1) * (size_t)T::PX * WYm * WZ;
  // Few tiles (a coarse level): split each queued tile over four workgroups.
  a.msplit = nt < 2048 ? 4 : 1;
  if (scr.prof) a.work = scr.prof->begin("ncc_tile_masked", kProfMasked, s);
  hipLaunchKernelGGL((ncc_tile_masked<T, E>), dim3((unsigned)(nt * a.msplit)), dim3(kThreads),
                     lds_masked, s, a);
  MIRORR_CHECK_LAUNCH();
  if (scr.prof) scr.prof->end(kProfMasked, s);
}

// Per-pair arguments of a batched search, packed into one device buffer:
// tile arguments, the flat item starts of the fast and masked launches,
// the re-score items and the counters to zero.
struct BatchLayout {
  size_t args, fstart, mstart, items, ctr, bytes;
  explicit BatchLayout(int n) {
    auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
    args = 0;
    fstart = al(args + sizeof(TileArgs) * n);
    mstart = al(fstart + sizeof(int) * (n + 1));
    items = al(mstart + sizeof(int) * (n + 1));
    ctr = al(items + sizeof(RescoreItem) * n);
    bytes = al(ctr + sizeof(int32_t*) * 2 * n);
  }
};

template <class T, class E>
void run_tiled_multi(const std::vector<SearchJob>& jobs, unsigned char* dev, hipStream_t s) {
  const int n = (int)jobs.size();
  const BatchLayout lay(n);
  std::vector<unsigned char> host(lay.bytes, 0);
  TileArgs* A = reinterpret_cast<TileArgs*>(host.data() + lay.args);
  int* fs = reinterpret_cast<int*>(host.data() + lay.fstart);
  int* ms = reinterpret_cast<int*>(host.data() + lay.mstart);
  RescoreItem* R = reinterpret_cast<RescoreItem*>(host.data() + lay.items);
  int32_t** C = reinterpret_cast<int32_t**>(host.data() + lay.ctr);
  int64_t nf = 0, nm = 0;
  bool rescore = false;
  for (int p = 0; p < n; ++p) {
    const SearchJob& j = jobs[p];
    TileArgs a{};
    a.y = j.y;
    a.sm = j.smask;
    a.bm = j.bmask;
    a.g = j.g;
    a.xt = j.bp.xt;
    a.sxx = j.bp.sxx;
    a.sumxt = j.bp.sumxt;
    a.bflags = j.bp.flags;
    a.stats = j.st.ms;
    a.out_best = j.res.best;
    a.out_second = j.res.second;
    a.out_match = j.res.match;
    a.ntx = j.g.nb[0];
    a.nty = (j.g.nb[1] + T::TY - 1) / T::TY;
    a.ntz = (j.g.nb[2] + T::TZ - 1) / T::TZ;
    const int64_t nt = (int64_t)a.ntx * a.nty * a.ntz;
    if (nt > j.scr.ntiles_cap) throw Error(kErrState, "tile queue too small");
    a.queue = j.scr.queue;
    a.qbits = j.scr.qbits;
    a.qcount = j.scr.count;
    a.rlist = j.scr.rlist;
    a.rcount = j.scr.rcount;
    a.rties = j.scr.rescore_ties;
    a.msplit = nt < 2048 ? 4 : 1;
    A[p] = a;
    fs[p] = (int)nf;
    ms[p] = (int)nm;
    nf += nt;
    nm += nt * a.msplit;
    R[p] = RescoreItem{j.g, j.res.best, j.res.second, j.res.match,
                       RescoreArgs{j.x, j.bmask, j.y, j.smask, j.st.ms, j.scr.rescored,
                                   j.scr.exact_records},
                       j.scr.rlist, j.scr.rcount};
    rescore = rescore || j.scr.rlist != nullptr;
    C[2 * p] = j.scr.count;
    C[2 * p + 1] = j.scr.rcount;
  }
  if (nf > INT_MAX || nm > INT_MAX) throw Error(kErrArg, "batched search too large");
  fs[n] = (int)nf;
  ms[n] = (int)nm;
  // Pageable source: the copy has read it when the call returns.
  MIRORR_HIP(hipMemcpyAsync(dev, host.data(), lay.bytes, hipMemcpyHostToDevice, s));
  const TileArgs* dA = reinterpret_cast<const TileArgs*>(dev + lay.args);
  const int* dfs = reinterpret_cast<const int*>(dev + lay.fstart);
  const int* dms = reinterpret_cast<const int*>(dev + lay.mstart);
  hipLaunchKernelGGL(reset_counters_multi, dim3((unsigned)((2 * n + 63) / 64)), dim3(64), 0, s,
                     reinterpret_cast<int32_t* const*>(dev + lay.ctr), 2 * n);
  const int ng = jobs[0].g.ng;
  const int WYf = T::fast_pitch((T::TY - 1) * ng + T::WX + T::YPAD);
  const int WZ = (T::TZ - 1) * ng + T::WX;
  const size_t lds_fast = sizeof(float) * (size_t)T::PX * WYf * WZ;
  if (nf > 0)
    hipLaunchKernelGGL((ncc_tile_fast_multi<T, E>), dim3((unsigned)nf), dim3(kThreads), lds_fast,
                       s, dA, dfs, n);
  const int WYm = (T::TY - 1) * ng + T::WX;
  const size_t lds_masked = (sizeof(float) + 1) * (size_t)T::PX * WYm * WZ;
  if (nm > 0)
    hipLaunchKernelGGL((ncc_tile_masked_multi<T, E>), dim3((unsigned)nm), dim3(kThreads),
                       lds_masked, s, dA, dms, n);
  if (rescore)
    hipLaunchKernelGGL(rescore_multi_kernel<E>, dim3(1024), dim3(kThreads), 0, s,
                       reinterpret_cast<const RescoreItem*>(dev + lay.items), n);
  MIRORR_CHECK_LAUNCH();
}

}  // namespace

template <class T>
bool lds_fits(int ng) {
  // Dynamic windows plus the static arrays of each kernel, within 64 KB; the
  // window spans at most kMaxWindow rows and planes (64-bit row masks).
  const size_t wyf = T::fast_pitch((T::TY - 1) * ng + T::WX + T::YPAD),
               wym = (T::TY - 1) * ng + T::WX;
  const size_t wz = (T::TZ - 1) * ng + T::WX;
  if (wyf > (size_t)kMaxWindow || wz > (size_t)kMaxWindow) return false;
  const size_t fast = 4 * T::PX * wyf * wz + T::NB * (16 * 16 + 12 + 16) + 16;
  const size_t masked = 5 * T::PX * wym * wz + masked_static_lds<T>();
  return fast <= 65536 && masked <= 65536;
}

// ---------------------------------------------------------------------------
// Tile shapes and their autotuner. Every shape of one neighbourhood width
// scores every candidate with the same operations in the same order, so the
// records do not depend on the shape; only the time does. The first search
// of a (width, gap, storage, grid) in the process runs each shape that fits
// its LDS once, timed with HIP events on the caller's stream, and keeps the
// fastest for the later searches of that key. MIRORR_TILE_AUTOTUNE=0 keeps
// the first (hand-picked) shape of each width.
namespace {

struct TileVariant {
  int nw;
  bool (*fits)(int ng);
  int64_t (*tiles)(const BMGrid&);
  void (*run_f32)(const BMGrid&, TileArgs, BMScratch&, hipStream_t);
  void (*run_f16)(const BMGrid&, TileArgs, BMScratch&, hipStream_t);
  void (*multi_f32)(const std::vector<SearchJob>&, unsigned char*, hipStream_t);
  void (*multi_f16)(const std::vector<SearchJob>&, unsigned char*, hipStream_t);
  const char* name;  // NW, candidate rows per lane, blocks per tile in y and z
};

#define MIRORR_TILE(NW, KY, TY, TZ)                                                       \
  TileVariant {                                                                             \
    NW, lds_fits<TileCfg<NW, KY, TY, TZ>>, tiles_of<TileCfg<NW, KY, TY, TZ>>,               \
        run_tiled<TileCfg<NW, KY, TY, TZ>, float>, run_tiled<TileCfg<NW, KY, TY, TZ>, half_t>, \
        run_tiled_multi<TileCfg<NW, KY, TY, TZ>, float>,                                    \
        run_tiled_multi<TileCfg<NW, KY, TY, TZ>, half_t>, #NW "/" #KY "/" #TY "x" #TZ        \
  }
// The first shape of each width is the default.
const TileVariant kTileVariants[] = {
    MIRORR_TILE(5, 5, 5, 10),  MIRORR_TILE(5, 5, 10, 5), MIRORR_TILE(5, 5, 6, 6),
    MIRORR_TILE(7, 4, 3, 6),   MIRORR_TILE(7, 4, 6, 3),  MIRORR_TILE(7, 4, 3, 3),
    MIRORR_TILE(9, 3, 3, 3),   MIRORR_TILE(9, 3, 2, 3),  MIRORR_TILE(9, 5, 2, 7),
    MIRORR_TILE(11, 4, 7, 1),  MIRORR_TILE(11, 4, 1, 7), MIRORR_TILE(11, 4, 3, 2),
    MIRORR_TILE(13, 5, 2, 3),  MIRORR_TILE(13, 5, 3, 2), MIRORR_TILE(13, 5, 1, 6),
};
#undef MIRORR_TILE
constexpr int kNumTileVariants = (int)(sizeof(kTileVariants) / sizeof(kTileVariants[0]));

int default_variant(int nw) {
  for (int i = 0; i < kNumTileVariants; ++i)
    if (kTileVariants[i].nw == nw) return i;
  return -1;
}

struct TuneKey {
  int nw, ng, storage, dims[3];
  bool operator<(const TuneKey& o) const {
    return std::tie(nw, ng, storage, dims[0], dims[1], dims[2]) <
           std::tie(o.nw, o.ng, o.storage, o.dims[0], o.dims[1], o.dims[2]);
  }
};
struct TuneResult {
  int variant;
  std::vector<std::pair<int, float>> times;  // (variant, ms) of the trial
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneResult> g_tune;
std::map<int, int> g_forced;  // width -> variant pinned by tile_force

bool autotune_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("MIRORR_TILE_AUTOTUNE");
    return !(e && e[0] == '0');
  }();
  return on;
}

void run_variant(int v, int storage, const BMGrid& g, const TileArgs& a, BMScratch& scr,
                 hipStream_t s) {
  if (storage == kStoreF16)
    kTileVariants[v].run_f16(g, a, scr, s);
  else
    kTileVariants[v].run_f32(g, a, scr, s);
}

// Runs the search with the tuned shape for this key, tuning it first when
// the key is new (each fitting shape once; the records of every trial are
// the same, so the last trial's stand).
void tiled_search(const BMGrid& g, const TileArgs& a, BMScratch& scr, hipStream_t s,
                  int storage) {
  const TuneKey key{g.nw, g.ng, storage, {g.dims[0], g.dims[1], g.dims[2]}};
  int chosen = -1;  // -1: tune now
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto f = g_forced.find(g.nw);
    auto it = g_tune.find(key);
    if (f != g_forced.end() && kTileVariants[f->second].fits(g.ng) &&
        kTileVariants[f->second].tiles(g) <= scr.ntiles_cap)
      chosen = f->second;
    else if (it != g_tune.end())
      chosen = it->second.variant;
    else if (!autotune_enabled() || (scr.prof && scr.prof->on()))
      chosen = default_variant(g.nw);  // not tuned now
  }
  if (chosen < 0) {
    // A stream under graph capture cannot be timed: the default shape is
    // recorded and the key stays untuned.
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    MIRORR_HIP(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone) chosen = default_variant(g.nw);
  }
  if (chosen >= 0) {
    run_variant(chosen, storage, g, a, scr, s);
    return;
  }
  TuneResult res{default_variant(g.nw), {}};
  float best = 0.0f;
  hipEvent_t e0, e1;
  MIRORR_HIP(hipEventCreate(&e0));
  MIRORR_HIP(hipEventCreate(&e1));
  try {
    for (int v = 0; v < kNumTileVariants; ++v) {
      if (kTileVariants[v].nw != g.nw || !kTileVariants[v].fits(g.ng)) continue;
      if (kTileVariants[v].tiles(g) > scr.ntiles_cap) continue;
      MIRORR_HIP(hipEventRecord(e0, s));
      run_variant(v, storage, g, a, scr, s);
      MIRORR_HIP(hipEventRecord(e1, s));
      MIRORR_HIP(hipEventSynchronize(e1));
      float ms = 0.0f;
      MIRORR_HIP(hipEventElapsedTime(&ms, e0, e1));
      res.times.emplace_back(v, ms);
      if (res.times.size() == 1 || ms < best) {
        best = ms;
        res.variant = v;
      }
    }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune.emplace(key, res);
}

}  // namespace

int tile_choice(const BMGrid& g, int storage) {
  const TuneKey key{g.nw, g.ng, storage, {g.dims[0], g.dims[1], g.dims[2]}};
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto f = g_forced.find(g.nw);
  if (f != g_forced.end()) return 1000 + f->second;  // pinned (the cap is checked per launch)
  auto it = g_tune.find(key);
  if (it != g_tune.end()) return it->second.variant;
  return autotune_enabled() ? -1 : default_variant(g.nw);
}

bool tiled_supported(const BMGrid& g) {
  if (g.bw != 4 || g.bg != 1 || g.ng < 1) return false;
  const int v = default_variant(g.nw);
  return v >= 0 && kTileVariants[v].fits(g.ng);
}

int64_t tile_count(const BMGrid& g) {
  int64_t n = 0;
  for (int v = 0; v < kNumTileVariants; ++v)
    if (kTileVariants[v].nw == g.nw) n = std::max(n, kTileVariants[v].tiles(g));
  return n;
}

bool tile_force(int nw, int index) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (index < 0) {
    g_forced.erase(nw);
    return true;
  }
  for (int v = 0, k = 0; v < kNumTileVariants; ++v) {
    if (kTileVariants[v].nw != nw) continue;
    if (k++ == index) {
      g_forced[nw] = v;
      return true;
    }
  }
  return false;
}

std::string tile_autotune_report() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  std::string out;
  for (const auto& kv : g_tune) {
    const TuneKey& k = kv.first;
    char line[256];
    std::snprintf(line, sizeof(line), "nw %d ng %d %s grid %dx%dx%d: %s;", k.nw, k.ng,
                  k.storage == kStoreF16 ? "f16" : "f32", k.dims[0], k.dims[1], k.dims[2],
                  kTileVariants[kv.second.variant].name);
    out += line;
    for (const auto& t : kv.second.times) {
      std::snprintf(line, sizeof(line), " %s %.1f us", kTileVariants[t.first].name,
                    1e3 * t.second);
      out += line;
    }
    out += "\n";
  }
  return out;
}

namespace {
template <class E>
void search_stats4_as(const void* y, const uint8_t* smask, int nx, int ny, int nz, SearchStats st,
                      hipStream_t s) {
  dim3 grid((nx - 3 + 63) / 64, (ny - 3 + 3) / 4, (nz - 3 + kStatsZ - 1) / kStatsZ);
  const E* ye = static_cast<const E*>(y);
  if (smask)
    hipLaunchKernelGGL((search_stats4_kernel<true, E>), grid, dim3(kThreads), 0, s, ye, smask, nx,
                       ny, nz, st.ms);
  else
    hipLaunchKernelGGL((search_stats4_kernel<false, E>), grid, dim3(kThreads), 0, s, ye, smask,
                       nx, ny, nz, st.ms);
}
}  // namespace

void launch_search_stats(const void* y, const uint8_t* smask, const BMGrid& g,
                         SearchStats st, hipStream_t s, int storage) {
  const int nx = g.dims[0], ny = g.dims[1], nz = g.dims[2];
  if (nx < g.bw || ny < g.bw || nz < g.bw) return;
  if (g.bw == 4) {
    if (storage == kStoreF16)
      search_stats4_as<half_t>(y, smask, nx, ny, nz, st, s);
    else
      search_stats4_as<float>(y, smask, nx, ny, nz, st, s);
  } else {
    if (storage == kStoreF16) throw Error(kErrArg, "half-precision images need 4-voxel blocks");
    dim3 grid((nx + kThreads - 1) / kThreads, ny - g.bw + 1, nz - g.bw + 1);
    hipLaunchKernelGGL(search_stats_kernel<0>, grid, dim3(kThreads), 0, s,
                       static_cast<const float*>(y), smask, nx, ny, nz, g.bw, st.ms);
  }
  MIRORR_CHECK_LAUNCH();
}

bool launch_variance_select(const BMGrid& g, const BlockPrep& bp, double portion,
                            VarSelScratch& scr, hipStream_t s) {
  const int64_t n = g.nblocks;
  if (n <= 0) return false;
  int64_t K = (int64_t)std::ceil(portion * (double)n);
  K = std::min<int64_t>(std::max<int64_t>(K, 1), n);
  if (K >= n) return false;
  const unsigned grid = (unsigned)std::min<int64_t>((n + kThreads - 1) / kThreads, 1024);
  hipLaunchKernelGGL(vsel_init, dim3(1), dim3(kThreads), 0, s, scr.st, scr.hist,
                     (unsigned long long)(n - K));
  hipLaunchKernelGGL(vsel_hist<0>, dim3(grid), dim3(kThreads), 0, s, bp.sxx, n, scr.st, scr.hist);
  hipLaunchKernelGGL(vsel_scan<0>, dim3(1), dim3(kThreads), 0, s, scr.st, scr.hist);
  hipLaunchKernelGGL(vsel_hist<1>, dim3(grid), dim3(kThreads), 0, s, bp.sxx, n, scr.st, scr.hist);
  hipLaunchKernelGGL(vsel_scan<1>, dim3(1), dim3(kThreads), 0, s, scr.st, scr.hist);
  hipLaunchKernelGGL(vsel_hist<2>, dim3(grid), dim3(kThreads), 0, s, bp.sxx, n, scr.st, scr.hist);
  hipLaunchKernelGGL(vsel_scan<2>, dim3(1), dim3(kThreads), 0, s, scr.st, scr.hist);
  hipLaunchKernelGGL(vsel_mark, dim3(grid), dim3(kThreads), 0, s, bp.sxx, bp.flags, g, scr.st,
                     scr.evals);
  MIRORR_CHECK_LAUNCH();
  return true;
}

void launch_block_prep(const void* x, const uint8_t* bmask, const BMGrid& g,
                       BlockPrep bp, hipStream_t s, int storage) {
  if (g.nblocks == 0) return;
  if (g.nblocks > (int64_t)INT_MAX / 2) throw Error(kErrArg, "block grid too large");
  if (g.bw == 4) {
    dim3 grid((unsigned)((g.nblocks * 4 + kThreads - 1) / kThreads));
    if (storage == kStoreF16)
      hipLaunchKernelGGL(block_prep4_kernel<half_t>, grid, dim3(kThreads), 0, s,
                         static_cast<const half_t*>(x), bmask, g, bp.xt, bp.sxx, bp.sumxt,
                         bp.flags);
    else
      hipLaunchKernelGGL(block_prep4_kernel<float>, grid, dim3(kThreads), 0, s,
                         static_cast<const float*>(x), bmask, g, bp.xt, bp.sxx, bp.sumxt,
                         bp.flags);
  } else {
    if (storage == kStoreF16) throw Error(kErrArg, "half-precision images need 4-voxel blocks");
    const int waves = kThreads / 64;
    dim3 grid((unsigned)((g.nblocks + waves - 1) / waves));
    hipLaunchKernelGGL(block_prep_kernel<0>, grid, dim3(kThreads), 0, s,
                       static_cast<const float*>(x), bmask, g, bp.xt, bp.sxx, bp.sumxt, bp.flags);
  }
  MIRORR_CHECK_LAUNCH();
}

void launch_ncc_generic(const float* x, const uint8_t* bmask, const float* y,
                        const uint8_t* smask, const BMGrid& g, BMResult res,
                        hipStream_t s, const uint32_t* flags) {
  if (g.nblocks == 0) return;
  hipLaunchKernelGGL(ncc_generic_kernel, dim3((unsigned)g.nblocks), dim3(kThreads), 0, s, x,
                     bmask, y, smask, g, res.best, res.second, res.match, flags);
  MIRORR_CHECK_LAUNCH();
}

void launch_metric_search(int metric, const float* x, const uint8_t* bmask, const float* y,
                          const uint8_t* smask, const BMGrid& g, BMResult res, hipStream_t s,
                          const uint32_t* flags) {
  if (g.nblocks == 0) return;
  if (g.bw > kMaxMetricBW) throw Error(kErrArg, "block width above 16 for this block metric");
  const size_t lds = (size_t)g.bw * g.bw * g.bw * (sizeof(float) + 1);
  if (metric == MIRORR_METRIC_SSD) {
    hipLaunchKernelGGL(ssd_generic_kernel, dim3((unsigned)g.nblocks), dim3(kThreads), lds, s, x,
                       bmask, y, smask, g, res.best, res.second, res.match, flags);
  } else if (metric == MIRORR_METRIC_CR) {
    hipLaunchKernelGGL(cr_generic_kernel, dim3((unsigned)g.nblocks), dim3(kThreads), lds, s, x,
                       bmask, y, smask, g, res.best, res.second, res.match, flags);
  } else {
    launch_ncc_generic(x, bmask, y, smask, g, res, s, flags);
    return;
  }
  MIRORR_CHECK_LAUNCH();
}

void launch_ncc_search(const void* x, const uint8_t* bmask, const void* y,
                       const uint8_t* smask, const BMGrid& g, const BlockPrep& bp,
                       const SearchStats& st, BMScratch& scr, BMResult res,
                       hipStream_t s, int storage) {
  if (g.nblocks == 0) return;
  if (!tiled_supported(g)) {
    if (storage == kStoreF16)
      throw Error(kErrArg, "half-precision images need the tiled search (blocks of width 4, "
                           "block gap 1, nhood width 5 to 13)");
    launch_ncc_generic(static_cast<const float*>(x), bmask, static_cast<const float*>(y), smask,
                       g, res, s, bp.flags);
    return;
  }
  TileArgs a{};
  a.y = y;
  a.sm = smask;
  a.bm = bmask;
  a.g = g;
  a.xt = bp.xt;
  a.sxx = bp.sxx;
  a.sumxt = bp.sumxt;
  a.bflags = bp.flags;
  a.stats = st.ms;
  a.out_best = res.best;
  a.out_second = res.second;
  a.out_match = res.match;
  tiled_search(g, a, scr, s, storage);
  if (scr.rlist) {
    // Blocks listed by the search kernels are searched again in double.
    RescoreArgs ra{x, bmask, y, smask, st.ms, scr.rescored, scr.exact_records};
    const int64_t grid = std::min<int64_t>(g.nblocks, 1024);
    if (storage == kStoreF16)
      hipLaunchKernelGGL(rescore_list_kernel<half_t>, dim3((unsigned)grid), dim3(kThreads), 0, s,
                         g, res.best, res.second, res.match, ra, scr.rlist, scr.rcount);
    else
      hipLaunchKernelGGL(rescore_list_kernel<float>, dim3((unsigned)grid), dim3(kThreads), 0, s,
                         g, res.best, res.second, res.match, ra, scr.rlist, scr.rcount);
    MIRORR_CHECK_LAUNCH();
  }
}

size_t search_batch_bytes(int njobs) { return BatchLayout(std::max(njobs, 1)).bytes; }

void launch_ncc_search_batch(const std::vector<SearchJob>& jobs, void* dev, size_t dev_bytes,
                             hipStream_t s, int storage) {
  if (jobs.empty()) return;
  if (dev_bytes < search_batch_bytes((int)jobs.size()))
    throw Error(kErrState, "batched search argument buffer too small");
  const BMGrid& g0 = jobs[0].g;
  for (const SearchJob& j : jobs) {
    if (j.g.bw != g0.bw || j.g.bg != g0.bg || j.g.nw != g0.nw || j.g.ng != g0.ng)
      throw Error(kErrArg, "a batched search needs one block and neighbourhood geometry");
    if (!tiled_supported(j.g)) throw Error(kErrArg, "a batched search needs the tiled search");
    if (j.g.nblocks > (int64_t)INT_MAX / 2) throw Error(kErrArg, "block grid too large");
  }
  // One shape for the whole batch: the pinned one where it fits every pair,
  // else the width's default.
  int v = default_variant(g0.nw);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto f = g_forced.find(g0.nw);
    if (f != g_forced.end() && kTileVariants[f->second].fits(g0.ng)) {
      bool ok = true;
      for (const SearchJob& j : jobs) ok = ok && kTileVariants[f->second].tiles(j.g) <= j.scr.ntiles_cap;
      if (ok) v = f->second;
    }
  }
  if (storage == kStoreF16)
    kTileVariants[v].multi_f16(jobs, static_cast<unsigned char*>(dev), s);
  else
    kTileVariants[v].multi_f32(jobs, static_cast<unsigned char*>(dev), s);
}

void launch_postprocess(const BMGrid& g, const Geom& geom, const BMResult& res,
                        uint8_t* kept, double* base_mm, double* match_mm,
                        hipStream_t s) {
  if (g.nblocks == 0) return;
  hipLaunchKernelGGL(postprocess_kernel, dim3((unsigned)((g.nblocks + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, s, g, geom, res.best, res.second, res.match, kept,
                     base_mm, match_mm);
  MIRORR_CHECK_LAUNCH();
}

void launch_postprocess_pairs(const BMGrid& g, const Geom& geom, const BMResult& res, int swap,
                              uint8_t* kept, double* X, double* Y, unsigned long long* count,
                              hipStream_t s) {
  if (g.nblocks == 0) return;
  const int64_t grid = std::min<int64_t>((g.nblocks + kThreads - 1) / kThreads, 512);
  hipLaunchKernelGGL(postprocess_pairs_kernel, dim3((unsigned)grid), dim3(kThreads), 0, s, g, geom,
                     res.best, res.second, res.match, swap, kept, X, Y, count);
  MIRORR_CHECK_LAUNCH();
}

int64_t compact_chunks(int64_t n) { return (n + kChunk - 1) / kChunk; }

void launch_compact_index(const uint8_t* flags, int64_t n, Compactor& c,
                          int32_t* out_index, hipStream_t s) {
  const int64_t nc = compact_chunks(n);
  if (nc == 0) return;
  if (nc > c.cap_chunks) throw Error(kErrState, "compaction scratch too small");
  hipLaunchKernelGGL(compact_count, dim3((unsigned)nc), dim3(kThreads), 0, s, flags, n,
                     c.chunk_counts);
  hipLaunchKernelGGL(compact_scan, dim3(1), dim3(1024), 0, s, c.chunk_counts, (int)nc,
                     c.chunk_offsets);
  hipLaunchKernelGGL(compact_scatter, dim3((unsigned)nc), dim3(kThreads), 0, s, flags, n,
                     c.chunk_offsets, out_index);
  MIRORR_CHECK_LAUNCH();
}

}  // namespace mirorr
