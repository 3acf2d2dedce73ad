// Least-trimmed-squares regression of block-match pairs on gfx950.
//
// Restates OptimizeTransform of the classic (itkMirorrRegistrationMethod.txx:758-1106)
// and symmetric (itkSymMirorrRegistrationMethod.txx:729-1442) methods with the
// per-point work on the GPU and the tiny FP64 solves in a one-thread tail.
// Per inner iteration (the reductions finish in one-workgroup launches or
// are repeated by every workgroup of the next kernel: kernel boundaries
// publish the partial results, where a per-workgroup agent-scope fence
// would write back L2 for every workgroup):
//  1. residual kernel: r_i = |F(x_i) - B(y_i)|^2 (classic |F(x_i) - y_i|^2),
//     +inf for culled blocks; the first radix digit (bits 63..52 of the
//     non-negative double) is histogrammed in LDS and merged globally (not
//     launched in a run's first inner iteration, which keeps every kept
//     pair: the moments read the kept flags then);
//  2. exact k-th smallest residual (std::nth_element in the reference): the
//     bin of rank k, a second 12-bit digit inside it, a compaction of the
//     candidates sharing the 24-bit prefix and a one-workgroup finish by rank
//     counting or 8-bit radix passes;
//  3. trimmed moments in one pass (count, sums and the raw second moments
//     x x^T, y y^T, x y^T), block-reduced in double, totalled in a fixed
//     order and centred in the solve; the rigid quaternion matrix sum Q^T Q
//     is assembled from those moments;
//  4. a one-thread solve (normal equations / quaternion eigenvector / polar
//     factor / Denman-Beavers roots) updating the device state.
// Every kernel is predicated on the state's `active` flag so the host can
// queue several inner iterations before reading that flag back.
#include <cfloat>
#include <climits>

#include "lts.hpp"
#include "select.hpp"

namespace mirorr {
namespace {

constexpr int kT = sel::kSelT;
// Points per moment workgroup: kPart (4 per thread), or twice that (8 per
// thread, half the partials for the solve to total) for runs of more than
// kBigRun pairs; fewer, larger workgroups are slower on the smaller runs.
constexpr int kPart = 1024;
constexpr int64_t kBigRun = 262144;
__host__ __device__ __forceinline__ int part_points(int64_t npts) {
  return npts > kBigRun ? 2 * kPart : kPart;
}

__device__ __forceinline__ int64_t lts_parts_dev(int64_t npts) {
  const int p = part_points(npts);
  return (npts + p - 1) / p;
}

using namespace xf;
using sel::block_scan;

__device__ __forceinline__ unsigned long long key_of(double r) {
  return (unsigned long long)__double_as_longlong(r);
}

// ---- residuals + first digit ----------------------------------------------
__device__ __forceinline__ void residual_body(const double* __restrict__ X,
                                              const double* __restrict__ Y,
                                              const uint8_t* __restrict__ kept, int64_t n,
                                              LtsState* st, double* __restrict__ r,
                                              unsigned int* hist1, int do_hist) {
  if (!st->active) return;
  __shared__ unsigned int h[4096];
  if (do_hist)
    for (int i = threadIdx.x; i < 4096; i += kT) h[i] = 0;
  __syncthreads();
  const Xf F = st->F;
  const Xf B = st->B;
  const bool sym = st->symmetric != 0;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    double v = __longlong_as_double(0x7ff0000000000000ll);  // +inf: culled block
    if (kept[i]) {
      double p[3], q[3];
      apply(F, X + 3 * i, p);
      if (sym) {
        apply(B, Y + 3 * i, q);
      } else {
        q[0] = Y[3 * i];
        q[1] = Y[3 * i + 1];
        q[2] = Y[3 * i + 2];
      }
      v = 0.0;
      for (int k = 0; k < 3; ++k) {
        const double d = p[k] - q[k];
        v += d * d;
      }
      if (do_hist) atomicAdd(&h[(unsigned)(key_of(v) >> 52)], 1u);
    }
    r[i] = v;
  }
  if (!do_hist) return;
  __syncthreads();
  for (int i = threadIdx.x; i < 4096; i += kT)
    if (h[i]) atomicAdd(&hist1[i], h[i]);
}

// Second digit (bits 51..40) among residuals in the first digit's bin. Each
// workgroup finds that bin (rank k in the first-digit histogram) itself;
// workgroup 0 records it.
__device__ __forceinline__ void hist2_body(const double* __restrict__ r, int64_t n, LtsState* st,
                                           const unsigned int* hist1, unsigned int* hist2) {
  if (!st->active) return;
  __shared__ unsigned int h[4096];
  __shared__ unsigned int s_bin;
  __shared__ unsigned long long s_krem;
  for (int i = threadIdx.x; i < 4096; i += kT) h[i] = 0;
  sel::find_bin<4096 / kT>(hist1, st->k, &s_bin, &s_krem);
  __syncthreads();
  const unsigned int b1 = s_bin;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->bin1 = b1;
    st->k2 = s_krem;
    st->ncand = 0;
  }
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    const unsigned long long k = key_of(r[i]);
    if ((unsigned)(k >> 52) == b1) atomicAdd(&h[(unsigned)(k >> 40) & 0xfffu], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 4096; i += kT)
    if (h[i]) atomicAdd(&hist2[i], h[i]);
}

// Rank k3 among the candidate keys (all share the 24-bit prefix): by rank
// counting when they are few, else by 8-bit radix passes on bits 39..0.
__device__ void final_select(const unsigned long long* cand, LtsState* st) {
  constexpr int kSmall = 512;
  __shared__ unsigned long long sc[kSmall];
  __shared__ unsigned int hb[kT];
  __shared__ unsigned long long s_prefix, s_k;
  const int t = threadIdx.x;
  const unsigned int nc = __hip_atomic_load(&st->ncand, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
  const unsigned long long k = st->k3;
  if (nc <= (unsigned)kSmall) {
    for (unsigned int i = t; i < nc; i += kT) sc[i] = cand[i];
    __syncthreads();
    for (unsigned int i = t; i < nc; i += kT) {
      const unsigned long long v = sc[i];
      unsigned int rank = 0;
      for (unsigned int j = 0; j < nc; ++j) {
        const unsigned long long u = sc[j];
        rank += (u < v || (u == v && j < i)) ? 1u : 0u;
      }
      if (rank == k) st->thr = __longlong_as_double((long long)v);
    }
    return;
  }
  if (t == 0) {
    s_prefix = (((unsigned long long)st->bin1 << 12) | st->bin2) << 40;
    s_k = k;
  }
  for (int shift = 32; shift >= 0; shift -= 8) {
    hb[t] = 0;
    __syncthreads();
    const unsigned long long pre = s_prefix, kk = s_k;
    const unsigned long long hi_mask = ~((1ull << (shift + 8)) - 1ull);
    for (unsigned int i = t; i < nc; i += kT) {
      const unsigned long long c = cand[i];
      if ((c & hi_mask) == pre) atomicAdd(&hb[(unsigned)(c >> shift) & 0xffu], 1u);
    }
    __syncthreads();
    const unsigned long long cnt = hb[t];
    const unsigned long long inc = block_scan(cnt);
    const unsigned long long before = inc - cnt;
    if (kk >= before && kk < inc) {
      s_prefix = pre | ((unsigned long long)t << shift);
      s_k = kk - before;
    }
    __syncthreads();
  }
  if (t == 0) st->thr = __longlong_as_double((long long)s_prefix);
}

// Candidates sharing the 24-bit prefix of the k-th residual. Each workgroup
// finds the second digit's bin (rank k2 in the second histogram) itself;
// workgroup 0 records it.
__device__ __forceinline__ void compact_body(const double* __restrict__ r, int64_t n,
                                             LtsState* st, const unsigned int* hist2,
                                             unsigned long long* __restrict__ cand) {
  if (!st->active) return;
  __shared__ unsigned int s_bin;
  __shared__ unsigned long long s_krem;
  sel::find_bin<4096 / kT>(hist2, st->k2, &s_bin, &s_krem);
  __syncthreads();
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->bin2 = s_bin;
    st->k3 = s_krem;
  }
  const unsigned long long pre = ((unsigned long long)st->bin1 << 12) | s_bin;
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) {
    const unsigned long long k = key_of(r[i]);
    if ((k >> 40) == pre) cand[atomicAdd(&st->ncand, 1u)] = k;
  }
}

__global__ __launch_bounds__(kT) void final_select_kernel(const unsigned long long* cand,
                                                          LtsState* st) {
  if (!st->active) return;
  final_select(cand, st);
}

// ---- trimmed moments ---------------------------------------------------------
// Butterfly over the wave with the step outermost: the N shuffles of a step
// are independent, so they issue together and wait on LDS once per step
// (the sums per value are the same as one value at a time).
template <int N>
__device__ void block_reduce(double (&v)[N], double* out) {
  __shared__ double sh[kT / 64][N];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int j = 0; j < N; ++j) v[j] += __shfl_xor(v[j], o, 64);
  if (lane == 0)
#pragma unroll
    for (int j = 0; j < N; ++j) sh[w][j] = v[j];
  __syncthreads();
  if (threadIdx.x < N) {
    double s = 0.0;
    for (int k = 0; k < kT / 64; ++k) s += sh[k][threadIdx.x];
    out[threadIdx.x] = s;
  }
}

// Sums per-workgroup partials in a fixed order (deterministic). Thread t
// handles value t % 32 of the partials t / 32, t / 32 + 8, ... (eight loads
// in flight at a time), then the eight groups are added in order.
template <int N>
__device__ void reduce_parts(const double* part, int stride, int nparts, double* out) {
  static_assert(N <= 32 && kT == 256, "32 values x 8 groups");
  __shared__ double grp[8][32];
  const int j = threadIdx.x & 31, gi = threadIdx.x >> 5;
  double s = 0.0;
  if (j < N) {
    int p = gi;
    for (; p + 56 < nparts; p += 64) {  // eight independent loads per step
      double a[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) a[u] = part[(int64_t)stride * (p + 8 * u) + j];
#pragma unroll
      for (int u = 0; u < 8; ++u) s += a[u];
    }
    for (; p < nparts; p += 8) s += part[(int64_t)stride * p + j];
  }
  grp[gi][j] = s;
  __syncthreads();
  if (threadIdx.x < N) {
    double t = 0.0;
    for (int k = 0; k < 8; ++k) t += grp[k][threadIdx.x];
    out[threadIdx.x] = t;
  }
  __syncthreads();
}

// One pass over the trimmed pairs: count, coordinate sums and the raw
// second moments x x^T (6), y y^T (6), x y^T (9), block-reduced in double.
// They are centred in the solve (sum (x-mx)(x-mx)^T = sum x x^T - n mx mx^T):
// coordinates are mm within a few hundred of the origin, so the subtraction
// costs a few digits of the 16 FP64 carries.
constexpr int kMoments = 28;   // values per workgroup partial
constexpr int kPartStride = 32;

// first: the run's first inner iteration, which keeps every kept pair (the
// reference keeps the pairs with r <= DBL_MAX, i.e. the finite residuals,
// which are the kept pairs'): the kept flags stand for the residuals, whose
// kernel is not launched then.
template <int PER>  // points per thread
__device__ __forceinline__ void moments_body(const double* __restrict__ X,
                                             const double* __restrict__ Y,
                                             const uint8_t* __restrict__ kept,
                                             const double* __restrict__ r, int64_t n,
                                             LtsState* st, double* part, int first) {
  if (!st->active || n < 1) return;
  const double thr = st->fraction > 0.99999 ? DBL_MAX : st->thr;
  double v[kMoments];
#pragma unroll
  for (int j = 0; j < kMoments; ++j) v[j] = 0.0;
  const int64_t base = (int64_t)blockIdx.x * (PER * kT);
  // The thread's points are loaded together (one round trip), then summed
  // in order.
  constexpr int kPer = PER;
  double xs[kPer][3], ys[kPer][3];
  bool in[kPer];
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int64_t i = min(base + threadIdx.x + (int64_t)u * kT, n - 1);
    in[u] = base + threadIdx.x + (int64_t)u * kT < n && (first ? kept[i] != 0 : r[i] <= thr);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      xs[u][k] = X[3 * i + k];
      ys[u][k] = Y[3 * i + k];
    }
  }
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    if (in[u]) {
      const double* x = xs[u];
      const double* y = ys[u];
      v[0] += 1.0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        v[1 + k] += x[k];
        v[4 + k] += y[k];
      }
      int q = 0;
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = a; b < 3; ++b, ++q) {
          v[7 + q] += x[a] * x[b];
          v[13 + q] += y[a] * y[b];
        }
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) v[19 + 3 * a + b] += x[a] * y[b];
    }
  }
  block_reduce<kMoments>(v, part + kPartStride * blockIdx.x);
}

// ---- the one-thread solves -------------------------------------------------
struct Moments {
  double n, mx[3], my[3];
  double xx[3][3], yy[3][3], xy[3][3];  // centred sums (not divided by n)
};

__device__ void unpack(const double* s1, const double* s2, Moments& m) {
  m.n = s1[0];
  const double in = 1.0 / m.n;  // the solve runs on one lane: one division per divisor
  for (int k = 0; k < 3; ++k) {
    m.mx[k] = s1[1 + k] * in;
    m.my[k] = s1[4 + k] * in;
  }
  int q = 0;
  for (int a = 0; a < 3; ++a)
    for (int b = a; b < 3; ++b, ++q) {
      m.xx[a][b] = m.xx[b][a] = s2[q];
      m.yy[a][b] = m.yy[b][a] = s2[6 + q];
    }
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) m.xy[a][b] = s2[12 + 3 * a + b];
}

// CalculateRigidQuaternion from moments: sum Q^T Q with a = x - y, b = x + y.
__device__ Quat rigid_quat(const Moments& m) {
  double aa[3][3], bb[3][3], ab[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      aa[i][j] = m.xx[i][j] - m.xy[i][j] - m.xy[j][i] + m.yy[i][j];
      bb[i][j] = m.xx[i][j] + m.xy[i][j] + m.xy[j][i] + m.yy[i][j];
      ab[i][j] = m.xx[i][j] + m.xy[i][j] - m.xy[j][i] - m.yy[i][j];
    }
  double T[4][4];
  T[0][0] = aa[0][0] + aa[1][1] + aa[2][2];
  T[0][1] = -ab[1][2] + ab[2][1];
  T[0][2] = ab[0][2] - ab[2][0];
  T[0][3] = -ab[0][1] + ab[1][0];
  T[1][1] = aa[0][0] + bb[2][2] + bb[1][1];
  T[1][2] = aa[0][1] - bb[0][1];
  T[1][3] = aa[0][2] - bb[0][2];
  T[2][2] = aa[1][1] + bb[2][2] + bb[0][0];
  T[2][3] = aa[1][2] - bb[1][2];
  T[3][3] = aa[2][2] + bb[1][1] + bb[0][0];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < i; ++j) T[i][j] = T[j][i];
  double ev[4], V[4][4];
  eigen_sym<4>(T, ev, V);
  return Quat{V[1][0], V[2][0], V[3][0], V[0][0]};
}

// CovarianceOfPoints(a, b) of centred sets: sum a b^T / n.
__device__ M3 cov(const double s[3][3], double n, bool transpose_) {
  const double in = 1.0 / n;
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = (transpose_ ? s[j][i] : s[i][j]) * in;
  return r;
}

__device__ double delta_of(const double* a, const double* b, int np) {
  double s = 0.0;
  for (int i = 0; i < np; ++i) s += (a[i] - b[i]) * (a[i] - b[i]);
  return s;
}

__device__ M4 to4(const M3& m, const double t[3]) {
  M4 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) r.a[i][j] = (i == j && i == 3) ? 1.0 : 0.0;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) r.a[i][j] = m.a[i][j];
    r.a[i][3] = t[i];
  }
  return r;
}
__device__ M3 top3(const M4& m) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = m.a[i][j];
  return r;
}

// Symmetric update (itkSymMirorrRegistrationMethod.txx:866-1379).
// Run by lanes 0 and 1 of one wave on identical inputs: they compute the
// same values, except that the two Denman-Beavers roots run one per lane;
// lane 0 writes the state. The state's parameters and delta are read once,
// at the start (no reload after the fits are stored); *dlast and *dnew get
// the previous and the new parameter change.
__device__ void solve_symmetric(LtsState* st, const Moments& m, int lane, double* dlast,
                                double* dnew) {
  const int kind = st->kind, np = kind == kRigid ? 6 : 12;
  const double delta0 = st->delta;
  double neu[12], neuF[12], neuB[12], last0[12];
  for (int i = 0; i < np; ++i) {
    last0[i] = st->last[i];
    neu[i] = neuF[i] = neuB[i] = last0[i];
  }
  double tr[3];
  for (int k = 0; k < 3; ++k) tr[k] = m.my[k] - m.mx[k];
  Xf fF = st->F, fB = st->B;
  if (kind == kAffine) {
    const M3 oxy = cov(m.xy, m.n, true), oxx = cov(m.xx, m.n, false);
    M3 rF = oxy;
    if (fabs(det(oxx)) >= 1e-4) rF = mul(oxy, inverse(oxx));
    const M3 oyx = cov(m.xy, m.n, false), oyy = cov(m.yy, m.n, false);
    M3 rB = oyx;
    if (fabs(det(oyy)) >= 1e-4) rB = mul(oyx, inverse(oyy));
    for (int i = 0; i < 9; ++i) {
      neu[i] = neuF[i] = rF.a[i / 3][i % 3];
      neuB[i] = rB.a[i / 3][i % 3];
    }
    for (int k = 0; k < 3; ++k) {
      neu[9 + k] = neuF[9 + k] = tr[k];
      neuB[9 + k] = -tr[k];
    }
    set_center(fF, m.mx);
    set_params(fF, neuF);
    set_center(fB, m.my);
    set_params(fB, neuB);
    M3 K = fF.M, Mi = inverse(fF.M);
    const M3 L = inverse(fB.M), N = fB.M;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        K.a[i][j] = (K.a[i][j] + L.a[i][j]) / 2.0;
        Mi.a[i][j] = (Mi.a[i][j] + N.a[i][j]) / 2.0;
      }
    set_matrix(fF, K);
    set_matrix(fB, Mi);
  } else {
    const Quat q = rigid_quat(m);
    double eu[3], euF[3], euB[3];
    quat_euler_zyx(q, eu);
    const double vm = sqrt(q.x * q.x + q.y * q.y + q.z * q.z);
    double ax[3] = {0.0, 0.0, 1.0};
    if (vm != 0.0) {
      ax[0] = q.x / vm;
      ax[1] = q.y / vm;
      ax[2] = q.z / vm;
    }
    const double angle = 2.0 * atan2(vm, q.r);
    quat_euler_zyx(quat_axis_angle(ax, angle), euF);
    quat_euler_zyx(quat_axis_angle(ax, -angle), euB);
    for (int k = 0; k < 3; ++k) {
      neu[k] = eu[k];
      neuF[k] = euF[k];
      neuB[k] = euB[k];
      neu[3 + k] = neuF[3 + k] = tr[k];
      neuB[3 + k] = -tr[k];
    }
    set_center(fF, m.mx);
    set_params(fF, neuF);
    set_center(fB, m.my);
    set_params(fB, neuB);
    M3 K = fF.M;
    const M3 L = inverse(fB.M);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) K.a[i][j] = (K.a[i][j] + L.a[i][j]) / 2.0;
    K = orthogonalize(K);
    set_matrix(fF, K);
    set_matrix(fB, transpose(K));
  }
  const double zero[3] = {0.0, 0.0, 0.0};
  {
    const double oF[3] = {fF.off[0], fF.off[1], fF.off[2]};
    const double oB[3] = {fB.off[0], fB.off[1], fB.off[2]};
    set_center(fF, zero);
    set_center(fB, zero);
    set_translation(fF, oF);
    set_translation(fB, oB);
  }
  const Xf iF = inverse(fF), iB = inverse(fB);
  M3 K, K2;
  double tF[3], tB[3];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      K.a[i][j] = (fF.M.a[i][j] + iB.M.a[i][j]) / 2.0;
      K2.a[i][j] = (iF.M.a[i][j] + fB.M.a[i][j]) / 2.0;
    }
    tF[i] = (fF.t[i] + iB.t[i]) / 2.0;
    tB[i] = (iF.t[i] + fB.t[i]) / 2.0;
  }
  if (kind == kRigid) {
    K = orthogonalize(K);
    K2 = orthogonalize(K2);
  }
  __shared__ M4 roots[2];
  roots[lane] = sqrt_db(lane == 0 ? to4(K, tF) : to4(K2, tB), 1e-14, 50);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  const M4 m12 = roots[0];
  const M4 im12 = roots[1];
  M3 A = top3(m12), Bm = top3(im12);
  Xf sF = fF, sB = fB;
  if (!set_matrix(sF, A) || !set_matrix(sB, Bm)) {
    // the square roots were not orthogonal enough for a rigid transform
    A = orthogonalize(A);
    Bm = orthogonalize(Bm);
    sF = fF;
    sB = fB;
    set_matrix(sF, A);
    set_matrix(sB, Bm);
    if (lane == 0) st->status |= 1;
  }
  const double t1[3] = {m12.a[0][3], m12.a[1][3], m12.a[2][3]};
  const double t2[3] = {im12.a[0][3], im12.a[1][3], im12.a[2][3]};
  set_translation(sF, t1);
  set_translation(sB, t2);
  if (lane != 0) return;
  st->F = sF;
  st->B = sB;
  *dlast = delta0;
  *dnew = delta_of(neu, last0, np);
  st->last_delta = *dlast;
  st->delta = *dnew;
  for (int i = 0; i < np; ++i) st->last[i] = neu[i];
}

// Classic update (itkMirorrRegistrationMethod.txx:841-999); last_parameters
// keeps the identity as the reference never updates it.
__device__ void solve_classic(LtsState* st, const Moments& m, double* dlast, double* dnew) {
  const int kind = st->kind, np = kind == kRigid ? 6 : 12;
  const double delta0 = st->delta;
  double neu[12], last0[12];
  for (int i = 0; i < np; ++i) {
    last0[i] = st->last[i];
    neu[i] = last0[i];
  }
  double a[3], b[3], ra[3];
  for (int k = 0; k < 3; ++k) {
    a[k] = m.my[k] - st->rc[k];
    b[k] = m.mx[k] - st->rc[k];
  }
  if (kind == kAffine) {
    const M3 oxy = cov(m.xy, m.n, true), oxx = cov(m.xx, m.n, false);
    M3 rot = oxy;
    if (fabs(det(oxx)) >= 1e-4) rot = mul(rot, inverse(oxx));
    mulv(rot, b, ra);
    for (int i = 0; i < 9; ++i) neu[i] = rot.a[i / 3][i % 3];
    for (int k = 0; k < 3; ++k) neu[9 + k] = a[k] - ra[k];
  } else {
    const Quat q = rigid_quat(m);
    const M3 rot = quat_rotation(q);
    double ang[3];
    mulv(rot, b, ra);
    quat_euler_zyx(q, ang);
    for (int k = 0; k < 3; ++k) {
      neu[k] = ang[k];
      neu[3 + k] = a[k] - ra[k];
    }
  }
  *dlast = delta0;
  *dnew = delta_of(neu, last0, np);
  st->last_delta = *dlast;
  st->delta = *dnew;
  for (int k = 0; k < 3; ++k) st->rc[k] = m.mx[k];
  set_center(st->F, st->rc);
  set_params(st->F, neu);
}

// Totals of the moments pass (four waves), cleared histograms for the next
// iteration, then the solve on lanes 0 and 1 (its register footprint stays
// out of the streaming kernels).
__device__ __forceinline__ void solve_body(LtsState* st, const double* part, int nparts,
                                           unsigned int* hist1, unsigned int* hist2) {
  if (!st->active) return;
  const int iter0 = st->iter;
  const double tol = st->tol;
  __shared__ double t[kMoments];
  reduce_parts<kMoments>(part, kPartStride, nparts, t);
  for (int i = threadIdx.x; i < 4096; i += blockDim.x) {
    hist1[i] = 0;
    hist2[i] = 0;
  }
  const int lane = threadIdx.x;
  if (lane >= 2 || (lane == 1 && !st->symmetric)) return;
  // Centred second moments in the layout unpack() reads: xx (6), yy (6), xy (9).
  double s2[21];
  const double cnt = t[0], icnt = 1.0 / cnt;
  int q = 0;
  for (int a = 0; a < 3; ++a)
    for (int b = a; b < 3; ++b, ++q) {
      s2[q] = t[7 + q] - t[1 + a] * t[1 + b] * icnt;
      s2[6 + q] = t[13 + q] - t[4 + a] * t[4 + b] * icnt;
    }
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) s2[12 + 3 * a + b] = t[19 + 3 * a + b] - t[1 + a] * t[4 + b] * icnt;
  Moments m;
  unpack(t, s2, m);
  double dlast = 0.0, dnew = 0.0;
  if (st->symmetric)
    solve_symmetric(st, m, lane, &dlast, &dnew);
  else
    solve_classic(st, m, &dlast, &dnew);
  if (lane != 0) return;
  const int iter = iter0 + 1;  // the loop head of the reference
  st->iter = iter;
  st->fraction = 0.5;
  st->active = (dnew > tol && iter < 100 && fabs(dnew - dlast) > tol) ? 1 : 0;
}


__device__ __forceinline__ void init_body(LtsState* st, const OuterState* outer, int kind,
                                          int symmetric, double tol, unsigned int* hist1,
                                          unsigned int* hist2) {
  // Runs with one workgroup; thread 0 writes the state.
  for (int i = threadIdx.x; i < 4096; i += blockDim.x) {
    hist1[i] = 0;
    hist2[i] = 0;
  }
  if (threadIdx.x != 0) return;
  st->kind = kind;
  st->symmetric = symmetric;
  st->tol = tol;
  const double* c = outer->T.c;
  st->F = identity(kind);
  set_center(st->F, c);
  st->B = st->F;
  for (int k = 0; k < 3; ++k) st->rc[k] = c[k];
  get_params(st->F, st->last);
  st->fraction = 1.0;
  st->delta = 1.0;
  st->last_delta = 0.0;
  st->iter = -1;
  st->status = 0;
  const unsigned long long n = st->n_kept;
  unsigned long long k = (unsigned long long)(0.5 * (double)n);
  if (n > 0 && k > n - 1) k = n - 1;
  st->k = k;
  st->ncand = 0;
  st->thr = DBL_MAX;
  st->finished = 0;
  st->active = n > 0 ? 1 : 0;
}

__device__ void centered(const Xf& t12, const Xf& shift, Xf& out) {
  double par[12];
  get_params(t12, par);
  out = identity(t12.kind);
  set_params(out, par);
  compose(out, shift, false);
}

// The resampling maps of the outer transforms T (classic) or t12F / t12B
// (symmetric), given in registers (the same in every lane). Lane l of the
// first wave computes one map, on the fixed (l < 2) or moving (l >= 2) side,
// at the level's grid (l even) or the full-resolution one (l odd): the four
// maps of the symmetric method take the time of one.
__device__ void prepare_maps_of(OuterState* o, int symmetric, const Xf& T, const Xf& t12F,
                                const Xf& t12B) {
  const int l = threadIdx.x;
  if (l >= (symmetric ? 4 : 2)) return;
  const bool side = (l >> 1) != 0, full = (l & 1) != 0;
  const Geom gin = side ? (full ? o->gMfull : o->gM) : (full ? o->gFfull : o->gF);
  const Geom gout = side ? o->gHM : o->gHF;
  Xf c = T;
  if (symmetric) centered(side ? t12B : t12F, side ? o->movT : o->fixT, c);
  const IndexMap map = make_index_map(c.M, c.off, gin, gout);
  IndexMap* dst = side ? (full ? &o->mmapM : &o->mapM) : (full ? &o->mmapF : &o->mapF);
  *dst = map;
}

__device__ void prepare_maps(OuterState* o, int symmetric) {
  const Xf T = o->T, t12F = o->t12F, t12B = o->t12B;
  prepare_maps_of(o, symmetric, T, t12F, t12B);
}

__device__ void finish_update(LtsState* st, OuterState* o);

// Run by the lanes of one wave (at least 4): each computes the update, lane
// 0 stores the state and the flags, lanes 0..3 the maps.
__device__ __forceinline__ void finish_body(LtsState* st, OuterState* o, int* flags) {
  if (!st->active && !st->finished) {
    if (threadIdx.x == 0) st->finished = 1;
    finish_update(st, o);
  }
  if (threadIdx.x != 0) return;
  flags[0] = st->active;
  flags[1] = o->stop;
  flags[2] = st->iter + 1;  // inner iterations run so far
}

__device__ void finish_update(LtsState* st, OuterState* o) {
  // The transforms are composed in registers and stored once (see
  // prepare_maps_of).
  const int kind = st->kind, symmetric = st->symmetric;
  Xf T = o->T, t12F = o->t12F, t12B = o->t12B;
  if (symmetric) {
    const Xf F = st->F, B = st->B;
    compose(t12F, F, true);
    compose(t12B, B, true);
    const Xf ib = inverse(t12B);
    T = identity(kind);
    compose(T, ib, true);
    compose(T, t12F, false);
  } else {
    const Xf F = st->F;
    compose(T, F, true);
  }
  double cur[12], lastp[12];
  get_params(T, cur);
  const int np = kind == kRigid ? 6 : 12;
  for (int i = 0; i < np; ++i) lastp[i] = o->lastp[i];
  const double sq = delta_of(lastp, cur, np);
  const int inner = o->inner_total + st->iter + 1;
  const int stop = sq < o->stop_tol ? 1 : 0;
  if (threadIdx.x == 0) {
    if (symmetric) {
      o->t12F = t12F;
      o->t12B = t12B;
    }
    o->T = T;
    o->sqdiff = sq;
    for (int i = 0; i < np; ++i) o->lastp[i] = cur[i];
    o->stop = stop;
    o->inner_total = inner;
  }
  prepare_maps_of(o, symmetric, T, t12F, t12B);
}

__global__ __launch_bounds__(64) void prepare_maps_kernel(OuterState* o, int symmetric) { prepare_maps(o, symmetric); }

__global__ void reset_count_kernel(LtsState* st) { st->n_kept = 0ull; }

// One pair per launch.
__global__ __launch_bounds__(kT) void residual_kernel(const double* __restrict__ X,
                                                      const double* __restrict__ Y,
                                                      const uint8_t* __restrict__ kept,
                                                      int64_t n, LtsState* st,
                                                      double* __restrict__ r,
                                                      unsigned int* hist1, int do_hist) {
  residual_body(X, Y, kept, n, st, r, hist1, do_hist);
}
__global__ __launch_bounds__(kT) void hist2_kernel(const double* __restrict__ r, int64_t n,
                                                   LtsState* st, const unsigned int* hist1,
                                                   unsigned int* hist2) {
  hist2_body(r, n, st, hist1, hist2);
}
__global__ __launch_bounds__(kT) void compact_kernel(const double* __restrict__ r, int64_t n,
                                                     LtsState* st, const unsigned int* hist2,
                                                     unsigned long long* __restrict__ cand) {
  compact_body(r, n, st, hist2, cand);
}
template <int PER>
__global__ __launch_bounds__(kT) void moments_kernel(const double* __restrict__ X,
                                                     const double* __restrict__ Y,
                                                     const uint8_t* __restrict__ kept,
                                                     const double* __restrict__ r, int64_t n,
                                                     LtsState* st, double* part, int first) {
  moments_body<PER>(X, Y, kept, r, n, st, part, first);
}
__global__ __launch_bounds__(kT) void solve_kernel(LtsState* st, const double* part, int nparts,
                                                   unsigned int* hist1, unsigned int* hist2) {
  solve_body(st, part, nparts, hist1, hist2);
}
__global__ void init_kernel(LtsState* st, const OuterState* outer, int kind, int symmetric,
                            double tol, unsigned int* hist1, unsigned int* hist2) {
  init_body(st, outer, kind, symmetric, tol, hist1, hist2);
}
__global__ __launch_bounds__(64) void finish_kernel(LtsState* st, OuterState* o, int* flags) {
  finish_body(st, o, flags);
}

// Several pairs per launch: pair blockIdx.y, with the grid in x sized for
// the largest pair (a pair's surplus workgroups find no points or parts).
// Every result is that of the one-pair launches.
__global__ __launch_bounds__(kT) void residual_multi(const LtsItem* __restrict__ items,
                                                     int do_hist) {
  const LtsItem it = items[blockIdx.y];
  residual_body(it.X, it.Y, it.kept, it.npts, it.st, it.r, it.hist1, do_hist);
}
__global__ __launch_bounds__(kT) void hist2_multi(const LtsItem* __restrict__ items) {
  const LtsItem it = items[blockIdx.y];
  hist2_body(it.r, it.npts, it.st, it.hist1, it.hist2);
}
__global__ __launch_bounds__(kT) void compact_multi(const LtsItem* __restrict__ items) {
  const LtsItem it = items[blockIdx.y];
  compact_body(it.r, it.npts, it.st, it.hist2, it.cand);
}
__global__ __launch_bounds__(kT) void final_select_multi(const LtsItem* __restrict__ items) {
  const LtsItem it = items[blockIdx.y];
  if (!it.st->active) return;
  final_select(it.cand, it.st);
}
// PER 8 takes the pairs of more than kBigRun points, PER 4 the others.
template <int PER>
__global__ __launch_bounds__(kT) void moments_multi(const LtsItem* __restrict__ items,
                                                    int first) {
  const LtsItem it = items[blockIdx.y];
  if ((part_points(it.npts) == 2 * kPart) != (PER == 8)) return;
  if ((int64_t)blockIdx.x >= lts_parts_dev(it.npts)) return;
  moments_body<PER>(it.X, it.Y, it.kept, it.r, it.npts, it.st, it.part, first);
}
__global__ __launch_bounds__(kT) void solve_multi(const LtsItem* __restrict__ items) {
  const LtsItem it = items[blockIdx.y];
  solve_body(it.st, it.part, (int)lts_parts_dev(it.npts), it.hist1, it.hist2);
}
__global__ void init_multi(const LtsItem* __restrict__ items, int kind, int symmetric,
                           double tol) {
  const LtsItem it = items[blockIdx.y];
  init_body(it.st, it.outer, kind, symmetric, tol, it.hist1, it.hist2);
}
__global__ __launch_bounds__(64) void finish_multi(const LtsItem* __restrict__ items) {
  const LtsItem it = items[blockIdx.y];
  finish_body(it.st, it.outer, it.flags);
}
__global__ void reset_count_multi(const LtsItem* __restrict__ items, int n) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) items[i].st->n_kept = 0ull;
}

}  // namespace

int64_t lts_parts(int64_t npts) {
  const int p = part_points(npts);
  return (npts + p - 1) / p;
}

void lts_enqueue_iteration(const double* X, const double* Y, const uint8_t* kept, int64_t npts,
                           LtsState* st, OuterState* outer, LtsBuffers& b, int iteration,
                           hipStream_t s) {
  (void)outer;
  const int nparts = (int)lts_parts(npts);
  if (nparts > b.nparts_cap) throw Error(kErrState, "LTS partial buffer too small");
  const int grid = (int)std::min<int64_t>((npts + kT - 1) / kT, 2048);
  const bool select = iteration > 0;  // the first pass keeps every pair
  // With the histogram, at most 256 workgroups: the residuals crowd into a
  // few bins, and each workgroup merges its bins with global atomics.
  const int g2 = std::min(grid, 256);
  if (select) {
    hipLaunchKernelGGL(residual_kernel, dim3(g2), dim3(kT), 0, s, X, Y, kept, npts, st, b.r,
                       b.hist1, 1);
    // At most 256 workgroups: each re-reads a 16 KB histogram to find its bin.
    hipLaunchKernelGGL(hist2_kernel, dim3(g2), dim3(kT), 0, s, b.r, npts, st, b.hist1, b.hist2);
    hipLaunchKernelGGL(compact_kernel, dim3(g2), dim3(kT), 0, s, b.r, npts, st, b.hist2, b.cand);
    hipLaunchKernelGGL(final_select_kernel, dim3(1), dim3(kT), 0, s, b.cand, st);
  }
  if (part_points(npts) == 2 * kPart)
    hipLaunchKernelGGL(moments_kernel<8>, dim3(nparts), dim3(kT), 0, s, X, Y, kept, b.r, npts, st,
                       b.part, select ? 0 : 1);
  else
    hipLaunchKernelGGL(moments_kernel<4>, dim3(nparts), dim3(kT), 0, s, X, Y, kept, b.r, npts, st,
                       b.part, select ? 0 : 1);
  hipLaunchKernelGGL(solve_kernel, dim3(1), dim3(kT), 0, s, st, b.part, nparts, b.hist1,
                     b.hist2);
  MIRORR_CHECK_LAUNCH();
}

void lts_start(LtsState* st, const OuterState* outer, int kind, int symmetric, double tol,
               LtsBuffers& b, hipStream_t s) {
  hipLaunchKernelGGL(init_kernel, dim3(1), dim3(kT), 0, s, st, outer, kind, symmetric, tol,
                     b.hist1, b.hist2);
  MIRORR_CHECK_LAUNCH();
}

void lts_finish(LtsState* st, OuterState* outer, int* flags, hipStream_t s) {
  if (!st || !outer || !flags) throw Error(kErrState, "lts_finish: null device buffer");
  hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(4), 0, s, st, outer, flags);
  MIRORR_CHECK_LAUNCH();
}

void outer_prepare_maps(OuterState* outer, int symmetric, hipStream_t s) {
  hipLaunchKernelGGL(prepare_maps_kernel, dim3(1), dim3(4), 0, s, outer, symmetric);
  MIRORR_CHECK_LAUNCH();
}

void lts_reset_count(LtsState* st, hipStream_t s) {
  hipLaunchKernelGGL(reset_count_kernel, dim3(1), dim3(1), 0, s, st);
  MIRORR_CHECK_LAUNCH();
}

void lts_start_multi(const LtsItem* items, int n, int kind, int symmetric, double tol,
                     hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(init_multi, dim3(1, (unsigned)n), dim3(kT), 0, s, items, kind, symmetric,
                     tol);
  MIRORR_CHECK_LAUNCH();
}

void lts_enqueue_iteration_multi(const LtsItem* items, int n, int64_t max_npts,
                                 int64_t min_npts, int iteration, hipStream_t s) {
  if (n <= 0) return;
  const int64_t npts = std::max<int64_t>(max_npts, 1);
  const int nparts = (int)lts_parts(npts);
  const int grid = (int)std::min<int64_t>((npts + kT - 1) / kT, 2048);
  const bool select = iteration > 0;
  const int g2 = std::min(grid, 256);
  const unsigned ny = (unsigned)n;
  if (select) {
    hipLaunchKernelGGL(residual_multi, dim3((unsigned)g2, ny), dim3(kT), 0, s, items, 1);
    hipLaunchKernelGGL(hist2_multi, dim3((unsigned)g2, ny), dim3(kT), 0, s, items);
    hipLaunchKernelGGL(compact_multi, dim3((unsigned)g2, ny), dim3(kT), 0, s, items);
    hipLaunchKernelGGL(final_select_multi, dim3(1, ny), dim3(kT), 0, s, items);
  }
  // nparts: the most partials of any pair (part_points grows with npts, so
  // the largest pair of each kind has the most of its kind)
  if (part_points(npts) == 2 * kPart)
    hipLaunchKernelGGL(moments_multi<8>, dim3((unsigned)nparts, ny), dim3(kT), 0, s, items,
                       select ? 0 : 1);
  if (part_points(std::max<int64_t>(min_npts, 1)) == kPart)
    hipLaunchKernelGGL(moments_multi<4>,
                       dim3((unsigned)lts_parts(std::min<int64_t>(npts, kBigRun)), ny), dim3(kT),
                       0, s, items, select ? 0 : 1);
  hipLaunchKernelGGL(solve_multi, dim3(1, ny), dim3(kT), 0, s, items);
  MIRORR_CHECK_LAUNCH();
}

void lts_finish_multi(const LtsItem* items, int n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(finish_multi, dim3(1, (unsigned)n), dim3(4), 0, s, items);
  MIRORR_CHECK_LAUNCH();
}

void lts_reset_count_multi(const LtsItem* items, int n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(reset_count_multi, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, items, n);
  MIRORR_CHECK_LAUNCH();
}

}  // namespace mirorr
