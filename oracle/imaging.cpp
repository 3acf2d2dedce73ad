// Oracle: resampling, Gaussian pyramid and pyramid schedule
// (test infrastructure only; see oracle.h).
//
// ITK is not available here, so its filters are restated from their
// documented algorithms (parity with a real ITK build is unpinned):
//  * LinearInterpolateImageFunction + ResampleImageFilter: a point maps to a
//    continuous index; inside [start-0.5, end+0.5] per axis it is linearly
//    interpolated with neighbours clamped to the buffer (edge replication),
//    else the default value 0 is written; uint8 outputs are truncated.
//  * RecursiveMultiResolutionPyramidImageFilter (MirorrPyramidImplement.h:94,
//    .txx:575-609): level l from level l+1 by DiscreteGaussianImageFilter
//    (variance (f/2)^2 in voxels, f = ratio of shrink factors, maximum error
//    0.1, kernel from modified Bessel functions, zero-flux Neumann edges)
//    then a linear resample onto the shrunk grid whose origin moves by
//    D (s_out - s_in) / 2.
//  * PyramidScheduleTuner::UpdateSymmetric (itkPyramidScheduleTuner.txx:299-414)
//    and the level cropping (itkPyramidScheduleTuner.h:148-161, .txx:23-126).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "oracle.h"
#include "xform.hpp"

namespace {

inline int64_t lin(const int d[3], int x, int y, int z) {
  return (int64_t)x + (int64_t)d[0] * ((int64_t)y + (int64_t)d[1] * z);
}

// Index -> physical of a grid (origin + D diag(s) idx) and its inverse.
struct Grid {
  int dims[3];
  double s[3], o[3], D[9];
  void to_phys(const double idx[3], double p[3]) const {
    for (int r = 0; r < 3; ++r) {
      double v = o[r];
      for (int c = 0; c < 3; ++c) v += D[3 * r + c] * s[c] * idx[c];
      p[r] = v;
    }
  }
  void to_index(const double p[3], double idx[3]) const {
    // (D diag(s))^-1 (p - o) = diag(1/s) D^-1 (p - o)
    orc::M3 m;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) m.a[r][c] = D[3 * r + c];
    const orc::M3 inv = orc::m3_inverse(m);
    double d[3] = {p[0] - o[0], p[1] - o[1], p[2] - o[2]};
    for (int r = 0; r < 3; ++r)
      idx[r] = (inv.a[r][0] * d[0] + inv.a[r][1] * d[1] + inv.a[r][2] * d[2]) / s[r];
  }
};

Grid grid_of(const orc_geom* g) {
  Grid r;
  for (int i = 0; i < 3; ++i) {
    r.dims[i] = g->dims[i];
    r.s[i] = g->spacing[i];
    r.o[i] = g->origin[i];
  }
  for (int i = 0; i < 9; ++i) r.D[i] = g->direction[i];
  return r;
}

// Linear interpolation with clamped neighbours; `in` false outside the buffer.
template <class T>
double interp(const T* img, const int d[3], const double ci[3], bool* in) {
  for (int a = 0; a < 3; ++a)
    if (!(ci[a] >= -0.5 && ci[a] < d[a] - 0.5)) {
      *in = false;
      return 0.0;
    }
  *in = true;
  int b[3];
  double f[3];
  for (int a = 0; a < 3; ++a) {
    int fl = (int)std::floor(ci[a]);
    double fr = ci[a] - fl;
    if (fl < 0) {  // basei clamped to the start; distance <= 0 -> no blend
      fl = 0;
      fr = 0.0;
    }
    b[a] = fl;
    f[a] = fr;
  }
  auto at = [&](int dx, int dy, int dz) -> double {
    const int x = std::min(b[0] + dx, d[0] - 1);
    const int y = std::min(b[1] + dy, d[1] - 1);
    const int z = std::min(b[2] + dz, d[2] - 1);
    return (double)img[lin(d, x, y, z)];
  };
  const double v000 = at(0, 0, 0), v100 = at(1, 0, 0), v010 = at(0, 1, 0), v110 = at(1, 1, 0);
  const double v001 = at(0, 0, 1), v101 = at(1, 0, 1), v011 = at(0, 1, 1), v111 = at(1, 1, 1);
  const double x00 = v000 + (v100 - v000) * f[0];
  const double x10 = v010 + (v110 - v010) * f[0];
  const double x01 = v001 + (v101 - v001) * f[0];
  const double x11 = v011 + (v111 - v011) * f[0];
  const double xx0 = x00 + (x10 - x00) * f[1];
  const double xx1 = x01 + (x11 - x01) * f[1];
  return xx0 + (xx1 - xx0) * f[2];
}

}  // namespace

extern "C" {

// Resample `in` (float) and optionally its uint8 mask onto the output grid
// under y = M x + off (the transform maps output physical points into the
// input's physical space). Outputs with NULL pointers are skipped.
void orc_resample(const float* in, const uint8_t* in_mask, const orc_geom* gin,
                  const double M[9], const double off[3], const orc_geom* gout,
                  float* out, uint8_t* out_mask) {
  const Grid gi = grid_of(gin), go = grid_of(gout);
  for (int z = 0; z < go.dims[2]; ++z)
    for (int y = 0; y < go.dims[1]; ++y)
      for (int x = 0; x < go.dims[0]; ++x) {
        const double idx[3] = {(double)x, (double)y, (double)z};
        double p[3], q[3], ci[3];
        go.to_phys(idx, p);
        for (int r = 0; r < 3; ++r) q[r] = M[3 * r] * p[0] + M[3 * r + 1] * p[1] + M[3 * r + 2] * p[2] + off[r];
        gi.to_index(q, ci);
        const int64_t o = lin(go.dims, x, y, z);
        bool inside;
        if (out) {
          const double v = interp(in, gi.dims, ci, &inside);
          out[o] = inside ? (float)v : 0.0f;
        }
        if (out_mask) {
          if (in_mask) {
            const double v = interp(in_mask, gi.dims, ci, &inside);
            out_mask[o] = inside ? (uint8_t)std::min(255.0, std::max(0.0, v)) : 0;
          } else {
            // A default mask is all 255; interpolating it gives 255 inside.
            bool in2;
            (void)interp(in, gi.dims, ci, &in2);
            out_mask[o] = in2 ? 255 : 0;
          }
        }
      }
}

// Modified Bessel functions as ITK's GaussianOperator evaluates them:
// Abramowitz & Stegun 9.8.1-9.8.4 polynomial approximations for I0 and I1
// and Miller's downward recurrence (normalised by I0) for I_n, n >= 2.
static double bessel_i0(double y) {
  const double d = std::fabs(y);
  if (d < 3.75) {
    double m = y / 3.75;
    m *= m;
    return 1.0 + m * (3.5156229 + m * (3.0899424 + m * (1.2067492 + m * (0.2659732 + m * (0.360768e-1 + m * 0.45813e-2)))));
  }
  const double m = 3.75 / d;
  return (std::exp(d) / std::sqrt(d)) *
         (0.39894228 + m * (0.1328592e-1 + m * (0.225319e-2 + m * (-0.157565e-2 + m * (0.916281e-2 + m * (-0.2057706e-1 + m * (0.2635537e-1 + m * (-0.1647633e-1 + m * 0.392377e-2))))))));
}
static double bessel_i1(double y) {
  const double d = std::fabs(y);
  double r;
  if (d < 3.75) {
    double m = y / 3.75;
    m *= m;
    r = d * (0.5 + m * (0.87890594 + m * (0.51498869 + m * (0.15084934 + m * (0.2658733e-1 + m * (0.301532e-2 + m * 0.32411e-3))))));
  } else {
    const double m = 3.75 / d;
    r = 0.2282967e-1 + m * (-0.2895312e-1 + m * (0.1787654e-1 - m * 0.420059e-2));
    r = 0.39894228 + m * (-0.3988024e-1 + m * (-0.362018e-2 + m * (0.163801e-2 + m * (-0.1031555e-1 + m * r))));
    r *= std::exp(d) / std::sqrt(d);
  }
  return y < 0.0 ? -r : r;
}
static double bessel_in(int n, double y) {
  if (y == 0.0) return 0.0;
  const double acc = 40.0, big = 1e10, small_ = 1e-10;
  const double tox = 2.0 / std::fabs(y);
  double bip = 0.0, ans = 0.0, bi = 1.0;
  for (int j = 2 * (n + (int)std::sqrt(acc * n)); j > 0; --j) {
    const double bim = bip + j * tox * bi;
    bip = bi;
    bi = bim;
    if (std::fabs(bi) > big) {
      ans *= small_;
      bi *= small_;
      bip *= small_;
    }
    if (j == n) ans = bip;
  }
  ans *= bessel_i0(y) / bi;
  return (y < 0.0 && (n & 1)) ? -ans : ans;
}

// GaussianOperator coefficients (centre first, one side), normalised to
// unit sum over the symmetric kernel. Returns the radius.
int orc_gaussian_kernel(double variance, double max_error, int max_width, double* out,
                        int cap) {
  std::vector<double> c;
  const double et = std::exp(-variance);
  const double cap_sum = 1.0 - max_error;
  double sum = 0.0;
  c.push_back(et * bessel_i0(variance));
  sum += c[0];
  c.push_back(et * bessel_i1(variance));
  sum += 2.0 * c[1];
  for (int i = 2; sum < cap_sum; ++i) {
    c.push_back(et * bessel_in(i, variance));
    sum += 2.0 * c[i];
    if (c[i] < sum * 2.220446049250313e-16) break;
    if ((int)c.size() > max_width) break;
  }
  for (double& v : c) v /= sum;
  const int n = (int)c.size();
  for (int i = 0; i < n && i < cap; ++i) out[i] = c[i];
  return n - 1;
}

// Separable discrete Gaussian (per-axis variance in voxels, 0 = none),
// zero-flux Neumann boundary, double intermediates, float output. Axes are
// filtered z, then y, then x.
void orc_gaussian_smooth(const float* in, const int dims[3], const double var[3],
                         double max_error, float* out) {
  const int64_t n = (int64_t)dims[0] * dims[1] * dims[2];
  std::vector<double> a(in, in + n), b(n);
  for (int ax = 2; ax >= 0; --ax) {
    double k[64];
    const int r = orc_gaussian_kernel(var[ax], max_error, 32, k, 64);
    for (int z = 0; z < dims[2]; ++z)
      for (int y = 0; y < dims[1]; ++y)
        for (int x = 0; x < dims[0]; ++x) {
          const int p[3] = {x, y, z};
          double s = 0.0;
          for (int j = -r; j <= r; ++j) {
            int q[3] = {p[0], p[1], p[2]};
            q[ax] = std::min(std::max(p[ax] + j, 0), dims[ax] - 1);
            s += k[j < 0 ? -j : j] * a[lin(dims, q[0], q[1], q[2])];
          }
          b[lin(dims, x, y, z)] = s;
        }
    std::swap(a, b);
  }
  for (int64_t i = 0; i < n; ++i) out[i] = (float)a[i];
}

// Output geometry of a pyramid level for a total shrink factor per axis
// (MultiResolutionPyramidImageFilter::GenerateOutputInformation).
void orc_pyramid_level_geom(const orc_geom* in, const int factors[3], orc_geom* out) {
  *out = *in;
  double ds[3];
  for (int a = 0; a < 3; ++a) {
    out->spacing[a] = in->spacing[a] * factors[a];
    out->dims[a] = std::max(1, (int)std::floor((double)in->dims[a] / factors[a]));
    ds[a] = out->spacing[a] - in->spacing[a];
  }
  for (int r = 0; r < 3; ++r) {
    double v = 0.0;
    for (int c = 0; c < 3; ++c) v += in->direction[3 * r + c] * ds[c];
    out->origin[r] = in->origin[r] + 0.5 * v;
  }
}

// One level of the recursive pyramid: smooth the finer level `prev` (with
// factor ratio f) and resample it onto `out_geom`.
void orc_pyramid_step(const float* prev, const orc_geom* prev_geom, const int ratio[3],
                      const orc_geom* out_geom, float* out) {
  const int64_t n = (int64_t)prev_geom->dims[0] * prev_geom->dims[1] * prev_geom->dims[2];
  std::vector<float> sm(n);
  double var[3];
  bool all_ones = true;
  for (int a = 0; a < 3; ++a) {
    var[a] = ratio[a] != 1 ? (0.5 * ratio[a]) * (0.5 * ratio[a]) : 0.0;
    if (ratio[a] != 1) all_ones = false;
  }
  if (all_ones) {  // the filter copies the data (same grid)
    std::memcpy(out, prev, n * sizeof(float));
    return;
  }
  orc_gaussian_smooth(prev, prev_geom->dims, var, 0.1, sm.data());
  const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  const double z[3] = {0, 0, 0};
  orc_resample(sm.data(), nullptr, prev_geom, I, z, out_geom, out, nullptr);
}

// ---- PyramidScheduleTuner ---------------------------------------------------

static double n_octaves(double dimension, double base) {
  return std::floor(std::log(double(dimension / base)) / std::log(2.0));
}
static unsigned n_levels(double dimension, double base) {
  return 1u + (unsigned)std::max(0.0, n_octaves(dimension, base));
}
static unsigned best_rate(double big, double small_) {
  const double MAGIC = 4. / 3.;
  const unsigned e = (unsigned)std::max(0.0, n_octaves(big * MAGIC, small_));
  return 1u << e;
}

// UpdateSymmetric: writes the moving and fixed schedules (rows x 3, coarsest
// first) and returns the number of rows (<= cap).
int orc_schedule_symmetric(const int msize[3], const double mspacing[3], const int fsize[3],
                           const double fspacing[3], unsigned min_length, unsigned max_levels,
                           int* msched, int* fsched, int cap) {
  const unsigned lm = *std::max_element(msize, msize + 3);
  const unsigned lf = *std::max_element(fsize, fsize + 3);
  const unsigned maxp = 1 + std::max(n_levels(lm, min_length), n_levels(lf, min_length));
  std::vector<unsigned> SM(maxp * 3, (unsigned)-1), SF(maxp * 3, (unsigned)-1);
  auto sm = [&](int l, int d) -> unsigned& { return SM[l * 3 + d]; };
  auto sf = [&](int l, int d) -> unsigned& { return SF[l * 3 + d]; };
  for (int d = 0; d < 3; ++d) {
    unsigned r = 1;
    sm(maxp - r, d) = 1u;
    sf(maxp - r, d) = 1u;
    const double ms = mspacing[d], fs = fspacing[d];
    const unsigned maxem = 1u << (n_levels(msize[d], min_length) - 1u);
    const unsigned maxef = 1u << (n_levels(fsize[d], min_length) - 1u);
    r = 2;
    if (ms < fs) {
      const unsigned rate = std::min(best_rate(fs, ms), maxem);
      for (unsigned s = rate >= 2u ? 2u : 1u; s <= rate; s *= 2) {
        sm(maxp - r, d) = s;
        sf(maxp - r, d) = 1u;
        ++r;
      }
    } else if (ms > fs) {
      const unsigned rate = std::min(best_rate(ms, fs), maxef);
      for (unsigned s = rate >= 2u ? 2u : 1u; s <= rate; s *= 2) {
        sm(maxp - r, d) = 1u;
        sf(maxp - r, d) = s;
        ++r;
      }
    } else {
      sm(maxp - r, d) = 1u;
      sf(maxp - r, d) = 1u;
      ++r;
    }
    for (int l = (int)maxp - (int)r; l >= 0; --l) {
      if (2 * sm(l + 1, d) > maxem || 2 * sf(l + 1, d) > maxef) {
        sm(l, d) = sm(l + 1, d);
        sf(l, d) = sf(l + 1, d);
      } else {
        sm(l, d) = 2 * sm(l + 1, d);
        sf(l, d) = 2 * sf(l + 1, d);
      }
    }
  }
  unsigned first = 0;
  {
    int nd = 0;
    for (int d = 0; d < 3; ++d) {
      if (sm(maxp - 1, d) != sm(maxp - 2, d)) ++nd;
      if (sf(maxp - 1, d) != sf(maxp - 2, d)) ++nd;
    }
    if (nd == 0) first = 1;
  }
  unsigned last = maxp;
  for (unsigned l = 2; l < maxp; ++l) {
    int nd = 0;
    for (int d = 0; d < 3; ++d) {
      if (sm(maxp - l - 1, d) != sm(maxp - l, d)) ++nd;
      if (sf(maxp - l - 1, d) != sf(maxp - l, d)) ++nd;
    }
    if (nd < 3) {
      last = l;
      break;
    }
  }
  unsigned n = last - first;
  if (n > max_levels) n = max_levels;
  if ((int)n > cap) return -1;
  for (unsigned l = 0; l < n; ++l)
    for (int d = 0; d < 3; ++d) {
      msched[3 * (n - l - 1) + d] = (int)sm(maxp - first - l - 1, d);
      fsched[3 * (n - l - 1) + d] = (int)sf(maxp - first - l - 1, d);
    }
  return (int)n;
}

// GetCroppedLevel (itkPyramidScheduleTuner.txx:23-52).
int orc_cropped_level(int level, int default_level, int total_levels) {
  int f = level;
  if (f < -total_levels) f = 0;
  if (f > total_levels) f = 0;
  if (f == 0) return default_level;
  if (f < 0) return total_levels + 1 + f;
  return f;
}

// GetCroppedMinMax (itkPyramidScheduleTuner.txx:54-126).
void orc_cropped_min_max(int level_min, int level_max, int total, int* min_out, int* max_out) {
  int fmin = level_min, fmax = level_max;
  if (fmin < -total || fmin > total) fmin = 0;
  int mn = fmin == 0 ? 1 : (fmin < 0 ? total + 1 + fmin : fmin);
  if (fmax < -total || fmax > total) fmax = 0;
  int mx = fmax == 0 ? total : (fmax < 0 ? total + 1 + fmax : fmax);
  if (mn > mx) mx = mn;
  *min_out = mn;
  *max_out = mx;
}

}  // extern "C"
