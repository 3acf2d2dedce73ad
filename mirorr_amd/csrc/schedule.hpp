// Pyramid schedule (host): PyramidScheduleTuner::UpdateSymmetric and the
// level cropping of the reference (itkPyramidScheduleTuner.txx:23-126,299-414,
// itkPyramidScheduleTuner.h:148-161).
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

namespace mirorr {

struct Schedule {
  std::vector<int> moving, fixed;  // rows x 3, coarsest first
  int rows = 0;
};

inline double sched_octaves(double dimension, double base) {
  return std::floor(std::log(double(dimension / base)) / std::log(2.0));
}
inline unsigned sched_levels(double dimension, double base) {
  return 1u + (unsigned)std::max(0.0, sched_octaves(dimension, base));
}

// Integer shrink factors per level and axis so that the coarsest level is
// about min_length voxels long, with a first crude alignment of unequal
// resolutions; levels that barely differ are dropped.
inline Schedule schedule_symmetric(const int msize[3], const double msp[3], const int fsize[3],
                                   const double fsp[3], unsigned min_length,
                                   unsigned max_levels) {
  const unsigned lm = (unsigned)*std::max_element(msize, msize + 3);
  const unsigned lf = (unsigned)*std::max_element(fsize, fsize + 3);
  const unsigned P = 1 + std::max(sched_levels(lm, min_length), sched_levels(lf, min_length));
  std::vector<unsigned> M(3 * P, 0u), F(3 * P, 0u);
  auto m = [&](int l, int d) -> unsigned& { return M[3 * l + d]; };
  auto f = [&](int l, int d) -> unsigned& { return F[3 * l + d]; };
  for (int d = 0; d < 3; ++d) {
    m(P - 1, d) = f(P - 1, d) = 1u;
    const unsigned em = 1u << (sched_levels(msize[d], min_length) - 1u);
    const unsigned ef = 1u << (sched_levels(fsize[d], min_length) - 1u);
    unsigned r = 2;
    if (msp[d] != fsp[d]) {
      const bool moving_finer = msp[d] < fsp[d];
      const double big = moving_finer ? fsp[d] : msp[d], small = moving_finer ? msp[d] : fsp[d];
      const unsigned best = 1u << (unsigned)std::max(0.0, sched_octaves(big * 4.0 / 3.0, small));
      const unsigned rate = std::min(best, moving_finer ? em : ef);
      for (unsigned s = rate >= 2u ? 2u : 1u; s <= rate; s *= 2, ++r) {
        m(P - r, d) = moving_finer ? s : 1u;
        f(P - r, d) = moving_finer ? 1u : s;
      }
    } else {
      m(P - r, d) = f(P - r, d) = 1u;
      ++r;
    }
    for (int l = (int)P - (int)r; l >= 0; --l) {
      const bool grow = !(2 * m(l + 1, d) > em || 2 * f(l + 1, d) > ef);
      m(l, d) = grow ? 2 * m(l + 1, d) : m(l + 1, d);
      f(l, d) = grow ? 2 * f(l + 1, d) : f(l + 1, d);
    }
  }
  auto ndiff = [&](int a, int b) {
    int n = 0;
    for (int d = 0; d < 3; ++d) n += (m(a, d) != m(b, d)) + (f(a, d) != f(b, d));
    return n;
  };
  const unsigned first = ndiff(P - 1, P - 2) == 0 ? 1u : 0u;
  unsigned last = P;
  for (unsigned l = 2; l < P; ++l)
    if (ndiff(P - l - 1, P - l) < 3) {
      last = l;
      break;
    }
  unsigned n = std::min(last - first, max_levels);
  Schedule s;
  s.rows = (int)n;
  s.moving.resize(3 * n);
  s.fixed.resize(3 * n);
  for (unsigned l = 0; l < n; ++l)
    for (int d = 0; d < 3; ++d) {
      s.moving[3 * (n - l - 1) + d] = (int)m(P - first - l - 1, d);
      s.fixed[3 * (n - l - 1) + d] = (int)f(P - first - l - 1, d);
    }
  return s;
}

// GetCroppedLevel: 0 -> default, negative counts back from the top.
inline int cropped_level(int level, int default_level, int total) {
  int f = level;
  if (f < -total || f > total) f = 0;
  if (f == 0) return default_level;
  return f < 0 ? total + 1 + f : f;
}

}  // namespace mirorr
