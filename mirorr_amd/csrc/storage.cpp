// Host-side encoding of float images into binary16 storage.
#include "storage.hpp"

#include <immintrin.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "../../include/mirorr_gpu.h"

namespace mirorr {

namespace {

// Minimum and maximum skipping NaNs (vminps/vmaxps return the second operand
// when the first is a NaN).
__attribute__((target("avx"))) void minmax_avx(const float* d, size_t n, float* lo, float* hi) {
  __m256 a = _mm256_set1_ps(INFINITY), b = _mm256_set1_ps(-INFINITY);
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    const __m256 v = _mm256_loadu_ps(d + i);
    a = _mm256_min_ps(v, a);
    b = _mm256_max_ps(v, b);
  }
  float la[8], hb[8];
  _mm256_storeu_ps(la, a);
  _mm256_storeu_ps(hb, b);
  float l = INFINITY, h = -INFINITY;
  for (int k = 0; k < 8; ++k) {
    l = std::min(l, la[k]);
    h = std::max(h, hb[k]);
  }
  for (; i < n; ++i) {
    l = fminf(l, d[i]);
    h = fmaxf(h, d[i]);
  }
  *lo = l;
  *hi = h;
}

}  // namespace

HalfEncoding half_encoding(const float* data, size_t n) {
  // NaNs are skipped; infinities take the slow pass.
  float lo = INFINITY, hi = -INFINITY;
  if (__builtin_cpu_supports("avx")) {
    minmax_avx(data, n, &lo, &hi);
  } else {
    for (size_t i = 0; i < n; ++i) {
      lo = fminf(lo, data[i]);
      hi = fmaxf(hi, data[i]);
    }
  }
  if (std::isinf(lo) || std::isinf(hi)) {
    lo = INFINITY;
    hi = -INFINITY;
    for (size_t i = 0; i < n; ++i) {
      const float v = data[i];
      if (!std::isfinite(v)) continue;
      lo = std::min(lo, v);
      hi = std::max(hi, v);
    }
  }
  HalfEncoding e;
  if (!(lo <= hi)) return e;  // no finite value
  e.offset = std::nearbyint(0.5 * ((double)lo + (double)hi));
  if (std::fabs(e.offset) >= 16777216.0) e.offset = 0.0;  // keep it exact in float
  const double h = std::max({(double)hi - e.offset, e.offset - (double)lo, std::fabs(e.offset)});
  int k = 0;
  if (h > 0.0) k = (int)std::floor(std::log2(2048.0 / h));
  k = std::min(std::max(k, -40), 40);
  e.scale = std::ldexp(1.0, k);
  return e;
}

namespace {

// (v - offset) * scale in float (offset exact in float, scale a power of
// two), rounded to binary16 to nearest even.
void encode_scalar(const float* data, size_t n, float off, float sc, half_t* out) {
  for (size_t i = 0; i < n; ++i) out[i] = (half_t)((data[i] - off) * sc);
}

__attribute__((target("avx,f16c"))) void encode_f16c(const float* data, size_t n, float off,
                                                      float sc, half_t* out) {
  const __m256 vo = _mm256_set1_ps(off), vs = _mm256_set1_ps(sc);
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    const __m256 v = _mm256_mul_ps(_mm256_sub_ps(_mm256_loadu_ps(data + i), vo), vs);
    _mm_storeu_si128(reinterpret_cast<__m128i*>(out + i),
                     _mm256_cvtps_ph(v, _MM_FROUND_TO_NEAREST_INT | _MM_FROUND_NO_EXC));
  }
  encode_scalar(data + i, n - i, off, sc, out + i);
}

}  // namespace

void encode_half(const float* data, size_t n, const HalfEncoding& e, half_t* out) {
  const float off = (float)e.offset, sc = (float)e.scale;
  if (__builtin_cpu_supports("f16c") && __builtin_cpu_supports("avx"))
    encode_f16c(data, n, off, sc, out);
  else
    encode_scalar(data, n, off, sc, out);
}

}  // namespace mirorr

extern "C" mirorr_status mirorr_half_encode(const float* data, size_t n, uint16_t* out,
                                            double* offset, double* scale) {
  if ((n > 0 && (!data || !out)) || !offset || !scale) return MIRORR_ERR_ARG;
  const mirorr::HalfEncoding e = mirorr::half_encoding(data, n);
  for (size_t i0 = 0; i0 < n; i0 += 4096) {
    mirorr::half_t h[4096];
    const size_t k = std::min<size_t>(4096, n - i0);
    mirorr::encode_half(data + i0, k, e, h);
    std::memcpy(out + i0, h, k * sizeof(uint16_t));
  }
  *offset = e.offset;
  *scale = e.scale;
  return MIRORR_OK;
}
