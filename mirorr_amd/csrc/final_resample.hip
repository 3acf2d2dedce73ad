// Final output resampling of the command line (--save-resampled-*,
// --save-fixed, --save-moving) for gfx950: ITK's ResampleImageFilter with
// default value 0 and the interpolator GetInterpolatorFromString picks
// (itkMirorrPyramidImplement.txx:364-428, 436-470):
//   nearest neighbour   NearestNeighborInterpolateImageFunction
//   linear              LinearInterpolateImageFunction
//   cubic B-spline      BSplineInterpolateImageFunction, order 3
//   windowed sinc       WindowedSincInterpolateImageFunction, radius 5,
//                       Welch window, zero-flux Neumann edges
// Every interpolator answers only inside [-0.5, n-0.5) per axis (ITK's
// IsInsideBuffer); outside the output is 0.
//
// The B-spline coefficients follow Unser's recursive prefilter (M. Unser,
// "Splines: a perfect fit for signal and image processing", IEEE Signal
// Processing Magazine 16(6), 1999, Box 2): pole z = sqrt(3) - 2, gain
// (1 - z)(1 - 1/z), mirror-symmetric causal initialisation truncated where
// |z|^n < 1e-10, anticausal initialisation z/(z^2 - 1) (z c[N-2] + c[N-1]),
// axes x, y, z in turn, in double. One thread per line (lines are
// independent recursions; consecutive threads take consecutive lines, so for
// the y and z passes the accesses coalesce). Evaluation gathers 4^3 double
// coefficients with mirrored indices; the sinc gathers 10^3 clamped voxels.
// This runs once per output image, after the registration.
#include <cmath>

#include "context.hpp"
#include "imaging.hpp"

namespace mirorr {
namespace {

constexpr int kThreads = 256;
constexpr int kSincRadius = 5;

__global__ __launch_bounds__(kThreads) void to_double_kernel(const float* __restrict__ in,
                                                             double* __restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

// Cubic B-spline prefilter along one axis, in place. `lines` lines of
// length N; line l starts at base(l) and steps by `stride`.
__global__ __launch_bounds__(kThreads) void bspline_prefilter_kernel(
    double* __restrict__ c, int nx, int ny, int nz, int axis, int horizon) {
  const int N = axis == 0 ? nx : axis == 1 ? ny : nz;
  const int64_t lines = (int64_t)nx * ny * nz / N;
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= lines || N < 2) return;
  int64_t base, stride;
  if (axis == 0) {  // l = y + ny * z
    base = l * nx;
    stride = 1;
  } else if (axis == 1) {  // l = x + nx * z
    const int64_t x = l % nx, z = l / nx;
    base = x + (int64_t)nx * ny * z;
    stride = nx;
  } else {  // l = x + nx * y
    base = l;
    stride = (int64_t)nx * ny;
  }
  double* p = c + base;
  const double z = sqrt(3.0) - 2.0;
  const double gain = (1.0 - z) * (1.0 - 1.0 / z);
  for (int n = 0; n < N; ++n) p[n * stride] *= gain;
  // Causal initial coefficient, mirror boundaries.
  if (horizon < N) {
    double sum = p[0], zn = z;
    for (int n = 1; n < horizon; ++n) {
      sum += zn * p[n * stride];
      zn *= z;
    }
    p[0] = sum;
  } else {
    const double iz = 1.0 / z;
    double zn = z, z2n = pow(z, (double)(N - 1));
    double sum = p[0] + z2n * p[(int64_t)(N - 1) * stride];
    z2n *= z2n * iz;
    for (int n = 1; n <= N - 2; ++n) {
      sum += (zn + z2n) * p[n * stride];
      zn *= z;
      z2n *= iz;
    }
    p[0] = sum / (1.0 - zn * zn);
  }
  double prev = p[0];
  for (int n = 1; n < N; ++n) {
    const double v = p[n * stride] + z * prev;
    p[n * stride] = v;
    prev = v;
  }
  // Anticausal initial coefficient and recursion.
  const int64_t last = (int64_t)(N - 1) * stride;
  double next = (z / (z * z - 1.0)) * (z * p[last - stride] + p[last]);
  p[last] = next;
  for (int n = N - 2; n >= 0; --n) {
    const double v = z * (next - p[n * stride]);
    p[n * stride] = v;
    next = v;
  }
}

// Mirror an index into [0, N-1] as BSplineInterpolateImageFunction does
// (i < 0 -> -i, i > N-1 -> 2(N-1) - i). The indices used lie in -2..N+1,
// where one reflection suffices for N >= 3; for N = 2 ITK's single pass
// leaves the buffer, and the reflection is repeated here instead (the
// whole-sample symmetric extension of period 2(N-1)).
__device__ __forceinline__ int mirror_index(int i, int N) {
  if (N == 1) return 0;
  for (int r = 0; r < 2; ++r) {
    if (i < 0) i = -i;
    if (i > N - 1) i = 2 * (N - 1) - i;
  }
  return i;
}

__device__ __forceinline__ void cubic_weights(double w, double o[4]) {
  const double w2 = w * w, w3 = w2 * w, u = 1.0 - w;
  o[0] = u * u * u / 6.0;
  o[1] = (4.0 - 6.0 * w2 + 3.0 * w3) / 6.0;
  o[2] = (1.0 + 3.0 * w + 3.0 * w2 - 3.0 * w3) / 6.0;
  o[3] = w3 / 6.0;
}

__device__ __forceinline__ double sinc_weight(double x) {
  const double px = M_PI * x;
  const double s = x == 0.0 ? 1.0 : sin(px) / px;
  return (1.0 - x * x / (double)(kSincRadius * kSincRadius)) * s;  // Welch window
}

template <int kInterp>
__global__ __launch_bounds__(kThreads) void final_resample_kernel(
    const float* __restrict__ img, const double* __restrict__ coef, int inx, int iny, int inz,
    IndexMap m, int nx, int ny, int nz, float* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  if (x >= nx) return;
  const double cx = m.A[0] * x + m.A[1] * y + m.A[2] * z + m.b[0];
  const double cy = m.A[3] * x + m.A[4] * y + m.A[5] * z + m.b[1];
  const double cz = m.A[6] * x + m.A[7] * y + m.A[8] * z + m.b[2];
  const int64_t o = lin3(nx, ny, x, y, z);
  if (!(cx >= -0.5 && cx < inx - 0.5 && cy >= -0.5 && cy < iny - 0.5 && cz >= -0.5 &&
        cz < inz - 0.5)) {
    out[o] = 0.0f;
    return;
  }
  double v = 0.0;
  if (kInterp == MIRORR_INTERP_NN) {
    // RoundHalfIntegerUp; inside the buffer test the result is in range.
    const int ix = min((int)floor(cx + 0.5), inx - 1), iy = min((int)floor(cy + 0.5), iny - 1),
              iz = min((int)floor(cz + 0.5), inz - 1);
    v = img[lin3(inx, iny, ix, iy, iz)];
  } else if (kInterp == MIRORR_INTERP_LINEAR) {
    bool inside;
    v = trilinear(img, inx, iny, inz, cx, cy, cz, &inside);
  } else if (kInterp == MIRORR_INTERP_BSPLINE) {
    const int bx = (int)floor(cx), by = (int)floor(cy), bz = (int)floor(cz);
    double wx[4], wy[4], wz[4];
    cubic_weights(cx - bx, wx);
    cubic_weights(cy - by, wy);
    cubic_weights(cz - bz, wz);
    int ix[4], iy[4];
    for (int k = 0; k < 4; ++k) {
      ix[k] = mirror_index(bx - 1 + k, inx);
      iy[k] = mirror_index(by - 1 + k, iny);
    }
    for (int kz = 0; kz < 4; ++kz) {
      const int iz = mirror_index(bz - 1 + kz, inz);
      double sy = 0.0;
      for (int ky = 0; ky < 4; ++ky) {
        const double* row = coef + lin3(inx, iny, 0, iy[ky], iz);
        double sx = 0.0;
        for (int kx = 0; kx < 4; ++kx) sx += wx[kx] * row[ix[kx]];
        sy += wy[ky] * sx;
      }
      v += wz[kz] * sy;
    }
  } else {  // windowed sinc, offsets -4..5 around floor(ci)
    const int bx = (int)floor(cx), by = (int)floor(cy), bz = (int)floor(cz);
    const double dx = cx - bx, dy = cy - by, dz = cz - bz;
    constexpr int W = 2 * kSincRadius;
    double wx[W], wy[W], wz[W];
    for (int i = 0; i < W; ++i) {
      const int off = i - (kSincRadius - 1);
      wx[i] = dx == 0.0 ? (off == 0 ? 1.0 : 0.0) : sinc_weight(dx - off);
      wy[i] = dy == 0.0 ? (off == 0 ? 1.0 : 0.0) : sinc_weight(dy - off);
      wz[i] = dz == 0.0 ? (off == 0 ? 1.0 : 0.0) : sinc_weight(dz - off);
    }
    int ix[W];
    for (int i = 0; i < W; ++i) ix[i] = min(max(bx + i - (kSincRadius - 1), 0), inx - 1);
    for (int k = 0; k < W; ++k) {
      const int iz = min(max(bz + k - (kSincRadius - 1), 0), inz - 1);
      double sy = 0.0;
      for (int j = 0; j < W; ++j) {
        const int iy = min(max(by + j - (kSincRadius - 1), 0), iny - 1);
        const float* row = img + lin3(inx, iny, 0, iy, iz);
        double sx = 0.0;
        for (int i = 0; i < W; ++i) sx += wx[i] * (double)row[ix[i]];
        sy += wy[j] * sx;
      }
      v += wz[k] * sy;
    }
  }
  out[o] = (float)v;
}

// Causal-initialisation horizon: the first n with |z|^n < 1e-10 (ITK's
// default tolerance of the decomposition filter).
int bspline_horizon() {
  const double z = std::sqrt(3.0) - 2.0;
  return (int)std::ceil(std::log(1e-10) / std::log(std::fabs(z)));
}

}  // namespace
}  // namespace mirorr

using namespace mirorr;

extern "C" mirorr_status mirorr_resample_final(mirorr_ctx* c, const float* data,
                                               const mirorr_geom* gin, const double matrix[9],
                                               const double offset[3], const mirorr_geom* gout,
                                               int interpolator, float* out) {
  if (!c || !data || !gin || !matrix || !offset || !gout || !out) return MIRORR_ERR_ARG;
  if (interpolator < MIRORR_INTERP_NN || interpolator > MIRORR_INTERP_SINC) {
    c->last_error = "unknown interpolator";
    return MIRORR_ERR_ARG;
  }
  for (int i = 0; i < 3; ++i)
    if (gin->dims[i] < 1 || gout->dims[i] < 1 || gout->dims[i] > 65535) {
      c->last_error = "grid dimensions must be in 1..65535";
      return MIRORR_ERR_ARG;
    }
  return guarded(c, [&] {
    hipStream_t s = c->stream;
    const Geom gi = to_geom(*gin), go = to_geom(*gout);
    const int64_t ni = (int64_t)gi.dims[0] * gi.dims[1] * gi.dims[2];
    const int64_t no = (int64_t)go.dims[0] * go.dims[1] * go.dims[2];
    DeviceBuffer din, dout, dcoef;
    din.ensure(sizeof(float) * ni, &c->budget);
    dout.ensure(sizeof(float) * no, &c->budget);
    MIRORR_HIP(hipMemcpyAsync(din.as<float>(), data, sizeof(float) * ni, hipMemcpyHostToDevice, s));
    xf::M3 M;
    for (int i = 0; i < 9; ++i) M.a[i / 3][i % 3] = matrix[i];
    const IndexMap map = make_index_map(M, offset, gi, go);
    double* coef = nullptr;
    if (interpolator == MIRORR_INTERP_BSPLINE) {
      dcoef.ensure(sizeof(double) * ni, &c->budget);
      coef = dcoef.as<double>();
      hipLaunchKernelGGL(to_double_kernel, dim3((unsigned)((ni + kThreads - 1) / kThreads)),
                         dim3(kThreads), 0, s, din.as<float>(), coef, ni);
      const int horizon = bspline_horizon();
      for (int axis = 0; axis < 3; ++axis) {
        const int64_t lines = ni / gi.dims[axis];
        hipLaunchKernelGGL(bspline_prefilter_kernel,
                           dim3((unsigned)((lines + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                           s, coef, gi.dims[0], gi.dims[1], gi.dims[2], axis, horizon);
      }
    }
    const dim3 grid((go.dims[0] + kThreads - 1) / kThreads, go.dims[1], go.dims[2]);
    const int* di = gi.dims;
    const int* dn = go.dims;
    switch (interpolator) {
      case MIRORR_INTERP_NN:
        hipLaunchKernelGGL((final_resample_kernel<MIRORR_INTERP_NN>), grid, dim3(kThreads), 0, s,
                           din.as<float>(), coef, di[0], di[1], di[2], map, dn[0], dn[1], dn[2],
                           dout.as<float>());
        break;
      case MIRORR_INTERP_LINEAR:
        hipLaunchKernelGGL((final_resample_kernel<MIRORR_INTERP_LINEAR>), grid, dim3(kThreads), 0,
                           s, din.as<float>(), coef, di[0], di[1], di[2], map, dn[0], dn[1], dn[2],
                           dout.as<float>());
        break;
      case MIRORR_INTERP_BSPLINE:
        hipLaunchKernelGGL((final_resample_kernel<MIRORR_INTERP_BSPLINE>), grid, dim3(kThreads),
                           0, s, din.as<float>(), coef, di[0], di[1], di[2], map, dn[0], dn[1],
                           dn[2], dout.as<float>());
        break;
      default:
        hipLaunchKernelGGL((final_resample_kernel<MIRORR_INTERP_SINC>), grid, dim3(kThreads), 0,
                           s, din.as<float>(), coef, di[0], di[1], di[2], map, dn[0], dn[1], dn[2],
                           dout.as<float>());
        break;
    }
    MIRORR_HIP(hipGetLastError());
    MIRORR_HIP(hipMemcpyAsync(out, dout.as<float>(), sizeof(float) * no, hipMemcpyDeviceToHost, s));
    MIRORR_HIP(hipStreamSynchronize(s));
  });
}
