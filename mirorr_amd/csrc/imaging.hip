// Trilinear resampling and the Gaussian pyramid for gfx950.
//
// Resampling restates ITK's ResampleImageFilter + LinearInterpolateImageFunction
// as the reference uses them (itkMirorrRegistrationMethod.txx:452-580,
// itkSymMirorrRegistrationMethod.txx:388-484): the output voxel index maps
// affinely to a continuous input index (double); inside [-0.5, n-0.5) per
// axis the value is the trilinear blend with neighbours clamped to the
// buffer, outside it is 0; masks are truncated to uint8. One thread per
// output voxel, x fastest so the writes coalesce; the gathers hit L2/MALL
// since consecutive x map to nearby inputs. HBM-bound.
//
// The pyramid restates RecursiveMultiResolutionPyramidImageFilter
// (itkMirorrPyramidImplement.h:94, .txx:575-609): separable discrete
// Gaussian (modified-Bessel kernel, maximum error 0.1, Neumann edges) along
// z, y, x in double, then a linear resample onto the shrunk grid.
#include <cmath>
#include <vector>

#include "imaging.hpp"

namespace mirorr {
namespace {

constexpr int kThreads = 256;

template <class Tin, class Tout>
__global__ __launch_bounds__(kThreads) void resample_kernel(
    const Tin* __restrict__ in, int inx, int iny, int inz, const IndexMap* __restrict__ map,
    const uint8_t* __restrict__ mask, int mnx, int mny, int mnz,
    const IndexMap* __restrict__ mmap, int mask_mode, int nx, int ny, int nz,
    Tout* __restrict__ out, uint8_t* __restrict__ out_mask, float outside) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  if (x >= nx) return;
  const int64_t o = lin3(nx, ny, x, y, z);
  if (out) {
    const IndexMap& m = *map;
    const double cx = m.A[0] * x + m.A[1] * y + m.A[2] * z + m.b[0];
    const double cy = m.A[3] * x + m.A[4] * y + m.A[5] * z + m.b[1];
    const double cz = m.A[6] * x + m.A[7] * y + m.A[8] * z + m.b[2];
    bool inside;
    const double v = trilinear(in, inx, iny, inz, cx, cy, cz, &inside);
    out[o] = inside ? (Tout)(float)v : (Tout)outside;
  }
  if (mask_mode != kMaskNone) {
    const IndexMap& m = *mmap;
    const double cx = m.A[0] * x + m.A[1] * y + m.A[2] * z + m.b[0];
    const double cy = m.A[3] * x + m.A[4] * y + m.A[5] * z + m.b[1];
    const double cz = m.A[6] * x + m.A[7] * y + m.A[8] * z + m.b[2];
    uint8_t r;
    if (mask_mode == kMaskInside) {
      r = (cx >= -0.5 && cx < mnx - 0.5 && cy >= -0.5 && cy < mny - 0.5 && cz >= -0.5 &&
           cz < mnz - 0.5) ? 255 : 0;
    } else {
      bool inside;
      const double v = trilinear(mask, mnx, mny, mnz, cx, cy, cz, &inside);
      r = inside ? (uint8_t)fmin(255.0, fmax(0.0, v)) : 0;
    }
    out_mask[o] = r;
  }
}

// 1-D convolution along `axis` with a symmetric kernel (centre first) and
// zero-flux Neumann edges (index clamped to the image).
template <class Tin, class Tout>
__global__ __launch_bounds__(kThreads) void conv_axis_kernel(
    const Tin* __restrict__ in, Tout* __restrict__ out, int nx, int ny, int nz, int axis,
    int radius, const double* __restrict__ k) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  if (x >= nx) return;
  const int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
  const int p = axis == 0 ? x : (axis == 1 ? y : z);
  const int64_t stride = axis == 0 ? 1 : (axis == 1 ? (int64_t)nx : (int64_t)nx * ny);
  const int64_t base = lin3(nx, ny, x, y, z) - (int64_t)p * stride;
  double s = 0.0;
  for (int j = -radius; j <= radius; ++j) {
    const int q = min(max(p + j, 0), n - 1);
    s += k[j < 0 ? -j : j] * (double)in[base + (int64_t)q * stride];
  }
  out[lin3(nx, ny, x, y, z)] = (Tout)s;
}

__global__ void identity_map_kernel(IndexMap* m, Geom gin, Geom gout) {
  const xf::M3 I = xf::eye3();
  const double z[3] = {0.0, 0.0, 0.0};
  *m = make_index_map(I, z, gin, gout);
}

// Modified Bessel functions as ITK's GaussianOperator evaluates them
// (Abramowitz & Stegun 9.8.1-9.8.4 polynomials for I0, I1; Miller's
// downward recurrence normalised by I0 for n >= 2).
double bessel_i0(double y) {
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
double bessel_i1(double y) {
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
double bessel_in(int n, double y) {
  if (y == 0.0) return 0.0;
  const double tox = 2.0 / std::fabs(y);
  double bip = 0.0, ans = 0.0, bi = 1.0;
  for (int j = 2 * (n + (int)std::sqrt(40.0 * n)); j > 0; --j) {
    const double bim = bip + j * tox * bi;
    bip = bi;
    bi = bim;
    if (std::fabs(bi) > 1e10) {
      ans *= 1e-10;
      bi *= 1e-10;
      bip *= 1e-10;
    }
    if (j == n) ans = bip;
  }
  ans *= bessel_i0(y) / bi;
  return (y < 0.0 && (n & 1)) ? -ans : ans;
}

}  // namespace

int gaussian_kernel(double variance, double max_error, int max_width, double* out, int cap) {
  std::vector<double> c;
  const double et = std::exp(-variance);
  double sum = 0.0;
  c.push_back(et * bessel_i0(variance));
  sum += c[0];
  c.push_back(et * bessel_i1(variance));
  sum += 2.0 * c[1];
  for (int i = 2; sum < 1.0 - max_error; ++i) {
    c.push_back(et * bessel_in(i, variance));
    sum += 2.0 * c[i];
    if (c[i] < sum * 2.220446049250313e-16) break;
    if ((int)c.size() > max_width) break;
  }
  const int n = (int)c.size();
  if (n > cap) throw Error(kErrArg, "Gaussian kernel wider than supported");
  for (int i = 0; i < n; ++i) out[i] = c[i] / sum;
  return n - 1;
}

namespace {
template <class Tin, class Tout>
void resample_as(const void* in, const int dims_in[3], const IndexMap* d_map,
                 const uint8_t* mask, const int dims_m[3], const IndexMap* d_mmap, int mask_mode,
                 const int dims_out[3], void* out, uint8_t* out_mask, hipStream_t s,
                 float outside) {
  dim3 grid((dims_out[0] + kThreads - 1) / kThreads, dims_out[1], dims_out[2]);
  const int mnx = dims_m ? dims_m[0] : 1, mny = dims_m ? dims_m[1] : 1, mnz = dims_m ? dims_m[2] : 1;
  hipLaunchKernelGGL((resample_kernel<Tin, Tout>), grid, dim3(kThreads), 0, s,
                     static_cast<const Tin*>(in), dims_in[0], dims_in[1], dims_in[2], d_map, mask,
                     mnx, mny, mnz, d_mmap, mask_mode, dims_out[0], dims_out[1], dims_out[2],
                     static_cast<Tout*>(out), out_mask, outside);
  MIRORR_CHECK_LAUNCH();
}
}  // namespace

void launch_resample(const void* in, const int dims_in[3], const IndexMap* d_map,
                     const uint8_t* mask, const int dims_m[3], const IndexMap* d_mmap,
                     int mask_mode, const int dims_out[3], void* out, uint8_t* out_mask,
                     hipStream_t s, int storage, float outside) {
  if (storage == kStoreF16)
    resample_as<half_t, half_t>(in, dims_in, d_map, mask, dims_m, d_mmap, mask_mode, dims_out,
                                out, out_mask, s, outside);
  else
    resample_as<float, float>(in, dims_in, d_map, mask, dims_m, d_mmap, mask_mode, dims_out, out,
                              out_mask, s, outside);
}

void launch_pyramid_step(const void* prev, const Geom& gprev, const int ratio[3],
                         const Geom& gout, void* out, double* tmp, IndexMap* d_map,
                         hipStream_t s, int storage, float outside) {
  const int nx = gprev.dims[0], ny = gprev.dims[1], nz = gprev.dims[2];
  const int64_t n = (int64_t)nx * ny * nz;
  bool all_ones = ratio[0] == 1 && ratio[1] == 1 && ratio[2] == 1;
  if (all_ones) {  // the filter copies the data onto the same grid
    MIRORR_HIP(hipMemcpyAsync(out, prev, n * storage_bytes(storage), hipMemcpyDeviceToDevice, s));
    return;
  }
  // Kernel coefficients per axis, uploaded once into the tail of tmp.
  double* a = tmp;
  double* b = tmp + n;
  static thread_local std::vector<double> hk;
  hk.assign(3 * 64, 0.0);
  int radius[3];
  for (int ax = 0; ax < 3; ++ax) {
    const double var = ratio[ax] != 1 ? (0.5 * ratio[ax]) * (0.5 * ratio[ax]) : 0.0;
    radius[ax] = gaussian_kernel(var, 0.1, 32, &hk[64 * ax], 64);
  }
  double* dk = tmp + 2 * n;  // 3 x 64 doubles reserved by the caller
  MIRORR_HIP(hipMemcpyAsync(dk, hk.data(), hk.size() * sizeof(double), hipMemcpyHostToDevice, s));
  dim3 grid((nx + kThreads - 1) / kThreads, ny, nz);
  // z, then y, then x; the last pass writes float into `out`'s scratch.
  if (storage == kStoreF16)
    hipLaunchKernelGGL((conv_axis_kernel<half_t, double>), grid, dim3(kThreads), 0, s,
                       static_cast<const half_t*>(prev), a, nx, ny, nz, 2, radius[2], dk + 128);
  else
    hipLaunchKernelGGL((conv_axis_kernel<float, double>), grid, dim3(kThreads), 0, s,
                       static_cast<const float*>(prev), a, nx, ny, nz, 2, radius[2], dk + 128);
  hipLaunchKernelGGL((conv_axis_kernel<double, double>), grid, dim3(kThreads), 0, s, a, b, nx,
                     ny, nz, 1, radius[1], dk + 64);
  float* smoothed = reinterpret_cast<float*>(a);  // a is free again
  hipLaunchKernelGGL((conv_axis_kernel<double, float>), grid, dim3(kThreads), 0, s, b, smoothed,
                     nx, ny, nz, 0, radius[0], dk);
  MIRORR_CHECK_LAUNCH();
  hipLaunchKernelGGL(identity_map_kernel, dim3(1), dim3(1), 0, s, d_map, gprev, gout);
  if (storage == kStoreF16)
    resample_as<float, half_t>(smoothed, gprev.dims, d_map, nullptr, nullptr, nullptr, kMaskNone,
                               gout.dims, out, nullptr, s, outside);
  else
    resample_as<float, float>(smoothed, gprev.dims, d_map, nullptr, nullptr, nullptr, kMaskNone,
                              gout.dims, out, nullptr, s, outside);
  // hk is reused by the next call: the copy above must finish first.
  MIRORR_HIP(hipStreamSynchronize(s));
}

}  // namespace mirorr
