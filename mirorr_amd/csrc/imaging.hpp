// Resampling and Gaussian pyramid on the device (implemented in imaging.hip).
#pragma once

#include "common.hpp"
#include "xform.hpp"

namespace mirorr {

// Output voxel index -> input continuous index, ci = A i + b (double).
struct IndexMap {
  double A[9];
  double b[3];
};

// Builds the index map of y = M x + off between two grids:
// A = S_in^-1 D_in^-1 M D_out S_out, b = S_in^-1 D_in^-1 (M o_out + off - o_in).
__host__ __device__ inline IndexMap make_index_map(const xf::M3& M, const double off[3],
                                                   const Geom& gin, const Geom& gout) {
  xf::M3 Din, Dout;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      Din.a[r][c] = gin.dir[3 * r + c];
      Dout.a[r][c] = gout.dir[3 * r + c] * gout.spacing[c];
    }
  const xf::M3 Dinv = xf::inverse(Din);
  const xf::M3 P = xf::mul(Dinv, xf::mul(M, Dout));
  double q[3];
  xf::mulv(M, gout.origin, q);
  for (int r = 0; r < 3; ++r) q[r] += off[r] - gin.origin[r];
  double u[3];
  xf::mulv(Dinv, q, u);
  IndexMap m;
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) m.A[3 * r + c] = P.a[r][c] / gin.spacing[r];
    m.b[r] = u[r] / gin.spacing[r];
  }
  return m;
}

__device__ __forceinline__ int64_t lin3(int nx, int ny, int x, int y, int z) {
  return (int64_t)x + (int64_t)nx * ((int64_t)y + (int64_t)ny * z);
}

// ITK LinearInterpolateImageFunction at continuous index (cx, cy, cz):
// *inside is false (and 0 returned) outside [-0.5, n-0.5) on any axis;
// otherwise the trilinear blend, neighbours clamped to the buffer.
template <class T>
__device__ __forceinline__ double trilinear(const T* __restrict__ img, int nx, int ny, int nz,
                                            double cx, double cy, double cz, bool* inside) {
  if (!(cx >= -0.5 && cx < nx - 0.5 && cy >= -0.5 && cy < ny - 0.5 && cz >= -0.5 &&
        cz < nz - 0.5)) {
    *inside = false;
    return 0.0;
  }
  *inside = true;
  int bx = (int)floor(cx), by = (int)floor(cy), bz = (int)floor(cz);
  double fx = cx - bx, fy = cy - by, fz = cz - bz;
  if (bx < 0) { bx = 0; fx = 0.0; }
  if (by < 0) { by = 0; fy = 0.0; }
  if (bz < 0) { bz = 0; fz = 0.0; }
  const int x1 = min(bx + 1, nx - 1), y1 = min(by + 1, ny - 1), z1 = min(bz + 1, nz - 1);
  const double v000 = img[lin3(nx, ny, bx, by, bz)], v100 = img[lin3(nx, ny, x1, by, bz)];
  const double v010 = img[lin3(nx, ny, bx, y1, bz)], v110 = img[lin3(nx, ny, x1, y1, bz)];
  const double v001 = img[lin3(nx, ny, bx, by, z1)], v101 = img[lin3(nx, ny, x1, by, z1)];
  const double v011 = img[lin3(nx, ny, bx, y1, z1)], v111 = img[lin3(nx, ny, x1, y1, z1)];
  const double x00 = v000 + (v100 - v000) * fx, x10 = v010 + (v110 - v010) * fx;
  const double x01 = v001 + (v101 - v001) * fx, x11 = v011 + (v111 - v011) * fx;
  const double xx0 = x00 + (x10 - x00) * fy, xx1 = x01 + (x11 - x01) * fy;
  return xx0 + (xx1 - xx0) * fz;
}

enum MaskMode : int {
  kMaskNone = 0,     // no mask output
  kMaskInside = 1,   // the reference's default all-255 mask: 255 inside the source grid
  kMaskLinear = 2,   // interpolate a uint8 mask and truncate
};

// Resample `in` (dims_in) onto an output grid of dims_out under `map`
// (linear interpolation, neighbours clamped to the buffer, `outside` outside
// [-0.5, n-0.5): the encoding of 0). A mask is written under `mmap` from
// the mask grid (dims_m; a separate grid so full-resolution masks serve
// every level). Maps are read from device memory (a transform the LTS
// updates on device). `in` and `out` hold elements of `storage`.
void launch_resample(const void* in, const int dims_in[3], const IndexMap* d_map,
                     const uint8_t* mask, const int dims_m[3], const IndexMap* d_mmap,
                     int mask_mode, const int dims_out[3], void* out, uint8_t* out_mask,
                     hipStream_t s, int storage = kStoreF32, float outside = 0.0f);

// ITK GaussianOperator coefficients (centre first), normalised to unit sum.
int gaussian_kernel(double variance, double max_error, int max_width, double* out, int cap);

// One recursive pyramid step: smooth `prev` per axis with variance
// (ratio/2)^2 voxels (zero-flux Neumann edges, double intermediates) and
// resample onto `gout`. `tmp` holds 2 * nvox(prev) doubles.
void launch_pyramid_step(const void* prev, const Geom& gprev, const int ratio[3],
                         const Geom& gout, void* out, double* tmp, IndexMap* d_map,
                         hipStream_t s, int storage = kStoreF32, float outside = 0.0f);

}  // namespace mirorr
