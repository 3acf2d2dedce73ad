// FP64 small-matrix algebra and ITK-convention transforms, usable on the host
// and inside the one-thread solve kernels (no exceptions, fixed sizes).
//
// Transform convention (itk::MatrixOffsetTransformBase): y = M (x - c) + c + t,
// offset = t + c - M c. Euler3D parameters (ax, ay, az, t) with M = Rz Rx Ry;
// affine parameters (M row-major, t). Quaternions are (x, y, z, r).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace mirorr {
namespace xf {

#define XF_HD __host__ __device__ inline

struct M3 {
  double a[3][3];
};
struct M4 {
  double a[4][4];
};

XF_HD M3 eye3() {
  M3 m;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) m.a[i][j] = i == j ? 1.0 : 0.0;
  return m;
}
XF_HD M3 mul(const M3& x, const M3& y) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += x.a[i][k] * y.a[k][j];
      r.a[i][j] = s;
    }
  return r;
}
XF_HD M4 mul(const M4& x, const M4& y) {
  M4 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += x.a[i][k] * y.a[k][j];
      r.a[i][j] = s;
    }
  return r;
}
XF_HD void mulv(const M3& m, const double* v, double* out) {
  double r[3];
  for (int i = 0; i < 3; ++i) r[i] = m.a[i][0] * v[0] + m.a[i][1] * v[1] + m.a[i][2] * v[2];
  out[0] = r[0];
  out[1] = r[1];
  out[2] = r[2];
}
XF_HD M3 transpose(const M3& m) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = m.a[j][i];
  return r;
}
XF_HD double det(const M3& m) {
  const auto& a = m.a;
  return a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
         a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
         a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
}
XF_HD M3 inverse(const M3& m) {
  const auto& a = m.a;
  const double id = 1.0 / det(m);  // one division, nine products
  M3 r;
  r.a[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) * id;
  r.a[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * id;
  r.a[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * id;
  r.a[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) * id;
  r.a[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * id;
  r.a[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * id;
  r.a[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) * id;
  r.a[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * id;
  r.a[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * id;
  return r;
}
// Closed-form inverse by 2x2 minors (Laplace expansion along the first two
// rows): no pivoting and only fixed indices, so a device thread keeps it in
// registers. The matrices inverted here (homogeneous transforms and their
// Denman-Beavers iterates) are well conditioned.
XF_HD M4 inverse(const M4& m) {
  const auto& a = m.a;
  const double s0 = a[0][0] * a[1][1] - a[1][0] * a[0][1];
  const double s1 = a[0][0] * a[1][2] - a[1][0] * a[0][2];
  const double s2 = a[0][0] * a[1][3] - a[1][0] * a[0][3];
  const double s3 = a[0][1] * a[1][2] - a[1][1] * a[0][2];
  const double s4 = a[0][1] * a[1][3] - a[1][1] * a[0][3];
  const double s5 = a[0][2] * a[1][3] - a[1][2] * a[0][3];
  const double c5 = a[2][2] * a[3][3] - a[3][2] * a[2][3];
  const double c4 = a[2][1] * a[3][3] - a[3][1] * a[2][3];
  const double c3 = a[2][1] * a[3][2] - a[3][1] * a[2][2];
  const double c2 = a[2][0] * a[3][3] - a[3][0] * a[2][3];
  const double c1 = a[2][0] * a[3][2] - a[3][0] * a[2][2];
  const double c0 = a[2][0] * a[3][1] - a[3][0] * a[2][1];
  const double det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
  const double id = 1.0 / det;
  M4 r;
  r.a[0][0] = (a[1][1] * c5 - a[1][2] * c4 + a[1][3] * c3) * id;
  r.a[0][1] = (-a[0][1] * c5 + a[0][2] * c4 - a[0][3] * c3) * id;
  r.a[0][2] = (a[3][1] * s5 - a[3][2] * s4 + a[3][3] * s3) * id;
  r.a[0][3] = (-a[2][1] * s5 + a[2][2] * s4 - a[2][3] * s3) * id;
  r.a[1][0] = (-a[1][0] * c5 + a[1][2] * c2 - a[1][3] * c1) * id;
  r.a[1][1] = (a[0][0] * c5 - a[0][2] * c2 + a[0][3] * c1) * id;
  r.a[1][2] = (-a[3][0] * s5 + a[3][2] * s2 - a[3][3] * s1) * id;
  r.a[1][3] = (a[2][0] * s5 - a[2][2] * s2 + a[2][3] * s1) * id;
  r.a[2][0] = (a[1][0] * c4 - a[1][1] * c2 + a[1][3] * c0) * id;
  r.a[2][1] = (-a[0][0] * c4 + a[0][1] * c2 - a[0][3] * c0) * id;
  r.a[2][2] = (a[3][0] * s4 - a[3][1] * s2 + a[3][3] * s0) * id;
  r.a[2][3] = (-a[2][0] * s4 + a[2][1] * s2 - a[2][3] * s0) * id;
  r.a[3][0] = (-a[1][0] * c3 + a[1][1] * c1 - a[1][2] * c0) * id;
  r.a[3][1] = (a[0][0] * c3 - a[0][1] * c1 + a[0][2] * c0) * id;
  r.a[3][2] = (-a[3][0] * s3 + a[3][1] * s1 - a[3][2] * s0) * id;
  r.a[3][3] = (a[2][0] * s3 - a[2][1] * s1 + a[2][2] * s0) * id;
  return r;
}

// Denman-Beavers matrix square root: Y <- (Y + Z^-1)/2, Z <- (Z + Y^-1)/2
// until max|Y Y - M| <= eps or max_iter iterations.
XF_HD M4 sqrt_db(const M4& M, double eps, int max_iter) {
  M4 Y = M, Z;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) Z.a[i][j] = i == j ? 1.0 : 0.0;
  double d = 1e300;
  for (int it = 0; d > eps && it < max_iter; ++it) {
    const M4 iz = inverse(Z), iy = inverse(Y);
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) {
        Y.a[r][c] = 0.5 * (Y.a[r][c] + iz.a[r][c]);
        Z.a[r][c] = 0.5 * (Z.a[r][c] + iy.a[r][c]);
      }
    const M4 yy = mul(Y, Y);
    d = 0.0;
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) d = fmax(d, fabs(yy.a[r][c] - M.a[r][c]));
  }
  return Y;
}

// Cyclic Jacobi on a symmetric N x N; eigenvalues ascending, vectors in columns.
template <int N>
XF_HD void eigen_sym(const double (&A)[N][N], double (&ev)[N], double (&V)[N][N]) {
  double a[N][N];
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      a[i][j] = A[i][j];
      V[i][j] = i == j ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = 0.0;
#pragma unroll
    for (int p = 0; p < N; ++p)
#pragma unroll
      for (int q = p + 1; q < N; ++q) off += a[p][q] * a[p][q];
    if (off < 1e-300) break;
#pragma unroll
    for (int p = 0; p < N; ++p)
#pragma unroll
      for (int q = p + 1; q < N; ++q) {
        if (a[p][q] == 0.0) continue;
        const double th = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const double x = a[k][p], y = a[k][q];
          a[k][p] = c * x - s * y;
          a[k][q] = s * x + c * y;
        }
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const double x = a[p][k], y = a[q][k];
          a[p][k] = c * x - s * y;
          a[q][k] = s * x + c * y;
        }
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const double x = V[k][p], y = V[k][q];
          V[k][p] = c * x - s * y;
          V[k][q] = s * x + c * y;
        }
      }
  }
#pragma unroll
  for (int i = 0; i < N; ++i) ev[i] = a[i][i];
  // Ascending order by compare-exchange on fixed indices (the order of
  // distinct eigenvalues and their vectors is what any sort gives).
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = i + 1; j < N; ++j) {
      if (ev[j] < ev[i]) {
        const double t = ev[i];
        ev[i] = ev[j];
        ev[j] = t;
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const double u = V[k][i];
          V[k][i] = V[k][j];
          V[k][j] = u;
        }
      }
    }
}

// Orthogonal polar factor U V^T of a non-singular m, as m (m^T m)^(-1/2).
XF_HD M3 orthogonalize(const M3& m) {
  const M3 mtm = mul(transpose(m), m);
  double ev[3], V[3][3];
  eigen_sym<3>(mtm.a, ev, V);
  M3 is;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += V[i][k] * V[j][k] / sqrt(ev[k]);
      is.a[i][j] = s;
    }
  return mul(m, is);
}

// ---- quaternions (x, y, z, r) ----------------------------------------------
struct Quat {
  double x, y, z, r;
};
XF_HD M3 quat_rotation(const Quat& q) {
  const double x2 = q.x * q.x, y2 = q.y * q.y, z2 = q.z * q.z, r2 = q.r * q.r;
  const double xy = q.x * q.y, yz = q.y * q.z, zx = q.z * q.x;
  const double rx = q.r * q.x, ry = q.r * q.y, rz = q.r * q.z;
  M3 R;
  R.a[0][0] = r2 + x2 - y2 - z2;
  R.a[1][1] = r2 - x2 + y2 - z2;
  R.a[2][2] = r2 - x2 - y2 + z2;
  R.a[1][0] = 2 * (xy + rz);
  R.a[2][0] = 2 * (zx - ry);
  R.a[2][1] = 2 * (yz + rx);
  R.a[0][1] = 2 * (xy - rz);
  R.a[0][2] = 2 * (zx + ry);
  R.a[1][2] = 2 * (yz - rx);
  return R;
}
// Angles (about X, Y, Z) of R = Rz Ry Rx, the convention of vnl_quaternion.
XF_HD void quat_euler_zyx(const Quat& q, double ang[3]) {
  const M3 R = quat_rotation(q);
  const double r00 = R.a[0][0], r10 = R.a[1][0], r20 = R.a[2][0];
  const double r21 = R.a[2][1], r22 = R.a[2][2], r12 = R.a[1][2], r11 = R.a[1][1];
  const double cry = sqrt(r00 * r00 + r10 * r10);
  if (cry > 1e-12) {
    ang[0] = atan2(r21, r22);
    ang[1] = atan2(-r20, cry);
    ang[2] = atan2(r10, r00);
  } else {
    ang[0] = atan2(-r12, r11);
    ang[1] = atan2(-r20, cry);
    ang[2] = 0.0;
  }
}
XF_HD Quat quat_axis_angle(const double ax[3], double angle) {
  const double h = 0.5 * angle, s = sin(h);
  return Quat{s * ax[0], s * ax[1], s * ax[2], cos(h)};
}

// ---- ITK transforms -------------------------------------------------------
enum : int { kRigid = 0, kAffine = 1 };

struct Xf {
  int kind;
  M3 M;
  double c[3], t[3], off[3], ang[3];
};

XF_HD void compute_offset(Xf& x) {
  double mc[3];
  mulv(x.M, x.c, mc);
  for (int i = 0; i < 3; ++i) x.off[i] = x.t[i] + x.c[i] - mc[i];
}
XF_HD void compute_translation(Xf& x) {
  double mc[3];
  mulv(x.M, x.c, mc);
  for (int i = 0; i < 3; ++i) x.t[i] = x.off[i] - x.c[i] + mc[i];
}
XF_HD M3 euler_zxy(const double ang[3]) {
  const double cx = cos(ang[0]), sx = sin(ang[0]);
  const double cy = cos(ang[1]), sy = sin(ang[1]);
  const double cz = cos(ang[2]), sz = sin(ang[2]);
  // Rz * Rx * Ry expanded
  M3 m;
  m.a[0][0] = cz * cy - sz * sx * sy;
  m.a[0][1] = -sz * cx;
  m.a[0][2] = cz * sy + sz * sx * cy;
  m.a[1][0] = sz * cy + cz * sx * sy;
  m.a[1][1] = cz * cx;
  m.a[1][2] = sz * sy - cz * sx * cy;
  m.a[2][0] = -cx * sy;
  m.a[2][1] = sx;
  m.a[2][2] = cx * cy;
  return m;
}
XF_HD void angles_from_matrix(Xf& x) {
  x.ang[0] = asin(x.M.a[2][1]);
  const double A = cos(x.ang[0]);
  if (fabs(A) > 0.00005) {
    x.ang[1] = atan2(-x.M.a[2][0] / A, x.M.a[2][2] / A);
    x.ang[2] = atan2(-x.M.a[0][1] / A, x.M.a[1][1] / A);
  } else {
    x.ang[2] = 0.0;
    x.ang[1] = atan2(x.M.a[1][0], x.M.a[0][0]);
  }
}
XF_HD Xf identity(int kind) {
  Xf x;
  x.kind = kind;
  x.M = eye3();
  for (int i = 0; i < 3; ++i) x.c[i] = x.t[i] = x.off[i] = x.ang[i] = 0.0;
  return x;
}
XF_HD int nparams(const Xf& x) { return x.kind == kRigid ? 6 : 12; }
XF_HD void set_center(Xf& x, const double c[3]) {
  for (int i = 0; i < 3; ++i) x.c[i] = c[i];
  compute_offset(x);
}
XF_HD void set_translation(Xf& x, const double t[3]) {
  for (int i = 0; i < 3; ++i) x.t[i] = t[i];
  compute_offset(x);
}
// Returns false when a rigid transform is given a matrix that is not
// orthogonal within 1e-10 (Rigid3DTransform::SetMatrix refuses it).
XF_HD bool set_matrix(Xf& x, const M3& m) {
  if (x.kind == kRigid) {
    const M3 p = mul(m, transpose(m));
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        if (fabs(p.a[i][j] - (i == j ? 1.0 : 0.0)) > 1e-10) return false;
  }
  x.M = m;
  compute_offset(x);
  if (x.kind == kRigid) angles_from_matrix(x);
  return true;
}
XF_HD void set_params(Xf& x, const double* p) {
  if (x.kind == kRigid) {
    for (int i = 0; i < 3; ++i) x.ang[i] = p[i];
    x.M = euler_zxy(x.ang);
    for (int i = 0; i < 3; ++i) x.t[i] = p[3 + i];
  } else {
    for (int i = 0; i < 9; ++i) x.M.a[i / 3][i % 3] = p[i];
    for (int i = 0; i < 3; ++i) x.t[i] = p[9 + i];
  }
  compute_offset(x);
}
XF_HD void get_params(const Xf& x, double* p) {
  if (x.kind == kRigid) {
    for (int i = 0; i < 3; ++i) p[i] = x.ang[i];
    for (int i = 0; i < 3; ++i) p[3 + i] = x.t[i];
  } else {
    for (int i = 0; i < 9; ++i) p[i] = x.M.a[i / 3][i % 3];
    for (int i = 0; i < 3; ++i) p[9 + i] = x.t[i];
  }
}
XF_HD void apply(const Xf& x, const double* p, double* out) {
  double r[3];
  mulv(x.M, p, r);
  for (int i = 0; i < 3; ++i) out[i] = r[i] + x.off[i];
}
// pre = true: x <- x o o (o first); else x <- o o x.
XF_HD void compose(Xf& x, const Xf& o, bool pre) {
  double v[3];
  if (pre) {
    mulv(x.M, o.off, v);
    for (int i = 0; i < 3; ++i) x.off[i] += v[i];
    x.M = mul(x.M, o.M);
  } else {
    mulv(o.M, x.off, v);
    for (int i = 0; i < 3; ++i) x.off[i] = v[i] + o.off[i];
    x.M = mul(o.M, x.M);
  }
  compute_translation(x);
  if (x.kind == kRigid) angles_from_matrix(x);
}
XF_HD Xf inverse(const Xf& x) {
  Xf r = x;
  r.M = inverse(x.M);
  double v[3];
  mulv(r.M, x.off, v);
  for (int i = 0; i < 3; ++i) r.off[i] = -v[i];
  compute_translation(r);
  if (r.kind == kRigid) angles_from_matrix(r);
  return r;
}
XF_HD void set_center_to_origin(Xf& x) {
  const double off[3] = {x.off[0], x.off[1], x.off[2]};
  const double z[3] = {0.0, 0.0, 0.0};
  set_center(x, z);
  set_translation(x, off);
}
XF_HD void update_center(Xf& x, const double c2[3]) {
  set_center_to_origin(x);
  const double ic[3] = {-c2[0], -c2[1], -c2[2]};
  set_center(x, ic);
  const double off[3] = {x.off[0], x.off[1], x.off[2]};
  set_center(x, c2);
  set_translation(x, off);
}
XF_HD Xf translation(const double t[3]) {
  Xf x = identity(kRigid);
  const double p[6] = {0.0, 0.0, 0.0, t[0], t[1], t[2]};
  set_params(x, p);
  return x;
}

#undef XF_HD

}  // namespace xf
}  // namespace mirorr
