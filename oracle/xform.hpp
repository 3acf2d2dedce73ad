// Oracle: ITK/VNL transform algebra restated for the registration loop
// (test infrastructure only; see oracle.h).
//
// ITK MatrixOffsetTransformBase semantics (used throughout
// itkMirorrRegistrationMethod.txx and itkSymMirorrRegistrationMethod.txx):
//   y = M (x - c) + c + t,   offset = t + c - M c
//   SetCenter / SetTranslation keep the other and recompute the offset;
//   Compose(other, pre): pre -> self(other(x)), else other(self(x)).
// Euler3DTransform: parameters (ax, ay, az, tx, ty, tz), M = Rz Rx Ry.
// AffineTransform: parameters (M row-major, t).
// VNL pieces: vnl_quaternion (x, y, z, r) with rotation_matrix_transpose,
// rotation_euler_angles (angles of R = Rz Ry Rx), axis, angle;
// vnl_symmetric_eigensystem (ascending eigenvalues); orthogonalize = U V^T
// (vnl_utils.h:124-135); Denman-Beavers square root (vnl_utils.h:46-119).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace orc {

struct M3 {
  double a[3][3];
};
struct M4 {
  double a[4][4];
};

inline M3 m3_identity() {
  M3 m{};
  for (int i = 0; i < 3; ++i) m.a[i][i] = 1.0;
  return m;
}
inline M3 m3_mul(const M3& x, const M3& y) {
  M3 r{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += x.a[i][k] * y.a[k][j];
      r.a[i][j] = s;
    }
  return r;
}
inline void m3_vec(const M3& m, const double v[3], double out[3]) {
  double r[3];
  for (int i = 0; i < 3; ++i) r[i] = m.a[i][0] * v[0] + m.a[i][1] * v[1] + m.a[i][2] * v[2];
  std::memcpy(out, r, sizeof r);
}
inline M3 m3_transpose(const M3& m) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = m.a[j][i];
  return r;
}
inline double m3_det(const M3& m) {
  const auto& a = m.a;
  return a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
         a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
         a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
}
// Inverse by cofactors (vnl_inverse / vnl_matrix_inverse of a 3x3).
inline M3 m3_inverse(const M3& m) {
  const auto& a = m.a;
  const double d = m3_det(m);
  M3 r;
  r.a[0][0] = (a[1][1] * a[2][2] - a[1][2] * a[2][1]) / d;
  r.a[0][1] = (a[0][2] * a[2][1] - a[0][1] * a[2][2]) / d;
  r.a[0][2] = (a[0][1] * a[1][2] - a[0][2] * a[1][1]) / d;
  r.a[1][0] = (a[1][2] * a[2][0] - a[1][0] * a[2][2]) / d;
  r.a[1][1] = (a[0][0] * a[2][2] - a[0][2] * a[2][0]) / d;
  r.a[1][2] = (a[0][2] * a[1][0] - a[0][0] * a[1][2]) / d;
  r.a[2][0] = (a[1][0] * a[2][1] - a[1][1] * a[2][0]) / d;
  r.a[2][1] = (a[0][1] * a[2][0] - a[0][0] * a[2][1]) / d;
  r.a[2][2] = (a[0][0] * a[1][1] - a[0][1] * a[1][0]) / d;
  return r;
}

// 4x4 inverse by Gauss-Jordan with partial pivoting.
inline M4 m4_inverse(const M4& m) {
  double w[4][8];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      w[i][j] = m.a[i][j];
      w[i][j + 4] = (i == j) ? 1.0 : 0.0;
    }
  for (int c = 0; c < 4; ++c) {
    int p = c;
    for (int r = c + 1; r < 4; ++r)
      if (std::fabs(w[r][c]) > std::fabs(w[p][c])) p = r;
    if (p != c)
      for (int j = 0; j < 8; ++j) std::swap(w[p][j], w[c][j]);
    const double d = w[c][c];
    for (int j = 0; j < 8; ++j) w[c][j] /= d;
    for (int r = 0; r < 4; ++r) {
      if (r == c) continue;
      const double f = w[r][c];
      if (f == 0.0) continue;
      for (int j = 0; j < 8; ++j) w[r][j] -= f * w[c][j];
    }
  }
  M4 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) r.a[i][j] = w[i][j + 4];
  return r;
}
inline M4 m4_mul(const M4& x, const M4& y) {
  M4 r{};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += x.a[i][k] * y.a[k][j];
      r.a[i][j] = s;
    }
  return r;
}

// vnl_matrix_sqrt by Denman-Beavers (vnl_utils.h:46-95).
inline M4 m4_sqrt_db(const M4& M, double eps, int max_iter) {
  M4 Y = M, Z{};
  for (int i = 0; i < 4; ++i) Z.a[i][i] = 1.0;
  double d = 1e300;
  int i = 0;
  while (d > eps && i < max_iter) {
    const M4 iz = m4_inverse(Z), iy = m4_inverse(Y);
    M4 nY, nZ;
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) {
        nY.a[r][c] = 0.5 * (Y.a[r][c] + iz.a[r][c]);
        nZ.a[r][c] = 0.5 * (Z.a[r][c] + iy.a[r][c]);
      }
    Y = nY;
    Z = nZ;
    ++i;
    const M4 YY = m4_mul(Y, Y);
    d = 0.0;
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) d = std::max(d, std::fabs(YY.a[r][c] - M.a[r][c]));
  }
  return Y;
}

// Cyclic Jacobi eigen-decomposition of a symmetric NxN matrix; eigenvalues
// ascending with eigenvectors as columns (vnl_symmetric_eigensystem order).
template <int N>
inline void sym_eigen(const double A[N][N], double evals[N], double evecs[N][N]) {
  double a[N][N], v[N][N];
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      a[i][j] = A[i][j];
      v[i][j] = (i == j) ? 1.0 : 0.0;
    }
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < N; ++p)
      for (int q = p + 1; q < N; ++q) off += a[p][q] * a[p][q];
    if (off < 1e-300) break;
    for (int p = 0; p < N; ++p)
      for (int q = p + 1; q < N; ++q) {
        if (a[p][q] == 0.0) continue;
        const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < N; ++k) {
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < N; ++k) {
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < N; ++k) {
          const double vkp = v[k][p], vkq = v[k][q];
          v[k][p] = c * vkp - s * vkq;
          v[k][q] = s * vkp + c * vkq;
        }
      }
  }
  int idx[N];
  for (int i = 0; i < N; ++i) idx[i] = i;
  std::sort(idx, idx + N, [&](int x, int y) { return a[x][x] < a[y][y]; });
  for (int i = 0; i < N; ++i) {
    evals[i] = a[idx[i]][idx[i]];
    for (int k = 0; k < N; ++k) evecs[k][i] = v[k][idx[i]];
  }
}

// orthogonalize (vnl_utils.h:124-135): U V^T of the SVD of m. Computed as
// m (m^T m)^(-1/2) from the eigen-decomposition of m^T m, which equals U V^T
// for a non-singular m.
inline M3 orthogonalize(const M3& m) {
  const M3 mtm = m3_mul(m3_transpose(m), m);
  double ev[3], V[3][3];
  sym_eigen<3>(mtm.a, ev, V);
  M3 inv_sqrt{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += V[i][k] * V[j][k] / std::sqrt(ev[k]);
      inv_sqrt.a[i][j] = s;
    }
  return m3_mul(m, inv_sqrt);
}

// ---- vnl_quaternion ---------------------------------------------------------
struct Quat {
  double x, y, z, r;
};
// rotation_matrix_transpose().transpose(): the rotation matrix of q.
inline M3 quat_rotation(const Quat& q) {
  const double x2 = q.x * q.x, y2 = q.y * q.y, z2 = q.z * q.z, r2 = q.r * q.r;
  const double xy = q.x * q.y, yz = q.y * q.z, zx = q.z * q.x;
  const double rx = q.r * q.x, ry = q.r * q.y, rz = q.r * q.z;
  M3 t;  // transpose as vnl builds it
  t.a[0][0] = r2 + x2 - y2 - z2;
  t.a[1][1] = r2 - x2 + y2 - z2;
  t.a[2][2] = r2 - x2 - y2 + z2;
  t.a[0][1] = 2 * (xy + rz);
  t.a[0][2] = 2 * (zx - ry);
  t.a[1][2] = 2 * (yz + rx);
  t.a[1][0] = 2 * (xy - rz);
  t.a[2][0] = 2 * (zx + ry);
  t.a[2][1] = 2 * (yz - rx);
  return m3_transpose(t);
}
// rotation_euler_angles: angles (about X, Y, Z) of R = Rz Ry Rx read from the
// transposed matrix as vnl does.
inline void quat_euler(const Quat& q, double ang[3]) {
  const M3 R = quat_rotation(q);
  const M3 t = m3_transpose(R);
  const double xx = t.a[0][0], xy = t.a[0][1], xz = t.a[0][2];
  const double yy = t.a[1][1], yz = t.a[1][2];
  const double zy = t.a[2][1], zz = t.a[2][2];
  const double cry = std::sqrt(xx * xx + xy * xy);
  if (cry > 1e-12) {
    ang[0] = std::atan2(yz, zz);
    ang[1] = std::atan2(-xz, cry);
    ang[2] = std::atan2(xy, xx);
  } else {
    ang[0] = std::atan2(-zy, yy);
    ang[1] = std::atan2(-xz, cry);
    ang[2] = 0.0;
  }
}
inline void quat_axis(const Quat& q, double ax[3]) {
  const double m = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z);
  if (m == 0.0) {
    ax[0] = ax[1] = 0.0;
    ax[2] = 1.0;
  } else {
    ax[0] = q.x / m;
    ax[1] = q.y / m;
    ax[2] = q.z / m;
  }
}
inline double quat_angle(const Quat& q) {
  return 2.0 * std::atan2(std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z), q.r);
}
inline Quat quat_from_axis_angle(const double ax[3], double angle) {
  const double a = 0.5 * angle, s = std::sin(a);
  return Quat{s * ax[0], s * ax[1], s * ax[2], std::cos(a)};
}

// ---- ITK transforms ----------------------------------------------------------
enum Kind { kEuler3D = 0, kAffine = 1 };

struct Xf {
  Kind kind = kEuler3D;
  M3 M = m3_identity();
  double c[3] = {0, 0, 0};    // center (fixed parameters)
  double t[3] = {0, 0, 0};    // translation
  double off[3] = {0, 0, 0};  // offset
  double ang[3] = {0, 0, 0};  // Euler angles (Euler3D only)

  void compute_offset() {
    double mc[3];
    m3_vec(M, c, mc);
    for (int i = 0; i < 3; ++i) off[i] = t[i] + c[i] - mc[i];
  }
  void compute_translation() {
    double mc[3];
    m3_vec(M, c, mc);
    for (int i = 0; i < 3; ++i) t[i] = off[i] - c[i] + mc[i];
  }
  // Euler3DTransform::ComputeMatrix (Z*X*Y).
  void compute_matrix_from_angles() {
    const double cx = std::cos(ang[0]), sx = std::sin(ang[0]);
    const double cy = std::cos(ang[1]), sy = std::sin(ang[1]);
    const double cz = std::cos(ang[2]), sz = std::sin(ang[2]);
    M3 rx{{{1, 0, 0}, {0, cx, -sx}, {0, sx, cx}}};
    M3 ry{{{cy, 0, sy}, {0, 1, 0}, {-sy, 0, cy}}};
    M3 rz{{{cz, -sz, 0}, {sz, cz, 0}, {0, 0, 1}}};
    M = m3_mul(m3_mul(rz, rx), ry);
  }
  // Euler3DTransform::ComputeMatrixParameters (Z*X*Y branch).
  void compute_angles_from_matrix() {
    ang[0] = std::asin(M.a[2][1]);
    const double A = std::cos(ang[0]);
    if (std::fabs(A) > 0.00005) {
      ang[1] = std::atan2(-M.a[2][0] / A, M.a[2][2] / A);
      ang[2] = std::atan2(-M.a[0][1] / A, M.a[1][1] / A);
    } else {
      ang[2] = 0.0;
      ang[1] = std::atan2(M.a[1][0], M.a[0][0]);
    }
  }
  void set_identity() {
    M = m3_identity();
    for (int i = 0; i < 3; ++i) c[i] = t[i] = off[i] = ang[i] = 0.0;
  }
  void set_center(const double cc[3]) {
    for (int i = 0; i < 3; ++i) c[i] = cc[i];
    compute_offset();
  }
  void set_translation(const double tt[3]) {
    for (int i = 0; i < 3; ++i) t[i] = tt[i];
    compute_offset();
  }
  // SetMatrix; Rigid3DTransform rejects non-orthogonal matrices (tol 1e-10).
  void set_matrix(const M3& m, bool check_orthogonal = true) {
    if (kind == kEuler3D && check_orthogonal) {
      const M3 mtm = m3_mul(m, m3_transpose(m));
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          if (std::fabs(mtm.a[i][j] - (i == j ? 1.0 : 0.0)) > 1e-10)
            throw std::runtime_error("Attempting to set a non-orthogonal rotation matrix");
    }
    M = m;
    compute_offset();
    if (kind == kEuler3D) compute_angles_from_matrix();
  }
  int nparams() const { return kind == kEuler3D ? 6 : 12; }
  void set_params(const double* p) {
    if (kind == kEuler3D) {
      for (int i = 0; i < 3; ++i) ang[i] = p[i];
      compute_matrix_from_angles();
      for (int i = 0; i < 3; ++i) t[i] = p[3 + i];
    } else {
      for (int i = 0; i < 9; ++i) M.a[i / 3][i % 3] = p[i];
      for (int i = 0; i < 3; ++i) t[i] = p[9 + i];
    }
    compute_offset();
  }
  void get_params(double* p) const {
    if (kind == kEuler3D) {
      for (int i = 0; i < 3; ++i) p[i] = ang[i];
      for (int i = 0; i < 3; ++i) p[3 + i] = t[i];
    } else {
      for (int i = 0; i < 9; ++i) p[i] = M.a[i / 3][i % 3];
      for (int i = 0; i < 3; ++i) p[9 + i] = t[i];
    }
  }
  void apply(const double x[3], double y[3]) const {
    double r[3];
    m3_vec(M, x, r);
    for (int i = 0; i < 3; ++i) y[i] = r[i] + off[i];
  }
  // Compose(other, pre) of MatrixOffsetTransformBase.
  void compose(const Xf& o, bool pre) {
    double v[3];
    if (pre) {
      m3_vec(M, o.off, v);
      for (int i = 0; i < 3; ++i) off[i] = v[i] + off[i];
      M = m3_mul(M, o.M);
    } else {
      m3_vec(o.M, off, v);
      for (int i = 0; i < 3; ++i) off[i] = v[i] + o.off[i];
      M = m3_mul(o.M, M);
    }
    compute_translation();
    if (kind == kEuler3D) compute_angles_from_matrix();
  }
  // GetInverse: same center, inverse matrix, offset -M^-1 offset.
  Xf inverse() const {
    Xf r;
    r.kind = kind;
    for (int i = 0; i < 3; ++i) r.c[i] = c[i];
    r.M = m3_inverse(M);
    double v[3];
    m3_vec(r.M, off, v);
    for (int i = 0; i < 3; ++i) r.off[i] = -v[i];
    r.compute_translation();
    if (kind == kEuler3D) r.compute_angles_from_matrix();
    return r;
  }
};

// setCenterToOrigin / updateCenter (itkTransformUtils.h:22-57).
inline void set_center_to_origin(Xf& x) {
  const double off[3] = {x.off[0], x.off[1], x.off[2]};
  const double zero[3] = {0, 0, 0};
  x.set_center(zero);
  x.set_translation(off);
}
inline void update_center(Xf& x, const double c2[3]) {
  set_center_to_origin(x);
  const double ic2[3] = {-c2[0], -c2[1], -c2[2]};
  x.set_center(ic2);
  const double off[3] = {x.off[0], x.off[1], x.off[2]};
  x.set_center(c2);
  x.set_translation(off);
}

}  // namespace orc
