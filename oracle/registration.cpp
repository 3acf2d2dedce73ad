// Oracle: LTS regression and the registration loops
// (test infrastructure only; see oracle.h).
//
// Restates:
//   MirorrRegistrationMethod::OptimizeTransform  itkMirorrRegistrationMethod.txx:758-1003
//     GetBaryCentre :1005-1019, GetParametersDelta :1021-1036,
//     CovarianceOfPoints :1038-1061, CalculateRigidQuaternion :1063-1106
//   MirorrRegistrationMethod::StartOptimization  itkMirorrRegistrationMethod.txx:422-659
//   SymMirorrRegistrationMethod::ComputeHalfSpaceTransforms  itkSymMirorrRegistrationMethod.txx:139-201
//   SymMirorrRegistrationMethod::StartOptimization itkSymMirorrRegistrationMethod.txx:238-720
//   SymMirorrRegistrationMethod::OptimizeTransform itkSymMirorrRegistrationMethod.txx:729-1442
//   MirorrPyramidImplement::RunMirorr (level loop) itkMirorrPyramidImplement.txx:71-158
//   CenteredTransformInitializer use               itkMirorrPyramidWrapper.txx:611-623
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "oracle.h"
#include "oracle_reg.h"
#include "xform.hpp"

using namespace orc;

namespace {

typedef std::vector<double> Pts;  // n x 3

// CovarianceOfPoints(A, B)[r][c] = sum A_r B_c / n.
M3 covariance(const Pts& a, const Pts& b) {
  M3 m{};
  const size_t n = a.size() / 3;
  for (size_t i = 0; i < n; ++i)
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) m.a[r][c] += a[3 * i + r] * b[3 * i + c];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) m.a[r][c] /= (double)n;
  return m;
}

void barycentre(const Pts& p, double m[3]) {
  m[0] = m[1] = m[2] = 0.0;
  const size_t n = p.size() / 3;
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) m[k] += p[3 * i + k];
  for (int k = 0; k < 3; ++k) m[k] /= (double)n;
}

double params_delta(const double* a, const double* b, int n) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) {
    const double d = a[i] - b[i];
    s += d * d;
  }
  return s;
}

// CalculateRigidQuaternion: eigenvector of the smallest eigenvalue of
// sum Q^T Q; quaternion (x, y, z, r) = (e1, e2, e3, e0).
Quat rigid_quaternion(const Pts& x, const Pts& y) {
  double T[4][4] = {};
  const size_t n = x.size() / 3;
  for (size_t i = 0; i < n; ++i) {
    double q[4][4] = {};
    for (int k = 0; k < 3; ++k) q[0][1 + k] = x[3 * i + k] - y[3 * i + k];
    q[1][2] = -x[3 * i + 2] - y[3 * i + 2];
    q[1][3] = x[3 * i + 1] + y[3 * i + 1];
    q[2][3] = -x[3 * i + 0] - y[3 * i + 0];
    for (int a = 0; a < 4; ++a)
      for (int b = a + 1; b < 4; ++b) q[b][a] = -q[a][b];
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        double s = 0.0;
        for (int k = 0; k < 4; ++k) s += q[k][a] * q[k][b];
        T[a][b] += s;
      }
  }
  double ev[4], V[4][4];
  sym_eigen<4>(T, ev, V);
  return Quat{V[1][0], V[2][0], V[3][0], V[0][0]};
}

// Residual subset selection shared by both LTS variants
// (itkMirorrRegistrationMethod.txx:809-839): fraction 1 keeps all points,
// else r_i <= the len-th smallest residual, len = floor(fraction * n).
double nth_residual(const std::vector<double>& r, double fraction) {
  if (fraction > 0.99999) return *std::max_element(r.begin(), r.end());
  std::vector<double> s(r);
  unsigned len = (unsigned)(fraction * s.size());
  len = std::min<unsigned>(len, (unsigned)s.size() - 1);
  std::nth_element(s.begin(), s.begin() + len, s.end());
  return s[len];
}

M4 to4(const M3& m, const double t[3]) {
  M4 r{};
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) r.a[i][j] = m.a[i][j];
    r.a[i][3] = t[i];
  }
  r.a[3][3] = 1.0;
  return r;
}
M3 top3(const M4& m) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[i][j] = m.a[i][j];
  return r;
}

}  // namespace

namespace orc {

int g_verbosity = 0;

// Classic LTS (itkMirorrRegistrationMethod.txx:758-1003). T is composed with
// the fitted transform (pre-multiplied). Returns the inner iteration count.
int lts_classic(Xf& T, const Pts& X, const Pts& Y, double tol) {
  Xf fin;
  fin.kind = T.kind;
  double rc[3] = {T.c[0], T.c[1], T.c[2]};
  fin.set_identity();
  fin.set_center(rc);
  const int np = fin.nparams();
  double last[12], neu[12];
  fin.get_params(last);  // never updated below, as in the reference
  double fraction = 1.0, last_delta = 0.0, delta = 1.0;
  int iter = -1;
  const size_t n = X.size() / 3;
  while (delta > tol && iter < 100 && std::fabs(delta - last_delta) > tol) {
    ++iter;
    if (n == 0) break;
    std::vector<double> r(n, 0.0);
    for (size_t i = 0; i < n; ++i) {
      double p[3];
      fin.apply(&X[3 * i], p);
      for (int k = 0; k < 3; ++k) {
        const double d = p[k] - Y[3 * i + k];
        r[i] += d * d;
      }
    }
    const double nth = nth_residual(r, fraction);
    Pts xs, ys;
    for (size_t i = 0; i < n; ++i)
      if (r[i] <= nth) {
        xs.insert(xs.end(), &X[3 * i], &X[3 * i] + 3);
        ys.insert(ys.end(), &Y[3 * i], &Y[3 * i] + 3);
      }
    fraction = 0.5;
    double mx[3], my[3];
    barycentre(xs, mx);
    barycentre(ys, my);
    for (size_t i = 0; i < xs.size() / 3; ++i)
      for (int k = 0; k < 3; ++k) {
        xs[3 * i + k] -= mx[k];
        ys[3 * i + k] -= my[k];
      }
    std::memcpy(neu, last, sizeof(double) * np);
    if (T.kind == kAffine) {
      const M3 oxy = covariance(ys, xs), oxx = covariance(xs, xs);
      M3 rot = oxy;
      if (std::fabs(m3_det(oxx)) < 1e-4) {
        if (g_verbosity) std::printf("ERROR: Problem inverting outer_xx as its determinant is low\n");
      } else {
        rot = m3_mul(rot, m3_inverse(oxx));
      }
      double a[3], b[3], ra[3];
      for (int k = 0; k < 3; ++k) {
        a[k] = my[k] - rc[k];
        b[k] = mx[k] - rc[k];
      }
      m3_vec(rot, b, ra);
      for (int i = 0; i < 9; ++i) neu[i] = rot.a[i / 3][i % 3];
      for (int k = 0; k < 3; ++k) neu[9 + k] = a[k] - ra[k];
      last_delta = delta;
      delta = params_delta(last, neu, np);
    } else {
      const Quat q = rigid_quaternion(xs, ys);
      const M3 rot = quat_rotation(q);
      double a[3], b[3], ra[3], ang[3];
      for (int k = 0; k < 3; ++k) {
        a[k] = my[k] - rc[k];
        b[k] = mx[k] - rc[k];
      }
      m3_vec(rot, b, ra);
      quat_euler(q, ang);
      for (int k = 0; k < 3; ++k) {
        neu[k] = ang[k];
        neu[3 + k] = a[k] - ra[k];
      }
      last_delta = delta;
      delta = params_delta(neu, last, np);
    }
    for (int k = 0; k < 3; ++k) rc[k] = mx[k];
    fin.set_center(rc);
    fin.set_params(neu);
  }
  T.compose(fin, true);
  return iter + 1;
}

// Symmetric LTS (itkSymMirorrRegistrationMethod.txx:729-1442). Updates the
// half-space transforms and T = t12F o t12B^-1. Returns inner iterations.
int lts_symmetric(Xf& T, Xf& t12F, Xf& t12B, const Pts& X, const Pts& Y, double tol) {
  Xf fF, fB;
  fF.kind = fB.kind = T.kind;
  double rc[3] = {T.c[0], T.c[1], T.c[2]};
  fF.set_identity();
  fF.set_center(rc);
  fB.set_identity();
  fB.set_center(rc);
  const int np = fF.nparams();
  double last[12], neu[12], neuF[12], neuB[12];
  {
    Xf id;
    id.kind = T.kind;
    id.set_identity();
    id.set_center(rc);
    id.get_params(last);
  }
  double fraction = 1.0, last_delta = 0.0, delta = 1.0;
  int iter = -1;
  const size_t n = X.size() / 3;
  const double zero[3] = {0, 0, 0};
  while (delta > tol && iter < 100 && std::fabs(delta - last_delta) > tol) {
    ++iter;
    if (n == 0) break;
    std::vector<double> r(n, 0.0);
    for (size_t i = 0; i < n; ++i) {
      double pf[3], pb[3];
      fF.apply(&X[3 * i], pf);
      fB.apply(&Y[3 * i], pb);
      for (int k = 0; k < 3; ++k) {
        const double d = pf[k] - pb[k];
        r[i] += d * d;
      }
    }
    const double nth = nth_residual(r, fraction);
    Pts xs, ys;
    for (size_t i = 0; i < n; ++i)
      if (r[i] <= nth) {
        xs.insert(xs.end(), &X[3 * i], &X[3 * i] + 3);
        ys.insert(ys.end(), &Y[3 * i], &Y[3 * i] + 3);
      }
    fraction = 0.5;
    double mx[3], my[3];
    barycentre(xs, mx);
    barycentre(ys, my);
    for (size_t i = 0; i < xs.size() / 3; ++i)
      for (int k = 0; k < 3; ++k) {
        xs[3 * i + k] -= mx[k];
        ys[3 * i + k] -= my[k];
      }
    std::memcpy(neu, last, sizeof(double) * np);
    std::memcpy(neuF, last, sizeof(double) * np);
    std::memcpy(neuB, last, sizeof(double) * np);
    double tr[3];
    for (int k = 0; k < 3; ++k) tr[k] = my[k] - mx[k];
    if (T.kind == kAffine) {
      M3 oxy = covariance(ys, xs), oxx = covariance(xs, xs);
      M3 rF = oxy;
      if (std::fabs(m3_det(oxx)) >= 1e-4) rF = m3_mul(oxy, m3_inverse(oxx));
      oxy = covariance(xs, ys);
      oxx = covariance(ys, ys);
      M3 rB = oxy;
      if (std::fabs(m3_det(oxx)) >= 1e-4) rB = m3_mul(oxy, m3_inverse(oxx));
      for (int i = 0; i < 9; ++i) {
        neu[i] = neuF[i] = rF.a[i / 3][i % 3];
        neuB[i] = rB.a[i / 3][i % 3];
      }
      for (int k = 0; k < 3; ++k) {
        neu[9 + k] = neuF[9 + k] = tr[k];
        neuB[9 + k] = -tr[k];
      }
      fF.set_center(mx);
      fF.set_params(neuF);
      fB.set_center(my);
      fB.set_params(neuB);
      // average forward and inverse-backward matrices
      M3 K = fF.M, L = m3_inverse(fB.M), Mi = m3_inverse(fF.M), N = fB.M;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          K.a[i][j] = (K.a[i][j] + L.a[i][j]) / 2.0;
          Mi.a[i][j] = (Mi.a[i][j] + N.a[i][j]) / 2.0;
        }
      fF.set_matrix(K);
      fB.set_matrix(Mi);
    } else {
      const Quat q = rigid_quaternion(xs, ys);
      double eu[3], euF[3], euB[3], ax[3];
      quat_euler(q, eu);
      quat_axis(q, ax);
      const double angle = quat_angle(q);
      quat_euler(quat_from_axis_angle(ax, angle), euF);
      quat_euler(quat_from_axis_angle(ax, -angle), euB);
      for (int k = 0; k < 3; ++k) {
        neu[k] = eu[k];
        neuF[k] = euF[k];
        neuB[k] = euB[k];
        neu[3 + k] = neuF[3 + k] = tr[k];
        neuB[3 + k] = -tr[k];
      }
      fF.set_center(mx);
      fF.set_params(neuF);
      fB.set_center(my);
      fB.set_params(neuB);
      M3 K = fF.M;
      const M3 L = m3_inverse(fB.M);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) K.a[i][j] = (K.a[i][j] + L.a[i][j]) / 2.0;
      K = orthogonalize(K);
      fF.set_matrix(K);
      fB.set_matrix(m3_transpose(K));
    }
    // Re-centre about the origin.
    {
      const double oF[3] = {fF.off[0], fF.off[1], fF.off[2]};
      const double oB[3] = {fB.off[0], fB.off[1], fB.off[2]};
      fF.set_center(zero);
      fB.set_center(zero);
      fF.set_translation(oF);
      fB.set_translation(oB);
    }
    const Xf iF = fF.inverse(), iB = fB.inverse();
    const bool rigid = T.kind == kEuler3D;
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
    if (rigid) {
      K = orthogonalize(K);
      K2 = orthogonalize(K2);
    }
    const M4 m12 = m4_sqrt_db(to4(K, tF), 1e-14, 50);
    const M4 im12 = m4_sqrt_db(to4(K2, tB), 1e-14, 50);
    M3 A = top3(m12), B = top3(im12);
    try {
      fF.set_matrix(A);
      fB.set_matrix(B);
    } catch (const std::exception&) {
      // rarely the square roots are not orthogonal enough for Euler3D
      A = orthogonalize(A);
      B = orthogonalize(B);
      fF.set_matrix(A);
      fB.set_matrix(B);
    }
    const double t1[3] = {m12.a[0][3], m12.a[1][3], m12.a[2][3]};
    const double t2[3] = {im12.a[0][3], im12.a[1][3], im12.a[2][3]};
    fF.set_translation(t1);
    fB.set_translation(t2);
    last_delta = delta;
    delta = rigid ? params_delta(neu, last, np) : params_delta(last, neu, np);
    std::memcpy(last, neu, sizeof(double) * np);
  }
  t12F.compose(fF, true);
  t12B.compose(fB, true);
  const Xf iB = t12B.inverse();
  T.set_identity();
  T.compose(iB, true);
  T.compose(t12F, false);
  return iter + 1;
}

}  // namespace orc

extern "C" {

// LTS entry points for tests: X, Y are n x 3 point lists; params/center
// describe the transform composed with the fit. kind 0 = Euler3D, 1 = affine.
int orc_lts_classic(int kind, const double* X, const double* Y, int64_t n, double* params,
                    double* center) {
  Xf T;
  T.kind = (Kind)kind;
  T.set_center(center);
  T.set_params(params);
  Pts x(X, X + 3 * n), y(Y, Y + 3 * n);
  const int it = lts_classic(T, x, y);
  T.get_params(params);
  for (int k = 0; k < 3; ++k) center[k] = T.c[k];
  return it;
}

// Symmetric LTS with identity half-space transforms: returns the fitted
// forward (t12F) and backward (t12B) half transforms as 3x4 [M | offset].
int orc_lts_symmetric(int kind, const double* X, const double* Y, int64_t n, double* fwd34,
                      double* bwd34, double* params) {
  Xf T, F, B;
  T.kind = F.kind = B.kind = (Kind)kind;
  Pts x(X, X + 3 * n), y(Y, Y + 3 * n);
  const int it = lts_symmetric(T, F, B, x, y);
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) {
      fwd34[4 * r + c] = F.M.a[r][c];
      bwd34[4 * r + c] = B.M.a[r][c];
    }
    fwd34[4 * r + 3] = F.off[r];
    bwd34[4 * r + 3] = B.off[r];
  }
  T.get_params(params);
  return it;
}

// Denman-Beavers square root of a 4x4 (row-major), vnl_utils.h:46-95.
void orc_sqrtm4(const double* m, double* out, double eps, int max_iter) {
  M4 a;
  for (int i = 0; i < 16; ++i) a.a[i / 4][i % 4] = m[i];
  const M4 r = m4_sqrt_db(a, eps, max_iter);
  for (int i = 0; i < 16; ++i) out[i] = r.a[i / 4][i % 4];
}

void orc_orthogonalize(const double* m, double* out) {
  M3 a;
  for (int i = 0; i < 9; ++i) a.a[i / 3][i % 3] = m[i];
  const M3 r = orthogonalize(a);
  for (int i = 0; i < 9; ++i) out[i] = r.a[i / 3][i % 3];
}

// Quaternion of the rigid fit of centred point sets (CalculateRigidQuaternion).
void orc_rigid_quaternion(const double* X, const double* Y, int64_t n, double q[4]) {
  Pts x(X, X + 3 * n), y(Y, Y + 3 * n);
  const Quat r = rigid_quaternion(x, y);
  q[0] = r.x;
  q[1] = r.y;
  q[2] = r.z;
  q[3] = r.r;
}

void orc_quat_euler(const double q[4], double ang[3]) {
  quat_euler(Quat{q[0], q[1], q[2], q[3]}, ang);
}

// Euler3D (ZXY) matrix of parameters, and the inverse extraction.
void orc_euler_matrix(const double ang[3], double m[9]) {
  Xf x;
  x.kind = kEuler3D;
  for (int k = 0; k < 3; ++k) x.ang[k] = ang[k];
  x.compute_matrix_from_angles();
  for (int i = 0; i < 9; ++i) m[i] = x.M.a[i / 3][i % 3];
}

}  // extern "C"
