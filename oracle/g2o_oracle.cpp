// FP64 CPU oracle of g2o's per-iteration Levenberg-Marquardt linear-system path
// for bundle adjustment (BlockSolver<BlockSolverTraits<6,3>> with Schur) and for
// 3D pose graphs (6x6 blocks, no Schur).
//
// TEST INFRASTRUCTURE ONLY. This file is the checker that the HIP library is
// compared against, and the CPU baseline that is timed beside it. Nothing in the
// product library links or calls it.
//
// It restates, without Eigen, the arithmetic of these reference files (paths are
// relative to the g2o source tree):
//   g2o/types/slam3d/se3quat.h:97-256            SE3Quat exp / product / map / normalize
//   g2o/types/slam3d/se3_ops.hpp:27-52           skew / project
//   g2o/types/sba/edge_project_xyz.cpp:44-101    EdgeSE3ProjectXYZ error + analytic Jacobians
//   g2o/types/sba/vertex_se3_expmap.cpp:48-51    VertexSE3Expmap::oplusImpl (left exp update)
//   g2o/types/slam3d/vertex_pointxyz.h:47-50     VertexPointXYZ::oplusImpl
//   g2o/core/robust_kernel_impl.cpp:60-74        RobustKernelHuber::robustify
//   g2o/core/base_edge.h:99,156-162              chi2, robustInformation
//   g2o/core/base_fixed_sized_edge.hpp:48-130    constructQuadraticForm
//   g2o/core/block_solver.hpp:331-606            Schur, back-substitution, setLambda
//   g2o/core/sparse_optimizer.cpp:78-130,175-200,457-468  errors, chi2, index map, update
//   g2o/core/optimization_algorithm_levenberg.cpp:60-176  LM control
//   g2o/solvers/pcg/linear_solver_pcg.hpp:87-205 block-Jacobi PCG and its stop rule
//   g2o/solvers/eigen/linear_solver_eigen.h:94-191 (semantics: FP64 Cholesky, false on non-PD)
//   BAL: g2o/examples/bal/bal_example.cpp:70-207 (9-DoF camera, additive
//        updates, the templated error functor with dual-number Jacobians
//        in place of auto_differentiation.h:155-302), BlockSolver<9, 3>
//   pose graph: g2o/types/slam3d/edge_se3.cpp:60-88, vertex_se3.h:99-108,
//               isometry3d_mappings.cpp:80-128, isometry3d_gradients.h:206-279
//
// Conventions of this file: 3x3 / 6x6 / 6x3 blocks are stored row-major.
// A camera pose is 7 doubles [tx ty tz qx qy qz qw] (SE3Quat::toVector order).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <unordered_map>
#include <vector>
#include <cstdio>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using std::vector;

inline double now_s() {
  return std::chrono::duration<double>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// ---------------------------------------------------------------- quaternions
struct Quat {
  double x, y, z, w;
};

// Hamilton product a*b (the product Eigen's Quaternion::operator* computes).
inline Quat qmul(const Quat& a, const Quat& b) {
  return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
          a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z,
          a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x,
          a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}

inline void qnormalize(Quat& q) {
  double n = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
  q.x /= n; q.y /= n; q.z /= n; q.w /= n;
}

// Rotation matrix (row-major) of a unit quaternion.
inline void q2R(const Quat& q, double R[9]) {
  double tx = 2 * q.x, ty = 2 * q.y, tz = 2 * q.z;
  double twx = tx * q.w, twy = ty * q.w, twz = tz * q.w;
  double txx = tx * q.x, txy = ty * q.x, txz = tz * q.x;
  double tyy = ty * q.y, tyz = tz * q.y, tzz = tz * q.z;
  R[0] = 1 - (tyy + tzz); R[1] = txy - twz;       R[2] = txz + twy;
  R[3] = txy + twz;       R[4] = 1 - (txx + tzz); R[5] = tyz - twx;
  R[6] = txz - twy;       R[7] = tyz + twx;       R[8] = 1 - (txx + tyy);
}

// Quaternion of a rotation matrix (row-major). Shepperd's method (J. Guidance
// and Control 1(3), 1978): take the largest of w^2, x^2, y^2, z^2 from the trace
// and the diagonal, then the other three components from the off-diagonals.
inline Quat R2q(const double R[9]) {
  Quat q;
  double tr = R[0] + R[4] + R[8];
  if (tr > 0) {
    double s = std::sqrt(tr + 1.0);
    q.w = 0.5 * s;
    s = 0.5 / s;
    q.x = (R[7] - R[5]) * s;
    q.y = (R[2] - R[6]) * s;
    q.z = (R[3] - R[1]) * s;
  } else {
    int i = 0;
    if (R[4] > R[0]) i = 1;
    if (R[8] > R[3 * i + i]) i = 2;
    int j = (i + 1) % 3, k = (j + 1) % 3;
    double s = std::sqrt(R[3 * i + i] - R[3 * j + j] - R[3 * k + k] + 1.0);
    double v[3];
    v[i] = 0.5 * s;
    s = 0.5 / s;
    q.w = (R[3 * k + j] - R[3 * j + k]) * s;
    v[j] = (R[3 * j + i] + R[3 * i + j]) * s;
    v[k] = (R[3 * k + i] + R[3 * i + k]) * s;
    q.x = v[0]; q.y = v[1]; q.z = v[2];
  }
  return q;
}

inline void qrot(const Quat& q, const double v[3], double out[3]) {
  double R[9];
  q2R(q, R);
  for (int r = 0; r < 3; ++r) out[r] = R[3 * r] * v[0] + R[3 * r + 1] * v[1] + R[3 * r + 2] * v[2];
}

// ---------------------------------------------------------------- SE3Quat
struct SE3 {
  Quat q{0, 0, 0, 1};
  double t[3]{0, 0, 0};
};

// se3quat.h:251-256: make w >= 0, then normalize.
inline void normalize_rotation(SE3& a) {
  if (a.q.w < 0) { a.q.x = -a.q.x; a.q.y = -a.q.y; a.q.z = -a.q.z; a.q.w = -a.q.w; }
  qnormalize(a.q);
}

// se3quat.h:97-103 (operator*)
inline SE3 se3_mul(const SE3& a, const SE3& b) {
  SE3 r;
  double rt[3];
  qrot(a.q, b.t, rt);
  for (int i = 0; i < 3; ++i) r.t[i] = a.t[i] + rt[i];
  r.q = qmul(a.q, b.q);
  normalize_rotation(r);
  return r;
}

inline void skew(const double v[3], double m[9]) {
  m[0] = 0;     m[1] = -v[2]; m[2] = v[1];
  m[3] = v[2];  m[4] = 0;     m[5] = -v[0];
  m[6] = -v[1]; m[7] = v[0];  m[8] = 0;
}

inline void mat3mul(const double a[9], const double b[9], double c[9]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      c[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
}

// se3quat.h:202-230 (SE3Quat::exp): update = [omega; upsilon].
inline SE3 se3_exp(const double upd[6]) {
  const double* w = upd;
  const double* u = upd + 3;
  double theta = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  double W[9], W2[9], R[9], V[9];
  skew(w, W);
  mat3mul(W, W, W2);
  double a, b, c, d;  // R = I + a W + b W^2, V = I + c W + d W^2
  if (theta < 0.00001) {
    a = 1.0; b = 0.5; c = 0.5; d = 1.0 / 6.0;
  } else {
    a = std::sin(theta) / theta;
    b = (1 - std::cos(theta)) / (theta * theta);
    c = (1 - std::cos(theta)) / (theta * theta);
    d = (theta - std::sin(theta)) / std::pow(theta, 3);
  }
  for (int i = 0; i < 9; ++i) {
    double I = (i % 4 == 0) ? 1.0 : 0.0;
    R[i] = I + a * W[i] + b * W2[i];
    V[i] = I + c * W[i] + d * W2[i];
  }
  SE3 r;
  r.q = R2q(R);
  for (int i = 0; i < 3; ++i) r.t[i] = V[3 * i] * u[0] + V[3 * i + 1] * u[1] + V[3 * i + 2] * u[2];
  normalize_rotation(r);  // SE3Quat(Quaternion, Vector3) constructor
  return r;
}

// se3quat.h:165-198 (SE3Quat::log)
inline void se3_log(const SE3& a, double out[6]) {
  double R[9];
  q2R(a.q, R);
  double d = 0.5 * (R[0] + R[4] + R[8] - 1);
  double dR[3] = {R[7] - R[5], R[2] - R[6], R[3] - R[1]};
  double omega[3], Om[9], Om2[9], Vinv[9];
  double cc;
  if (std::fabs(d) > 0.99999) {
    for (int i = 0; i < 3; ++i) omega[i] = 0.5 * dR[i];
    cc = 1.0 / 12.0;
  } else {
    double theta = std::acos(d);
    double s = theta / (2 * std::sqrt(1 - d * d));
    for (int i = 0; i < 3; ++i) omega[i] = s * dR[i];
    cc = (1 - theta / (2 * std::tan(theta / 2))) / (theta * theta);
  }
  skew(omega, Om);
  mat3mul(Om, Om, Om2);
  for (int i = 0; i < 9; ++i) Vinv[i] = ((i % 4 == 0) ? 1.0 : 0.0) - 0.5 * Om[i] + cc * Om2[i];
  for (int i = 0; i < 3; ++i) out[i] = omega[i];
  for (int i = 0; i < 3; ++i) out[3 + i] = Vinv[3 * i] * a.t[0] + Vinv[3 * i + 1] * a.t[1] + Vinv[3 * i + 2] * a.t[2];
}

inline SE3 se3_from7(const double* p) {
  SE3 r;
  r.t[0] = p[0]; r.t[1] = p[1]; r.t[2] = p[2];
  r.q = {p[3], p[4], p[5], p[6]};
  normalize_rotation(r);
  return r;
}
inline void se3_to7(const SE3& a, double* p) {
  p[0] = a.t[0]; p[1] = a.t[1]; p[2] = a.t[2];
  p[3] = a.q.x; p[4] = a.q.y; p[5] = a.q.z; p[6] = a.q.w;
}

// ---------------------------------------------------------------- Huber
// robust_kernel_impl.cpp:60-74
inline void huber(double e, double delta, double rho[3]) {
  double dsqr = delta * delta;
  if (e <= dsqr) {
    rho[0] = e; rho[1] = 1.; rho[2] = 0.;
  } else {
    double sqrte = std::sqrt(e);
    rho[0] = 2 * sqrte * delta - dsqr;
    rho[1] = delta / sqrte;
    rho[2] = -0.5 * rho[1] / e;
  }
}

// ------------------------------------------------- EdgeSE3ProjectXYZ (BA edge)
// edge_project_xyz.cpp:44-49 (computeError) and :57-93 (linearizeOplus).
// Jp is d e / d point (2x3 row-major), Jc is d e / d [omega; upsilon] (2x6).
inline void project_edge(const SE3& T, const double R[9], const double in[4], const double X[3],
                         const double z[2], double e[2], double* Jp, double* Jc) {
  double Xc[3];
  for (int i = 0; i < 3; ++i) Xc[i] = R[3 * i] * X[0] + R[3 * i + 1] * X[1] + R[3 * i + 2] * X[2] + T.t[i];
  const double fx = in[0], fy = in[1], cx = in[2], cy = in[3];
  double x = Xc[0], y = Xc[1], zz = Xc[2];
  e[0] = z[0] - (x / zz * fx + cx);
  e[1] = z[1] - (y / zz * fy + cy);
  if (!Jp) return;
  double z2 = zz * zz;
  double tmp[6] = {fx, 0, -x / zz * fx, 0, fy, -y / zz * fy};
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 3; ++c)
      Jp[3 * r + c] = -1. / zz * (tmp[3 * r] * R[c] + tmp[3 * r + 1] * R[3 + c] + tmp[3 * r + 2] * R[6 + c]);
  Jc[0] = x * y / z2 * fx;
  Jc[1] = -(1 + (x * x / z2)) * fx;
  Jc[2] = y / zz * fx;
  Jc[3] = -1. / zz * fx;
  Jc[4] = 0;
  Jc[5] = x / z2 * fx;
  Jc[6] = (1 + y * y / z2) * fy;
  Jc[7] = -x * y / z2 * fy;
  Jc[8] = -x / zz * fy;
  Jc[9] = 0;
  Jc[10] = -1. / zz * fy;
  Jc[11] = y / z2 * fy;
}

// ---------------------------------------------------------------- small dense
// 3x3 inverse by cofactors (the closed form Eigen uses for fixed 3x3).
inline bool inv3(const double m[9], double out[9]) {
  double c0 = m[4] * m[8] - m[5] * m[7];
  double c1 = m[5] * m[6] - m[3] * m[8];
  double c2 = m[3] * m[7] - m[4] * m[6];
  double det = m[0] * c0 + m[1] * c1 + m[2] * c2;
  double id = 1.0 / det;
  out[0] = c0 * id;
  out[1] = (m[2] * m[7] - m[1] * m[8]) * id;
  out[2] = (m[1] * m[5] - m[2] * m[4]) * id;
  out[3] = c1 * id;
  out[4] = (m[0] * m[8] - m[2] * m[6]) * id;
  out[5] = (m[2] * m[3] - m[0] * m[5]) * id;
  out[6] = c2 * id;
  out[7] = (m[1] * m[6] - m[0] * m[7]) * id;
  out[8] = (m[0] * m[4] - m[1] * m[3]) * id;
  return std::isfinite(id);
}

// n x n inverse (row-major, n <= 9) by Gauss-Jordan elimination with
// partial pivoting.
inline bool invN(int n, const double* m, double* out) {
  double a[9 * 18];
  if (n > 9) return false;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < 2 * n; ++j)
      a[i * 2 * n + j] = (j < n) ? m[i * n + j] : ((j - n == i) ? 1.0 : 0.0);
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[r * 2 * n + c]) > std::fabs(a[p * 2 * n + c])) p = r;
    if (a[p * 2 * n + c] == 0.0) return false;
    if (p != c)
      for (int j = 0; j < 2 * n; ++j) std::swap(a[p * 2 * n + j], a[c * 2 * n + j]);
    double piv = 1.0 / a[c * 2 * n + c];
    for (int j = 0; j < 2 * n; ++j) a[c * 2 * n + j] *= piv;
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      double f = a[r * 2 * n + c];
      if (f == 0.0) continue;
      for (int j = 0; j < 2 * n; ++j) a[r * 2 * n + j] -= f * a[c * 2 * n + j];
    }
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) out[i * n + j] = a[i * 2 * n + n + j];
  return true;
}

// Dense LL^T of an SPD matrix (row-major, full storage) and solve. Returns
// false when a pivot is not positive, as LinearSolverEigen::solve does when
// SimplicialLLT reports failure (linear_solver_eigen.h:121-139).
bool chol_solve(int n, const double* A, const double* b, double* x) {
  vector<double> L(static_cast<size_t>(n) * n, 0.0);
  for (int j = 0; j < n; ++j) {
    double s = A[static_cast<size_t>(j) * n + j];
    for (int k = 0; k < j; ++k) s -= L[static_cast<size_t>(j) * n + k] * L[static_cast<size_t>(j) * n + k];
    if (!(s > 0.0) || !std::isfinite(s)) return false;
    double ljj = std::sqrt(s);
    L[static_cast<size_t>(j) * n + j] = ljj;
#pragma omp parallel for if (n > 512)
    for (int i = j + 1; i < n; ++i) {
      double t = A[static_cast<size_t>(i) * n + j];
      const double* li = &L[static_cast<size_t>(i) * n];
      const double* lj = &L[static_cast<size_t>(j) * n];
      for (int k = 0; k < j; ++k) t -= li[k] * lj[k];
      L[static_cast<size_t>(i) * n + j] = t / ljj;
    }
  }
  vector<double> y(n);
  for (int i = 0; i < n; ++i) {
    double s = b[i];
    for (int k = 0; k < i; ++k) s -= L[static_cast<size_t>(i) * n + k] * y[k];
    y[i] = s / L[static_cast<size_t>(i) * n + i];
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = y[i];
    for (int k = i + 1; k < n; ++k) s -= L[static_cast<size_t>(k) * n + i] * x[k];
    x[i] = s / L[static_cast<size_t>(i) * n + i];
  }
  return true;
}

// ------------------------------------------------ block-sparse symmetric matrix
// Upper-triangle block storage of a symmetric matrix with uniform block size
// bd (the layout PCG flattens, linear_solver_pcg.hpp:95-115).
struct BlockSym {
  int bd = 6, nb = 0;
  vector<int> row, col;    // upper blocks (row <= col)
  vector<double> val;      // bd*bd row-major per block
  vector<int> diag;        // block index of (i,i), -1 when absent
};

// linear_solver_pcg.hpp:185-205 (mult): dest = A src over the upper blocks and
// their transposes. Parallel over block rows through a full-row index.
struct FullRows {
  vector<int> ptr, blk;
  vector<char> trans;
};

FullRows make_full_rows(const BlockSym& A) {
  FullRows F;
  F.ptr.assign(A.nb + 1, 0);
  for (size_t k = 0; k < A.row.size(); ++k) {
    F.ptr[A.row[k] + 1]++;
    if (A.row[k] != A.col[k]) F.ptr[A.col[k] + 1]++;
  }
  for (int i = 0; i < A.nb; ++i) F.ptr[i + 1] += F.ptr[i];
  F.blk.resize(F.ptr[A.nb]);
  F.trans.resize(F.ptr[A.nb]);
  vector<int> fill(F.ptr.begin(), F.ptr.end() - 1);
  for (size_t k = 0; k < A.row.size(); ++k) {
    int i = A.row[k], j = A.col[k];
    F.blk[fill[i]] = static_cast<int>(k); F.trans[fill[i]++] = 0;
    if (i != j) { F.blk[fill[j]] = static_cast<int>(k); F.trans[fill[j]++] = 1; }
  }
  return F;
}

void sym_mult(const BlockSym& A, const FullRows& F, const double* src, double* dest) {
  const int bd = A.bd;
#pragma omp parallel for schedule(static) if (A.nb > 256)
  for (int i = 0; i < A.nb; ++i) {
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // bd <= 9
    for (int p = F.ptr[i]; p < F.ptr[i + 1]; ++p) {
      int k = F.blk[p];
      const double* m = &A.val[static_cast<size_t>(k) * bd * bd];
      if (!F.trans[p]) {
        int j = A.col[k];
        for (int r = 0; r < bd; ++r)
          for (int c = 0; c < bd; ++c) acc[r] += m[r * bd + c] * src[j * bd + c];
      } else {
        int j = A.row[k];
        for (int r = 0; r < bd; ++r)
          for (int c = 0; c < bd; ++c) acc[r] += m[c * bd + r] * src[j * bd + c];
      }
    }
    for (int r = 0; r < bd; ++r) dest[i * bd + r] = acc[r];
  }
}

// LinearSolverPCG::solve (linear_solver_pcg.hpp:87-161): block-Jacobi PCG from
// x = 0. `residual` carries LinearSolverPCG::_residual between calls (-1 after
// init()); stop when dn <= max(tol * dn0, residual) (absoluteTolerance = true,
// linear_solver_pcg.h:45-50), maxIter < 0 means n. Always succeeds.
struct PCGState {
  double tol = 1e-6;
  int max_iter = -1;
  double residual = -1.0;
  int last_iters = 0;
};

bool pcg_solve(const BlockSym& A, const double* b, double* x, PCGState& st) {
  const int bd = A.bd, n = A.nb * bd;
  FullRows F = make_full_rows(A);
  vector<double> J(static_cast<size_t>(A.nb) * bd * bd);
  for (int i = 0; i < A.nb; ++i) {
    if (A.diag[i] < 0) return false;
    if (!invN(bd, &A.val[static_cast<size_t>(A.diag[i]) * bd * bd], &J[static_cast<size_t>(i) * bd * bd])) return false;
  }
  auto mdiag = [&](const vector<double>& src, vector<double>& dst) {
    for (int i = 0; i < A.nb; ++i)
      for (int r = 0; r < bd; ++r) {
        double s = 0;
        for (int c = 0; c < bd; ++c) s += J[static_cast<size_t>(i) * bd * bd + r * bd + c] * src[i * bd + c];
        dst[i * bd + r] = s;
      }
  };
  auto dot = [&](const vector<double>& a, const vector<double>& b2) {
    double s = 0;
    for (int i = 0; i < n; ++i) s += a[i] * b2[i];
    return s;
  };
  for (int i = 0; i < n; ++i) x[i] = 0;
  vector<double> r(b, b + n), d(n, 0.0), q(n, 0.0), s(n, 0.0);
  mdiag(r, d);
  double dn = dot(r, d);
  double d0 = st.tol * dn;
  if (st.residual > 0.0 && st.residual > d0) d0 = st.residual;
  int maxIter = st.max_iter < 0 ? n : st.max_iter;
  int it;
  for (it = 0; it < maxIter; ++it) {
    if (dn <= d0) break;
    sym_mult(A, F, d.data(), q.data());
    double a = dn / dot(d, q);
    for (int i = 0; i < n; ++i) { x[i] += a * d[i]; r[i] -= a * q[i]; }
    mdiag(r, s);
    double dold = dn;
    dn = dot(r, s);
    double ba = dn / dold;
    for (int i = 0; i < n; ++i) d[i] = s[i] + ba * d[i];
  }
  st.residual = 0.5 * dn;
  st.last_iters = it;
  return true;
}

bool chol_solve_blocks(const BlockSym& A, const double* b, double* x) {
  const int bd = A.bd, n = A.nb * bd;
  vector<double> D(static_cast<size_t>(n) * n, 0.0);
  for (size_t k = 0; k < A.row.size(); ++k) {
    int i = A.row[k], j = A.col[k];
    const double* m = &A.val[k * bd * bd];
    for (int r = 0; r < bd; ++r)
      for (int c = 0; c < bd; ++c) {
        D[static_cast<size_t>(i * bd + r) * n + j * bd + c] = m[r * bd + c];
        D[static_cast<size_t>(j * bd + c) * n + i * bd + r] = m[r * bd + c];
      }
  }
  return chol_solve(n, D.data(), b, x);
}

// ---------------------------------------------------------------- BA problem
struct BA {
  int C = 0, P = 0, E = 0;
  vector<SE3> cams;
  vector<double> intr;   // 4 per camera
  vector<double> pts;    // 3 per point
  vector<int> ecam, ept;
  vector<double> meas, w;
  vector<double> om;     // 3 per edge: upper triangle [O00 O01 O11] of Omega (empty: I)
  vector<uint8_t> cfix, pfix;
  double huber = 0;

  // index mapping (sparse_optimizer.cpp:175-200; block_solver.hpp:122-137)
  vector<int> camH, ptH;  // free index or -1
  int Cf = 0, Pf = 0;
  // active edges: not all vertices fixed (sparse_optimizer.cpp:209-287)
  vector<int> active;
  // Hpl blocks: one per distinct (free cam, free point) pair; per point column
  // sorted by camera (HplCCS, block_solver.hpp:235)
  vector<int> colPtr;            // Pf+1
  vector<int> colCam;            // free camera index per block
  vector<int> edgeBlk;           // Hpl block per edge or -1
  // Hschur pattern (block_solver.hpp:237-266): per row sorted columns
  vector<int> sPtr, sCol;
  // per-iteration system
  vector<double> Hpp, Hll, Hpl, b;
  vector<double> err;            // 2 per edge
  // state stack (base_vertex.h:111-122)
  vector<vector<SE3>> stackC;
  vector<vector<double>> stackP;
  // solver workspace
  BlockSym S;
  vector<double> bschur, Dinv, x, coeff;
  double lambda = 0;
  double user_lambda = 0;  // "initialLambda" property (<= 0: tau * max diag)
  std::vector<double> trials;  // per LM trial of the last optimize: LM_TRIAL_FIELDS values
  PCGState pcg;
  int solver_kind = 0;  // 0 Cholesky, 1 PCG
  double times[6] = {0, 0, 0, 0, 0, 0};
};

void ba_structure(BA& p) {
  p.camH.assign(p.C, -1);
  p.ptH.assign(p.P, -1);
  p.Cf = p.Pf = 0;
  // active vertices: those with an active edge (sparse_optimizer.cpp:217-285);
  // every edge of a free vertex is active
  vector<char> cu(p.C, 0), pu(p.P, 0);
  for (int k = 0; k < p.E; ++k) { cu[p.ecam[k]] = 1; pu[p.ept[k]] = 1; }
  for (int c = 0; c < p.C; ++c) if (!p.cfix[c] && cu[c]) p.camH[c] = p.Cf++;
  for (int i = 0; i < p.P; ++i) if (!p.pfix[i] && pu[i]) p.ptH[i] = p.Pf++;
  p.active.clear();
  for (int k = 0; k < p.E; ++k)
    if (!(p.cfix[p.ecam[k]] && p.pfix[p.ept[k]])) p.active.push_back(k);
  // Hpl blocks per point column
  vector<vector<int>> colCams(p.Pf);
  for (int k : p.active) {
    int c = p.camH[p.ecam[k]], l = p.ptH[p.ept[k]];
    if (c >= 0 && l >= 0) colCams[l].push_back(c);
  }
  p.colPtr.assign(p.Pf + 1, 0);
  for (int l = 0; l < p.Pf; ++l) {
    auto& v = colCams[l];
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    p.colPtr[l + 1] = p.colPtr[l] + static_cast<int>(v.size());
  }
  p.colCam.resize(p.colPtr[p.Pf]);
  for (int l = 0; l < p.Pf; ++l) std::copy(colCams[l].begin(), colCams[l].end(), p.colCam.begin() + p.colPtr[l]);
  p.edgeBlk.assign(p.E, -1);
  for (int k : p.active) {
    int c = p.camH[p.ecam[k]], l = p.ptH[p.ept[k]];
    if (c < 0 || l < 0) continue;
    auto b0 = p.colCam.begin() + p.colPtr[l], b1 = p.colCam.begin() + p.colPtr[l + 1];
    p.edgeBlk[k] = static_cast<int>(std::lower_bound(b0, b1, c) - p.colCam.begin());
  }
  // Schur pattern: diagonal of Hpp plus co-observing pairs
  vector<vector<int>> rows(p.Cf);
  for (int c = 0; c < p.Cf; ++c) rows[c].push_back(c);
  for (int l = 0; l < p.Pf; ++l)
    for (int a = p.colPtr[l]; a < p.colPtr[l + 1]; ++a)
      for (int bb = a; bb < p.colPtr[l + 1]; ++bb) rows[p.colCam[a]].push_back(p.colCam[bb]);
  p.sPtr.assign(p.Cf + 1, 0);
  for (int c = 0; c < p.Cf; ++c) {
    auto& v = rows[c];
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    p.sPtr[c + 1] = p.sPtr[c] + static_cast<int>(v.size());
  }
  p.sCol.resize(p.sPtr[p.Cf]);
  for (int c = 0; c < p.Cf; ++c) std::copy(rows[c].begin(), rows[c].end(), p.sCol.begin() + p.sPtr[c]);
  p.S.bd = 6;
  p.S.nb = p.Cf;
  p.S.row.resize(p.sCol.size());
  p.S.col = p.sCol;
  p.S.diag.assign(p.Cf, -1);
  for (int c = 0; c < p.Cf; ++c)
    for (int k = p.sPtr[c]; k < p.sPtr[c + 1]; ++k) {
      p.S.row[k] = c;
      if (p.sCol[k] == c) p.S.diag[c] = k;
    }
  p.S.val.assign(p.sCol.size() * 36, 0.0);
  p.err.assign(2 * static_cast<size_t>(p.E), 0.0);
}

// The edge's 2x2 information Omega (base_edge.h:165-180) as a full matrix
// [o00 o01; o01 o11]; identity unless set.
inline void edge_info(const BA& p, int k, double o[4]) {
  if (p.om.empty()) { o[0] = 1; o[1] = 0; o[2] = 0; o[3] = 1; return; }
  const double* v = &p.om[3 * static_cast<size_t>(k)];
  o[0] = v[0]; o[1] = v[1]; o[2] = v[1]; o[3] = v[2];
}
// e^T Omega e (base_edge.h:99, chi2 = error.dot(information() * error))
inline double edge_quad(const BA& p, int k, const double e[2]) {
  double o[4];
  edge_info(p, k, o);
  return e[0] * (o[0] * e[0] + o[1] * e[1]) + e[1] * (o[2] * e[0] + o[3] * e[1]);
}

// computeActiveErrors + activeRobustChi2 (sparse_optimizer.cpp:78-130)
double ba_chi2(BA& p) {
  double chi = 0;
  const int n = static_cast<int>(p.active.size());
#pragma omp parallel for reduction(+ : chi) schedule(static)
  for (int a = 0; a < n; ++a) {
    int k = p.active[a];
    const SE3& T = p.cams[p.ecam[k]];
    double R[9];
    q2R(T.q, R);
    double e[2];
    project_edge(T, R, &p.intr[4 * p.ecam[k]], &p.pts[3 * p.ept[k]], &p.meas[2 * k], e, nullptr, nullptr);
    p.err[2 * k] = e[0];
    p.err[2 * k + 1] = e[1];
    double c2 = p.w[k] * edge_quad(p, k, e);
    if (p.huber > 0) {
      double rho[3];
      huber(c2, p.huber, rho);
      chi += rho[0];
    } else {
      chi += c2;
    }
  }
  return chi;
}

// BlockSolver::buildSystem (block_solver.hpp:492-565) with the per-edge
// linearizeOplus + constructQuadraticForm (base_fixed_sized_edge.hpp:48-130).
// Uses the errors of the last ba_chi2 call, as the reference uses _error.
void ba_build_system(BA& p) {
  p.Hpp.assign(static_cast<size_t>(p.Cf) * 36, 0.0);
  p.Hll.assign(static_cast<size_t>(p.Pf) * 9, 0.0);
  p.Hpl.assign(p.colCam.size() * 18, 0.0);
  p.b.assign(6 * static_cast<size_t>(p.Cf) + 3 * static_cast<size_t>(p.Pf), 0.0);
  const int n = static_cast<int>(p.active.size());
  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
#endif
  // Per-thread camera accumulators stand in for the per-vertex QuadraticFormLock
  // of the reference's OpenMP build (base_edge.h:42-58).
  vector<vector<double>> hppT(nth), bpT(nth), hllT(nth), blT(nth);
#pragma omp parallel
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    auto& hpp = hppT[tid];
    auto& bp = bpT[tid];
    auto& hll = hllT[tid];
    auto& bl = blT[tid];
    hpp.assign(static_cast<size_t>(p.Cf) * 36, 0.0);
    bp.assign(static_cast<size_t>(p.Cf) * 6, 0.0);
    hll.assign(static_cast<size_t>(p.Pf) * 9, 0.0);
    bl.assign(static_cast<size_t>(p.Pf) * 3, 0.0);
#pragma omp for schedule(static)
    for (int a = 0; a < n; ++a) {
      int k = p.active[a];
      int ci = p.camH[p.ecam[k]], li = p.ptH[p.ept[k]];
      const SE3& T = p.cams[p.ecam[k]];
      double R[9];
      q2R(T.q, R);
      double e[2] = {p.err[2 * k], p.err[2 * k + 1]}, etmp[2];
      double Jp[6], Jc[12];
      project_edge(T, R, &p.intr[4 * p.ecam[k]], &p.pts[3 * p.ept[k]], &p.meas[2 * k], etmp, Jp, Jc);
      double O[4];
      edge_info(p, k, O);  // Omega (w I when unset: Omega = I)
      double wr_scale = 1.0;
      if (p.huber > 0) {
        double rho[3];
        huber(p.w[k] * edge_quad(p, k, e), p.huber, rho);
        wr_scale = rho[1];
      }
      // robustInformation: Omega' = rho' w Omega (base_edge.h:156-162)
      double Op[4];
      for (int i = 0; i < 4; ++i) Op[i] = wr_scale * p.w[k] * O[i];
      // omega_r = -Omega' e
      double wr[2] = {-(Op[0] * e[0] + Op[1] * e[1]), -(Op[2] * e[0] + Op[3] * e[1])};
      // A^T Omega' B for 2xN Jacobians A (column i) and B (column j)
      auto q = [&](const double* A, int na, int i, const double* B, int nb, int j) {
        return A[i] * (Op[0] * B[j] + Op[1] * B[nb + j]) + A[na + i] * (Op[2] * B[j] + Op[3] * B[nb + j]);
      };
      if (li >= 0) {
        for (int r = 0; r < 3; ++r) {
          bl[3 * li + r] += Jp[r] * wr[0] + Jp[3 + r] * wr[1];
          for (int c = 0; c < 3; ++c) hll[9 * li + 3 * r + c] += q(Jp, 3, r, Jp, 3, c);
        }
      }
      if (ci >= 0) {
        for (int r = 0; r < 6; ++r) {
          bp[6 * ci + r] += Jc[r] * wr[0] + Jc[6 + r] * wr[1];
          for (int c = 0; c < 6; ++c) hpp[36 * ci + 6 * r + c] += q(Jc, 6, r, Jc, 6, c);
        }
      }
      if (ci >= 0 && li >= 0) {
        // Hpl(cam, pt) += (Jc^T Omega') Jp, written through the transposed map
        // because the point is vertex 0 (block_solver.hpp:210-216).
        double* h = &p.Hpl[static_cast<size_t>(p.edgeBlk[k]) * 18];
        for (int r = 0; r < 6; ++r)
          for (int c = 0; c < 3; ++c) h[3 * r + c] += q(Jc, 6, r, Jp, 3, c);
      }
    }
  }
  for (int t = 0; t < nth; ++t) {
    for (size_t i = 0; i < p.Hpp.size(); ++i) p.Hpp[i] += hppT[t][i];
    for (size_t i = 0; i < p.Hll.size(); ++i) p.Hll[i] += hllT[t][i];
    for (int i = 0; i < 6 * p.Cf; ++i) p.b[i] += bpT[t][i];
    for (int i = 0; i < 3 * p.Pf; ++i) p.b[6 * p.Cf + i] += blT[t][i];
  }
}

// BlockSolver::solve with Schur (block_solver.hpp:331-477). lambda is applied
// to the diagonals as setLambda does (block_solver.hpp:567-590) and undone
// afterwards (restoreDiagonal, :592-606) by working on copies.
bool ba_solve(BA& p) {
  double t0 = now_s();
  const int Cf = p.Cf, Pf = p.Pf;
  const int sp = 6 * Cf;
  // Hschur = Hpp (+ lambda)
  std::fill(p.S.val.begin(), p.S.val.end(), 0.0);
  for (int c = 0; c < Cf; ++c) {
    double* d = &p.S.val[static_cast<size_t>(p.S.diag[c]) * 36];
    for (int i = 0; i < 36; ++i) d[i] = p.Hpp[36 * static_cast<size_t>(c) + i];
    for (int i = 0; i < 6; ++i) d[7 * i] += p.lambda;
  }
  p.Dinv.assign(static_cast<size_t>(Pf) * 9, 0.0);
  p.coeff.assign(sp + 3 * static_cast<size_t>(Pf), 0.0);
  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
#endif
  vector<vector<double>> Sth(nth), cth(nth);
  bool okD = true;
#pragma omp parallel reduction(&& : okD)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    auto& Sl = Sth[tid];
    auto& cl = cth[tid];
    Sl.assign(p.S.val.size(), 0.0);
    cl.assign(sp, 0.0);
#pragma omp for schedule(dynamic, 64)
    for (int l = 0; l < Pf; ++l) {
      double D[9];
      for (int i = 0; i < 9; ++i) D[i] = p.Hll[9 * static_cast<size_t>(l) + i];
      for (int i = 0; i < 3; ++i) D[4 * i] += p.lambda;
      double* Di = &p.Dinv[9 * static_cast<size_t>(l)];
      okD = inv3(D, Di) && okD;
      const double* bl = &p.b[sp + 3 * static_cast<size_t>(l)];
      double db[3];
      for (int r = 0; r < 3; ++r) db[r] = Di[3 * r] * bl[0] + Di[3 * r + 1] * bl[1] + Di[3 * r + 2] * bl[2];
      for (int a = p.colPtr[l]; a < p.colPtr[l + 1]; ++a) {
        int i1 = p.colCam[a];
        const double* Bi = &p.Hpl[static_cast<size_t>(a) * 18];
        double BD[18];
        for (int r = 0; r < 6; ++r)
          for (int c = 0; c < 3; ++c)
            BD[3 * r + c] = Bi[3 * r] * Di[c] + Bi[3 * r + 1] * Di[3 + c] + Bi[3 * r + 2] * Di[6 + c];
        for (int r = 0; r < 6; ++r) cl[6 * i1 + r] += Bi[3 * r] * db[0] + Bi[3 * r + 1] * db[1] + Bi[3 * r + 2] * db[2];
        int sk = p.sPtr[i1];
        for (int bb = a; bb < p.colPtr[l + 1]; ++bb) {
          int i2 = p.colCam[bb];
          while (p.sCol[sk] < i2) ++sk;
          const double* Bj = &p.Hpl[static_cast<size_t>(bb) * 18];
          double* h = &Sl[static_cast<size_t>(sk) * 36];
          for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c)
              h[6 * r + c] -= BD[3 * r] * Bj[3 * c] + BD[3 * r + 1] * Bj[3 * c + 1] + BD[3 * r + 2] * Bj[3 * c + 2];
        }
      }
    }
  }
  if (!okD) return false;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < static_cast<long>(p.S.val.size()); ++i)
    for (int t = 0; t < nth; ++t) p.S.val[i] += Sth[t][i];
  p.bschur.assign(sp, 0.0);
  for (int i = 0; i < sp; ++i) {
    double c = 0;
    for (int t = 0; t < nth; ++t) c += cth[t][i];
    p.coeff[i] = c;
    p.bschur[i] = p.b[i] - c;
  }
  double t1 = now_s();
  p.times[2] += t1 - t0;
  p.x.assign(sp + 3 * static_cast<size_t>(Pf), 0.0);
  bool ok = (p.solver_kind == 0) ? chol_solve_blocks(p.S, p.bschur.data(), p.x.data())
                                 : pcg_solve(p.S, p.bschur.data(), p.x.data(), p.pcg);
  double t2 = now_s();
  p.times[3] += t2 - t1;
  if (!ok) return false;
  // back-substitution (block_solver.hpp:452-476): xl = Dinv (bl - Hpl^T xp)
#pragma omp parallel for schedule(static)
  for (int l = 0; l < Pf; ++l) {
    double cl[3];
    for (int r = 0; r < 3; ++r) cl[r] = p.b[sp + 3 * static_cast<size_t>(l) + r];
    for (int a = p.colPtr[l]; a < p.colPtr[l + 1]; ++a) {
      const double* B = &p.Hpl[static_cast<size_t>(a) * 18];
      const double* xp = &p.x[6 * static_cast<size_t>(p.colCam[a])];
      for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 6; ++r) cl[c] -= B[3 * r + c] * xp[r];
    }
    const double* Di = &p.Dinv[9 * static_cast<size_t>(l)];
    for (int r = 0; r < 3; ++r)
      p.x[sp + 3 * static_cast<size_t>(l) + r] = Di[3 * r] * cl[0] + Di[3 * r + 1] * cl[1] + Di[3 * r + 2] * cl[2];
  }
  return true;
}

// SparseOptimizer::update (sparse_optimizer.cpp:457-468)
void ba_update(BA& p, const double* x) {
#pragma omp parallel for schedule(static)
  for (int c = 0; c < p.C; ++c) {
    int h = p.camH[c];
    if (h < 0) continue;
    p.cams[c] = se3_mul(se3_exp(&x[6 * h]), p.cams[c]);  // vertex_se3_expmap.cpp:48-51
  }
  const size_t off = 6 * static_cast<size_t>(p.Cf);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < p.P; ++i) {
    int h = p.ptH[i];
    if (h < 0) continue;
    for (int r = 0; r < 3; ++r) p.pts[3 * static_cast<size_t>(i) + r] += x[off + 3 * static_cast<size_t>(h) + r];
  }
}

void ba_push(BA& p) { p.stackC.push_back(p.cams); p.stackP.push_back(p.pts); }
void ba_pop(BA& p) { p.cams = p.stackC.back(); p.pts = p.stackP.back(); p.stackC.pop_back(); p.stackP.pop_back(); }
void ba_discard(BA& p) { p.stackC.pop_back(); p.stackP.pop_back(); }

// computeLambdaInit (optimization_algorithm_levenberg.cpp:155-167)
double ba_lambda_init(const BA& p) {
  double m = 0;
  for (int c = 0; c < p.Cf; ++c)
    for (int j = 0; j < 6; ++j) m = std::max(m, std::fabs(p.Hpp[36 * static_cast<size_t>(c) + 7 * j]));
  for (int l = 0; l < p.Pf; ++l)
    for (int j = 0; j < 3; ++j) m = std::max(m, std::fabs(p.Hll[9 * static_cast<size_t>(l) + 4 * j]));
  return 1e-5 * m;
}

// computeScale (optimization_algorithm_levenberg.cpp:169-176)
double ba_scale(const BA& p) {
  double s = 0;
  for (size_t j = 0; j < p.x.size(); ++j) s += p.x[j] * (p.lambda * p.x[j] + p.b[j]);
  return s;
}


// ------------------------------------------- exact second derivatives (adjoint)
// Hyper-dual numbers x = a + b e1 + sum_j (c_j e2_j + d_j e1 e2_j) with
// e1^2 = e2_j e2_k = 0: one evaluation of a function gives its value, the
// derivative along direction 1 (b), the gradient (c) and the row of mixed
// second derivatives d^2 f / dv_1 dv_j (d) exactly, without truncation error.
// The energy of one projection edge is differentiated this way (12 variables:
// the left exp-map update [omega; upsilon] of the camera, the point update,
// the measurement and the information weight), independently of the closed
// form the HIP kernels use.
constexpr int HDN = 12;
struct HD {
  double a = 0, b = 0;
  double c[HDN] = {}, d[HDN] = {};
};
inline HD hd_const(double v) { HD r; r.a = v; return r; }
inline HD operator+(const HD& x, const HD& y) {
  HD r; r.a = x.a + y.a; r.b = x.b + y.b;
  for (int j = 0; j < HDN; ++j) { r.c[j] = x.c[j] + y.c[j]; r.d[j] = x.d[j] + y.d[j]; }
  return r;
}
inline HD operator-(const HD& x, const HD& y) {
  HD r; r.a = x.a - y.a; r.b = x.b - y.b;
  for (int j = 0; j < HDN; ++j) { r.c[j] = x.c[j] - y.c[j]; r.d[j] = x.d[j] - y.d[j]; }
  return r;
}
inline HD operator*(const HD& x, const HD& y) {
  HD r; r.a = x.a * y.a; r.b = x.a * y.b + x.b * y.a;
  for (int j = 0; j < HDN; ++j) {
    r.c[j] = x.a * y.c[j] + x.c[j] * y.a;
    r.d[j] = x.a * y.d[j] + x.d[j] * y.a + x.b * y.c[j] + x.c[j] * y.b;
  }
  return r;
}
inline HD operator*(double s, const HD& x) {
  HD r; r.a = s * x.a; r.b = s * x.b;
  for (int j = 0; j < HDN; ++j) { r.c[j] = s * x.c[j]; r.d[j] = s * x.d[j]; }
  return r;
}
// f(x) for a scalar function with value f0, derivatives f1, f2 at x.a
inline HD hd_apply(const HD& x, double f0, double f1, double f2) {
  HD r; r.a = f0; r.b = f1 * x.b;
  for (int j = 0; j < HDN; ++j) { r.c[j] = f1 * x.c[j]; r.d[j] = f1 * x.d[j] + f2 * x.b * x.c[j]; }
  return r;
}
inline HD hd_inv(const HD& x) { return hd_apply(x, 1 / x.a, -1 / (x.a * x.a), 2 / (x.a * x.a * x.a)); }
inline HD hd_sqrt(const HD& x) {
  double s = std::sqrt(x.a);
  return hd_apply(x, s, 0.5 / s, -0.25 / (s * x.a));
}

// 0.5 * rho(w' |z' - proj(exp(omega, upsilon) T (X + dX))|^2) with w' = w + dw,
// z' = z + dz; v = [omega(3) upsilon(3) dX(3) dz(2) dw(1)] evaluated at v = 0.
// SE3Quat::exp (se3quat.h:202-230) enters through R = I + A W + B W^2 and
// V = I + C W + D W^2; at omega = 0 the hyper-dual theta^2 has only an e1 e2
// part, so the series A = 1 - th2/6, B = C = 1/2 - th2/24, D = 1/6 - th2/120
// are exact through second order.
HD edge_energy_hd(const SE3& T, const double X[3], const double z[2], double w, const double in[4], double delta,
                  const HD v[12], const double* info = nullptr) {
  double R0[9];
  q2R(T.q, R0);
  const HD* om = v;
  const HD* up = v + 3;
  HD th2 = om[0] * om[0] + om[1] * om[1] + om[2] * om[2];
  HD A = hd_const(1.0) - (1.0 / 6.0) * th2;
  HD B = hd_const(0.5) - (1.0 / 24.0) * th2;
  HD D = hd_const(1.0 / 6.0) - (1.0 / 120.0) * th2;
  HD W[9], W2[9];
  HD zero = hd_const(0.0);
  W[0] = zero; W[1] = hd_const(0.0) - om[2]; W[2] = om[1];
  W[3] = om[2]; W[4] = zero; W[5] = hd_const(0.0) - om[0];
  W[6] = hd_const(0.0) - om[1]; W[7] = om[0]; W[8] = zero;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) W2[3 * i + j] = W[3 * i] * W[j] + W[3 * i + 1] * W[3 + j] + W[3 * i + 2] * W[6 + j];
  HD Y[3];
  for (int i = 0; i < 3; ++i) {
    Y[i] = hd_const(T.t[i]);
    for (int k = 0; k < 3; ++k) Y[i] = Y[i] + R0[3 * i + k] * (hd_const(X[k]) + v[6 + k]);
  }
  HD P[3];
  for (int i = 0; i < 3; ++i) {
    HD acc = Y[i] + up[i];
    for (int k = 0; k < 3; ++k) acc = acc + (A * W[3 * i + k] + B * W2[3 * i + k]) * Y[k] + (B * W[3 * i + k] + D * W2[3 * i + k]) * up[k];
    P[i] = acc;
  }
  HD iz = hd_inv(P[2]);
  HD e0 = hd_const(z[0]) + v[9] - (in[0] * (P[0] * iz) + hd_const(in[2]));
  HD e1 = hd_const(z[1]) + v[10] - (in[1] * (P[1] * iz) + hd_const(in[3]));
  // e^T Omega e with Omega = [o00 o01; o01 o11] (identity when info is null)
  const double o00 = info ? info[0] : 1.0, o01 = info ? info[1] : 0.0, o11 = info ? info[2] : 1.0;
  HD s = (hd_const(w) + v[11]) * (o00 * (e0 * e0) + (2.0 * o01) * (e0 * e1) + o11 * (e1 * e1));
  HD rho = s;
  if (delta > 0 && s.a > delta * delta) rho = (2 * delta) * hd_sqrt(s) - hd_const(delta * delta);
  return 0.5 * rho;
}

// Gradient g[12] and Hessian H[12*12] (row-major) of edge_energy_hd at v = 0.
void edge_energy_derivs(const SE3& T, const double X[3], const double z[2], double w, const double in[4],
                        double delta, double* E, double* g, double* H, const double* om = nullptr) {
  for (int i = 0; i < HDN; ++i) {
    HD v[12];
    v[i].b = 1.0;
    for (int j = 0; j < HDN; ++j) v[j].c[j] = 1.0;
    HD f = edge_energy_hd(T, X, z, w, in, delta, v, om);
    if (i == 0) {
      if (E) *E = f.a;
      for (int j = 0; j < HDN; ++j) g[j] = f.c[j];
    }
    for (int j = 0; j < HDN; ++j) H[HDN * i + j] = f.d[j];
  }
}

// Exact Hessian of F/2 = sum 0.5 rho_e at the current state, assembled into the
// solver's blocks (Hpp, Hll, Hpl as in ba_build_system). With eg != null the
// per-edge mixed derivatives d^2 E_e / d state d (z, w) are kept (27 per edge:
// rows of the 9 state variables, columns dz0 dz1 dw).
void ba_build_exact(BA& p, vector<double>* mixed) {
  p.Hpp.assign(static_cast<size_t>(p.Cf) * 36, 0.0);
  p.Hll.assign(static_cast<size_t>(p.Pf) * 9, 0.0);
  p.Hpl.assign(p.colCam.size() * 18, 0.0);
  if (mixed) mixed->assign(static_cast<size_t>(p.E) * 27, 0.0);
  const int n = static_cast<int>(p.active.size());
  vector<double> He(static_cast<size_t>(n) * 144);
#pragma omp parallel for schedule(dynamic, 64)
  for (int a = 0; a < n; ++a) {
    int k = p.active[a];
    double g[12];
    edge_energy_derivs(p.cams[p.ecam[k]], &p.pts[3 * p.ept[k]], &p.meas[2 * k], p.w[k], &p.intr[4 * p.ecam[k]],
                       p.huber, nullptr, g, &He[144 * static_cast<size_t>(a)],
                       p.om.empty() ? nullptr : &p.om[3 * static_cast<size_t>(k)]);
  }
  for (int a = 0; a < n; ++a) {
    int k = p.active[a];
    const double* H = &He[144 * static_cast<size_t>(a)];
    int ci = p.camH[p.ecam[k]], li = p.ptH[p.ept[k]];
    if (ci >= 0)
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) p.Hpp[36 * static_cast<size_t>(ci) + 6 * r + c] += H[12 * r + c];
    if (li >= 0)
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) p.Hll[9 * static_cast<size_t>(li) + 3 * r + c] += H[12 * (6 + r) + 6 + c];
    if (ci >= 0 && li >= 0) {
      double* h = &p.Hpl[static_cast<size_t>(p.edgeBlk[k]) * 18];
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 3; ++c) h[3 * r + c] += H[12 * r + 6 + c];
    }
    if (mixed)
      for (int r = 0; r < 9; ++r)
        for (int c = 0; c < 3; ++c) (*mixed)[27 * static_cast<size_t>(k) + 3 * r + c] = H[12 * r + 9 + c];
  }
}


// ---------------------------------------------------------- 3D pose graph
// VertexSE3 / EdgeSE3 (vertex_se3.h:99-108, edge_se3.cpp:60-88,
// isometry3d_mappings.cpp:80-128). A pose is the Isometry3 x_world = R(q) x + t,
// stored [tx ty tz qx qy qz qw]. Increment: X <- X * fromVectorMQT(d), d =
// [dt; dqx dqy dqz], qw = sqrt(1 - |dq|^2) (identity rotation if negative).
// Error: toVectorMQT(Z^-1 Xi^-1 Xj) = [t; vec(q)] with q normalised to w >= 0.
// The Jacobians d e / d [d_i; d_j] at 0 come from dual-number evaluation of
// that map (the c parts of the hyper-dual numbers above), not from a closed
// form.
template <class S>
struct QT {
  S x, y, z, w;
};
inline HD operator*(const HD& x, double s) { return s * x; }
inline HD operator-(const HD& x) { return (-1.0) * x; }
template <class S>
QT<S> qt_mul(const QT<S>& a, const QT<S>& b) {
  return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
          a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
template <class S>
QT<S> qt_conj(const QT<S>& a) { return {-a.x, -a.y, -a.z, a.w}; }
// v' = R(q) v for a unit quaternion
template <class S>
void qt_rot(const QT<S>& q, const S v[3], S out[3]) {
  S c[3] = {q.y * v[2] - q.z * v[1], q.z * v[0] - q.x * v[2], q.x * v[1] - q.y * v[0]};
  S c2[3] = {q.y * c[2] - q.z * c[1], q.z * c[0] - q.x * c[2], q.x * c[1] - q.y * c[0]};
  for (int i = 0; i < 3; ++i) out[i] = v[i] + 2.0 * (q.w * c[i] + c2[i]);
}

struct PG {
  int N = 0, E = 0;
  vector<double> t;     // 3 per pose
  vector<Quat> q;       // per pose
  vector<int> from, to;
  vector<double> mt;    // measurement translation, 3 per edge
  vector<Quat> mq;      // measurement rotation
  vector<double> info;  // 36 per edge
  vector<uint8_t> fixed;
  vector<int> H;        // free index or -1
  int Nf = 0;
  vector<int> active;
  BlockSym S;           // upper blocks: diagonal per free pose, one per free pair
  vector<int> edgeBlk;  // off-diagonal block of each edge or -1
  vector<int> diagBlk;  // per free pose
  vector<double> Hdiag; // undamped diagonal blocks
  vector<double> b, x;
  double lambda = 0;
  PCGState pcg;
  int solver_kind = 1;
  vector<vector<double>> stackT;
  vector<vector<Quat>> stackQ;
  vector<double> trials;  // per LM trial of the last optimize: LM_TRIAL_FIELDS values
};

// error e[6] and, if J != null, J[6*12] = d e / d [d_i(6) d_j(6)] (row-major)
void pg_edge(const PG& p, int k, double e[6], double* J) {
  const int i = p.from[k], j = p.to[k];
  HD v[12];
  if (J)
    for (int a = 0; a < 12; ++a) v[a].c[a] = 1.0;
  auto incr = [&](int vtx, const HD* d, HD tout[3], QT<HD>& qout) {
    const Quat& q0 = p.q[vtx];
    QT<HD> qv{hd_const(q0.x), hd_const(q0.y), hd_const(q0.z), hd_const(q0.w)};
    HD n2 = d[3] * d[3] + d[4] * d[4] + d[5] * d[5];
    HD w = hd_sqrt(hd_const(1.0) - n2);  // n2 = 0 at the evaluation point
    QT<HD> dq{d[3], d[4], d[5], w};
    qout = qt_mul(qv, dq);
    HD dt[3] = {d[0], d[1], d[2]}, rdt[3];
    qt_rot(qv, dt, rdt);
    for (int a = 0; a < 3; ++a) tout[a] = hd_const(p.t[3 * vtx + a]) + rdt[a];
  };
  HD ti[3], tj[3];
  QT<HD> qi, qj;
  incr(i, v, ti, qi);
  incr(j, v + 6, tj, qj);
  const Quat& m = p.mq[k];
  QT<HD> qz{hd_const(m.x), hd_const(m.y), hd_const(m.z), hd_const(m.w)};
  QT<HD> qa = qt_mul(qt_conj(qz), qt_mul(qt_conj(qi), qj));
  HD d[3] = {tj[0] - ti[0], tj[1] - ti[1], tj[2] - ti[2]}, u[3], w3[3], ta[3];
  qt_rot(qt_conj(qi), d, u);
  for (int a = 0; a < 3; ++a) w3[a] = u[a] - hd_const(p.mt[3 * k + a]);
  qt_rot(qt_conj(qz), w3, ta);
  HD nq = hd_sqrt(qa.x * qa.x + qa.y * qa.y + qa.z * qa.z + qa.w * qa.w);
  const double sgn = qa.w.a < 0 ? -1.0 : 1.0;
  HD inv = sgn * hd_inv(nq);
  HD ev[6] = {ta[0], ta[1], ta[2], qa.x * inv, qa.y * inv, qa.z * inv};
  for (int r = 0; r < 6; ++r) {
    e[r] = ev[r].a;
    if (J)
      for (int a = 0; a < 12; ++a) J[12 * r + a] = ev[r].c[a];
  }
}

void pg_structure(PG& p) {
  p.H.assign(p.N, -1);
  p.Nf = 0;
  vector<char> used(p.N, 0);
  for (int k = 0; k < p.E; ++k) used[p.from[k]] = used[p.to[k]] = 1;
  for (int i = 0; i < p.N; ++i)
    if (!p.fixed[i] && used[i]) p.H[i] = p.Nf++;
  p.active.clear();
  for (int k = 0; k < p.E; ++k)
    if (!(p.fixed[p.from[k]] && p.fixed[p.to[k]])) p.active.push_back(k);
  p.S = BlockSym();
  p.S.bd = 6;
  p.S.nb = p.Nf;
  p.S.diag.assign(p.Nf, -1);
  p.diagBlk.assign(p.Nf, -1);
  for (int i = 0; i < p.Nf; ++i) {
    p.S.diag[i] = p.diagBlk[i] = static_cast<int>(p.S.row.size());
    p.S.row.push_back(i);
    p.S.col.push_back(i);
  }
  std::unordered_map<long long, int> pair;
  p.edgeBlk.assign(p.E, -1);
  for (int k : p.active) {
    int a = p.H[p.from[k]], c = p.H[p.to[k]];
    if (a < 0 || c < 0 || a == c) continue;
    int r = std::min(a, c), cc = std::max(a, c);
    long long key = static_cast<long long>(r) * p.Nf + cc;
    auto it = pair.find(key);
    if (it == pair.end()) {
      it = pair.emplace(key, static_cast<int>(p.S.row.size())).first;
      p.S.row.push_back(r);
      p.S.col.push_back(cc);
    }
    p.edgeBlk[k] = it->second;
  }
  p.S.val.assign(p.S.row.size() * 36, 0.0);
}

double pg_chi2(const PG& p) {
  double chi = 0;
  const int n = static_cast<int>(p.active.size());
#pragma omp parallel for reduction(+ : chi) schedule(static)
  for (int a = 0; a < n; ++a) {
    int k = p.active[a];
    double e[6];
    pg_edge(p, k, e, nullptr);
    const double* O = &p.info[36 * static_cast<size_t>(k)];
    double s = 0;
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c) s += e[r] * O[6 * r + c] * e[c];
    chi += s;
  }
  return chi;
}

// BlockSolver::buildSystem without Schur (block_solver.hpp:492-565, 333-343)
void pg_build(PG& p) {
  std::fill(p.S.val.begin(), p.S.val.end(), 0.0);
  p.b.assign(6 * static_cast<size_t>(p.Nf), 0.0);
  for (int k : p.active) {
    double e[6], J[72];
    pg_edge(p, k, e, J);
    const double* O = &p.info[36 * static_cast<size_t>(k)];
    int hv[2] = {p.H[p.from[k]], p.H[p.to[k]]};
    double OJ[2][36], Oe[6];  // Omega J_v (6x6), Omega e
    for (int r = 0; r < 6; ++r) {
      Oe[r] = 0;
      for (int c = 0; c < 6; ++c) Oe[r] += O[6 * r + c] * e[c];
    }
    for (int v = 0; v < 2; ++v)
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) {
          double s = 0;
          for (int m = 0; m < 6; ++m) s += O[6 * r + m] * J[12 * m + 6 * v + c];
          OJ[v][6 * r + c] = s;
        }
    for (int v = 0; v < 2; ++v) {
      if (hv[v] < 0) continue;
      double* bb = &p.b[6 * static_cast<size_t>(hv[v])];
      double* D = &p.S.val[36 * static_cast<size_t>(p.diagBlk[hv[v]])];
      for (int r = 0; r < 6; ++r) {
        for (int m = 0; m < 6; ++m) bb[r] -= J[12 * m + 6 * v + r] * Oe[m];
        for (int c = 0; c < 6; ++c) {
          double s = 0;
          for (int m = 0; m < 6; ++m) s += J[12 * m + 6 * v + r] * OJ[v][6 * m + c];
          D[6 * r + c] += s;
        }
      }
    }
    if (p.edgeBlk[k] >= 0) {
      // block (min, max): J_a^T Omega J_b with a the lower free index
      int va = hv[0] < hv[1] ? 0 : 1, vb = 1 - va;
      double* B = &p.S.val[36 * static_cast<size_t>(p.edgeBlk[k])];
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) {
          double s = 0;
          for (int m = 0; m < 6; ++m) s += J[12 * m + 6 * va + r] * OJ[vb][6 * m + c];
          B[6 * r + c] += s;
        }
    }
  }
  p.Hdiag.resize(36 * static_cast<size_t>(p.Nf));
  for (int i = 0; i < p.Nf; ++i)
    std::copy(&p.S.val[36 * static_cast<size_t>(p.diagBlk[i])], &p.S.val[36 * static_cast<size_t>(p.diagBlk[i])] + 36,
              &p.Hdiag[36 * static_cast<size_t>(i)]);
}

bool pg_solve(PG& p) {
  for (int i = 0; i < p.Nf; ++i) {
    double* D = &p.S.val[36 * static_cast<size_t>(p.diagBlk[i])];
    for (int c = 0; c < 36; ++c) D[c] = p.Hdiag[36 * static_cast<size_t>(i) + c];
    for (int r = 0; r < 6; ++r) D[7 * r] += p.lambda;
  }
  p.x.assign(6 * static_cast<size_t>(p.Nf), 0.0);
  return p.solver_kind == 0 ? chol_solve_blocks(p.S, p.b.data(), p.x.data())
                            : pcg_solve(p.S, p.b.data(), p.x.data(), p.pcg);
}

// VertexSE3::oplusImpl: X <- X * fromVectorMQT(d)
void pg_update(PG& p, const double* x) {
  for (int i = 0; i < p.N; ++i) {
    int h = p.H[i];
    if (h < 0) continue;
    const double* d = &x[6 * static_cast<size_t>(h)];
    double n2 = d[3] * d[3] + d[4] * d[4] + d[5] * d[5];
    Quat dq = n2 > 1 ? Quat{0, 0, 0, 1} : Quat{d[3], d[4], d[5], std::sqrt(1 - n2)};
    double rt[3];
    qrot(p.q[i], d, rt);
    for (int a = 0; a < 3; ++a) p.t[3 * i + a] += rt[a];
    p.q[i] = qmul(p.q[i], dq);
    qnormalize(p.q[i]);
  }
}

// Record of one LM trial (a pass of the do-loop of
// optimization_algorithm_levenberg.cpp:97-144): lambda used, currentChi,
// tempChi, the divisor of rho (computeScale() + 1e-3, or 1 after a failed
// solve), rho, accepted (1/0).
constexpr int LM_TRIAL_FIELDS = 6;
inline void log_trial(std::vector<double>* log, double lambda, double chi, double temp, double scale, double rho,
                      bool accepted) {
  if (!log) return;
  const double rec[LM_TRIAL_FIELDS] = {lambda, chi, temp, scale, rho, accepted ? 1.0 : 0.0};
  log->insert(log->end(), rec, rec + LM_TRIAL_FIELDS);
}

// OptimizationAlgorithmLevenberg::solve (optimization_algorithm_levenberg.cpp:
// 60-153) over a problem P providing chi2(), build() (linearise at the
// state), lambda_init(), solve(lambda, &pcg_iters) -> ok, update(), scale()
// (computeScale of the last solve), push(), pop(), discard() and times[6].
// trace[6 * it + (0..5)] = chi2 before, chi2 after, lambda, trials, PCG
// iterations, status (1 = Terminate). Returns the iterations run.
template <class P>
int lm_optimize(P& p, int iters, double* trace, double* times, std::vector<double>* log = nullptr) {
  if (log) log->clear();
  for (int k = 0; k < 6; ++k) p.times[k] = 0;
  double tall = now_s();
  double lambda = 0, ni = 2;
  int it;
  for (it = 0; it < iters; ++it) {
    double t0 = now_s();
    double currentChi = p.chi2();
    double t1 = now_s();
    p.times[0] += t1 - t0;
    p.build();
    p.times[1] += now_s() - t1;
    if (it == 0) { lambda = p.lambda_init(); ni = 2; }
    double chi_before = currentChi;
    double rho = 0;
    int q = 0, pcgsum = 0;
    do {
      p.push();
      int pcg_it = 0;
      bool ok = p.solve(lambda, &pcg_it);
      pcgsum += pcg_it;
      double tu = now_s();
      // A failed solve is rejected below and the state popped, so the update
      // the reference still applies (levenberg.cpp:112) has no effect.
      if (ok) p.update();
      p.times[4] += now_s() - tu;
      double tc = now_s();
      double tempChi = p.chi2();
      p.times[0] += now_s() - tc;
      if (!ok) tempChi = std::numeric_limits<double>::max();
      rho = currentChi - tempChi;
      double scale = ok ? p.scale(lambda) + 1e-3 : 1;
      rho /= scale;
      log_trial(log, lambda, currentChi, tempChi, scale, rho, rho > 0 && std::isfinite(tempChi) && ok);
      if (rho > 0 && std::isfinite(tempChi) && ok) {
        double alpha = 1. - std::pow((2 * rho - 1), 3);
        alpha = std::min(alpha, 2. / 3.);
        double sf = std::max(1. / 3., alpha);
        lambda *= sf;
        ni = 2;
        currentChi = tempChi;
        p.discard();
      } else {
        lambda *= ni;
        ni *= 2;
        p.pop();
        if (!std::isfinite(lambda)) break;
      }
      q++;
    } while (rho < 0 && q < 10);
    int status = (q == 10 || rho == 0 || !std::isfinite(lambda)) ? 1 : 0;
    if (trace) {
      trace[6 * it + 0] = chi_before;
      trace[6 * it + 1] = currentChi;
      trace[6 * it + 2] = lambda;
      trace[6 * it + 3] = q;
      trace[6 * it + 4] = pcgsum;
      trace[6 * it + 5] = status;
    }
    if (status) { ++it; break; }
  }
  p.times[5] = now_s() - tall;
  if (times) std::copy(p.times, p.times + 6, times);
  return it;
}

// LM adapter of the pose graph
struct PGLM {
  PG& p;
  double times[6] = {0, 0, 0, 0, 0, 0};
  double chi2() { return pg_chi2(p); }
  void build() { pg_build(p); }
  double lambda_init() {
    double m = 0;
    for (int i = 0; i < p.Nf; ++i)
      for (int r = 0; r < 6; ++r) m = std::max(m, std::fabs(p.Hdiag[36 * static_cast<size_t>(i) + 7 * r]));
    return 1e-5 * m;
  }
  bool solve(double lambda, int* pcg_it) {
    p.lambda = lambda;
    double t = now_s();
    bool ok = pg_solve(p);
    times[3] += now_s() - t;
    *pcg_it = p.solver_kind == 1 ? p.pcg.last_iters : 0;
    return ok;
  }
  void update() { pg_update(p, p.x.data()); }
  double scale(double lambda) {
    double s = 0;
    for (size_t j = 0; j < p.x.size(); ++j) s += p.x[j] * (lambda * p.x[j] + p.b[j]);
    return s;
  }
  void push() { p.stackT.push_back(p.t); p.stackQ.push_back(p.q); }
  void pop() { p.t = p.stackT.back(); p.q = p.stackQ.back(); discard(); }
  void discard() { p.stackT.pop_back(); p.stackQ.pop_back(); }
};


// ------------------------------------------------------------- BAL (9-DoF)
// g2o/examples/bal/bal_example.cpp:70-207: VertexCameraBAL holds 9 values
// [w (angle-axis), t, f, k1, k2] with oplus = plain addition (:86-89),
// VertexPointBAL a 3-vector (+), EdgeObservationBAL the error
// prediction - measurement of the templated functor (:154-204), whose
// Jacobians the reference takes by automatic differentiation
// (G2O_MAKE_AUTO_AD_FUNCTIONS, g2o/core/auto_differentiation.h:155-302).
// Here the same functor is evaluated on forward-mode dual numbers carrying
// the derivatives along the 12 variables (9 camera + 3 point): exact
// derivatives of the same arithmetic, independent of the closed form the HIP
// kernels use. Solved by BlockSolver<BlockSolverTraits<9, 3>> (:227-250):
// the Schur complement on 9x9 camera blocks, Cholesky or block-Jacobi PCG.

template <int N>
struct Jet {
  double a = 0;
  double v[N] = {};
};
template <int N> inline Jet<N> operator+(const Jet<N>& x, const Jet<N>& y) {
  Jet<N> r; r.a = x.a + y.a; for (int i = 0; i < N; ++i) r.v[i] = x.v[i] + y.v[i]; return r;
}
template <int N> inline Jet<N> operator-(const Jet<N>& x, const Jet<N>& y) {
  Jet<N> r; r.a = x.a - y.a; for (int i = 0; i < N; ++i) r.v[i] = x.v[i] - y.v[i]; return r;
}
template <int N> inline Jet<N> operator-(const Jet<N>& x) {
  Jet<N> r; r.a = -x.a; for (int i = 0; i < N; ++i) r.v[i] = -x.v[i]; return r;
}
template <int N> inline Jet<N> operator*(const Jet<N>& x, const Jet<N>& y) {
  Jet<N> r; r.a = x.a * y.a; for (int i = 0; i < N; ++i) r.v[i] = x.a * y.v[i] + x.v[i] * y.a; return r;
}
template <int N> inline Jet<N> operator/(const Jet<N>& x, const Jet<N>& y) {
  Jet<N> r; r.a = x.a / y.a;
  for (int i = 0; i < N; ++i) r.v[i] = (x.v[i] - r.a * y.v[i]) / y.a;
  return r;
}
template <int N> inline Jet<N> jsqrt(const Jet<N>& x) {
  Jet<N> r; r.a = std::sqrt(x.a); for (int i = 0; i < N; ++i) r.v[i] = x.v[i] / (2 * r.a); return r;
}
template <int N> inline Jet<N> jsin(const Jet<N>& x) {
  Jet<N> r; r.a = std::sin(x.a); const double c = std::cos(x.a); for (int i = 0; i < N; ++i) r.v[i] = c * x.v[i];
  return r;
}
template <int N> inline Jet<N> jcos(const Jet<N>& x) {
  Jet<N> r; r.a = std::cos(x.a); const double s = -std::sin(x.a); for (int i = 0; i < N; ++i) r.v[i] = s * x.v[i];
  return r;
}
template <int N> inline Jet<N> operator+(double s, const Jet<N>& x) { Jet<N> r = x; r.a += s; return r; }
template <int N> inline Jet<N> operator-(double s, const Jet<N>& x) { Jet<N> r = -x; r.a += s; return r; }
template <int N> inline Jet<N> operator-(const Jet<N>& x, double s) { Jet<N> r = x; r.a -= s; return r; }
inline double jsqrt(double x) { return std::sqrt(x); }
inline double jsin(double x) { return std::sin(x); }
inline double jcos(double x) { return std::cos(x); }
inline double jval(double x) { return x; }
template <int N> inline double jval(const Jet<N>& x) { return x.a; }

// EdgeObservationBAL::operator() (bal_example.cpp:154-204): rotate by
// Rodrigues' formula (first-order expansion at theta = 0), translate, divide
// by -z, scale by f (1 + k1 r^2 + k2 r^4); e = prediction - z.
template <typename T>
void bal_error(const T* cam, const T* X, const double z[2], T e[2]) {
  T p[3];
  const T theta = jsqrt(cam[0] * cam[0] + cam[1] * cam[1] + cam[2] * cam[2]);
  if (jval(theta) > 0) {
    const T v[3] = {cam[0] / theta, cam[1] / theta, cam[2] / theta};
    const T cth = jcos(theta), sth = jsin(theta);
    const T vxp[3] = {v[1] * X[2] - v[2] * X[1], v[2] * X[0] - v[0] * X[2], v[0] * X[1] - v[1] * X[0]};
    const T vdp = v[0] * X[0] + v[1] * X[1] + v[2] * X[2];
    const T omc = 1.0 - cth;
    for (int i = 0; i < 3; ++i) p[i] = X[i] * cth + vxp[i] * sth + v[i] * vdp * omc;
  } else {
    const T wxp[3] = {cam[1] * X[2] - cam[2] * X[1], cam[2] * X[0] - cam[0] * X[2], cam[0] * X[1] - cam[1] * X[0]};
    for (int i = 0; i < 3; ++i) p[i] = X[i] + wxp[i];
  }
  for (int i = 0; i < 3; ++i) p[i] = p[i] + cam[3 + i];
  const T px = -p[0] / p[2], py = -p[1] / p[2];
  const T r2 = px * px + py * py;
  const T rp = 1.0 + cam[7] * r2 + cam[8] * r2 * r2;
  e[0] = cam[6] * rp * px - z[0];
  e[1] = cam[6] * rp * py - z[1];
}

// e[2], Jc[2x9] = de/dcam, Jp[2x3] = de/dX (row-major), by dual numbers.
void bal_edge(const double* cam, const double* X, const double* z, double* e, double* Jc, double* Jp) {
  if (!Jc && !Jp) {
    bal_error<double>(cam, X, z, e);
    return;
  }
  using J12 = Jet<12>;
  J12 cj[9], xj[3], ej[2];
  for (int i = 0; i < 9; ++i) { cj[i].a = cam[i]; cj[i].v[i] = 1.0; }
  for (int i = 0; i < 3; ++i) { xj[i].a = X[i]; xj[i].v[9 + i] = 1.0; }
  bal_error<J12>(cj, xj, z, ej);
  for (int r = 0; r < 2; ++r) {
    e[r] = ej[r].a;
    if (Jc) for (int c = 0; c < 9; ++c) Jc[9 * r + c] = ej[r].v[c];
    if (Jp) for (int c = 0; c < 3; ++c) Jp[3 * r + c] = ej[r].v[9 + c];
  }
}

struct BAL {
  int C = 0, P = 0, E = 0;
  vector<double> cams;   // 9 per camera
  vector<double> pts;    // 3 per point
  vector<int> ecam, ept;
  vector<double> meas, w;
  vector<uint8_t> cfix, pfix;
  double huber = 0;
  vector<int> camH, ptH;
  int Cf = 0, Pf = 0;
  vector<int> active;
  vector<int> colPtr, colCam, edgeBlk;  // Hpl blocks per point column (distinct cameras)
  vector<int> ptPtr, ptEdge;            // active edges grouped by point (P + 1 CSR), edge order
  vector<int> sPtr, sCol;               // Hschur pattern
  vector<double> Hpp, Hll, Hpl, b, err;
  vector<vector<double>> stackC, stackP;
  BlockSym S;
  vector<double> bschur, Dinv, x;
  double lambda = 0;
  vector<double> trials;
  PCGState pcg;
  int solver_kind = 0;
  double times[6] = {0, 0, 0, 0, 0, 0};
};

// BlockSolver<9, 3>::buildStructure (block_solver.hpp:110-269) and the index
// mapping (sparse_optimizer.cpp:175-200), as for the 6-DoF problem.
void bal_structure(BAL& p) {
  p.camH.assign(p.C, -1);
  p.ptH.assign(p.P, -1);
  p.Cf = p.Pf = 0;
  vector<char> cu(p.C, 0), pu(p.P, 0);
  for (int k = 0; k < p.E; ++k) { cu[p.ecam[k]] = 1; pu[p.ept[k]] = 1; }
  for (int c = 0; c < p.C; ++c) if (!p.cfix[c] && cu[c]) p.camH[c] = p.Cf++;
  for (int i = 0; i < p.P; ++i) if (!p.pfix[i] && pu[i]) p.ptH[i] = p.Pf++;
  p.active.clear();
  for (int k = 0; k < p.E; ++k)
    if (!(p.cfix[p.ecam[k]] && p.pfix[p.ept[k]])) p.active.push_back(k);
  vector<vector<int>> colCams(p.Pf);
  for (int k : p.active) {
    int c = p.camH[p.ecam[k]], l = p.ptH[p.ept[k]];
    if (c >= 0 && l >= 0) colCams[l].push_back(c);
  }
  p.colPtr.assign(p.Pf + 1, 0);
  for (int l = 0; l < p.Pf; ++l) {
    auto& v = colCams[l];
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    p.colPtr[l + 1] = p.colPtr[l] + static_cast<int>(v.size());
  }
  p.colCam.resize(p.colPtr[p.Pf]);
  for (int l = 0; l < p.Pf; ++l) std::copy(colCams[l].begin(), colCams[l].end(), p.colCam.begin() + p.colPtr[l]);
  p.edgeBlk.assign(p.E, -1);
  for (int k : p.active) {
    int c = p.camH[p.ecam[k]], l = p.ptH[p.ept[k]];
    if (c < 0 || l < 0) continue;
    auto b0 = p.colCam.begin() + p.colPtr[l], b1 = p.colCam.begin() + p.colPtr[l + 1];
    p.edgeBlk[k] = static_cast<int>(std::lower_bound(b0, b1, c) - p.colCam.begin());
  }
  vector<vector<int>> rows(p.Cf);
  for (int c = 0; c < p.Cf; ++c) rows[c].push_back(c);
  for (int l = 0; l < p.Pf; ++l)
    for (int a = p.colPtr[l]; a < p.colPtr[l + 1]; ++a)
      for (int bb = a; bb < p.colPtr[l + 1]; ++bb) rows[p.colCam[a]].push_back(p.colCam[bb]);
  p.sPtr.assign(p.Cf + 1, 0);
  for (int c = 0; c < p.Cf; ++c) {
    auto& v = rows[c];
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    p.sPtr[c + 1] = p.sPtr[c] + static_cast<int>(v.size());
  }
  p.sCol.resize(p.sPtr[p.Cf]);
  for (int c = 0; c < p.Cf; ++c) std::copy(rows[c].begin(), rows[c].end(), p.sCol.begin() + p.sPtr[c]);
  p.S.bd = 9;
  p.S.nb = p.Cf;
  p.S.row.resize(p.sCol.size());
  p.S.col = p.sCol;
  p.S.diag.assign(p.Cf, -1);
  for (int c = 0; c < p.Cf; ++c)
    for (int k = p.sPtr[c]; k < p.sPtr[c + 1]; ++k) {
      p.S.row[k] = c;
      if (p.sCol[k] == c) p.S.diag[c] = k;
    }
  p.S.val.assign(p.sCol.size() * 81, 0.0);
  p.err.assign(2 * static_cast<size_t>(p.E), 0.0);
  p.ptPtr.assign(p.P + 1, 0);
  for (int k : p.active) p.ptPtr[p.ept[k] + 1]++;
  for (int i = 0; i < p.P; ++i) p.ptPtr[i + 1] += p.ptPtr[i];
  p.ptEdge.resize(p.active.size());
  vector<int> fill(p.ptPtr.begin(), p.ptPtr.end() - 1);
  for (int k : p.active) p.ptEdge[fill[p.ept[k]]++] = k;
}

// computeActiveErrors + activeRobustChi2 (sparse_optimizer.cpp:78-130);
// Omega = w I (bal_example.cpp:313 sets the identity).
double bal_chi2(BAL& p) {
  double chi = 0;
  const int n = static_cast<int>(p.active.size());
#pragma omp parallel for reduction(+ : chi) schedule(static)
  for (int a = 0; a < n; ++a) {
    const int k = p.active[a];
    double e[2];
    bal_edge(&p.cams[9 * static_cast<size_t>(p.ecam[k])], &p.pts[3 * static_cast<size_t>(p.ept[k])], &p.meas[2 * k], e,
             nullptr, nullptr);
    p.err[2 * k] = e[0];
    p.err[2 * k + 1] = e[1];
    const double c2 = p.w[k] * (e[0] * e[0] + e[1] * e[1]);
    if (p.huber > 0) {
      double rho[3];
      huber(c2, p.huber, rho);
      chi += rho[0];
    } else {
      chi += c2;
    }
  }
  return chi;
}

// buildSystem (block_solver.hpp:492-565) with constructQuadraticForm
// (base_fixed_sized_edge.hpp:48-130): Hpp += Jc^T O' Jc, Hll += Jp^T O' Jp,
// Hpl += Jc^T O' Jp, b -= J^T O' e with O' = rho' w I.
void bal_build_system(BAL& p) {
  p.Hpp.assign(static_cast<size_t>(p.Cf) * 81, 0.0);
  p.Hll.assign(static_cast<size_t>(p.Pf) * 9, 0.0);
  p.Hpl.assign(p.colCam.size() * 27, 0.0);
  p.b.assign(9 * static_cast<size_t>(p.Cf) + 3 * static_cast<size_t>(p.Pf), 0.0);
  const size_t sp = 9 * static_cast<size_t>(p.Cf);
  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
#endif
  // points are split over the threads (a point's Hll, b_l and Hpl blocks are
  // its own); the camera sums go through per-thread accumulators added in
  // thread order, as the per-vertex locks of the reference's OpenMP build
  // serialise them (base_edge.h:42-58)
  vector<vector<double>> hppT(nth);
#pragma omp parallel
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    auto& hpp = hppT[tid];
    hpp.assign(static_cast<size_t>(p.Cf) * 90, 0.0);  // 81 Hpp + 9 b_p per camera
#pragma omp for schedule(static)
    for (int i = 0; i < p.P; ++i) {
      for (int a = p.ptPtr[i]; a < p.ptPtr[i + 1]; ++a) {
        const int k = p.ptEdge[a];
        const int ci = p.camH[p.ecam[k]], li = p.ptH[p.ept[k]];
        double e[2], Jc[18], Jp[6];
        bal_edge(&p.cams[9 * static_cast<size_t>(p.ecam[k])], &p.pts[3 * static_cast<size_t>(p.ept[k])],
                 &p.meas[2 * k], e, Jc, Jp);
        double om = p.w[k];
        if (p.huber > 0) {
          double rho[3];
          huber(p.w[k] * (e[0] * e[0] + e[1] * e[1]), p.huber, rho);
          om *= rho[1];
        }
        if (li >= 0) {
          for (int r = 0; r < 3; ++r) {
            p.b[sp + 3 * li + r] -= om * (Jp[r] * e[0] + Jp[3 + r] * e[1]);
            for (int c = 0; c < 3; ++c) p.Hll[9 * li + 3 * r + c] += om * (Jp[r] * Jp[c] + Jp[3 + r] * Jp[3 + c]);
          }
        }
        if (ci >= 0) {
          double* h = &hpp[90 * static_cast<size_t>(ci)];
          for (int r = 0; r < 9; ++r) {
            h[81 + r] -= om * (Jc[r] * e[0] + Jc[9 + r] * e[1]);
            for (int c = 0; c < 9; ++c) h[9 * r + c] += om * (Jc[r] * Jc[c] + Jc[9 + r] * Jc[9 + c]);
          }
        }
        if (ci >= 0 && li >= 0) {
          double* h = &p.Hpl[static_cast<size_t>(p.edgeBlk[k]) * 27];
          for (int r = 0; r < 9; ++r)
            for (int c = 0; c < 3; ++c) h[3 * r + c] += om * (Jc[r] * Jp[c] + Jc[9 + r] * Jp[3 + c]);
        }
      }
    }
  }
  for (int t = 0; t < nth; ++t)
    for (int c = 0; c < p.Cf; ++c) {
      const double* h = &hppT[t][90 * static_cast<size_t>(c)];
      for (int i = 0; i < 81; ++i) p.Hpp[81 * static_cast<size_t>(c) + i] += h[i];
      for (int r = 0; r < 9; ++r) p.b[9 * static_cast<size_t>(c) + r] += h[81 + r];
    }
}

// BlockSolver<9, 3>::solve with Schur (block_solver.hpp:331-477), lambda on
// the diagonals (setLambda, :567-590) applied to copies.
bool bal_solve(BAL& p) {
  const double t0 = now_s();
  const int Cf = p.Cf, Pf = p.Pf;
  const size_t sp = 9 * static_cast<size_t>(Cf);
  std::fill(p.S.val.begin(), p.S.val.end(), 0.0);
  for (int c = 0; c < Cf; ++c) {
    double* d = &p.S.val[static_cast<size_t>(p.S.diag[c]) * 81];
    for (int i = 0; i < 81; ++i) d[i] = p.Hpp[81 * static_cast<size_t>(c) + i];
    for (int i = 0; i < 9; ++i) d[10 * i] += p.lambda;
  }
  p.Dinv.assign(static_cast<size_t>(Pf) * 9, 0.0);
  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
#endif
  // points split over the threads, each with its own Hschur and coefficient
  // accumulators, added in thread order
  vector<vector<double>> Sth(nth), cth(nth);
  bool okD = true;
#pragma omp parallel reduction(&& : okD)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    auto& Sl = Sth[tid];
    auto& coeff = cth[tid];
    Sl.assign(p.S.val.size(), 0.0);
    coeff.assign(sp, 0.0);
#pragma omp for schedule(static)
    for (int l = 0; l < Pf; ++l) {
      double D[9];
      for (int i = 0; i < 9; ++i) D[i] = p.Hll[9 * static_cast<size_t>(l) + i];
      for (int i = 0; i < 3; ++i) D[4 * i] += p.lambda;
      double* Di = &p.Dinv[9 * static_cast<size_t>(l)];
      okD = inv3(D, Di) && okD;
      const double* bl = &p.b[sp + 3 * static_cast<size_t>(l)];
      double db[3];
      for (int r = 0; r < 3; ++r) db[r] = Di[3 * r] * bl[0] + Di[3 * r + 1] * bl[1] + Di[3 * r + 2] * bl[2];
      for (int a = p.colPtr[l]; a < p.colPtr[l + 1]; ++a) {
        const int i1 = p.colCam[a];
        const double* Bi = &p.Hpl[static_cast<size_t>(a) * 27];
        double BD[27];
        for (int r = 0; r < 9; ++r)
          for (int c = 0; c < 3; ++c)
            BD[3 * r + c] = Bi[3 * r] * Di[c] + Bi[3 * r + 1] * Di[3 + c] + Bi[3 * r + 2] * Di[6 + c];
        for (int r = 0; r < 9; ++r)
          coeff[9 * i1 + r] += Bi[3 * r] * db[0] + Bi[3 * r + 1] * db[1] + Bi[3 * r + 2] * db[2];
        int sk = p.sPtr[i1];
        for (int bb = a; bb < p.colPtr[l + 1]; ++bb) {
          const int i2 = p.colCam[bb];
          while (p.sCol[sk] < i2) ++sk;
          const double* Bj = &p.Hpl[static_cast<size_t>(bb) * 27];
          double* h = &Sl[static_cast<size_t>(sk) * 81];
          for (int r = 0; r < 9; ++r)
            for (int c = 0; c < 9; ++c)
              h[9 * r + c] -= BD[3 * r] * Bj[3 * c] + BD[3 * r + 1] * Bj[3 * c + 1] + BD[3 * r + 2] * Bj[3 * c + 2];
        }
      }
    }
  }
  if (!okD) return false;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < static_cast<long>(p.S.val.size()); ++i)
    for (int t = 0; t < nth; ++t) p.S.val[i] += Sth[t][i];
  p.bschur.assign(sp, 0.0);
  for (size_t i = 0; i < sp; ++i) {
    double c = 0;
    for (int t = 0; t < nth; ++t) c += cth[t][i];
    p.bschur[i] = p.b[i] - c;
  }
  const double t1 = now_s();
  p.times[2] += t1 - t0;
  p.x.assign(sp + 3 * static_cast<size_t>(Pf), 0.0);
  const bool ok = (p.solver_kind == 0) ? chol_solve_blocks(p.S, p.bschur.data(), p.x.data())
                                       : pcg_solve(p.S, p.bschur.data(), p.x.data(), p.pcg);
  p.times[3] += now_s() - t1;
  if (!ok) return false;
  // back-substitution (block_solver.hpp:452-476)
#pragma omp parallel for schedule(static)
  for (int l = 0; l < Pf; ++l) {
    double cl[3];
    for (int r = 0; r < 3; ++r) cl[r] = p.b[sp + 3 * static_cast<size_t>(l) + r];
    for (int a = p.colPtr[l]; a < p.colPtr[l + 1]; ++a) {
      const double* B = &p.Hpl[static_cast<size_t>(a) * 27];
      const double* xp = &p.x[9 * static_cast<size_t>(p.colCam[a])];
      for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 9; ++r) cl[c] -= B[3 * r + c] * xp[r];
    }
    const double* Di = &p.Dinv[9 * static_cast<size_t>(l)];
    for (int r = 0; r < 3; ++r)
      p.x[sp + 3 * static_cast<size_t>(l) + r] = Di[3 * r] * cl[0] + Di[3 * r + 1] * cl[1] + Di[3 * r + 2] * cl[2];
  }
  return true;
}

// VertexCameraBAL / VertexPointBAL::oplusImpl (bal_example.cpp:86-89, 115-118)
void bal_update(BAL& p, const double* x) {
  for (int c = 0; c < p.C; ++c) {
    const int h = p.camH[c];
    if (h < 0) continue;
    for (int r = 0; r < 9; ++r) p.cams[9 * static_cast<size_t>(c) + r] += x[9 * static_cast<size_t>(h) + r];
  }
  const size_t off = 9 * static_cast<size_t>(p.Cf);
  for (int i = 0; i < p.P; ++i) {
    const int h = p.ptH[i];
    if (h < 0) continue;
    for (int r = 0; r < 3; ++r) p.pts[3 * static_cast<size_t>(i) + r] += x[off + 3 * static_cast<size_t>(h) + r];
  }
}

// LM adapter of the BAL problem (lm_optimize)
struct BALLM {
  BAL& p;
  double* times = p.times;
  double chi2() { return bal_chi2(p); }
  void build() { bal_build_system(p); }
  // computeLambdaInit (optimization_algorithm_levenberg.cpp:155-167)
  double lambda_init() {
    double m = 0;
    for (int c = 0; c < p.Cf; ++c)
      for (int j = 0; j < 9; ++j) m = std::max(m, std::fabs(p.Hpp[81 * static_cast<size_t>(c) + 10 * j]));
    for (int l = 0; l < p.Pf; ++l)
      for (int j = 0; j < 3; ++j) m = std::max(m, std::fabs(p.Hll[9 * static_cast<size_t>(l) + 4 * j]));
    return 1e-5 * m;
  }
  bool solve(double lambda, int* pcg_it) {
    p.lambda = lambda;
    const bool ok = bal_solve(p);
    *pcg_it = p.solver_kind == 1 ? p.pcg.last_iters : 0;
    return ok;
  }
  void update() { bal_update(p, p.x.data()); }
  double scale(double lambda) {
    double s = 0;
    for (size_t j = 0; j < p.x.size(); ++j) s += p.x[j] * (lambda * p.x[j] + p.b[j]);
    return s;
  }
  void push() { p.stackC.push_back(p.cams); p.stackP.push_back(p.pts); }
  void pop() { p.cams = p.stackC.back(); p.pts = p.stackP.back(); discard(); }
  void discard() { p.stackC.pop_back(); p.stackP.pop_back(); }
};

}  // namespace

// =============================================================== C interface
// SparseOptimizer::optimize(iters) with OptimizationAlgorithmGaussNewton
// (optimization_algorithm_gauss_newton.cpp:47-90, sparse_optimizer.cpp:
// 393-453): computeActiveErrors, buildSystem, solve without damping, update,
// and stop after the first failed solve. A failed solve leaves the state
// unchanged here (the reference applies whatever x the solver left).
// trace[iters*4]: chi2 before, chi2 after, pcg iterations, status (1 = Fail).
template <typename P>
int gn_optimize(P& p, int iters, double* trace) {
  int it;
  for (it = 0; it < iters; ++it) {
    const double before = p.chi2();
    p.build();
    int pcg = 0;
    const bool ok = p.solve(0.0, &pcg);
    if (ok) p.update();
    if (trace) {
      trace[4 * it + 0] = before;
      trace[4 * it + 1] = p.chi2();
      trace[4 * it + 2] = pcg;
      trace[4 * it + 3] = ok ? 0 : 1;
    }
    if (!ok) { ++it; break; }
  }
  return it;
}

struct BAGN {
  BA& p;
  double chi2() { return ba_chi2(p); }
  void build() { ba_build_system(p); }
  bool solve(double lambda, int* pcg_it) {
    p.lambda = lambda;
    const bool ok = ba_solve(p);
    *pcg_it = p.pcg.last_iters;
    return ok;
  }
  void update() { ba_update(p, p.x.data()); }
};


// SparseOptimizer::optimize(iters) with OptimizationAlgorithmDogleg
// (optimization_algorithm_dogleg.cpp:62-213) at its default properties
// (initialDelta 1e4, maxTrialsAfterFailure 100, initialLambda 1e-7,
// lambdaFactor 10). multiplyHessian is BlockSolver's (block_solver.h:149-151):
// the product with Hpp only, so for bundle adjustment the landmark part of
// H v is zero, as in the reference.
// trace[iters*6]: chi2 before, chi2 after, delta after, tries, pcg iterations,
// status (0 OK, 1 Terminate, 2 Fail).
template <typename P>
int dl_optimize(P& p, int iters, double* trace) {
  double delta = 1e4, lam = 1e-7;
  const double lf = 10.0;
  const int max_tries = 100;
  bool wasPD = true;
  auto dot = [](const vector<double>& a, const vector<double>& b) {
    double s = 0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
  };
  int it;
  for (it = 0; it < iters; ++it) {
    const double currentChi = p.chi2();
    p.build();
    const vector<double> b = p.b();
    const size_t n = b.size();
    vector<double> aux(n, 0.0), hsd(n), hdl(n), hgn;
    p.hess_mul(b, aux);
    const double alpha = dot(b, b) / dot(aux, b);
    for (size_t i = 0; i < n; ++i) hsd[i] = alpha * b[i];
    const double hsdNorm = std::sqrt(dot(hsd, hsd));
    double hgnNorm = -1.0, chiAfter = currentChi;
    bool solved = false, good = false;
    int tries = 0, pcg = 0, status = 0;
    do {
      ++tries;
      if (!solved) {
        solved = true;
        bool ok = false;
        while (!ok) {
          int pit = 0;
          ok = p.solve(wasPD ? 0.0 : lam, &pit);
          pcg += pit;
          wasPD = wasPD && ok;
          if (!wasPD) {
            if (ok) {
              lam = std::max(1e-12, lam / (0.5 * lf));
            } else {
              lam *= lf;
              if (lam > 1e3) { lam = 1e3; status = 2; break; }
            }
          }
        }
        if (status == 2) break;
        hgn = p.x();
        hgnNorm = std::sqrt(dot(hgn, hgn));
      }
      if (hgnNorm < delta) {
        hdl = hgn;
      } else if (hsdNorm > delta) {
        for (size_t i = 0; i < n; ++i) hdl[i] = delta / hsdNorm * hsd[i];
      } else {
        for (size_t i = 0; i < n; ++i) aux[i] = hgn[i] - hsd[i];
        const double c = dot(hsd, aux), bma = dot(aux, aux), hsd2 = dot(hsd, hsd);
        const double disc = std::sqrt(c * c + bma * (delta * delta - hsd2));
        const double beta = c <= 0 ? (-c + disc) / bma : (delta * delta - hsd2) / (c + disc);
        for (size_t i = 0; i < n; ++i) hdl[i] = hsd[i] + beta * (hgn[i] - hsd[i]);
      }
      std::fill(aux.begin(), aux.end(), 0.0);
      p.hess_mul(hdl, aux);
      double lin = -dot(aux, hdl) + 2 * dot(b, hdl);
      p.push();
      p.update(hdl.data());
      const double newChi = p.chi2();
      if (std::fabs(lin) < 1e-12) lin = 1e-12;
      const double rho = (currentChi - newChi) / lin;
      if (rho > 0) {
        p.discard();
        good = true;
        chiAfter = newChi;
      } else {
        p.pop();
      }
      if (rho > 0.75)
        delta = std::max(delta, 3 * std::sqrt(dot(hdl, hdl)));
      else if (rho < 0.25)
        delta *= 0.5;
    } while (!good && tries < max_tries);
    if (status == 0 && !good) status = 1;
    if (trace) {
      double* t = trace + 6 * it;
      t[0] = currentChi; t[1] = chiAfter; t[2] = delta; t[3] = tries; t[4] = pcg; t[5] = status;
    }
    if (status) { ++it; break; }
  }
  return it;
}

struct BADL {
  BA& p;
  double chi2() { return ba_chi2(p); }
  void build() { ba_build_system(p); }
  vector<double> b() { return p.b; }
  vector<double> x() { return p.x; }
  // Hpp v in the pose part (Hpp is block diagonal), zero in the landmark part
  void hess_mul(const vector<double>& v, vector<double>& out) {
    for (int c = 0; c < p.Cf; ++c)
      for (int r = 0; r < 6; ++r) {
        double s = 0;
        for (int k = 0; k < 6; ++k) s += p.Hpp[36 * static_cast<size_t>(c) + 6 * r + k] * v[6 * static_cast<size_t>(c) + k];
        out[6 * static_cast<size_t>(c) + r] = s;
      }
  }
  bool solve(double lambda, int* pcg_it) {
    p.lambda = lambda;
    const bool ok = ba_solve(p);
    *pcg_it = p.pcg.last_iters;
    return ok;
  }
  void update(const double* v) { ba_update(p, v); }
  void push() { ba_push(p); }
  void pop() { ba_pop(p); }
  void discard() { ba_discard(p); }
};

struct PGDL {
  PG& p;
  double chi2() { return pg_chi2(p); }
  void build() { pg_build(p); }
  vector<double> b() { return p.b; }
  vector<double> x() { return p.x; }
  // H v over the upper blocks and their transposes, with the undamped diagonal
  void hess_mul(const vector<double>& v, vector<double>& out) {
    for (size_t k = 0; k < p.S.row.size(); ++k) {
      const int r = p.S.row[k], c = p.S.col[k];
      const double* B = r == c ? &p.Hdiag[36 * static_cast<size_t>(r)] : &p.S.val[36 * k];
      for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
          out[6 * static_cast<size_t>(r) + i] += B[6 * i + j] * v[6 * static_cast<size_t>(c) + j];
          if (r != c) out[6 * static_cast<size_t>(c) + j] += B[6 * i + j] * v[6 * static_cast<size_t>(r) + i];
        }
    }
  }
  bool solve(double lambda, int* pcg_it) {
    p.lambda = lambda;
    const bool ok = pg_solve(p);
    *pcg_it = p.solver_kind == 1 ? p.pcg.last_iters : 0;
    return ok;
  }
  void update(const double* v) { pg_update(p, v); }
  void push() { p.stackT.push_back(p.t); p.stackQ.push_back(p.q); }
  void pop() { p.t = p.stackT.back(); p.q = p.stackQ.back(); discard(); }
  void discard() { p.stackT.pop_back(); p.stackQ.pop_back(); }
};

extern "C" {

void* orc_ba_new(int C, int P, int E, const double* cam7, const double* intr4, const double* pts3,
                 const int* ecam, const int* ept, const double* meas2, const double* w,
                 const unsigned char* cam_fixed, const unsigned char* pt_fixed, double huber_delta) {
  BA* p = new BA();
  p->C = C; p->P = P; p->E = E;
  p->cams.resize(C);
  for (int c = 0; c < C; ++c) p->cams[c] = se3_from7(&cam7[7 * c]);
  p->intr.assign(intr4, intr4 + 4 * static_cast<size_t>(C));
  p->pts.assign(pts3, pts3 + 3 * static_cast<size_t>(P));
  p->ecam.assign(ecam, ecam + E);
  p->ept.assign(ept, ept + E);
  p->meas.assign(meas2, meas2 + 2 * static_cast<size_t>(E));
  if (w) p->w.assign(w, w + E); else p->w.assign(E, 1.0);
  p->cfix.assign(C, 0);
  p->pfix.assign(P, 0);
  if (cam_fixed) p->cfix.assign(cam_fixed, cam_fixed + C);
  if (pt_fixed) p->pfix.assign(pt_fixed, pt_fixed + P);
  p->huber = huber_delta;
  ba_structure(*p);
  return p;
}

void orc_ba_free(void* h) { delete static_cast<BA*>(h); }

void orc_ba_dims(void* h, int* cf, int* pf, int* nschur) {
  BA* p = static_cast<BA*>(h);
  *cf = p->Cf; *pf = p->Pf; *nschur = static_cast<int>(p->sCol.size());
}

void orc_ba_get_state(void* h, double* cam7, double* pts3) {
  BA* p = static_cast<BA*>(h);
  for (int c = 0; c < p->C; ++c) se3_to7(p->cams[c], &cam7[7 * c]);
  std::copy(p->pts.begin(), p->pts.end(), pts3);
}

void orc_ba_set_state(void* h, const double* cam7, const double* pts3) {
  BA* p = static_cast<BA*>(h);
  for (int c = 0; c < p->C; ++c) p->cams[c] = se3_from7(&cam7[7 * c]);
  p->pts.assign(pts3, pts3 + 3 * static_cast<size_t>(p->P));
}

double orc_ba_chi2(void* h) { return ba_chi2(*static_cast<BA*>(h)); }

// Linearize at the current state. Outputs are optional (nullptr skips):
// Hpp[Cf*36], Hll[Pf*9], Hpl_edge[E*18] (the Hpl block each edge maps to, zero
// for edges without one), b[6Cf+3Pf]. Returns chi2 at the state.
double orc_ba_linearize(void* h, double* Hpp, double* Hll, double* Hpl_edge, double* b) {
  BA* p = static_cast<BA*>(h);
  double chi = ba_chi2(*p);
  ba_build_system(*p);
  if (Hpp) std::copy(p->Hpp.begin(), p->Hpp.end(), Hpp);
  if (Hll) std::copy(p->Hll.begin(), p->Hll.end(), Hll);
  if (b) std::copy(p->b.begin(), p->b.end(), b);
  if (Hpl_edge) {
    for (int k = 0; k < p->E; ++k) {
      double* o = &Hpl_edge[18 * static_cast<size_t>(k)];
      if (p->edgeBlk[k] < 0) { std::fill(o, o + 18, 0.0); continue; }
      std::copy(&p->Hpl[18 * static_cast<size_t>(p->edgeBlk[k])], &p->Hpl[18 * static_cast<size_t>(p->edgeBlk[k])] + 18, o);
    }
  }
  return chi;
}

// Schur complement of the last linearization with damping lambda. Writes the
// upper blocks (rows/cols/36 values each, count = nschur of orc_ba_dims) and
// bschur[6Cf]. Returns 0, or 2 when an Hll block is singular.
int orc_ba_schur(void* h, double lambda, int* rows, int* cols, double* blocks, double* bschur) {
  BA* p = static_cast<BA*>(h);
  p->lambda = lambda;
  int saved = p->solver_kind;
  p->solver_kind = 0;
  // reuse ba_solve's Schur part by solving with Cholesky, then export
  bool ok = ba_solve(*p);
  p->solver_kind = saved;
  for (size_t k = 0; k < p->sCol.size(); ++k) {
    if (rows) rows[k] = p->S.row[k];
    if (cols) cols[k] = p->S.col[k];
    if (blocks) std::copy(&p->S.val[36 * k], &p->S.val[36 * k] + 36, &blocks[36 * k]);
  }
  if (bschur) std::copy(p->bschur.begin(), p->bschur.end(), bschur);
  return ok ? 0 : 2;
}

// One linear solve of the damped system at the last linearization:
// solver 0 = Cholesky of the reduced system, 1 = PCG (tol, maxiter; the PCG
// residual memory is reset first). dx[6Cf+3Pf]. Returns 0 or 2 (not PD).
int orc_ba_solve(void* h, double lambda, int solver, double tol, int maxiter, double* dx, int* pcg_iters) {
  BA* p = static_cast<BA*>(h);
  p->lambda = lambda;
  p->solver_kind = solver;
  p->pcg.tol = tol;
  p->pcg.max_iter = maxiter;
  p->pcg.residual = -1.0;
  bool ok = ba_solve(*p);
  if (pcg_iters) *pcg_iters = p->pcg.last_iters;
  if (!ok) return 2;
  std::copy(p->x.begin(), p->x.end(), dx);
  return 0;
}

// SparseOptimizer::optimize(iters) with OptimizationAlgorithmLevenberg
// (optimization_algorithm_levenberg.cpp:60-153). trace[iters*6] per iteration:
// chi2 before, chi2 after, lambda after, trials, pcg iterations (sum), status.
// times[6]: residuals, quadratic form, schur, linear solver, update, total.
// Returns the number of iterations run.
int orc_ba_optimize(void* h, int iters, int solver, double tol, int maxiter, double* trace, double* times) {
  BA* p = static_cast<BA*>(h);
  p->solver_kind = solver;
  p->pcg = PCGState();
  p->pcg.tol = tol;
  p->pcg.max_iter = maxiter;
  for (double& t : p->times) t = 0;
  p->trials.clear();
  double tall = now_s();
  double lambda = 0, ni = 2;
  int it;
  for (it = 0; it < iters; ++it) {
    double t0 = now_s();
    double currentChi = ba_chi2(*p);
    double t1 = now_s();
    p->times[0] += t1 - t0;
    ba_build_system(*p);
    p->times[1] += now_s() - t1;
    if (it == 0) { lambda = p->user_lambda > 0 ? p->user_lambda : ba_lambda_init(*p); ni = 2; }
    double chi_before = currentChi;
    double rho = 0;
    int q = 0, pcgsum = 0;
    do {
      ba_push(*p);
      p->lambda = lambda;
      bool ok = ba_solve(*p);
      pcgsum += p->pcg.last_iters;
      double tu = now_s();
      // A failed solve is rejected below and the state popped, so the update
      // the reference still applies (levenberg.cpp:112) has no effect.
      if (ok) ba_update(*p, p->x.data());
      p->times[4] += now_s() - tu;
      double tc = now_s();
      double tempChi = ba_chi2(*p);
      p->times[0] += now_s() - tc;
      if (!ok) tempChi = std::numeric_limits<double>::max();
      rho = currentChi - tempChi;
      double scale = ok ? ba_scale(*p) + 1e-3 : 1;
      rho /= scale;
      log_trial(&p->trials, lambda, currentChi, tempChi, scale, rho, rho > 0 && std::isfinite(tempChi) && ok);
      if (rho > 0 && std::isfinite(tempChi) && ok) {
        double alpha = 1. - std::pow((2 * rho - 1), 3);
        alpha = std::min(alpha, 2. / 3.);
        double sf = std::max(1. / 3., alpha);
        lambda *= sf;
        ni = 2;
        currentChi = tempChi;
        ba_discard(*p);
      } else {
        lambda *= ni;
        ni *= 2;
        ba_pop(*p);
        if (!std::isfinite(lambda)) break;
      }
      q++;
    } while (rho < 0 && q < 10);
    int status = (q == 10 || rho == 0 || !std::isfinite(lambda)) ? 1 : 0;
    if (trace) {
      trace[6 * it + 0] = chi_before;
      trace[6 * it + 1] = currentChi;
      trace[6 * it + 2] = lambda;
      trace[6 * it + 3] = q;
      trace[6 * it + 4] = pcgsum;
      trace[6 * it + 5] = status;
    }
    if (status) { ++it; break; }
  }
  p->times[5] = now_s() - tall;
  if (times) std::copy(p->times, p->times + 6, times);
  return it;
}

// OptimizationAlgorithmLevenberg's "initialLambda" property (<= 0: computeLambdaInit).
void orc_ba_set_user_lambda(void* h, double v) { static_cast<BA*>(h)->user_lambda = v; }

// The LM trials of the last orc_ba_optimize / orc_pg_optimize: up to `max`
// records of LM_TRIAL_FIELDS doubles (see log_trial). Returns the count.
static int copy_trials(const std::vector<double>& t, double* out, int max) {
  const int n = static_cast<int>(t.size()) / LM_TRIAL_FIELDS;
  const int m = std::min(n, max);
  if (out) std::copy(t.begin(), t.begin() + static_cast<size_t>(m) * LM_TRIAL_FIELDS, out);
  return n;
}
int orc_ba_trials(void* h, double* out, int max) { return copy_trials(static_cast<BA*>(h)->trials, out, max); }
int orc_pg_trials(void* h, double* out, int max) { return copy_trials(static_cast<PG*>(h)->trials, out, max); }

int orc_ba_gauss_newton(void* h, int iters, int solver, double tol, int maxiter, double* trace) {
  BA* p = static_cast<BA*>(h);
  p->solver_kind = solver;
  p->pcg = PCGState();
  p->pcg.tol = tol;
  p->pcg.max_iter = maxiter;
  BAGN gn{*p};
  return gn_optimize(gn, iters, trace);
}

// Replace the measurements / information weights (same structure).
// BaseEdge::setInformation per edge: om[E*3] = [O00 O01 O11]; null: identity
void orc_ba_set_info(void* h, const double* om) {
  BA* p = static_cast<BA*>(h);
  if (om) p->om.assign(om, om + 3 * static_cast<size_t>(p->E));
  else p->om.clear();
}

void orc_ba_set_meas(void* h, const double* meas, const double* w) {
  BA* p = static_cast<BA*>(h);
  std::copy(meas, meas + 2 * static_cast<size_t>(p->E), p->meas.begin());
  if (w) std::copy(w, w + p->E, p->w.begin());
}

// ----------------------------------------------------------------- adjoint
// Implicit differentiation of the converged solution theta*(z, w): the
// stationarity G(theta*, z, w) = sum_e dE_e/dtheta = 0 gives
// dL/dz = -(dG/dz)^T lambda with H lambda = dL/dtheta, H = dG/dtheta the exact
// Hessian (second-order projection and Huber terms included), solved by the
// Schur complement and a dense Cholesky of the reduced camera system.
// dL_dcam[C*6] is the loss gradient in the left exp-map coordinates
// [omega; upsilon] of each camera, dL_dpt[P*3] in point coordinates (entries of
// fixed vertices are ignored). Outputs dL_dmeas[E*2], dL_dw[E] (zero for
// inactive edges) and, if not null, lambda[6Cf+3Pf]. Returns 0, or 2 if H is
// not positive definite.
int orc_ba_adjoint(void* h, const double* dL_dcam, const double* dL_dpt, double* dL_dmeas, double* dL_dw,
                   double* lambda_out) {
  BA* p = static_cast<BA*>(h);
  ba_chi2(*p);
  vector<double> mixed;
  ba_build_exact(*p, &mixed);
  p->b.assign(6 * static_cast<size_t>(p->Cf) + 3 * static_cast<size_t>(p->Pf), 0.0);
  for (int c = 0; c < p->C; ++c)
    if (p->camH[c] >= 0)
      for (int r = 0; r < 6; ++r) p->b[6 * static_cast<size_t>(p->camH[c]) + r] = dL_dcam[6 * static_cast<size_t>(c) + r];
  for (int i = 0; i < p->P; ++i)
    if (p->ptH[i] >= 0)
      for (int r = 0; r < 3; ++r)
        p->b[6 * static_cast<size_t>(p->Cf) + 3 * static_cast<size_t>(p->ptH[i]) + r] = dL_dpt[3 * static_cast<size_t>(i) + r];
  const int saved_kind = p->solver_kind;
  const double saved_lambda = p->lambda;
  p->solver_kind = 0;
  p->lambda = 0.0;
  bool ok = ba_solve(*p);
  p->solver_kind = saved_kind;
  p->lambda = saved_lambda;
  if (!ok) return 2;
  for (int k = 0; k < p->E; ++k) {
    double gz[3] = {0, 0, 0};
    int ci = p->camH[p->ecam[k]], li = p->ptH[p->ept[k]];
    const double* m = &mixed[27 * static_cast<size_t>(k)];
    if (ci >= 0)
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 3; ++c) gz[c] -= m[3 * r + c] * p->x[6 * static_cast<size_t>(ci) + r];
    if (li >= 0)
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
          gz[c] -= m[3 * (6 + r) + c] * p->x[6 * static_cast<size_t>(p->Cf) + 3 * static_cast<size_t>(li) + r];
    dL_dmeas[2 * static_cast<size_t>(k)] = gz[0];
    dL_dmeas[2 * static_cast<size_t>(k) + 1] = gz[1];
    dL_dw[k] = gz[2];
  }
  if (lambda_out) std::copy(p->x.begin(), p->x.end(), lambda_out);
  return 0;
}

// Newton steps with the exact Hessian (gradient from ba_build_system's b =
// -grad(F/2)), damped as Levenberg-Marquardt on that Hessian: a step is kept
// if the robust chi2 does not increase, else the state is restored and the
// damping mu (0 at first) raised to max(1e-15 * max diag, 10 mu) (an increase
// within 1e-12 relative, rounding at the optimum, counts as none); after an
// accepted step mu drops by 10 (below 1e-15 * max diag it is 0); a step with
// a component above max_step (> 0) counts as rejected. Drives a
// converged LM solution to the stationary point to machine precision before
// it is differentiated. Returns 0, or 2 if no step was acceptable.
int orc_ba_newton(void* h, int steps, double max_step) {
  BA* p = static_cast<BA*>(h);
  const int saved_kind = p->solver_kind;
  const double saved_lambda = p->lambda;
  p->solver_kind = 0;
  int st = 0;
  double mu = 0.0;
  for (int it = 0; it < steps && st == 0; ++it) {
    const double chi0 = ba_chi2(*p);
    ba_build_system(*p);  // b
    ba_build_exact(*p, nullptr);
    double mx = 0;
    for (int c = 0; c < p->Cf; ++c)
      for (int k = 0; k < 6; ++k) mx = std::max(mx, std::fabs(p->Hpp[36 * static_cast<size_t>(c) + 7 * k]));
    for (int l = 0; l < p->Pf; ++l)
      for (int k = 0; k < 3; ++k) mx = std::max(mx, std::fabs(p->Hll[9 * static_cast<size_t>(l) + 4 * k]));
    const double mu_min = 1e-15 * mx;
    bool accepted = false;
    for (int trial = 0; trial < 30 && !accepted; ++trial) {
      p->lambda = mu;
      bool solved = ba_solve(*p);
      if (solved && max_step > 0) {  // trust region on the step
        double m = 0;
        for (double v : p->x) m = std::max(m, std::fabs(v));
        if (m > max_step) solved = false;
      }

      if (solved) {
        vector<SE3> c0 = p->cams;
        vector<double> x0 = p->pts;
        ba_update(*p, p->x.data());
        const double chi1 = ba_chi2(*p);

        if (std::isfinite(chi1) && chi1 <= chi0 + 1e-12 * std::fabs(chi0)) {
          accepted = true;
        } else {
          p->cams = c0;
          p->pts = x0;
        }
      }
      if (!accepted) mu = std::max(mu_min, 10 * mu);
    }
    if (!accepted) st = 2;

    mu = (mu / 10 < mu_min) ? 0.0 : mu / 10;
  }
  p->solver_kind = saved_kind;
  p->lambda = saved_lambda;
  return st;
}

// Exact-Hessian blocks at the current state (as orc_ba_linearize lays out the
// Gauss-Newton ones).
void orc_ba_exact_hessian(void* h, double* Hpp, double* Hll, double* Hpl_edge) {
  BA* p = static_cast<BA*>(h);
  ba_chi2(*p);
  ba_build_exact(*p, nullptr);
  if (Hpp) std::copy(p->Hpp.begin(), p->Hpp.end(), Hpp);
  if (Hll) std::copy(p->Hll.begin(), p->Hll.end(), Hll);
  if (Hpl_edge)
    for (int k = 0; k < p->E; ++k) {
      double* o = &Hpl_edge[18 * static_cast<size_t>(k)];
      if (p->edgeBlk[k] < 0) { std::fill(o, o + 18, 0.0); continue; }
      std::copy(&p->Hpl[18 * static_cast<size_t>(p->edgeBlk[k])], &p->Hpl[18 * static_cast<size_t>(p->edgeBlk[k])] + 18, o);
    }
}

// Energy 0.5 rho of one edge with its gradient g[12] and Hessian H[144] in
// the variables [omega upsilon dX dz dw] at zero.
double orc_edge_energy_derivs(const double* cam7, const double* X3, const double* z2, double w, const double* intr4,
                              double huber_delta, double* g12, double* H144) {
  double E = 0;
  edge_energy_derivs(se3_from7(cam7), X3, z2, w, intr4, huber_delta, &E, g12, H144);
  return E;
}

// ------------------------------------------------------------ pose graph API
// poses7[N*7] [tx ty tz qx qy qz qw], from/to[E], meas7[E*7], info36[E*36]
// (row-major), fixed[N] (0/1).
void* orc_pg_new(int N, int E, const double* poses7, const int* from, const int* to, const double* meas7,
                 const double* info36, const uint8_t* fixed) {
  PG* p = new PG();
  p->N = N;
  p->E = E;
  p->t.resize(3 * static_cast<size_t>(N));
  p->q.resize(N);
  for (int i = 0; i < N; ++i) {
    for (int a = 0; a < 3; ++a) p->t[3 * i + a] = poses7[7 * static_cast<size_t>(i) + a];
    const double* qq = poses7 + 7 * static_cast<size_t>(i) + 3;
    p->q[i] = {qq[0], qq[1], qq[2], qq[3]};
    qnormalize(p->q[i]);
  }
  p->from.assign(from, from + E);
  p->to.assign(to, to + E);
  p->mt.resize(3 * static_cast<size_t>(E));
  p->mq.resize(E);
  for (int k = 0; k < E; ++k) {
    for (int a = 0; a < 3; ++a) p->mt[3 * k + a] = meas7[7 * static_cast<size_t>(k) + a];
    const double* qq = meas7 + 7 * static_cast<size_t>(k) + 3;
    p->mq[k] = {qq[0], qq[1], qq[2], qq[3]};
    qnormalize(p->mq[k]);
  }
  p->info.assign(info36, info36 + 36 * static_cast<size_t>(E));
  p->fixed.assign(fixed, fixed + N);
  pg_structure(*p);
  return p;
}

void orc_pg_free(void* h) { delete static_cast<PG*>(h); }
int orc_pg_nfree(void* h) { return static_cast<PG*>(h)->Nf; }
int orc_pg_nblocks(void* h) { return static_cast<int>(static_cast<PG*>(h)->S.row.size()); }

void orc_pg_get_state(void* h, double* poses7) {
  PG* p = static_cast<PG*>(h);
  for (int i = 0; i < p->N; ++i) {
    for (int a = 0; a < 3; ++a) poses7[7 * static_cast<size_t>(i) + a] = p->t[3 * i + a];
    const Quat& q = p->q[i];
    double* o = poses7 + 7 * static_cast<size_t>(i) + 3;
    o[0] = q.x; o[1] = q.y; o[2] = q.z; o[3] = q.w;
  }
}

void orc_pg_set_state(void* h, const double* poses7) {
  PG* p = static_cast<PG*>(h);
  for (int i = 0; i < p->N; ++i) {
    for (int a = 0; a < 3; ++a) p->t[3 * i + a] = poses7[7 * static_cast<size_t>(i) + a];
    const double* qq = poses7 + 7 * static_cast<size_t>(i) + 3;
    p->q[i] = {qq[0], qq[1], qq[2], qq[3]};
    qnormalize(p->q[i]);
  }
}

double orc_pg_chi2(void* h) { return pg_chi2(*static_cast<PG*>(h)); }

// error e6 and Jacobian J72 = d e / d [d_from(6) d_to(6)] (6x12 row-major) of edge k
void orc_pg_edge(void* h, int k, double* e6, double* J72) { pg_edge(*static_cast<PG*>(h), k, e6, J72); }

// Undamped system at the state: rows/cols/blocks (count orc_pg_nblocks, upper,
// 36 row-major each) and b[6 Nf]. Returns chi2.
double orc_pg_linearize(void* h, int* rows, int* cols, double* blocks, double* b) {
  PG* p = static_cast<PG*>(h);
  double chi = pg_chi2(*p);
  pg_build(*p);
  for (size_t k = 0; k < p->S.row.size(); ++k) {
    if (rows) rows[k] = p->S.row[k];
    if (cols) cols[k] = p->S.col[k];
  }
  if (blocks) std::copy(p->S.val.begin(), p->S.val.end(), blocks);
  if (b) std::copy(p->b.begin(), p->b.end(), b);
  return chi;
}

int orc_pg_optimize(void* h, int iters, int solver, double tol, int maxiter, double* trace, double* times) {
  PG* p = static_cast<PG*>(h);
  p->solver_kind = solver;
  p->pcg = PCGState();
  p->pcg.tol = tol;
  p->pcg.max_iter = maxiter;
  PGLM lm{*p};
  return lm_optimize(lm, iters, trace, times, &p->trials);
}

int orc_pg_gauss_newton(void* h, int iters, int solver, double tol, int maxiter, double* trace) {
  PG* p = static_cast<PG*>(h);
  p->solver_kind = solver;
  p->pcg = PCGState();
  p->pcg.tol = tol;
  p->pcg.max_iter = maxiter;
  PGLM lm{*p};
  return gn_optimize(lm, iters, trace);
}

// -------------------------------------------------------- single functions
void orc_se3_exp(const double* upd6, double* out7) { se3_to7(se3_exp(upd6), out7); }
void orc_se3_log(const double* a7, double* out6) { se3_log(se3_from7(a7), out6); }
void orc_se3_mul(const double* a7, const double* b7, double* out7) {
  se3_to7(se3_mul(se3_from7(a7), se3_from7(b7)), out7);
}
void orc_quat_from_R(const double* R9, double* q4) {
  Quat q = R2q(R9);
  q4[0] = q.x; q4[1] = q.y; q4[2] = q.z; q4[3] = q.w;
}
void orc_quat_to_R(const double* q4, double* R9) { q2R({q4[0], q4[1], q4[2], q4[3]}, R9); }
void orc_huber(double e, double delta, double* rho3) { huber(e, delta, rho3); }
void orc_project_edge(const double* cam7, const double* intr4, const double* X3, const double* z2,
                      double* e2, double* Jp6, double* Jc12) {
  SE3 T = se3_from7(cam7);
  double R[9];
  q2R(T.q, R);
  project_edge(T, R, intr4, X3, z2, e2, Jp6, Jc12);
}

// Block-Jacobi PCG on a symmetric matrix given by its upper blocks
// (brow <= bcol, bd x bd row-major values). residual_io carries the PCG
// residual memory (-1 for a fresh solver). Returns iterations, or -1 when a
// diagonal block is missing or singular.
int orc_pcg_blocks(int nb, int bd, int nblk, const int* brow, const int* bcol, const double* blocks,
                   const double* b, double* x, double tol, int maxiter, double* residual_io) {
  BlockSym A;
  A.bd = bd; A.nb = nb;
  A.row.assign(brow, brow + nblk);
  A.col.assign(bcol, bcol + nblk);
  A.val.assign(blocks, blocks + static_cast<size_t>(nblk) * bd * bd);
  A.diag.assign(nb, -1);
  for (int k = 0; k < nblk; ++k) if (brow[k] == bcol[k]) A.diag[brow[k]] = k;
  PCGState st;
  st.tol = tol; st.max_iter = maxiter; st.residual = residual_io ? *residual_io : -1.0;
  if (!pcg_solve(A, b, x, st)) return -1;
  if (residual_io) *residual_io = st.residual;
  return st.last_iters;
}

// Dense Cholesky solve (row-major full A). Returns 0, or 2 when not PD.
int orc_chol_solve(int n, const double* A, const double* b, double* x) {
  return chol_solve(n, A, b, x) ? 0 : 2;
}

int orc_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int orc_ba_dogleg(void* h, int iters, int solver, double tol, int maxiter, double* trace) {
  BA* p = static_cast<BA*>(h);
  p->solver_kind = solver;
  p->pcg = PCGState();
  p->pcg.tol = tol;
  p->pcg.max_iter = maxiter;
  BADL dl{*p};
  return dl_optimize(dl, iters, trace);
}

int orc_pg_dogleg(void* h, int iters, int solver, double tol, int maxiter, double* trace) {
  PG* p = static_cast<PG*>(h);
  p->solver_kind = solver;
  p->pcg = PCGState();
  p->pcg.tol = tol;
  p->pcg.max_iter = maxiter;
  PGDL dl{*p};
  return dl_optimize(dl, iters, trace);
}


// ------------------------------------------------------------------ BAL (9-DoF)
void* orc_bal_new(int C, int P, int E, const double* cams9, const double* pts3, const int* ecam, const int* ept,
                  const double* meas2, const double* w, const unsigned char* cam_fixed,
                  const unsigned char* pt_fixed, double huber_delta) {
  BAL* p = new BAL();
  p->C = C; p->P = P; p->E = E;
  p->cams.assign(cams9, cams9 + 9 * static_cast<size_t>(C));
  p->pts.assign(pts3, pts3 + 3 * static_cast<size_t>(P));
  p->ecam.assign(ecam, ecam + E);
  p->ept.assign(ept, ept + E);
  p->meas.assign(meas2, meas2 + 2 * static_cast<size_t>(E));
  if (w) p->w.assign(w, w + E); else p->w.assign(E, 1.0);
  p->cfix.assign(C, 0);
  p->pfix.assign(P, 0);
  if (cam_fixed) p->cfix.assign(cam_fixed, cam_fixed + C);
  if (pt_fixed) p->pfix.assign(pt_fixed, pt_fixed + P);
  p->huber = huber_delta;
  bal_structure(*p);
  return p;
}

void orc_bal_free(void* h) { delete static_cast<BAL*>(h); }

void orc_bal_dims(void* h, int* cf, int* pf, int* nschur) {
  BAL* p = static_cast<BAL*>(h);
  *cf = p->Cf; *pf = p->Pf; *nschur = static_cast<int>(p->sCol.size());
}

void orc_bal_get_state(void* h, double* cams9, double* pts3) {
  BAL* p = static_cast<BAL*>(h);
  std::copy(p->cams.begin(), p->cams.end(), cams9);
  std::copy(p->pts.begin(), p->pts.end(), pts3);
}

void orc_bal_set_state(void* h, const double* cams9, const double* pts3) {
  BAL* p = static_cast<BAL*>(h);
  p->cams.assign(cams9, cams9 + 9 * static_cast<size_t>(p->C));
  p->pts.assign(pts3, pts3 + 3 * static_cast<size_t>(p->P));
}

// chi2 at the state; err[E*2] (optional) the edges' errors (0 for inactive)
double orc_bal_chi2(void* h, double* err) {
  BAL* p = static_cast<BAL*>(h);
  std::fill(p->err.begin(), p->err.end(), 0.0);
  const double chi = bal_chi2(*p);
  if (err) std::copy(p->err.begin(), p->err.end(), err);
  return chi;
}

// One edge function: e[2], Jc[18], Jp[6] (row-major; Jacobians optional)
void orc_bal_edge(const double* cam9, const double* X, const double* z, double* e, double* Jc, double* Jp) {
  bal_edge(cam9, X, z, e, Jc, Jp);
}

// Linearise at the state: Hpp[Cf*81], Hll[Pf*9], Hpl_edge[E*27], b[9Cf+3Pf]
// (each optional). Returns chi2.
double orc_bal_linearize(void* h, double* Hpp, double* Hll, double* Hpl_edge, double* b) {
  BAL* p = static_cast<BAL*>(h);
  const double chi = bal_chi2(*p);
  bal_build_system(*p);
  if (Hpp) std::copy(p->Hpp.begin(), p->Hpp.end(), Hpp);
  if (Hll) std::copy(p->Hll.begin(), p->Hll.end(), Hll);
  if (b) std::copy(p->b.begin(), p->b.end(), b);
  if (Hpl_edge) {
    for (int k = 0; k < p->E; ++k) {
      double* o = &Hpl_edge[27 * static_cast<size_t>(k)];
      if (p->edgeBlk[k] < 0) { std::fill(o, o + 27, 0.0); continue; }
      const double* s = &p->Hpl[27 * static_cast<size_t>(p->edgeBlk[k])];
      std::copy(s, s + 27, o);
    }
  }
  return chi;
}

// Schur complement of the last linearisation at damping lambda: upper blocks
// (rows, cols, 81 values each) and bschur[9Cf]. Returns 0, or 2 if a point
// block is singular.
int orc_bal_schur(void* h, double lambda, int* rows, int* cols, double* blocks, double* bschur) {
  BAL* p = static_cast<BAL*>(h);
  if (p->Hpp.empty()) bal_build_system(*p);  // not linearised yet
  p->lambda = lambda;
  const int saved = p->solver_kind;
  p->solver_kind = 0;
  const bool ok = bal_solve(*p);
  p->solver_kind = saved;
  for (size_t k = 0; k < p->sCol.size(); ++k) {
    if (rows) rows[k] = p->S.row[k];
    if (cols) cols[k] = p->S.col[k];
    if (blocks) std::copy(&p->S.val[81 * k], &p->S.val[81 * k] + 81, &blocks[81 * k]);
  }
  if (bschur) std::copy(p->bschur.begin(), p->bschur.end(), bschur);
  return ok ? 0 : 2;
}

// One damped solve at the last linearisation (solver 0 Cholesky, 1 PCG with
// the residual memory reset): dx[9Cf+3Pf]. Returns 0 or 2 (not PD).
int orc_bal_solve(void* h, double lambda, int solver, double tol, int maxiter, double* dx, int* pcg_iters) {
  BAL* p = static_cast<BAL*>(h);
  if (p->Hpp.empty()) bal_build_system(*p);  // not linearised yet
  p->lambda = lambda;
  p->solver_kind = solver;
  p->pcg = PCGState();
  p->pcg.tol = tol;
  p->pcg.max_iter = maxiter;
  const bool ok = bal_solve(*p);
  if (pcg_iters) *pcg_iters = p->pcg.last_iters;
  if (!ok) return 2;
  std::copy(p->x.begin(), p->x.end(), dx);
  return 0;
}

// SparseOptimizer::optimize(iters) with OptimizationAlgorithmLevenberg over
// BlockSolver<9, 3> (bal_example.cpp:227-250, 341): trace as orc_ba_optimize.
int orc_bal_optimize(void* h, int iters, int solver, double tol, int maxiter, double* trace, double* times) {
  BAL* p = static_cast<BAL*>(h);
  p->solver_kind = solver;
  p->pcg = PCGState();
  p->pcg.tol = tol;
  p->pcg.max_iter = maxiter;
  BALLM lm{*p};
  return lm_optimize(lm, iters, trace, times, &p->trials);
}

int orc_bal_trials(void* h, double* out, int max) { return copy_trials(static_cast<BAL*>(h)->trials, out, max); }

}  // extern "C"
