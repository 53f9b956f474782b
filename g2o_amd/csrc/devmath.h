// Device-side math shared by the kernels: SE3Quat pieces, the projection edge,
// Huber, small inverses and reductions. FP64 unless templated.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace g2ohip {

// Rotation matrix (row-major) of a unit quaternion [x y z w].
__host__ __device__ inline void quat_to_rot(const double* q, double* R) {
  double tx = 2 * q[0], ty = 2 * q[1], tz = 2 * q[2];
  double twx = tx * q[3], twy = ty * q[3], twz = tz * q[3];
  double txx = tx * q[0], txy = ty * q[0], txz = tz * q[0];
  double tyy = ty * q[1], tyz = tz * q[1], tzz = tz * q[2];
  R[0] = 1 - (tyy + tzz); R[1] = txy - twz;       R[2] = txz + twy;
  R[3] = txy + twz;       R[4] = 1 - (txx + tzz); R[5] = tyz - twx;
  R[6] = txz - twy;       R[7] = tyz + twx;       R[8] = 1 - (txx + tyy);
}

// Quaternion of a rotation matrix, Shepperd's method: the largest of
// w^2, x^2, y^2, z^2 is taken from the trace / diagonal, the others from the
// off-diagonal sums and differences.
__host__ __device__ inline void rot_to_quat(const double* R, double* q) {
  double tr = R[0] + R[4] + R[8];
  if (tr > 0) {
    double s = sqrt(tr + 1.0);
    q[3] = 0.5 * s;
    s = 0.5 / s;
    q[0] = (R[7] - R[5]) * s;
    q[1] = (R[2] - R[6]) * s;
    q[2] = (R[3] - R[1]) * s;
  } else {
    int i = 0;
    if (R[4] > R[0]) i = 1;
    if (R[8] > R[3 * i + i]) i = 2;
    int j = (i + 1) % 3, k = (j + 1) % 3;
    double s = sqrt(R[3 * i + i] - R[3 * j + j] - R[3 * k + k] + 1.0);
    double v[3];
    v[i] = 0.5 * s;
    s = 0.5 / s;
    q[3] = (R[3 * k + j] - R[3 * j + k]) * s;
    v[j] = (R[3 * j + i] + R[3 * i + j]) * s;
    v[k] = (R[3 * k + i] + R[3 * i + k]) * s;
    q[0] = v[0]; q[1] = v[1]; q[2] = v[2];
  }
}

// w >= 0 then unit norm: SE3Quat::normalizeRotation (se3quat.h:251-256).
__host__ __device__ inline void normalize_quat(double* q) {
  if (q[3] < 0) { q[0] = -q[0]; q[1] = -q[1]; q[2] = -q[2]; q[3] = -q[3]; }
  double n = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  q[0] /= n; q[1] /= n; q[2] /= n; q[3] /= n;
}

// Hamilton product a * b.
__host__ __device__ inline void quat_mul(const double* a, const double* b, double* r) {
  r[0] = a[3] * b[0] + a[0] * b[3] + a[1] * b[2] - a[2] * b[1];
  r[1] = a[3] * b[1] + a[1] * b[3] + a[2] * b[0] - a[0] * b[2];
  r[2] = a[3] * b[2] + a[2] * b[3] + a[0] * b[1] - a[1] * b[0];
  r[3] = a[3] * b[3] - a[0] * b[0] - a[1] * b[1] - a[2] * b[2];
}

// pose7 <- exp(upd) * pose7: VertexSE3Expmap::oplusImpl
// (vertex_se3_expmap.cpp:48-51) with SE3Quat::exp (se3quat.h:202-230) and
// SE3Quat::operator* (se3quat.h:97-103). upd = [omega; upsilon].
__host__ __device__ inline void se3_left_update(const double* upd, const double* in, double* out) {
  const double w0 = upd[0], w1 = upd[1], w2 = upd[2];
  double theta = sqrt(w0 * w0 + w1 * w1 + w2 * w2);
  double W[9] = {0, -w2, w1, w2, 0, -w0, -w1, w0, 0};
  double W2[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) W2[3 * i + j] = W[3 * i] * W[j] + W[3 * i + 1] * W[3 + j] + W[3 * i + 2] * W[6 + j];
  double a, b, c, d;
  if (theta < 0.00001) {
    a = 1.0; b = 0.5; c = 0.5; d = 1.0 / 6.0;
  } else {
    double st = sin(theta), ct = cos(theta);
    a = st / theta;
    b = (1 - ct) / (theta * theta);
    c = b;
    d = (theta - st) / (theta * theta * theta);
  }
  double R[9], V[9];
  for (int i = 0; i < 9; ++i) {
    double I = (i % 4 == 0) ? 1.0 : 0.0;
    R[i] = I + a * W[i] + b * W2[i];
    V[i] = I + c * W[i] + d * W2[i];
  }
  double qe[4], te[3];
  rot_to_quat(R, qe);
  normalize_quat(qe);
  for (int i = 0; i < 3; ++i) te[i] = V[3 * i] * upd[3] + V[3 * i + 1] * upd[4] + V[3 * i + 2] * upd[5];
  // exp * T: t = te + Re * t_T, q = qe * q_T
  double Re[9];
  quat_to_rot(qe, Re);
  const double* tT = in;
  double qT[4] = {in[3], in[4], in[5], in[6]};
  for (int i = 0; i < 3; ++i) out[i] = te[i] + Re[3 * i] * tT[0] + Re[3 * i + 1] * tT[1] + Re[3 * i + 2] * tT[2];
  double q[4];
  quat_mul(qe, qT, q);
  normalize_quat(q);
  out[3] = q[0]; out[4] = q[1]; out[5] = q[2]; out[6] = q[3];
}

// RobustKernelHuber::robustify (robust_kernel_impl.cpp:60-74); delta <= 0
// means no robust kernel (rho = (e, 1, 0)).
__host__ __device__ inline void huber_rho(double e, double delta, double* rho) {
  if (delta <= 0) { rho[0] = e; rho[1] = 1.0; rho[2] = 0.0; return; }
  double dsqr = delta * delta;
  if (e <= dsqr) {
    rho[0] = e; rho[1] = 1.0; rho[2] = 0.0;
  } else {
    double sqrte = sqrt(e);
    rho[0] = 2 * sqrte * delta - dsqr;
    rho[1] = delta / sqrte;
    rho[2] = -0.5 * rho[1] / e;
  }
}

// Xc = R X + t and e = z - (fx x/z + cx, fy y/z + cy)
// (EdgeSE3ProjectXYZ::computeError, edge_project_xyz.cpp:44-49).
__device__ inline void project_error(const double* Rt, const double* in, const double* X, const double* z,
                                     double* Xc, double* e) {
  for (int i = 0; i < 3; ++i) Xc[i] = Rt[3 * i] * X[0] + Rt[3 * i + 1] * X[1] + Rt[3 * i + 2] * X[2] + Rt[9 + i];
  e[0] = z[0] - (Xc[0] / Xc[2] * in[0] + in[2]);
  e[1] = z[1] - (Xc[1] / Xc[2] * in[1] + in[3]);
}

// An edge's 2x2 information Omega (BaseEdge::information, base_edge.h:165-180)
// enters through its upper-triangular factor S = [[s0, s1], [0, s2]],
// Omega = S^T S: with e' = S e and J' = S J, e^T Omega e = |e'|^2,
// J^T Omega J = J'^T J' and J^T Omega e = J'^T e', so every Gauss-Newton
// kernel works on whitened errors and Jacobians. S == nullptr: Omega = I.
__host__ __device__ inline void whiten_vec(const double* S, double* e) {
  e[0] = S[0] * e[0] + S[1] * e[1];
  e[1] = S[2] * e[1];
}
// rows of a 2xN row-major Jacobian
template <int N>
__host__ __device__ inline void whiten_rows(const double* S, double* J) {
#pragma unroll
  for (int c = 0; c < N; ++c) {
    J[c] = S[0] * J[c] + S[1] * J[N + c];
    J[N + c] = S[2] * J[N + c];
  }
}
// Omega = S^T S as its upper triangle [o00 o01 o11]
__host__ __device__ inline void info_of(const double* S, double* om) {
  om[0] = S[0] * S[0];
  om[1] = S[0] * S[1];
  om[2] = S[1] * S[1] + S[2] * S[2];
}

// Analytic Jacobians of EdgeSE3ProjectXYZ::linearizeOplus
// (edge_project_xyz.cpp:57-93) in precision T: Jp = d e / d X (2x3 row-major),
// Jc = d e / d [omega; upsilon] (2x6 row-major), from the camera-frame point.
template <typename T>
__host__ __device__ inline void project_jacobians(const T* R, T fx, T fy, T x, T y, T z, T* Jp, T* Jc) {
  T iz = T(1) / z;
  T z2 = z * z;
  T t00 = fx, t02 = -x * iz * fx, t11 = fy, t12 = -y * iz * fy;
  for (int c = 0; c < 3; ++c) {
    Jp[c] = -iz * (t00 * R[c] + t02 * R[6 + c]);
    Jp[3 + c] = -iz * (t11 * R[3 + c] + t12 * R[6 + c]);
  }
  Jc[0] = x * y / z2 * fx;
  Jc[1] = -(T(1) + (x * x / z2)) * fx;
  Jc[2] = y * iz * fx;
  Jc[3] = -iz * fx;
  Jc[4] = T(0);
  Jc[5] = x / z2 * fx;
  Jc[6] = (T(1) + y * y / z2) * fy;
  Jc[7] = -x * y / z2 * fy;
  Jc[8] = -x * iz * fy;
  Jc[9] = T(0);
  Jc[10] = -iz * fy;
  Jc[11] = y / z2 * fy;
}

// Exact Hessian of one projection edge's energy E = 0.5 rho(w |e|^2),
// e = z - proj(exp(omega, upsilon) T (X + dX)), in the variables
// v = [omega upsilon dX] at v = 0 (9x9, row-major):
//   rho' w Je^T Je + 2 rho'' w^2 (Je^T e)(Je^T e)^T - JP^T M JP - S(g)
// with Je = [Jc Jp] the edge Jacobian, Y = R X + t the camera-frame point,
// JP = dY/dv = [-[Y]x, I, R], a = rho' w e, M = sum_k a_k d2proj_k/dY2,
// g = dproj/dY^T a, and S(g) the Hessian of g . Y(v) through the second-order
// terms of exp(omega, upsilon) Y = Y + omega x Y + upsilon + 1/2 omega x
// (omega x Y) + 1/2 omega x upsilon and omega x (R dX):
//   S_ww = 1/2 (g Y^T + Y g^T) - (g . Y) I, S_wu = -1/2 [g]x, S_wX = -[g]x R.
// Also returns rho', rho'' and s for the mixed derivatives.
// With an information factor S (Omega = S^T S, see whiten_vec; null = I) the
// energy is 0.5 rho(w e^T Omega e): every e above reads Omega e where it is
// contracted, and Je^T Je reads Je^T Omega Je.
__host__ __device__ inline void edge_exact_hessian(const double* R, double fx, double fy, const double* Y,
                                                   const double* e, double w, double huber, const double* S,
                                                   double* H, double* rho_out) {
  double om[3] = {1.0, 0.0, 1.0};
  if (S) info_of(S, om);
  const double oe[2] = {om[0] * e[0] + om[1] * e[1], om[1] * e[0] + om[2] * e[1]};  // Omega e
  double rho[3];
  const double s = w * (e[0] * oe[0] + e[1] * oe[1]);
  huber_rho(s, huber, rho);
  rho_out[0] = rho[1];
  rho_out[1] = rho[2];
  rho_out[2] = s;
  double Jp[6], Jc[12];
  project_jacobians<double>(R, fx, fy, Y[0], Y[1], Y[2], Jp, Jc);
  double J[2][9];
  for (int c = 0; c < 6; ++c) { J[0][c] = Jc[c]; J[1][c] = Jc[6 + c]; }
  for (int c = 0; c < 3; ++c) { J[0][6 + c] = Jp[c]; J[1][6 + c] = Jp[3 + c]; }
  double u[9], OJ[2][9];
  for (int i = 0; i < 9; ++i) {
    u[i] = J[0][i] * oe[0] + J[1][i] * oe[1];
    OJ[0][i] = om[0] * J[0][i] + om[1] * J[1][i];
    OJ[1][i] = om[1] * J[0][i] + om[2] * J[1][i];
  }
  const double gn = rho[1] * w, hb = 2.0 * rho[2] * w * w;
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) H[9 * i + j] = gn * (J[0][i] * OJ[0][j] + J[1][i] * OJ[1][j]) + hb * u[i] * u[j];
  // second-order terms
  const double x = Y[0], y = Y[1], z = Y[2], iz = 1.0 / z, iz2 = iz * iz, iz3 = iz2 * iz;
  const double a0 = gn * oe[0], a1 = gn * oe[1];
  const double g[3] = {a0 * fx * iz, a1 * fy * iz, -(a0 * fx * x + a1 * fy * y) * iz2};
  double M[9] = {0, 0, -a0 * fx * iz2, 0, 0, -a1 * fy * iz2, -a0 * fx * iz2, -a1 * fy * iz2,
                 2.0 * (a0 * fx * x + a1 * fy * y) * iz3};
  // JP columns: omega k -> -[Y]x e_k = Y x e_k ... as a 3x9 matrix
  double JP[3][9];
  const double Yx[9] = {0, -z, y, z, 0, -x, -y, x, 0};  // [Y]x
  for (int r = 0; r < 3; ++r) {
    for (int k = 0; k < 3; ++k) {
      JP[r][k] = -Yx[3 * r + k];
      JP[r][3 + k] = (r == k) ? 1.0 : 0.0;
      JP[r][6 + k] = R[3 * r + k];
    }
  }
  double MJ[3][9];
  for (int r = 0; r < 3; ++r)
    for (int j = 0; j < 9; ++j) MJ[r][j] = M[3 * r] * JP[0][j] + M[3 * r + 1] * JP[1][j] + M[3 * r + 2] * JP[2][j];
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) H[9 * i + j] -= JP[0][i] * MJ[0][j] + JP[1][i] * MJ[1][j] + JP[2][i] * MJ[2][j];
  const double gY = g[0] * x + g[1] * y + g[2] * z;
  const double gx[9] = {0, -g[2], g[1], g[2], 0, -g[0], -g[1], g[0], 0};  // [g]x
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const double sww = 0.5 * (g[i] * Y[j] + Y[i] * g[j]) - (i == j ? gY : 0.0);
      H[9 * i + j] -= sww;
      H[9 * i + 3 + j] -= -0.5 * gx[3 * i + j];  // S_wu
      H[9 * (3 + j) + i] -= -0.5 * gx[3 * i + j];  // S_uw = S_wu^T
      double swx = 0;
      for (int k = 0; k < 3; ++k) swx -= gx[3 * i + k] * R[3 * k + j];
      H[9 * i + 6 + j] -= swx;
      H[9 * (6 + j) + i] -= swx;
    }
}

// ------------------------------------------------------------ pose graph
// EdgeSE3::computeError (edge_se3.cpp:60-66): e = toVectorMQT(Z^-1 Xi^-1 Xj)
// = [t_E; s vec(q_E)] with s = sign(w_E) (quaternion normalised to w >= 0,
// isometry3d_mappings.cpp:40-46, 80-85), and its Jacobians with respect to the
// right-multiplied increments Xi fromVectorMQT(d_i), Xj fromVectorMQT(d_j)
// (vertex_se3.h:99-108), d = [dt; dq], at d = 0. Poses are [t; q] with
// q = (x, y, z, w). With u = Rz^T dq (to first order, Z^-1 (Xi D)^-1 Xj =
// (I - 2[u]x, -Rz^T (dt + 2 dq x tz)) E and Xi^-1 (Xj D) = E D):
//   J_j = [[R_E, 0], [0, s (w_E I + [v_E]x)]]
//   J_i = [[-Rz^T, 2 ([t_E]x Rz^T + Rz^T [tz]x)], [0, s (-w_E I + [v_E]x) Rz^T]]
// Ji, Jj row-major 6x6.
__host__ __device__ inline void pg_quat_mul(const double* a, const double* b, double* o) {
  o[0] = a[3] * b[0] + a[0] * b[3] + a[1] * b[2] - a[2] * b[1];
  o[1] = a[3] * b[1] - a[0] * b[2] + a[1] * b[3] + a[2] * b[0];
  o[2] = a[3] * b[2] + a[0] * b[1] - a[1] * b[0] + a[2] * b[3];
  o[3] = a[3] * b[3] - a[0] * b[0] - a[1] * b[1] - a[2] * b[2];
}
__host__ __device__ inline void pg_quat_R(const double* q, double* R) {
  const double x = q[0], y = q[1], z = q[2], w = q[3];
  R[0] = 1 - 2 * (y * y + z * z); R[1] = 2 * (x * y - z * w); R[2] = 2 * (x * z + y * w);
  R[3] = 2 * (x * y + z * w); R[4] = 1 - 2 * (x * x + z * z); R[5] = 2 * (y * z - x * w);
  R[6] = 2 * (x * z - y * w); R[7] = 2 * (y * z + x * w); R[8] = 1 - 2 * (x * x + y * y);
}
__host__ __device__ inline void pg_edge_error(const double* Xi, const double* Xj, const double* Z, double* e,
                                              double* Ji, double* Jj) {
  const double qi_c[4] = {-Xi[3], -Xi[4], -Xi[5], Xi[6]};
  const double qz_c[4] = {-Z[3], -Z[4], -Z[5], Z[6]};
  double tmp[4], qe[4];
  pg_quat_mul(qi_c, Xj + 3, tmp);
  pg_quat_mul(qz_c, tmp, qe);
  const double nq = sqrt(qe[0] * qe[0] + qe[1] * qe[1] + qe[2] * qe[2] + qe[3] * qe[3]);
  for (int k = 0; k < 4; ++k) qe[k] /= nq;
  double Ri[9], Rz[9];
  pg_quat_R(Xi + 3, Ri);
  pg_quat_R(Z + 3, Rz);
  const double d[3] = {Xj[0] - Xi[0], Xj[1] - Xi[1], Xj[2] - Xi[2]};
  double u[3], te[3];
  for (int r = 0; r < 3; ++r) u[r] = Ri[r] * d[0] + Ri[3 + r] * d[1] + Ri[6 + r] * d[2] - Z[r];
  for (int r = 0; r < 3; ++r) te[r] = Rz[r] * u[0] + Rz[3 + r] * u[1] + Rz[6 + r] * u[2];
  const double s = qe[3] < 0 ? -1.0 : 1.0;
  for (int r = 0; r < 3; ++r) { e[r] = te[r]; e[3 + r] = s * qe[r]; }
  if (!Ji) return;
  double Re[9];
  pg_quat_R(qe, Re);
  const double w = qe[3];
  const double vx[9] = {0, -qe[2], qe[1], qe[2], 0, -qe[0], -qe[1], qe[0], 0};  // [v_E]x
  const double tx[9] = {0, -te[2], te[1], te[2], 0, -te[0], -te[1], te[0], 0};  // [t_E]x
  const double zx[9] = {0, -Z[2], Z[1], Z[2], 0, -Z[0], -Z[1], Z[0], 0};         // [tz]x
  for (int k = 0; k < 36; ++k) { Ji[k] = 0; Jj[k] = 0; }
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      Jj[6 * r + c] = Re[3 * r + c];
      Jj[6 * (3 + r) + 3 + c] = s * ((r == c ? w : 0.0) + vx[3 * r + c]);
      const double RzT = Rz[3 * c + r];
      Ji[6 * r + c] = -RzT;
      double a = 0, bq = 0, m = 0;
      for (int k = 0; k < 3; ++k) {
        a += tx[3 * r + k] * Rz[3 * c + k];  // [t_E]x Rz^T
        bq += Rz[3 * k + r] * zx[3 * k + c];  // Rz^T [tz]x
        m += ((r == k ? -w : 0.0) + vx[3 * r + k]) * Rz[3 * c + k];
      }
      Ji[6 * r + 3 + c] = 2.0 * (a + bq);
      Ji[6 * (3 + r) + 3 + c] = s * m;
    }
}

// 3x3 inverse by cofactors; symmetric input given by its 6 unique entries
// [xx xy xz yy yz zz] plus lambda on the diagonal. Returns false if singular.
__host__ __device__ inline bool inv3_sym(const double* h, double lambda, double* out) {
  double a = h[0] + lambda, b = h[1], c = h[2], d = h[3] + lambda, e = h[4], f = h[5] + lambda;
  double c0 = d * f - e * e;
  double c1 = e * c - b * f;
  double c2 = b * e - d * c;
  double det = a * c0 + b * c1 + c * c2;
  if (!(det != 0.0) || !isfinite(det)) return false;
  double id = 1.0 / det;
  out[0] = c0 * id; out[1] = c1 * id; out[2] = c2 * id;
  out[3] = c1 * id; out[4] = (a * f - c * c) * id; out[5] = (b * c - a * e) * id;
  out[6] = c2 * id; out[7] = (b * c - a * e) * id; out[8] = (a * d - b * b) * id;
  return true;
}

// Cholesky factor L L^T = A + lambda I of a symmetric 3x3 A given by its six
// unique entries [xx xy xz yy yz zz]; writes the lower-triangular L^-1 as
// [i00 i10 i11 i20 i21 i22]. Returns false unless A + lambda I is positive
// definite (and finite).
__host__ __device__ inline bool chol3_inv(const double* h, double lambda, double* m) {
  const double a00 = h[0] + lambda;
  if (!(a00 > 0.0) || !isfinite(a00)) return false;
  const double l00 = sqrt(a00), l10 = h[1] / l00, l20 = h[2] / l00;
  const double d1 = h[3] + lambda - l10 * l10;
  if (!(d1 > 0.0) || !isfinite(d1)) return false;
  const double l11 = sqrt(d1), l21 = (h[4] - l20 * l10) / l11;
  const double d2 = h[5] + lambda - l20 * l20 - l21 * l21;
  if (!(d2 > 0.0) || !isfinite(d2)) return false;
  const double l22 = sqrt(d2);
  const double i00 = 1.0 / l00, i11 = 1.0 / l11, i22 = 1.0 / l22;
  const double i10 = -l10 * i00 * i11;
  m[0] = i00;
  m[1] = i10;
  m[2] = i11;
  m[3] = -(l20 * i00 + l21 * i10) * i22;
  m[4] = -l21 * i11 * i22;
  m[5] = i22;
  return true;
}

// 6x6 inverse (row-major, in place) by Gauss-Jordan with partial pivoting.
__device__ inline bool inv6(double* m) {
  double a[6][12];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 12; ++j) a[i][j] = (j < 6) ? m[6 * i + j] : ((j - 6 == i) ? 1.0 : 0.0);
  for (int c = 0; c < 6; ++c) {
    int p = c;
    for (int r = c + 1; r < 6; ++r)
      if (fabs(a[r][c]) > fabs(a[p][c])) p = r;
    if (!(a[p][c] != 0.0) || !isfinite(a[p][c])) return false;
    if (p != c)
      for (int j = 0; j < 12; ++j) { double t = a[p][j]; a[p][j] = a[c][j]; a[c][j] = t; }
    double piv = 1.0 / a[c][c];
    for (int j = 0; j < 12; ++j) a[c][j] *= piv;
    for (int r = 0; r < 6; ++r) {
      if (r == c) continue;
      double f = a[r][c];
      for (int j = 0; j < 12; ++j) a[r][j] -= f * a[c][j];
    }
  }
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) m[6 * i + j] = a[i][6 + j];
  return true;
}

// Neumaier (compensated) accumulation of a double into (s, c).
__device__ inline void neumaier_add(double& s, double& c, double v) {
  double t = s + v;
  if (fabs(s) >= fabs(v)) c += (s - t) + v;
  else c += (v - t) + s;
  s = t;
}

}  // namespace g2ohip
