// HIP kernels of the BAL path: g2o's BAL example (g2o/examples/bal/
// bal_example.cpp:70-207) solved by BlockSolver<BlockSolverTraits<9, 3>>
// with Schur. A camera is 9 values [w (angle-axis), t, f, k1, k2] updated by
// addition (VertexCameraBAL::oplusImpl, :86-89); a point 3 values (+).
//
// Layout in HBM (FP64 throughout):
//  * state: cameras [C][9], points [P][3]; camera table [C][BAL_TAB] with
//    R, t, f, k1, k2, w and the rotation-derivative coefficients;
//  * slots = active edges, point-major (structure.cpp); per slot one int4
//    {raw camera, raw point, Hpl index q or -1, free camera or -1}, the free
//    point index per slot, measurement (2) and information weight;
//  * Hpl: one row-major 9x3 block per observation with a free camera and a
//    free point, camera-major (q order), so a camera's blocks are one run;
//  * Hpp [Cf][81], Hll 6 unique per free point, b [9 Cf + 3 Pf];
//  * Hschur: full block-CSR (both triangles) of 9x9 blocks, F_val[nF][81].
// Every reduction has a fixed order: camera sums are one workgroup's
// reduction over the camera's own observation list, point sums a thread's
// loop over the point's slots, off-diagonal Schur blocks a wavefront's loop
// over the block's observation pairs; no floating-point atomics.
#include <hip/hip_runtime.h>

#include "devmath.h"
#include "kernels.h"

namespace g2ohip {

namespace {

inline int bal_blocks(long long n, int t) { return static_cast<int>((n + t - 1) / t); }

__device__ __forceinline__ double bal_wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// upper-triangle entries (row, col) of a 9x9 block, row-major order
__constant__ int c_up9_r[45] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 2,
                                3, 3, 3, 3, 3, 3, 4, 4, 4, 4, 4, 5, 5, 5, 5, 6, 6, 6, 7, 7, 8};
__constant__ int c_up9_c[45] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 1, 2, 3, 4, 5, 6, 7, 8, 2, 3, 4, 5, 6, 7, 8,
                                3, 4, 5, 6, 7, 8, 4, 5, 6, 7, 8, 5, 6, 7, 8, 6, 7, 8, 7, 8, 8};

// One observation: e = prediction - z (bal_example.cpp:154-204) and, if
// asked, Jc = de/d[w t f k1 k2] (2x9) and Jp = de/dX (2x3), row-major.
// P = R X + t; p = -P.xy / P.z; r = 1 + k1 |p|^2 + k2 |p|^4; u = f r p.
// du/dp = f (r I + 2 (k1 + 2 k2 |p|^2) p p^T), dp/dP = -[[1, 0, px], [0, 1,
// py]] / P.z, dP/dX = R, dP/dt = I, and dP/dw the derivative of Rodrigues'
// formula R(w) X = cos(th) X + a (w x X) + b (w.X) w (a = sin(th)/th,
// b = (1 - cos(th))/th^2):
//   dP/dw = u w^T - a [X]_x + b (w X^T + (w.X) I),
//   u = -a X + c1 (w x X) + c2 (w.X) w, c1 = a'/th, c2 = b'/th,
// which at th = 0 is -[X]_x, the derivative of the reference's first-order
// branch X + w x X.
template <bool JAC>
__device__ __forceinline__ void bal_obs(const double* __restrict__ tb, const double* X, const double* z, double* e,
                                        double* Jc, double* Jp) {
  const double* R = tb;
  double P[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) P[i] = fma(R[3 * i], X[0], fma(R[3 * i + 1], X[1], fma(R[3 * i + 2], X[2], tb[9 + i])));
  const double f = tb[12], k1 = tb[13], k2 = tb[14];
  const double iz = 1.0 / P[2];
  const double px = -P[0] * iz, py = -P[1] * iz;
  const double r2 = px * px + py * py;
  const double rp = 1.0 + k1 * r2 + k2 * r2 * r2;
  e[0] = f * rp * px - z[0];
  e[1] = f * rp * py - z[1];
  if (!JAC) return;
  const double g = 2.0 * (k1 + 2.0 * k2 * r2);
  const double m00 = f * (rp + g * px * px), m01 = f * g * px * py, m11 = f * (rp + g * py * py);
  // A = du/dP (2x3)
  const double A[6] = {-m00 * iz, -m01 * iz, -(m00 * px + m01 * py) * iz,
                       -m01 * iz, -m11 * iz, -(m01 * px + m11 * py) * iz};
  // Jp = A R, Jt = A
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      Jp[3 * r + c] = A[3 * r] * R[c] + A[3 * r + 1] * R[3 + c] + A[3 * r + 2] * R[6 + c];
      Jc[9 * r + 3 + c] = A[3 * r + c];
    }
  // dP/dw
  const double w0 = tb[15], w1 = tb[16], w2 = tb[17];
  const double a = tb[18], b = tb[19], c1 = tb[20], c2 = tb[21];
  const double wx[3] = {w1 * X[2] - w2 * X[1], w2 * X[0] - w0 * X[2], w0 * X[1] - w1 * X[0]};
  const double wdx = w0 * X[0] + w1 * X[1] + w2 * X[2];
  const double wv[3] = {w0, w1, w2};
  double u[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) u[i] = -a * X[i] + c1 * wx[i] + c2 * wdx * wv[i];
  double D[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) D[3 * i + j] = u[i] * wv[j] + b * wv[i] * X[j] + (i == j ? b * wdx : 0.0);
  // - a [X]_x
  D[1] += a * X[2]; D[2] -= a * X[1];
  D[3] -= a * X[2]; D[5] += a * X[0];
  D[6] += a * X[1]; D[7] -= a * X[0];
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) Jc[9 * r + c] = A[3 * r] * D[c] + A[3 * r + 1] * D[3 + c] + A[3 * r + 2] * D[6 + c];
  Jc[6] = rp * px;       Jc[15] = rp * py;
  Jc[7] = f * r2 * px;   Jc[16] = f * r2 * py;
  Jc[8] = f * r2 * r2 * px; Jc[17] = f * r2 * r2 * py;
}

// rho' w of an observation with error e (Huber when delta > 0) and rho
__device__ __forceinline__ double bal_weight(double w, const double* e, double delta, double* rho0) {
  double rho[3];
  huber_rho(w * (e[0] * e[0] + e[1] * e[1]), delta, rho);
  *rho0 = rho[0];
  return rho[1] * w;
}

}  // namespace

// ------------------------------------------------------------ camera table
// R of Rodrigues' formula (bal_example.cpp:163-180; R = I at w = 0, the
// first-order branch), and the coefficients a, b, c1, c2 of its derivative
// (series below th = 1e-2, where the closed forms cancel).
__global__ void k_bal_cam_table(int C, const double* __restrict__ cams, double* __restrict__ tab) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  const double* v = cams + 9 * static_cast<size_t>(c);
  double* o = tab + BAL_TAB * static_cast<size_t>(c);
  const double w0 = v[0], w1 = v[1], w2 = v[2];
  const double th2 = w0 * w0 + w1 * w1 + w2 * w2;
  const double th = sqrt(th2);
  double a, b, c1, c2, cth;
  if (th < 1e-2) {
    const double t4 = th2 * th2;
    a = 1.0 - th2 / 6.0 + t4 / 120.0;
    b = 0.5 - th2 / 24.0 + t4 / 720.0;
    c1 = -1.0 / 3.0 + th2 / 30.0 - t4 / 840.0;
    c2 = -1.0 / 12.0 + th2 / 180.0 - t4 / 6720.0;
    cth = 1.0 - th2 / 2.0 + t4 / 24.0;
  } else {
    const double s = sin(th);
    cth = cos(th);
    a = s / th;
    b = (1.0 - cth) / th2;
    c1 = (th * cth - s) / (th2 * th);
    c2 = (th * s - 2.0 * (1.0 - cth)) / (th2 * th2);
  }
  if (th > 0.0) {
    // R = cos(th) I + a [w]_x + b w w^T
    o[0] = cth + b * w0 * w0;      o[1] = -a * w2 + b * w0 * w1;  o[2] = a * w1 + b * w0 * w2;
    o[3] = a * w2 + b * w1 * w0;   o[4] = cth + b * w1 * w1;      o[5] = -a * w0 + b * w1 * w2;
    o[6] = -a * w1 + b * w2 * w0;  o[7] = a * w0 + b * w2 * w1;   o[8] = cth + b * w2 * w2;
  } else {
    for (int k = 0; k < 9; ++k) o[k] = (k % 4 == 0) ? 1.0 : 0.0;
  }
  for (int k = 0; k < 3; ++k) o[9 + k] = v[3 + k];
  o[12] = v[6]; o[13] = v[7]; o[14] = v[8];
  o[15] = w0; o[16] = w1; o[17] = w2;
  o[18] = a; o[19] = b; o[20] = c1; o[21] = c2;
}

hipError_t launch_bal_cam_table(int C, const double* cams, double* tab, hipStream_t st) {
  hipLaunchKernelGGL(k_bal_cam_table, dim3(bal_blocks(C, 256)), dim3(256), 0, st, C, cams, tab);
  return hipGetLastError();
}

// ------------------------------------------------------------------ chi2
// computeActiveErrors + activeRobustChi2 (sparse_optimizer.cpp:78-130): one
// partial per workgroup of a fixed grid (launch_sum adds them in order).
constexpr int BAL_CHI2_BLOCKS = 1024;
__global__ __launch_bounds__(256) void k_bal_chi2(DevBAL g, const double* __restrict__ tab,
                                                  const double* __restrict__ pts, double* __restrict__ part,
                                                  double* __restrict__ err_out) {
  __shared__ double s_w[4];
  double chi = 0.0;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < g.ns; s += gridDim.x * blockDim.x) {
    const int4 rec = g.slot_rec[s];
    double X[3], z[2], e[2];
#pragma unroll
    for (int k = 0; k < 3; ++k) X[k] = pts[3 * static_cast<size_t>(rec.y) + k];
    z[0] = g.meas[2 * static_cast<size_t>(s)];
    z[1] = g.meas[2 * static_cast<size_t>(s) + 1];
    bal_obs<false>(tab + BAL_TAB * static_cast<size_t>(rec.x), X, z, e, nullptr, nullptr);
    double rho0;
    (void)bal_weight(g.w[s], e, g.huber, &rho0);
    chi += rho0;
    if (err_out) {
      err_out[2 * static_cast<size_t>(s)] = e[0];
      err_out[2 * static_cast<size_t>(s) + 1] = e[1];
    }
  }
  chi = bal_wave_sum(chi);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = chi;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
}

int bal_chi2_blocks() { return BAL_CHI2_BLOCKS; }

hipError_t launch_bal_chi2(const DevBAL& g, const double* tab, const double* pts, double* part, double* out,
                           double* err_out, hipStream_t st) {
  hipLaunchKernelGGL(k_bal_chi2, dim3(BAL_CHI2_BLOCKS), dim3(256), 0, st, g, tab, pts, part, err_out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_sum(part, BAL_CHI2_BLOCKS, out, st);
}

// ------------------------------------------------------------ linearisation
// constructQuadraticForm (base_fixed_sized_edge.hpp:48-130) with
// Omega' = rho' w I (robustInformation, base_edge.h:156-162), in two passes
// that each own what they sum:
//  * point pass, one thread per point position: Hll and b_l of a free point
//    from its slots in order, and the chi2 of every slot (per-workgroup
//    partials);
//  * camera pass, one workgroup per free camera: Hpp and b_p from the
//    camera's slots (threads stride the list, then wave and workgroup
//    reductions), and the Hpl block of each of its observations with a free
//    point.
__global__ __launch_bounds__(256) void k_bal_lin_points(DevBAL g, const double* __restrict__ tab,
                                                        const double* __restrict__ pts, double* __restrict__ Hll,
                                                        double* __restrict__ b, double* __restrict__ part) {
  __shared__ double s_w[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double chi = 0.0;
  if (i < g.npos) {
    const int s0 = g.pos_ptr[i], s1 = g.pos_ptr[i + 1];
    const int ph = g.pos_h[i];
    double h[6] = {0, 0, 0, 0, 0, 0}, bl[3] = {0, 0, 0};
    double X[3];
    const int praw = g.slot_rec[s0].y;
#pragma unroll
    for (int k = 0; k < 3; ++k) X[k] = pts[3 * static_cast<size_t>(praw) + k];
    for (int s = s0; s < s1; ++s) {
      const int cam = g.slot_rec[s].x;
      const double z[2] = {g.meas[2 * static_cast<size_t>(s)], g.meas[2 * static_cast<size_t>(s) + 1]};
      double e[2], Jc[18], Jp[6];
      bal_obs<true>(tab + BAL_TAB * static_cast<size_t>(cam), X, z, e, Jc, Jp);
      double rho0;
      const double om = bal_weight(g.w[s], e, g.huber, &rho0);
      chi += rho0;
      if (ph >= 0) {
        int k = 0;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = r; c < 3; ++c) h[k++] += om * (Jp[r] * Jp[c] + Jp[3 + r] * Jp[3 + c]);
#pragma unroll
        for (int r = 0; r < 3; ++r) bl[r] -= om * (Jp[r] * e[0] + Jp[3 + r] * e[1]);
      }
    }
    if (ph >= 0) {
#pragma unroll
      for (int k = 0; k < 6; ++k) Hll[6 * static_cast<size_t>(ph) + k] = h[k];
#pragma unroll
      for (int r = 0; r < 3; ++r) b[9 * static_cast<size_t>(g.Cf) + 3 * static_cast<size_t>(ph) + r] = bl[r];
    }
  }
  chi = bal_wave_sum(chi);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = chi;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
}

int bal_point_blocks(int npos) { return bal_blocks(npos > 0 ? npos : 1, 256); }

__global__ __launch_bounds__(256) void k_bal_lin_cams(DevBAL g, const double* __restrict__ tab,
                                                      const double* __restrict__ pts, double* __restrict__ Hpp,
                                                      double* __restrict__ b, double* __restrict__ Hpl) {
  __shared__ double s_acc[4][54];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int c = blockIdx.x;
  const int j0 = g.cs_ptr[c], j1 = g.cs_ptr[c + 1];
  const double* tb = tab + BAL_TAB * static_cast<size_t>(g.cam_raw[c]);
  double acc[54];
#pragma unroll
  for (int k = 0; k < 54; ++k) acc[k] = 0.0;
  for (int j = j0 + threadIdx.x; j < j1; j += 256) {
    const int s = g.cs_slot[j];
    const int4 rec = g.slot_rec[s];
    double X[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) X[k] = pts[3 * static_cast<size_t>(rec.y) + k];
    const double z[2] = {g.meas[2 * static_cast<size_t>(s)], g.meas[2 * static_cast<size_t>(s) + 1]};
    double e[2], Jc[18], Jp[6];
    bal_obs<true>(tb, X, z, e, Jc, Jp);
    double rho0;
    const double om = bal_weight(g.w[s], e, g.huber, &rho0);
#pragma unroll
    for (int k = 0; k < 45; ++k) {
      const int r = c_up9_r[k], cc = c_up9_c[k];
      acc[k] += om * (Jc[r] * Jc[cc] + Jc[9 + r] * Jc[9 + cc]);
    }
#pragma unroll
    for (int r = 0; r < 9; ++r) acc[45 + r] -= om * (Jc[r] * e[0] + Jc[9 + r] * e[1]);
    if (rec.z >= 0) {
      double* o = Hpl + 27 * static_cast<size_t>(rec.z);
#pragma unroll
      for (int r = 0; r < 9; ++r)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) o[3 * r + cc] = om * (Jc[r] * Jp[cc] + Jc[9 + r] * Jp[3 + cc]);
    }
  }
#pragma unroll
  for (int k = 0; k < 54; ++k) acc[k] = bal_wave_sum(acc[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 54; ++k) s_acc[wv][k] = acc[k];
  }
  __syncthreads();
  const int t = threadIdx.x;
  if (t < 54) {
    const double v = (s_acc[0][t] + s_acc[1][t]) + (s_acc[2][t] + s_acc[3][t]);
    if (t < 45) {
      const int r = c_up9_r[t], cc = c_up9_c[t];
      Hpp[81 * static_cast<size_t>(c) + 9 * r + cc] = v;
      Hpp[81 * static_cast<size_t>(c) + 9 * cc + r] = v;
    } else {
      b[9 * static_cast<size_t>(c) + (t - 45)] = v;
    }
  }
}

hipError_t launch_bal_linearize(const DevBAL& g, const double* tab, const double* pts, double* Hpp, double* Hll,
                                double* Hpl, double* b, double* part, double* chi_out, hipStream_t st) {
  const int nb = bal_point_blocks(g.npos);
  hipLaunchKernelGGL(k_bal_lin_points, dim3(nb), dim3(256), 0, st, g, tab, pts, Hll, b, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bal_lin_cams, dim3(g.Cf), dim3(256), 0, st, g, tab, pts, Hpp, b, Hpl);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  return launch_sum(part, nb, chi_out, st);
}

// ------------------------------------------------------------- LM scalars
// computeLambdaInit (optimization_algorithm_levenberg.cpp:155-167): the
// largest |diagonal entry| of Hpp and Hll; a grid-stride pass writes one
// maximum per workgroup, one workgroup takes the maximum of those.
constexpr int BAL_MAX_BLOCKS = 256;
__device__ __forceinline__ double bal_block_max(double m, double* s_m) {
  s_m[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) s_m[threadIdx.x] = fmax(s_m[threadIdx.x], s_m[threadIdx.x + s]);
    __syncthreads();
  }
  return s_m[0];
}
__global__ __launch_bounds__(256) void k_bal_max_diag(DevBAL g, const double* __restrict__ Hpp,
                                                      const double* __restrict__ Hll, double* __restrict__ part) {
  __shared__ double s_m[256];
  double m = 0.0;
  const long long stride = static_cast<long long>(gridDim.x) * 256;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < 9LL * g.Cf; i += stride)
    m = fmax(m, fabs(Hpp[81 * (i / 9) + 10 * (i % 9)]));
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < 3LL * g.Pf; i += stride) {
    const int k = static_cast<int>(i % 3);
    m = fmax(m, fabs(Hll[6 * (i / 3) + (k == 0 ? 0 : (k == 1 ? 3 : 5))]));
  }
  m = bal_block_max(m, s_m);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}
__global__ __launch_bounds__(256) void k_bal_max_final(const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double s_m[256];
  const double m = bal_block_max(threadIdx.x < BAL_MAX_BLOCKS ? part[threadIdx.x] : 0.0, s_m);
  if (threadIdx.x == 0) *out = m;
}

hipError_t launch_bal_max_diag(const DevBAL& g, const double* Hpp, const double* Hll, double* part, double* out,
                               hipStream_t st) {
  hipLaunchKernelGGL(k_bal_max_diag, dim3(BAL_MAX_BLOCKS), dim3(256), 0, st, g, Hpp, Hll, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bal_max_final, dim3(1), dim3(256), 0, st, part, out);
  return hipGetLastError();
}

// ------------------------------------------------------- Schur complement
// (Hll + lambda I)^-1 per free point, 9 values row-major (BlockSolver::solve
// inverts with Eigen's closed form, block_solver.hpp:371), and L^-1 of its
// Cholesky factor L L^T = Hll + lambda I, 6 values [i00 i10 i11 i20 i21 i22];
// err |= 1 unless positive definite.
__global__ void k_bal_point_dinv(int Pf, const double* __restrict__ Hll, double lambda, double* __restrict__ Dinv,
                                 double* __restrict__ Linv, int* __restrict__ err) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= Pf) return;
  double h[6], out[9], li[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) h[k] = Hll[6 * static_cast<size_t>(l) + k];
  const bool ok = chol3_inv(h, lambda, li) && inv3_sym(h, lambda, out);
  if (!ok) {
    atomicOr(err, 1);
#pragma unroll
    for (int k = 0; k < 9; ++k) out[k] = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) li[k] = 0.0;
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) Dinv[9 * static_cast<size_t>(l) + k] = out[k];
#pragma unroll
  for (int k = 0; k < 6; ++k) Linv[6 * static_cast<size_t>(l) + k] = li[k];
}

// Diagonal blocks and reduced right-hand side, one workgroup per free camera
// over its observations (block_solver.hpp:345-450), in the Gram form: with
// G = W L^-T (W the 9x3 Hpl block, L L^T = Hll + lambda I of the point),
// S_ii = Hpp_ii + lambda I - sum G G^T and bschur_i = b_i - sum G (L^-1 b_l)
// (W Dinv W^T = G G^T, W Dinv b_l = G L^-1 b_l). With G_out the G blocks are
// also stored for the off-diagonal products.
__global__ __launch_bounds__(256) void k_bal_schur_diag(DevBAL g, const double* __restrict__ Hpp,
                                                        const double* __restrict__ Hpl, const double* __restrict__ Linv,
                                                        const double* __restrict__ b, double lambda,
                                                        double* __restrict__ F_val, double* __restrict__ bschur,
                                                        double* __restrict__ G_out) {
  __shared__ double s_acc[4][54];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int c = blockIdx.x;
  double acc[54];
#pragma unroll
  for (int k = 0; k < 54; ++k) acc[k] = 0.0;
  const size_t boff = 9 * static_cast<size_t>(g.Cf);
  for (int q = g.cam_qptr[c] + threadIdx.x; q < g.cam_qptr[c + 1]; q += 256) {
    const int ph = g.q_ph[q];
    double W[27], L[6], bl[3];
#pragma unroll
    for (int k = 0; k < 27; ++k) W[k] = Hpl[27 * static_cast<size_t>(q) + k];
#pragma unroll
    for (int k = 0; k < 6; ++k) L[k] = Linv[6 * static_cast<size_t>(ph) + k];
#pragma unroll
    for (int k = 0; k < 3; ++k) bl[k] = b[boff + 3 * static_cast<size_t>(ph) + k];
    const double y[3] = {L[0] * bl[0], L[1] * bl[0] + L[2] * bl[1], L[3] * bl[0] + L[4] * bl[1] + L[5] * bl[2]};
    double G[27];
#pragma unroll
    for (int r = 0; r < 9; ++r) {
      G[3 * r] = W[3 * r] * L[0];
      G[3 * r + 1] = W[3 * r] * L[1] + W[3 * r + 1] * L[2];
      G[3 * r + 2] = W[3 * r] * L[3] + W[3 * r + 1] * L[4] + W[3 * r + 2] * L[5];
    }
    if (G_out) {
#pragma unroll
      for (int k = 0; k < 27; ++k) G_out[27 * static_cast<size_t>(q) + k] = G[k];
    }
#pragma unroll
    for (int k = 0; k < 45; ++k) {
      const int r = c_up9_r[k], cc = c_up9_c[k];
      acc[k] += G[3 * r] * G[3 * cc] + G[3 * r + 1] * G[3 * cc + 1] + G[3 * r + 2] * G[3 * cc + 2];
    }
#pragma unroll
    for (int r = 0; r < 9; ++r) acc[45 + r] += G[3 * r] * y[0] + G[3 * r + 1] * y[1] + G[3 * r + 2] * y[2];
  }
#pragma unroll
  for (int k = 0; k < 54; ++k) acc[k] = bal_wave_sum(acc[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 54; ++k) s_acc[wv][k] = acc[k];
  }
  __syncthreads();
  const int t = threadIdx.x;
  if (t < 54) {
    const double v = (s_acc[0][t] + s_acc[1][t]) + (s_acc[2][t] + s_acc[3][t]);
    if (t < 45) {
      const int r = c_up9_r[t], cc = c_up9_c[t];
      const double h = Hpp[81 * static_cast<size_t>(c) + 9 * r + cc] + (r == cc ? lambda : 0.0) - v;
      double* S = F_val + 81 * static_cast<size_t>(g.F_diag[c]);
      S[9 * r + cc] = h;
      S[9 * cc + r] = h;
    } else {
      const int r = t - 45;
      bschur[9 * static_cast<size_t>(c) + r] = b[9 * static_cast<size_t>(c) + r] - v;
    }
  }
}

// Workgroups go to the 8 XCDs round robin; the blocks are taken so that each
// XCD works through a contiguous range of them (rows in order), whose rows
// share cameras and points and so hit in that XCD's L2. The grid is 8 x
// per workgroups, per = the workgroups of four blocks an XCD takes.
__device__ __forceinline__ int bal_xcd_block(int nO) {
  const int per = (nO + 31) / 32;
  return (blockIdx.x & 7) * per + (blockIdx.x >> 3);
}
inline int bal_xcd_grid(int nO) { return 8 * ((nO + 31) / 32); }

// Off-diagonal blocks, one wavefront per off-diagonal upper block (i < j):
// S_ij = - sum over the observation pairs (a by camera i, b by camera j, of
// one point) of W_a Dinv W_b^T, the pairs in a fixed order. Lane t forms
// entry t (and entry 64 + t for t < 17) of the row-major 9x9 block; the
// result goes to both triangles of the full block rows.
__global__ __launch_bounds__(256) void k_bal_schur_offdiag(DevBAL g, const double* __restrict__ Hpl,
                                                           const double* __restrict__ Dinv,
                                                           double* __restrict__ F_val) {
  const int lane = threadIdx.x & 63;
  const int k = bal_xcd_block(g.nO) * 4 + (threadIdx.x >> 6);
  if (k >= g.nO) return;
  const int e0 = lane, e1 = 64 + (lane < 17 ? lane : 0);
  const int r0 = e0 / 9, c0 = e0 % 9, r1 = e1 / 9, c1 = e1 % 9;
  double v0 = 0.0, v1 = 0.0;
  for (int p = g.od_ptr[k]; p < g.od_ptr[k + 1]; ++p) {
    const int2 qq = g.od_pair[p];
    const int ph = g.q_ph[qq.x];
    const double* Wa = Hpl + 27 * static_cast<size_t>(qq.x);
    const double* Wb = Hpl + 27 * static_cast<size_t>(qq.y);
    const double* D = Dinv + 9 * static_cast<size_t>(ph);
    double d[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) d[t] = D[t];
    const double a0[3] = {Wa[3 * r0], Wa[3 * r0 + 1], Wa[3 * r0 + 2]};
    const double bb0[3] = {Wb[3 * c0], Wb[3 * c0 + 1], Wb[3 * c0 + 2]};
    const double a1[3] = {Wa[3 * r1], Wa[3 * r1 + 1], Wa[3 * r1 + 2]};
    const double bb1[3] = {Wb[3 * c1], Wb[3 * c1 + 1], Wb[3 * c1 + 2]};
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      const double u0 = a0[0] * d[t] + a0[1] * d[3 + t] + a0[2] * d[6 + t];
      const double u1 = a1[0] * d[t] + a1[1] * d[3 + t] + a1[2] * d[6 + t];
      v0 = fma(u0, bb0[t], v0);
      v1 = fma(u1, bb1[t], v1);
    }
  }
  const int blk = g.od_blk[k];
  double* S = F_val + 81 * static_cast<size_t>(g.up_pos[blk]);
  double* ST = F_val + 81 * static_cast<size_t>(g.up_posT[blk]);
  S[e0] = -v0;
  ST[9 * c0 + r0] = -v0;
  if (lane < 17) {
    S[e1] = -v1;
    ST[9 * c1 + r1] = -v1;
  }
}

// The same blocks on the matrix cores. With G = W L^-T (9x3 per observation,
// L^-1 of its point's factor; k_bal_schur_diag stores it), S_ij = - sum
// over the pairs of G_a G_b^T is one GEMM of K = 3 x pairs per block,
// accumulated by v_mfma_f64_16x16x4f64 (FP64 in and out). One wavefront per
// off-diagonal upper block, one pair per instruction: lane t holds
// A[t % 16][t / 16] = G_a[row][k] and B[t / 16][t % 16] = G_b[col][k] for
// row / col t % 16 < 9 and k = t / 16 < 3, zero elsewhere (9 x 9 x 3 of the
// 16 x 16 x 4 product is used). The accumulator holds D[t / 16 + 4 r][t % 16]
// in its register r. The pair list is read 64 pairs at a time, one per lane,
// and each pair's observation indices reach the wavefront by readlane.
typedef double bal_d4 __attribute__((ext_vector_type(4)));
constexpr int BAL_MFMA_UNROLL = 16;
__global__ __launch_bounds__(256) void k_bal_schur_offdiag_mfma(DevBAL g, const double* __restrict__ G,
                                                                double* __restrict__ F_val) {
  const int lane = threadIdx.x & 63;
  const int k = bal_xcd_block(g.nO) * 4 + (threadIdx.x >> 6);
  if (k >= g.nO) return;  // uniform per wavefront
  const int ri = lane & 15, kc = lane >> 4;
  const bool on = ri < 9 && kc < 3;
  const int off = on ? 3 * ri + kc : 0;  // the lane's entry of a 9x3 block
  bal_d4 acc = {0.0, 0.0, 0.0, 0.0};
  const int p0 = g.od_ptr[k], p1 = g.od_ptr[k + 1];
  for (int c0 = p0; c0 < p1; c0 += 64) {
    const int n = min(64, p1 - c0);
    const int2 mine = g.od_pair[c0 + (lane < n ? lane : 0)];
    for (int u0 = 0; u0 < n; u0 += BAL_MFMA_UNROLL) {
      double ga[BAL_MFMA_UNROLL], gb[BAL_MFMA_UNROLL];
#pragma unroll
      for (int u = 0; u < BAL_MFMA_UNROLL; ++u) {
        const int src = min(u0 + u, n - 1);
        const int qa = __builtin_amdgcn_readlane(mine.x, src);
        const int qb = __builtin_amdgcn_readlane(mine.y, src);
        ga[u] = G[27 * static_cast<size_t>(qa) + off];
        gb[u] = G[27 * static_cast<size_t>(qb) + off];
      }
#pragma unroll
      for (int u = 0; u < BAL_MFMA_UNROLL; ++u) {
        if (u0 + u >= n) break;  // uniform per wavefront
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(on ? ga[u] : 0.0, on ? gb[u] : 0.0, acc, 0, 0, 0);
      }
    }
  }
  const int blk = g.od_blk[k];
  double* S = F_val + 81 * static_cast<size_t>(g.up_pos[blk]);
  double* ST = F_val + 81 * static_cast<size_t>(g.up_posT[blk]);
  const int col = ri;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = kc + 4 * r;
    if (row < 9 && col < 9) {
      S[9 * row + col] = -acc[r];
      ST[9 * col + row] = -acc[r];
    }
  }
}

hipError_t launch_bal_schur_diag(const DevBAL& g, const double* Hpp, const double* Hll, const double* Hpl,
                                 const double* b, double lambda, double* Dinv, double* Linv, double* G,
                                 double* F_val, double* bschur, int* err, hipStream_t st) {
  hipLaunchKernelGGL(k_bal_point_dinv, dim3(bal_blocks(g.Pf > 0 ? g.Pf : 1, 256)), dim3(256), 0, st, g.Pf, Hll,
                     lambda, Dinv, Linv, err);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bal_schur_diag, dim3(g.Cf), dim3(256), 0, st, g, Hpp, Hpl, Linv, b, lambda, F_val, bschur, G);
  return hipGetLastError();
}

hipError_t launch_bal_schur_offdiag(const DevBAL& g, const double* Hpl, const double* Dinv, const double* G,
                                    double* F_val, bool mfma, hipStream_t st) {
  if (g.nO == 0) return hipSuccess;
  if (mfma)
    hipLaunchKernelGGL(k_bal_schur_offdiag_mfma, dim3(bal_xcd_grid(g.nO)), dim3(256), 0, st, g, G, F_val);
  else
    hipLaunchKernelGGL(k_bal_schur_offdiag, dim3(bal_xcd_grid(g.nO)), dim3(256), 0, st, g, Hpl, Dinv, F_val);
  return hipGetLastError();
}

// ------------------------------------------------------ back-substitution
// x_l = Dinv (b_l - sum W^T x_p) per free point (block_solver.hpp:452-476),
// one thread per point position, its observations in slot order.
__global__ __launch_bounds__(256) void k_bal_backsub(DevBAL g, const double* __restrict__ Hpl,
                                                     const double* __restrict__ Dinv, const double* __restrict__ b,
                                                     double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.npos) return;
  const int ph = g.pos_h[i];
  if (ph < 0) return;
  const size_t boff = 9 * static_cast<size_t>(g.Cf);
  double cl[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) cl[k] = b[boff + 3 * static_cast<size_t>(ph) + k];
  for (int s = g.pos_ptr[i]; s < g.pos_ptr[i + 1]; ++s) {
    const int4 rec = g.slot_rec[s];
    if (rec.z < 0) continue;
    const double* W = Hpl + 27 * static_cast<size_t>(rec.z);
    const double* xp = x + 9 * static_cast<size_t>(rec.w);
#pragma unroll
    for (int r = 0; r < 9; ++r) {
      const double xr = xp[r];
#pragma unroll
      for (int c = 0; c < 3; ++c) cl[c] -= W[3 * r + c] * xr;
    }
  }
  const double* D = Dinv + 9 * static_cast<size_t>(ph);
#pragma unroll
  for (int r = 0; r < 3; ++r) x[boff + 3 * static_cast<size_t>(ph) + r] = D[3 * r] * cl[0] + D[3 * r + 1] * cl[1] + D[3 * r + 2] * cl[2];
}

hipError_t launch_bal_backsub(const DevBAL& g, const double* Hpl, const double* Dinv, const double* b, double* x,
                              hipStream_t st) {
  if (g.npos == 0) return hipSuccess;
  hipLaunchKernelGGL(k_bal_backsub, dim3(bal_blocks(g.npos, 256)), dim3(256), 0, st, g, Hpl, Dinv, b, x);
  return hipGetLastError();
}

// ------------------------------------------------------------------ update
// VertexCameraBAL / VertexPointBAL::oplusImpl (bal_example.cpp:86-89,
// 115-118): x added to the free vertices' values.
__global__ void k_bal_update(DevBAL g, const double* __restrict__ x, double* __restrict__ cams,
                             double* __restrict__ pts) {
  const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < g.C) {
    const int h = g.camH[t];
    if (h < 0) return;
#pragma unroll
    for (int k = 0; k < 9; ++k) cams[9 * t + k] += x[9 * static_cast<size_t>(h) + k];
  } else if (t < static_cast<long long>(g.C) + g.P) {
    const long long p = t - g.C;
    const int h = g.ptH[p];
    if (h < 0) return;
    const size_t off = 9 * static_cast<size_t>(g.Cf) + 3 * static_cast<size_t>(h);
#pragma unroll
    for (int k = 0; k < 3; ++k) pts[3 * p + k] += x[off + k];
  }
}

hipError_t launch_bal_update(const DevBAL& g, const double* x, double* cams, double* pts, hipStream_t st) {
  hipLaunchKernelGGL(k_bal_update, dim3(bal_blocks(static_cast<long long>(g.C) + g.P, 256)), dim3(256), 0, st, g, x,
                     cams, pts);
  return hipGetLastError();
}

}  // namespace g2ohip
