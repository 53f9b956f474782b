// Exact-Hessian kernels: the implicit-function adjoint of a converged bundle
// adjustment (no counterpart in g2o, which stops at the solution) and the
// undamped Newton step that drives an LM solution to the stationary point the
// adjoint differentiates. FP64 throughout.
#include <hip/hip_runtime.h>

#include "devmath.h"
#include "kernels.h"

namespace g2ohip {

namespace {

__device__ __forceinline__ int lower_bound_dev(const int* a, int n, int key) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

inline int nblk(long long n, int t) { return static_cast<int>((n + t - 1) / t); }

}  // namespace

// Exact Hessian of F/2 = sum_e 0.5 rho_e (edge_exact_hessian) assembled like
// k_linearize: Hpl per observation at its camera-major position, Hll per
// point, Hpp per (chunk, camera) partial (27 entries, the 6 right-hand-side
// entries left at zero). The gradient b is not touched.
__global__ __launch_bounds__(LIN_THREADS) void k_linearize_exact(DevGraph g, const double* __restrict__ camRt,
                                                                const double* __restrict__ pts,
                                                                double* __restrict__ Hpl, double* __restrict__ Hll,
                                                                double* __restrict__ hpp_part) {
  // camera accumulators (27 per chunk camera) and the chunk's camera list
  extern __shared__ double s_dyn[];
  double* const s_cacc = s_dyn;
  int* const s_cam = reinterpret_cast<int*>(s_dyn + 27 * g.lc_maxc);
  __shared__ double s_pacc[LIN_PTS * 6];
  const int ch = blockIdx.x, tid = threadIdx.x;
  const int p0 = g.lc_pos0[ch], p1 = g.lc_pos0[ch + 1];
  const int s0 = g.pos_ptr[p0], s1 = g.pos_ptr[p1];
  const int c0 = g.lc_cam_ptr[ch], nc = g.lc_cam_ptr[ch + 1] - c0;
  const int np = p1 - p0;
  for (int i = tid; i < nc; i += LIN_THREADS) s_cam[i] = g.lc_cams[c0 + i];
  for (int i = tid; i < nc * 27; i += LIN_THREADS) s_cacc[i] = 0.0;
  for (int i = tid; i < np * 6; i += LIN_THREADS) s_pacc[i] = 0.0;
  __syncthreads();
  for (int s = s0 + tid; s < s1; s += LIN_THREADS) {
    const int pos = g.slot_pos[s], lp = pos - p0;
    const int ph = g.pos_h[pos];
    const int cam = g.slot_cam[s], chh = g.camH[cam];
    const double* Rt = camRt + 12 * cam;
    const double* in = g.intr + 4 * cam;
    const double X[3] = {pts[3 * pos], pts[3 * pos + 1], pts[3 * pos + 2]};  // position-ordered state
    const double z[2] = {g.slot_meas[2 * s], g.slot_meas[2 * s + 1]};
    double Y[3], e[2], H[81], rho[3];
    project_error(Rt, in, X, z, Y, e);
    const double* S = g.slot_S ? g.slot_S + 3 * static_cast<size_t>(s) : nullptr;
    edge_exact_hessian(Rt, in[0], in[1], Y, e, g.slot_w[s], g.huber, S, H, rho);
    if (ph >= 0) {
      double* pa = s_pacc + 6 * lp;
      int k = 0;
      for (int r = 0; r < 3; ++r)
        for (int c = r; c < 3; ++c) atomicAdd(&pa[k++], H[9 * (6 + r) + 6 + c]);
    }
    if (chh >= 0) {
      double* ca = s_cacc + 27 * lower_bound_dev(s_cam, nc, chh);
      int k = 0;
      for (int r = 0; r < 6; ++r)
        for (int c = r; c < 6; ++c) atomicAdd(&ca[k++], H[9 * r + c]);
      if (ph >= 0) {
        double* h = Hpl + 18 * static_cast<size_t>(g.slot_q[s]);
        for (int r = 0; r < 6; ++r)
          for (int c = 0; c < 3; ++c) h[3 * r + c] = H[9 * r + 6 + c];
      }
    }
  }
  __syncthreads();
  for (int lp = tid; lp < np; lp += LIN_THREADS) {
    const int ph = g.pos_h[p0 + lp];
    if (ph < 0) continue;
    for (int k = 0; k < 6; ++k) Hll[6 * static_cast<size_t>(ph) + k] = s_pacc[6 * lp + k];
  }
  for (int i = tid; i < nc * 27; i += LIN_THREADS) hpp_part[static_cast<size_t>(c0) * 27 + i] = s_cacc[i];
}

hipError_t launch_linearize_exact(const DevGraph& g, const double* camRt, const double* pts, double* Hpl,
                                  double* Hll, double* hpp_part, hipStream_t st) {
  if (g.nlc == 0) return hipSuccess;
  const size_t lds = (27 * sizeof(double) + sizeof(int)) * static_cast<size_t>(g.lc_maxc);
  if (!lds_fits(k_linearize_exact, lds)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_linearize_exact, dim3(g.nlc), dim3(LIN_THREADS), lds, st, g, camRt, pts, Hpl, Hll, hpp_part);
  return hipGetLastError();
}

// Back-substitution from the stored blocks (block_solver.hpp:452-476):
// x_l = Dinv_l (b_l - sum Hpl^T x_c), one thread per point.
template <typename HT>
__global__ void k_backsub_blocks(DevGraph g, const HT* __restrict__ Hpl, const HT* __restrict__ Dinv,
                                 const double* __restrict__ b, double* __restrict__ x) {
  const int pos = blockIdx.x * blockDim.x + threadIdx.x;
  if (pos >= g.npos) return;
  const int ph = g.pos_h[pos];
  if (ph < 0) return;
  const size_t off = 6 * static_cast<size_t>(g.Cf) + 3 * static_cast<size_t>(ph);
  double v[3] = {b[off], b[off + 1], b[off + 2]};
  for (int s = g.pos_ptr[pos]; s < g.pos_ptr[pos + 1]; ++s) {
    const int chh = g.slot_camh[s];
    if (chh < 0) continue;
    const HT* h = Hpl + 18 * static_cast<size_t>(g.slot_q[s]);
    const double* xc = x + 6 * static_cast<size_t>(chh);
    for (int c = 0; c < 3; ++c)
      for (int r = 0; r < 6; ++r) v[c] -= static_cast<double>(h[3 * r + c]) * xc[r];
  }
  const HT* D = Dinv + 9 * static_cast<size_t>(ph);
  for (int r = 0; r < 3; ++r)
    x[off + r] = static_cast<double>(D[3 * r]) * v[0] + static_cast<double>(D[3 * r + 1]) * v[1] +
                 static_cast<double>(D[3 * r + 2]) * v[2];
}

template <typename HT>
hipError_t launch_backsub_blocks(const DevGraph& g, const HT* Hpl, const HT* Dinv, const double* b, double* x,
                                 hipStream_t st) {
  if (g.npos == 0) return hipSuccess;
  hipLaunchKernelGGL(k_backsub_blocks<HT>, dim3(nblk(g.npos, 256)), dim3(256), 0, st, g, Hpl, Dinv, b, x);
  return hipGetLastError();
}
template hipError_t launch_backsub_blocks<double>(const DevGraph&, const double*, const double*, const double*,
                                                  double*, hipStream_t);

// Adjoint outputs per slot from lam = H^-1 dL/dtheta: with v = Je lam (the
// edge's camera and point parts; zero for fixed vertices) and the edge's
// information Omega (identity unless set),
//   dL/dz = -(rho' w Omega v + 2 rho'' w^2 Omega e (Omega e . v)),
//   dL/dw = -(rho' + rho'' s)(Omega e . v),  s = w e^T Omega e
// (the mixed derivatives d grad / dz and d grad / dw of the edge energy).
__global__ __launch_bounds__(256) void k_adjoint_out(DevGraph g, const double* __restrict__ camRt,
                                                     const double* __restrict__ pts, const double* __restrict__ lam,
                                                     double* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= g.ns) return;
  const int pos = g.slot_pos[s];
  const int ph = g.pos_h[pos];
  const int cam = g.slot_cam[s], chh = g.camH[cam];
  const double* Rt = camRt + 12 * cam;
  const double* in = g.intr + 4 * cam;
  const double X[3] = {pts[3 * pos], pts[3 * pos + 1], pts[3 * pos + 2]};  // position-ordered state
  const double z[2] = {g.slot_meas[2 * s], g.slot_meas[2 * s + 1]};
  const double w = g.slot_w[s];
  double Y[3], e[2];
  project_error(Rt, in, X, z, Y, e);
  double Jp[6], Jc[12];
  project_jacobians<double>(Rt, in[0], in[1], Y[0], Y[1], Y[2], Jp, Jc);
  double v[2] = {0, 0};
  if (chh >= 0) {
    const double* lc = lam + 6 * static_cast<size_t>(chh);
    for (int k = 0; k < 6; ++k) { v[0] += Jc[k] * lc[k]; v[1] += Jc[6 + k] * lc[k]; }
  }
  if (ph >= 0) {
    const double* ll = lam + 6 * static_cast<size_t>(g.Cf) + 3 * static_cast<size_t>(ph);
    for (int k = 0; k < 3; ++k) { v[0] += Jp[k] * ll[k]; v[1] += Jp[3 + k] * ll[k]; }
  }
  double om[3] = {1.0, 0.0, 1.0};
  if (g.slot_S) info_of(g.slot_S + 3 * static_cast<size_t>(s), om);
  const double oe[2] = {om[0] * e[0] + om[1] * e[1], om[1] * e[0] + om[2] * e[1]};
  const double ov[2] = {om[0] * v[0] + om[1] * v[1], om[1] * v[0] + om[2] * v[1]};
  double rho[3];
  const double sq = w * (e[0] * oe[0] + e[1] * oe[1]);
  huber_rho(sq, g.huber, rho);
  const double ev = oe[0] * v[0] + oe[1] * v[1];
  const double hb = 2.0 * rho[2] * w * w * ev;
  out[3 * static_cast<size_t>(s)] = -(rho[1] * w * ov[0] + hb * oe[0]);
  out[3 * static_cast<size_t>(s) + 1] = -(rho[1] * w * ov[1] + hb * oe[1]);
  out[3 * static_cast<size_t>(s) + 2] = -(rho[1] + rho[2] * sq) * ev;
}

hipError_t launch_adjoint_out(const DevGraph& g, const double* camRt, const double* pts, const double* lam,
                              double* out, hipStream_t st) {
  if (g.ns == 0) return hipSuccess;
  hipLaunchKernelGGL(k_adjoint_out, dim3(nblk(g.ns, 256)), dim3(256), 0, st, g, camRt, pts, lam, out);
  return hipGetLastError();
}

}  // namespace g2ohip
