// HIP kernels of the g2o LM linear-system path on gfx950.
//
// Layout in HBM (all built once by structure.cpp, see structure.h):
//  * slots = active edges, point-major: slot_cam / slot_pos (int32),
//    slot_meas (2 x f64), slot_w (f64);
//  * Hpl: one row-major 6x3 block per observation with a free camera and a
//    free point, in camera-major order (q index: camera by camera, points in
//    processing order), Hpl[18 * q + 3 * row + col]: the blocks one Hschur
//    block row needs are contiguous runs;
//  * Hll: 6 unique f64 per free point; b: f64 [6 Cf + 3 Pf] (g2o's layout);
//  * Hschur: full symmetric block-CSR (both triangles) F_val[nF][36].
// Camera-side sums of the linearisation and of the matrix-free passes go
// through per-chunk partials that a second kernel sums in a fixed order; the
// Schur complement is formed by the workgroup that owns each block row.
#include <hip/hip_runtime.h>

#include "devmath.h"
#include "kernels.h"

namespace g2ohip {

namespace {

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Sum of a block's per-thread values (blockDim.x multiple of 64, <= 1024).
__device__ double block_sum(double v, double* s_red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane == 0) s_red[w] = v;
  __syncthreads();
  double t = 0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < nw; ++i) t += s_red[i];
    s_red[0] = t;
  }
  __syncthreads();
  t = s_red[0];
  __syncthreads();
  return t;
}

// Compensated sum of part[0..n) computed identically by every block.
__device__ double block_sum_array(const double* __restrict__ part, int n, double* s_red) {
  double s = 0, c = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) neumaier_add(s, c, part[i]);
  // combine (s, c) pairs in a fixed tree
  double* ss = s_red;
  double* cc = s_red + blockDim.x;
  ss[threadIdx.x] = s;
  cc[threadIdx.x] = c;
  __syncthreads();
  for (int stride = blockDim.x / 2; stride > 0; stride >>= 1) {
    if (threadIdx.x < stride) {
      double a = ss[threadIdx.x], b = ss[threadIdx.x + stride];
      double t = a + b;
      double e = (fabs(a) >= fabs(b)) ? ((a - t) + b) : ((b - t) + a);
      ss[threadIdx.x] = t;
      cc[threadIdx.x] += cc[threadIdx.x + stride] + e;
    }
    __syncthreads();
  }
  double r = ss[0] + cc[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ int lower_bound_i(const int* a, int n, int key) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Sum of v over the run of consecutive lanes that share this lane's key (the
// keys of one run sit in consecutive lanes): a segmented inclusive scan, after
// which the last lane of each run holds the run's total; returns true there.
// Every lane of the wavefront must take part (inactive lanes pass a key of
// their own).
template <int N>
__device__ __forceinline__ bool run_sum(int key, double (&v)[N]) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int k2 = __shfl_up(key, off, 64);
    const bool take = lane >= off && k2 == key;
#pragma unroll
    for (int n = 0; n < N; ++n) {
      const double t = __shfl_up(v[n], off, 64);
      v[n] = take ? v[n] + t : v[n];
    }
  }
  const int kn = __shfl_down(key, 1, 64);
  return lane == 63 || kn != key;
}

__constant__ int c_up6_r[21] = {0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 4, 4, 5};
__constant__ int c_up6_c[21] = {0, 1, 2, 3, 4, 5, 1, 2, 3, 4, 5, 2, 3, 4, 5, 3, 4, 5, 4, 5, 5};

inline int nblocks(long long n, int t) { return static_cast<int>((n + t - 1) / t); }

}  // namespace

// =================================================================== state
__global__ void k_cam_table(int C, const double* __restrict__ cams7, double* __restrict__ camRt,
                            int* __restrict__ err4) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (err4 && c < 4) err4[c] = 0;
  if (c >= C) return;
  const double* p = cams7 + 7 * c;
  double q[4] = {p[3], p[4], p[5], p[6]};
  double R[9];
  quat_to_rot(q, R);
  double* o = camRt + 12 * c;
  for (int i = 0; i < 9; ++i) o[i] = R[i];
  o[9] = p[0]; o[10] = p[1]; o[11] = p[2];
}

hipError_t launch_cam_table(const DevGraph& g, const double* cams7, double* camRt, hipStream_t st, int* err4) {
  hipLaunchKernelGGL(k_cam_table, dim3(nblocks(g.C > 4 ? g.C : 4, 256)), dim3(256), 0, st, g.C, cams7, camRt, err4);
  return hipGetLastError();
}

// computeActiveErrors + activeRobustChi2 (sparse_optimizer.cpp:78-130):
// per-slot error, robust chi2, fixed-grid partial sums.
__global__ __launch_bounds__(256) void k_chi2(DevGraph g, const double* __restrict__ camRt,
                                              const double* __restrict__ pts, double* __restrict__ partial,
                                              double* __restrict__ err_out) {
  __shared__ double s_red[4];
  double chi = 0;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < g.ns; s += gridDim.x * blockDim.x) {
    const int4 rec = g.slot_rec[s];
    const int cam = rec.z, pp = rec.x;  // the point's index in the position-ordered state
    double X[3] = {pts[3 * pp], pts[3 * pp + 1], pts[3 * pp + 2]};
    double z[2] = {g.slot_meas[2 * s], g.slot_meas[2 * s + 1]};
    double Xc[3], e[2];
    project_error(camRt + 12 * cam, g.intr + 4 * cam, X, z, Xc, e);
    if (err_out) { err_out[2 * s] = e[0]; err_out[2 * s + 1] = e[1]; }
    if (g.slot_S) whiten_vec(g.slot_S + 3 * static_cast<size_t>(s), e);
    double rho[3];
    huber_rho(g.slot_w[s] * (e[0] * e[0] + e[1] * e[1]), g.huber, rho);
    chi += rho[0];
  }
  chi = block_sum(chi, s_red);
  if (threadIdx.x == 0) partial[blockIdx.x] = chi;
}

__global__ __launch_bounds__(256) void k_sum(const double* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ double s_red[512];
  double v = block_sum_array(part, n, s_red);
  if (threadIdx.x == 0) *out = v;
}

hipError_t launch_sum(const double* partial, int n, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, st, partial, n, out);
  return hipGetLastError();
}

hipError_t launch_chi2(const DevGraph& g, const double* camRt, const double* pts, double* partial, double* out,
                       double* err_out, hipStream_t st) {
  hipLaunchKernelGGL(k_chi2, dim3(CHI2_BLOCKS), dim3(256), 0, st, g, camRt, pts, partial, err_out);
  return launch_sum(partial, CHI2_BLOCKS, out, st);
}

// SparseOptimizer::update (sparse_optimizer.cpp:457-468): exp-map left update
// of free cameras (vertex_se3_expmap.cpp:48-51) and additive point update
// (vertex_pointxyz.h:47-50), FP64.
__global__ void k_update(DevGraph g, const double* __restrict__ dx, double* __restrict__ cams7,
                         double* __restrict__ pts) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < g.C) {
    int h = g.camH[t];
    if (h < 0) return;
    double in[7], out[7];
    for (int i = 0; i < 7; ++i) in[i] = cams7[7 * t + i];
    se3_left_update(dx + 6 * h, in, out);
    for (int i = 0; i < 7; ++i) cams7[7 * t + i] = out[i];
  } else if (t < g.C + g.P) {
    // the point state is in position order: index p holds position p's point
    int p = t - g.C;
    int h = p < g.npos ? g.pos_h[p] : -1;
    if (h < 0) return;
    const double* d = dx + 6 * static_cast<size_t>(g.Cf) + 3 * static_cast<size_t>(h);
    pts[3 * p] += d[0];
    pts[3 * p + 1] += d[1];
    pts[3 * p + 2] += d[2];
  }
}

hipError_t launch_update(const DevGraph& g, const double* dx, double* cams7, double* pts, hipStream_t st) {
  hipLaunchKernelGGL(k_update, dim3(nblocks(static_cast<long long>(g.C) + g.P, 256)), dim3(256), 0, st, g, dx,
                     cams7, pts);
  return hipGetLastError();
}

hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t st) {
  return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st);
}

// ============================================================ linearisation

// Operands of one observation slot: its indices and everything its error and
// Jacobians need. Every load is unconditional (callers pass a valid slot),
// so a thread can have the operands of several slots in flight at once
// instead of walking one slot's chain of dependent loads
// (slot -> point / camera -> coordinates) at a time.
struct SlotOps {
  int pos, cam, ph, chh, q;
  double Rt[12], in[4], X[3], z[2], w;
  double S[3];  // information factor (identity without g.slot_S)
};
// e and the Jacobians of one observation, whitened by its information factor
__device__ __forceinline__ void whiten_obs(const DevGraph& g, const double* S, double* e, double* Jp, double* Jc) {
  if (!g.slot_S) return;
  whiten_vec(S, e);
  whiten_rows<3>(S, Jp);
  whiten_rows<6>(S, Jc);
}
__device__ __forceinline__ void load_slot_ops(const DevGraph& g, const double* __restrict__ camRt,
                                              const double* __restrict__ pts, int s, SlotOps& o) {
  // {position, free point index, camera, free camera index} in one load
  const int4 rec = g.slot_rec[s];
  o.pos = rec.x;
  o.ph = rec.y;
  o.cam = rec.z;
  o.chh = rec.w;
  o.q = g.slot_q[s];
  o.z[0] = g.slot_meas[2 * static_cast<size_t>(s)];
  o.z[1] = g.slot_meas[2 * static_cast<size_t>(s) + 1];
  o.w = g.slot_w[s];
  if (g.slot_S) {
#pragma unroll
    for (int k = 0; k < 3; ++k) o.S[k] = g.slot_S[3 * static_cast<size_t>(s) + k];
  } else {
    o.S[0] = 1.0; o.S[1] = 0.0; o.S[2] = 1.0;
  }
#pragma unroll
  for (int k = 0; k < 12; ++k) o.Rt[k] = camRt[12 * static_cast<size_t>(o.cam) + k];
#pragma unroll
  for (int k = 0; k < 4; ++k) o.in[k] = g.intr[4 * static_cast<size_t>(o.cam) + k];
#pragma unroll
  for (int k = 0; k < 3; ++k) o.X[k] = pts[3 * static_cast<size_t>(o.pos) + k];  // position-ordered state
}
// Slots whose operands one thread of a linearisation chunk loads together:
// two for the linearisation; one (all of a slot's loads before its first
// branch) where more would cost occupancy (the residual and Schur-diagonal
// passes hold more registers per slot).
constexpr int LIN_PF = 2;
constexpr int LIN_PF1 = 1;
// BlockSolver::buildSystem (block_solver.hpp:492-565) for EdgeSE3ProjectXYZ:
// errors, Huber weights, Jacobians (edge_project_xyz.cpp:57-93) and the
// Hll / Hpp / b sums in FP64, the stored Hpl blocks in HT; constructQuadraticForm
// (base_fixed_sized_edge.hpp:48-130) with Omega' = rho' w I (robustInformation,
// base_edge.h:156-162) and b -= ... as +J^T omega_r, omega_r = -rho' w e.
// One workgroup per chunk of consecutive points: the slot's Hpl block goes to
// its own slot; Hll / b_l are summed per point and Hpp / b_p per chunk camera
// with LDS atomics (f64), then written once.
template <typename HT>
__global__ __launch_bounds__(LIN_THREADS) void k_linearize(DevGraph g, const double* __restrict__ camRt,
                                                          const double* __restrict__ pts, HT* __restrict__ Hpl,
                                                          double* __restrict__ Hll, double* __restrict__ b,
                                                          double* __restrict__ hpp_part,
                                                          double* __restrict__ chunk_chi) {
  __shared__ double s_red[LIN_THREADS / 64];
  // camera accumulators sized for the plan's largest chunk (27 per camera),
  // then the point accumulators and the chunk's camera list: a smaller
  // footprint keeps more workgroups resident per CU
  extern __shared__ double s_dyn[];
  double* const s_cacc = s_dyn;
  double* const s_pacc = s_dyn + 27 * g.lc_maxc;
  int* const s_cam = reinterpret_cast<int*>(s_pacc + 9 * LIN_PTS);
  const int ch = blockIdx.x, tid = threadIdx.x;
  const int p0 = g.lc_pos0[ch], p1 = g.lc_pos0[ch + 1];
  const int s0 = g.pos_ptr[p0], s1 = g.pos_ptr[p1];
  const int c0 = g.lc_cam_ptr[ch], nc = g.lc_cam_ptr[ch + 1] - c0;
  const int np = p1 - p0;
  for (int i = tid; i < nc; i += LIN_THREADS) s_cam[i] = g.lc_cams[c0 + i];
  for (int i = tid; i < nc * 27; i += LIN_THREADS) s_cacc[i] = 0.0;
  for (int i = tid; i < np * 9; i += LIN_THREADS) s_pacc[i] = 0.0;
  __syncthreads();
  double chi = 0;
  // point-major slots: the lanes of one point are consecutive, so its Hll / b_l
  // terms are reduced across the run before one LDS atomic per run and entry;
  // the camera-side terms go to the chunk's camera accumulators directly
  for (int base = s0; base < s1; base += LIN_PF * LIN_THREADS) {
    SlotOps op[LIN_PF];
#pragma unroll
    for (int u = 0; u < LIN_PF; ++u) load_slot_ops(g, camRt, pts, min(base + tid + u * LIN_THREADS, s1 - 1), op[u]);
#pragma unroll
    for (int u = 0; u < LIN_PF; ++u) {
      const bool valid = base + tid + u * LIN_THREADS < s1;
      const SlotOps& o = op[u];
      const int lp = o.pos - p0, ph = o.ph, chh = o.chh;
      const double* Rt = o.Rt;
      const double* in = o.in;
      const double w = o.w;
      double Xc[3], e[2];
      project_error(Rt, in, o.X, o.z, Xc, e);
      // Jacobians and the Hll / Hpp / b sums in FP64 for either HT; only the
      // stored Hpl block is rounded to HT (the FP32 path keeps Hpp and Hll
      // accurate, which its FP64 Schur diagonal relies on)
      double Jp[6], Jc[12];
      project_jacobians<double>(Rt, in[0], in[1], Xc[0], Xc[1], Xc[2], Jp, Jc);
      whiten_obs(g, o.S, e, Jp, Jc);
      double rho[3];
      huber_rho(w * (e[0] * e[0] + e[1] * e[1]), g.huber, rho);
      const double om = rho[1] * w;
      const double wr0 = -om * e[0], wr1 = -om * e[1];
      if (valid) {
        chi += rho[0];
        if (chh >= 0) {
          double* ca = s_cacc + 27 * lower_bound_i(s_cam, nc, chh);
          for (int k = 0; k < 21; ++k) {
            const int r = c_up6_r[k], c = c_up6_c[k];
            atomicAdd(&ca[k], om * (Jc[r] * Jc[c] + Jc[6 + r] * Jc[6 + c]));
          }
          for (int r = 0; r < 6; ++r) atomicAdd(&ca[21 + r], Jc[r] * wr0 + Jc[6 + r] * wr1);
          if (ph >= 0) {
            // the 18 entries as 9 two-element stores (a block is 8-byte aligned)
            HT v[18];
            for (int r = 0; r < 6; ++r)
              for (int c = 0; c < 3; ++c)
                v[3 * r + c] = static_cast<HT>(om * (Jc[r] * Jp[c] + Jc[6 + r] * Jp[3 + c]));
            using HT2 = typename std::conditional<sizeof(HT) == 4, float2, double2>::type;
            HT2* h2 = reinterpret_cast<HT2*>(Hpl + 18 * static_cast<size_t>(o.q));
#pragma unroll
            for (int k = 0; k < 9; ++k) h2[k] = HT2{v[2 * k], v[2 * k + 1]};
          }
        }
      }
      const bool pt_on = valid && ph >= 0;
      double pv[9];
      {
        int k = 0;
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = r; c < 3; ++c) pv[k++] = pt_on ? om * (Jp[r] * Jp[c] + Jp[3 + r] * Jp[3 + c]) : 0.0;
#pragma unroll
        for (int r = 0; r < 3; ++r) pv[6 + r] = pt_on ? Jp[r] * wr0 + Jp[3 + r] * wr1 : 0.0;
      }
      const int key = pt_on ? lp : -1 - (threadIdx.x & 63);
      if (run_sum<9>(key, pv) && pt_on) {
        double* pa = s_pacc + 9 * lp;
#pragma unroll
        for (int k = 0; k < 9; ++k) atomicAdd(&pa[k], pv[k]);
      }
    }
  }
  chi = block_sum(chi, s_red);
  if (tid == 0) chunk_chi[ch] = chi;
  // block_sum ends with a barrier, so the LDS accumulators are complete
  const size_t boff = 6 * static_cast<size_t>(g.Cf);
  for (int lp = tid; lp < np; lp += LIN_THREADS) {
    const int ph = g.pos_h[p0 + lp];
    if (ph < 0) continue;
    for (int k = 0; k < 6; ++k) Hll[6 * static_cast<size_t>(ph) + k] = s_pacc[9 * lp + k];
    for (int r = 0; r < 3; ++r) b[boff + 3 * static_cast<size_t>(ph) + r] = s_pacc[9 * lp + 6 + r];
  }
  for (int i = tid; i < nc * 27; i += LIN_THREADS) hpp_part[static_cast<size_t>(c0) * 27 + i] = s_cacc[i];
}

template <typename HT>
hipError_t launch_linearize(const DevGraph& g, const double* camRt, const double* pts, HT* Hpl, double* Hll,
                            double* b, double* hpp_part, double* chunk_chi, hipStream_t st) {
  if (g.nlc == 0) return hipSuccess;
  const size_t lds = sizeof(double) * (27 * static_cast<size_t>(g.lc_maxc) + 9 * LIN_PTS) +
                     sizeof(int) * static_cast<size_t>(g.lc_maxc);
  if (!lds_fits(k_linearize<HT>, lds)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_linearize<HT>, dim3(g.nlc), dim3(LIN_THREADS), lds, st, g, camRt, pts, Hpl, Hll, b, hpp_part,
                     chunk_chi);
  return hipGetLastError();
}
template hipError_t launch_linearize<float>(const DevGraph&, const double*, const double*, float*, double*, double*,
                                            double*, double*, hipStream_t);
template hipError_t launch_linearize<double>(const DevGraph&, const double*, const double*, double*, double*,
                                             double*, double*, double*, hipStream_t);

// Sum over free camera c's chunk partial rows of entry k (rows of `stride`
// doubles): four running sums keep four loads in flight; their fixed
// combination keeps the result deterministic.
__device__ __forceinline__ double camp_sum(const DevGraph& g, const double* __restrict__ part, int stride, int c,
                                           int k) {
  double v4[4] = {0.0, 0.0, 0.0, 0.0};
  const int q0 = g.camp_ptr[c], q1 = g.camp_ptr[c + 1];
  int q = q0;
  for (; q + 4 <= q1; q += 4)
#pragma unroll
    for (int j = 0; j < 4; ++j) v4[j] += part[static_cast<size_t>(g.camp_idx[q + j]) * stride + k];
  for (int j = 0; q < q1; ++q, ++j) v4[j] += part[static_cast<size_t>(g.camp_idx[q]) * stride + k];
  return (v4[0] + v4[1]) + (v4[2] + v4[3]);
}

// Hpp and b_p (b == null: Hpp only): sum of the chunk partials of each free
// camera, fixed order.
__global__ void k_reduce_hpp(DevGraph g, const double* __restrict__ hpp_part, double* __restrict__ Hpp,
                             double* __restrict__ b) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.Cf * 27) return;
  const int c = t / 27, k = t % 27;
  const double v = camp_sum(g, hpp_part, 27, c, k);
  if (k < 21) {
    const int r = c_up6_r[k], cc = c_up6_c[k];
    Hpp[36 * static_cast<size_t>(c) + 6 * r + cc] = v;
    Hpp[36 * static_cast<size_t>(c) + 6 * cc + r] = v;
  } else if (b) {
    b[6 * static_cast<size_t>(c) + (k - 21)] = v;
  }
}

hipError_t launch_reduce_hpp(const DevGraph& g, const double* hpp_part, double* Hpp, double* b, hipStream_t st) {
  hipLaunchKernelGGL(k_reduce_hpp, dim3(nblocks(static_cast<long long>(g.Cf) * 27, 256)), dim3(256), 0, st, g,
                     hpp_part, Hpp, b);
  return hipGetLastError();
}

// computeLambdaInit (optimization_algorithm_levenberg.cpp:155-167): max |diag|.
__global__ __launch_bounds__(256) void k_max_diag(DevGraph g, const double* __restrict__ Hpp,
                                                  const double* __restrict__ Hll, double* __restrict__ partial) {
  __shared__ double s_red[4];
  double m = 0;
  const long long n = 6LL * g.Cf + 3LL * g.Pf;
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    double v;
    if (i < 6LL * g.Cf) v = Hpp[36 * (i / 6) + 7 * (i % 6)];
    else {
      long long j = i - 6LL * g.Cf;
      const int dmap[3] = {0, 3, 5};
      v = Hll[6 * (j / 3) + dmap[j % 3]];
    }
    m = fmax(m, fabs(v));
  }
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int i = 0; i < 4; ++i) t = fmax(t, s_red[i]);
    partial[blockIdx.x] = t;
  }
}

__global__ void k_max_final(const double* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ double s_red[4];
  double m = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, part[i]);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int i = 0; i < 4; ++i) t = fmax(t, s_red[i]);
    *out = t;
  }
}

hipError_t launch_max_diag(const DevGraph& g, const double* Hpp, const double* Hll, double* partial, double* out,
                           hipStream_t st) {
  hipLaunchKernelGGL(k_max_diag, dim3(RED_BLOCKS), dim3(256), 0, st, g, Hpp, Hll, partial);
  hipLaunchKernelGGL(k_max_final, dim3(1), dim3(256), 0, st, partial, RED_BLOCKS, out);
  return hipGetLastError();
}

// computeScale (optimization_algorithm_levenberg.cpp:169-176)
__global__ __launch_bounds__(256) void k_scale(int n, const double* __restrict__ x, const double* __restrict__ b,
                                               double lambda, double* __restrict__ partial) {
  __shared__ double s_red[4];
  double s = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    s += x[i] * (lambda * x[i] + b[i]);
  s = block_sum(s, s_red);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

hipError_t launch_scale(int n, const double* x, const double* b, double lambda, double* partial, double* out,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_scale, dim3(RED_BLOCKS), dim3(256), 0, st, n, x, b, lambda, partial);
  return launch_sum(partial, RED_BLOCKS, out, st);
}

// ========================================================== explicit Schur
// BlockSolver::solve, Schur part (block_solver.hpp:345-433):
// Hschur(i, j) = Hpp(i, i) [i == j] + lambda I [i == j]
//              - sum over free points p seen by i and j of Hpl_ip Dinv_p Hpl_jp^T,
// Dinv_p = (Hll_p + lambda I)^-1, and bschur_i = b_i - sum_p Hpl_ip Dinv_p b_p.
// Off-diagonal blocks (i < j) come from point windows: with L_p L_p^T =
// Hll_p + lambda I and G = Hpl L_p^-T, the point's term is G_ip G_jp^T. A
// workgroup stages its window's G blocks in LDS once (every observation is
// read from HBM once per solve) and one lane forms each partial sum of up
// to SCHUR_WIN_PCH pair products in registers, no atomics; a second pass sums
// each block's partials in a fixed order. Diagonal blocks and bschur are
// formed per camera in FP64 (k_schur_diag).
template <typename HT>
struct GStride;
template <>
struct GStride<float> { static constexpr int v = 20; };  // 80-byte rows: 16-byte aligned
template <>
struct GStride<double> { static constexpr int v = 18; };

__device__ __forceinline__ void lds_load18(const float* p, float* o) {
  const float4* p4 = reinterpret_cast<const float4*>(p);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float4 v = p4[k];
    o[4 * k] = v.x; o[4 * k + 1] = v.y; o[4 * k + 2] = v.z; o[4 * k + 3] = v.w;
  }
  const float2 v = reinterpret_cast<const float2*>(p)[8];
  o[16] = v.x; o[17] = v.y;
}
__device__ __forceinline__ void lds_load18(const double* p, double* o) {
  const double2* p2 = reinterpret_cast<const double2*>(p);
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const double2 v = p2[k];
    o[2 * k] = v.x; o[2 * k + 1] = v.y;
  }
}
__device__ __forceinline__ void store36(float* o, const float* a) {
  float4* o4 = reinterpret_cast<float4*>(o);
#pragma unroll
  for (int k = 0; k < 9; ++k) o4[k] = make_float4(a[4 * k], a[4 * k + 1], a[4 * k + 2], a[4 * k + 3]);
}
__device__ __forceinline__ void store36(double* o, const double* a) {
  double2* o2 = reinterpret_cast<double2*>(o);
#pragma unroll
  for (int k = 0; k < 18; ++k) o2[k] = make_double2(a[2 * k], a[2 * k + 1]);
}

// Vector chunk of an Hpl block for the staging copy: 9 chunks per block.
template <typename HT>
struct Chunk;
template <>
struct Chunk<float> { using T = float2; static constexpr int per_block = 9; };   // 72-byte block
template <>
struct Chunk<double> { using T = double2; static constexpr int per_block = 9; };  // 144-byte block

// Multiply-accumulate of one pair: acc += Ga Gb^T (6x3 times 3x6).
template <typename HT>
__device__ __forceinline__ void pair_mac(const HT* __restrict__ sG, unsigned pr, HT* acc) {
  constexpr int GS = GStride<HT>::v;
  HT Ga[18], Gb[18];
  lds_load18(sG + (pr & 0xFFFFu) * GS, Ga);
  lds_load18(sG + (pr >> 16) * GS, Gb);
  // three fused multiply-adds into the accumulator per entry (108 per pair)
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      HT a = acc[6 * i + j];
      a = fma(Ga[3 * i], Gb[3 * j], a);
      a = fma(Ga[3 * i + 1], Gb[3 * j + 1], a);
      acc[6 * i + j] = fma(Ga[3 * i + 2], Gb[3 * j + 2], a);
    }
}

// L^-1 (lower triangular, [i00 i10 i11 i20 i21 i22]) of L L^T = Hll + lambda I
// per free point; err |= 1 unless positive definite.
__global__ __launch_bounds__(256) void k_point_linv(int Pf, const double* __restrict__ Hll, double lambda,
                                                    double* __restrict__ Linv, int* __restrict__ err) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= Pf) return;
  double m[6];
  if (!chol3_inv(Hll + 6 * static_cast<size_t>(p), lambda, m)) {
    atomicOr(err, 1);
    for (int k = 0; k < 6; ++k) m[k] = 0.0;
  }
  for (int k = 0; k < 6; ++k) Linv[6 * static_cast<size_t>(p) + k] = m[k];
}

hipError_t launch_point_linv(const DevGraph& g, const double* Hll, double lambda, double* Linv, int* err,
                             hipStream_t st) {
  if (g.Pf == 0) return hipSuccess;
  hipLaunchKernelGGL(k_point_linv, dim3(nblocks(g.Pf, 256)), dim3(256), 0, st, g.Pf, Hll, lambda, Linv, err);
  return hipGetLastError();
}

template <typename HT, int OBS>
__global__ __launch_bounds__(256) void k_schur_win(DevGraph g, DevWinPlan wp, const HT* __restrict__ Hpl,
                                                  const double* __restrict__ Linv, HT* __restrict__ part) {
  constexpr int GS = GStride<HT>::v;
  using CT = typename Chunk<HT>::T;
  constexpr int NCH = Chunk<HT>::per_block;
  __shared__ __align__(16) HT s_G[OBS * GS];
  // q of each window observation: in the padding of the FP32 rows (keeps the
  // FP32 footprint at 40 KiB, four workgroups per CU), else its own array
  constexpr bool QPAD = GS * sizeof(HT) >= 18 * sizeof(HT) + sizeof(int);
  __shared__ int s_qa[QPAD ? 1 : OBS];
  auto s_q = [&](int l) -> int& {
    if constexpr (QPAD) return *reinterpret_cast<int*>(s_G + l * GS + 18);
    else return s_qa[l];
  };
  const int w = blockIdx.x, tid = threadIdx.x;
  const int o0 = wp.win_obs0[w], nob = wp.win_obs0[w + 1] - o0;
  if (nob <= 0) return;
  // Each staging phase issues all of a thread's loads before its first LDS
  // write, from clamped (valid) indices: the q indices share the LDS array
  // with the rows being written, so a load-then-write loop would wait for
  // every load in turn.
  constexpr int UQ = OBS / 256;  // observations per thread
  {
    int qv[UQ];
#pragma unroll
    for (int u = 0; u < UQ; ++u) qv[u] = wp.win_q[o0 + min(tid + 256 * u, nob - 1)];
#pragma unroll
    for (int u = 0; u < UQ; ++u)
      if (tid + 256 * u < nob) s_q(tid + 256 * u) = qv[u];
  }
  __syncthreads();
  // the window's Hpl blocks in 8/16-byte chunks (chunk e = (l, k) = (e / NCH,
  // e % NCH)); its observations ascend in q, so runs of one camera are
  // contiguous in HBM
  if constexpr (!QPAD) {
    // q in its own array: the unrolled loop keeps several loads in flight
    // (chunk e walked incrementally: 256 = 28 NCH + 4 for NCH = 9)
    static_assert(NCH == 9, "the staging walk assumes 9 chunks per block");
    int l = tid / NCH, k = tid - NCH * (tid / NCH);
#pragma unroll 6
    for (int e = tid; e < nob * NCH; e += 256) {
      const CT v = reinterpret_cast<const CT*>(Hpl + 18 * static_cast<size_t>(s_q(l)))[k];
      reinterpret_cast<CT*>(s_G + l * GS)[k] = v;
      l += 28;
      k += 4;
      if (k >= NCH) { k -= NCH; ++l; }
    }
  } else {
    constexpr int UE = (OBS * NCH + 255) / 256;  // chunks per thread
    const int ne = nob * NCH;
    int qv[UE];
#pragma unroll
    for (int u = 0; u < UE; ++u) qv[u] = s_q(min(tid + 256 * u, ne - 1) / NCH);
    CT v[UE];
#pragma unroll
    for (int u = 0; u < UE; ++u) {
      const int e = min(tid + 256 * u, ne - 1);
      v[u] = reinterpret_cast<const CT*>(Hpl + 18 * static_cast<size_t>(qv[u]))[e - NCH * (e / NCH)];
    }
#pragma unroll
    for (int u = 0; u < UE; ++u) {
      const int e = tid + 256 * u;
      if (e < ne) reinterpret_cast<CT*>(s_G + (e / NCH) * GS)[e - NCH * (e / NCH)] = v[u];
    }
  }
  __syncthreads();
  // G = Hpl L^-T with the point's L^-1 (k_point_linv)
  {
    int ph[UQ];
#pragma unroll
    for (int u = 0; u < UQ; ++u) ph[u] = g.q_ph[s_q(min(tid + 256 * u, nob - 1))];
    double m[UQ][6];
#pragma unroll
    for (int u = 0; u < UQ; ++u)
#pragma unroll
      for (int k = 0; k < 6; ++k) m[u][k] = Linv[6 * static_cast<size_t>(ph[u]) + k];
#pragma unroll
    for (int u = 0; u < UQ; ++u) {
      const int l = tid + 256 * u;
      if (l >= nob) continue;
      HT* G = s_G + l * GS;
      for (int r = 0; r < 6; ++r) {
        const double h0 = G[3 * r], h1 = G[3 * r + 1], h2 = G[3 * r + 2];
        G[3 * r] = static_cast<HT>(h0 * m[u][0]);
        G[3 * r + 1] = static_cast<HT>(h0 * m[u][1] + h1 * m[u][2]);
        G[3 * r + 2] = static_cast<HT>(h0 * m[u][3] + h1 * m[u][4] + h2 * m[u][5]);
      }
    }
  }
  __syncthreads();
  const int wv = tid >> 6, lane = tid & 63;
  const int ge = wp.win_grp0[w + 1];
  constexpr unsigned NONE = 0xFFFFFFFFu;
  for (int grp = wp.win_grp0[w] + wv; grp < ge; grp += 4) {
    const int len = wp.grp_len[grp];
    const unsigned* rows = wp.pair_rows + 64 * static_cast<size_t>(wp.grp_row0[grp]) + lane;
    const int slot = wp.grp_slot[64 * static_cast<size_t>(grp) + lane];
    HT acc[36];
#pragma unroll
    for (int k = 0; k < 36; ++k) acc[k] = HT(0);
    // pair entries four rows at a time, the next four in flight while the
    // current ones are multiplied
    unsigned cur[4], nxt[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) cur[j] = j < len ? rows[64 * static_cast<size_t>(j)] : NONE;
    for (int r0 = 0; r0 < len; r0 += 4) {
#pragma unroll
      for (int j = 0; j < 4; ++j) nxt[j] = r0 + 4 + j < len ? rows[64 * static_cast<size_t>(r0 + 4 + j)] : NONE;
      if (cur[0] == NONE) break;  // this lane's list has ended
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (cur[j] != NONE) pair_mac<HT>(s_G, cur[j], acc);
#pragma unroll
      for (int j = 0; j < 4; ++j) cur[j] = nxt[j];
    }
    if (slot >= 0) store36(part + 36 * static_cast<size_t>(slot), acc);
  }
}

template <typename HT>
__global__ __launch_bounds__(256) void k_schur_offdiag(DevGraph g, DevWinPlan wp, const HT* __restrict__ part,
                                                      HT* __restrict__ F_val) {
  const long long t = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  const int u = static_cast<int>(t / 36), e = static_cast<int>(t % 36);
  if (u >= wp.nod) return;
  // four independent running sums keep four loads in flight; their fixed
  // combination order keeps the result deterministic
  double s4[4] = {0.0, 0.0, 0.0, 0.0};
  const int s0 = wp.od_ptr[u], s1 = wp.od_ptr[u + 1];
  int sl = s0;
  for (; sl + 4 <= s1; sl += 4)
#pragma unroll
    for (int j = 0; j < 4; ++j) s4[j] += static_cast<double>(part[36 * static_cast<size_t>(sl + j) + e]);
  for (int j = 0; sl < s1; ++sl, ++j) s4[j] += static_cast<double>(part[36 * static_cast<size_t>(sl) + e]);
  const HT v = static_cast<HT>(-((s4[0] + s4[1]) + (s4[2] + s4[3])));
  const int k = wp.od_blk[u];
  F_val[36 * static_cast<size_t>(g.up_pos[k]) + e] = v;
  F_val[36 * static_cast<size_t>(g.up_posT[k]) + (e % 6) * 6 + e / 6] = v;
}

template <typename HT>
hipError_t launch_schur_win(const DevGraph& g, const DevWinPlan& wp, const HT* Hpl, const double* Linv, HT* part,
                            hipStream_t st) {
  if (wp.nwin == 0) return hipSuccess;
  if (wp.obs_max <= 256)
    hipLaunchKernelGGL((k_schur_win<HT, 256>), dim3(wp.nwin), dim3(256), 0, st, g, wp, Hpl, Linv, part);
  else if (wp.obs_max <= 512)
    hipLaunchKernelGGL((k_schur_win<HT, 512>), dim3(wp.nwin), dim3(256), 0, st, g, wp, Hpl, Linv, part);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}
template <typename HT>
hipError_t launch_schur_offdiag(const DevGraph& g, const DevWinPlan& wp, const HT* part, HT* F_val, hipStream_t st) {
  if (wp.nod == 0) return hipSuccess;
  const long long n = 36LL * wp.nod;
  hipLaunchKernelGGL(k_schur_offdiag<HT>, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, g, wp,
                     part, F_val);
  return hipGetLastError();
}
template hipError_t launch_schur_win<float>(const DevGraph&, const DevWinPlan&, const float*, const double*, float*,
                                            hipStream_t);
template hipError_t launch_schur_win<double>(const DevGraph&, const DevWinPlan&, const double*, const double*,
                                             double*, hipStream_t);
template hipError_t launch_schur_offdiag<float>(const DevGraph&, const DevWinPlan&, const float*, float*, hipStream_t);
template hipError_t launch_schur_offdiag<double>(const DevGraph&, const DevWinPlan&, const double*, double*,
                                                 hipStream_t);

// ================================================= row-owner Schur blocks
// The off-diagonal Schur blocks (a, b), a < b, one workgroup per camera row
// a (or per column chunk of a long row): it walks camera a's observations
// (a, p) in batches, forms G_a = Hpl_ap L_p^-T once per observation, then
// takes every later free camera b of point p - the point's free slots are
// sorted by camera, so they are a contiguous run after (a, p) - and adds
// G_a G_b^T into an LDS accumulator for block (a, b). Nothing is staged in
// HBM, every observation of a track of n cameras is touched by the n - 1
// rows it pairs with, and a track of any length costs n (n - 1) / 2 block
// products. The LDS sums are FP64 atomic adds (their order, and so the
// last bits of the blocks, may vary from run to run).
// 16 wavefronts per workgroup: a long row's accumulators fill most of a
// CU's LDS, so one workgroup per CU has to bring the waves that hide the
// gathers' latency
constexpr int SR_THREADS = 1024;
constexpr int SR_OBS = 256;  // observations per batch (threads 0..255 set them up)
static_assert(SR_OBS % 64 == 0 && SR_OBS <= SR_THREADS, "batch setup: one thread per observation");

// Exclusive prefix sum of SR_OBS counts by wavefront 0 (four per lane);
// pre[SR_OBS] = total.
__device__ __forceinline__ void sr_prefix(const int* __restrict__ cnt, int* __restrict__ pre) {
  constexpr int PL = SR_OBS / 64;
  const int lane = threadIdx.x & 63;
  if (threadIdx.x < 64) {
    int c[PL];
    int v = 0;
#pragma unroll
    for (int i = 0; i < PL; ++i) {
      c[i] = cnt[PL * lane + i];
      v += c[i];
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int o = __shfl_up(v, off, 64);
      if (lane >= off) v += o;
    }
    int ex = v;
#pragma unroll
    for (int i = 0; i < PL; ++i) ex -= c[i];
#pragma unroll
    for (int i = 0; i < PL; ++i) {
      pre[PL * lane + i] = ex;
      ex += c[i];
    }
    if (lane == 63) pre[SR_OBS] = v;
  }
}

template <typename HT>
__global__ __launch_bounds__(SR_THREADS) void k_schur_rows(DevGraph g, const int4* __restrict__ items, int nitems,
                                                           const int* __restrict__ q_slot,
                                                           const HT* __restrict__ Hpl,
                                                           const double* __restrict__ Linv,
                                                           HT* __restrict__ F_val) {
  extern __shared__ double s_dyn[];  // [nblk * 36] accumulators, then [nblk] column cameras
  __shared__ HT s_Ga[SR_OBS][18];
  __shared__ double s_L[SR_OBS][6];
  __shared__ int s_lo[SR_OBS], s_cnt[SR_OBS], s_pre[SR_OBS + 1];
  // consecutive rows on one XCD (workgroups go to the 8 XCDs round robin):
  // neighbouring rows share points, so their Hpl reads hit the same L2
  const int per = (nitems + 7) / 8;
  const int item = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (item >= nitems) return;
  const int4 it = items[item];
  const int a = it.x, k0 = it.y, k1 = it.z;
  const bool full = it.w != 0;
  const int nblk = k1 - k0;
  double* s_acc = s_dyn;
  int* s_col = reinterpret_cast<int*>(s_dyn + 36 * static_cast<size_t>(nblk));
  const int tid = threadIdx.x;
  for (int e = tid; e < 36 * nblk; e += SR_THREADS) s_acc[e] = 0.0;
  for (int j = tid; j < nblk; j += SR_THREADS) s_col[j] = g.S_col[k0 + j];
  const int cb_lo = g.S_col[k0], cb_hi = g.S_col[k1 - 1];
  const int qa0 = g.cam_qptr[a], qa1 = g.cam_qptr[a + 1];
  for (int qb = qa0; qb < qa1; qb += SR_OBS) {
    __syncthreads();  // the previous batch's pairs are done (and the zeroing above)
    if (tid < SR_OBS) {
      const int q = qb + tid;
      const bool valid = q < qa1;
      const int qc = valid ? q : qa0;
      const int s = q_slot[qc];
      const int ph = g.q_ph[qc];
      const int pos = g.slot_pos[s];
      HT h[18];
#pragma unroll
      for (int k = 0; k < 18; ++k) h[k] = Hpl[18 * static_cast<size_t>(qc) + k];
      double m[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) m[k] = Linv[6 * static_cast<size_t>(ph) + k];
      const int se = g.pos_ptr[pos] + g.pos_nfree[pos];
      int lo = s + 1;
      while (lo < se && g.slot_camh[lo] == a) ++lo;  // a second edge to camera a
      int hi = se;
      if (!full) {
        // the chunk's cameras cb_lo..cb_hi: a sub-run of the sorted slots
        int l = lo, r = hi;
        while (l < r) { const int md = (l + r) >> 1; if (g.slot_camh[md] < cb_lo) l = md + 1; else r = md; }
        lo = l;
        r = hi;
        while (l < r) { const int md = (l + r) >> 1; if (g.slot_camh[md] <= cb_hi) l = md + 1; else r = md; }
        hi = l;
      }
      s_lo[tid] = lo;
      s_cnt[tid] = valid ? max(hi - lo, 0) : 0;
      // G_a = Hpl L^-T (L^-1 packed lower triangular)
#pragma unroll
      for (int r = 0; r < 6; ++r) {
        const double h0 = h[3 * r], h1 = h[3 * r + 1], h2 = h[3 * r + 2];
        s_Ga[tid][3 * r] = static_cast<HT>(h0 * m[0]);
        s_Ga[tid][3 * r + 1] = static_cast<HT>(h0 * m[1] + h1 * m[2]);
        s_Ga[tid][3 * r + 2] = static_cast<HT>(h0 * m[3] + h1 * m[4] + h2 * m[5]);
      }
#pragma unroll
      for (int k = 0; k < 6; ++k) s_L[tid][k] = m[k];
    }
    __syncthreads();
    sr_prefix(s_cnt, s_pre);
    __syncthreads();
    const int total = s_pre[SR_OBS];
    for (int k = tid; k < total; k += SR_THREADS) {
      // the batch observation t whose run holds pair k
      int l = 0, r = SR_OBS - 1;
      while (l < r) { const int md = (l + r + 1) >> 1; if (s_pre[md] <= k) l = md; else r = md - 1; }
      const int t = l;
      const int s2 = s_lo[t] + (k - s_pre[t]);
      const int b = g.slot_camh[s2];
      const int q2 = g.slot_q[s2];
      HT hb[18];
#pragma unroll
      for (int e = 0; e < 18; ++e) hb[e] = Hpl[18 * static_cast<size_t>(q2) + e];
      // column of b in the chunk
      int jl = 0, jr = nblk - 1;
      while (jl < jr) { const int md = (jl + jr) >> 1; if (s_col[md] < b) jl = md + 1; else jr = md; }
      double m[6];
#pragma unroll
      for (int e = 0; e < 6; ++e) m[e] = s_L[t][e];
      HT gb[18];
#pragma unroll
      for (int rr = 0; rr < 6; ++rr) {
        const double h0 = hb[3 * rr], h1 = hb[3 * rr + 1], h2 = hb[3 * rr + 2];
        gb[3 * rr] = static_cast<HT>(h0 * m[0]);
        gb[3 * rr + 1] = static_cast<HT>(h0 * m[1] + h1 * m[2]);
        gb[3 * rr + 2] = static_cast<HT>(h0 * m[3] + h1 * m[4] + h2 * m[5]);
      }
      double* acc = s_acc + 36 * jl;
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        const HT a0 = s_Ga[t][3 * i], a1 = s_Ga[t][3 * i + 1], a2 = s_Ga[t][3 * i + 2];
#pragma unroll
        for (int j = 0; j < 6; ++j) {
          const HT v = a0 * gb[3 * j] + a1 * gb[3 * j + 1] + a2 * gb[3 * j + 2];
          atomicAdd(acc + 6 * i + j, static_cast<double>(v));
        }
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < 36 * nblk; e += SR_THREADS) {
    const int j = e / 36, r = e - 36 * j;
    const int k = k0 + j;
    const HT v = static_cast<HT>(-s_acc[e]);
    F_val[36 * static_cast<size_t>(g.up_pos[k]) + r] = v;
    F_val[36 * static_cast<size_t>(g.up_posT[k]) + (r % 6) * 6 + r / 6] = v;
  }
}

template <typename HT>
hipError_t launch_schur_rows(const DevGraph& g, const int4* items, int nitems, int max_blk, const int* q_slot,
                             const HT* Hpl, const double* Linv, HT* F_val, hipStream_t st) {
  if (nitems == 0) return hipSuccess;
  const size_t lds = static_cast<size_t>(max_blk) * (36 * sizeof(double) + sizeof(int));
  if (!lds_fits(k_schur_rows<HT>, lds)) return hipErrorInvalidValue;
  const int per = (nitems + 7) / 8;
  hipLaunchKernelGGL(k_schur_rows<HT>, dim3(8 * per), dim3(SR_THREADS), lds, st, g, items, nitems, q_slot, Hpl,
                     Linv, F_val);
  return hipGetLastError();
}
template hipError_t launch_schur_rows<float>(const DevGraph&, const int4*, int, int, const int*, const float*,
                                             const double*, float*, hipStream_t);
template hipError_t launch_schur_rows<double>(const DevGraph&, const int4*, int, int, const int*, const double*,
                                              const double*, double*, hipStream_t);

// ========================================================== implicit Schur
template <typename HT>
__global__ void k_point_dinv(DevGraph g, const double* __restrict__ Hll, double lambda, HT* __restrict__ Dinv,
                             int* __restrict__ err) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.Pf) return;
  double D[9];
  if (!inv3_sym(Hll + 6 * static_cast<size_t>(p), lambda, D)) {
    atomicOr(err, 1);
    for (int i = 0; i < 9; ++i) D[i] = 0;
  }
  for (int i = 0; i < 9; ++i) Dinv[9 * static_cast<size_t>(p) + i] = static_cast<HT>(D[i]);
}

template <typename HT>
hipError_t launch_point_dinv(const DevGraph& g, const double* Hll, double lambda, HT* Dinv, int* err, hipStream_t st) {
  hipLaunchKernelGGL(k_point_dinv<HT>, dim3(nblocks(g.Pf, 256)), dim3(256), 0, st, g, Hll, lambda, Dinv, err);
  return hipGetLastError();
}
template hipError_t launch_point_dinv<float>(const DevGraph&, const double*, double, float*, int*, hipStream_t);
template hipError_t launch_point_dinv<double>(const DevGraph&, const double*, double, double*, int*, hipStream_t);

// One pass of the matrix-free reduced operator over the linearisation chunks:
// with p != null: v_l = sum Hpl^T p_c, y = Dinv v_l, part_c += Hpl y;
// with p == null: y = Dinv b_l (the Schur right-hand side).
// sc != null: PCG iteration `it` (skipped once the solve is over).
template <typename HT>
__global__ __launch_bounds__(LIN_THREADS) void k_implicit_pass(DevGraph g, const HT* __restrict__ Hpl,
                                                              const HT* __restrict__ Dinv,
                                                              const double* __restrict__ p,
                                                              const double* __restrict__ b,
                                                              double* __restrict__ cam_part, PcgScalars* sc, int it,
                                                              int exclude_own, const double* __restrict__ Linv) {
  if (sc && sc->done) return;
  // y = (Hll + lambda I)^-1 v: from the FP64 factor L^-1 (packed lower
  // triangle, k_point_linv) as L^-T (L^-1 v) when given - the Gram form keeps
  // the product accurate for ill-conditioned points, where rounding the
  // inverse's entries to FP32 would not - else from the stored inverse
  auto dinv_apply = [&](int ph, const double* v, double* y) {
    if (Linv) {
      const double* m = Linv + 6 * static_cast<size_t>(ph);
      const double u0 = m[0] * v[0], u1 = m[1] * v[0] + m[2] * v[1], u2 = m[3] * v[0] + m[4] * v[1] + m[5] * v[2];
      y[0] = m[0] * u0 + m[1] * u1 + m[3] * u2;
      y[1] = m[2] * u1 + m[4] * u2;
      y[2] = m[5] * u2;
    } else {
      const HT* D = Dinv + 9 * static_cast<size_t>(ph);
      for (int r = 0; r < 3; ++r)
        y[r] = static_cast<double>(D[3 * r]) * v[0] + static_cast<double>(D[3 * r + 1]) * v[1] +
               static_cast<double>(D[3 * r + 2]) * v[2];
    }
  };
  // camera accumulators (6 per chunk camera) and the chunk's camera list
  extern __shared__ double s_dyn[];
  double* const s_cacc = s_dyn;
  int* const s_cam = reinterpret_cast<int*>(s_dyn + 6 * g.lc_maxc);
  __shared__ double s_v[LIN_PTS * 3];
  const int ch = blockIdx.x, tid = threadIdx.x;
  const int p0 = g.lc_pos0[ch], p1 = g.lc_pos0[ch + 1];
  const int s0 = g.pos_ptr[p0], s1 = g.pos_ptr[p1];
  const int c0 = g.lc_cam_ptr[ch], nc = g.lc_cam_ptr[ch + 1] - c0;
  const int np = p1 - p0;
  for (int i = tid; i < nc; i += LIN_THREADS) s_cam[i] = g.lc_cams[c0 + i];
  for (int i = tid; i < nc * 6; i += LIN_THREADS) s_cacc[i] = 0.0;
  for (int i = tid; i < np * 3; i += LIN_THREADS) s_v[i] = 0.0;
  __syncthreads();
  if (p) {
    for (int s = s0 + tid; s < s1; s += LIN_THREADS) {
      const int pos = g.slot_pos[s];
      const int chh = g.camH[g.slot_cam[s]];
      if (g.pos_h[pos] < 0 || chh < 0) continue;
      const double* pc = p + 6 * static_cast<size_t>(chh);
      const HT* h = Hpl + 18 * static_cast<size_t>(g.slot_q[s]);
      for (int c = 0; c < 3; ++c) {
        double v = 0;
        for (int r = 0; r < 6; ++r) v += static_cast<double>(h[3 * r + c]) * pc[r];
        atomicAdd(&s_v[3 * (pos - p0) + c], v);
      }
    }
    __syncthreads();
  }
  for (int lp = tid; lp < np; lp += LIN_THREADS) {
    const int ph = g.pos_h[p0 + lp];
    if (ph < 0) continue;
    double v[3];
    if (p) {
      v[0] = s_v[3 * lp]; v[1] = s_v[3 * lp + 1]; v[2] = s_v[3 * lp + 2];
    } else {
      const double* bl = b + 6 * static_cast<size_t>(g.Cf) + 3 * static_cast<size_t>(ph);
      v[0] = bl[0]; v[1] = bl[1]; v[2] = bl[2];
    }
    double y[3];
    dinv_apply(ph, v, y);
    for (int r = 0; r < 3; ++r) s_v[3 * lp + r] = y[r];
  }
  __syncthreads();
  for (int s = s0 + tid; s < s1; s += LIN_THREADS) {
    const int pos = g.slot_pos[s];
    const int chh = g.camH[g.slot_cam[s]];
    if (g.pos_h[pos] < 0 || chh < 0) continue;
    double y[3] = {s_v[3 * (pos - p0)], s_v[3 * (pos - p0) + 1], s_v[3 * (pos - p0) + 2]};
    if (exclude_own && p) {
      // only the other cameras' terms: y = Dinv (sum_j V_j^T p_j - sum over
      // this camera's slots of the point V^T p_i); the camera's own block is
      // applied from the FP64 diagonal (k_implicit_reduce). The point's
      // slots are sorted by camera, so its slots of this camera are adjacent.
      const double* pc = p + 6 * static_cast<size_t>(chh);
      const int sb = g.pos_ptr[pos], se = g.pos_ptr[pos + 1];
      int a = s, e = s + 1;
      while (a > sb && g.camH[g.slot_cam[a - 1]] == chh) --a;
      while (e < se && g.camH[g.slot_cam[e]] == chh) ++e;
      double own[3] = {0.0, 0.0, 0.0};
      for (int s2 = a; s2 < e; ++s2) {
        const HT* h2 = Hpl + 18 * static_cast<size_t>(g.slot_q[s2]);
        for (int c = 0; c < 3; ++c) {
          double v = 0;
          for (int r = 0; r < 6; ++r) v += static_cast<double>(h2[3 * r + c]) * pc[r];
          own[c] += v;
        }
      }
      double yo[3];
      dinv_apply(g.pos_h[pos], own, yo);
      for (int r = 0; r < 3; ++r) y[r] -= yo[r];
    }
    const int lc = lower_bound_i(s_cam, nc, chh);
    for (int r = 0; r < 6; ++r) {
      const HT* h = Hpl + 18 * static_cast<size_t>(g.slot_q[s]) + 3 * r;
      const double v = static_cast<double>(h[0]) * y[0] + static_cast<double>(h[1]) * y[1] +
                       static_cast<double>(h[2]) * y[2];
      atomicAdd(&s_cacc[6 * lc + r], v);
    }
  }
  __syncthreads();
  for (int i = tid; i < nc * 6; i += LIN_THREADS) cam_part[static_cast<size_t>(c0) * 6 + i] = s_cacc[i];
}

template <typename HT>
hipError_t launch_implicit_pass(const DevGraph& g, const HT* Hpl, const HT* Dinv, const double* p, const double* b,
                                double* cam_part, const PcgScalars* sc, int it, hipStream_t st, bool exclude_own,
                                const double* Linv) {
  const size_t lds = (6 * sizeof(double) + sizeof(int)) * static_cast<size_t>(g.lc_maxc);
  if (!lds_fits(k_implicit_pass<HT>, lds)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_implicit_pass<HT>, dim3(g.nlc), dim3(LIN_THREADS), lds, st, g, Hpl, Dinv, p, b, cam_part,
                     const_cast<PcgScalars*>(sc), it, exclude_own ? 1 : 0, Linv);
  return hipGetLastError();
}
template hipError_t launch_implicit_pass<float>(const DevGraph&, const float*, const float*, const double*,
                                                const double*, double*, const PcgScalars*, int, hipStream_t, bool,
                                                const double*);
template hipError_t launch_implicit_pass<double>(const DevGraph&, const double*, const double*, const double*,
                                                 const double*, double*, const PcgScalars*, int, hipStream_t, bool,
                                                 const double*);

// q = (Hpp + lambda I) p - sum of chunk partials, or with Sdiag (the FP64
// diagonal blocks of the damped reduced system, the passes then carrying
// only the other cameras' terms) q = Sdiag p - sum of chunk partials;
// dq_part[block] = sum of q_c . p_c
template <typename HT>
__global__ __launch_bounds__(64) void k_implicit_reduce(DevGraph g, const double* __restrict__ cam_part,
                                                        const double* __restrict__ Hpp, double lambda,
                                                        const double* __restrict__ Sdiag,
                                                        const double* __restrict__ p, double* __restrict__ q,
                                                        double* __restrict__ dq_part, const PcgScalars* sc) {
  if (sc->done) return;
  const int c = blockIdx.x * 64 + threadIdx.x;
  double dq = 0;
  if (c < g.Cf) {
    double pc[6], out[6];
    for (int r = 0; r < 6; ++r) pc[r] = p[6 * static_cast<size_t>(c) + r];
    const double* A = Sdiag ? Sdiag : Hpp;
    for (int r = 0; r < 6; ++r) {
      double v = Sdiag ? 0.0 : lambda * pc[r];
      for (int k = 0; k < 6; ++k) v += A[36 * static_cast<size_t>(c) + 6 * r + k] * pc[k];
      out[r] = v;
    }
    for (int q2 = g.camp_ptr[c]; q2 < g.camp_ptr[c + 1]; ++q2) {
      const double* pp = cam_part + 6 * static_cast<size_t>(g.camp_idx[q2]);
      for (int r = 0; r < 6; ++r) out[r] -= pp[r];
    }
    for (int r = 0; r < 6; ++r) {
      q[6 * static_cast<size_t>(c) + r] = out[r];
      dq += out[r] * pc[r];
    }
  }
  dq = wave_sum(dq);
  if (threadIdx.x == 0) dq_part[blockIdx.x] = dq;
}

template <typename HT>
hipError_t launch_implicit_reduce(const DevGraph& g, const double* cam_part, const double* Hpp, double lambda,
                                  const double* p, double* q, double* dq_part, PcgScalars* sc, int it,
                                  hipStream_t st, const double* Sdiag) {
  (void)it;
  hipLaunchKernelGGL(k_implicit_reduce<HT>, dim3(implicit_reduce_blocks(g.Cf)), dim3(64), 0, st, g, cam_part, Hpp,
                     lambda, Sdiag, p, q, dq_part, sc);
  return hipGetLastError();
}
template hipError_t launch_implicit_reduce<float>(const DevGraph&, const double*, const double*, double,
                                                  const double*, double*, double*, PcgScalars*, int, hipStream_t,
                                                  const double*);
template hipError_t launch_implicit_reduce<double>(const DevGraph&, const double*, const double*, double,
                                                   const double*, double*, double*, PcgScalars*, int, hipStream_t,
                                                   const double*);

__global__ void k_implicit_rhs_reduce(DevGraph g, const double* __restrict__ cam_part, const double* __restrict__ b,
                                      double* __restrict__ bschur) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * g.Cf) return;
  const int c = t / 6, r = t % 6;
  double v = b[t];
  for (int q = g.camp_ptr[c]; q < g.camp_ptr[c + 1]; ++q) v -= cam_part[6 * static_cast<size_t>(g.camp_idx[q]) + r];
  bschur[t] = v;
}

hipError_t launch_implicit_rhs_reduce(const DevGraph& g, const double* cam_part, const double* b, double* bschur,
                                      hipStream_t st) {
  hipLaunchKernelGGL(k_implicit_rhs_reduce, dim3(nblocks(6LL * g.Cf, 256)), dim3(256), 0, st, g, cam_part, b, bschur);
  return hipGetLastError();
}

// ======================================= FP64 matrix-free residual / back-sub
// Re-linearises every edge in FP64 from the state. For the camera update xp:
//   u_l = (Hll + lambda I)^-1 sum_e -w' Jp^T (e + Jc xp_c)   (= x_l, the
//         back-substitution of block_solver.hpp:452-476, written to xl),
//   part_c += sum_e -w' Jc^T (e + Jc xp_c + Jp u_l)           (if cam_part)
// so that r_c = -lambda xp_c + sum part_c is the camera block of
// b - (H + lambda I) [xp; x_l] with the point block exactly zero: the FP64
// residual of the reduced system used for iterative refinement.
__global__ __launch_bounds__(LIN_THREADS) void k_refine(DevGraph g, const double* __restrict__ camRt,
                                                       const double* __restrict__ pts, double lambda,
                                                       const double* __restrict__ xp, double* __restrict__ xl,
                                                       double* __restrict__ cam_part, int* __restrict__ err) {
  // camera accumulators (6 per chunk camera) and the chunk's camera list
  extern __shared__ double s_dyn[];
  double* const s_cacc = s_dyn;
  int* const s_cam = reinterpret_cast<int*>(s_dyn + 6 * g.lc_maxc);
  __shared__ double s_pacc[LIN_PTS * 9];
  const int ch = blockIdx.x, tid = threadIdx.x;
  const int p0 = g.lc_pos0[ch], p1 = g.lc_pos0[ch + 1];
  const int s0 = g.pos_ptr[p0], s1 = g.pos_ptr[p1];
  const int c0 = g.lc_cam_ptr[ch], nc = g.lc_cam_ptr[ch + 1] - c0;
  const int np = p1 - p0;
  for (int i = tid; i < nc; i += LIN_THREADS) s_cam[i] = g.lc_cams[c0 + i];
  for (int i = tid; i < nc * 6; i += LIN_THREADS) s_cacc[i] = 0.0;
  for (int i = tid; i < np * 9; i += LIN_THREADS) s_pacc[i] = 0.0;
  __syncthreads();
  for (int sb = s0 + tid; sb < s1; sb += LIN_PF1 * LIN_THREADS) {
    SlotOps op[LIN_PF1];
    double xv[LIN_PF1][6];
#pragma unroll
    for (int u = 0; u < LIN_PF1; ++u) {
      load_slot_ops(g, camRt, pts, min(sb + u * LIN_THREADS, s1 - 1), op[u]);
#pragma unroll
      for (int k = 0; k < 6; ++k) xv[u][k] = xp ? xp[6 * static_cast<size_t>(max(op[u].chh, 0)) + k] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < LIN_PF1; ++u) {
      if (sb + u * LIN_THREADS >= s1) break;
      const SlotOps& o = op[u];
      if (o.ph < 0) continue;
      double Xc[3], e[2];
      project_error(o.Rt, o.in, o.X, o.z, Xc, e);
      double Jp[6], Jc[12];
      project_jacobians<double>(o.Rt, o.in[0], o.in[1], Xc[0], Xc[1], Xc[2], Jp, Jc);
      whiten_obs(g, o.S, e, Jp, Jc);
      double rho[3];
      huber_rho(o.w * (e[0] * e[0] + e[1] * e[1]), g.huber, rho);
      const double om = rho[1] * o.w;
      double t0 = e[0], t1 = e[1];
      if (o.chh >= 0 && xp)
        for (int k = 0; k < 6; ++k) { t0 += Jc[k] * xv[u][k]; t1 += Jc[6 + k] * xv[u][k]; }
      double* pa = s_pacc + 9 * (o.pos - p0);
      int k = 0;
      for (int r = 0; r < 3; ++r)
        for (int c = r; c < 3; ++c) atomicAdd(&pa[k++], om * (Jp[r] * Jp[c] + Jp[3 + r] * Jp[3 + c]));
      for (int r = 0; r < 3; ++r) atomicAdd(&pa[6 + r], -om * (Jp[r] * t0 + Jp[3 + r] * t1));
    }
  }
  __syncthreads();
  for (int lp = tid; lp < np; lp += LIN_THREADS) {
    const int ph = g.pos_h[p0 + lp];
    double* pa = s_pacc + 9 * lp;
    if (ph < 0) { pa[6] = pa[7] = pa[8] = 0.0; continue; }
    double D[9];
    if (!inv3_sym(pa, lambda, D)) {
      atomicOr(err, 1);
      for (int i = 0; i < 9; ++i) D[i] = 0;
    }
    double u[3];
    for (int r = 0; r < 3; ++r) u[r] = D[3 * r] * pa[6] + D[3 * r + 1] * pa[7] + D[3 * r + 2] * pa[8];
    for (int r = 0; r < 3; ++r) {
      xl[3 * static_cast<size_t>(ph) + r] = u[r];
      pa[6 + r] = u[r];
    }
  }
  if (!cam_part) return;
  __syncthreads();
  for (int sb = s0 + tid; sb < s1; sb += LIN_PF1 * LIN_THREADS) {
    SlotOps op[LIN_PF1];
    double xv[LIN_PF1][6];
#pragma unroll
    for (int u = 0; u < LIN_PF1; ++u) {
      load_slot_ops(g, camRt, pts, min(sb + u * LIN_THREADS, s1 - 1), op[u]);
#pragma unroll
      for (int k = 0; k < 6; ++k) xv[u][k] = xp ? xp[6 * static_cast<size_t>(max(op[u].chh, 0)) + k] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < LIN_PF1; ++u) {
      if (sb + u * LIN_THREADS >= s1) break;
      const SlotOps& o = op[u];
      if (o.chh < 0) continue;
      double Xc[3], e[2];
      project_error(o.Rt, o.in, o.X, o.z, Xc, e);
      double Jp[6], Jc[12];
      project_jacobians<double>(o.Rt, o.in[0], o.in[1], Xc[0], Xc[1], Xc[2], Jp, Jc);
      whiten_obs(g, o.S, e, Jp, Jc);
      double rho[3];
      huber_rho(o.w * (e[0] * e[0] + e[1] * e[1]), g.huber, rho);
      const double om = rho[1] * o.w;
      double t0 = e[0], t1 = e[1];
      if (xp)
        for (int k = 0; k < 6; ++k) { t0 += Jc[k] * xv[u][k]; t1 += Jc[6 + k] * xv[u][k]; }
      const double* uu = s_pacc + 9 * (o.pos - p0) + 6;
      for (int k = 0; k < 3; ++k) { t0 += Jp[k] * uu[k]; t1 += Jp[3 + k] * uu[k]; }
      const int lc = lower_bound_i(s_cam, nc, o.chh);
      for (int r = 0; r < 6; ++r) atomicAdd(&s_cacc[6 * lc + r], -om * (Jc[r] * t0 + Jc[6 + r] * t1));
    }
  }
  __syncthreads();
  for (int i = tid; i < nc * 6; i += LIN_THREADS) cam_part[static_cast<size_t>(c0) * 6 + i] = s_cacc[i];
}

hipError_t launch_refine(const DevGraph& g, const double* camRt, const double* pts, double lambda, const double* xp,
                         double* xl, double* cam_part, int* err, hipStream_t st) {
  if (g.nlc == 0) return hipSuccess;
  const size_t lds = (6 * sizeof(double) + sizeof(int)) * static_cast<size_t>(g.lc_maxc);
  if (!lds_fits(k_refine, lds)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_refine, dim3(g.nlc), dim3(LIN_THREADS), lds, st, g, camRt, pts, lambda, xp, xl, cam_part,
                     err);
  return hipGetLastError();
}

__global__ void k_refine_reduce(DevGraph g, const double* __restrict__ cam_part, const double* __restrict__ xp,
                                double lambda, double* __restrict__ r) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * g.Cf) return;
  const int c = t / 6, k = t % 6;
  r[t] = (xp ? -lambda * xp[t] : 0.0) + camp_sum(g, cam_part, 6, c, k);
}

hipError_t launch_refine_reduce(const DevGraph& g, const double* cam_part, const double* xp, double lambda, double* r,
                                hipStream_t st) {
  hipLaunchKernelGGL(k_refine_reduce, dim3(nblocks(6LL * g.Cf, 256)), dim3(256), 0, st, g, cam_part, xp, lambda, r);
  return hipGetLastError();
}

// ============================================================ block-Jacobi PCG
// LinearSolverPCG::solve (linear_solver_pcg.hpp:87-161) as a stream of
// kernels per iteration. Vectors and the matrix in HT, dot products in f64 with
// per-row partials summed (compensated, fixed order) by every consumer block.
// Inverse of a symmetric positive-definite BD x BD block (row-major, in
// place): A = L L^T, then A^-1 = L^-T L^-1. Constant loop bounds keep
// everything in registers. Returns false unless every pivot is positive.
template <int BD>
__device__ __forceinline__ bool inv_spd(double* m) {
  double L[BD][BD];
#pragma unroll
  for (int i = 0; i < BD; ++i)
#pragma unroll
    for (int j = 0; j < BD; ++j) L[i][j] = 0.0;
  bool ok = true;
#pragma unroll
  for (int j = 0; j < BD; ++j) {
    double s = m[BD * j + j];
#pragma unroll
    for (int k = 0; k < j; ++k) s -= L[j][k] * L[j][k];
    ok = ok && (s > 0.0) && isfinite(s);
    const double ljj = sqrt(fmax(s, 1e-300));
    L[j][j] = ljj;
    const double il = 1.0 / ljj;
#pragma unroll
    for (int i = j + 1; i < BD; ++i) {
      double t = m[BD * i + j];
#pragma unroll
      for (int k = 0; k < j; ++k) t -= L[i][k] * L[j][k];
      L[i][j] = t * il;
    }
  }
  // N = L^-1 (lower triangular)
  double N[BD][BD];
#pragma unroll
  for (int i = 0; i < BD; ++i)
#pragma unroll
    for (int j = 0; j < BD; ++j) N[i][j] = 0.0;
#pragma unroll
  for (int j = 0; j < BD; ++j) {
    N[j][j] = 1.0 / L[j][j];
#pragma unroll
    for (int i = j + 1; i < BD; ++i) {
      double t = 0.0;
#pragma unroll
      for (int k = j; k < i; ++k) t -= L[i][k] * N[k][j];
      N[i][j] = t / L[i][i];
    }
  }
  // A^-1 = N^T N
#pragma unroll
  for (int i = 0; i < BD; ++i)
#pragma unroll
    for (int j = i; j < BD; ++j) {
      double t = 0.0;
#pragma unroll
      for (int k = j; k < BD; ++k) t += N[k][i] * N[k][j];
      m[BD * i + j] = t;
      m[BD * j + i] = t;
    }
  return ok;
}

template <typename HT, int BD>
__global__ void k_jacobi(int nb, const HT* __restrict__ F_val, const int* __restrict__ F_diag, HT* __restrict__ Jinv,
                         int* __restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  double m[BD * BD];
  const HT* src = F_val + static_cast<size_t>(F_diag[i]) * BD * BD;
#pragma unroll
  for (int k = 0; k < BD * BD; ++k) m[k] = static_cast<double>(src[k]);
  if (!inv_spd<BD>(m)) {
    atomicOr(err, 2);
#pragma unroll
    for (int k = 0; k < BD * BD; ++k) m[k] = 0;
  }
#pragma unroll
  for (int k = 0; k < BD * BD; ++k) Jinv[static_cast<size_t>(i) * BD * BD + k] = static_cast<HT>(m[k]);
}

template <typename HT, int BD>
hipError_t launch_jacobi(int nb, const HT* F_val, const int* F_diag, HT* Jinv, int* err, hipStream_t st) {
  hipLaunchKernelGGL((k_jacobi<HT, BD>), dim3(nblocks(nb, 64)), dim3(64), 0, st, nb, F_val, F_diag, Jinv, err);
  return hipGetLastError();
}

// ------------------------------------------ FP64 diagonal Schur blocks
// S_ii = Hpp_ii + lambda I - sum_p V_ip Dinv_p V_ip^T with V the Hpl block
// subtracts terms of the size of Hpp (~1e9 on long trajectories) to leave a
// block whose smallest eigenvalue can be ~1e2, so it is formed in FP64 for
// every precision: Hpp_ii and Hll_p are FP64 sums (k_linearize), and V is
// re-linearised in FP64 from the state, V = omega Jc^T Jp
// (edge_project_xyz.cpp:57-93 via base_fixed_sized_edge.hpp:49-140), or read from
// the given blocks (the exact Hessian). One workgroup per free camera walks
// its observations in camera-major order (two observations of one point by
// the camera are adjacent and enter as one summed V), sums the 21 upper
// entries and the 6 right-hand-side terms V Dinv b_p in registers and
// reduces them over the workgroup: no atomics, fixed order. Writes the block
// into F_val (if not null), its inverse into Jinv and b_i - sum V Dinv b_p
// into bschur (if not null).
// V += omega Jc^T Jp of the observation (slot, point position pp) by the
// camera whose [R | t] and intrinsics are Rt, in.
__device__ __forceinline__ void obs_V64(const DevGraph& g, const double* __restrict__ Rt,
                                       const double* __restrict__ in, const double* __restrict__ pts, int slot,
                                       int pp, double* V) {
  const double X[3] = {pts[3 * static_cast<size_t>(pp)], pts[3 * static_cast<size_t>(pp) + 1],
                       pts[3 * static_cast<size_t>(pp) + 2]};
  const double z[2] = {g.slot_meas[2 * static_cast<size_t>(slot)], g.slot_meas[2 * static_cast<size_t>(slot) + 1]};
  const double w = g.slot_w[slot];
  double Xc[3], e[2], rho[3], Jp[6], Jc[12];
  project_error(Rt, in, X, z, Xc, e);
  project_jacobians<double>(Rt, in[0], in[1], Xc[0], Xc[1], Xc[2], Jp, Jc);
  if (g.slot_S) whiten_obs(g, g.slot_S + 3 * static_cast<size_t>(slot), e, Jp, Jc);
  huber_rho(w * (e[0] * e[0] + e[1] * e[1]), g.huber, rho);
  const double om = rho[1] * w;
  for (int r = 0; r < 6; ++r)
    for (int cc = 0; cc < 3; ++cc) V[3 * r + cc] += om * (Jc[r] * Jp[cc] + Jc[6 + r] * Jp[3 + cc]);
}

// The same from operands already loaded (load_slot_ops).
__device__ __forceinline__ void obs_V64_ops(const DevGraph& g, const SlotOps& o, double* V) {
  double Xc[3], e[2], rho[3], Jp[6], Jc[12];
  project_error(o.Rt, o.in, o.X, o.z, Xc, e);
  project_jacobians<double>(o.Rt, o.in[0], o.in[1], Xc[0], Xc[1], Xc[2], Jp, Jc);
  whiten_obs(g, o.S, e, Jp, Jc);
  huber_rho(o.w * (e[0] * e[0] + e[1] * e[1]), g.huber, rho);
  const double om = rho[1] * o.w;
  for (int r = 0; r < 6; ++r)
    for (int cc = 0; cc < 3; ++cc) V[3 * r + cc] += om * (Jc[r] * Jp[cc] + Jc[6 + r] * Jp[3 + cc]);
}

template <typename HT>
__global__ __launch_bounds__(256) void k_schur_diag(DevGraph g, const int* __restrict__ q_slot,
                                                    const double* __restrict__ camRt, const double* __restrict__ pts,
                                                    const HT* __restrict__ Hpl, const double* __restrict__ Hll,
                                                    const double* __restrict__ Hpp, const double* __restrict__ b,
                                                    double lambda, HT* __restrict__ F_val, HT* __restrict__ Jinv,
                                                    double* __restrict__ bschur, int* __restrict__ err) {
  __shared__ double s_red[4][27];
  const int c = blockIdx.x, tid = threadIdx.x;
  const int qb = g.cam_qptr[c], qe = g.cam_qptr[c + 1];
  const int craw = g.cam_raw[c];
  const double* Rt = camRt + 12 * static_cast<size_t>(craw);
  const double* in = g.intr + 4 * static_cast<size_t>(craw);
  const double* bl = b ? b + 6 * static_cast<size_t>(g.Cf) : nullptr;
  double acc[27];
  for (int k = 0; k < 27; ++k) acc[k] = 0;
  for (int q = qb + tid; q < qe; q += 256) {
    // indices of this observation and of its predecessor, loaded together
    const int ph = g.q_ph[q];
    const int ph_prev = q > qb ? g.q_ph[q - 1] : -1;
    const int slot = q_slot[q], pp = g.q_pt[q];
    if (ph_prev == ph) continue;  // summed by the run's first observation
    double V[18];
    for (int k = 0; k < 18; ++k) V[k] = 0;
    if (Hpl) {
      for (int q2 = q; q2 < qe && g.q_ph[q2] == ph; ++q2) {
        const HT* h = Hpl + 18 * static_cast<size_t>(q2);
        for (int k = 0; k < 18; ++k) V[k] += static_cast<double>(h[k]);
      }
    } else {
      obs_V64(g, Rt, in, pts, slot, pp, V);
      for (int q2 = q + 1; q2 < qe && g.q_ph[q2] == ph; ++q2) obs_V64(g, Rt, in, pts, q_slot[q2], g.q_pt[q2], V);
    }
    double D[9];
    if (!inv3_sym(Hll + 6 * static_cast<size_t>(ph), lambda, D)) {
      atomicOr(err, 1);
      continue;
    }
    double VD[18];
    for (int r = 0; r < 6; ++r)
      for (int cc = 0; cc < 3; ++cc)
        VD[3 * r + cc] = V[3 * r] * D[cc] + V[3 * r + 1] * D[3 + cc] + V[3 * r + 2] * D[6 + cc];
    int k = 0;
    for (int r = 0; r < 6; ++r)
      for (int cc = r; cc < 6; ++cc)
        acc[k++] += VD[3 * r] * V[3 * cc] + VD[3 * r + 1] * V[3 * cc + 1] + VD[3 * r + 2] * V[3 * cc + 2];
    if (bl) {
      const double* y = bl + 3 * static_cast<size_t>(ph);
      for (int r = 0; r < 6; ++r) acc[21 + r] += VD[3 * r] * y[0] + VD[3 * r + 1] * y[1] + VD[3 * r + 2] * y[2];
    }
  }
  const int lane = tid & 63, wv = tid >> 6;
  for (int k = 0; k < 27; ++k) {
    const double v = wave_sum(acc[k]);
    if (lane == 0) s_red[wv][k] = v;
  }
  __syncthreads();
  if (bl && bschur && tid >= 64 && tid < 70) {
    const int r = tid - 64;
    const double v = ((s_red[0][21 + r] + s_red[1][21 + r]) + s_red[2][21 + r]) + s_red[3][21 + r];
    bschur[6 * static_cast<size_t>(c) + r] = b[6 * static_cast<size_t>(c) + r] - v;
  }
  if (tid != 0) return;
  double m[36];
  for (int k = 0; k < 21; ++k) {
    const double v = ((s_red[0][k] + s_red[1][k]) + s_red[2][k]) + s_red[3][k];
    const int r = c_up6_r[k], cc = c_up6_c[k];
    const double h = Hpp[36 * static_cast<size_t>(c) + 6 * r + cc] - v;
    m[6 * r + cc] = h;
    m[6 * cc + r] = h;
  }
  for (int k = 0; k < 6; ++k) m[7 * k] += lambda;
  if (F_val) {
    HT* dst = F_val + 36 * static_cast<size_t>(g.F_diag[c]);
    for (int k = 0; k < 36; ++k) dst[k] = static_cast<HT>(m[k]);
  }
  if (!inv_spd<6>(m)) {
    atomicOr(err, 2);
    for (int k = 0; k < 36; ++k) m[k] = 0;
  }
  for (int k = 0; k < 36; ++k) Jinv[36 * static_cast<size_t>(c) + k] = static_cast<HT>(m[k]);
}

// The same diagonal terms point-major (the LM path): one workgroup per
// linearisation chunk, whose slots, points and measurements are contiguous.
// Per point D = (Hll + lambda I)^-1 and y = D b_p go to LDS; per observation
// run (one point, one camera) V is re-linearised in FP64 from the state and
// the 21 upper entries of V D V^T and the 6 entries of V y are added to the
// chunk camera's LDS accumulator (f64 atomics); the chunk writes one partial
// row per camera in the layout of the linearisation's Hpp partials.
__global__ __launch_bounds__(LIN_THREADS) void k_schur_diag_pm(DevGraph g, const double* __restrict__ camRt,
                                                               const double* __restrict__ pts,
                                                               const double* __restrict__ Hll,
                                                               const double* __restrict__ b, double lambda,
                                                               double* __restrict__ part, int* __restrict__ err) {
  // camera accumulators (27 per chunk camera) and the chunk's camera list
  extern __shared__ double s_dyn[];
  double* const s_cacc = s_dyn;
  int* const s_cam = reinterpret_cast<int*>(s_dyn + 27 * g.lc_maxc);
  __shared__ double s_pd[LIN_PTS * 9];
  const int ch = blockIdx.x, tid = threadIdx.x;
  const int p0 = g.lc_pos0[ch], p1 = g.lc_pos0[ch + 1];
  const int s0 = g.pos_ptr[p0], s1 = g.pos_ptr[p1];
  const int c0 = g.lc_cam_ptr[ch], nc = g.lc_cam_ptr[ch + 1] - c0;
  const int np = p1 - p0;
  const double* bl = b ? b + 6 * static_cast<size_t>(g.Cf) : nullptr;
  for (int i = tid; i < nc; i += LIN_THREADS) s_cam[i] = g.lc_cams[c0 + i];
  for (int i = tid; i < nc * 27; i += LIN_THREADS) s_cacc[i] = 0.0;
  for (int lp = tid; lp < np; lp += LIN_THREADS) {
    const int ph = g.pos_h[p0 + lp];
    double* pd = s_pd + 9 * lp;
    if (ph < 0) continue;
    double D[9];
    if (!inv3_sym(Hll + 6 * static_cast<size_t>(ph), lambda, D)) {
      atomicOr(err, 1);
      for (int k = 0; k < 9; ++k) D[k] = 0.0;
    }
    pd[0] = D[0]; pd[1] = D[1]; pd[2] = D[2]; pd[3] = D[4]; pd[4] = D[5]; pd[5] = D[8];
    for (int r = 0; r < 3; ++r) {
      const double* y = bl ? bl + 3 * static_cast<size_t>(ph) : nullptr;
      pd[6 + r] = y ? D[3 * r] * y[0] + D[3 * r + 1] * y[1] + D[3 * r + 2] * y[2] : 0.0;
    }
  }
  __syncthreads();
  for (int sb = s0 + tid; sb < s1; sb += LIN_PF1 * LIN_THREADS) {
    // operands of this thread's next slots, and the neighbouring slots'
    // indices that delimit runs of one (point, camera) pair, all in flight
    SlotOps op[LIN_PF1];
    int nb_pos[LIN_PF1][2], nb_cam[LIN_PF1][2];
#pragma unroll
    for (int u = 0; u < LIN_PF1; ++u) {
      const int sc = min(sb + u * LIN_THREADS, s1 - 1);
      load_slot_ops(g, camRt, pts, sc, op[u]);
      const int sp = max(sc - 1, s0), sn = min(sc + 1, s1 - 1);
      nb_pos[u][0] = g.slot_pos[sp];
      nb_cam[u][0] = g.slot_cam[sp];
      nb_pos[u][1] = g.slot_pos[sn];
      nb_cam[u][1] = g.slot_cam[sn];
    }
#pragma unroll
    for (int u = 0; u < LIN_PF1; ++u) {
      const int s = sb + u * LIN_THREADS;
      if (s >= s1) break;
      const SlotOps& o = op[u];
      const int pos = o.pos, cam = o.cam, chh = o.chh;
      if (o.ph < 0 || chh < 0) continue;
      if (s > s0 && nb_pos[u][0] == pos && nb_cam[u][0] == cam) continue;  // the run's first slot sums it
      const double* Rt = o.Rt;
      const double* in = o.in;
      double V[18];
      for (int k = 0; k < 18; ++k) V[k] = 0.0;
      obs_V64_ops(g, o, V);
      if (s + 1 < s1 && nb_pos[u][1] == pos && nb_cam[u][1] == cam)
        for (int s2 = s + 1; s2 < s1 && g.slot_pos[s2] == pos && g.slot_cam[s2] == cam; ++s2)
          obs_V64(g, Rt, in, pts, s2, pos, V);
      const double* pd = s_pd + 9 * (pos - p0);
      const double D[9] = {pd[0], pd[1], pd[2], pd[1], pd[3], pd[4], pd[2], pd[4], pd[5]};
      double VD[18];
      for (int r = 0; r < 6; ++r)
        for (int cc = 0; cc < 3; ++cc)
          VD[3 * r + cc] = V[3 * r] * D[cc] + V[3 * r + 1] * D[3 + cc] + V[3 * r + 2] * D[6 + cc];
      double* ca = s_cacc + 27 * lower_bound_i(s_cam, nc, chh);
      int k = 0;
      for (int r = 0; r < 6; ++r)
        for (int cc = r; cc < 6; ++cc)
          atomicAdd(&ca[k++], VD[3 * r] * V[3 * cc] + VD[3 * r + 1] * V[3 * cc + 1] + VD[3 * r + 2] * V[3 * cc + 2]);
      if (bl)
        for (int r = 0; r < 6; ++r) atomicAdd(&ca[21 + r], V[3 * r] * pd[6] + V[3 * r + 1] * pd[7] + V[3 * r + 2] * pd[8]);
    }
  }
  __syncthreads();
  for (int i = tid; i < nc * 27; i += LIN_THREADS) part[static_cast<size_t>(c0) * 27 + i] = s_cacc[i];
}

// Per free camera (64 lanes): the chunk partials summed in a fixed order
// (lanes 0..26), then S_ii = Hpp_ii + lambda I - sum into F_val (if not
// null), its inverse into Jinv (err |= 2 if not PD) and bschur_i = b_i - sum
// (if bschur).
// With Sdiag the block goes there in FP64 and F_val's diagonal block is
// zeroed (the SpMV applies it from Sdiag).
template <typename HT>
__global__ __launch_bounds__(64) void k_schur_diag_finish(DevGraph g, const double* __restrict__ part,
                                                          const double* __restrict__ Hpp,
                                                          const double* __restrict__ b, double lambda,
                                                          HT* __restrict__ F_val, HT* __restrict__ Jinv,
                                                          double* __restrict__ bschur, double* __restrict__ Sdiag,
                                                          int* __restrict__ err) {
  __shared__ double s_v[27];
  const int c = blockIdx.x, lane = threadIdx.x;
  if (lane < 27) s_v[lane] = camp_sum(g, part, 27, c, lane);
  __syncthreads();
  if (bschur && lane >= 21 && lane < 27)
    bschur[6 * static_cast<size_t>(c) + lane - 21] = b[6 * static_cast<size_t>(c) + lane - 21] - s_v[lane];
  if (lane != 0) return;
  double m[36];
  for (int k = 0; k < 21; ++k) {
    const int r = c_up6_r[k], cc = c_up6_c[k];
    const double h = Hpp[36 * static_cast<size_t>(c) + 6 * r + cc] - s_v[k];
    m[6 * r + cc] = h;
    m[6 * cc + r] = h;
  }
  for (int k = 0; k < 6; ++k) m[7 * k] += lambda;
  if (F_val) {
    HT* dst = F_val + 36 * static_cast<size_t>(g.F_diag[c]);
    for (int k = 0; k < 36; ++k) dst[k] = Sdiag ? HT(0) : static_cast<HT>(m[k]);
  }
  if (Sdiag)
    for (int k = 0; k < 36; ++k) Sdiag[36 * static_cast<size_t>(c) + k] = m[k];
  if (!inv_spd<6>(m)) {
    atomicOr(err, 2);
    for (int k = 0; k < 36; ++k) m[k] = 0;
  }
  for (int k = 0; k < 36; ++k) Jinv[36 * static_cast<size_t>(c) + k] = static_cast<HT>(m[k]);
}

template <typename HT>
hipError_t launch_schur_diag_pm(const DevGraph& g, const double* camRt, const double* pts, const double* Hll,
                                const double* Hpp, const double* b, double lambda, double* part, HT* F_val, HT* Jinv,
                                double* bschur, double* Sdiag, int* err, hipStream_t st) {
  if (g.Cf == 0) return hipSuccess;
  const size_t lds = (27 * sizeof(double) + sizeof(int)) * static_cast<size_t>(g.lc_maxc);
  if (!lds_fits(k_schur_diag_pm, lds)) return hipErrorInvalidValue;
  if (g.nlc > 0)
    hipLaunchKernelGGL(k_schur_diag_pm, dim3(g.nlc), dim3(LIN_THREADS), lds, st, g, camRt, pts, Hll, b, lambda, part,
                       err);
  hipLaunchKernelGGL(k_schur_diag_finish<HT>, dim3(g.Cf), dim3(64), 0, st, g, part, Hpp, b, lambda, F_val, Jinv,
                     bschur, Sdiag, err);
  return hipGetLastError();
}
template hipError_t launch_schur_diag_pm<float>(const DevGraph&, const double*, const double*, const double*,
                                                const double*, const double*, double, double*, float*, float*,
                                                double*, double*, int*, hipStream_t);
template hipError_t launch_schur_diag_pm<double>(const DevGraph&, const double*, const double*, const double*,
                                                 const double*, const double*, double, double*, double*, double*,
                                                 double*, double*, int*, hipStream_t);

template <typename HT>
hipError_t launch_schur_diag(const DevGraph& g, const int* q_slot, const double* camRt, const double* pts,
                             const HT* Hpl, const double* Hll, const double* Hpp, const double* b, double lambda,
                             HT* F_val, HT* Jinv, double* bschur, int* err, hipStream_t st) {
  if (g.Cf == 0) return hipSuccess;
  hipLaunchKernelGGL(k_schur_diag<HT>, dim3(g.Cf), dim3(256), 0, st, g, q_slot, camRt, pts, Hpl, Hll, Hpp, b, lambda,
                     F_val, Jinv, bschur, err);
  return hipGetLastError();
}
template hipError_t launch_schur_diag<float>(const DevGraph&, const int*, const double*, const double*, const float*,
                                             const double*, const double*, const double*, double, float*, float*,
                                             double*, int*, hipStream_t);
template hipError_t launch_schur_diag<double>(const DevGraph&, const int*, const double*, const double*,
                                              const double*, const double*, const double*, const double*, double,
                                              double*, double*, double*, int*, hipStream_t);

// ----------------------------------------------------- block-Jacobi PCG
// LinearSolverPCG::solve (linear_solver_pcg.hpp:112-155) as two launches per
// iteration. Direction update, convergence test and SpMV are fused: the new
// direction d = s + beta d_old is formed on the fly from s and the previous
// direction (double-buffered by iteration parity), so no kernel has to wait
// for a separate direction pass. Dot products are per-workgroup FP64 partials
// that the next kernel sums with compensation, in a fixed order, in
// wavefront 0 of each of its workgroups (which broadcasts the scalar through
// LDS). Few, wide workgroups keep that redundant summation small: every
// workgroup of one kernel reads every partial of the other, so its cost
// grows with the product of the two grid sizes.

// Cross-lane moves within rows of 16 lanes (DPP) and reads of single lanes:
// the scalar reductions run on the VALU without an LDS round trip.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp(static_cast<int>(b), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(static_cast<int>(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
__device__ __forceinline__ double lane_value(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), l);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), l);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
// DPP controls: quad_perm [1,0,3,2] and [2,3,0,1] (partners 1 and 2 apart),
// half-row mirror (the other quad of 8 lanes), row mirror (the other half of
// a row of 16).
constexpr int DPP_QUAD_X1 = 0xB1, DPP_QUAD_X2 = 0x4E, DPP_HALF_MIRROR = 0x141, DPP_ROW_MIRROR = 0x140;

// Sum over the wavefront with the same bits in every lane: each DPP step
// adds two equal-shaped partial sums (commutative, so both partners agree),
// then the four row sums are added in a fixed order.
__device__ __forceinline__ double wave_total(double v) {
  v += dpp_move<DPP_QUAD_X1>(v);
  v += dpp_move<DPP_QUAD_X2>(v);
  v += dpp_move<DPP_HALF_MIRROR>(v);
  v += dpp_move<DPP_ROW_MIRROR>(v);
  return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
}

// (s, c) += (os, oc) with the rounding error of s + os kept in c; symmetric
// in its two operands, so partner lanes compute the same bits.
__device__ __forceinline__ void neumaier_combine(double& s, double& c, double os, double oc) {
  const double t = s + os;
  const double e = (fabs(s) >= fabs(os)) ? ((s - t) + os) : ((os - t) + s);
  c = (c + oc) + e;
  s = t;
}
template <int CTRL>
__device__ __forceinline__ void neumaier_dpp(double& s, double& c) {
  neumaier_combine(s, c, dpp_move<CTRL>(s), dpp_move<CTRL>(c));
}
__device__ __forceinline__ double wave_total_compensated(double s, double c) {
  neumaier_dpp<DPP_QUAD_X1>(s, c);
  neumaier_dpp<DPP_QUAD_X2>(s, c);
  neumaier_dpp<DPP_HALF_MIRROR>(s, c);
  neumaier_dpp<DPP_ROW_MIRROR>(s, c);
  double S = lane_value(s, 0), C = lane_value(c, 0);
  neumaier_combine(S, C, lane_value(s, 16), lane_value(c, 16));
  neumaier_combine(S, C, lane_value(s, 32), lane_value(c, 32));
  neumaier_combine(S, C, lane_value(s, 48), lane_value(c, 48));
  return S + C;
}

// The other kernel's partials, loaded by the caller before anything that
// depends on them so that their latency overlaps its other loads: one batch
// of PCG_PART_BATCH loads per lane covers 64 * PCG_PART_BATCH partials.
constexpr int PCG_PART_BATCH = 8;
struct PcgPartials {
  double v[PCG_PART_BATCH];
};
__device__ __forceinline__ PcgPartials pcg_part_load(const double* __restrict__ part, int n) {
  const int lane = threadIdx.x & 63;
  PcgPartials p;
#pragma unroll
  for (int j = 0; j < PCG_PART_BATCH; ++j) p.v[j] = part[min(lane + 64 * j, n - 1)];
  return p;
}
// Compensated sum of part[0..n), the same value in every lane: lane l takes
// partials l, l + 64, ... in index order, then the lanes combine.
__device__ double pcg_part_sum(const PcgPartials& p, const double* __restrict__ part, int n) {
  const int lane = threadIdx.x & 63;
  double s = 0, c = 0;
#pragma unroll
  for (int j = 0; j < PCG_PART_BATCH; ++j) neumaier_add(s, c, lane + 64 * j < n ? p.v[j] : 0.0);
  for (int i0 = lane + 64 * PCG_PART_BATCH; i0 < n; i0 += 64 * PCG_PART_BATCH) {
    double v[PCG_PART_BATCH];
#pragma unroll
    for (int j = 0; j < PCG_PART_BATCH; ++j) v[j] = part[min(i0 + 64 * j, n - 1)];
#pragma unroll
    for (int j = 0; j < PCG_PART_BATCH; ++j) neumaier_add(s, c, i0 + 64 * j < n ? v[j] : 0.0);
  }
  return wave_total_compensated(s, c);
}

// Orders LDS traffic within one wavefront (written by some lanes, read by
// others of the same wavefront).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double pcg_dir_val(double s, double d_old, double beta, bool first) {
  return first ? s : fma(beta, d_old, s);
}

// Vectors (x, r, d, q, s) and all recurrences are FP64; the matrix F_val and
// the block-Jacobi inverses J are stored in MT (FP32 or FP64). With MT =
// float the iteration is FP64 PCG on the FP32-rounded operator: the Krylov
// process keeps its FP64 orthogonality, and the bandwidth-heavy data (the
// matrix) moves at half width.

// Row-parallel vector kernels (init, update) give each block row LPR lanes
// (8 for blocks of up to 8, else 16; lane k < BD owns component k), RPW rows
// per workgroup of PCG_ROW_THREADS threads; the preconditioner product gathers
// the row's r components by lane shuffles. One partial per workgroup.
constexpr int PCG_ROW_THREADS = 512;
constexpr int PCG_ROW_WAVES = PCG_ROW_THREADS / 64;
template <int BD>
struct PcgRowShape {
  static_assert(BD >= 1 && BD <= 16, "block size");
  static constexpr int LPR = BD <= 8 ? 8 : 16;
  static constexpr int RPW = PCG_ROW_THREADS / LPR;
};

// Sum of one value per wavefront of a PCG_ROW_THREADS workgroup, in
// wavefront order, by thread 0.
__device__ __forceinline__ void pcg_block_partial(double wave_value, double* s_w, double* part) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) s_w[wv] = wave_value;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < PCG_ROW_WAVES; ++w) t += s_w[w];
    part[blockIdx.x] = t;
  }
}

// Stores of the solve's progress into host-mapped memory (system scope), so
// the host can follow the iteration without synchronising the stream.
__device__ __forceinline__ void progress_store(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
// The final values, then done = 1 (release: the values are visible first).
__device__ __forceinline__ void progress_done(PcgProgress* p, int iters, int err, int ab, double dn, double d0,
                                              double dn0) {
  p->iters = iters;
  p->err = err;
  p->abort = ab;
  p->dn = dn;
  p->d0 = d0;
  p->dn0 = dn0;
  __hip_atomic_store(&p->done, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ void k_publish_scalars(const double* __restrict__ scal, const int* __restrict__ err,
                                  TrialScalars* __restrict__ out, int seq) {
  if (threadIdx.x != 0) return;
  for (int k = 0; k < 4; ++k) out->v[k] = scal[k];
  out->err = *err;
  __hip_atomic_store(&out->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

hipError_t launch_publish_scalars(const double* scal, const int* err, TrialScalars* out, int seq, hipStream_t st) {
  hipLaunchKernelGGL(k_publish_scalars, dim3(1), dim3(64), 0, st, scal, err, out, seq);
  return hipGetLastError();
}

// x = 0, r = b, s = M^-1 r, part[block] = r . s   (linear_solver_pcg.hpp:120-130);
// workgroup 0 also resets the solve's scalars.
template <typename MT, int BD>
__global__ __launch_bounds__(PCG_ROW_THREADS) void k_pcg_init(int nb, const double* __restrict__ b,
                                                              const MT* __restrict__ J, double* __restrict__ x,
                                                              double* __restrict__ r, double* __restrict__ s,
                                                              double* __restrict__ part,
                                                              PcgScalars* __restrict__ sc) {
  __shared__ double s_w[PCG_ROW_WAVES];
  const int lane = threadIdx.x & 63;
  if (blockIdx.x == 0 && threadIdx.x == 0) *sc = PcgScalars{};
  constexpr int LPR = PcgRowShape<BD>::LPR;
  const int i = blockIdx.x * PcgRowShape<BD>::RPW + (threadIdx.x / LPR), k = threadIdx.x & (LPR - 1);
  const bool on = i < nb && k < BD;
  // loads from clamped indices (all in flight together), values selected
  const int ic = on ? i : 0, kc = on ? k : 0;
  const size_t o = BD * static_cast<size_t>(ic) + kc;
  double jr[BD];
#pragma unroll
  for (int c = 0; c < BD; ++c) jr[c] = static_cast<double>(J[static_cast<size_t>(ic) * BD * BD + BD * kc + c]);
  const double bv = b[o];
  const double rv = on ? bv : 0.0;
  if (on) {
    r[o] = rv;
    x[o] = 0.0;
  }
  double v = 0.0;
#pragma unroll
  for (int c = 0; c < BD; ++c) v = fma(jr[c], __shfl(rv, (lane & ~(LPR - 1)) + c, 64), v);
  double dot = 0.0;
  if (on) {
    s[o] = v;
    dot = rv * v;
  }
  pcg_block_partial(wave_total(dot), s_w, part);
}

// Head of iteration `it`: dn = r . s from the previous update's partials
// (loaded by the caller); the stop test `dn <= d0` of
// linear_solver_pcg.hpp:141 (d0 fixed at it == 0 as tol * dn0 floored by the
// residual memory, or d0_abs when >= 0); then beta = dn / dn_old. Returns
// false (and records the iteration count) when the solve is over. Run by one
// wavefront per workgroup, identically in every workgroup; workgroup 0
// publishes the progress (iteration reached, done) to host-mapped memory.
__device__ bool pcg_head(const PcgPartials& pr, const double* __restrict__ part_rs, int n_rs, PcgScalars* sc,
                         PcgProgress* prog, const int* abort, double tol, double resid_mem, double d0_abs, int it,
                         int last, double* beta) {
  const int done0 = sc->done;
  const int ab = abort ? *abort : 0;
  const int err0 = sc->err;
  const double d0_prev = sc->d0, dn0_prev = sc->dn0;
  const double dn_old = sc->dn[(it + 1) & 1];
  const double dn = pcg_part_sum(pr, part_rs, n_rs);
  if (done0) return false;
  const bool first = blockIdx.x == 0 && threadIdx.x == 0;
  if (ab) {
    if (first) {
      sc->done = 1;
      sc->iters = it;
      if (prog) progress_done(prog, it, err0, ab, dn, d0_prev, dn0_prev);
    }
    return false;
  }
  double d0;
  if (it == 0) {
    if (d0_abs >= 0.0) {
      d0 = d0_abs;
    } else {
      d0 = tol * dn;
      if (resid_mem > 0.0 && resid_mem > d0) d0 = resid_mem;
    }
  } else {
    d0 = d0_prev;
  }
  if (first) {
    if (it == 0) { sc->d0 = d0; sc->dn0 = dn; }
    sc->dn[it & 1] = dn;  // the other slot still holds dn_old for this launch
    if (!(dn >= 0.0)) sc->err = 1;  // indefinite preconditioner (reported, not acted on)
  }
  if (dn <= d0 || last) {
    if (first) {
      sc->done = 1;
      sc->iters = it;
      if (prog)
        progress_done(prog, it, err0 | (dn >= 0.0 ? 0 : 1), 0, dn, d0, it == 0 ? dn : dn0_prev);
    }
    return false;
  }
  if (first && prog) progress_store(&prog->it, it);
  *beta = it == 0 ? 0.0 : dn / dn_old;
  return true;
}

// Explicit operator: d = s + beta d_old for the owned rows and q = A d over
// the full block rows (both triangles). PCG_SPMV_ROWS block rows per
// workgroup, one wavefront per row. A row is taken in chunks of CH blocks
// (up to 128): lane t of the wavefront loads the columns of blocks t and
// t + 64 and gathers s and d_old there once, forms d = s + beta d_old into
// LDS, and the lanes, as
// groups of BD (lane k of a group owns output row k), then multiply their
// blocks' rows by d read back from LDS. The matrix entries of the first
// chunk and its gathered vectors are loaded before the iteration head (run
// by wavefront 0, whose partial loads are issued first of all), so the
// matrix, the gathers and the head's reduction are in flight together.
// part_dq[workgroup] = d . q over its rows.
// HS: block slots per group per chunk (CH = min(128, HS * groups)): 13 for
// BA rows (tens of blocks), 1 for pose-graph rows (a handful).
constexpr int PCG_SPMV_ROWS = 8;
constexpr int PCG_SPMV_THREADS = 64 * PCG_SPMV_ROWS;
template <int BD, int HS>
struct SpmvShape {
  static constexpr int GPW = 64 / BD;  // groups per wavefront
  static constexpr int CH = 128 < GPW * HS ? 128 : GPW * HS;
  static constexpr int LB = (CH + 63) / 64;  // column slots per lane
};
// FIRST: iteration 0 (d = s; no d_old, no beta), a separate instantiation so
// the other iterations carry no run-time branch on it.
// D64: the diagonal blocks are applied from Ddiag (FP64, BD*BD per row) and
// F_val's diagonal blocks hold zeros. The FP32 operator keeps its diagonal in
// FP64 this way: Hpp_ii - sum V Dinv V^T is the largest entry of a row, and
// rounding it to FP32 is the largest perturbation of the reduced system.
template <typename MT, int BD, bool FIRST, bool D64, int HS>
__global__ __launch_bounds__(PCG_SPMV_THREADS) void k_pcg_dir_spmv(
    int nb, const int* __restrict__ F_ptr, const int* __restrict__ F_col, const MT* __restrict__ F_val,
    const double* __restrict__ Ddiag, const double* __restrict__ s, const double* __restrict__ d_old,
    double* __restrict__ d, double* __restrict__ q, const double* __restrict__ part_rs, int n_rs,
    double* __restrict__ part_dq, PcgScalars* sc, PcgProgress* prog, const int* abort, double tol, double resid_mem,
    double d0_abs, int it, int last) {
  constexpr int GPW = SpmvShape<BD, HS>::GPW;
  constexpr int CH = SpmvShape<BD, HS>::CH;
  constexpr int LB = SpmvShape<BD, HS>::LB;
  constexpr int R = PCG_SPMV_ROWS;
  __shared__ double s_dd[R][CH][BD];
  __shared__ double s_acc[R][64];
  __shared__ double s_dot[R];
  __shared__ double s_beta;
  __shared__ int s_go;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  PcgPartials pr;
  if (wv == 0) pr = pcg_part_load(part_rs, n_rs);
  const int row = blockIdx.x * R + wv;
  const bool act = row < nb;
  const int rw = act ? row : 0;
  const int gw = lane / BD, k = lane % BD;
  // own-row operands of the final stage (lane kk owns component kk)
  const int kk = lane < BD ? lane : 0;
  const size_t o_own = BD * static_cast<size_t>(rw) + kk;
  const double s_own = s[o_own];
  const double dold_own = FIRST ? 0.0 : d_old[o_own];
  double Dg[BD];
  if (D64) {
#pragma unroll
    for (int c = 0; c < BD; ++c) Dg[c] = Ddiag[static_cast<size_t>(rw) * BD * BD + BD * kk + c];
  }
  const int p0 = F_ptr[rw], p1 = act ? F_ptr[rw + 1] : p0;
  const int len = p1 - p0;
  int nc = min(CH, len);
  double sg[LB][BD], dg[LB][BD];
  MT v[HS][BD];
  // every load unconditional from a clamped (valid) index, values selected
  // afterwards, so all of them are in flight together
  auto load_chunk = [&](int c0) {
#pragma unroll
    for (int lb = 0; lb < LB; ++lb) {
      const int t = lane + 64 * lb;
      const size_t j = BD * static_cast<size_t>(F_col[p0 + c0 + (t < nc ? t : 0)]);
#pragma unroll
      for (int c = 0; c < BD; ++c) {
        sg[lb][c] = s[j + c];
        if (!FIRST) dg[lb][c] = d_old[j + c];
      }
    }
#pragma unroll
    for (int h = 0; h < HS; ++h) {
      const int bk = gw + h * GPW;
      const bool on = gw < GPW && bk < nc;
      const MT* V = F_val + static_cast<size_t>(p0 + c0 + (on ? bk : 0)) * BD * BD + BD * k;
#pragma unroll
      for (int c = 0; c < BD; ++c) {
        const MT vv = V[c];
        v[h][c] = on ? vv : MT(0);
      }
    }
  };
  load_chunk(0);
  if (wv == 0) {
    double beta = 0.0;
    const bool go = pcg_head(pr, part_rs, n_rs, sc, prog, abort, tol, resid_mem, d0_abs, it, last, &beta);
    if (lane == 0) {
      s_go = go;
      s_beta = beta;
    }
  }
  __syncthreads();
  if (!s_go) return;
  const double beta = s_beta;
  double acc = 0.0;
  for (int c0 = 0; c0 < len; c0 += CH) {
    if (c0 > 0) {
      nc = min(CH, len - c0);
      load_chunk(c0);
      wave_sync();  // the previous chunk's reads of s_dd are done
    }
#pragma unroll
    for (int lb = 0; lb < LB; ++lb) {
      const int t = lane + 64 * lb;
      if (t < nc) {
#pragma unroll
        for (int c = 0; c < BD; ++c) s_dd[wv][t][c] = FIRST ? sg[lb][c] : fma(beta, dg[lb][c], sg[lb][c]);
      }
    }
    wave_sync();
#pragma unroll
    for (int h = 0; h < HS; ++h) {
      const int bk = gw + h * GPW;
      const int bc = bk < nc ? bk : 0;
#pragma unroll
      for (int c = 0; c < BD; ++c) acc = fma(static_cast<double>(v[h][c]), s_dd[wv][bc][c], acc);
    }
  }
  s_acc[wv][lane] = acc;
  wave_sync();
  double vs = 0.0;
  if (lane < BD)
    for (int g2 = 0; g2 < GPW; ++g2) vs += s_acc[wv][g2 * BD + lane];
  const double dd = FIRST ? s_own : fma(beta, dold_own, s_own);
  if (D64) {
    // every lane of the wavefront takes part in the shuffles
    double dgv = 0.0;
#pragma unroll
    for (int c = 0; c < BD; ++c) dgv = fma(Dg[c], __shfl(dd, c, 64), dgv);
    vs += dgv;
  }
  double dot = 0.0;
  if (act && lane < BD) {
    q[o_own] = vs;
    d[o_own] = dd;
    dot = vs * dd;
  }
  dot = wave_total(dot);
  if (lane == 0) s_dot[wv] = dot;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < R; ++w) t += s_dot[w];
    part_dq[blockIdx.x] = t;
  }
}

// Direction only (matrix-free operator): d = s + beta d_old, thread per row.
template <int BD>
__global__ __launch_bounds__(64) void k_pcg_dir(int nb, const double* __restrict__ s,
                                                const double* __restrict__ d_old, double* __restrict__ d,
                                                const double* __restrict__ part_rs, int n_rs, PcgScalars* sc,
                                                PcgProgress* prog, const int* abort, double tol, double resid_mem,
                                                double d0_abs, int it, int last) {
  const PcgPartials pr = pcg_part_load(part_rs, n_rs);
  double beta;
  if (!pcg_head(pr, part_rs, n_rs, sc, prog, abort, tol, resid_mem, d0_abs, it, last, &beta)) return;
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= nb) return;
  for (int k = 0; k < BD; ++k) {
    const size_t o = BD * static_cast<size_t>(i) + k;
    d[o] = pcg_dir_val(s[o], it == 0 ? 0.0 : d_old[o], beta, it == 0);
  }
}

// ---------------------------------------------- one launch per iteration
// LinearSolverPCG::solve (linear_solver_pcg.hpp:137-153) on the explicit
// operator with the update of iteration L-1 folded into launch L, so an
// iteration is one launch instead of two. With w = M^-1 q (M symmetric),
// r' = r - alpha q and s' = M^-1 r' = s - alpha w, so
//   dn' = r'.s' = r.s - 2 alpha s.q + alpha^2 q.w
// follows from dot products of the vectors launch L-1 already holds. Launch
// L >= 1 sums launch L-1's partials d.q, s.q, q.w and r.s, forms
// alpha = r.s / d.q and dn', applies x += alpha d, r -= alpha q,
// s -= alpha w to its own rows, takes the stop test on dn' and
// beta = dn' / r.s, gathers s - alpha w and d at the columns it needs for
// d' = s' + beta d, then forms q' = A d' and w' = M^-1 q' for its rows and
// the partials d'.q', s'.q', q'.w', r'.s'. The r.s passed on is summed from
// the updated vectors, so dn's formula never feeds on its own output.
// Launch 0 takes dn = r.s of k_pcg_init's partials and d = s.
// s, w and d alternate between two buffers by launch parity (launch L
// writes index L & 1 and gathers from the other); x, r and q are updated in
// place for the workgroup's own rows; the partials alternate as well
// ([parity][4][workgroups]). A solve that ended at launch L'
// (sc->done = L' + 1) makes later launches return at once, while the other
// workgroups of launch L' itself still apply their rows' update.
template <int NB>
struct PcgPartialsN {
  double v[NB];
};
template <int NB>
__device__ __forceinline__ PcgPartialsN<NB> pcg_part_load_n(const double* __restrict__ part, int n) {
  const int lane = threadIdx.x & 63;
  PcgPartialsN<NB> p;
#pragma unroll
  for (int j = 0; j < NB; ++j) p.v[j] = part[min(lane + 64 * j, n - 1)];
  return p;
}
template <int NB>
__device__ double pcg_part_sum_n(const PcgPartialsN<NB>& p, const double* __restrict__ part, int n) {
  const int lane = threadIdx.x & 63;
  double s = 0, c = 0;
#pragma unroll
  for (int j = 0; j < NB; ++j) neumaier_add(s, c, lane + 64 * j < n ? p.v[j] : 0.0);
  for (int i = lane + 64 * NB; i < n; i += 64) neumaier_add(s, c, part[i]);
  return wave_total_compensated(s, c);
}
constexpr int PCG_FUSED_PB = 4;  // partial batches per lane loaded ahead (256 workgroups)

template <typename MT, int BD, bool FIRST, bool D64, int HS>
__global__ __launch_bounds__(PCG_SPMV_THREADS) void k_pcg_fused(
    int nb, const int* __restrict__ F_head, const int* __restrict__ F_col, const MT* __restrict__ F_val,
    const double* __restrict__ Ddiag, const MT* __restrict__ J, double* __restrict__ x, double* __restrict__ r,
    double* __restrict__ q, const double* __restrict__ s_prev, double* __restrict__ s_next,
    const double* __restrict__ w_prev, double* __restrict__ w_next, const double* __restrict__ d_prev,
    double* __restrict__ d_next, const double* __restrict__ part_prev, int n_prev, double* __restrict__ part_next,
    PcgScalars* sc, PcgProgress* prog, const int* abort, double tol, double resid_mem, double d0_abs, int it,
    int last) {
  constexpr int GPW = SpmvShape<BD, HS>::GPW;
  constexpr int CH = SpmvShape<BD, HS>::CH;
  constexpr int LB = SpmvShape<BD, HS>::LB;
  constexpr int R = PCG_SPMV_ROWS;
  constexpr int NP = PCG_FUSED_PB;
  __shared__ double s_dd[R][CH][BD];
  __shared__ double s_acc[R][64];
  __shared__ double s_dot[R][4];
  __shared__ double s_ab[2];
  __shared__ int s_state;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // launch L-1's partials (FIRST: k_pcg_init's r.s), first of all
  PcgPartialsN<NP> p_dq, p_sq, p_qw, p_rs;
  if (wv == 0) {
    if (FIRST) {
      p_rs = pcg_part_load_n<NP>(part_prev, n_prev);
    } else {
      p_dq = pcg_part_load_n<NP>(part_prev, n_prev);
      p_sq = pcg_part_load_n<NP>(part_prev + n_prev, n_prev);
      p_qw = pcg_part_load_n<NP>(part_prev + 2 * static_cast<size_t>(n_prev), n_prev);
      p_rs = pcg_part_load_n<NP>(part_prev + 3 * static_cast<size_t>(n_prev), n_prev);
    }
  }
  const int row = blockIdx.x * R + wv;
  const bool act = row < nb;
  const int rw = act ? row : 0;
  const int gw = lane / BD, k = lane % BD;
  // own-row operands (lane kk < BD owns component kk)
  const int kk = lane < BD ? lane : 0;
  const size_t o_own = BD * static_cast<size_t>(rw) + kk;
  const double s_own = s_prev[o_own];
  const double r_own = r[o_own];
  const double w_own = FIRST ? 0.0 : w_prev[o_own];
  const double d_own = FIRST ? 0.0 : d_prev[o_own];
  const double q_own = FIRST ? 0.0 : q[o_own];
  const double x_own = FIRST ? 0.0 : x[o_own];
  double Jr[BD];
#pragma unroll
  for (int c = 0; c < BD; ++c) Jr[c] = static_cast<double>(J[static_cast<size_t>(rw) * BD * BD + BD * kk + c]);
  double Dg[BD];
  if (D64) {
#pragma unroll
    for (int c = 0; c < BD; ++c) Dg[c] = Ddiag[static_cast<size_t>(rw) * BD * BD + BD * kk + c];
  }
  // the row's head (first block, length, first columns) in one round trip:
  // the matrix entries and the gathers both depend only on it
  const int* hd = F_head + static_cast<size_t>(rw) * PCG_HEAD_STRIDE;
  const int p0 = hd[0];
  const int len = act ? hd[1] : 0;
  int col0[LB];
#pragma unroll
  for (int lb = 0; lb < LB; ++lb) col0[lb] = hd[2 + min(lane + 64 * lb, PCG_HEAD_COLS - 1)];
  int nc = min(CH, len);
  double sg[LB][BD], wg[LB][BD], dg[LB][BD];
  MT v[HS][BD];
  auto load_chunk = [&](int c0) {
#pragma unroll
    for (int lb = 0; lb < LB; ++lb) {
      const int t = lane + 64 * lb;
      const size_t j = BD * static_cast<size_t>(c0 == 0 ? col0[lb] : F_col[p0 + c0 + (t < nc ? t : 0)]);
#pragma unroll
      for (int c = 0; c < BD; ++c) {
        sg[lb][c] = s_prev[j + c];
        if (!FIRST) {
          wg[lb][c] = w_prev[j + c];
          dg[lb][c] = d_prev[j + c];
        }
      }
    }
#pragma unroll
    for (int h = 0; h < HS; ++h) {
      const int bk = gw + h * GPW;
      const bool on = gw < GPW && bk < nc;
      const MT* V = F_val + static_cast<size_t>(p0 + c0 + (on ? bk : 0)) * BD * BD + BD * k;
#pragma unroll
      for (int c = 0; c < BD; ++c) {
        const MT vv = V[c];
        v[h][c] = on ? vv : MT(0);
      }
    }
  };
  load_chunk(0);
  if (wv == 0) {
    const int dit = sc->done;
    const int ab = abort ? *abort : 0;
    const int err0 = sc->err;
    const double d0_prev = sc->d0, dn0_prev = sc->dn0;
    double rs, alpha = 0.0, dn, dq = 1.0;
    if (FIRST) {
      rs = pcg_part_sum_n<NP>(p_rs, part_prev, n_prev);
      dn = rs;
    } else {
      dq = pcg_part_sum_n<NP>(p_dq, part_prev, n_prev);
      const double sq = pcg_part_sum_n<NP>(p_sq, part_prev + n_prev, n_prev);
      const double qw = pcg_part_sum_n<NP>(p_qw, part_prev + 2 * static_cast<size_t>(n_prev), n_prev);
      rs = pcg_part_sum_n<NP>(p_rs, part_prev + 3 * static_cast<size_t>(n_prev), n_prev);
      alpha = rs / dq;
      dn = (rs - 2.0 * alpha * sq) + alpha * alpha * qw;
    }
    const bool first = blockIdx.x == 0 && lane == 0;
    int state;
    double beta = 0.0;
    if (dit != 0 && dit - 1 < it) {
      state = 0;  // ended at an earlier launch
    } else if (ab) {
      state = 0;
      if (first) {
        sc->done = it + 1;
        sc->iters = it;
        if (prog) progress_done(prog, it, err0, ab, dn, d0_prev, dn0_prev);
      }
    } else {
      double d0;
      if (FIRST) {
        if (d0_abs >= 0.0) {
          d0 = d0_abs;
        } else {
          d0 = tol * dn;
          if (resid_mem > 0.0 && resid_mem > d0) d0 = resid_mem;
        }
      } else {
        d0 = d0_prev;
      }
      if (first) {
        if (FIRST) { sc->d0 = d0; sc->dn0 = dn; }
        sc->dn[it & 1] = dn;
        if (!(dn >= 0.0)) sc->err = 1;  // indefinite preconditioner (reported, not acted on)
        // d.q <= 0: the operator is not positive definite along d (reported
        // in sc->err; the iteration itself follows linear_solver_pcg.hpp)
        if (!FIRST && !(dq > 0.0)) sc->err = 1;
      }
      if (dn <= d0 || last) {
        state = 1;  // this launch's update, then the end
        if (first) {
          sc->done = it + 1;
          sc->iters = it;
          if (prog) progress_done(prog, it, err0 | (dn >= 0.0 ? 0 : 1), 0, dn, d0, FIRST ? dn : dn0_prev);
        }
      } else {
        state = 2;
        beta = FIRST ? 0.0 : dn / rs;
        if (first && prog) progress_store(&prog->it, it);
      }
    }
    if (lane == 0) {
      s_state = state;
      s_ab[0] = alpha;
      s_ab[1] = beta;
    }
  }
  __syncthreads();
  const int state = s_state;
  if (state == 0) return;
  const double alpha = s_ab[0], beta = s_ab[1];
  // this workgroup's rows: the update of launch L-1
  const bool own = act && lane < BD;
  double r_new = r_own, s_new = s_own;
  if (!FIRST) {
    r_new = fma(-alpha, q_own, r_own);
    s_new = fma(-alpha, w_own, s_own);
    if (own) {
      x[o_own] = fma(alpha, d_own, x_own);
      r[o_own] = r_new;
      s_next[o_own] = s_new;
    }
  }
  if (state == 1) return;
  double acc = 0.0;
  for (int c0 = 0; c0 < len; c0 += CH) {
    if (c0 > 0) {
      nc = min(CH, len - c0);
      load_chunk(c0);
      wave_sync();  // the previous chunk's reads of s_dd are done
    }
#pragma unroll
    for (int lb = 0; lb < LB; ++lb) {
      const int t = lane + 64 * lb;
      if (t < nc) {
#pragma unroll
        for (int c = 0; c < BD; ++c)
          s_dd[wv][t][c] = FIRST ? sg[lb][c] : fma(beta, dg[lb][c], fma(-alpha, wg[lb][c], sg[lb][c]));
      }
    }
    wave_sync();
#pragma unroll
    for (int h = 0; h < HS; ++h) {
      const int bk = gw + h * GPW;
      const int bc = bk < nc ? bk : 0;
#pragma unroll
      for (int c = 0; c < BD; ++c) acc = fma(static_cast<double>(v[h][c]), s_dd[wv][bc][c], acc);
    }
  }
  s_acc[wv][lane] = acc;
  wave_sync();
  double vs = 0.0;
  if (lane < BD)
    for (int g2 = 0; g2 < GPW; ++g2) vs += s_acc[wv][g2 * BD + lane];
  const double dd = FIRST ? s_own : fma(beta, d_own, s_new);
  if (D64) {
    // every lane of the wavefront takes part in the shuffles
    double dgv = 0.0;
#pragma unroll
    for (int c = 0; c < BD; ++c) dgv = fma(Dg[c], __shfl(dd, c, 64), dgv);
    vs += dgv;
  }
  // w' = M^-1 q' for the row (q' components from lanes 0..BD-1)
  double wn = 0.0;
#pragma unroll
  for (int c = 0; c < BD; ++c) wn = fma(Jr[c], __shfl(vs, c, 64), wn);
  double t_dq = 0.0, t_sq = 0.0, t_qw = 0.0, t_rs = 0.0;
  if (own) {
    q[o_own] = vs;
    d_next[o_own] = dd;
    w_next[o_own] = wn;
    t_dq = dd * vs;
    t_sq = s_new * vs;
    t_qw = vs * wn;
    t_rs = r_new * s_new;
  }
  t_dq = wave_total(t_dq);
  t_sq = wave_total(t_sq);
  t_qw = wave_total(t_qw);
  t_rs = wave_total(t_rs);
  if (lane == 0) {
    s_dot[wv][0] = t_dq;
    s_dot[wv][1] = t_sq;
    s_dot[wv][2] = t_qw;
    s_dot[wv][3] = t_rs;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double t = 0.0;
    for (int w = 0; w < R; ++w) t += s_dot[w][threadIdx.x];
    part_next[static_cast<size_t>(threadIdx.x) * gridDim.x + blockIdx.x] = t;
  }
}

// a = dn / d.q; x += a d; r -= a q; s = M^-1 r; part_rs[block] = r . s
template <typename MT, int BD>
__global__ __launch_bounds__(PCG_ROW_THREADS) void k_pcg_update(
    int nb, const MT* __restrict__ J, double* __restrict__ x, double* __restrict__ r, const double* __restrict__ d,
    const double* __restrict__ q, double* __restrict__ s, const double* __restrict__ part_dq, int n_dq,
    double* __restrict__ part_rs, PcgScalars* sc, int it) {
  __shared__ double s_w[PCG_ROW_WAVES];
  __shared__ double s_a;
  __shared__ int s_go;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  PcgPartials pr;
  if (wv == 0) pr = pcg_part_load(part_dq, n_dq);
  constexpr int LPR = PcgRowShape<BD>::LPR;
  const int i = blockIdx.x * PcgRowShape<BD>::RPW + (threadIdx.x / LPR), k = threadIdx.x & (LPR - 1);
  const bool on = i < nb && k < BD;
  // operands first (independent of alpha), then the reduction of d . q;
  // loads from clamped indices so they are all in flight together
  const int ic = on ? i : 0, kc = on ? k : 0;
  const size_t o = BD * static_cast<size_t>(ic) + kc;
  const double dv = d[o], qv = q[o], rv0 = r[o], xv = x[o];
  double jr[BD];
#pragma unroll
  for (int c = 0; c < BD; ++c) jr[c] = static_cast<double>(J[static_cast<size_t>(ic) * BD * BD + BD * kc + c]);
  if (wv == 0) {
    const int done0 = sc->done;
    const double dn = sc->dn[it & 1];
    const double dq = pcg_part_sum(pr, part_dq, n_dq);
    // d . q <= 0: the operator is not positive definite along d (reported in
    // sc->err for callers that need it; the iteration itself follows
    // linear_solver_pcg.hpp, which does not check)
    if (!done0 && blockIdx.x == 0 && lane == 0 && !(dq > 0.0)) sc->err = 1;
    if (lane == 0) {
      s_go = !done0;
      s_a = dn / dq;
    }
  }
  __syncthreads();
  if (!s_go) return;
  const double a = s_a;
  double rv = 0.0;
  if (on) {
    x[o] = fma(a, dv, xv);
    rv = fma(-a, qv, rv0);
    r[o] = rv;
  }
  double v = 0.0;
#pragma unroll
  for (int c = 0; c < BD; ++c) v += (on ? jr[c] : 0.0) * __shfl(rv, (lane & ~(LPR - 1)) + c, 64);
  double dot = 0.0;
  if (on) {
    s[o] = v;
    dot = rv * v;
  }
  pcg_block_partial(wave_total(dot), s_w, part_rs);
}

int pcg_row_blocks(int nb, int bd) { return nblocks(nb, bd <= 8 ? PCG_ROW_THREADS / 8 : PCG_ROW_THREADS / 16); }
int pcg_spmv_blocks(int nb) { return nblocks(nb, PCG_SPMV_ROWS); }
int implicit_reduce_blocks(int Cf) { return nblocks(Cf, 64); }

template <typename MT, int BD>
hipError_t launch_pcg_init(int nb, const double* b, const MT* Jinv, double* x, double* r, double* s,
                           double* part_rs, PcgScalars* sc, hipStream_t st) {
  hipLaunchKernelGGL((k_pcg_init<MT, BD>), dim3(pcg_row_blocks(nb, BD)), dim3(PCG_ROW_THREADS), 0, st, nb, b, Jinv, x,
                     r, s, part_rs, sc);
  return hipGetLastError();
}

// Rows averaging fewer blocks than this take the one-slot shape (HS = 1).
constexpr int PCG_SHORT_ROW = 16;

template <typename MT, int BD>
hipError_t launch_pcg_dir_spmv(int nb, const int* F_ptr, const int* F_col, const MT* F_val, const double* s,
                               const double* d_old, double* d, double* q, const double* part_rs, double* part_dq,
                               PcgScalars* sc, const PcgControl& ctl, int it, int last, hipStream_t st,
                               const double* Ddiag, long long nnz) {
  // block slots per group for long rows: FP32 blocks a chunk of 128 blocks
  // (two column slots per lane, so a BA row of ~70 blocks is one round trip),
  // FP64 blocks 64 (the register footprint of 128 spills), fewer for blocks
  // above 8
  constexpr int CH_LONG = sizeof(MT) == 4 ? 128 : 64;
  constexpr int HS_LONG = BD <= 8 ? (CH_LONG + 64 / BD - 1) / (64 / BD) : 5;
  const bool short_rows = nnz >= 0 && nnz < static_cast<long long>(PCG_SHORT_ROW) * nb;
#define G2OHIP_SPMV_LAUNCH(FIRST, D64, HS)                                                                        \
  hipLaunchKernelGGL((k_pcg_dir_spmv<MT, BD, FIRST, D64, HS>), dim3(pcg_spmv_blocks(nb)), dim3(PCG_SPMV_THREADS), \
                     0, st, nb, F_ptr, F_col, F_val, Ddiag, s, d_old, d, q, part_rs, pcg_row_blocks(nb, BD), part_dq, \
                     sc, ctl.prog, ctl.abort, ctl.tol, ctl.resid_mem, ctl.d0_abs, it, last)
#define G2OHIP_SPMV_SHAPE(HS)                         \
  if (Ddiag) {                                        \
    if (it == 0) G2OHIP_SPMV_LAUNCH(true, true, HS);  \
    else G2OHIP_SPMV_LAUNCH(false, true, HS);         \
  } else {                                            \
    if (it == 0) G2OHIP_SPMV_LAUNCH(true, false, HS); \
    else G2OHIP_SPMV_LAUNCH(false, false, HS);        \
  }
  if (short_rows) {
    G2OHIP_SPMV_SHAPE(1)
  } else {
    G2OHIP_SPMV_SHAPE(HS_LONG)
  }
#undef G2OHIP_SPMV_SHAPE
#undef G2OHIP_SPMV_LAUNCH
  return hipGetLastError();
}

template <int BD>
hipError_t launch_pcg_dir(int nb, const double* s, const double* d_old, double* d, const double* part_rs,
                          PcgScalars* sc, const PcgControl& ctl, int it, int last, hipStream_t st) {
  hipLaunchKernelGGL((k_pcg_dir<BD>), dim3(nblocks(nb, 64)), dim3(64), 0, st, nb, s, d_old, d, part_rs,
                     pcg_row_blocks(nb, BD), sc, ctl.prog, ctl.abort, ctl.tol, ctl.resid_mem, ctl.d0_abs, it, last);
  return hipGetLastError();
}
template hipError_t launch_pcg_dir<6>(int, const double*, const double*, double*, const double*, PcgScalars*,
                                      const PcgControl&, int, int, hipStream_t);

template <typename MT, int BD>
hipError_t launch_pcg_update(int nb, const MT* Jinv, double* x, double* r, const double* d, const double* q,
                             double* s, const double* part_dq, int n_dq, double* part_rs, PcgScalars* sc, int it,
                             hipStream_t st) {
  hipLaunchKernelGGL((k_pcg_update<MT, BD>), dim3(pcg_row_blocks(nb, BD)), dim3(PCG_ROW_THREADS), 0, st, nb, Jinv, x,
                     r, d, q, s, part_dq, n_dq, part_rs, sc, it);
  return hipGetLastError();
}

#define G2OHIP_PCG_INST(MT, BD)                                                                                    \
  template hipError_t launch_jacobi<MT, BD>(int, const MT*, const int*, MT*, int*, hipStream_t);                  \
  template hipError_t launch_pcg_init<MT, BD>(int, const double*, const MT*, double*, double*, double*, double*, \
                                              PcgScalars*, hipStream_t);                                         \
  template hipError_t launch_pcg_dir_spmv<MT, BD>(int, const int*, const int*, const MT*, const double*,         \
                                                  const double*, double*, double*, const double*, double*,       \
                                                  PcgScalars*, const PcgControl&, int, int, hipStream_t,         \
                                                  const double*, long long);                                      \
  template hipError_t launch_pcg_update<MT, BD>(int, const MT*, double*, double*, const double*, const double*,  \
                                                double*, const double*, int, double*, PcgScalars*, int,          \
                                                hipStream_t);
G2OHIP_PCG_INST(float, 6)
G2OHIP_PCG_INST(double, 6)
G2OHIP_PCG_INST(double, 3)
G2OHIP_PCG_INST(double, 9)
#undef G2OHIP_PCG_INST

template <typename MT, int BD>
hipError_t launch_pcg_fused(int nb, const int* F_head, const int* F_col, const MT* F_val, const double* Ddiag,
                            const MT* J, double* x, double* r, double* q, double* const s[2], double* const w[2],
                            double* const d[2], const double* part_init, double* part, PcgScalars* sc,
                            const PcgControl& ctl, int it, int last, hipStream_t st, long long nnz) {
  const int G = pcg_spmv_blocks(nb);
  const double* s_prev = it == 0 ? s[0] : s[(it + 1) & 1];
  const double* w_prev = w[(it + 1) & 1];
  const double* d_prev = d[(it + 1) & 1];
  const double* p_prev = it == 0 ? part_init : part + 4 * static_cast<size_t>(G) * ((it + 1) & 1);
  const int n_prev = it == 0 ? pcg_row_blocks(nb, BD) : G;
  double* p_next = part + 4 * static_cast<size_t>(G) * (it & 1);
  // FP32 rows in chunks of 110 blocks (11 block slots per lane group): a
  // bundle-adjustment row of up to 110 blocks in one round trip without
  // spilling the extra gathers of this kernel
  constexpr int CH_LONG = sizeof(MT) == 4 ? 110 : 64;
  constexpr int HS_LONG = BD <= 8 ? (CH_LONG + 64 / BD - 1) / (64 / BD) : 5;
  const bool short_rows = nnz >= 0 && nnz < static_cast<long long>(PCG_SHORT_ROW) * nb;
#define G2OHIP_FUSED_LAUNCH(FIRST, D64, HS)                                                                       \
  hipLaunchKernelGGL((k_pcg_fused<MT, BD, FIRST, D64, HS>), dim3(G), dim3(PCG_SPMV_THREADS), 0, st, nb, F_head,    \
                     F_col, F_val, Ddiag, J, x, r, q, s_prev, s[it & 1], w_prev, w[it & 1], d_prev, d[it & 1],     \
                     p_prev, n_prev, p_next, sc, ctl.prog, ctl.abort, ctl.tol, ctl.resid_mem, ctl.d0_abs, it, last)
// the FP64 diagonal beside the matrix only for FP32 blocks
#define G2OHIP_FUSED_SHAPE(HS)                          \
  if constexpr (sizeof(MT) == 4) {                      \
    if (!Ddiag) return hipErrorInvalidValue;            \
    if (it == 0) G2OHIP_FUSED_LAUNCH(true, true, HS);   \
    else G2OHIP_FUSED_LAUNCH(false, true, HS);          \
  } else {                                              \
    if (Ddiag) return hipErrorInvalidValue;             \
    if (it == 0) G2OHIP_FUSED_LAUNCH(true, false, HS);  \
    else G2OHIP_FUSED_LAUNCH(false, false, HS);         \
  }
  if (short_rows) {
    G2OHIP_FUSED_SHAPE(1)
  } else {
    G2OHIP_FUSED_SHAPE(HS_LONG)
  }
#undef G2OHIP_FUSED_SHAPE
#undef G2OHIP_FUSED_LAUNCH
  return hipGetLastError();
}
template hipError_t launch_pcg_fused<float, 6>(int, const int*, const int*, const float*, const double*, const float*,
                                               double*, double*, double*, double* const[2], double* const[2],
                                               double* const[2], const double*, double*, PcgScalars*,
                                               const PcgControl&, int, int, hipStream_t, long long);
template hipError_t launch_pcg_fused<double, 6>(int, const int*, const int*, const double*, const double*,
                                                const double*, double*, double*, double*, double* const[2],
                                                double* const[2], double* const[2], const double*, double*,
                                                PcgScalars*, const PcgControl&, int, int, hipStream_t, long long);
template hipError_t launch_pcg_fused<double, 9>(int, const int*, const int*, const double*, const double*,
                                                const double*, double*, double*, double*, double* const[2],
                                                double* const[2], double* const[2], const double*, double*,
                                                PcgScalars*, const PcgControl&, int, int, hipStream_t, long long);
int pcg_fused_blocks(int nb) { return pcg_spmv_blocks(nb); }
bool pcg_fused_fits(int nb) { return pcg_fused_blocks(nb) <= 64 * PCG_FUSED_PB; }

// y = A x over 6x6 blocks: block-CSR (F_ptr, F_col, F_val) or, with F_ptr
// null, block diagonal (block i at F_val + 36 i); one thread per row of y.
__global__ void k_block_mv6(int nb, const int* __restrict__ F_ptr, const int* __restrict__ F_col,
                            const double* __restrict__ F_val, const double* __restrict__ x, double* __restrict__ y) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * nb) return;
  const int i = t / 6, r = t - 6 * i;
  const int k0 = F_ptr ? F_ptr[i] : i, k1 = F_ptr ? F_ptr[i + 1] : i + 1;
  double acc = 0.0;
  for (int k = k0; k < k1; ++k) {
    const int c = F_ptr ? F_col[k] : i;
    const double* B = F_val + 36 * static_cast<size_t>(k) + 6 * r;
    const double* v = x + 6 * static_cast<size_t>(c);
#pragma unroll
    for (int j = 0; j < 6; ++j) acc = fma(B[j], v[j], acc);
  }
  y[t] = acc;
}

hipError_t launch_block_mv6(int nb, const int* F_ptr, const int* F_col, const double* F_val, const double* x,
                            double* y, hipStream_t st) {
  if (nb <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_block_mv6, dim3(nblocks(6 * nb, 256)), dim3(256), 0, st, nb, F_ptr, F_col, F_val, x, y);
  return hipGetLastError();
}

template <typename HT>
__global__ void k_to_double(int n, const HT* __restrict__ src, double* __restrict__ dst, double scale_add) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // scale_add == 0: copy; == 1: accumulate
  dst[i] = (scale_add != 0.0 ? dst[i] : 0.0) + static_cast<double>(src[i]);
}

template <typename HT>
hipError_t launch_to_double(int n, const HT* src, double* dst, double scale_add, hipStream_t st) {
  hipLaunchKernelGGL(k_to_double<HT>, dim3(nblocks(n, 256)), dim3(256), 0, st, n, src, dst, scale_add);
  return hipGetLastError();
}
template hipError_t launch_to_double<float>(int, const float*, double*, double, hipStream_t);
template hipError_t launch_to_double<double>(int, const double*, double*, double, hipStream_t);

__global__ void k_axpy_d(int n, double a, const double* __restrict__ x, double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}

hipError_t launch_axpy_d(int n, double a, const double* x, double* y, hipStream_t st) {
  hipLaunchKernelGGL(k_axpy_d, dim3(nblocks(n, 256)), dim3(256), 0, st, n, a, x, y);
  return hipGetLastError();
}

}  // namespace g2ohip
