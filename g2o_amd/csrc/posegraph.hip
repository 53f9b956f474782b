// 3D pose-graph kernels (VertexSE3 / EdgeSE3, FP64, no Schur): per-edge
// errors and Jacobians, deterministic block assembly through fixed CSR lists,
// damping, the right-multiplied pose update and chi2.
#include <hip/hip_runtime.h>

#include "devmath.h"
#include "kernels.h"

namespace g2ohip {

namespace {
inline int nblk(long long n, int t) { return static_cast<int>((n + t - 1) / t); }

__device__ __forceinline__ double wave_sum_pg(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
}  // namespace

// Per active edge a (edge k = act[a]): e, Ji, Jj (edge_se3.cpp:60-88); the
// record rec[PG_REC * a] holds Ji^T O Ji (36), Jj^T O Jj (36), Ja^T O Jb (36,
// a = the end with the lower free index), -Ji^T O e (6), -Jj^T O e (6)
// (constructQuadraticForm, base_fixed_sized_edge.hpp:48-130, with b -= J^T O e).
// part[block] = sum of e^T O e over the block's edges.
__global__ __launch_bounds__(64) void k_pg_edges(DevPG pg, const double* __restrict__ poses,
                                                 double* __restrict__ rec, double* __restrict__ part) {
  const int a = blockIdx.x * 64 + threadIdx.x;
  double chi = 0;
  if (a < pg.A) {
    const int k = pg.act[a];
    const int i = pg.from[k], j = pg.to[k];
    double e[6], Ji[36], Jj[36];
    pg_edge_error(poses + 7 * static_cast<size_t>(i), poses + 7 * static_cast<size_t>(j),
                  pg.meas + 7 * static_cast<size_t>(k), e, Ji, Jj);
    const double* O = pg.info + 36 * static_cast<size_t>(k);
    double Oe[6], OJi[36], OJj[36];
    for (int r = 0; r < 6; ++r) {
      double s = 0;
      for (int c = 0; c < 6; ++c) s += O[6 * r + c] * e[c];
      Oe[r] = s;
      chi += e[r] * s;
      for (int c = 0; c < 6; ++c) {
        double si = 0, sj = 0;
        for (int m = 0; m < 6; ++m) {
          si += O[6 * r + m] * Ji[6 * m + c];
          sj += O[6 * r + m] * Jj[6 * m + c];
        }
        OJi[6 * r + c] = si;
        OJj[6 * r + c] = sj;
      }
    }
    double* R = rec + PG_REC * static_cast<size_t>(a);
    const bool swap = pg.edge_swap[a] != 0;
    for (int r = 0; r < 6; ++r) {
      for (int c = 0; c < 6; ++c) {
        double hii = 0, hjj = 0, hab = 0;
        for (int m = 0; m < 6; ++m) {
          hii += Ji[6 * m + r] * OJi[6 * m + c];
          hjj += Jj[6 * m + r] * OJj[6 * m + c];
          hab += swap ? Jj[6 * m + r] * OJi[6 * m + c] : Ji[6 * m + r] * OJj[6 * m + c];
        }
        R[6 * r + c] = hii;
        R[36 + 6 * r + c] = hjj;
        R[72 + 6 * r + c] = hab;
      }
      double bi = 0, bj = 0;
      for (int m = 0; m < 6; ++m) {
        bi += Ji[6 * m + r] * Oe[m];
        bj += Jj[6 * m + r] * Oe[m];
      }
      R[108 + r] = -bi;
      R[114 + r] = -bj;
    }
  }
  chi = wave_sum_pg(chi);
  if (threadIdx.x == 0 && part) part[blockIdx.x] = chi;
}

// Sums the records into the blocks: diagonal blocks into Hdiag and b, every
// upper block into both triangles of F_val (diagonal blocks undamped).
__global__ void k_pg_assemble(DevPG pg, const double* __restrict__ rec, double* __restrict__ F_val,
                              double* __restrict__ Hdiag, double* __restrict__ b) {
  const long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  const long long nblkel = 36LL * pg.nU;
  if (t < nblkel) {
    const int k = static_cast<int>(t / 36), m = static_cast<int>(t % 36);
    double v = 0;
    if (k < pg.Nf) {
      for (int q = pg.end_ptr[k]; q < pg.end_ptr[k + 1]; ++q) {
        const int ref = pg.end_ref[q];
        v += rec[PG_REC * static_cast<size_t>(ref >> 1) + ((ref & 1) ? 36 : 0) + m];
      }
      Hdiag[36 * static_cast<size_t>(k) + m] = v;
      F_val[36 * static_cast<size_t>(pg.F_diag[k]) + m] = v;
    } else {
      const int o = k - pg.Nf;
      for (int q = pg.blk_ptr[o]; q < pg.blk_ptr[o + 1]; ++q)
        v += rec[PG_REC * static_cast<size_t>(pg.blk_ref[q]) + 72 + m];
      F_val[36 * static_cast<size_t>(pg.up_pos[k]) + m] = v;
      F_val[36 * static_cast<size_t>(pg.up_posT[k]) + (m % 6) * 6 + m / 6] = v;
    }
  } else if (t < nblkel + 6LL * pg.Nf) {
    const int i = static_cast<int>((t - nblkel) / 6), r = static_cast<int>((t - nblkel) % 6);
    double v = 0;
    for (int q = pg.end_ptr[i]; q < pg.end_ptr[i + 1]; ++q) {
      const int ref = pg.end_ref[q];
      v += rec[PG_REC * static_cast<size_t>(ref >> 1) + 108 + 6 * (ref & 1) + r];
    }
    b[6 * static_cast<size_t>(i) + r] = v;
  }
}

// BlockSolver::setLambda for the diagonal blocks (block_solver.hpp:567-590)
__global__ void k_pg_damp(DevPG pg, const double* __restrict__ Hdiag, double lambda, double* __restrict__ F_val) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 36 * pg.Nf) return;
  const int i = t / 36, m = t % 36;
  F_val[36 * static_cast<size_t>(pg.F_diag[i]) + m] = Hdiag[t] + ((m % 7 == 0) ? lambda : 0.0);
}

// VertexSE3::oplusImpl (vertex_se3.h:99-108): X <- X fromVectorMQT(d), the
// rotation part (sqrt(1 - |dq|^2), dq), identity if |dq| > 1.
__global__ void k_pg_update(DevPG pg, const double* __restrict__ x, double* __restrict__ poses) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= pg.N) return;
  const int h = pg.H[i];
  if (h < 0) return;
  const double* d = x + 6 * static_cast<size_t>(h);
  double* P = poses + 7 * static_cast<size_t>(i);
  const double n2 = d[3] * d[3] + d[4] * d[4] + d[5] * d[5];
  const double dq[4] = {n2 > 1.0 ? 0.0 : d[3], n2 > 1.0 ? 0.0 : d[4], n2 > 1.0 ? 0.0 : d[5],
                        n2 > 1.0 ? 1.0 : sqrt(1.0 - n2)};
  double R[9];
  pg_quat_R(P + 3, R);
  for (int r = 0; r < 3; ++r) P[r] += R[3 * r] * d[0] + R[3 * r + 1] * d[1] + R[3 * r + 2] * d[2];
  double q[4];
  pg_quat_mul(P + 3, dq, q);
  const double n = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  for (int k = 0; k < 4; ++k) P[3 + k] = q[k] / n;
}

// computeActiveErrors + chi2 (sparse_optimizer.cpp:78-130); err_out[6 * a] if
// not null.
__global__ __launch_bounds__(64) void k_pg_chi2(DevPG pg, const double* __restrict__ poses,
                                                double* __restrict__ part, double* __restrict__ err_out) {
  const int a = blockIdx.x * 64 + threadIdx.x;
  double chi = 0;
  if (a < pg.A) {
    const int k = pg.act[a];
    double e[6];
    pg_edge_error(poses + 7 * static_cast<size_t>(pg.from[k]), poses + 7 * static_cast<size_t>(pg.to[k]),
                  pg.meas + 7 * static_cast<size_t>(k), e, nullptr, nullptr);
    const double* O = pg.info + 36 * static_cast<size_t>(k);
    for (int r = 0; r < 6; ++r) {
      double s = 0;
      for (int c = 0; c < 6; ++c) s += O[6 * r + c] * e[c];
      chi += e[r] * s;
    }
    if (err_out)
      for (int r = 0; r < 6; ++r) err_out[6 * static_cast<size_t>(a) + r] = e[r];
  }
  chi = wave_sum_pg(chi);
  if (threadIdx.x == 0) part[blockIdx.x] = chi;
}

// max |diag| of the undamped diagonal blocks (computeLambdaInit)
__global__ void k_pg_max_diag(int Nf, const double* __restrict__ Hdiag, double* __restrict__ out) {
  __shared__ double s[256];
  double m = 0;
  for (int t = threadIdx.x; t < 6 * Nf; t += 256) m = fmax(m, fabs(Hdiag[36 * (t / 6) + 7 * (t % 6)]));
  s[threadIdx.x] = m;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (static_cast<int>(threadIdx.x) < st) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

int pg_edge_blocks(int A) { return nblk(A, 64); }

hipError_t launch_pg_edges(const DevPG& pg, const double* poses, double* rec, double* part, hipStream_t st) {
  if (pg.A == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pg_edges, dim3(nblk(pg.A, 64)), dim3(64), 0, st, pg, poses, rec, part);
  return hipGetLastError();
}
hipError_t launch_pg_assemble(const DevPG& pg, const double* rec, double* F_val, double* Hdiag, double* b,
                              hipStream_t st) {
  const long long n = 36LL * pg.nU + 6LL * pg.Nf;
  hipLaunchKernelGGL(k_pg_assemble, dim3(nblk(n, 256)), dim3(256), 0, st, pg, rec, F_val, Hdiag, b);
  return hipGetLastError();
}
hipError_t launch_pg_damp(const DevPG& pg, const double* Hdiag, double lambda, double* F_val, hipStream_t st) {
  hipLaunchKernelGGL(k_pg_damp, dim3(nblk(36LL * pg.Nf, 256)), dim3(256), 0, st, pg, Hdiag, lambda, F_val);
  return hipGetLastError();
}
hipError_t launch_pg_update(const DevPG& pg, const double* x, double* poses, hipStream_t st) {
  hipLaunchKernelGGL(k_pg_update, dim3(nblk(pg.N, 256)), dim3(256), 0, st, pg, x, poses);
  return hipGetLastError();
}
hipError_t launch_pg_chi2(const DevPG& pg, const double* poses, double* part, double* err_out, hipStream_t st) {
  if (pg.A == 0) return hipSuccess;
  hipLaunchKernelGGL(k_pg_chi2, dim3(nblk(pg.A, 64)), dim3(64), 0, st, pg, poses, part, err_out);
  return hipGetLastError();
}
hipError_t launch_pg_max_diag(int Nf, const double* Hdiag, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_pg_max_diag, dim3(1), dim3(256), 0, st, Nf, Hdiag, out);
  return hipGetLastError();
}

}  // namespace g2ohip
