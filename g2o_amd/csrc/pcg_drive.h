// Host driver of one device PCG solve (shared by the bundle-adjustment and
// pose-graph solvers).
#pragma once
#include <algorithm>
#include <cstring>

#include <hip/hip_runtime.h>

#include "context.h"
#include "kernels.h"

namespace g2ohip {

// One LinearSolverPCG::solve (linear_solver_pcg.hpp:112-155) on the device:
// init, then `op` (direction + operator) and the update kernel per
// iteration. The stop test runs on the device and later launches return at
// once, so the host only polls the scalars: first after `hint` iterations
// (the previous solve's count), then every 4. If the loop reaches max_iter
// the head runs once more to record the final residual.
template <typename MT, int BD, typename OpFn>
inline hipError_t pcg_drive(hipStream_t stream, void* h_pinned, int nb, const double* rhs, const MT* J, double* x,
                            double* r, double* s, double* const dirs[2], double* q, double* part_rs,
                            double* part_dq, int n_dq, PcgScalars* sc, int max_iter, int hint, OpFn op,
                            PcgRun* out) {
  hipError_t e;
  PcgScalars hs{};
  auto poll = [&]() -> hipError_t {
    hipError_t e2 = hipMemcpyAsync(h_pinned, sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, stream);
    if (e2 == hipSuccess) e2 = hipStreamSynchronize(stream);
    if (e2 == hipSuccess) std::memcpy(&hs, h_pinned, sizeof(PcgScalars));
    return e2;
  };
  if ((e = hipMemsetAsync(sc, 0, sizeof(PcgScalars), stream)) != hipSuccess) return e;
  if ((e = launch_pcg_init<MT, BD>(nb, rhs, J, x, r, s, part_rs, stream)) != hipSuccess) return e;
  int next = std::max(4, hint);
  bool done = false;
  for (int it = 0; it < max_iter; ++it) {
    if ((e = op(it, 0, dirs[(it + 1) & 1], dirs[it & 1])) != hipSuccess) return e;
    if ((e = launch_pcg_update<MT, BD>(nb, J, x, r, dirs[it & 1], q, s, part_dq, n_dq, part_rs, sc, it, stream)) !=
        hipSuccess)
      return e;
    if (it + 1 == next) {
      if ((e = poll()) != hipSuccess) return e;
      if (hs.done) { done = true; break; }
      next += 4;
    }
  }
  if (!done) {
    if ((e = op(max_iter, 1, dirs[(max_iter + 1) & 1], dirs[max_iter & 1])) != hipSuccess) return e;
    if ((e = poll()) != hipSuccess) return e;
  }
  out->iters = hs.iters;
  out->dn_final = hs.dn[hs.iters & 1];
  out->d0 = hs.d0;
  out->dn0 = hs.dn0;
  out->breakdown = hs.err != 0;
  return hipSuccess;
}


}  // namespace g2ohip
