// Host driver of one device PCG solve (shared by the bundle-adjustment and
// pose-graph solvers).
#pragma once
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>

#include <hip/hip_runtime.h>

#include "context.h"
#include "kernels.h"

namespace g2ohip {

// One LinearSolverPCG::solve (linear_solver_pcg.hpp:112-155) on the device:
// init (which also resets the scalars), then `op` (direction + operator)
// and the update kernel per iteration. The stop test runs on the device and
// later launches return at once. With `prog` (host-mapped PcgProgress,
// written by the iteration heads) the host never waits on the stream: it
// launches until it is PCG_LAG iterations ahead of the device's progress,
// then spins on the progress word, stops launching once the device reports
// done - so the queue does not drain while the host looks, and at most
// PCG_LAG iterations are launched past the end - and reads the final values
// from the same record. If the word does not advance for PCG_STALL_MS the
// host synchronises the stream instead (which always ends the wait) and
// reports prog_failed. Without `prog` the host polls the device scalars
// every 4 iterations. If the loop reaches max_iter the head runs once more
// to record the final residual.
constexpr int PCG_LAG = 2;
constexpr double PCG_STALL_MS = 50.0;

// Host loop shared by both arrangements: init() once, then step(it, last)
// per iteration (step(max_iter, 1) only runs the final head).
template <typename InitFn, typename StepFn>
inline hipError_t pcg_loop(hipStream_t stream, void* h_pinned, volatile PcgProgress* prog, PcgScalars* sc,
                           int max_iter, InitFn init, StepFn step, PcgRun* out, bool* prog_failed,
                           const int* abort) {
  hipError_t e;
  PcgScalars hs{};
  int hw = 0;
  static_assert(sizeof(PcgScalars) <= 64, "scalars and the abort word share the pinned buffer");
  char* hp = static_cast<char*>(h_pinned);
  auto poll = [&]() -> hipError_t {
    hipError_t e2 = hipMemcpyAsync(hp, sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, stream);
    if (e2 == hipSuccess && abort) e2 = hipMemcpyAsync(hp + 64, abort, sizeof(int), hipMemcpyDeviceToHost, stream);
    if (e2 == hipSuccess) e2 = hipStreamSynchronize(stream);
    if (e2 == hipSuccess) {
      std::memcpy(&hs, hp, sizeof(PcgScalars));
      if (abort) std::memcpy(&hw, hp + 64, sizeof(int));
    }
    return e2;
  };
  // spin until the device reports done; false if the record stalled (then
  // the stream has been synchronised and hs holds the scalars)
  auto wait_done = [&](bool* stalled) -> hipError_t {
    *stalled = false;
    int seen = prog->it;
    auto last_move = std::chrono::steady_clock::now();
    for (int k = 0; !prog->done; ++k) {
      if ((k & 63) != 63) continue;
      const auto now = std::chrono::steady_clock::now();
      if (prog->it != seen) {
        seen = prog->it;
        last_move = now;
      } else if (std::chrono::duration<double, std::milli>(now - last_move).count() > PCG_STALL_MS) {
        *stalled = true;
        return poll();
      }
    }
    return hipSuccess;
  };
  if (prog) {
    prog->it = -1;
    prog->done = 0;
  }
  if ((e = init()) != hipSuccess) return e;
  int next = 4;
  bool done = false;
  bool have_final = false;  // hs already holds the final scalars
  for (int it = 0; it < max_iter; ++it) {
    if ((e = step(it, 0)) != hipSuccess) return e;
    if (prog) {
      if (prog->done) { done = true; break; }
      if (it - prog->it < PCG_LAG) continue;
      const auto t0 = std::chrono::steady_clock::now();
      int seen = prog->it;
      auto last_move = t0;
      while (!prog->done && it - prog->it >= PCG_LAG) {
        const auto now = std::chrono::steady_clock::now();
        if (prog->it != seen) {
          seen = prog->it;
          last_move = now;
        } else if (std::chrono::duration<double, std::milli>(now - last_move).count() > PCG_STALL_MS) {
          // the progress word is not reaching the host: synchronise instead
          if (prog_failed) *prog_failed = true;
          prog = nullptr;
          if ((e = poll()) != hipSuccess) return e;
          if (hs.done) done = have_final = true;
          next = it + 1 + 4;
          break;
        }
      }
      if (done) break;
      if (prog && prog->done) { done = true; break; }
    } else if (it + 1 == next) {
      if ((e = poll()) != hipSuccess) return e;
      if (hs.done) { done = have_final = true; break; }
      next += 4;
    }
  }
  if (!done) {
    if ((e = step(max_iter, 1)) != hipSuccess) return e;
  }
  if (!have_final && prog) {
    bool stalled = false;
    if ((e = wait_done(&stalled)) != hipSuccess) return e;
    if (stalled) {
      if (prog_failed) *prog_failed = true;
      have_final = true;
    } else {
      // acquire: the record's final values were stored before done
      std::atomic_thread_fence(std::memory_order_acquire);
      out->iters = prog->iters;
      out->dn_final = prog->dn;
      out->d0 = prog->d0;
      out->dn0 = prog->dn0;
      out->breakdown = prog->err != 0;
      out->err_word = prog->abort;
      return hipSuccess;
    }
  }
  if (!have_final && (e = poll()) != hipSuccess) return e;
  out->iters = hs.iters;
  out->dn_final = hs.dn[hs.iters & 1];
  out->d0 = hs.d0;
  out->dn0 = hs.dn0;
  out->breakdown = hs.err != 0;
  out->err_word = hw;
  return hipSuccess;
}

// Two-launch arrangement: `op` (direction + operator) and the update kernel.
template <typename MT, int BD, typename OpFn>
inline hipError_t pcg_drive(hipStream_t stream, void* h_pinned, volatile PcgProgress* prog, int nb,
                            const double* rhs, const MT* J, double* x, double* r, double* s, double* const dirs[2],
                            double* q, double* part_rs, double* part_dq, int n_dq, PcgScalars* sc, int max_iter,
                            OpFn op, PcgRun* out, bool* prog_failed, const int* abort = nullptr) {
  auto init = [&]() { return launch_pcg_init<MT, BD>(nb, rhs, J, x, r, s, part_rs, sc, stream); };
  auto step = [&](int it, int last) -> hipError_t {
    hipError_t e = op(it, last, dirs[(it + 1) & 1], dirs[it & 1]);
    if (e != hipSuccess || last) return e;
    return launch_pcg_update<MT, BD>(nb, J, x, r, dirs[it & 1], q, s, part_dq, n_dq, part_rs, sc, it, stream);
  };
  return pcg_loop(stream, h_pinned, prog, sc, max_iter, init, step, out, prog_failed, abort);
}

}  // namespace g2ohip
