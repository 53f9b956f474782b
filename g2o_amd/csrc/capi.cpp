// extern "C" entry points of include/g2o_hip.h. Each wraps one context
// operation; errors become status codes plus the context's last-error text.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>

#include "context.h"

namespace {

// Runs one entry point's body with the context's device current. No C++
// exception leaves the ABI: a host allocation failure (std::bad_alloc from a
// container of the structure builder, the window planner or a host copy)
// becomes G2OHIP_ERR_OOM and anything else G2OHIP_ERR_INVALID, each with a
// last-error message.
template <typename F>
int guarded(g2ohip_ctx* ctx, F&& body) noexcept {
  if (!ctx) return G2OHIP_ERR_INVALID;
  if (ctx->device >= 0) (void)hipSetDevice(ctx->device);
  try {
    return body();
  } catch (const std::bad_alloc&) {
    return ctx->fail_nothrow(G2OHIP_ERR_OOM, "out of host memory");
  } catch (...) {
    return ctx->fail_nothrow(G2OHIP_ERR_INVALID, "unexpected host exception");
  }
}

}  // namespace

extern "C" {

void g2ohip_default_options(g2ohip_options* opt) {
  if (!opt) return;
  opt->precision = G2OHIP_FP64;
  opt->schur_mode = G2OHIP_SCHUR_EXPLICIT;
  opt->pcg_tolerance = 1e-6;
  opt->pcg_max_iter = -1;
  opt->refine_steps = 4;
  opt->max_trials = 10;
  opt->user_lambda_init = 0.0;
  opt->tau = 1e-5;
}

int g2ohip_create(int device, g2ohip_ctx** out) try {
  if (!out) return G2OHIP_ERR_INVALID;
  *out = nullptr;
  g2ohip_ctx* c = new (std::nothrow) g2ohip_ctx();
  if (!c) return G2OHIP_ERR_OOM;
  g2ohip_default_options(&c->opt);
  if (device >= 0) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device >= n) {
      delete c;
      return G2OHIP_ERR_NO_DEVICE;
    }
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&c->h_pinned), 256, hipHostMallocDefault) != hipSuccess) {
      delete c;
      return G2OHIP_ERR_HIP;
    }
    c->device = device;
    // the no-allocation entry point's scratch array: its address on this device
    if (g2ohip::noalloc_init_device(device) != hipSuccess) {
      delete c;
      return G2OHIP_ERR_HIP;
    }
    // host-mapped PCG progress record and LM trial scalars (optional:
    // without them, or with G2OHIP_HOST_MAPPED=0 in the environment, the host
    // copies and synchronises the stream)
    static_assert(sizeof(g2ohip::PcgProgress) <= 128 && sizeof(g2ohip::TrialScalars) <= 128, "mapped layout");
    if (const char* sp = std::getenv("G2OHIP_SCHUR_PATH")) {
      if (std::strcmp(sp, "windows") == 0) c->schur_path = 1;
      else if (std::strcmp(sp, "rows") == 0) c->schur_path = 2;
    }
    if (const char* bs = std::getenv("G2OHIP_BAL_SCHUR")) c->bal_mfma = std::strcmp(bs, "valu") != 0;
    if (const char* pf = std::getenv("G2OHIP_PCG_FUSED")) c->pcg_fused = pf[0] != '0';
    // the concurrent Schur diagonal pass (G2OHIP_SCHUR_FORK=0: one stream)
    const char* fork_env = std::getenv("G2OHIP_SCHUR_FORK");
    if (!(fork_env && fork_env[0] == '0')) {
      if (hipStreamCreate(&c->side) != hipSuccess ||
          hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
          hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) {
        delete c;
        return G2OHIP_ERR_HIP;
      }
    }
    const char* mapped_env = std::getenv("G2OHIP_HOST_MAPPED");
    const bool mapped = !(mapped_env && mapped_env[0] == '0');
    void* hp = nullptr;
    void* dp = nullptr;
    if (mapped && hipHostMalloc(&hp, 256, hipHostMallocMapped | hipHostMallocCoherent) == hipSuccess) {
      if (hipHostGetDevicePointer(&dp, hp, 0) == hipSuccess && dp) {
        c->h_prog = static_cast<g2ohip::PcgProgress*>(hp);
        c->d_prog = static_cast<g2ohip::PcgProgress*>(dp);
        c->h_trial = reinterpret_cast<g2ohip::TrialScalars*>(static_cast<char*>(hp) + 128);
        c->d_trial = reinterpret_cast<g2ohip::TrialScalars*>(static_cast<char*>(dp) + 128);
      } else {
        (void)hipHostFree(hp);
      }
    }
    (void)hipGetLastError();
  }
  *out = c;
  return G2OHIP_OK;
} catch (...) {
  return G2OHIP_ERR_OOM;
}

void g2ohip_destroy(g2ohip_ctx* ctx) {
  if (!ctx) return;
  if (ctx->device >= 0) (void)hipSetDevice(ctx->device);
  delete ctx;
}

const char* g2ohip_last_error(const g2ohip_ctx* ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

int g2ohip_set_options(g2ohip_ctx* ctx, const g2ohip_options* opt) {
  return guarded(ctx, [&]() -> int {
    if (!opt) return ctx->fail(G2OHIP_ERR_INVALID, "null options");
    if (opt->precision != G2OHIP_FP64 && opt->precision != G2OHIP_FP32)
      return ctx->fail(G2OHIP_ERR_INVALID, "unknown precision");
    if (opt->schur_mode != G2OHIP_SCHUR_EXPLICIT && opt->schur_mode != G2OHIP_SCHUR_IMPLICIT)
      return ctx->fail(G2OHIP_ERR_INVALID, "unknown Schur mode");
    if (!(opt->pcg_tolerance >= 0) || opt->max_trials < 1 || opt->refine_steps < 0 || opt->refine_steps > 16)
      return ctx->fail(G2OHIP_ERR_INVALID, "invalid option value");
    ctx->opt = *opt;
    ctx->pcg_residual_mem = -1.0;  // LinearSolverPCG::init
    ctx->fp32_off = false;
    return G2OHIP_OK;
  });
}

int g2ohip_get_options(const g2ohip_ctx* ctx, g2ohip_options* opt) {
  if (!ctx || !opt) return G2OHIP_ERR_INVALID;
  *opt = ctx->opt;
  return G2OHIP_OK;
}

int g2ohip_ba_set_structure(g2ohip_ctx* ctx, int num_cams, int num_pts, int num_edges, const int* edge_cam,
                            const int* edge_pt, const uint8_t* cam_fixed, const uint8_t* pt_fixed) {
  return guarded(ctx, [&]() -> int {
    return ctx->set_structure(num_cams, num_pts, num_edges, edge_cam, edge_pt, cam_fixed, pt_fixed);
  });
}

int g2ohip_ba_structure_info(g2ohip_ctx* ctx, long long* info) {
  return guarded(ctx, [&]() -> int {
    if (!ctx->have_structure || !info) return ctx->fail(G2OHIP_ERR_INVALID, "no structure");
    if (ctx->mode != 0) return ctx->fail(G2OHIP_ERR_INVALID, "6-DoF bundle-adjustment call on a pose-graph or BAL context");
    const bool rows = ctx->use_schur_rows();
    if (!rows) {
      const int rc = ctx->ensure_win(0);
      if (rc) return rc;
    }
    const auto& st = ctx->st;
    int maxn = 0;
    for (int i = 0; i < st.npos; ++i) maxn = std::max(maxn, st.pos_ptr[i + 1] - st.pos_ptr[i]);
    info[0] = st.Cf;
    info[1] = st.Pf;
    info[2] = st.ns;
    info[3] = st.nS;
    info[4] = st.nlc;
    const auto& wp = st.win[0];
    info[5] = rows ? 0 : static_cast<long long>(wp.win_obs0.size()) - 1;
    info[6] = g2ohip::schur_pair_count(st);
    info[7] = rows ? 0 : wp.od_ptr.back();
    info[8] = maxn;
    info[9] = st.nF;
    info[10] = static_cast<long long>(st.lc_cams.size());
    info[11] = st.npos;
    info[12] = st.nq;
    return G2OHIP_OK;
  });
}

int g2ohip_ba_set_measurements(g2ohip_ctx* ctx, const double* meas, const double* weight, const double* intr,
                               double huber_delta) {
  return guarded(ctx, [&]() -> int {
    return ctx->set_measurements(meas, weight, intr, huber_delta);
  });
}

int g2ohip_ba_set_information(g2ohip_ctx* ctx, const double* omega) {
  return guarded(ctx, [&]() -> int { return ctx->set_information(omega); });
}

int g2ohip_ba_set_state(g2ohip_ctx* ctx, const double* cams, const double* pts) {
  return guarded(ctx, [&]() -> int {
    if (!cams || (!pts && ctx->st.P > 0)) return ctx->fail(G2OHIP_ERR_INVALID, "null state");
    return ctx->set_state(cams, pts);
  });
}

int g2ohip_ba_get_state(g2ohip_ctx* ctx, double* cams, double* pts) {
  return guarded(ctx, [&]() -> int {
    return ctx->get_state(cams, pts);
  });
}

int g2ohip_compute_active_errors(g2ohip_ctx* ctx, double* chi2) {
  return guarded(ctx, [&]() -> int {
    if (!chi2) return ctx->fail(G2OHIP_ERR_INVALID, "null output");
    return ctx->compute_active_errors(chi2);
  });
}

int g2ohip_get_errors(g2ohip_ctx* ctx, double* err) {
  return guarded(ctx, [&]() -> int {
    if (!err) return ctx->fail(G2OHIP_ERR_INVALID, "null output");
    double chi;
    return ctx->compute_active_errors(&chi, err);
  });
}

int g2ohip_build_system(g2ohip_ctx* ctx) {
  return guarded(ctx, [&]() -> int {
    return ctx->build_system();
  });
}

int g2ohip_set_lambda(g2ohip_ctx* ctx, double lambda) {
  return guarded(ctx, [&]() -> int {
    if (!(lambda >= 0) && !(lambda < 0)) return ctx->fail(G2OHIP_ERR_INVALID, "lambda is NaN");
    ctx->lambda = lambda;
    return G2OHIP_OK;
  });
}

int g2ohip_restore_diagonal(g2ohip_ctx* ctx) {
  return guarded(ctx, [&]() -> int {
    // the damping lives only inside the solve kernels; nothing to restore
    return G2OHIP_OK;
  });
}

int g2ohip_solve(g2ohip_ctx* ctx, int* ok) {
  return guarded(ctx, [&]() -> int {
    if (!ok) return ctx->fail(G2OHIP_ERR_INVALID, "null output");
    return ctx->solve(ok);
  });
}

int g2ohip_get_x(g2ohip_ctx* ctx, double* x) {
  return guarded(ctx, [&]() -> int {
    if (!x || !ctx->have_solution) return ctx->fail(G2OHIP_ERR_INVALID, "no solution");
    size_t n = ctx->system_size();
    if (hipStreamSynchronize(ctx->stream) != hipSuccess ||
        hipMemcpy(x, ctx->d_x.p, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      return ctx->fail(G2OHIP_ERR_HIP, "copy of x failed");
    ctx->points_to_pub(x);
    return G2OHIP_OK;
  });
}

int g2ohip_get_b(g2ohip_ctx* ctx, double* b) {
  return guarded(ctx, [&]() -> int {
    if (!b || !ctx->have_system) return ctx->fail(G2OHIP_ERR_INVALID, "no system");
    size_t n = ctx->system_size();
    if (hipStreamSynchronize(ctx->stream) != hipSuccess ||
        hipMemcpy(b, ctx->d_b.p, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      return ctx->fail(G2OHIP_ERR_HIP, "copy of b failed");
    ctx->points_to_pub(b);
    return G2OHIP_OK;
  });
}

int g2ohip_multiply_hessian(g2ohip_ctx* ctx, const double* src, double* dest) {
  return guarded(ctx, [&]() -> int {
    if (!src || !dest) return ctx->fail(G2OHIP_ERR_INVALID, "null vector");
    return ctx->multiply_hessian(src, dest);
  });
}

int g2ohip_update(g2ohip_ctx* ctx) {
  return guarded(ctx, [&]() -> int {
    return ctx->update();
  });
}

int g2ohip_push(g2ohip_ctx* ctx) {
  return guarded(ctx, [&]() -> int {
    return ctx->push();
  });
}

int g2ohip_pop(g2ohip_ctx* ctx) {
  return guarded(ctx, [&]() -> int {
    return ctx->pop();
  });
}

int g2ohip_discard_top(g2ohip_ctx* ctx) {
  return guarded(ctx, [&]() -> int {
    return ctx->discard_top();
  });
}

int g2ohip_lambda_init(g2ohip_ctx* ctx, double* lambda) {
  return guarded(ctx, [&]() -> int {
    if (!lambda) return ctx->fail(G2OHIP_ERR_INVALID, "null output");
    return ctx->lambda_init(lambda);
  });
}

int g2ohip_compute_scale(g2ohip_ctx* ctx, double* scale) {
  return guarded(ctx, [&]() -> int {
    if (!scale) return ctx->fail(G2OHIP_ERR_INVALID, "null output");
    return ctx->compute_scale(scale);
  });
}

int g2ohip_optimize(g2ohip_ctx* ctx, int iterations, g2ohip_iter_stats* stats, int* done) {
  return guarded(ctx, [&]() -> int {
    int d = 0;
    int rc = ctx->optimize(iterations, stats, &d);
    if (done) *done = d;
    return rc;
  });
}

int g2ohip_compute_marginals(g2ohip_ctx* ctx, int nblocks, const int* rows, const int* cols, double* blocks,
                             double tol) {
  return guarded(ctx, [&]() -> int {
    return ctx->marginals(nblocks, rows, cols, blocks, tol);
  });
}

int g2ohip_set_algorithm(g2ohip_ctx* ctx, int algorithm) {
  return guarded(ctx, [&]() -> int {
    if (algorithm != G2OHIP_LEVENBERG && algorithm != G2OHIP_GAUSS_NEWTON && algorithm != G2OHIP_DOGLEG)
      return ctx->fail(G2OHIP_ERR_INVALID, "unknown algorithm");
    ctx->algorithm = algorithm;
    return G2OHIP_OK;
  });
}

int g2ohip_set_force_stop_flag(g2ohip_ctx* ctx, const volatile int* flag) {
  return guarded(ctx, [&]() -> int {
    ctx->force_stop = flag;
    return G2OHIP_OK;
  });
}

int g2ohip_ba_get_hessian(g2ohip_ctx* ctx, double* Hpp, double* Hll, double* Hpl_edge) {
  return guarded(ctx, [&]() -> int {
    return ctx->get_hessian(Hpp, Hll, Hpl_edge);
  });
}

int g2ohip_ba_get_schur(g2ohip_ctx* ctx, int* rows, int* cols, double* blocks, double* bschur) {
  return guarded(ctx, [&]() -> int {
    return ctx->get_schur(rows, cols, blocks, bschur);
  });
}

int g2ohip_last_solve_info(g2ohip_ctx* ctx, double* info) {
  return guarded(ctx, [&]() -> int {
    if (!info) return ctx->fail(G2OHIP_ERR_INVALID, "null output");
    for (int i = 0; i < 4; ++i) info[i] = ctx->last_info[i];
    return G2OHIP_OK;
  });
}

int g2ohip_lm_trials(g2ohip_ctx* ctx, double* out, int max, int* count) {
  return guarded(ctx, [&]() -> int {
    if (!count) return ctx->fail(G2OHIP_ERR_INVALID, "null output");
    const int n = static_cast<int>(ctx->trials.size() / G2OHIP_TRIAL_FIELDS);
    *count = n;
    if (out && max > 0) {
      const size_t m = static_cast<size_t>(std::min(n, max)) * G2OHIP_TRIAL_FIELDS;
      std::copy(ctx->trials.begin(), ctx->trials.begin() + m, out);
    }
    return G2OHIP_OK;
  });
}

int g2ohip_set_profiling(g2ohip_ctx* ctx, int on) {
  return guarded(ctx, [&]() -> int {
    return ctx->set_profiling(on);
  });
}

int g2ohip_get_profile(g2ohip_ctx* ctx, int kernel, double* total_ms, int* launches) {
  return guarded(ctx, [&]() -> int {
    return ctx->get_profile(kernel, total_ms, launches);
  });
}

int g2ohip_replay_pcg(g2ohip_ctx* ctx, int reps, double* us_spmv, double* us_iteration) {
  return guarded(ctx, [&]() -> int { return ctx->replay_pcg(reps, us_spmv, us_iteration); });
}

int g2ohip_linear_solve_pcg(g2ohip_ctx* ctx, int nb, int bd, int nblk, const int* brow, const int* bcol,
                            const double* blocks, const double* b, double* x, double tol, int max_iter,
                            double* residual_io, int* iterations) {
  return guarded(ctx, [&]() -> int {
    return ctx->linear_solve_pcg(nb, bd, nblk, brow, bcol, blocks, b, x, tol, max_iter, residual_io, iterations);
  });
}

int g2ohip_linear_solve_pcg_device(g2ohip_ctx* ctx, int nb, int bd, const int* F_ptr, const int* F_col,
                                   const double* F_val, const int* F_diag, const double* b, double* x, double tol,
                                   int max_iter, double* residual_io, int* iterations) {
  return guarded(ctx, [&]() -> int {
    return ctx->linear_solve_pcg_device(nb, bd, F_ptr, F_col, F_val, F_diag, b, x, tol, max_iter, residual_io,
                                        iterations);
  });
}

int g2ohip_ba_adjoint(g2ohip_ctx* ctx, const double* dL_dcam, const double* dL_dpt, double* dL_dmeas,
                      double* dL_dweight, double tol, int max_iter, int* iterations) {
  return guarded(ctx, [&]() -> int {
    return ctx->adjoint(dL_dcam, dL_dpt, dL_dmeas, dL_dweight, tol < 0 ? 1e-24 : tol, max_iter, iterations);
  });
}

int g2ohip_ba_newton(g2ohip_ctx* ctx, int steps, double tol, double max_step, int* ok) {
  return guarded(ctx, [&]() -> int {
    if (!ok) return ctx->fail(G2OHIP_ERR_INVALID, "ok must not be NULL");
    return ctx->newton(steps, tol < 0 ? 1e-24 : tol, max_step, ok);
  });
}

// ------------------------------------------------------------ pose graph
int g2ohip_pg_set_structure(g2ohip_ctx* ctx, int num_poses, int num_edges, const int* edge_from, const int* edge_to,
                            const uint8_t* fixed) {
  return guarded(ctx, [&]() -> int {
    return ctx->pg_set_structure(num_poses, num_edges, edge_from, edge_to, fixed);
  });
}

int g2ohip_pg_structure_info(g2ohip_ctx* ctx, long long* info) {
  return guarded(ctx, [&]() -> int {
    if (!info) return ctx->fail(G2OHIP_ERR_INVALID, "null output");
    if (!ctx->have_structure || ctx->mode != 1) return ctx->fail(G2OHIP_ERR_INVALID, "no pose-graph structure");
    info[0] = ctx->pgst.Nf;
    info[1] = static_cast<long long>(ctx->pgst.act.size());
    info[2] = ctx->pgst.nU;
    info[3] = ctx->pgst.nF;
    return G2OHIP_OK;
  });
}

int g2ohip_pg_set_measurements(g2ohip_ctx* ctx, const double* meas, const double* information) {
  return guarded(ctx, [&]() -> int {
    return ctx->pg_set_measurements(meas, information);
  });
}

int g2ohip_pg_set_state(g2ohip_ctx* ctx, const double* poses) {
  return guarded(ctx, [&]() -> int {
    if (!poses) return ctx->fail(G2OHIP_ERR_INVALID, "null poses");
    return ctx->pg_set_state(poses);
  });
}

int g2ohip_pg_get_state(g2ohip_ctx* ctx, double* poses) {
  return guarded(ctx, [&]() -> int {
    if (!poses) return ctx->fail(G2OHIP_ERR_INVALID, "null output");
    return ctx->pg_get_state(poses);
  });
}

int g2ohip_pg_get_system(g2ohip_ctx* ctx, int* rows, int* cols, double* blocks, double* b) {
  return guarded(ctx, [&]() -> int {
    return ctx->pg_get_system(rows, cols, blocks, b);
  });
}

// -------------------------------------------------------------------- BAL
int g2ohip_bal_set_structure(g2ohip_ctx* ctx, int num_cams, int num_pts, int num_edges, const int* edge_cam,
                             const int* edge_pt, const uint8_t* cam_fixed, const uint8_t* pt_fixed) {
  return guarded(ctx, [&]() -> int {
    return ctx->bal_set_structure(num_cams, num_pts, num_edges, edge_cam, edge_pt, cam_fixed, pt_fixed);
  });
}

int g2ohip_bal_structure_info(g2ohip_ctx* ctx, long long* info) {
  return guarded(ctx, [&]() -> int {
    if (!info) return ctx->fail(G2OHIP_ERR_INVALID, "null output");
    if (!ctx->have_structure || ctx->mode != 2) return ctx->fail(G2OHIP_ERR_INVALID, "no BAL structure");
    const g2ohip::BAStructure& st = ctx->st;
    info[0] = st.Cf;
    info[1] = st.Pf;
    info[2] = st.ns;
    info[3] = st.nS;
    info[4] = st.nF;
    info[5] = ctx->bal_pairs;
    info[6] = st.nq;
    return G2OHIP_OK;
  });
}

int g2ohip_bal_set_measurements(g2ohip_ctx* ctx, const double* meas, const double* weight, double huber_delta) {
  return guarded(ctx, [&]() -> int {
    return ctx->bal_set_measurements(meas, weight, huber_delta);
  });
}

int g2ohip_bal_set_state(g2ohip_ctx* ctx, const double* cams, const double* pts) {
  return guarded(ctx, [&]() -> int {
    return ctx->bal_set_state(cams, pts);
  });
}

int g2ohip_bal_get_state(g2ohip_ctx* ctx, double* cams, double* pts) {
  return guarded(ctx, [&]() -> int {
    return ctx->bal_get_state(cams, pts);
  });
}

int g2ohip_bal_get_hessian(g2ohip_ctx* ctx, double* Hpp, double* Hll, double* Hpl_edge) {
  return guarded(ctx, [&]() -> int {
    return ctx->bal_get_hessian(Hpp, Hll, Hpl_edge);
  });
}

int g2ohip_bal_get_schur(g2ohip_ctx* ctx, int* rows, int* cols, double* blocks, double* bschur) {
  return guarded(ctx, [&]() -> int {
    return ctx->bal_get_schur(rows, cols, blocks, bschur);
  });
}

}  // extern "C"
