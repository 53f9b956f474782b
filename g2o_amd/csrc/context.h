// Device context behind the C ABI: one bundle-adjustment graph on one GPU.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/g2o_hip.h"
#include "kernels.h"
#include "structure.h"

namespace g2ohip {

template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t count) {
    if (count <= n && p) return hipSuccess;
    release();
    size_t bytes = (count ? count : 1) * sizeof(T);
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), bytes);
    if (e == hipSuccess) n = count;
    else p = nullptr;
    return e;
  }
  hipError_t upload(const std::vector<T>& v, hipStream_t st) {
    hipError_t e = alloc(v.size());
    if (e != hipSuccess) return e;
    if (v.empty()) return hipSuccess;
    return hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st);
  }
};

// Buffers of one Hessian precision.
template <typename HT>
struct PrecBufs {
  DBuf<HT> Hpl;       // 18 * nq, one row-major 6x3 block per observation, camera-major
  DBuf<HT> F_val;     // nF * 36
  DBuf<HT> Jinv;      // Cf * 36
  DBuf<HT> Dinv;      // Pf * 9, (Hll + lambda I)^-1 (implicit operator, exact back-substitution)
  DBuf<HT> part;      // 36 per window-plan slot: off-diagonal Schur partials
  DBuf<double> x, r, d, d2, q, s;  // PCG vectors, FP64 (d, d2: directions of odd / even iterations)
  DBuf<double> s2, w, w2;          // one-launch iteration: the second s buffer, w = M^-1 q (two buffers)
};

// Outcome of one PCG solve (read back from PcgScalars).
struct PcgRun {
  int iters = 0;
  double dn_final = 0;  // r . M^-1 r at exit
  double d0 = 0;        // the absolute stop threshold used
  double dn0 = 0;       // r . M^-1 r of the right-hand side
  bool breakdown = false;  // d.q <= 0 or r.M^-1 r < 0 seen: operator or preconditioner not PD
  int err_word = 0;        // the PcgControl::abort word at the end (0 without one)
};

struct WinPlanBufs {
  DBuf<int> win_obs0, win_q, win_grp0, grp_len, grp_row0, grp_slot, od_blk, od_ptr;
  DBuf<unsigned> pair_rows;
};

// noalloc.hip: the per-device scratch array of g2ohip_linear_solve_pcg_device
hipError_t noalloc_init_device(int device);
size_t noalloc_bytes(int nb, int bd);

}  // namespace g2ohip

struct g2ohip_ctx {
  int device = -1;
  hipStream_t stream = nullptr;
  // second stream (default priority) for the Schur diagonal pass, which runs
  // beside the off-diagonal blocks; fork / join events on the two streams
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // explicit-operator PCG in one launch per iteration (G2OHIP_PCG_FUSED=0:
  // direction + product and update as two launches)
  bool pcg_fused = true;
  std::string last_error;
  g2ohip_options opt{};
  const volatile int* force_stop = nullptr;

  g2ohip::BAStructure st;
  int mode = 0;  // 0: bundle adjustment, 1: 3D pose graph, 2: BAL (the last *_set_structure)
  bool have_structure = false, have_meas = false, have_state = false;
  bool have_system = false;   // build_system ran since the last state change
  bool have_lin = false;      // the blocks hold a Gauss-Newton system (kept across state updates,
                              // as the reference's solver keeps its last Hpp after optimize)
  int system_prec = 0;        // precision of the last build_system
  int last_solve_prec = 0;    // precision of the last BA solve_impl
  bool have_solution = false;
  bool have_exact = false;    // the FP64 blocks hold the exact Hessian (adjoint / Newton)
  double lambda = 0.0;
  int algorithm = 0;  // g2ohip_algorithm
  double pcg_residual_mem = -1.0;  // LinearSolverPCG::_residual
  double last_info[4] = {0, 0, -1, 0};
  // one record per LM trial of the last optimize (g2ohip_lm_trials)
  std::vector<double> trials;
  // FP32 blocks with FP64 refinement: once the FP32 solution's FP64 residual
  // misses the first solve's threshold by more than 10x, or a correction
  // stops above it (the FP32-rounded reduced system no longer resolves its
  // weakest modes), the solve is redone in FP64 and FP64 is kept until the
  // next optimize() / set_options()
  bool fp32_off = false;
  bool refine_stalled = false;
  double chi2_lin = 0.0;

  g2ohip::DevGraph g{};
  g2ohip::DBuf<int> d_camH, d_ptH, d_pos_pt, d_pos_h, d_pos_ptr, d_pos_nfree, d_slot_cam, d_slot_pos, d_slot_camh, d_slot_ph, d_slot_q, d_cam_qptr, d_q_ph, d_q_slot, d_q_pt, d_cam_raw, d_lc_pos0,
      d_lc_cam_ptr, d_lc_cams, d_camp_ptr, d_camp_idx, d_S_row, d_S_col, d_S_diag, d_up_pos, d_up_posT, d_F_ptr,
      d_F_col, d_F_diag, d_iota;
  g2ohip::DBuf<int4> d_slot_rec;
  g2ohip::DBuf<double> d_meas, d_w, d_intr;
  g2ohip::DBuf<double> d_S;  // per-slot information factors (empty: identity)
  g2ohip::WinPlanBufs win_bufs[2];
  g2ohip::DevWinPlan dwins[2]{};
  bool win_ready[2] = {false, false};
  // off-diagonal Schur blocks: 0 = choose per structure, 1 = window plan
  // (partials staged in HBM), 2 = camera rows (G2OHIP_SCHUR_PATH in the
  // environment at context creation: "windows" or "rows")
  int schur_path = 0;
  int schur_rows_chosen = -1;  // the choice made for the current structure
  g2ohip::DBuf<int4> d_row_items;
  int row_nitems = 0, row_max_blk = 0;

  g2ohip::DBuf<double> d_cams, d_pts, d_camRt;
  std::vector<g2ohip::DBuf<double>*> stack_c, stack_p;
  int stack_depth = 0;

  g2ohip::DBuf<double> d_linv;  // L^-1 of the point blocks (Schur windows)
  g2ohip::DBuf<double> d_Sdiag;  // FP64 diagonal blocks of the FP32 reduced system [Cf][36]
  g2ohip::DBuf<double> d_Hll, d_b, d_Hpp, d_hpp_part, d_diag_part, d_chunk_chi, d_coef_part, d_bschur, d_x, d_red, d_scal,
      d_cam_part, d_res, d_xp, d_err64;
  g2ohip::DBuf<double> d_dq_part, d_dn_part;
  g2ohip::DBuf<double> d_fpart;  // one-launch PCG partials: [2][4][pcg_fused_blocks(Cf)]
  g2ohip::DBuf<int> d_F_head;    // row heads of the reduced system (row_heads)
  g2ohip::DBuf<int> d_err;
  // host-mapped PCG progress record (PcgControl::prog); prog_ok drops to
  // false if it ever fails to reach the host, and the PCG then polls the
  // stream. h_trial / d_trial: the LM trial scalars (launch_publish_scalars),
  // tagged with trial_seq.
  g2ohip::PcgProgress* h_prog = nullptr;
  g2ohip::PcgProgress* d_prog = nullptr;
  bool prog_ok = true;
  g2ohip::TrialScalars* h_trial = nullptr;
  g2ohip::TrialScalars* d_trial = nullptr;
  int trial_seq = 0;
  volatile g2ohip::PcgProgress* progress() const { return prog_ok && h_prog ? h_prog : nullptr; }
  g2ohip::PcgControl with_progress(g2ohip::PcgControl c) const {
    c.prog = prog_ok && h_prog ? d_prog : nullptr;
    return c;
  }
  g2ohip::DBuf<g2ohip::PcgScalars> d_sc;
  g2ohip::PrecBufs<float> f32;
  g2ohip::PrecBufs<double> f64;
  double* h_pinned = nullptr;  // small pinned host buffer for scalars

  // 3D pose graph (posegraph.cpp)
  g2ohip::PGStructure pgst;
  g2ohip::DevPG pg{};
  g2ohip::DBuf<int> d_pg_H, d_pg_act, d_pg_from, d_pg_to, d_pg_swap, d_pg_end_ptr, d_pg_end_ref, d_pg_blk_ptr,
      d_pg_blk_ref, d_pg_up_pos, d_pg_up_posT, d_pg_F_diag, d_pg_F_ptr, d_pg_F_col,
      d_pg_F_head;
  g2ohip::DBuf<double> d_pg_meas, d_pg_info, d_pg_rec, d_pg_F_val, d_pg_Hdiag, d_pg_part;

  // BAL (9-DoF cameras, bal.cpp): the structure is `st` with the lists below
  g2ohip::DevBAL bal{};
  g2ohip::DBuf<int> d_bal_cs_ptr, d_bal_cs_slot, d_bal_od_ptr, d_bal_od_blk;
  g2ohip::DBuf<int2> d_bal_od_pair;
  g2ohip::DBuf<int4> d_bal_rec;
  g2ohip::DBuf<double> d_bal_tab, d_bal_Hpp, d_bal_Hpl, d_bal_Dinv, d_bal_F_val, d_bal_part, d_bal_G;
  long long bal_pairs = 0;
  // off-diagonal BAL Schur blocks on the matrix cores (default) or the vector
  // units (G2OHIP_BAL_SCHUR=valu in the environment at context creation)
  bool bal_mfma = true;

  // kernel timing (HIP events on `stream`)
  bool profiling = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<std::pair<int, int>> ev_marks[G2OHIP_K_COUNT];
  int prof_begin();
  void prof_end(int kernel, int begin);
  int set_profiling(int on);
  int get_profile(int kernel, double* total_ms, int* launches);

  ~g2ohip_ctx();
  int fail(int code, const std::string& msg);
  int fail_nothrow(int code, const char* msg) noexcept;
  int require_device();
  int hip(hipError_t e, const char* what);

  // operations (return a g2ohip_status)
  int set_structure(int C, int P, int E, const int* ec, const int* ep, const uint8_t* cf, const uint8_t* pf);
  int set_measurements(const double* meas, const double* w, const double* intr, double huber);
  int set_state(const double* cams, const double* pts);
  int set_information(const double* omega);
  int get_state(double* cams, double* pts);
  int ensure_win(int which);
  bool use_schur_rows();
  int ensure_rows();
  // off-diagonal blocks of the reduced system into B.F_val (the Hll^-1
  // factors d_linv must be current)
  template <typename HT>
  int schur_offdiag(g2ohip::PrecBufs<HT>& B);
  // the FP64 diagonal pass of the reduced system on stream st
  template <typename HT>
  int schur_diag(g2ohip::PrecBufs<HT>& B, bool implicit, hipStream_t st);
  int compute_active_errors(double* chi2, double* err_out = nullptr);
  int build_system(bool read_chi = true);
  int solve(int* ok, bool defer_err = false);
  int update();
  int push();
  int pop();
  int discard_top();
  int lambda_init(double* out);
  int compute_scale(double* out);
  int optimize(int iterations, g2ohip_iter_stats* stats, int* done);
  int optimize_gauss_newton(int iterations, g2ohip_iter_stats* stats, int* done);
  int optimize_dogleg(int iterations, g2ohip_iter_stats* stats, int* done);
  int get_hessian(double* Hpp, double* Hll, double* Hpl_edge);
  int replay_pcg(int reps, double* us_spmv, double* us_iter);
  int multiply_hessian(const double* src, double* dest);
  int get_schur(int* rows, int* cols, double* blocks, double* bschur);
  int linear_solve_pcg_device(int nb, int bd, const int* F_ptr, const int* F_col, const double* F_val,
                              const int* F_diag, const double* b, double* x, double tol, int max_iter,
                              double* residual_io, int* iterations);
  int linear_solve_pcg(int nb, int bd, int nblk, const int* brow, const int* bcol, const double* blocks,
                       const double* b, double* x, double tol, int max_iter, double* residual_io, int* iterations);
  int adjoint(const double* dL_dcam, const double* dL_dpt, double* dL_dmeas, double* dL_dw, double tol,
              int max_iter, int* iterations);

  // state saved by push / pop, in doubles: 7 per camera or pose (9 per BAL
  // camera), 3 per point
  size_t cams_len() const {
    return mode == 1 ? 7 * static_cast<size_t>(pgst.N) : (mode == 2 ? 9 : 7) * static_cast<size_t>(st.C);
  }
  size_t pts_len() const { return mode == 1 ? 0 : 3 * static_cast<size_t>(st.P); }
  // camera block size of the reduced system
  int cam_dim() const { return mode == 2 ? 9 : 6; }
  // the point part of a system-sized vector (x, b), from the device's point
  // order to hessianIndex order
  void points_to_pub(double* v) const;
  size_t system_size() const {
    return mode == 1 ? 6 * static_cast<size_t>(pgst.Nf)
                     : cam_dim() * static_cast<size_t>(st.Cf) + 3 * static_cast<size_t>(st.Pf);
  }
  int bal_set_structure(int C, int P, int E, const int* ec, const int* ep, const uint8_t* cf, const uint8_t* pf);
  int bal_set_measurements(const double* meas, const double* w, double huber);
  int bal_set_state(const double* cams9, const double* pts);
  int bal_get_state(double* cams9, double* pts);
  int bal_chi2(double* chi2, double* err_out);
  int bal_build();
  int bal_lambda_init(double* out);
  int bal_solve(int* ok);
  int bal_update();
  int bal_get_hessian(double* Hpp, double* Hll, double* Hpl_edge);
  int bal_get_schur(int* rows, int* cols, double* blocks, double* bschur);
  int pg_set_structure(int N, int E, const int* from, const int* to, const uint8_t* fixed);
  int pg_set_measurements(const double* meas7, const double* info36);
  int pg_set_state(const double* poses7);
  int pg_get_state(double* poses7);
  int pg_chi2(double* chi2, double* err_out);
  int pg_build();
  int pg_solve(int* ok);
  hipError_t pg_pcg_fused(const double* rhs, const g2ohip::PcgControl& ctl, int max_iter, g2ohip::PcgRun* out,
                          bool* prog_failed);
  int pg_marginals(int nblk, const int* rows, const int* cols, double* out, double tol);
  int marginals(int nblk, const int* rows, const int* cols, double* out, double tol);
  int pg_update();
  int pg_lambda_init(double* out);
  int pg_get_system(int* rows, int* cols, double* blocks, double* b);

  int exact_system();
  int solve_exact(double lam, double tol, int max_iter, int* iters, int* ok, double* max_abs);
  int newton(int steps, double tol, double max_step, int* ok);

  template <typename HT>
  int solve_impl(int* ok, bool defer_err);
  template <typename HT>
  int run_pcg(g2ohip::PrecBufs<HT>& B, const double* rhs, const g2ohip::PcgControl& ctl0, bool implicit, int max_iter,
              g2ohip::PcgRun* out);
  int read_scalar(const double* dptr, double* out);
  int read_trial_scalars(double* sc4, int* err);
  int read_err(int* out);
};

// status-propagating helpers of the context implementation files
#define CK(...)                                  \
  do {                                           \
    int rc_ = (__VA_ARGS__);                     \
    if (rc_ != G2OHIP_OK) return rc_;            \
  } while (0)
#define HK(...) CK(hip((__VA_ARGS__), #__VA_ARGS__))

// HIP-event timing of one kernel group (no-op unless profiling is on)
struct ProfScope {
  g2ohip_ctx* c;
  int k, b;
  ProfScope(g2ohip_ctx* c_, int k_) : c(c_), k(k_), b(c_->prof_begin()) {}
  ~ProfScope() { c->prof_end(k, b); }
};
