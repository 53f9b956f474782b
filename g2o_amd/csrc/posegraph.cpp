// 3D pose graph on the device: VertexSE3 / EdgeSE3 with 6x6 blocks and no
// Schur complement (BlockSolver::solve's non-Schur branch,
// block_solver.hpp:333-343, with LinearSolverPCG). The LM loop of
// context.cpp drives these through the generic Solver calls.
#include <algorithm>
#include <cmath>
#include <vector>

#include "context.h"
#include "pcg_drive.h"

using namespace g2ohip;

int g2ohip_ctx::pg_set_structure(int N, int E, const int* from, const int* to, const uint8_t* fixed) {
  if (!from || !to) return fail(G2OHIP_ERR_INVALID, "edge index arrays are required");
  std::string msg = build_pg_structure(pgst, N, E, from, to, fixed);
  if (!msg.empty()) {
    have_structure = false;
    return fail(G2OHIP_ERR_INVALID, msg);
  }
  mode = 1;
  have_structure = true;
  have_meas = have_state = have_system = have_solution = have_exact = have_lin = false;
  pcg_residual_mem = -1.0;
  stack_depth = 0;
  if (device < 0) return G2OHIP_OK;
  std::vector<int> fr(from, from + E), tt(to, to + E);
  HK(d_pg_H.upload(pgst.H, stream));
  HK(d_pg_act.upload(pgst.act, stream));
  HK(d_pg_from.upload(fr, stream));
  HK(d_pg_to.upload(tt, stream));
  HK(d_pg_swap.upload(pgst.edge_swap, stream));
  HK(d_pg_end_ptr.upload(pgst.end_ptr, stream));
  HK(d_pg_end_ref.upload(pgst.end_ref, stream));
  HK(d_pg_blk_ptr.upload(pgst.blk_ptr, stream));
  HK(d_pg_blk_ref.upload(pgst.blk_ref, stream));
  HK(d_pg_up_pos.upload(pgst.up_pos, stream));
  HK(d_pg_up_posT.upload(pgst.up_posT, stream));
  HK(d_pg_F_diag.upload(pgst.F_diag, stream));
  HK(d_pg_F_ptr.upload(pgst.F_ptr, stream));
  HK(d_pg_F_col.upload(pgst.F_col, stream));
  pgst.F_head = row_heads(pgst.F_ptr, pgst.F_col);
  HK(d_pg_F_head.upload(pgst.F_head, stream));
  const int A = static_cast<int>(pgst.act.size());
  const size_t n = 6 * static_cast<size_t>(pgst.Nf);
  HK(d_cams.alloc(7 * static_cast<size_t>(N)));  // the poses
  HK(d_pg_rec.alloc(PG_REC * static_cast<size_t>(std::max(A, 1))));
  HK(d_pg_F_val.alloc(36 * static_cast<size_t>(pgst.nF)));
  HK(d_pg_Hdiag.alloc(36 * static_cast<size_t>(pgst.Nf)));
  HK(d_pg_part.alloc(std::max(pg_edge_blocks(A), 1)));
  HK(d_b.alloc(n));
  HK(d_x.alloc(n));
  HK(d_red.alloc(RED_BLOCKS));
  HK(d_scal.alloc(8));
  HK(d_dq_part.alloc(std::max(pcg_spmv_blocks(pgst.Nf), 1)));
  HK(d_dn_part.alloc(std::max(pcg_row_blocks(pgst.Nf), 1)));
  HK(d_fpart.alloc(8 * static_cast<size_t>(pcg_fused_blocks(std::max(pgst.Nf, 1)))));
  HK(d_err.alloc(4));
  HK(d_sc.alloc(1));
  pg = DevPG{};
  pg.N = N; pg.Nf = pgst.Nf; pg.A = A; pg.nU = pgst.nU;
  pg.H = d_pg_H.p; pg.act = d_pg_act.p; pg.from = d_pg_from.p; pg.to = d_pg_to.p; pg.edge_swap = d_pg_swap.p;
  pg.end_ptr = d_pg_end_ptr.p; pg.end_ref = d_pg_end_ref.p; pg.blk_ptr = d_pg_blk_ptr.p; pg.blk_ref = d_pg_blk_ref.p;
  pg.up_pos = d_pg_up_pos.p; pg.up_posT = d_pg_up_posT.p; pg.F_diag = d_pg_F_diag.p;
  HK(hipStreamSynchronize(stream));
  return G2OHIP_OK;
}

int g2ohip_ctx::pg_set_measurements(const double* meas7, const double* info36) {
  CK(require_device());
  if (!have_structure || mode != 1) return fail(G2OHIP_ERR_INVALID, "pose-graph structure first");
  if (!meas7 || !info36) return fail(G2OHIP_ERR_INVALID, "measurements and information are required");
  const size_t E = static_cast<size_t>(pgst.E);
  std::vector<double> m(meas7, meas7 + 7 * E), inf(info36, info36 + 36 * E);
  for (size_t k = 0; k < E; ++k) {  // unit quaternions, w >= 0 (as the Isometry3 measurement holds them)
    double* q = &m[7 * k + 3];
    const double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (!(n > 0)) return fail(G2OHIP_ERR_INVALID, "zero measurement quaternion");
    for (int a = 0; a < 4; ++a) q[a] /= n;
  }
  HK(d_pg_meas.upload(m, stream));
  HK(d_pg_info.upload(inf, stream));
  pg.meas = d_pg_meas.p;
  pg.info = d_pg_info.p;
  HK(hipStreamSynchronize(stream));
  have_meas = true;
  have_system = have_solution = false;
  return G2OHIP_OK;
}

int g2ohip_ctx::pg_set_state(const double* poses7) {
  CK(require_device());
  if (!have_structure || mode != 1) return fail(G2OHIP_ERR_INVALID, "pose-graph structure first");
  std::vector<double> p(poses7, poses7 + 7 * static_cast<size_t>(pgst.N));
  for (int i = 0; i < pgst.N; ++i) {
    double* q = &p[7 * static_cast<size_t>(i) + 3];
    const double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (!(n > 0)) return fail(G2OHIP_ERR_INVALID, "zero pose quaternion");
    for (int a = 0; a < 4; ++a) q[a] /= n;
  }
  HK(hipMemcpyAsync(d_cams.p, p.data(), p.size() * sizeof(double), hipMemcpyHostToDevice, stream));
  HK(hipStreamSynchronize(stream));
  have_state = true;
  have_system = have_solution = false;
  return G2OHIP_OK;
}

int g2ohip_ctx::pg_get_state(double* poses7) {
  CK(require_device());
  if (!have_state || mode != 1) return fail(G2OHIP_ERR_INVALID, "no pose-graph state");
  HK(hipStreamSynchronize(stream));
  HK(hipMemcpy(poses7, d_cams.p, 7 * sizeof(double) * pgst.N, hipMemcpyDeviceToHost));
  return G2OHIP_OK;
}

int g2ohip_ctx::pg_chi2(double* chi2, double* err_out) {
  const int A = static_cast<int>(pgst.act.size());
  DBuf<double> es;
  if (err_out) HK(es.alloc(6 * static_cast<size_t>(std::max(A, 1))));
  {
    ProfScope ps(this, G2OHIP_K_CHI2);
    HK(launch_pg_chi2(pg, d_cams.p, d_pg_part.p, err_out ? es.p : nullptr, stream));
  }
  HK(launch_sum(d_pg_part.p, pg_edge_blocks(A), d_scal.p, stream));
  CK(read_scalar(d_scal.p, chi2));
  if (err_out) {
    std::vector<double> h(6 * static_cast<size_t>(A));
    HK(hipMemcpy(h.data(), es.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::fill(err_out, err_out + 6 * static_cast<size_t>(pgst.E), 0.0);
    for (int a = 0; a < A; ++a)
      for (int r = 0; r < 6; ++r) err_out[6 * static_cast<size_t>(pgst.act[a]) + r] = h[6 * static_cast<size_t>(a) + r];
  }
  return G2OHIP_OK;
}

int g2ohip_ctx::pg_build() {
  const int A = static_cast<int>(pgst.act.size());
  {
    ProfScope ps(this, G2OHIP_K_LINEARIZE);
    HK(launch_pg_edges(pg, d_cams.p, d_pg_rec.p, d_pg_part.p, stream));
    HK(launch_pg_assemble(pg, d_pg_rec.p, d_pg_F_val.p, d_pg_Hdiag.p, d_b.p, stream));
  }
  HK(launch_sum(d_pg_part.p, pg_edge_blocks(A), d_scal.p + 1, stream));
  CK(read_scalar(d_scal.p + 1, &chi2_lin));
  have_system = have_lin = true;
  have_exact = false;
  system_prec = G2OHIP_FP64;
  have_solution = false;
  return G2OHIP_OK;
}

int g2ohip_ctx::pg_lambda_init(double* out) {
  if (opt.user_lambda_init > 0) { *out = opt.user_lambda_init; return G2OHIP_OK; }
  HK(launch_pg_max_diag(pgst.Nf, d_pg_Hdiag.p, d_scal.p + 2, stream));
  double m;
  CK(read_scalar(d_scal.p + 2, &m));
  *out = opt.tau * m;
  return G2OHIP_OK;
}

// One PCG solve on the pose graph's full block rows (B.Jinv current), one
// launch per iteration.
hipError_t g2ohip_ctx::pg_pcg_fused(const double* rhs, const PcgControl& ctl, int max_iter, PcgRun* out,
                                    bool* prog_failed) {
  const int nb = pgst.Nf;
  const size_t n = 6 * static_cast<size_t>(nb);
  PrecBufs<double>& B = f64;
  hipError_t e;
  if ((e = B.s2.alloc(n)) != hipSuccess || (e = B.w.alloc(n)) != hipSuccess || (e = B.w2.alloc(n)) != hipSuccess)
    return e;
  double* const s2[2] = {B.s.p, B.s2.p};
  double* const w2[2] = {B.w.p, B.w2.p};
  double* const dirs[2] = {B.d.p, B.d2.p};
  auto init = [&]() {
    return launch_pcg_init<double, 6>(nb, rhs, B.Jinv.p, B.x.p, B.r.p, B.s.p, d_dn_part.p, d_sc.p, stream);
  };
  auto step = [&](int it, int last) {
    return launch_pcg_fused<double, 6>(nb, d_pg_F_head.p, d_pg_F_col.p, d_pg_F_val.p, nullptr, B.Jinv.p, B.x.p, B.r.p,
                                       B.q.p, s2, w2, dirs, d_dn_part.p, d_fpart.p, d_sc.p, ctl, it, last, stream,
                                       pgst.nF);
  };
  return pcg_loop(stream, h_pinned, progress(), d_sc.p, max_iter, init, step, out, prog_failed, ctl.abort);
}

// (H + lambda I) x = b by block-Jacobi PCG on the full block rows
// (linear_solver_pcg.hpp:87-205, with the residual memory).
int g2ohip_ctx::pg_solve(int* ok) {
  const int nb = pgst.Nf;
  const int n = 6 * nb;
  PrecBufs<double>& B = f64;
  *ok = 1;
  HK(hipMemsetAsync(d_err.p, 0, sizeof(int) * 4, stream));
  HK(B.Jinv.alloc(36 * static_cast<size_t>(nb)));
  HK(B.x.alloc(n)); HK(B.r.alloc(n)); HK(B.d.alloc(n)); HK(B.d2.alloc(n)); HK(B.q.alloc(n)); HK(B.s.alloc(n));
  HK(launch_pg_damp(pg, d_pg_Hdiag.p, lambda, d_pg_F_val.p, stream));
  HK((launch_jacobi<double, 6>(nb, d_pg_F_val.p, d_pg_F_diag.p, B.Jinv.p, d_err.p, stream)));
  int err = 0;
  CK(read_err(&err));
  if (err) { *ok = 0; return G2OHIP_OK; }
  const int max_iter = opt.pcg_max_iter < 0 ? n : opt.pcg_max_iter;
  const PcgControl ctl = with_progress(PcgControl{opt.pcg_tolerance, pcg_residual_mem, -1.0});
  double* const dirs[2] = {B.d.p, B.d2.p};
  auto op = [&](int it, int last, const double* d_old, double* d) -> hipError_t {
    return launch_pcg_dir_spmv<double, 6>(nb, d_pg_F_ptr.p, d_pg_F_col.p, d_pg_F_val.p, B.s.p, d_old, d, B.q.p,
                                          d_dn_part.p, d_dq_part.p, d_sc.p, ctl, it, last, stream, nullptr,
                                          pgst.nF);
  };
  PcgRun pr;
  {
    ProfScope ps(this, G2OHIP_K_PCG);
    bool prog_failed = false;
    if (pcg_fused && pcg_fused_fits(nb)) {
      HK(pg_pcg_fused(d_b.p, ctl, max_iter, &pr, &prog_failed));
    } else {
      HK((pcg_drive<double, 6>(stream, h_pinned, progress(), nb, d_b.p, B.Jinv.p, B.x.p, B.r.p, B.s.p, dirs, B.q.p,
                               d_dn_part.p, d_dq_part.p, pcg_spmv_blocks(nb), d_sc.p, max_iter, op, &pr,
                               &prog_failed)));
    }
    if (prog_failed) prog_ok = false;
  }
  pcg_residual_mem = 0.5 * pr.dn_final;
  HK(launch_to_double<double>(n, B.x.p, d_x.p, 0.0, stream));
  last_info[0] = pr.iters;
  last_info[1] = 0;
  last_info[2] = pcg_residual_mem;
  have_solution = true;
  return G2OHIP_OK;
}

// Blocks of H^-1 for a pose graph, whose Hpp is the whole Hessian
// (block_solver.hpp:479-490; the reference's Cholesky-based linear solvers
// compute them from the factor, marginal_covariance_cholesky.h:42-80). Here
// each distinct block column c is found by six block-Jacobi PCG solves
// H X = e_(6c+j) on the undamped system, stopped at r.M^-1 r <= tol * r0.M^-1 r0;
// the requested rows of X are the blocks.
int g2ohip_ctx::pg_marginals(int nblk, const int* rows, const int* cols, double* out, double tol) {
  const int nb = pgst.Nf;
  const int n = 6 * nb;
  PrecBufs<double>& B = f64;
  HK(hipMemsetAsync(d_err.p, 0, sizeof(int) * 4, stream));
  HK(B.Jinv.alloc(36 * static_cast<size_t>(nb)));
  HK(B.x.alloc(n)); HK(B.r.alloc(n)); HK(B.d.alloc(n)); HK(B.d2.alloc(n)); HK(B.q.alloc(n)); HK(B.s.alloc(n));
  HK(launch_pg_damp(pg, d_pg_Hdiag.p, 0.0, d_pg_F_val.p, stream));  // the next solve damps again
  HK((launch_jacobi<double, 6>(nb, d_pg_F_val.p, d_pg_F_diag.p, B.Jinv.p, d_err.p, stream)));
  int err = 0;
  CK(read_err(&err));
  if (err) return fail(G2OHIP_ERR_NOT_PD, "a diagonal block of H is not positive definite");
  const PcgControl ctl = with_progress(PcgControl{tol > 0 ? tol : 1e-24, -1.0, -1.0});
  double* const dirs[2] = {B.d.p, B.d2.p};
  auto op = [&](int it, int last, const double* d_old, double* d) -> hipError_t {
    return launch_pcg_dir_spmv<double, 6>(nb, d_pg_F_ptr.p, d_pg_F_col.p, d_pg_F_val.p, B.s.p, d_old, d, B.q.p,
                                          d_dn_part.p, d_dq_part.p, d_sc.p, ctl, it, last, stream, nullptr,
                                          pgst.nF);
  };
  g2ohip::DBuf<double> rhs;
  HK(rhs.alloc(n));
  std::vector<double> e(n, 0.0), x(n);
  std::vector<int> ucols(cols, cols + nblk);
  std::sort(ucols.begin(), ucols.end());
  ucols.erase(std::unique(ucols.begin(), ucols.end()), ucols.end());
  for (int c : ucols) {
    for (int j = 0; j < 6; ++j) {
      e[6 * static_cast<size_t>(c) + j] = 1.0;
      HK(hipMemcpyAsync(rhs.p, e.data(), n * sizeof(double), hipMemcpyHostToDevice, stream));
      e[6 * static_cast<size_t>(c) + j] = 0.0;
      PcgRun pr;
      {
        ProfScope ps(this, G2OHIP_K_PCG);
        bool prog_failed = false;
        if (pcg_fused && pcg_fused_fits(nb)) {
          HK(pg_pcg_fused(rhs.p, ctl, n, &pr, &prog_failed));
        } else {
          HK((pcg_drive<double, 6>(stream, h_pinned, progress(), nb, rhs.p, B.Jinv.p, B.x.p, B.r.p, B.s.p, dirs,
                                   B.q.p, d_dn_part.p, d_dq_part.p, pcg_spmv_blocks(nb), d_sc.p, n, op, &pr,
                                   &prog_failed)));
        }
        if (prog_failed) prog_ok = false;
      }
      HK(hipMemcpyAsync(x.data(), B.x.p, n * sizeof(double), hipMemcpyDeviceToHost, stream));
      HK(hipStreamSynchronize(stream));
      for (int k = 0; k < nblk; ++k) {
        if (cols[k] != c) continue;
        for (int i = 0; i < 6; ++i)
          out[36 * static_cast<size_t>(k) + 6 * i + j] = x[6 * static_cast<size_t>(rows[k]) + i];
      }
    }
  }
  have_solution = false;
  return G2OHIP_OK;
}

int g2ohip_ctx::pg_update() {
  {
    ProfScope ps(this, G2OHIP_K_UPDATE);
    HK(launch_pg_update(pg, d_x.p, d_cams.p, stream));
  }
  HK(hipStreamSynchronize(stream));
  return G2OHIP_OK;
}

int g2ohip_ctx::pg_get_system(int* rows, int* cols, double* blocks, double* b) {
  if (!have_system || mode != 1) return fail(G2OHIP_ERR_INVALID, "build_system first");
  HK(hipStreamSynchronize(stream));
  std::vector<double> F(36 * static_cast<size_t>(pgst.nF)), Hd(36 * static_cast<size_t>(pgst.Nf));
  HK(hipMemcpy(F.data(), d_pg_F_val.p, F.size() * sizeof(double), hipMemcpyDeviceToHost));
  HK(hipMemcpy(Hd.data(), d_pg_Hdiag.p, Hd.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int k = 0; k < pgst.nU; ++k) {
    if (rows) rows[k] = pgst.U_row[k];
    if (cols) cols[k] = pgst.U_col[k];
    if (blocks) {
      const double* src = k < pgst.Nf ? &Hd[36 * static_cast<size_t>(k)] : &F[36 * static_cast<size_t>(pgst.up_pos[k])];
      std::copy(src, src + 36, blocks + 36 * static_cast<size_t>(k));
    }
  }
  if (b) HK(hipMemcpy(b, d_b.p, 6 * sizeof(double) * pgst.Nf, hipMemcpyDeviceToHost));
  return G2OHIP_OK;
}
