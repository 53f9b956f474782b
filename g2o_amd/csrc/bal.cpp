// BAL problems on the device: g2o's BAL example (g2o/examples/bal/
// bal_example.cpp:70-207, 227-341) with BlockSolver<BlockSolverTraits<9, 3>>,
// Schur, block-Jacobi PCG on the 9x9 reduced camera system and LM. The LM
// loop of context.cpp drives these through the generic Solver calls.
#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "context.h"
#include "pcg_drive.h"

using namespace g2ohip;

// BlockSolver<9, 3>::buildStructure (block_solver.hpp:110-269) and the index
// mapping (sparse_optimizer.cpp:175-200) for EdgeObservationBAL edges between
// VertexCameraBAL and marginalised VertexPointBAL vertices. Besides the
// shared structure (slots, camera-major Hpl order, Hschur pattern), BAL keeps
// per free camera the list of its slots and, per off-diagonal upper Hschur
// block, the observation pairs that form it.
int g2ohip_ctx::bal_set_structure(int C, int P, int E, const int* ec, const int* ep, const uint8_t* cf,
                                  const uint8_t* pf) {
  if (!ec || !ep) return fail(G2OHIP_ERR_INVALID, "edge index arrays are required");
  LinPlanCaps caps;
  caps.max_point_cams = INT_MAX;  // the linearisation chunks are not used here
  std::string msg = build_structure(st, C, P, E, ec, ep, cf, pf, true, &caps);
  if (!msg.empty()) {
    have_structure = false;
    return fail(msg.find("range") != std::string::npos || msg.find("invalid") != std::string::npos
                    ? G2OHIP_ERR_INVALID
                    : G2OHIP_ERR_UNSUPPORTED,
                msg);
  }
  // one observation per (free camera, free point): the diagonal Schur blocks
  // are summed per observation
  for (int i = 0; i < st.npos; ++i) {
    if (st.pos_h[i] < 0) continue;
    for (int s = st.pos_ptr[i] + 1; s < st.pos_ptr[i] + st.pos_nfree[i]; ++s)
      if (st.slot_camh[s] == st.slot_camh[s - 1]) {
        have_structure = false;
        return fail(G2OHIP_ERR_UNSUPPORTED, "a camera observes a point twice");
      }
  }
  // off-diagonal upper blocks in S order and their observation pairs
  std::vector<int> od_of(st.nS, -1), od_blk;
  for (int k = 0; k < st.nS; ++k)
    if (st.S_row[k] != st.S_col[k]) { od_of[k] = static_cast<int>(od_blk.size()); od_blk.push_back(k); }
  const int nO = static_cast<int>(od_blk.size());
  auto block_of = [&](int a, int b) {
    const int* b0 = st.S_col.data() + st.S_rowptr[a];
    const int* b1 = st.S_col.data() + st.S_rowptr[a + 1];
    return static_cast<int>(std::lower_bound(b0, b1, b) - st.S_col.data());
  };
  std::vector<long long> cnt(static_cast<size_t>(nO) + 1, 0);
  for (int i = 0; i < st.npos; ++i) {
    if (st.pos_h[i] < 0) continue;
    const int s0 = st.pos_ptr[i], s1 = s0 + st.pos_nfree[i];
    for (int a = s0; a < s1; ++a)
      for (int b = a + 1; b < s1; ++b) cnt[od_of[block_of(st.slot_camh[a], st.slot_camh[b])] + 1]++;
  }
  for (int o = 0; o < nO; ++o) cnt[o + 1] += cnt[o];
  const long long npairs = cnt[nO];
  if (npairs >= (1LL << 31)) {
    have_structure = false;
    return fail(G2OHIP_ERR_UNSUPPORTED, "more than 2^31 off-diagonal Schur pair products");
  }
  mode = 2;
  have_structure = true;
  have_meas = have_state = have_system = have_solution = have_exact = have_lin = false;
  pcg_residual_mem = -1.0;
  stack_depth = 0;
  bal_pairs = npairs;
  if (device < 0) return G2OHIP_OK;
  std::vector<int2> pairs(static_cast<size_t>(npairs));
  {
    std::vector<long long> fill(cnt.begin(), cnt.end() - 1);
    for (int i = 0; i < st.npos; ++i) {
      if (st.pos_h[i] < 0) continue;
      const int s0 = st.pos_ptr[i], s1 = s0 + st.pos_nfree[i];
      for (int a = s0; a < s1; ++a)
        for (int b = a + 1; b < s1; ++b) {
          const int o = od_of[block_of(st.slot_camh[a], st.slot_camh[b])];
          pairs[fill[o]++] = make_int2(st.slot_q[a], st.slot_q[b]);
        }
    }
  }
  std::vector<int> od_ptr(static_cast<size_t>(nO) + 1);
  for (int o = 0; o <= nO; ++o) od_ptr[o] = static_cast<int>(cnt[o]);
  // per free camera: its slots (with any point), slot order
  std::vector<int> cs_ptr(static_cast<size_t>(st.Cf) + 1, 0), cs_slot;
  for (int s = 0; s < st.ns; ++s)
    if (st.slot_camh[s] >= 0) cs_ptr[st.slot_camh[s] + 1]++;
  for (int c = 0; c < st.Cf; ++c) cs_ptr[c + 1] += cs_ptr[c];
  cs_slot.resize(cs_ptr[st.Cf]);
  {
    std::vector<int> fill(cs_ptr.begin(), cs_ptr.end() - 1);
    for (int s = 0; s < st.ns; ++s)
      if (st.slot_camh[s] >= 0) cs_slot[fill[st.slot_camh[s]]++] = s;
  }
  std::vector<int4> rec(st.ns);
  for (int s = 0; s < st.ns; ++s)
    rec[s] = make_int4(st.slot_cam[s], st.pos_pt[st.slot_pos[s]], st.slot_q[s], st.slot_camh[s]);
  std::vector<int> fdiag(st.Cf);
  for (int c = 0; c < st.Cf; ++c) fdiag[c] = st.up_pos[st.S_diag[c]];
  HK(d_camH.upload(st.camH, stream));
  HK(d_ptH.upload(st.ptH, stream));
  HK(d_cam_raw.upload(st.free_cam_raw, stream));
  HK(d_bal_rec.upload(rec, stream));
  HK(d_pos_ptr.upload(st.pos_ptr, stream));
  HK(d_pos_h.upload(st.pos_h, stream));
  HK(d_bal_cs_ptr.upload(cs_ptr, stream));
  HK(d_bal_cs_slot.upload(cs_slot, stream));
  HK(d_cam_qptr.upload(st.cam_qptr, stream));
  HK(d_q_ph.upload(st.q_ph, stream));
  HK(d_bal_od_ptr.upload(od_ptr, stream));
  HK(d_bal_od_blk.upload(od_blk, stream));
  HK(d_bal_od_pair.upload(pairs, stream));
  HK(d_F_ptr.upload(st.F_ptr, stream));
  HK(d_F_col.upload(st.F_col, stream));
  st.F_head = row_heads(st.F_ptr, st.F_col);
  HK(d_F_head.upload(st.F_head, stream));
  HK(d_F_diag.upload(fdiag, stream));
  HK(d_up_pos.upload(st.up_pos, stream));
  HK(d_up_posT.upload(st.up_posT, stream));
  const size_t nx = 9 * static_cast<size_t>(st.Cf) + 3 * static_cast<size_t>(st.Pf);
  HK(d_cams.alloc(9 * static_cast<size_t>(C)));
  HK(d_pts.alloc(3 * static_cast<size_t>(std::max(P, 1))));
  HK(d_bal_tab.alloc(BAL_TAB * static_cast<size_t>(C)));
  HK(d_Hll.alloc(6 * static_cast<size_t>(std::max(st.Pf, 1))));
  HK(d_b.alloc(nx));
  HK(d_x.alloc(nx));
  HK(d_bal_Hpp.alloc(81 * static_cast<size_t>(st.Cf)));
  HK(d_bal_Hpl.alloc(27 * static_cast<size_t>(std::max(st.nq, 1))));
  HK(d_bal_Dinv.alloc(9 * static_cast<size_t>(std::max(st.Pf, 1))));
  HK(d_linv.alloc(6 * static_cast<size_t>(std::max(st.Pf, 1))));
  HK(d_bal_F_val.alloc(81 * static_cast<size_t>(st.nF)));
  HK(d_bschur.alloc(9 * static_cast<size_t>(st.Cf)));
  HK(d_bal_part.alloc(std::max(bal_point_blocks(st.npos), bal_chi2_blocks())));
  HK(d_red.alloc(RED_BLOCKS));  // computeScale's partials
  HK(d_scal.alloc(8));
  HK(d_err.alloc(4));
  HK(d_sc.alloc(1));
  HK(d_dq_part.alloc(std::max(pcg_spmv_blocks(st.Cf), 1)));
  HK(d_dn_part.alloc(std::max(pcg_row_blocks(st.Cf, 9), 1)));
  HK(d_fpart.alloc(8 * static_cast<size_t>(pcg_fused_blocks(std::max(st.Cf, 1)))));
  HK(hipMemsetAsync(d_b.p, 0, nx * sizeof(double), stream));
  HK(hipMemsetAsync(d_x.p, 0, nx * sizeof(double), stream));
  bal = DevBAL{};
  bal.C = st.C; bal.P = st.P; bal.Cf = st.Cf; bal.Pf = st.Pf; bal.ns = st.ns; bal.npos = st.npos; bal.nq = st.nq;
  bal.nO = nO;
  bal.camH = d_camH.p; bal.ptH = d_ptH.p; bal.cam_raw = d_cam_raw.p;
  bal.slot_rec = d_bal_rec.p;
  bal.pos_ptr = d_pos_ptr.p; bal.pos_h = d_pos_h.p;
  bal.cs_ptr = d_bal_cs_ptr.p; bal.cs_slot = d_bal_cs_slot.p;
  bal.cam_qptr = d_cam_qptr.p; bal.q_ph = d_q_ph.p;
  bal.od_ptr = d_bal_od_ptr.p; bal.od_blk = d_bal_od_blk.p; bal.od_pair = d_bal_od_pair.p;
  bal.F_diag = d_F_diag.p; bal.up_pos = d_up_pos.p; bal.up_posT = d_up_posT.p;
  HK(hipStreamSynchronize(stream));  // the host vectors go out of scope
  return G2OHIP_OK;
}

// EdgeObservationBAL measurements (bal_example.cpp:305-315): meas[E*2],
// information weight * I2 (weight NULL = 1, the example's identity), Huber
// delta > 0 attaches RobustKernelHuber to every edge.
int g2ohip_ctx::bal_set_measurements(const double* meas, const double* w, double huber) {
  if (mode != 2 || !have_structure) return fail(G2OHIP_ERR_INVALID, "BAL structure first");
  if (!meas) return fail(G2OHIP_ERR_INVALID, "measurements are required");
  CK(require_device());
  std::vector<double> m(2 * static_cast<size_t>(st.ns)), ww(st.ns);
  for (int s = 0; s < st.ns; ++s) {
    const int k = st.slot_edge[s];
    m[2 * static_cast<size_t>(s)] = meas[2 * static_cast<size_t>(k)];
    m[2 * static_cast<size_t>(s) + 1] = meas[2 * static_cast<size_t>(k) + 1];
    ww[s] = w ? w[k] : 1.0;
    if (!(ww[s] > 0.0) || !std::isfinite(ww[s])) return fail(G2OHIP_ERR_INVALID, "information weights must be > 0");
  }
  HK(d_meas.upload(m, stream));
  HK(d_w.upload(ww, stream));
  bal.meas = d_meas.p;
  bal.w = d_w.p;
  bal.huber = huber;
  HK(hipStreamSynchronize(stream));
  have_meas = true;
  have_system = have_solution = false;
  return G2OHIP_OK;
}

int g2ohip_ctx::bal_set_state(const double* cams9, const double* pts) {
  if (mode != 2 || !have_structure) return fail(G2OHIP_ERR_INVALID, "BAL structure first");
  if (!cams9 || (!pts && st.P > 0)) return fail(G2OHIP_ERR_INVALID, "camera and point values are required");
  CK(require_device());
  HK(hipMemcpyAsync(d_cams.p, cams9, 9 * sizeof(double) * st.C, hipMemcpyHostToDevice, stream));
  if (st.P > 0) HK(hipMemcpyAsync(d_pts.p, pts, 3 * sizeof(double) * st.P, hipMemcpyHostToDevice, stream));
  HK(hipStreamSynchronize(stream));
  have_state = true;
  have_system = have_solution = false;
  return G2OHIP_OK;
}

int g2ohip_ctx::bal_get_state(double* cams9, double* pts) {
  if (mode != 2 || !have_state) return fail(G2OHIP_ERR_INVALID, "no BAL state");
  CK(require_device());
  if (cams9) HK(hipMemcpyAsync(cams9, d_cams.p, 9 * sizeof(double) * st.C, hipMemcpyDeviceToHost, stream));
  if (pts && st.P > 0) HK(hipMemcpyAsync(pts, d_pts.p, 3 * sizeof(double) * st.P, hipMemcpyDeviceToHost, stream));
  HK(hipStreamSynchronize(stream));
  return G2OHIP_OK;
}

int g2ohip_ctx::bal_chi2(double* chi2, double* err_out) {
  HK(launch_bal_cam_table(st.C, d_cams.p, d_bal_tab.p, stream));
  DBuf<double> es;
  if (err_out) HK(es.alloc(2 * static_cast<size_t>(std::max(st.ns, 1))));
  {
    ProfScope ps(this, G2OHIP_K_CHI2);
    HK(launch_bal_chi2(bal, d_bal_tab.p, d_pts.p, d_bal_part.p, d_scal.p, err_out ? es.p : nullptr, stream));
  }
  CK(read_scalar(d_scal.p, chi2));
  if (err_out) {
    std::vector<double> h(2 * static_cast<size_t>(st.ns));
    HK(hipMemcpy(h.data(), es.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::fill(err_out, err_out + 2 * static_cast<size_t>(st.E), 0.0);
    for (int s = 0; s < st.ns; ++s) {
      const int k = st.slot_edge[s];
      err_out[2 * static_cast<size_t>(k)] = h[2 * static_cast<size_t>(s)];
      err_out[2 * static_cast<size_t>(k) + 1] = h[2 * static_cast<size_t>(s) + 1];
    }
  }
  return G2OHIP_OK;
}

// buildSystem (block_solver.hpp:492-565): the linearisation of every active
// edge, FP64; the BAL path has no FP32 or implicit-Schur variant.
int g2ohip_ctx::bal_build() {
  if (opt.precision != G2OHIP_FP64 || opt.schur_mode != G2OHIP_SCHUR_EXPLICIT)
    return fail(G2OHIP_ERR_UNSUPPORTED, "BAL problems run in FP64 with the explicit Schur complement");
  HK(launch_bal_cam_table(st.C, d_cams.p, d_bal_tab.p, stream));
  {
    ProfScope ps(this, G2OHIP_K_LINEARIZE);
    HK(launch_bal_linearize(bal, d_bal_tab.p, d_pts.p, d_bal_Hpp.p, d_Hll.p, d_bal_Hpl.p, d_b.p, d_bal_part.p,
                            d_scal.p + 1, stream));
  }
  CK(read_scalar(d_scal.p + 1, &chi2_lin));
  have_system = have_lin = true;
  have_exact = false;
  system_prec = G2OHIP_FP64;
  have_solution = false;
  return G2OHIP_OK;
}

int g2ohip_ctx::bal_lambda_init(double* out) {
  if (opt.user_lambda_init > 0) { *out = opt.user_lambda_init; return G2OHIP_OK; }
  HK(launch_bal_max_diag(bal, d_bal_Hpp.p, d_Hll.p, d_red.p, d_scal.p + 2, stream));
  double m;
  CK(read_scalar(d_scal.p + 2, &m));
  *out = opt.tau * m;
  return G2OHIP_OK;
}

// BlockSolver<9, 3>::solve (block_solver.hpp:331-477) at the current lambda:
// Schur complement, block-Jacobi PCG on the reduced camera system
// (linear_solver_pcg.hpp:87-205, with the residual memory), back-substitution.
int g2ohip_ctx::bal_solve(int* ok) {
  const int nb = st.Cf;
  const int n = 9 * nb;
  PrecBufs<double>& B = f64;
  *ok = 1;
  HK(hipMemsetAsync(d_err.p, 0, sizeof(int) * 4, stream));
  HK(B.Jinv.alloc(81 * static_cast<size_t>(nb)));
  HK(B.x.alloc(n)); HK(B.r.alloc(n)); HK(B.d.alloc(n)); HK(B.d2.alloc(n)); HK(B.q.alloc(n)); HK(B.s.alloc(n));
  if (bal_mfma) HK(d_bal_G.alloc(27 * static_cast<size_t>(std::max(st.nq, 1))));
  {
    ProfScope ps(this, G2OHIP_K_SCHUR);
    HK(launch_bal_schur_diag(bal, d_bal_Hpp.p, d_Hll.p, d_bal_Hpl.p, d_b.p, lambda, d_bal_Dinv.p, d_linv.p,
                             bal_mfma ? d_bal_G.p : nullptr, d_bal_F_val.p, d_bschur.p, d_err.p, stream));
    {
      // the off-diagonal products alone, also under their own timer
      ProfScope po(this, G2OHIP_K_SCHUR_WINDOWS);
      HK(launch_bal_schur_offdiag(bal, d_bal_Hpl.p, d_bal_Dinv.p, d_bal_G.p, d_bal_F_val.p, bal_mfma, stream));
    }
    HK((launch_jacobi<double, 9>(nb, d_bal_F_val.p, d_F_diag.p, B.Jinv.p, d_err.p, stream)));
  }
  int err = 0;
  CK(read_err(&err));
  if (err) {
    *ok = 0;
    have_solution = true;
    last_info[0] = 0;
    return G2OHIP_OK;
  }
  const int max_iter = opt.pcg_max_iter < 0 ? n : opt.pcg_max_iter;
  const PcgControl ctl = with_progress(PcgControl{opt.pcg_tolerance, pcg_residual_mem, -1.0});
  double* const dirs[2] = {B.d.p, B.d2.p};
  auto op = [&](int it, int last, const double* d_old, double* d) -> hipError_t {
    return launch_pcg_dir_spmv<double, 9>(nb, d_F_ptr.p, d_F_col.p, d_bal_F_val.p, B.s.p, d_old, d, B.q.p,
                                          d_dn_part.p, d_dq_part.p, d_sc.p, ctl, it, last, stream, nullptr, st.nF);
  };
  PcgRun pr;
  {
    ProfScope ps(this, G2OHIP_K_PCG);
    bool prog_failed = false;
    if (pcg_fused && pcg_fused_fits(nb)) {
      // one launch per iteration
      HK(B.s2.alloc(n)); HK(B.w.alloc(n)); HK(B.w2.alloc(n));
      double* const s2[2] = {B.s.p, B.s2.p};
      double* const w2[2] = {B.w.p, B.w2.p};
      auto init = [&]() {
        return launch_pcg_init<double, 9>(nb, d_bschur.p, B.Jinv.p, B.x.p, B.r.p, B.s.p, d_dn_part.p, d_sc.p, stream);
      };
      auto step = [&](int it, int last) {
        return launch_pcg_fused<double, 9>(nb, d_F_head.p, d_F_col.p, d_bal_F_val.p, nullptr, B.Jinv.p, B.x.p, B.r.p,
                                           B.q.p, s2, w2, dirs, d_dn_part.p, d_fpart.p, d_sc.p, ctl, it, last,
                                           stream, st.nF);
      };
      HK(pcg_loop(stream, h_pinned, progress(), d_sc.p, max_iter, init, step, &pr, &prog_failed, ctl.abort));
    } else {
      HK((pcg_drive<double, 9>(stream, h_pinned, progress(), nb, d_bschur.p, B.Jinv.p, B.x.p, B.r.p, B.s.p, dirs,
                               B.q.p, d_dn_part.p, d_dq_part.p, pcg_spmv_blocks(nb), d_sc.p, max_iter, op, &pr,
                               &prog_failed)));
    }
    if (prog_failed) prog_ok = false;
  }
  pcg_residual_mem = 0.5 * pr.dn_final;
  HK(launch_copy(d_x.p, B.x.p, sizeof(double) * n, stream));
  {
    ProfScope ps(this, G2OHIP_K_REFINE);
    HK(launch_bal_backsub(bal, d_bal_Hpl.p, d_bal_Dinv.p, d_b.p, d_x.p, stream));
  }
  last_info[0] = pr.iters;
  last_info[1] = 0;
  last_info[2] = pcg_residual_mem;
  have_solution = true;
  return G2OHIP_OK;
}

int g2ohip_ctx::bal_update() {
  {
    ProfScope ps(this, G2OHIP_K_UPDATE);
    HK(launch_bal_update(bal, d_x.p, d_cams.p, d_pts.p, stream));
  }
  HK(hipStreamSynchronize(stream));
  return G2OHIP_OK;
}

// The blocks of the last linearisation: Hpp [Cf][9][9], Hll [Pf][3][3],
// Hpl_edge [E][9][3] (the edge's block, zero without one).
int g2ohip_ctx::bal_get_hessian(double* Hpp, double* Hll, double* Hpl_edge) {
  if (mode != 2 || !have_lin) return fail(G2OHIP_ERR_INVALID, "build_system first");
  HK(hipStreamSynchronize(stream));
  if (Hpp) HK(hipMemcpy(Hpp, d_bal_Hpp.p, 81 * sizeof(double) * st.Cf, hipMemcpyDeviceToHost));
  if (Hll) {
    std::vector<double> h(6 * static_cast<size_t>(st.Pf));
    if (st.Pf) HK(hipMemcpy(h.data(), d_Hll.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    static const int full[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
    for (int l = 0; l < st.Pf; ++l)
      for (int k = 0; k < 9; ++k)
        Hll[9 * static_cast<size_t>(st.ph_pub[l]) + k] = h[6 * static_cast<size_t>(l) + full[k]];
  }
  if (Hpl_edge) {
    std::vector<double> h(27 * static_cast<size_t>(st.nq));
    if (st.nq) HK(hipMemcpy(h.data(), d_bal_Hpl.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::fill(Hpl_edge, Hpl_edge + 27 * static_cast<size_t>(st.E), 0.0);
    for (int s = 0; s < st.ns; ++s) {
      const int q = st.slot_q[s];
      if (q < 0) continue;
      std::copy(&h[27 * static_cast<size_t>(q)], &h[27 * static_cast<size_t>(q)] + 27,
                Hpl_edge + 27 * static_cast<size_t>(st.slot_edge[s]));
    }
  }
  return G2OHIP_OK;
}

// The reduced system of the last solve: upper blocks in (row, col) order
// (rows, cols, 81 values each) and bschur [9 Cf].
int g2ohip_ctx::bal_get_schur(int* rows, int* cols, double* blocks, double* bschur) {
  if (mode != 2 || !have_solution) return fail(G2OHIP_ERR_INVALID, "solve first");
  HK(hipStreamSynchronize(stream));
  std::vector<double> F(81 * static_cast<size_t>(st.nF));
  HK(hipMemcpy(F.data(), d_bal_F_val.p, F.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int k = 0; k < st.nS; ++k) {
    if (rows) rows[k] = st.S_row[k];
    if (cols) cols[k] = st.S_col[k];
    if (blocks)
      std::copy(&F[81 * static_cast<size_t>(st.up_pos[k])], &F[81 * static_cast<size_t>(st.up_pos[k])] + 81,
                blocks + 81 * static_cast<size_t>(k));
  }
  if (bschur) HK(hipMemcpy(bschur, d_bschur.p, 9 * sizeof(double) * st.Cf, hipMemcpyDeviceToHost));
  return G2OHIP_OK;
}
