// Device context: structure upload, the g2o Solver operations on the device
// and the Levenberg-Marquardt loop (see context.h).
#include "context.h"
#include "pcg_drive.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <type_traits>

using namespace g2ohip;

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <typename HT>
PrecBufs<HT>& bufs_of(g2ohip_ctx& c);
template <>
PrecBufs<float>& bufs_of<float>(g2ohip_ctx& c) { return c.f32; }
template <>
PrecBufs<double>& bufs_of<double>(g2ohip_ctx& c) { return c.f64; }

}  // namespace

g2ohip_ctx::~g2ohip_ctx() {
  for (auto e : ev_pool) (void)hipEventDestroy(e);
  for (auto* b : stack_c) delete b;
  for (auto* b : stack_p) delete b;
  if (h_pinned) (void)hipHostFree(h_pinned);
  if (h_prog) (void)hipHostFree(h_prog);
  if (ev_fork) (void)hipEventDestroy(ev_fork);
  if (ev_join) (void)hipEventDestroy(ev_join);
  if (side) (void)hipStreamDestroy(side);
  if (stream) (void)hipStreamDestroy(stream);
}

int g2ohip_ctx::fail(int code, const std::string& msg) {
  return fail_nothrow(code, msg.c_str());
}

int g2ohip_ctx::fail_nothrow(int code, const char* msg) noexcept {
  try {
    last_error.assign(msg);
  } catch (...) {
    last_error.clear();  // no memory for the message: the code still reports the failure
  }
  return code;
}

int g2ohip_ctx::require_device() {
  if (device < 0) return fail(G2OHIP_ERR_NO_DEVICE, "context has no device");
  return G2OHIP_OK;
}

int g2ohip_ctx::hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return G2OHIP_OK;
  if (e == hipErrorOutOfMemory) return fail(G2OHIP_ERR_OOM, std::string(what) + ": out of device memory");
  return fail(G2OHIP_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}


int g2ohip_ctx::read_scalar(const double* dptr, double* out) {
  HK(hipMemcpyAsync(h_pinned, dptr, sizeof(double), hipMemcpyDeviceToHost, stream));
  HK(hipStreamSynchronize(stream));
  *out = h_pinned[0];
  return G2OHIP_OK;
}

int g2ohip_ctx::read_err(int* out) {
  HK(hipMemcpyAsync(h_pinned, d_err.p, sizeof(int), hipMemcpyDeviceToHost, stream));
  HK(hipStreamSynchronize(stream));
  std::memcpy(out, h_pinned, sizeof(int));
  return G2OHIP_OK;
}

// d_scal[0..3] (chi2, linearisation chi2, max diagonal, scale) and the error
// word in one round trip: published by a kernel into host-mapped memory
// (the host spins on the sequence number, no stream synchronisation), or
// copied after a synchronisation if that is unavailable or stalls.
int g2ohip_ctx::read_trial_scalars(double* sc4, int* err) {
  if (h_trial && prog_ok) {
    const int seq = ++trial_seq;
    HK(launch_publish_scalars(d_scal.p, d_err.p, d_trial, seq, stream));
    volatile TrialScalars* t = h_trial;
    const auto t0 = std::chrono::steady_clock::now();
    bool have = false;
    for (int k = 0;; ++k) {
      if (t->seq == seq) { have = true; break; }
      if ((k & 255) == 255 &&
          std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() > 200.0)
        break;
    }
    if (!have) {
      HK(hipStreamSynchronize(stream));
      have = t->seq == seq;
    }
    if (have) {
      for (int k = 0; k < 4; ++k) sc4[k] = t->v[k];
      *err = t->err;
      return G2OHIP_OK;
    }
    prog_ok = false;  // the mapped memory does not reach the host
  }
  char* hp = reinterpret_cast<char*>(h_pinned);
  HK(hipMemcpyAsync(hp, d_scal.p, 4 * sizeof(double), hipMemcpyDeviceToHost, stream));
  HK(hipMemcpyAsync(hp + 64, d_err.p, sizeof(int), hipMemcpyDeviceToHost, stream));
  HK(hipStreamSynchronize(stream));
  std::memcpy(sc4, hp, 4 * sizeof(double));
  std::memcpy(err, hp + 64, sizeof(int));
  return G2OHIP_OK;
}

// ---------------------------------------------------------------- timing
int g2ohip_ctx::prof_begin() {
  if (!profiling) return -1;
  if (ev_used + 2 > ev_pool.size()) {
    for (int i = 0; i < 256; ++i) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return -1;
      ev_pool.push_back(e);
    }
  }
  int i = static_cast<int>(ev_used++);
  (void)hipEventRecord(ev_pool[i], stream);
  return i;
}

void g2ohip_ctx::prof_end(int kernel, int begin) {
  if (begin < 0 || !profiling) return;
  int i = static_cast<int>(ev_used++);
  (void)hipEventRecord(ev_pool[i], stream);
  ev_marks[kernel].push_back({begin, i});
}

int g2ohip_ctx::set_profiling(int on) {
  CK(require_device());
  HK(hipStreamSynchronize(stream));
  profiling = on != 0;
  ev_used = 0;
  for (auto& m : ev_marks) m.clear();
  return G2OHIP_OK;
}

int g2ohip_ctx::get_profile(int kernel, double* total_ms, int* launches) {
  if (kernel < 0 || kernel >= G2OHIP_K_COUNT) return fail(G2OHIP_ERR_INVALID, "unknown kernel id");
  HK(hipStreamSynchronize(stream));
  double t = 0;
  for (auto& m : ev_marks[kernel]) {
    float ms = 0;
    HK(hipEventElapsedTime(&ms, ev_pool[m.first], ev_pool[m.second]));
    t += ms;
  }
  if (total_ms) *total_ms = t;
  if (launches) *launches = static_cast<int>(ev_marks[kernel].size());
  return G2OHIP_OK;
}


// ------------------------------------------------------------- structure
int g2ohip_ctx::set_structure(int C, int P, int E, const int* ec, const int* ep, const uint8_t* cf,
                              const uint8_t* pf) {
  if (!ec || !ep) return fail(G2OHIP_ERR_INVALID, "edge index arrays are required");
  std::string msg = build_structure(st, C, P, E, ec, ep, cf, pf, true);
  if (!msg.empty()) {
    have_structure = false;
    return fail(msg.find("range") != std::string::npos || msg.find("invalid") != std::string::npos
                    ? G2OHIP_ERR_INVALID
                    : G2OHIP_ERR_UNSUPPORTED,
                msg);
  }
  have_structure = true;
  mode = 0;
  have_meas = have_state = have_system = have_solution = have_exact = have_lin = false;
  win_ready[0] = win_ready[1] = false;
  schur_rows_chosen = -1;
  row_nitems = row_max_blk = 0;
  pcg_residual_mem = -1.0;
  stack_depth = 0;
  if (device < 0) return G2OHIP_OK;
  HK(d_camH.upload(st.camH, stream));
  HK(d_ptH.upload(st.ptH, stream));
  HK(d_pos_pt.upload(st.pos_pt, stream));
  HK(d_pos_h.upload(st.pos_h, stream));
  HK(d_pos_ptr.upload(st.pos_ptr, stream));
  HK(d_pos_nfree.upload(st.pos_nfree, stream));
  HK(d_slot_cam.upload(st.slot_cam, stream));
  HK(d_slot_pos.upload(st.slot_pos, stream));
  HK(d_slot_camh.upload(st.slot_camh, stream));
  HK(d_slot_ph.upload(st.slot_ph, stream));
  HK(d_slot_q.upload(st.slot_q, stream));
  HK(d_cam_qptr.upload(st.cam_qptr, stream));
  HK(d_q_ph.upload(st.q_ph, stream));
  HK(d_q_slot.upload(st.q_slot, stream));
  HK(d_q_pt.upload(st.q_pt, stream));
  {
    std::vector<int4> rec(st.ns);
    for (int s = 0; s < st.ns; ++s) {
      const int pos = st.slot_pos[s], cam = st.slot_cam[s];
      rec[s] = make_int4(pos, st.pos_h[pos], cam, st.camH[cam]);
    }
    HK(d_slot_rec.upload(rec, stream));
    HK(hipStreamSynchronize(stream));  // the host vectors go out of scope
  }
  HK(d_cam_raw.upload(st.free_cam_raw, stream));
  HK(d_lc_pos0.upload(st.lc_pos0, stream));
  HK(d_lc_cam_ptr.upload(st.lc_cam_ptr, stream));
  HK(d_lc_cams.upload(st.lc_cams, stream));
  HK(d_camp_ptr.upload(st.camp_ptr, stream));
  HK(d_camp_idx.upload(st.camp_idx, stream));
  HK(d_S_row.upload(st.S_row, stream));
  HK(d_S_col.upload(st.S_col, stream));
  HK(d_S_diag.upload(st.S_diag, stream));
  HK(d_up_pos.upload(st.up_pos, stream));
  HK(d_up_posT.upload(st.up_posT, stream));
  HK(d_F_ptr.upload(st.F_ptr, stream));
  HK(d_F_col.upload(st.F_col, stream));
  st.F_head = row_heads(st.F_ptr, st.F_col);
  HK(d_F_head.upload(st.F_head, stream));
  std::vector<int> fdiag(st.Cf), iota(st.Cf);
  for (int c = 0; c < st.Cf; ++c) { fdiag[c] = st.up_pos[st.S_diag[c]]; iota[c] = c; }
  HK(d_F_diag.upload(fdiag, stream));
  HK(d_iota.upload(iota, stream));
  const size_t nx = 6 * static_cast<size_t>(st.Cf) + 3 * static_cast<size_t>(st.Pf);
  HK(d_cams.alloc(7 * static_cast<size_t>(C)));
  HK(d_pts.alloc(3 * static_cast<size_t>(std::max(P, 1))));
  HK(d_camRt.alloc(12 * static_cast<size_t>(C)));
  HK(d_Hll.alloc(6 * static_cast<size_t>(std::max(st.Pf, 1))));
  HK(d_b.alloc(nx));
  HK(d_x.alloc(nx));
  HK(d_Hpp.alloc(36 * static_cast<size_t>(st.Cf)));
  HK(d_hpp_part.alloc(27 * std::max<size_t>(st.lc_cams.size(), 1)));
  HK(d_diag_part.alloc(27 * std::max<size_t>(st.lc_cams.size(), 1)));
  HK(d_cam_part.alloc(6 * std::max<size_t>(st.lc_cams.size(), 1)));
  HK(d_chunk_chi.alloc(std::max(st.nlc, 1)));
  HK(d_bschur.alloc(6 * static_cast<size_t>(st.Cf)));
  HK(d_Sdiag.alloc(36 * static_cast<size_t>(std::max(st.Cf, 1))));
  HK(d_res.alloc(6 * static_cast<size_t>(st.Cf)));
  HK(d_xp.alloc(6 * static_cast<size_t>(st.Cf)));
  HK(d_red.alloc(std::max(RED_BLOCKS, CHI2_BLOCKS)));
  HK(d_scal.alloc(8));
  HK(d_dq_part.alloc(std::max(st.Cf, 1)));
  HK(d_dn_part.alloc(std::max(st.Cf, 1)));
  HK(d_fpart.alloc(8 * static_cast<size_t>(pcg_fused_blocks(std::max(st.Cf, 1)))));
  HK(d_err.alloc(4));
  HK(d_sc.alloc(1));
  HK(hipMemsetAsync(d_b.p, 0, nx * sizeof(double), stream));
  HK(hipMemsetAsync(d_x.p, 0, nx * sizeof(double), stream));
  g = DevGraph{};
  g.C = st.C; g.P = st.P; g.Cf = st.Cf; g.Pf = st.Pf; g.npos = st.npos; g.ns = st.ns; g.nlc = st.nlc;
  g.lc_maxc = std::max(st.lc_maxc, 1);
  g.nS = st.nS; g.nF = st.nF;
  g.camH = d_camH.p; g.ptH = d_ptH.p;
  g.pos_pt = d_pos_pt.p; g.pos_h = d_pos_h.p; g.pos_ptr = d_pos_ptr.p; g.pos_nfree = d_pos_nfree.p;
  g.slot_cam = d_slot_cam.p; g.slot_pos = d_slot_pos.p; g.slot_camh = d_slot_camh.p; g.slot_ph = d_slot_ph.p;
  g.nq = st.nq; g.slot_q = d_slot_q.p; g.cam_qptr = d_cam_qptr.p; g.q_ph = d_q_ph.p;
  g.q_pt = d_q_pt.p; g.cam_raw = d_cam_raw.p;
  g.slot_rec = d_slot_rec.p;
  g.lc_pos0 = d_lc_pos0.p; g.lc_cam_ptr = d_lc_cam_ptr.p; g.lc_cams = d_lc_cams.p;
  g.camp_ptr = d_camp_ptr.p; g.camp_idx = d_camp_idx.p;
  g.S_row = d_S_row.p; g.S_col = d_S_col.p; g.S_diag = d_S_diag.p; g.up_pos = d_up_pos.p; g.up_posT = d_up_posT.p;
  g.F_ptr = d_F_ptr.p; g.F_col = d_F_col.p; g.F_diag = d_F_diag.p;
  d_S.release();  // information back to identity
  g.slot_S = nullptr;
  HK(hipStreamSynchronize(stream));
  return G2OHIP_OK;
}

// which = 0: the plan of the FP32 blocks, 1: of the FP64 blocks
int g2ohip_ctx::ensure_win(int which) {
  if (win_ready[which]) return G2OHIP_OK;
  WinPlan& wp = st.win[which];
  if (wp.win_obs0.empty()) {
    const int obs = which == 0 || max_point_obs(st) > SCHUR_WIN_OBS64 ? SCHUR_WIN_OBS32 : SCHUR_WIN_OBS64;
    std::string msg = build_win_plan(st, wp, obs, SCHUR_WIN_PCH);
    if (!msg.empty()) return fail(G2OHIP_ERR_UNSUPPORTED, msg);
  }
  if (device < 0) return G2OHIP_OK;
  WinPlanBufs& B = win_bufs[which];
  DevWinPlan& dwin = dwins[which];
  HK(B.win_obs0.upload(wp.win_obs0, stream));
  HK(B.win_q.upload(wp.win_q, stream));
  HK(B.win_grp0.upload(wp.win_grp0, stream));
  HK(B.grp_len.upload(wp.grp_len, stream));
  HK(B.grp_row0.upload(wp.grp_row0, stream));
  HK(B.grp_slot.upload(wp.grp_slot, stream));
  HK(B.od_blk.upload(wp.od_blk, stream));
  HK(B.od_ptr.upload(wp.od_ptr, stream));
  HK(B.pair_rows.upload(wp.pair_rows, stream));
  dwin.nwin = static_cast<int>(wp.win_obs0.size()) - 1;
  dwin.nod = static_cast<int>(wp.od_blk.size());
  dwin.obs_max = wp.obs_max;
  dwin.win_obs0 = B.win_obs0.p; dwin.win_q = B.win_q.p; dwin.win_grp0 = B.win_grp0.p;
  dwin.grp_len = B.grp_len.p; dwin.grp_row0 = B.grp_row0.p; dwin.grp_slot = B.grp_slot.p;
  dwin.od_blk = B.od_blk.p; dwin.od_ptr = B.od_ptr.p; dwin.pair_rows = B.pair_rows.p;
  HK(hipStreamSynchronize(stream));
  win_ready[which] = true;
  return G2OHIP_OK;
}

// Window plan or camera rows for the off-diagonal Schur blocks. The window
// plan stages one partial block per (window, block) in HBM: cheap when
// tracks are short (a window's pairs hit few blocks), but a track of n
// cameras adds n (n - 1) / 2 partials of its own, and its host-built pair
// list grows the same way. The row kernel stages nothing. Rows are chosen
// when a point has more free observations than a window holds, or when
// tracks longer than SCHUR_LONG_TRACK cameras make most of the pair
// products.
bool g2ohip_ctx::use_schur_rows() {
  if (schur_path == 1) return false;
  if (schur_path == 2) return true;
  if (schur_rows_chosen < 0) {
    long long pairs = 0, long_pairs = 0;
    for (int p = 0; p < st.npos; ++p) {
      const long long n = st.pos_h[p] >= 0 ? st.pos_nfree[p] : 0;
      pairs += n * (n - 1) / 2;
      if (n > SCHUR_LONG_TRACK) long_pairs += n * (n - 1) / 2;
    }
    schur_rows_chosen = (max_point_obs(st) > SCHUR_WIN_OBS32 || 2 * long_pairs > pairs) ? 1 : 0;
  }
  return schur_rows_chosen == 1;
}

// Row items of the camera-row Schur kernel: each free camera's off-diagonal
// upper blocks in runs of at most SCHUR_ROW_CAP.
int g2ohip_ctx::ensure_rows() {
  if (row_nitems > 0 || st.nS == st.Cf) return G2OHIP_OK;
  std::vector<int4> items;
  int mx = 0;
  for (int a = 0; a < st.Cf; ++a) {
    const int b0 = st.S_rowptr[a] + 1, b1 = st.S_rowptr[a + 1];
    for (int k = b0; k < b1; k += SCHUR_ROW_CAP) {
      const int e = std::min(b1, k + SCHUR_ROW_CAP);
      items.push_back(make_int4(a, k, e, (k == b0 && e == b1) ? 1 : 0));
      mx = std::max(mx, e - k);
    }
  }
  HK(d_row_items.upload(items, stream));
  HK(hipStreamSynchronize(stream));
  row_nitems = static_cast<int>(items.size());
  row_max_blk = mx;
  return G2OHIP_OK;
}

template <typename HT>
int g2ohip_ctx::schur_offdiag(PrecBufs<HT>& B) {
  const int which = std::is_same<HT, float>::value ? 0 : 1;
  if (use_schur_rows()) {
    CK(ensure_rows());
    ProfScope pr(this, G2OHIP_K_SCHUR_WINDOWS);
    HK(launch_schur_rows<HT>(g, d_row_items.p, row_nitems, row_max_blk, d_q_slot.p, B.Hpl.p, d_linv.p, B.F_val.p,
                             stream));
    return G2OHIP_OK;
  }
  CK(ensure_win(which));
  HK(B.part.alloc(36 * static_cast<size_t>(std::max(st.win[which].od_ptr.back(), 1))));
  {
    ProfScope pr(this, G2OHIP_K_SCHUR_WINDOWS);
    HK(launch_schur_win<HT>(g, dwins[which], B.Hpl.p, d_linv.p, B.part.p, stream));
  }
  HK(launch_schur_offdiag<HT>(g, dwins[which], B.part.p, B.F_val.p, stream));
  return G2OHIP_OK;
}
template int g2ohip_ctx::schur_offdiag<float>(PrecBufs<float>&);
template int g2ohip_ctx::schur_offdiag<double>(PrecBufs<double>&);

int g2ohip_ctx::set_measurements(const double* meas, const double* w, const double* intr, double huber) {
  if (mode != 0) return fail(G2OHIP_ERR_INVALID, "6-DoF bundle-adjustment call on a pose-graph or BAL context");
  if (!have_structure) return fail(G2OHIP_ERR_INVALID, "set_structure first");
  if (!meas || !intr) return fail(G2OHIP_ERR_INVALID, "measurements and intrinsics are required");
  CK(require_device());
  std::vector<double> m(2 * static_cast<size_t>(st.ns)), ww(st.ns);
  for (int s = 0; s < st.ns; ++s) {
    int k = st.slot_edge[s];
    m[2 * static_cast<size_t>(s)] = meas[2 * static_cast<size_t>(k)];
    m[2 * static_cast<size_t>(s) + 1] = meas[2 * static_cast<size_t>(k) + 1];
    ww[s] = w ? w[k] : 1.0;
  }
  std::vector<double> in(intr, intr + 4 * static_cast<size_t>(st.C));
  HK(d_meas.upload(m, stream));
  HK(d_w.upload(ww, stream));
  HK(d_intr.upload(in, stream));
  HK(hipStreamSynchronize(stream));
  g.slot_meas = d_meas.p;
  g.slot_w = d_w.p;
  g.intr = d_intr.p;
  g.huber = huber;
  have_meas = true;
  have_system = have_lin = false;
  return G2OHIP_OK;
}

// BaseEdge::setInformation for every projection edge (base_edge.h:165-180):
// omega[E*3] = [O00 O01 O11], the upper triangle of each edge's 2x2
// information matrix, as the g2o file format stores it. The kernels take the
// factor S (Omega = S^T S, upper triangular) per slot; all-identity input, or
// omega == nullptr, drops it (the unwhitened fast path).
int g2ohip_ctx::set_information(const double* omega) {
  if (mode != 0) return fail(G2OHIP_ERR_INVALID, "6-DoF bundle-adjustment call on a pose-graph or BAL context");
  if (!have_structure) return fail(G2OHIP_ERR_INVALID, "set_structure first");
  CK(require_device());
  std::vector<double> S;
  bool identity = true;
  if (omega) {
    S.resize(3 * static_cast<size_t>(st.ns));
    for (int s = 0; s < st.ns; ++s) {
      const double* o = omega + 3 * static_cast<size_t>(st.slot_edge[s]);
      if (!(o[0] == 1.0 && o[1] == 0.0 && o[2] == 1.0)) identity = false;
      const double s0 = std::sqrt(o[0]);
      const double s1 = o[1] / s0;
      const double d = o[2] - s1 * s1;
      if (!(o[0] > 0.0) || !(d > 0.0) || !std::isfinite(o[0]) || !std::isfinite(o[1]) || !std::isfinite(o[2]))
        return fail(G2OHIP_ERR_INVALID, "information matrix of edge " + std::to_string(st.slot_edge[s]) +
                                            " is not positive definite");
      S[3 * static_cast<size_t>(s)] = s0;
      S[3 * static_cast<size_t>(s) + 1] = s1;
      S[3 * static_cast<size_t>(s) + 2] = std::sqrt(d);
    }
  }
  if (identity) {
    d_S.release();
    g.slot_S = nullptr;
  } else {
    HK(d_S.upload(S, stream));
    HK(hipStreamSynchronize(stream));
    g.slot_S = d_S.p;
  }
  have_system = have_lin = false;
  return G2OHIP_OK;
}

int g2ohip_ctx::set_state(const double* cams, const double* pts) {
  if (mode != 0) return fail(G2OHIP_ERR_INVALID, "6-DoF bundle-adjustment call on a pose-graph or BAL context");
  if (!have_structure) return fail(G2OHIP_ERR_INVALID, "set_structure first");
  CK(require_device());
  std::vector<double> c(cams, cams + 7 * static_cast<size_t>(st.C));
  for (int i = 0; i < st.C; ++i) {
    double* q = &c[7 * static_cast<size_t>(i) + 3];
    if (q[3] < 0) { q[0] = -q[0]; q[1] = -q[1]; q[2] = -q[2]; q[3] = -q[3]; }
    double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    if (!(n > 0)) return fail(G2OHIP_ERR_INVALID, "zero quaternion");
    for (int k = 0; k < 4; ++k) q[k] /= n;
  }
  // the device keeps the points in position order (st.pt_dev)
  std::vector<double> p(3 * static_cast<size_t>(st.P));
  for (int i = 0; i < st.P; ++i)
    for (int k = 0; k < 3; ++k) p[3 * static_cast<size_t>(st.pt_dev[i]) + k] = pts[3 * static_cast<size_t>(i) + k];
  HK(d_cams.upload(c, stream));
  if (st.P > 0) HK(d_pts.upload(p, stream));
  HK(hipStreamSynchronize(stream));
  have_state = true;
  have_system = false;
  stack_depth = 0;
  return G2OHIP_OK;
}

int g2ohip_ctx::get_state(double* cams, double* pts) {
  if (mode != 0) return fail(G2OHIP_ERR_INVALID, "6-DoF bundle-adjustment call on a pose-graph or BAL context");
  if (!have_state) return fail(G2OHIP_ERR_INVALID, "no state");
  if (cams) HK(hipMemcpyAsync(cams, d_cams.p, 7 * sizeof(double) * st.C, hipMemcpyDeviceToHost, stream));
  std::vector<double> p(pts ? 3 * static_cast<size_t>(st.P) : 0);
  if (pts && st.P > 0) HK(hipMemcpyAsync(p.data(), d_pts.p, 3 * sizeof(double) * st.P, hipMemcpyDeviceToHost, stream));
  HK(hipStreamSynchronize(stream));
  if (pts)
    for (int i = 0; i < st.P; ++i)
      for (int k = 0; k < 3; ++k) pts[3 * static_cast<size_t>(i) + k] = p[3 * static_cast<size_t>(st.pt_dev[i]) + k];
  return G2OHIP_OK;
}

// ------------------------------------------------------------- residuals
int g2ohip_ctx::compute_active_errors(double* chi2, double* err_out) {
  if (mode == 2) {
    if (!(have_structure && have_meas && have_state)) return fail(G2OHIP_ERR_INVALID, "structure/measurements/state missing");
    return bal_chi2(chi2, err_out);
  }
  if (mode == 1) {
    if (!(have_structure && have_meas && have_state)) return fail(G2OHIP_ERR_INVALID, "structure/measurements/state missing");
    return pg_chi2(chi2, err_out);
  }
  if (!(have_structure && have_meas && have_state)) return fail(G2OHIP_ERR_INVALID, "structure/measurements/state missing");
  HK(launch_cam_table(g, d_cams.p, d_camRt.p, stream));
  DBuf<double> eslots;
  if (err_out) HK(eslots.alloc(2 * static_cast<size_t>(st.ns)));
  {
    ProfScope ps(this, G2OHIP_K_CHI2);
    HK(launch_chi2(g, d_camRt.p, d_pts.p, d_red.p, d_scal.p, err_out ? eslots.p : nullptr, stream));
  }
  CK(read_scalar(d_scal.p, chi2));
  if (err_out) {
    std::vector<double> es(2 * static_cast<size_t>(st.ns));
    HK(hipMemcpy(es.data(), eslots.p, es.size() * sizeof(double), hipMemcpyDeviceToHost));
    std::fill(err_out, err_out + 2 * static_cast<size_t>(st.E), 0.0);
    for (int s = 0; s < st.ns; ++s) {
      int k = st.slot_edge[s];
      err_out[2 * static_cast<size_t>(k)] = es[2 * static_cast<size_t>(s)];
      err_out[2 * static_cast<size_t>(k) + 1] = es[2 * static_cast<size_t>(s) + 1];
    }
  }
  return G2OHIP_OK;
}

int g2ohip_ctx::build_system(bool read_chi) {
  if (mode == 2) {
    if (!(have_structure && have_meas && have_state)) return fail(G2OHIP_ERR_INVALID, "structure/measurements/state missing");
    return bal_build();
  }
  if (mode == 1) {
    if (!(have_structure && have_meas && have_state)) return fail(G2OHIP_ERR_INVALID, "structure/measurements/state missing");
    return pg_build();
  }
  if (!(have_structure && have_meas && have_state)) return fail(G2OHIP_ERR_INVALID, "structure/measurements/state missing");
  HK(launch_cam_table(g, d_cams.p, d_camRt.p, stream));
  const int prec = (opt.precision == G2OHIP_FP32 && !fp32_off) ? G2OHIP_FP32 : G2OHIP_FP64;
  if (prec == G2OHIP_FP32) {
    HK(f32.Hpl.alloc(18 * static_cast<size_t>(std::max(st.nq, 1))));
    ProfScope ps(this, G2OHIP_K_LINEARIZE);
    HK(launch_linearize<float>(g, d_camRt.p, d_pts.p, f32.Hpl.p, d_Hll.p, d_b.p, d_hpp_part.p, d_chunk_chi.p, stream));
  } else {
    HK(f64.Hpl.alloc(18 * static_cast<size_t>(std::max(st.nq, 1))));
    ProfScope ps(this, G2OHIP_K_LINEARIZE);
    HK(launch_linearize<double>(g, d_camRt.p, d_pts.p, f64.Hpl.p, d_Hll.p, d_b.p, d_hpp_part.p, d_chunk_chi.p,
                                stream));
  }
  HK(launch_reduce_hpp(g, d_hpp_part.p, d_Hpp.p, d_b.p, stream));
  HK(launch_sum(d_chunk_chi.p, st.nlc, d_scal.p + 1, stream));
  if (read_chi) CK(read_scalar(d_scal.p + 1, &chi2_lin));
  have_system = have_lin = true;
  have_exact = false;
  system_prec = prec;
  have_solution = false;
  return G2OHIP_OK;
}

int g2ohip_ctx::lambda_init(double* out) {
  if (mode == 2) {
    if (!have_system) return fail(G2OHIP_ERR_INVALID, "build_system first");
    return bal_lambda_init(out);
  }
  if (mode == 1) {
    if (!have_system) return fail(G2OHIP_ERR_INVALID, "build_system first");
    return pg_lambda_init(out);
  }
  if (!have_system) return fail(G2OHIP_ERR_INVALID, "build_system first");
  if (opt.user_lambda_init > 0) { *out = opt.user_lambda_init; return G2OHIP_OK; }
  HK(launch_max_diag(g, d_Hpp.p, d_Hll.p, d_red.p, d_scal.p + 2, stream));
  double m;
  CK(read_scalar(d_scal.p + 2, &m));
  *out = opt.tau * m;
  return G2OHIP_OK;
}

int g2ohip_ctx::compute_scale(double* out) {
  if (!have_solution) return fail(G2OHIP_ERR_INVALID, "solve first");
  const int n = static_cast<int>(system_size());
  HK(d_red.alloc(RED_BLOCKS));  // a no-op once set_structure sized it
  HK(launch_scale(n, d_x.p, d_b.p, lambda, d_red.p, d_scal.p + 3, stream));
  return read_scalar(d_scal.p + 3, out);
}

// ------------------------------------------------------------------- PCG

template <typename HT>
int g2ohip_ctx::run_pcg(PrecBufs<HT>& B, const double* rhs, const PcgControl& ctl0, bool implicit, int max_iter,
                        PcgRun* out) {
  const int nb = st.Cf;
  const PcgControl ctl = with_progress(ctl0);
  bool prog_failed = false;
  const int n = 6 * nb;
  HK(B.x.alloc(n)); HK(B.r.alloc(n)); HK(B.d.alloc(n)); HK(B.d2.alloc(n)); HK(B.q.alloc(n)); HK(B.s.alloc(n));
  double* const dirs[2] = {B.d.p, B.d2.p};
  ProfScope ps(this, G2OHIP_K_PCG);
  if (implicit) {
    auto op = [&](int it, int last, const double* d_old, double* d) -> hipError_t {
      hipError_t e = launch_pcg_dir<6>(nb, B.s.p, d_old, d, d_dn_part.p, d_sc.p, ctl, it, last, stream);
      if (e != hipSuccess || last) return e;
      // FP32: each camera's own block from the FP64 diagonal (as the explicit
      // FP32 operator keeps it), the FP32 passes carry the cross terms only
      constexpr bool f32 = std::is_same<HT, float>::value;
      e = launch_implicit_pass<HT>(g, B.Hpl.p, B.Dinv.p, d, nullptr, d_cam_part.p, d_sc.p, it, stream, f32,
                                   f32 ? d_linv.p : nullptr);
      if (e != hipSuccess) return e;
      return launch_implicit_reduce<HT>(g, d_cam_part.p, d_Hpp.p, lambda, d, B.q.p, d_dq_part.p, d_sc.p, it, stream,
                                        f32 ? d_Sdiag.p : nullptr);
    };
    HK((pcg_drive<HT, 6>(stream, h_pinned, progress(), nb, rhs, B.Jinv.p, B.x.p, B.r.p, B.s.p, dirs, B.q.p,
                         d_dn_part.p, d_dq_part.p, implicit_reduce_blocks(nb), d_sc.p, max_iter, op, out,
                         &prog_failed, ctl.abort)));
  } else if (pcg_fused && pcg_fused_fits(nb)) {
    HK(B.s2.alloc(n)); HK(B.w.alloc(n)); HK(B.w2.alloc(n));
    double* const s2[2] = {B.s.p, B.s2.p};
    double* const w2[2] = {B.w.p, B.w2.p};
    auto init = [&]() { return launch_pcg_init<HT, 6>(nb, rhs, B.Jinv.p, B.x.p, B.r.p, B.s.p, d_dn_part.p, d_sc.p,
                                                      stream); };
    auto step = [&](int it, int last) -> hipError_t {
      const int pb = prof_begin();
      const hipError_t e = launch_pcg_fused<HT, 6>(nb, d_F_head.p, d_F_col.p, B.F_val.p,
                                                   std::is_same<HT, float>::value ? d_Sdiag.p : nullptr, B.Jinv.p,
                                                   B.x.p, B.r.p, B.q.p, s2, w2, dirs, d_dn_part.p, d_fpart.p,
                                                   d_sc.p, ctl, it, last, stream, st.nF);
      prof_end(G2OHIP_K_PCG_SPMV, pb);
      return e;
    };
    HK(pcg_loop(stream, h_pinned, progress(), d_sc.p, max_iter, init, step, out, &prog_failed, ctl.abort));
  } else {
    auto op = [&](int it, int last, const double* d_old, double* d) -> hipError_t {
      const int pb = prof_begin();
      const hipError_t e = launch_pcg_dir_spmv<HT, 6>(nb, d_F_ptr.p, d_F_col.p, B.F_val.p, B.s.p, d_old, d, B.q.p,
                                                      d_dn_part.p, d_dq_part.p, d_sc.p, ctl, it, last, stream,
                                                      std::is_same<HT, float>::value ? d_Sdiag.p : nullptr, st.nF);
      prof_end(G2OHIP_K_PCG_SPMV, pb);
      return e;
    };
    HK((pcg_drive<HT, 6>(stream, h_pinned, progress(), nb, rhs, B.Jinv.p, B.x.p, B.r.p, B.s.p, dirs, B.q.p,
                         d_dn_part.p, d_dq_part.p, pcg_spmv_blocks(nb), d_sc.p, max_iter, op, out, &prog_failed,
                         ctl.abort)));
  }
  if (prog_failed) prog_ok = false;
  return G2OHIP_OK;
}

// BlockSolver::solve with Schur (block_solver.hpp:331-477).
// defer_err: the final error word (back-substitution) is left for the
// caller to read with its next scalars (d_err).
// Diagonal blocks, block-Jacobi inverses and (explicit) the reduced
// right-hand side in FP64, re-linearised from the state point-major (FP32
// blocks are too coarse for the cancellation in Hpp - sum V Dinv V^T), on
// stream `st`. FP32 blocks: the diagonal blocks stay FP64 (d_Sdiag) for the
// SpMV and for the implicit operator (the FP32 implicit path takes its
// reduced right-hand side from here too: b_p - sum V Dinv b_l cancels towards
// convergence, and an FP32 Hpl would leave it mostly rounding error).
template <typename HT>
int g2ohip_ctx::schur_diag(PrecBufs<HT>& B, bool implicit, hipStream_t st) {
  constexpr bool f32 = std::is_same<HT, float>::value;
  double* sdiag = f32 ? d_Sdiag.p : nullptr;
  const bool rhs_here = !implicit || f32;
  HK(launch_schur_diag_pm<HT>(g, d_camRt.p, d_pts.p, d_Hll.p, d_Hpp.p, rhs_here ? d_b.p : nullptr, lambda,
                              d_diag_part.p, implicit ? nullptr : B.F_val.p, B.Jinv.p,
                              rhs_here ? d_bschur.p : nullptr, sdiag, d_err.p, st));
  return G2OHIP_OK;
}

template <typename HT>
int g2ohip_ctx::solve_impl(int* ok, bool defer_err) {
  const int which = std::is_same<HT, float>::value ? 0 : 1;
  PrecBufs<HT>& B = bufs_of<HT>(*this);
  const int nb = st.Cf;
  const bool implicit = opt.schur_mode == G2OHIP_SCHUR_IMPLICIT;
  *ok = 1;
  // the FP64 matrix-free passes below re-linearise at the current state: the
  // camera table must follow it (a rejected LM trial is popped after its chi2
  // pass left the table at the trial state)
  HK(launch_cam_table(g, d_cams.p, d_camRt.p, stream, d_err.p));  // and d_err[0..3] = 0
  HK(B.Jinv.alloc(36 * static_cast<size_t>(nb)));
  {
    ProfScope ps(this, G2OHIP_K_SCHUR);
    // explicit operator: the FP64 diagonal pass (below) shares no data with
    // the off-diagonal blocks and runs beside them on the side stream
    const bool fork = !implicit && side != nullptr;
    if (fork) {
      HK(hipEventRecord(ev_fork, stream));
      HK(hipStreamWaitEvent(side, ev_fork, 0));
    }
    if (implicit) {
      HK(B.Dinv.alloc(9 * static_cast<size_t>(std::max(st.Pf, 1))));
      HK(launch_point_dinv<HT>(g, d_Hll.p, lambda, B.Dinv.p, d_err.p, stream));
      if (std::is_same<HT, float>::value) {  // the FP32 operator's FP64 point factors
        HK(d_linv.alloc(6 * static_cast<size_t>(std::max(st.Pf, 1))));
        HK(launch_point_linv(g, d_Hll.p, lambda, d_linv.p, d_err.p, stream));
      }
    } else {
      HK(B.F_val.alloc(36 * static_cast<size_t>(std::max(st.nF, 1))));
      HK(d_linv.alloc(6 * static_cast<size_t>(std::max(st.Pf, 1))));
      if (fork) CK(schur_diag<HT>(B, implicit, side));
      HK(launch_point_linv(g, d_Hll.p, lambda, d_linv.p, d_err.p, stream));
      CK(schur_offdiag<HT>(B));
    }
    if (!fork) CK(schur_diag<HT>(B, implicit, stream));
    if (fork) {
      HK(hipEventRecord(ev_join, side));
      HK(hipStreamWaitEvent(stream, ev_join, 0));
    }
  }
  if (implicit && !std::is_same<HT, float>::value) {
    // matrix-free reduced operator: right-hand side by one pass
    HK(launch_implicit_pass<HT>(g, B.Hpl.p, B.Dinv.p, nullptr, d_b.p, d_cam_part.p, nullptr, 0, stream));
    HK(launch_implicit_rhs_reduce(g, d_cam_part.p, d_b.p, d_bschur.p, stream));
  }
  // a failed point or camera block (d_err) stops the PCG at its first head
  // and is read back with the PCG's own scalars: no separate round trip
  const bool refine = (which == 0) && opt.refine_steps > 0;
  refine_stalled = false;
  const int max_iter = opt.pcg_max_iter < 0 ? 6 * nb : opt.pcg_max_iter;
  const double* xp64 = d_x.p;
  double* xl64 = d_x.p + 6 * static_cast<size_t>(nb);
  // reduced system on the Schur right-hand side, with the residual memory of
  // the previous solve (linear_solver_pcg.hpp:128-136)
  bool xl_current = false;  // x_l matches the current x_p
  PcgControl ctl_main{opt.pcg_tolerance, pcg_residual_mem, -1.0};
  ctl_main.abort = d_err.p;
  PcgRun pr;
  CK(run_pcg<HT>(B, d_bschur.p, ctl_main, implicit, max_iter, &pr));
  if (pr.err_word) { *ok = 0; return G2OHIP_OK; }
  int total_it = pr.iters;
  double dn_final = pr.dn_final;
  HK(launch_to_double<double>(6 * nb, B.x.p, d_x.p, 0.0, stream));
  int steps_run = 0;
  if (refine) {
    // FP64 iterative refinement: r = g - S x evaluated in FP64 matrix-free from
    // the state (no FP32 block enters it), then an FP32 PCG correction S dx = r
    // that stops on the first solve's absolute threshold d0. Corrections are
    // repeated until one finds the FP64 residual already at the threshold
    // (zero iterations), at most refine_steps times: the refined x then meets
    // the reference stop rule measured on the FP64 residual, not only on the
    // FP32 recurrence. A correction runs at most twice the first solve's
    // iterations plus 10.
    for (int step = 0; step < opt.refine_steps; ++step) {
      ++steps_run;
      xl_current = true;  // the residual pass writes x_l for the current x_p
      {
        ProfScope ps(this, G2OHIP_K_REFINE);
        HK(launch_refine(g, d_camRt.p, d_pts.p, lambda, xp64, xl64, d_cam_part.p, d_err.p, stream));
      }
      HK(launch_refine_reduce(g, d_cam_part.p, xp64, lambda, d_res.p, stream));
      PcgControl ctl_corr{opt.pcg_tolerance, -1.0, pr.d0};
      ctl_corr.abort = d_err.p;
      PcgRun pc;
      CK(run_pcg<HT>(B, d_res.p, ctl_corr, implicit, std::min(max_iter, 2 * pr.iters + 10), &pc));
      if (pc.err_word) { *ok = 0; return G2OHIP_OK; }
      total_it += pc.iters;
      // the FP32-rounded system stops resolving the FP64 problem: its
      // solution misses the FP64 threshold by more than 10x, or the
      // correction stops above it
      if (pc.dn0 > 10.0 * pr.d0 || pc.dn_final > pr.d0 || pc.breakdown) refine_stalled = true;
      // the residual memory stays the main solve's: half the recurrence's
      // r.M^-1 r at its exit, as the FP64 solver leaves it after the same
      // iterations (a correction that runs ends far below it and would
      // tighten the next solve's threshold beyond the FP64 path's)
      if (pc.iters == 0) break;  // the FP64 residual meets the threshold
      HK(launch_to_double<double>(6 * nb, B.x.p, d_x.p, 1.0, stream));
      xl_current = false;
    }
  }
  pcg_residual_mem = 0.5 * dn_final;
  // back-substitution x_l = Dinv (b_l - Hpl^T x_p), re-linearised in FP64
  // (block_solver.hpp:452-476); already done by the last residual pass if
  // no correction changed x_p after it
  if (!xl_current) {
    ProfScope ps(this, G2OHIP_K_REFINE);
    HK(launch_refine(g, d_camRt.p, d_pts.p, lambda, xp64, xl64, nullptr, d_err.p, stream));
  }
  last_info[1] = steps_run;
  if (!defer_err) {
    int err = 0;
    CK(read_err(&err));
    if (err) *ok = 0;
  }
  last_info[0] = total_it;
  last_info[2] = pcg_residual_mem;
  last_solve_prec = which == 0 ? G2OHIP_FP32 : G2OHIP_FP64;
  have_solution = true;
  return G2OHIP_OK;
}

int g2ohip_ctx::solve(int* ok, bool defer_err) {
  if (mode == 2) {
    if (!have_system) return fail(G2OHIP_ERR_INVALID, "build_system first");
    return bal_solve(ok);
  }
  if (mode == 1) {
    if (!have_system) return fail(G2OHIP_ERR_INVALID, "build_system first");
    return pg_solve(ok);
  }
  if (!have_system) return fail(G2OHIP_ERR_INVALID, "build_system first");
  last_info[3] = 0;
  if (system_prec != G2OHIP_FP32) return solve_impl<double>(ok, defer_err);
  const double mem = pcg_residual_mem;
  CK(solve_impl<float>(ok, defer_err));
  if (!refine_stalled) return G2OHIP_OK;
  // FP64 fallback at the same state (the linearisation point: the LM solves
  // before it updates)
  fp32_off = true;
  pcg_residual_mem = mem;
  CK(build_system());
  CK(solve_impl<double>(ok, defer_err));
  last_info[3] = 1;
  return G2OHIP_OK;
}

// BlockSolverBase::multiplyHessian (block_solver.h:149-151): the pose rows of
// dest = Hpp src, with Hpp the undamped pose-pose blocks of the last
// build_system (block diagonal for bundle adjustment, the whole Hessian of a
// pose graph).
int g2ohip_ctx::multiply_hessian(const double* src, double* dest) {
  if (!have_system) return fail(G2OHIP_ERR_INVALID, "build_system first");
  if (mode == 2) return fail(G2OHIP_ERR_UNSUPPORTED, "multiplyHessian is not provided for BAL problems");
  const int nb = mode == 1 ? pgst.Nf : st.Cf;
  const size_t n = 6 * static_cast<size_t>(nb);
  DBuf<double> dx, dy;
  std::vector<double> hx(src, src + n);
  HK(dx.upload(hx, stream));
  HK(dy.alloc(n));
  if (mode == 1) HK(launch_block_mv6(nb, d_pg_F_ptr.p, d_pg_F_col.p, d_pg_F_val.p, dx.p, dy.p, stream));
  else HK(launch_block_mv6(nb, nullptr, nullptr, d_Hpp.p, dx.p, dy.p, stream));
  HK(hipMemcpyAsync(dest, dy.p, n * sizeof(double), hipMemcpyDeviceToHost, stream));
  HK(hipStreamSynchronize(stream));
  return G2OHIP_OK;
}

// --------------------------------------------------------- update / stack
int g2ohip_ctx::update() {
  if (mode == 2) {
    if (!have_solution) return fail(G2OHIP_ERR_INVALID, "solve first");
    return bal_update();
  }
  if (mode == 1) {
    if (!have_solution) return fail(G2OHIP_ERR_INVALID, "solve first");
    return pg_update();
  }
  if (!have_solution) return fail(G2OHIP_ERR_INVALID, "solve first");
  {
    ProfScope ps(this, G2OHIP_K_UPDATE);
    HK(launch_update(g, d_x.p, d_cams.p, d_pts.p, stream));
  }
  HK(hipStreamSynchronize(stream));
  return G2OHIP_OK;
}

int g2ohip_ctx::push() {
  if (!have_state) return fail(G2OHIP_ERR_INVALID, "no state");
  if (stack_depth >= 4) return fail(G2OHIP_ERR_INVALID, "state stack is full (depth 4)");
  if (static_cast<int>(stack_c.size()) <= stack_depth) {
    stack_c.push_back(new DBuf<double>());
    stack_p.push_back(new DBuf<double>());
  }
  DBuf<double>& c = *stack_c[stack_depth];
  DBuf<double>& p = *stack_p[stack_depth];
  HK(c.alloc(cams_len()));
  HK(p.alloc(std::max<size_t>(pts_len(), 1)));
  HK(launch_copy(c.p, d_cams.p, sizeof(double) * cams_len(), stream));
  if (pts_len()) HK(launch_copy(p.p, d_pts.p, sizeof(double) * pts_len(), stream));
  stack_depth++;
  return G2OHIP_OK;
}

int g2ohip_ctx::pop() {
  if (stack_depth <= 0) return fail(G2OHIP_ERR_INVALID, "state stack is empty");
  stack_depth--;
  HK(launch_copy(d_cams.p, stack_c[stack_depth]->p, sizeof(double) * cams_len(), stream));
  if (pts_len()) HK(launch_copy(d_pts.p, stack_p[stack_depth]->p, sizeof(double) * pts_len(), stream));
  return G2OHIP_OK;
}

int g2ohip_ctx::discard_top() {
  if (stack_depth <= 0) return fail(G2OHIP_ERR_INVALID, "state stack is empty");
  stack_depth--;
  return G2OHIP_OK;
}

// ------------------------------------------------------------------- LM
// SparseOptimizer::optimize (sparse_optimizer.cpp:393-455) driving
// OptimizationAlgorithmLevenberg::solve (optimization_algorithm_levenberg.cpp:60-153).
int g2ohip_ctx::optimize(int iterations, g2ohip_iter_stats* stats, int* done) {
  if (!(have_structure && have_meas && have_state)) return fail(G2OHIP_ERR_INVALID, "structure/measurements/state missing");
  if (algorithm == G2OHIP_GAUSS_NEWTON) return optimize_gauss_newton(iterations, stats, done);
  if (algorithm == G2OHIP_DOGLEG) return optimize_dogleg(iterations, stats, done);
  *done = 0;
  pcg_residual_mem = -1.0;  // LinearSolverPCG::init() via algorithm->init()
  fp32_off = false;
  trials.clear();
  double ni = 2.0;
  for (int i = 0; i < iterations; ++i) {
    if (force_stop && *force_stop) break;
    g2ohip_iter_stats s{};
    const double t_it = now_s();
    double t = now_s();
    // computeActiveErrors + activeRobustChi2 + buildSystem, fused: the
    // linearisation pass returns the chi2 of the same state (read back with
    // the first trial's scalars unless lambda_init reads first).
    CK(build_system(i == 0 || mode != 0));
    double currentChi0 = chi2_lin;
    bool have_chi0 = i == 0 || mode != 0;
    s.time_quadratic_form = now_s() - t;
    if (i == 0) {
      CK(lambda_init(&lambda));
      ni = 2.0;
    }
    double currentChi = currentChi0;
    double rho = 0;
    int q = 0;
    do {
      CK(push());
      t = now_s();
      int ok = 0;
      CK(solve(&ok, true));
      s.time_linear_solver += now_s() - t;
      s.iterations_linear_solver += static_cast<int>(last_info[0]);
      s.refine_steps += static_cast<int>(last_info[1]);
      s.fp64_fallback |= static_cast<int>(last_info[3]);
      double tempChi = std::numeric_limits<double>::max();
      double scale = 1.0;
      // update, new chi2 and computeScale enqueued together; one read of
      // chi2, the linearisation chi2, the scale and the error word
      t = now_s();
      if (ok && mode == 0) {
        HK(launch_scale(static_cast<int>(system_size()), d_x.p, d_b.p, lambda, d_red.p, d_scal.p + 3, stream));
        {
          ProfScope ps(this, G2OHIP_K_UPDATE);
          HK(launch_update(g, d_x.p, d_cams.p, d_pts.p, stream));
        }
        HK(launch_cam_table(g, d_cams.p, d_camRt.p, stream));
        ProfScope ps(this, G2OHIP_K_CHI2);
        HK(launch_chi2(g, d_camRt.p, d_pts.p, d_red.p, d_scal.p, nullptr, stream));
      }
      double sc4[4] = {0, 0, 0, 0};
      int err = 0;
      if (mode == 0) {
        CK(read_trial_scalars(sc4, &err));
      } else if (ok) {  // pose graph: its own state-side calls
        CK(update());
        CK(compute_active_errors(&sc4[0]));
        CK(compute_scale(&sc4[3]));
      }
      s.time_residuals += now_s() - t;
      if (!have_chi0) {
        chi2_lin = currentChi0 = currentChi = sc4[1];
        have_chi0 = true;
      }
      if (err) ok = 0;
      if (ok) {
        tempChi = sc4[0];
        scale = sc4[3] + 1e-3;
        have_solution = true;
      }
      rho = (currentChi - tempChi) / scale;
      {
        const bool acc = rho > 0 && std::isfinite(tempChi) && ok;
        const double rec[G2OHIP_TRIAL_FIELDS] = {lambda, currentChi, tempChi, scale, rho, acc ? 1.0 : 0.0,
                                                 last_info[0], last_info[1], last_info[3], last_info[2]};
        trials.insert(trials.end(), rec, rec + G2OHIP_TRIAL_FIELDS);
      }
      if (rho > 0 && std::isfinite(tempChi) && ok) {
        double alpha = 1. - std::pow((2 * rho - 1), 3);
        alpha = std::min(alpha, 2. / 3.);
        double sf = std::max(1. / 3., alpha);
        lambda *= sf;
        ni = 2;
        currentChi = tempChi;
        CK(discard_top());
      } else {
        lambda *= ni;
        ni *= 2;
        CK(pop());
        if (!std::isfinite(lambda)) break;
      }
      q++;
    } while (rho < 0 && q < opt.max_trials && !(force_stop && *force_stop));
    s.chi2_before = currentChi0;
    s.chi2 = currentChi;
    s.lambda = lambda;
    s.levenberg_iterations = q;
    s.status = (q == opt.max_trials || rho == 0 || !std::isfinite(lambda)) ? 1 : 0;
    s.time_iteration = now_s() - t_it;
    if (stats) stats[i] = s;
    *done = i + 1;
    have_system = false;
    if (s.status) break;
  }
  // the last pop / copies and any asynchronous error
  HK(hipStreamSynchronize(stream));
  return G2OHIP_OK;
}

// OptimizationAlgorithmGaussNewton::solve (optimization_algorithm_gauss_newton.cpp:
// 47-90) under SparseOptimizer::optimize: computeActiveErrors, buildSystem,
// the undamped solve, update; the loop ends after the first failed solve.
int g2ohip_ctx::optimize_gauss_newton(int iterations, g2ohip_iter_stats* stats, int* done) {
  *done = 0;
  pcg_residual_mem = -1.0;  // LinearSolverPCG::init() via algorithm->init()
  fp32_off = false;
  for (int i = 0; i < iterations; ++i) {
    if (force_stop && *force_stop) break;
    g2ohip_iter_stats s{};
    const double t_it = now_s();
    double t = now_s();
    CK(build_system(true));
    s.chi2_before = chi2_lin;
    s.time_quadratic_form = now_s() - t;
    lambda = 0.0;
    t = now_s();
    int ok = 0;
    CK(solve(&ok));
    s.time_linear_solver = now_s() - t;
    s.iterations_linear_solver = static_cast<int>(last_info[0]);
    s.refine_steps = static_cast<int>(last_info[1]);
    s.fp64_fallback = static_cast<int>(last_info[3]);
    t = now_s();
    if (ok) CK(update());
    s.time_update = now_s() - t;
    // the chi2 SparseOptimizer records after the iteration (batch statistics)
    t = now_s();
    CK(compute_active_errors(&s.chi2));
    s.time_residuals = now_s() - t;
    s.status = ok ? 0 : 1;
    s.time_iteration = now_s() - t_it;
    if (stats) stats[i] = s;
    *done = i + 1;
    have_system = false;
    if (!ok) break;
  }
  HK(hipStreamSynchronize(stream));
  return G2OHIP_OK;
}

// OptimizationAlgorithmDogleg::solve (optimization_algorithm_dogleg.cpp:62-213)
// at its default properties (initialDelta 1e4, maxTrialsAfterFailure 100,
// initialLambda 1e-7, lambdaFactor 10). Linearisation, the (damped) solve,
// the state update and chi2 run on the device; the step algebra over the
// n-vectors b, h_gn, h_sd runs here in FP64. multiplyHessian is BlockSolver's
// (block_solver.h:149-151), the product with Hpp alone: block diagonal for
// bundle adjustment, with a zero landmark part as in the reference; the whole
// Hessian for pose graphs. The Hpp blocks are copied once per iteration, right
// after the linearisation (undamped).
int g2ohip_ctx::optimize_dogleg(int iterations, g2ohip_iter_stats* stats, int* done) {
  *done = 0;
  if (mode == 2) return fail(G2OHIP_ERR_UNSUPPORTED, "Dogleg needs multiplyHessian, not provided for BAL problems");
  pcg_residual_mem = -1.0;
  fp32_off = false;
  double delta = 1e4, lam = 1e-7;
  const double lf = 10.0;
  const int max_tries = 100;
  bool wasPD = true;
  const size_t n = system_size();
  std::vector<double> b(n), aux(n), hsd(n), hdl(n), hgn(n), H;
  auto dot = [n](const std::vector<double>& x, const std::vector<double>& y) {
    double t = 0;
    for (size_t i = 0; i < n; ++i) t += x[i] * y[i];
    return t;
  };
  auto hess_mul = [&](const std::vector<double>& v, std::vector<double>& out) {
    std::fill(out.begin(), out.end(), 0.0);
    if (mode == 0) {
      for (int c = 0; c < st.Cf; ++c)
        for (int r = 0; r < 6; ++r) {
          double t = 0;
          for (int k = 0; k < 6; ++k) t += H[36 * static_cast<size_t>(c) + 6 * r + k] * v[6 * static_cast<size_t>(c) + k];
          out[6 * static_cast<size_t>(c) + r] = t;
        }
    } else {
      for (int r = 0; r < pgst.Nf; ++r)
        for (int k = pgst.F_ptr[r]; k < pgst.F_ptr[r + 1]; ++k) {
          const int c = pgst.F_col[k];
          const double* B = &H[36 * static_cast<size_t>(k)];
          for (int i = 0; i < 6; ++i) {
            double t = 0;
            for (int j = 0; j < 6; ++j) t += B[6 * i + j] * v[6 * static_cast<size_t>(c) + j];
            out[6 * static_cast<size_t>(r) + i] += t;
          }
        }
    }
  };
  for (int it = 0; it < iterations; ++it) {
    if (force_stop && *force_stop) break;
    g2ohip_iter_stats s{};
    const double t_it = now_s();
    double t = now_s();
    CK(build_system(true));
    const double currentChi = chi2_lin;
    s.time_quadratic_form = now_s() - t;
    HK(hipMemcpy(b.data(), d_b.p, n * sizeof(double), hipMemcpyDeviceToHost));
    if (mode == 0) {
      H.resize(36 * static_cast<size_t>(st.Cf));
      HK(hipMemcpy(H.data(), d_Hpp.p, H.size() * sizeof(double), hipMemcpyDeviceToHost));
    } else {
      H.resize(36 * static_cast<size_t>(pgst.nF));
      HK(hipMemcpy(H.data(), d_pg_F_val.p, H.size() * sizeof(double), hipMemcpyDeviceToHost));
    }
    hess_mul(b, aux);
    const double alpha = dot(b, b) / dot(aux, b);
    for (size_t i = 0; i < n; ++i) hsd[i] = alpha * b[i];
    const double hsdNorm = std::sqrt(dot(hsd, hsd));
    double hgnNorm = -1.0, chiAfter = currentChi;
    bool solved = false, good = false;
    int tries = 0, status = 0;
    do {
      ++tries;
      if (!solved) {
        solved = true;
        int ok = 0;
        while (!ok) {
          lambda = wasPD ? 0.0 : lam;
          t = now_s();
          CK(solve(&ok));
          s.time_linear_solver += now_s() - t;
          s.iterations_linear_solver += static_cast<int>(last_info[0]);
          s.refine_steps += static_cast<int>(last_info[1]);
          s.fp64_fallback |= static_cast<int>(last_info[3]);
          wasPD = wasPD && ok;
          if (!wasPD) {
            if (ok) {
              lam = std::max(1e-12, lam / (0.5 * lf));
            } else {
              lam *= lf;
              if (lam > 1e3) { lam = 1e3; status = 2; break; }
            }
          }
        }
        if (status == 2) break;
        HK(hipMemcpy(hgn.data(), d_x.p, n * sizeof(double), hipMemcpyDeviceToHost));
        hgnNorm = std::sqrt(dot(hgn, hgn));
      }
      if (hgnNorm < delta) {
        hdl = hgn;
      } else if (hsdNorm > delta) {
        for (size_t i = 0; i < n; ++i) hdl[i] = delta / hsdNorm * hsd[i];
      } else {
        for (size_t i = 0; i < n; ++i) aux[i] = hgn[i] - hsd[i];
        const double c = dot(hsd, aux), bma = dot(aux, aux), hsd2 = dot(hsd, hsd);
        const double disc = std::sqrt(c * c + bma * (delta * delta - hsd2));
        const double beta = c <= 0 ? (-c + disc) / bma : (delta * delta - hsd2) / (c + disc);
        for (size_t i = 0; i < n; ++i) hdl[i] = hsd[i] + beta * (hgn[i] - hsd[i]);
      }
      hess_mul(hdl, aux);
      double lin = -dot(aux, hdl) + 2 * dot(b, hdl);
      t = now_s();
      CK(push());
      HK(hipMemcpy(d_x.p, hdl.data(), n * sizeof(double), hipMemcpyHostToDevice));
      have_solution = true;
      CK(update());
      s.time_update += now_s() - t;
      t = now_s();
      double newChi = 0;
      CK(compute_active_errors(&newChi));
      s.time_residuals += now_s() - t;
      if (std::fabs(lin) < 1e-12) lin = 1e-12;
      const double rho = (currentChi - newChi) / lin;
      if (rho > 0) {
        CK(discard_top());
        good = true;
        chiAfter = newChi;
      } else {
        CK(pop());
      }
      if (rho > 0.75)
        delta = std::max(delta, 3 * std::sqrt(dot(hdl, hdl)));
      else if (rho < 0.25)
        delta *= 0.5;
    } while (!good && tries < max_tries && !(force_stop && *force_stop));
    if (status == 0 && !good) status = 1;
    s.chi2_before = currentChi;
    s.chi2 = chiAfter;
    s.lambda = delta;  // the trust-region radius after the iteration
    s.levenberg_iterations = tries;
    s.status = status;
    s.time_iteration = now_s() - t_it;
    if (stats) stats[it] = s;
    *done = it + 1;
    have_system = false;
    if (status) break;
  }
  HK(hipStreamSynchronize(stream));
  return G2OHIP_OK;
}

// BlockSolver::computeMarginals (block_solver.hpp:479-490): the requested
// blocks of Hpp^-1, Hpp as the last buildSystem left it (undamped; after
// optimize that is the last iteration's linearisation). Projection edges join
// one camera and one point, so Hpp is block diagonal: a diagonal block is the
// inverse of that camera's 6x6 block and every other block is zero. Pose
// graphs: Hpp is the whole Hessian (pg_marginals).
int g2ohip_ctx::marginals(int nblk, const int* rows, const int* cols, double* out, double tol) {
  if (!have_lin) return fail(G2OHIP_ERR_INVALID, "no linearised system: build_system or optimize first");
  if (mode == 2) return fail(G2OHIP_ERR_UNSUPPORTED, "marginals are not provided for BAL problems");
  const int nb = mode == 0 ? st.Cf : pgst.Nf;
  if (nblk < 0 || (nblk > 0 && (!rows || !cols || !out))) return fail(G2OHIP_ERR_INVALID, "bad block list");
  for (int k = 0; k < nblk; ++k)
    if (rows[k] < 0 || rows[k] >= nb || cols[k] < 0 || cols[k] >= nb)
      return fail(G2OHIP_ERR_INVALID, "block index out of range");
  if (nblk == 0) return G2OHIP_OK;
  if (mode == 1) return pg_marginals(nblk, rows, cols, out, tol);
  g2ohip::DBuf<int> idx;
  g2ohip::DBuf<double> inv;
  std::vector<int> id(nb);
  for (int i = 0; i < nb; ++i) id[i] = i;
  HK(idx.upload(id, stream));
  HK(inv.alloc(36 * static_cast<size_t>(nb)));
  HK(hipMemsetAsync(d_err.p, 0, sizeof(int) * 4, stream));
  HK((launch_jacobi<double, 6>(nb, d_Hpp.p, idx.p, inv.p, d_err.p, stream)));
  int err = 0;
  CK(read_err(&err));
  if (err) return fail(G2OHIP_ERR_NOT_PD, "a camera block of Hpp is not positive definite");
  std::vector<double> h(36 * static_cast<size_t>(nb));
  HK(hipMemcpy(h.data(), inv.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int k = 0; k < nblk; ++k) {
    double* o = out + 36 * static_cast<size_t>(k);
    if (rows[k] == cols[k])
      std::memcpy(o, &h[36 * static_cast<size_t>(rows[k])], 36 * sizeof(double));
    else
      std::fill(o, o + 36, 0.0);
  }
  return G2OHIP_OK;
}

// ------------------------------------------------------------------ timing
// Kernel time of the PCG iteration on the last solve's operator: `reps`
// direction + SpMV launches back to back between two HIP events on the
// context's stream (us_spmv), then `reps` iterations of both launches
// (us_iter); with the one-launch iteration, `reps` launches of k_pcg_fused
// (both figures). The launches continue the last solve's recurrences on its
// buffers with a stop threshold that never triggers, so every launch does a
// full iteration's work; the buffers are scratch once the solution has been
// copied out, and the next solve re-initialises them.
int g2ohip_ctx::replay_pcg(int reps, double* us_spmv, double* us_iter) {
  if (mode != 0 || !have_solution || opt.schur_mode != G2OHIP_SCHUR_EXPLICIT)
    return fail(G2OHIP_ERR_INVALID, "needs a solve of the explicit bundle-adjustment system first");
  if (reps < 1) return fail(G2OHIP_ERR_INVALID, "reps must be positive");
  const bool single = last_solve_prec == G2OHIP_FP32;
  const int nb = st.Cf;
  DBuf<PcgScalars> sc;
  HK(sc.alloc(1));
  PcgScalars init{};
  init.dn[0] = init.dn[1] = 1.0;
  init.d0 = -1e300;  // never reached: every launch runs its whole iteration
  PcgControl ctl{0.0, -1.0, -1e300};
  hipEvent_t e[3];
  for (auto& ev : e) HK(hipEventCreate(&ev));
  auto body = [&](auto& B, const double* diag) -> int {
    HK(hipMemcpyAsync(sc.p, &init, sizeof(init), hipMemcpyHostToDevice, stream));
    HK(hipStreamSynchronize(stream));
    using MT = typename std::remove_reference<decltype(*B.F_val.p)>::type;
    if (pcg_fused && pcg_fused_fits(nb)) {
      // one launch per iteration: both figures are the launch of k_pcg_fused
      double* const s2[2] = {B.s.p, B.s2.p};
      double* const w2[2] = {B.w.p, B.w2.p};
      double* const dirs[2] = {B.d.p, B.d2.p};
      HK(hipEventRecord(e[0], stream));
      for (int k = 0; k < reps; ++k)
        HK((launch_pcg_fused<MT, 6>(nb, d_F_head.p, d_F_col.p, B.F_val.p, diag, B.Jinv.p, B.x.p, B.r.p, B.q.p, s2,
                                    w2, dirs, d_dn_part.p, d_fpart.p, sc.p, ctl, 1 + k, 0, stream, st.nF)));
      HK(hipEventRecord(e[1], stream));
      HK(hipEventSynchronize(e[1]));
      float a = 0;
      HK(hipEventElapsedTime(&a, e[0], e[1]));
      if (us_spmv) *us_spmv = 1e3 * a / reps;
      if (us_iter) *us_iter = 1e3 * a / reps;
      return G2OHIP_OK;
    }
    HK(hipEventRecord(e[0], stream));
    for (int k = 0; k < reps; ++k)
      HK((launch_pcg_dir_spmv<MT, 6>(nb, d_F_ptr.p, d_F_col.p, B.F_val.p, B.s.p, B.d2.p, B.d.p, B.q.p, d_dn_part.p,
                                     d_dq_part.p, sc.p, ctl, 1, 0, stream, diag, st.nF)));
    HK(hipEventRecord(e[1], stream));
    for (int k = 0; k < reps; ++k) {
      double* dn = (k & 1) ? B.d2.p : B.d.p;
      double* dold = (k & 1) ? B.d.p : B.d2.p;
      HK((launch_pcg_dir_spmv<MT, 6>(nb, d_F_ptr.p, d_F_col.p, B.F_val.p, B.s.p, dold, dn, B.q.p, d_dn_part.p,
                                     d_dq_part.p, sc.p, ctl, 1 + k, 0, stream, diag, st.nF)));
      HK((launch_pcg_update<MT, 6>(nb, B.Jinv.p, B.x.p, B.r.p, dn, B.q.p, B.s.p, d_dq_part.p, pcg_spmv_blocks(nb),
                                   d_dn_part.p, sc.p, 1 + k, stream)));
    }
    HK(hipEventRecord(e[2], stream));
    HK(hipEventSynchronize(e[2]));
    float a = 0, b = 0;
    HK(hipEventElapsedTime(&a, e[0], e[1]));
    HK(hipEventElapsedTime(&b, e[1], e[2]));
    if (us_spmv) *us_spmv = 1e3 * a / reps;
    if (us_iter) *us_iter = 1e3 * b / reps;
    return G2OHIP_OK;
  };
  const int rc = single ? body(f32, d_Sdiag.p) : body(f64, nullptr);
  for (auto& ev : e) (void)hipEventDestroy(ev);
  have_solution = rc == G2OHIP_OK && have_solution;
  return rc;
}

// --------------------------------------------------------------- inspection
void g2ohip_ctx::points_to_pub(double* v) const {
  if (mode == 1 || st.Pf == 0) return;
  const size_t off = cam_dim() * static_cast<size_t>(st.Cf);
  std::vector<double> tmp(v + off, v + off + 3 * static_cast<size_t>(st.Pf));
  for (int i = 0; i < st.Pf; ++i)
    for (int r = 0; r < 3; ++r) v[off + 3 * static_cast<size_t>(st.ph_pub[i]) + r] = tmp[3 * static_cast<size_t>(i) + r];
}

int g2ohip_ctx::get_hessian(double* Hpp, double* Hll, double* Hpl_edge) {
  if (mode != 0) return fail(G2OHIP_ERR_INVALID, "6-DoF bundle-adjustment call on a pose-graph or BAL context");
  if (!have_system && !have_exact) return fail(G2OHIP_ERR_INVALID, "build_system first");
  HK(hipStreamSynchronize(stream));
  if (Hpp) HK(hipMemcpy(Hpp, d_Hpp.p, 36 * sizeof(double) * st.Cf, hipMemcpyDeviceToHost));
  if (Hll) {
    std::vector<double> h(6 * static_cast<size_t>(st.Pf));
    if (st.Pf) HK(hipMemcpy(h.data(), d_Hll.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    const int map[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
    for (int p = 0; p < st.Pf; ++p)
      for (int k = 0; k < 9; ++k)
        Hll[9 * static_cast<size_t>(st.ph_pub[p]) + k] = h[6 * static_cast<size_t>(p) + map[k]];
  }
  if (Hpl_edge) {
    std::vector<double> h(18 * static_cast<size_t>(st.nq));
    if (system_prec == G2OHIP_FP32 && !have_exact) {
      std::vector<float> hf(h.size());
      if (st.nq) HK(hipMemcpy(hf.data(), f32.Hpl.p, hf.size() * sizeof(float), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < h.size(); ++i) h[i] = hf[i];
    } else if (st.nq) {
      HK(hipMemcpy(h.data(), f64.Hpl.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    }
    std::fill(Hpl_edge, Hpl_edge + 18 * static_cast<size_t>(st.E), 0.0);
    for (int s = 0; s < st.ns; ++s) {
      const int k = st.slot_edge[s], q = st.slot_q[s];
      if (q < 0) continue;
      for (int e = 0; e < 18; ++e) Hpl_edge[18 * static_cast<size_t>(k) + e] = h[18 * static_cast<size_t>(q) + e];
    }
  }
  return G2OHIP_OK;
}

int g2ohip_ctx::get_schur(int* rows, int* cols, double* blocks, double* bschur) {
  if (mode != 0) return fail(G2OHIP_ERR_INVALID, "6-DoF bundle-adjustment call on a pose-graph or BAL context");
  if (!have_solution) return fail(G2OHIP_ERR_INVALID, "solve first");
  if (opt.schur_mode != G2OHIP_SCHUR_EXPLICIT) return fail(G2OHIP_ERR_INVALID, "explicit Schur mode only");
  HK(hipStreamSynchronize(stream));
  for (int k = 0; k < st.nS; ++k) {
    if (rows) rows[k] = st.S_row[k];
    if (cols) cols[k] = st.S_col[k];
  }
  if (blocks) {
    std::vector<double> F(36 * static_cast<size_t>(st.nF));
    if (system_prec == G2OHIP_FP32) {
      std::vector<float> Ff(F.size());
      HK(hipMemcpy(Ff.data(), f32.F_val.p, Ff.size() * sizeof(float), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < F.size(); ++i) F[i] = Ff[i];
      // the FP32 operator keeps its diagonal blocks in FP64
      std::vector<double> D(36 * static_cast<size_t>(st.Cf));
      HK(hipMemcpy(D.data(), d_Sdiag.p, D.size() * sizeof(double), hipMemcpyDeviceToHost));
      for (int c = 0; c < st.Cf; ++c)
        std::copy(&D[36 * static_cast<size_t>(c)], &D[36 * static_cast<size_t>(c)] + 36,
                  &F[36 * static_cast<size_t>(st.up_pos[st.S_diag[c]])]);
    } else {
      HK(hipMemcpy(F.data(), f64.F_val.p, F.size() * sizeof(double), hipMemcpyDeviceToHost));
    }
    for (int k = 0; k < st.nS; ++k)
      std::copy(&F[36 * static_cast<size_t>(st.up_pos[k])], &F[36 * static_cast<size_t>(st.up_pos[k])] + 36,
                &blocks[36 * static_cast<size_t>(k)]);
  }
  if (bschur) HK(hipMemcpy(bschur, d_bschur.p, 6 * sizeof(double) * st.Cf, hipMemcpyDeviceToHost));
  return G2OHIP_OK;
}

// ------------------------------------------------ LinearSolverPCG on a host A
int g2ohip_ctx::linear_solve_pcg(int nb, int bd, int nblk, const int* brow, const int* bcol, const double* blocks,
                                 const double* b, double* x, double tol, int max_iter, double* residual_io,
                                 int* iterations) {
  CK(require_device());
  if (bd != 3 && bd != 6) return fail(G2OHIP_ERR_UNSUPPORTED, "block size must be 3 or 6");
  if (nb <= 0 || nblk <= 0 || !brow || !bcol || !blocks || !b || !x)
    return fail(G2OHIP_ERR_INVALID, "invalid linear system");
  // full block rows: both triangles, sorted by column
  std::vector<std::vector<std::pair<int, int>>> rows(nb);  // (col, signed block idx+1: <0 transposed)
  std::vector<int> diag(nb, -1);
  for (int k = 0; k < nblk; ++k) {
    int i = brow[k], j = bcol[k];
    if (i < 0 || j < 0 || i >= nb || j >= nb || i > j) return fail(G2OHIP_ERR_INVALID, "blocks must be upper (row <= col)");
    rows[i].push_back({j, k + 1});
    if (i != j) rows[j].push_back({i, -(k + 1)});
    else diag[i] = k;
  }
  for (int i = 0; i < nb; ++i)
    if (diag[i] < 0) return fail(G2OHIP_ERR_NOT_PD, "missing diagonal block");
  std::vector<int> Fptr(nb + 1, 0), Fcol, Fdiag(nb);
  std::vector<double> Fval;
  const int bb = bd * bd;
  for (int i = 0; i < nb; ++i) {
    std::sort(rows[i].begin(), rows[i].end());
    for (auto& e : rows[i]) {
      int k = std::abs(e.second) - 1;
      if (e.first == i && e.second > 0) Fdiag[i] = static_cast<int>(Fcol.size());
      Fcol.push_back(e.first);
      for (int r = 0; r < bd; ++r)
        for (int c = 0; c < bd; ++c)
          Fval.push_back(e.second > 0 ? blocks[static_cast<size_t>(k) * bb + r * bd + c]
                                      : blocks[static_cast<size_t>(k) * bb + c * bd + r]);
    }
    Fptr[i + 1] = static_cast<int>(Fcol.size());
  }
  DBuf<int> dFptr, dFcol, dFdiag;
  DBuf<double> dFval, dJ, db, dx, dr, dd, dd2, dq, ds, ddq, ddn;
  DBuf<int> derr;
  DBuf<PcgScalars> dsc;
  HK(dFptr.upload(Fptr, stream)); HK(dFcol.upload(Fcol, stream)); HK(dFdiag.upload(Fdiag, stream));
  HK(dFval.upload(Fval, stream));
  std::vector<double> bv(b, b + static_cast<size_t>(nb) * bd);
  HK(db.upload(bv, stream));
  const size_t n = static_cast<size_t>(nb) * bd;
  HK(dJ.alloc(static_cast<size_t>(nb) * bb)); HK(dx.alloc(n)); HK(dr.alloc(n)); HK(dd.alloc(n)); HK(dd2.alloc(n));
  HK(dq.alloc(n)); HK(ds.alloc(n)); HK(ddq.alloc(pcg_spmv_blocks(nb))); HK(ddn.alloc(pcg_row_blocks(nb)));
  HK(derr.alloc(1)); HK(dsc.alloc(1));
  HK(hipMemsetAsync(derr.p, 0, sizeof(int), stream));
  const PcgControl ctl = with_progress(PcgControl{tol, residual_io ? *residual_io : -1.0, -1.0});
  double* const dirs[2] = {dd.p, dd2.p};
  auto run = [&](auto bdc) -> int {
    constexpr int BD = decltype(bdc)::value;
    HK((launch_jacobi<double, BD>(nb, dFval.p, dFdiag.p, dJ.p, derr.p, stream)));
    const int mi = max_iter < 0 ? static_cast<int>(n) : max_iter;
    auto op = [&](int it, int last, const double* d_old, double* d) -> hipError_t {
      return launch_pcg_dir_spmv<double, BD>(nb, dFptr.p, dFcol.p, dFval.p, ds.p, d_old, d, dq.p, ddn.p, ddq.p,
                                             dsc.p, ctl, it, last, stream, nullptr, Fptr[nb]);
    };
    PcgRun pr;
    bool prog_failed = false;
    HK((pcg_drive<double, BD>(stream, h_pinned, progress(), nb, db.p, dJ.p, dx.p, dr.p, ds.p, dirs, dq.p, ddn.p,
                              ddq.p, pcg_spmv_blocks(nb), dsc.p, mi, op, &pr, &prog_failed)));
    if (prog_failed) prog_ok = false;
    if (iterations) *iterations = pr.iters;
    if (residual_io) *residual_io = 0.5 * pr.dn_final;
    return G2OHIP_OK;
  };
  if (bd == 3) CK(run(std::integral_constant<int, 3>()));
  else CK(run(std::integral_constant<int, 6>()));
  int err = 0;
  HK(hipStreamSynchronize(stream));
  HK(hipMemcpy(&err, derr.p, sizeof(int), hipMemcpyDeviceToHost));
  if (err) return fail(G2OHIP_ERR_NOT_PD, "singular diagonal block");
  HK(hipMemcpy(x, dx.p, n * sizeof(double), hipMemcpyDeviceToHost));
  return G2OHIP_OK;
}

// ------------------------------------------------ exact Hessian: Newton, adjoint
// Exact Hessian of F/2 in the solver's FP64 blocks (Hpp, Hll, Hpl); b is left
// as it is.
int g2ohip_ctx::exact_system() {
  HK(launch_cam_table(g, d_cams.p, d_camRt.p, stream));
  HK(f64.Hpl.alloc(18 * static_cast<size_t>(std::max(st.nq, 1))));
  HK(launch_linearize_exact(g, d_camRt.p, d_pts.p, f64.Hpl.p, d_Hll.p, d_hpp_part.p, stream));
  HK(launch_reduce_hpp(g, d_hpp_part.p, d_Hpp.p, nullptr, stream));
  have_system = have_lin = false;  // the blocks no longer hold the Gauss-Newton system
  have_exact = true;
  return G2OHIP_OK;
}

// x = (H + lam I)^-1 b for the current FP64 blocks: explicit Schur, PCG to the
// relative tolerance tol (no residual memory), back-substitution from the
// blocks. *ok = 0 if a point block or a preconditioner block is singular or
// the solve produced non-finite values, or the PCG met a direction of
// non-positive curvature (the role of a failed Cholesky).
int g2ohip_ctx::solve_exact(double lam, double tol, int max_iter, int* iters, int* ok, double* max_abs) {
  PrecBufs<double>& B = f64;
  const int nb = st.Cf;
  *ok = 1;
  *iters = 0;
  HK(hipMemsetAsync(d_err.p, 0, sizeof(int) * 4, stream));
  HK(B.Jinv.alloc(36 * static_cast<size_t>(nb)));
  HK(B.Dinv.alloc(9 * static_cast<size_t>(std::max(st.Pf, 1))));
  HK(B.F_val.alloc(36 * static_cast<size_t>(std::max(st.nF, 1))));
  HK(launch_point_dinv<double>(g, d_Hll.p, lam, B.Dinv.p, d_err.p, stream));
  HK(d_linv.alloc(6 * static_cast<size_t>(std::max(st.Pf, 1))));
  HK(launch_point_linv(g, d_Hll.p, lam, d_linv.p, d_err.p, stream));
  CK(schur_offdiag<double>(B));
  HK(launch_schur_diag<double>(g, d_q_slot.p, d_camRt.p, d_pts.p, B.Hpl.p, d_Hll.p, d_Hpp.p, d_b.p, lam, B.F_val.p,
                               B.Jinv.p, d_bschur.p, d_err.p, stream));
  int err = 0;
  CK(read_err(&err));
  if (err) { *ok = 0; return G2OHIP_OK; }
  PcgRun pr;
  // default iteration cap 20 n: the exact reduced systems of long camera
  // chains are ill-conditioned (cond ~1e11) and the solve is run to a tight
  // tolerance
  CK(run_pcg<double>(B, d_bschur.p, PcgControl{tol, -1.0, -1.0}, false, max_iter < 0 ? 120 * nb : max_iter, &pr));
  *iters = pr.iters;
  if (pr.breakdown) { *ok = 0; return G2OHIP_OK; }  // (H + lam I) is not positive definite
  HK(launch_to_double<double>(6 * nb, B.x.p, d_x.p, 0.0, stream));
  HK(launch_backsub_blocks<double>(g, B.Hpl.p, B.Dinv.p, d_b.p, d_x.p, stream));
  const size_t nx = 6 * static_cast<size_t>(st.Cf) + 3 * static_cast<size_t>(st.Pf);
  std::vector<double> hx(nx);
  HK(hipMemcpyAsync(hx.data(), d_x.p, nx * sizeof(double), hipMemcpyDeviceToHost, stream));
  HK(hipStreamSynchronize(stream));
  double m = 0;
  for (double v : hx) {
    if (!std::isfinite(v)) { *ok = 0; break; }
    m = std::max(m, std::fabs(v));
  }
  if (max_abs) *max_abs = m;
  return G2OHIP_OK;
}

// Newton steps with the exact Hessian, damped as Levenberg-Marquardt on that
// Hessian: a step is kept if the robust chi2 does not increase (within 1e-12
// relative: rounding at the optimum), else the state is restored and the
// damping mu (0 at first) raised to max(1e-15 * max diag, 10 mu); after an
// accepted step mu drops by 10 (to 0 below 1e-15 * max diag). A step whose
// largest component exceeds max_step (> 0) counts as rejected. At most 30
// trials per step; *ok = 0 if none was acceptable.
int g2ohip_ctx::newton(int steps, double tol, double max_step, int* ok) {
  if (mode != 0) return fail(G2OHIP_ERR_INVALID, "6-DoF bundle-adjustment call on a pose-graph or BAL context");
  CK(require_device());
  if (!(have_structure && have_meas && have_state)) return fail(G2OHIP_ERR_INVALID, "structure/measurements/state missing");
  *ok = 1;
  double mu = 0.0;
  for (int it = 0; it < steps && *ok; ++it) {
    double chi0 = 0;
    CK(compute_active_errors(&chi0));
    // b = -grad(F/2) from the Gauss-Newton linearisation, then the exact blocks
    HK(f64.Hpl.alloc(18 * static_cast<size_t>(std::max(st.nq, 1))));
    HK(launch_linearize<double>(g, d_camRt.p, d_pts.p, f64.Hpl.p, d_Hll.p, d_b.p, d_hpp_part.p, d_chunk_chi.p,
                                stream));
    HK(launch_reduce_hpp(g, d_hpp_part.p, d_Hpp.p, d_b.p, stream));
    CK(exact_system());
    HK(launch_max_diag(g, d_Hpp.p, d_Hll.p, d_red.p, d_scal.p + 2, stream));
    double mx = 0;
    CK(read_scalar(d_scal.p + 2, &mx));
    const double mu_min = 1e-15 * mx;
    bool accepted = false;
    for (int trial = 0; trial < 30 && !accepted; ++trial) {
      int iters = 0, solved = 0;
      double step = 0;
      CK(solve_exact(mu, tol, -1, &iters, &solved, &step));
      if (solved && max_step > 0 && step > max_step) solved = 0;  // outside the trust region
      if (solved) {
        CK(push());
        HK(launch_update(g, d_x.p, d_cams.p, d_pts.p, stream));
        double chi1 = 0;
        CK(compute_active_errors(&chi1));
        if (std::isfinite(chi1) && chi1 <= chi0 + 1e-12 * std::fabs(chi0)) {
          accepted = true;
          CK(discard_top());
        } else {
          CK(pop());
        }
      }
      if (!accepted) mu = std::max(mu_min, 10 * mu);
    }
    if (!accepted) *ok = 0;
    mu = (mu / 10 < mu_min) ? 0.0 : mu / 10;
  }
  HK(hipStreamSynchronize(stream));
  have_system = have_lin = false;
  have_solution = false;
  return G2OHIP_OK;
}

int g2ohip_ctx::adjoint(const double* dL_dcam, const double* dL_dpt, double* dL_dmeas, double* dL_dw, double tol,
                        int max_iter, int* iterations) {
  if (mode != 0) return fail(G2OHIP_ERR_INVALID, "6-DoF bundle-adjustment call on a pose-graph or BAL context");
  CK(require_device());
  if (!(have_structure && have_meas && have_state)) return fail(G2OHIP_ERR_INVALID, "structure/measurements/state missing");
  if (!dL_dcam || !dL_dmeas) return fail(G2OHIP_ERR_INVALID, "dL_dcam and dL_dmeas are required");
  CK(exact_system());
  // right-hand side dL/dtheta in the solver's ordering (free cameras, then free points)
  const size_t nx = 6 * static_cast<size_t>(st.Cf) + 3 * static_cast<size_t>(st.Pf);
  std::vector<double> rhs(nx, 0.0);
  for (int c = 0; c < st.C; ++c)
    if (st.camH[c] >= 0)
      for (int r = 0; r < 6; ++r) rhs[6 * static_cast<size_t>(st.camH[c]) + r] = dL_dcam[6 * static_cast<size_t>(c) + r];
  if (dL_dpt)
    for (int i = 0; i < st.P; ++i)
      if (st.ptH[i] >= 0)
        for (int r = 0; r < 3; ++r)
          rhs[6 * static_cast<size_t>(st.Cf) + 3 * static_cast<size_t>(st.ptH[i]) + r] = dL_dpt[3 * static_cast<size_t>(i) + r];
  HK(hipMemcpyAsync(d_b.p, rhs.data(), nx * sizeof(double), hipMemcpyHostToDevice, stream));
  int iters = 0, ok = 1;
  CK(solve_exact(0.0, tol, max_iter, &iters, &ok, nullptr));
  if (iterations) *iterations = iters;
  if (!ok) return fail(G2OHIP_ERR_NOT_PD, "exact Hessian is not positive definite at this state");
  DBuf<double> out;
  HK(out.alloc(3 * static_cast<size_t>(std::max(st.ns, 1))));
  HK(launch_adjoint_out(g, d_camRt.p, d_pts.p, d_x.p, out.p, stream));
  std::vector<double> h(3 * static_cast<size_t>(st.ns));
  HK(hipMemcpyAsync(h.data(), out.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
  HK(hipStreamSynchronize(stream));
  std::fill(dL_dmeas, dL_dmeas + 2 * static_cast<size_t>(st.E), 0.0);
  if (dL_dw) std::fill(dL_dw, dL_dw + st.E, 0.0);
  for (int s = 0; s < st.ns; ++s) {
    const size_t k = static_cast<size_t>(st.slot_edge[s]);
    dL_dmeas[2 * k] = h[3 * static_cast<size_t>(s)];
    dL_dmeas[2 * k + 1] = h[3 * static_cast<size_t>(s) + 1];
    if (dL_dw) dL_dw[k] = h[3 * static_cast<size_t>(s) + 2];
  }
  have_solution = false;
  return G2OHIP_OK;
}
