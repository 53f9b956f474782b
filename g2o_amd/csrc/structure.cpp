// Host-side structure builder (see structure.h).
#include "structure.h"

#include <algorithm>
#include <cstdint>
#include <numeric>
#include <unordered_set>

namespace g2ohip {

namespace {

template <typename T>
void csr_from_counts(std::vector<T>& ptr) {
  T acc = 0;
  for (auto& v : ptr) {
    T c = v;
    v = acc;
    acc += c;
  }
}

}  // namespace

std::string build_structure(BAStructure& st, int C, int P, int E, const int* edge_cam, const int* edge_pt,
                            const uint8_t* cam_fixed, const uint8_t* pt_fixed, bool schur) {
  st = BAStructure();
  if (C <= 0 || P < 0 || E < 0) return "invalid sizes";
  st.C = C; st.P = P; st.E = E;
  for (int k = 0; k < E; ++k) {
    if (edge_cam[k] < 0 || edge_cam[k] >= C) return "edge camera index out of range";
    if (edge_pt[k] < 0 || edge_pt[k] >= P) return "edge point index out of range";
  }
  // hessianIndex: free vertices in input order (sparse_optimizer.cpp:175-200)
  st.camH.assign(C, -1);
  st.ptH.assign(P, -1);
  for (int c = 0; c < C; ++c)
    if (!(cam_fixed && cam_fixed[c])) { st.camH[c] = st.Cf++; st.free_cam_raw.push_back(c); }
  for (int p = 0; p < P; ++p)
    if (!(pt_fixed && pt_fixed[p])) st.ptH[p] = st.Pf++;
  if (st.Cf == 0) return "all cameras are fixed";
  if (static_cast<long long>(st.Cf) * st.Cf >= (1LL << 31)) return "too many free cameras for 32-bit block keys";

  // active edges grouped by point
  std::vector<int> pcount(P + 1, 0);
  for (int k = 0; k < E; ++k) {
    bool cf = st.camH[edge_cam[k]] < 0, pf = st.ptH[edge_pt[k]] < 0;
    if (!(cf && pf)) pcount[edge_pt[k]]++;
  }
  std::vector<int> pptr(P + 1, 0);
  for (int p = 0; p < P; ++p) pptr[p + 1] = pptr[p] + pcount[p];
  std::vector<int> pedges(pptr[P]);
  {
    std::vector<int> fill(pptr.begin(), pptr.end() - 1);
    for (int k = 0; k < E; ++k) {
      bool cf = st.camH[edge_cam[k]] < 0, pf = st.ptH[edge_pt[k]] < 0;
      if (!(cf && pf)) pedges[fill[edge_pt[k]]++] = k;
    }
  }
  auto camkey = [&](int c) { return st.camH[c] >= 0 ? st.camH[c] : st.Cf + c; };
  // locality order of the points: (first camera, last camera, id)
  struct PK { int lo, hi, p; };
  std::vector<PK> order;
  order.reserve(P);
  for (int p = 0; p < P; ++p) {
    if (pptr[p + 1] == pptr[p]) continue;
    int lo = 1 << 30, hi = -1;
    for (int q = pptr[p]; q < pptr[p + 1]; ++q) {
      int k = camkey(edge_cam[pedges[q]]);
      lo = std::min(lo, k);
      hi = std::max(hi, k);
    }
    order.push_back({lo, hi, p});
  }
  std::sort(order.begin(), order.end(), [](const PK& a, const PK& b) {
    if (a.lo != b.lo) return a.lo < b.lo;
    if (a.hi != b.hi) return a.hi < b.hi;
    return a.p < b.p;
  });
  st.npos = static_cast<int>(order.size());
  st.pos_pt.resize(st.npos);
  st.pos_h.resize(st.npos);
  st.pos_ptr.assign(st.npos + 1, 0);
  st.pos_nfree.assign(st.npos, 0);
  st.edge_slot.assign(E, -1);
  st.ns = pptr[P];
  st.slot_edge.resize(st.ns);
  st.slot_cam.resize(st.ns);
  st.slot_pos.resize(st.ns);
  int s = 0;
  std::vector<int> tmp;
  for (int i = 0; i < st.npos; ++i) {
    int p = order[i].p;
    st.pos_pt[i] = p;
    st.pos_h[i] = st.ptH[p];
    st.pos_ptr[i] = s;
    tmp.assign(pedges.begin() + pptr[p], pedges.begin() + pptr[p + 1]);
    std::stable_sort(tmp.begin(), tmp.end(), [&](int a, int b) { return camkey(edge_cam[a]) < camkey(edge_cam[b]); });
    for (int k : tmp) {
      st.slot_edge[s] = k;
      st.slot_cam[s] = edge_cam[k];
      st.slot_pos[s] = i;
      st.edge_slot[k] = s;
      if (st.camH[edge_cam[k]] >= 0) st.pos_nfree[i]++;
      ++s;
    }
  }
  st.pos_ptr[st.npos] = s;
  st.slot_camh.resize(st.ns);
  st.slot_ph.resize(st.ns);
  for (int q = 0; q < st.ns; ++q) {
    st.slot_camh[q] = st.camH[st.slot_cam[q]];
    st.slot_ph[q] = st.pos_h[st.slot_pos[q]];
  }
  st.cam_qptr.assign(st.Cf + 1, 0);
  for (int q = 0; q < st.ns; ++q)
    if (st.slot_camh[q] >= 0 && st.slot_ph[q] >= 0) st.cam_qptr[st.slot_camh[q] + 1]++;
  for (int c = 0; c < st.Cf; ++c) st.cam_qptr[c + 1] += st.cam_qptr[c];
  st.nq = st.cam_qptr[st.Cf];
  st.q_slot.resize(st.nq);
  st.q_ph.resize(st.nq);
  st.slot_q.assign(st.ns, -1);
  {
    std::vector<int> fill(st.cam_qptr.begin(), st.cam_qptr.end() - 1);
    for (int q = 0; q < st.ns; ++q) {  // slots ascend: each camera lists its points in processing order
      const int c = st.slot_camh[q];
      if (c < 0 || st.slot_ph[q] < 0) continue;
      const int d = fill[c]++;
      st.q_slot[d] = q;
      st.q_ph[d] = st.slot_ph[q];
      st.slot_q[q] = d;
    }
  }

  // linearisation chunks: consecutive points, bounded slots / cameras / points
  const LinPlanCaps& lc = st.lcaps;
  std::vector<int> stamp(st.Cf, -1);
  std::vector<int> cams;
  st.lc_pos0.push_back(0);
  st.lc_cam_ptr.push_back(0);
  int cur_slots = 0, cur_pts = 0;
  auto close_chunk = [&](int end_pos) {
    std::sort(cams.begin(), cams.end());
    st.lc_cams.insert(st.lc_cams.end(), cams.begin(), cams.end());
    st.lc_pos0.push_back(end_pos);
    st.lc_cam_ptr.push_back(static_cast<int>(st.lc_cams.size()));
    cams.clear();
    cur_slots = cur_pts = 0;
    st.nlc++;
  };
  for (int i = 0; i < st.npos; ++i) {
    int n = st.pos_ptr[i + 1] - st.pos_ptr[i];
    int nf = st.pos_nfree[i];
    if (n > lc.max_slots || nf > lc.max_cams)
      return "a point has more observations than one workgroup handles (" + std::to_string(n) + ")";
    int newc = 0;
    for (int q = st.pos_ptr[i]; q < st.pos_ptr[i] + nf; ++q) {
      int c = st.camH[st.slot_cam[q]];
      if (stamp[c] != st.nlc) newc++;
    }
    // duplicates of one camera inside the point count once
    if (cur_pts > 0 && (cur_slots + n > lc.max_slots || static_cast<int>(cams.size()) + newc > lc.max_cams ||
                        cur_pts + 1 > lc.max_pts))
      close_chunk(i);
    for (int q = st.pos_ptr[i]; q < st.pos_ptr[i] + nf; ++q) {
      int c = st.camH[st.slot_cam[q]];
      if (stamp[c] != st.nlc) { stamp[c] = st.nlc; cams.push_back(c); }
    }
    cur_slots += n;
    cur_pts += 1;
  }
  if (cur_pts > 0) close_chunk(st.npos);
  // per free camera: the chunk rows that hold its partial Hpp / b
  st.camp_ptr.assign(st.Cf + 1, 0);
  for (int c : st.lc_cams) st.camp_ptr[c + 1]++;
  for (int c = 0; c < st.Cf; ++c) st.camp_ptr[c + 1] += st.camp_ptr[c];
  st.camp_idx.resize(st.lc_cams.size());
  {
    std::vector<int> fill(st.camp_ptr.begin(), st.camp_ptr.end() - 1);
    for (size_t r = 0; r < st.lc_cams.size(); ++r) st.camp_idx[fill[st.lc_cams[r]]++] = static_cast<int>(r);
  }

  // Hschur pattern (or the Hpp diagonal only without Schur)
  std::vector<uint64_t> keys;
  for (int c = 0; c < st.Cf; ++c) keys.push_back(static_cast<uint64_t>(c) * st.Cf + c);
  if (schur) {
    for (int i = 0; i < st.npos; ++i) {
      if (st.pos_h[i] < 0) continue;
      int a0 = st.pos_ptr[i], nf = st.pos_nfree[i];
      for (int a = 0; a < nf; ++a) {
        uint64_t ca = st.camH[st.slot_cam[a0 + a]];
        for (int b = a; b < nf; ++b) keys.push_back(ca * st.Cf + st.camH[st.slot_cam[a0 + b]]);
      }
    }
  }
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  st.nS = static_cast<int>(keys.size());
  st.S_row.resize(st.nS);
  st.S_col.resize(st.nS);
  st.S_diag.assign(st.Cf, -1);
  st.S_rowptr.assign(st.Cf + 1, 0);
  for (int k = 0; k < st.nS; ++k) {
    st.S_row[k] = static_cast<int>(keys[k] / st.Cf);
    st.S_col[k] = static_cast<int>(keys[k] % st.Cf);
    if (st.S_row[k] == st.S_col[k]) st.S_diag[st.S_row[k]] = k;
    st.S_rowptr[st.S_row[k] + 1]++;
  }
  for (int c = 0; c < st.Cf; ++c) st.S_rowptr[c + 1] += st.S_rowptr[c];
  // full symmetric block rows: row i holds (i, j) for every stored (i, j) or (j, i)
  st.F_ptr.assign(st.Cf + 1, 0);
  for (int k = 0; k < st.nS; ++k) {
    st.F_ptr[st.S_row[k] + 1]++;
    if (st.S_row[k] != st.S_col[k]) st.F_ptr[st.S_col[k] + 1]++;
  }
  for (int c = 0; c < st.Cf; ++c) st.F_ptr[c + 1] += st.F_ptr[c];
  st.nF = st.F_ptr[st.Cf];
  st.F_col.resize(st.nF);
  st.F_up.resize(st.nF);
  st.F_trans.resize(st.nF);
  st.up_pos.assign(st.nS, -1);
  st.up_posT.assign(st.nS, -1);
  {
    // lower entries (j < i) of row i come from upper blocks (j, i) with j < i;
    // emit them first in column order, then the upper entries: rows stay sorted.
    std::vector<int> fill(st.F_ptr.begin(), st.F_ptr.end() - 1);
    for (int k = 0; k < st.nS; ++k) {  // sorted by (row, col): columns j ascending per target row
      int i = st.S_row[k], j = st.S_col[k];
      if (i == j) continue;
      int f = fill[j]++;
      st.F_col[f] = i; st.F_up[f] = k; st.F_trans[f] = 1; st.up_posT[k] = f;
    }
    for (int k = 0; k < st.nS; ++k) {
      int i = st.S_row[k], j = st.S_col[k];
      int f = fill[i]++;
      st.F_col[f] = j; st.F_up[f] = k; st.F_trans[f] = 0; st.up_pos[k] = f;
    }
  }
  return "";
}

std::string build_row_plan(BAStructure& st, int max_cols, int chunk, int num_xcds) {
  RowPlan& rp = st.rows;
  rp = RowPlan();
  rp.max_cols = max_cols;
  rp.chunk = chunk;
  const int Cf = st.Cf;
  for (int i = 0; i < st.npos; ++i)
    if (st.pos_h[i] >= 0) rp.pair_instances += static_cast<long long>(st.pos_nfree[i]) * (st.pos_nfree[i] + 1) / 2;
  if (rp.pair_instances >= (1LL << 31)) return "more than 2^31 Schur pair products";
  for (int r = 0; r < Cf; ++r) {
    const int k0 = st.S_rowptr[r], k1 = st.S_rowptr[r + 1];
    if (k1 <= k0 || st.S_col[k0] != r) return "Hschur row without its diagonal block";
    rp.max_row_cols = std::max(rp.max_row_cols, k1 - k0);
    for (int k = k0; k < k1; k += max_cols) {
      rp.tile_row.push_back(r);
      rp.tile_k0.push_back(k);
      rp.tile_k1.push_back(std::min(k1, k + max_cols));
    }
  }
  // pair products grouped by upper block (counting sort by block)
  rp.blk_pair_ptr.assign(st.nS + 1, 0);
  std::vector<int> pair_blk;
  pair_blk.reserve(static_cast<size_t>(rp.pair_instances));
  for (int i = 0; i < st.npos; ++i) {
    if (st.pos_h[i] < 0) continue;
    const int a0 = st.pos_ptr[i], nf = st.pos_nfree[i];
    for (int a = 0; a < nf; ++a) {
      int k = st.S_rowptr[st.slot_camh[a0 + a]];
      for (int bb = a; bb < nf; ++bb) {
        const int cb = st.slot_camh[a0 + bb];
        while (st.S_col[k] < cb) ++k;  // cameras ascend within the point
        pair_blk.push_back(k);
        rp.blk_pair_ptr[k + 1]++;
      }
    }
  }
  for (int k = 0; k < st.nS; ++k) rp.blk_pair_ptr[k + 1] += rp.blk_pair_ptr[k];
  rp.pair_qa.resize(static_cast<size_t>(rp.pair_instances));
  rp.pair_qb.resize(static_cast<size_t>(rp.pair_instances));
  {
    std::vector<int> fill(rp.blk_pair_ptr.begin(), rp.blk_pair_ptr.end() - 1);
    size_t q = 0;
    for (int i = 0; i < st.npos; ++i) {
      if (st.pos_h[i] < 0) continue;
      const int a0 = st.pos_ptr[i], nf = st.pos_nfree[i];
      for (int a = 0; a < nf; ++a)
        for (int bb = a; bb < nf; ++bb) {
          const int dst = fill[pair_blk[q++]]++;
          rp.pair_qa[dst] = st.slot_q[a0 + a];
          rp.pair_qb[dst] = st.slot_q[a0 + bb];
        }
    }
  }
  // Workgroups are dealt to the XCDs round-robin by blockIdx; map blockIdx so
  // that each XCD walks a contiguous range of rows (neighbouring cameras share
  // points, so their Hpl blocks meet in that XCD's L2). blockIdx blk runs on
  // XCD blk % X as its (blk / X)-th workgroup.
  const int T = static_cast<int>(rp.tile_row.size());
  const int X = std::max(1, num_xcds);
  std::vector<int> start(X + 1, 0);
  for (int x = 0; x < X; ++x) start[x + 1] = start[x] + (T > x ? (T - x + X - 1) / X : 0);
  rp.tile_order.assign(T, 0);
  for (int blk = 0; blk < T; ++blk) rp.tile_order[blk] = start[blk % X] + blk / X;
  return "";
}

std::string build_pg_structure(PGStructure& st, int N, int E, const int* from, const int* to, const uint8_t* fixed) {
  st = PGStructure();
  if (N <= 0 || E < 0) return "invalid pose graph size";
  st.N = N;
  st.E = E;
  st.H.assign(N, -1);
  for (int i = 0; i < N; ++i)
    if (!(fixed && fixed[i])) st.H[i] = st.Nf++;
  if (st.Nf == 0) return "every pose is fixed";
  for (int k = 0; k < E; ++k) {
    if (from[k] < 0 || from[k] >= N || to[k] < 0 || to[k] >= N) return "edge vertex index out of range";
    if (from[k] == to[k]) return "edge joins a pose to itself";
    if (st.H[from[k]] >= 0 || st.H[to[k]] >= 0) st.act.push_back(k);
  }
  const int A = static_cast<int>(st.act.size());
  // upper pattern: diagonal, then distinct pairs sorted by (row, col)
  std::vector<std::pair<long long, int>> keys;  // (row * Nf + col, active position)
  st.edge_blk.assign(A, -1);
  st.edge_swap.assign(A, 0);
  for (int a = 0; a < A; ++a) {
    const int k = st.act[a];
    const int hi = st.H[from[k]], hj = st.H[to[k]];
    if (hi < 0 || hj < 0) continue;
    st.edge_swap[a] = hi > hj ? 1 : 0;
    const int r = std::min(hi, hj), c = std::max(hi, hj);
    keys.push_back({static_cast<long long>(r) * st.Nf + c, a});
  }
  std::sort(keys.begin(), keys.end());
  st.U_row.resize(st.Nf);
  st.U_col.resize(st.Nf);
  for (int i = 0; i < st.Nf; ++i) st.U_row[i] = st.U_col[i] = i;
  st.blk_ptr.push_back(0);
  for (size_t q = 0; q < keys.size(); ++q) {
    if (q == 0 || keys[q].first != keys[q - 1].first) {
      if (q) st.blk_ptr.push_back(static_cast<int>(st.blk_ref.size()));
      st.U_row.push_back(static_cast<int>(keys[q].first / st.Nf));
      st.U_col.push_back(static_cast<int>(keys[q].first % st.Nf));
    }
    st.edge_blk[keys[q].second] = static_cast<int>(st.U_row.size()) - 1;
    st.blk_ref.push_back(keys[q].second);
  }
  st.blk_ptr.push_back(static_cast<int>(st.blk_ref.size()));
  if (keys.empty()) st.blk_ptr.assign(1, 0);
  st.nU = static_cast<int>(st.U_row.size());
  // edge ends per free pose, in active-edge order
  st.end_ptr.assign(st.Nf + 1, 0);
  for (int a = 0; a < A; ++a) {
    const int k = st.act[a];
    if (st.H[from[k]] >= 0) st.end_ptr[st.H[from[k]] + 1]++;
    if (st.H[to[k]] >= 0) st.end_ptr[st.H[to[k]] + 1]++;
  }
  for (int i = 0; i < st.Nf; ++i) st.end_ptr[i + 1] += st.end_ptr[i];
  st.end_ref.resize(st.end_ptr[st.Nf]);
  {
    std::vector<int> fill(st.end_ptr.begin(), st.end_ptr.end() - 1);
    for (int a = 0; a < A; ++a) {
      const int k = st.act[a];
      if (st.H[from[k]] >= 0) st.end_ref[fill[st.H[from[k]]]++] = 2 * a;
      if (st.H[to[k]] >= 0) st.end_ref[fill[st.H[to[k]]]++] = 2 * a + 1;
    }
  }
  // full block rows
  std::vector<std::vector<std::pair<int, int>>> rows(st.Nf);  // (col, upper block, transposed?)
  for (int k = 0; k < st.nU; ++k) {
    rows[st.U_row[k]].push_back({st.U_col[k], 2 * k});
    if (st.U_row[k] != st.U_col[k]) rows[st.U_col[k]].push_back({st.U_row[k], 2 * k + 1});
  }
  st.F_ptr.assign(st.Nf + 1, 0);
  st.up_pos.assign(st.nU, -1);
  st.up_posT.assign(st.nU, -1);
  st.F_diag.assign(st.Nf, -1);
  for (int i = 0; i < st.Nf; ++i) {
    std::sort(rows[i].begin(), rows[i].end());
    for (auto& e : rows[i]) {
      const int pos = static_cast<int>(st.F_col.size());
      st.F_col.push_back(e.first);
      if (e.second & 1) st.up_posT[e.second >> 1] = pos;
      else st.up_pos[e.second >> 1] = pos;
      if (e.first == i) st.F_diag[i] = pos;
    }
    st.F_ptr[i + 1] = static_cast<int>(st.F_col.size());
  }
  st.nF = static_cast<int>(st.F_col.size());
  return "";
}

}  // namespace g2ohip
