// Host-side structure builder (see structure.h).
#include "structure.h"

#include "kernels.h"

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
                            const uint8_t* cam_fixed, const uint8_t* pt_fixed, bool schur,
                            const LinPlanCaps* caps) {
  st = BAStructure();
  if (caps) st.lcaps = *caps;
  if (C <= 0 || P < 0 || E < 0) return "invalid sizes";
  st.C = C; st.P = P; st.E = E;
  for (int k = 0; k < E; ++k) {
    if (edge_cam[k] < 0 || edge_cam[k] >= C) return "edge camera index out of range";
    if (edge_pt[k] < 0 || edge_pt[k] >= P) return "edge point index out of range";
  }
  // hessianIndex: the free vertices among the active ones, in input order
  // (sparse_optimizer.cpp:175-200). A vertex is active if one of its edges is
  // (sparse_optimizer.cpp:217-285: an edge is active unless all its vertices
  // are fixed), so a free vertex is active if it has any edge.
  std::vector<uint8_t> cam_used(C, 0), pt_used(P, 0);
  for (int k = 0; k < E; ++k) cam_used[edge_cam[k]] = pt_used[edge_pt[k]] = 1;
  st.camH.assign(C, -1);
  st.ptH.assign(P, -1);
  for (int c = 0; c < C; ++c)
    if (!(cam_fixed && cam_fixed[c]) && cam_used[c]) { st.camH[c] = st.Cf++; st.free_cam_raw.push_back(c); }
  for (int p = 0; p < P; ++p)
    if (!(pt_fixed && pt_fixed[p]) && pt_used[p]) st.ptH[p] = st.Pf++;
  if (st.Cf == 0) return "no free camera has an edge";
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
  // device point index: the free points renumbered in processing order, so
  // every per-point array (Hll, b_l, x_l, the point factors) is read in the
  // order the point-major kernels walk it; ph_pub maps it back to the
  // hessianIndex above for the interface's vectors and blocks
  st.ph_pub.assign(st.Pf, -1);
  {
    int nxt = 0;
    for (const PK& o : order)
      if (st.ptH[o.p] >= 0) {
        st.ph_pub[nxt] = st.ptH[o.p];
        st.ptH[o.p] = nxt++;
      }
    if (nxt != st.Pf) return "internal error: free point without an active edge";
  }
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
  // device order of the point state: the points with a position at their
  // position, then the rest in input order
  st.pt_dev.assign(P, -1);
  for (int i = 0; i < st.npos; ++i) st.pt_dev[st.pos_pt[i]] = i;
  {
    int nxt = st.npos;
    for (int p = 0; p < P; ++p)
      if (st.pt_dev[p] < 0) st.pt_dev[p] = nxt++;
  }
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
  st.q_pt.resize(st.nq);
  st.slot_q.assign(st.ns, -1);
  {
    std::vector<int> fill(st.cam_qptr.begin(), st.cam_qptr.end() - 1);
    for (int q = 0; q < st.ns; ++q) {  // slots ascend: each camera lists its points in processing order
      const int c = st.slot_camh[q];
      if (c < 0 || st.slot_ph[q] < 0) continue;
      const int d = fill[c]++;
      st.q_slot[d] = q;
      st.q_ph[d] = st.slot_ph[q];
      st.q_pt[d] = st.slot_pos[q];
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
    st.lc_maxc = std::max(st.lc_maxc, static_cast<int>(cams.size()));
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
    if (nf > lc.max_point_cams)
      return "a point is observed by more than " + std::to_string(lc.max_point_cams) + " free cameras (" +
             std::to_string(nf) + ")";
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

long long schur_pair_count(const BAStructure& st) {
  long long pairs = 0;
  for (int p = 0; p < st.npos; ++p) {
    if (st.pos_h[p] < 0) continue;
    // free slots sorted by camera: pairs of different cameras
    const int s0 = st.pos_ptr[p], n = st.pos_nfree[p];
    for (int i = 0; i < n;) {
      int j = i;
      while (j < n && st.slot_camh[s0 + j] == st.slot_camh[s0 + i]) ++j;
      pairs += static_cast<long long>(j - i) * (n - j);
      i = j;
    }
  }
  return pairs;
}

int max_point_obs(const BAStructure& st) {
  int m = 0;
  for (int p = 0; p < st.npos; ++p)
    if (st.pos_h[p] >= 0) m = std::max(m, st.pos_nfree[p]);
  return m;
}

std::string build_win_plan(const BAStructure& st, WinPlan& wp, int obs_max, int pch) {
  wp = WinPlan();
  if (obs_max < 1 || obs_max > 65535 || pch < 1) return "invalid window plan parameters";
  wp.obs_max = obs_max;
  wp.pch = pch;
  const int Cf = st.Cf;
  for (int r = 0; r < Cf; ++r) {
    const int k0 = st.S_rowptr[r], k1 = st.S_rowptr[r + 1];
    if (k1 <= k0 || st.S_col[k0] != r) return "Hschur row without its diagonal block";
  }
  std::vector<int> od_of(st.nS, -1);
  for (int k = 0; k < st.nS; ++k)
    if (st.S_row[k] != st.S_col[k]) {
      od_of[k] = static_cast<int>(wp.od_blk.size());
      wp.od_blk.push_back(k);
    }
  const int nod = static_cast<int>(wp.od_blk.size());
  std::vector<int> od_cnt(nod, 0);
  // partial id -> (off-diagonal block, rank within the block); converted to
  // block-major slots once every block's count is known
  std::vector<int> part_od, part_rank;
  auto nobs = [&](int pos) { return st.pos_h[pos] >= 0 ? st.pos_nfree[pos] : 0; };
  wp.win_obs0.push_back(0);
  wp.win_grp0.push_back(0);
  wp.grp_row0.push_back(0);
  std::vector<int> qs;
  std::vector<std::pair<int, uint32_t>> lst;  // (off-diagonal block, pair)
  struct Part { int od, beg, n; };
  std::vector<Part> parts;
  int pos = 0;
  while (pos < st.npos) {
    int pe = pos, cur = 0;
    while (pe < st.npos) {
      const int n = nobs(pe);
      if (n > obs_max)
        return "a point has more than " + std::to_string(obs_max) + " observations with a free camera";
      if (cur > 0 && cur + n > obs_max) break;
      cur += n;
      ++pe;
    }
    qs.clear();
    for (int p = pos; p < pe; ++p)
      for (int k = 0; k < nobs(p); ++k) qs.push_back(st.slot_q[st.pos_ptr[p] + k]);
    std::sort(qs.begin(), qs.end());
    wp.win_q.insert(wp.win_q.end(), qs.begin(), qs.end());
    wp.win_obs0.push_back(static_cast<int>(wp.win_q.size()));
    auto local = [&](int q) {
      return static_cast<uint32_t>(std::lower_bound(qs.begin(), qs.end(), q) - qs.begin());
    };
    lst.clear();
    for (int p = pos; p < pe; ++p) {
      const int s0 = st.pos_ptr[p], nf = nobs(p);
      for (int a = 0; a < nf; ++a) {
        const int ca = st.slot_camh[s0 + a];
        const uint32_t la = local(st.slot_q[s0 + a]);
        const int* cb0 = st.S_col.data() + st.S_rowptr[ca];
        const int* cb1 = st.S_col.data() + st.S_rowptr[ca + 1];
        for (int bb = a + 1; bb < nf; ++bb) {
          const int cb = st.slot_camh[s0 + bb];
          if (cb == ca) continue;  // a second edge to the same camera: diagonal block
          const int k = static_cast<int>(std::lower_bound(cb0, cb1, cb) - st.S_col.data());
          if (k >= st.S_rowptr[ca + 1] || st.S_col[k] != cb) return "Schur pattern misses a co-observed pair";
          lst.emplace_back(od_of[k], la | (local(st.slot_q[s0 + bb]) << 16));
        }
      }
    }
    wp.pairs += static_cast<long long>(lst.size());
    std::stable_sort(lst.begin(), lst.end(),
                     [](const std::pair<int, uint32_t>& x, const std::pair<int, uint32_t>& y) { return x.first < y.first; });
    parts.clear();
    for (size_t i = 0; i < lst.size();) {
      size_t j = i;
      while (j < lst.size() && lst[j].first == lst[i].first) ++j;
      for (size_t c = i; c < j; c += pch)
        parts.push_back({lst[i].first, static_cast<int>(c), static_cast<int>(std::min<size_t>(pch, j - c))});
      i = j;
    }
    std::stable_sort(parts.begin(), parts.end(), [](const Part& x, const Part& y) { return x.n > y.n; });
    for (size_t g0 = 0; g0 < parts.size(); g0 += 64) {
      const int len = parts[g0].n;
      wp.grp_len.push_back(len);
      for (int r = 0; r < len; ++r)
        for (int lane = 0; lane < 64; ++lane) {
          const size_t i = g0 + lane;
          wp.pair_rows.push_back(i < parts.size() && r < parts[i].n ? lst[parts[i].beg + r].second : 0xFFFFFFFFu);
        }
      wp.grp_row0.push_back(wp.grp_row0.back() + len);
      for (int lane = 0; lane < 64; ++lane) {
        const size_t i = g0 + lane;
        if (i < parts.size()) {
          wp.grp_slot.push_back(static_cast<int>(part_od.size()));
          part_od.push_back(parts[i].od);
          part_rank.push_back(od_cnt[parts[i].od]++);
        } else {
          wp.grp_slot.push_back(-1);
        }
      }
    }
    wp.win_grp0.push_back(static_cast<int>(wp.grp_len.size()));
    pos = pe;
  }
  if (static_cast<long long>(wp.pair_rows.size()) >= (1LL << 31)) return "more than 2^31 Schur pair entries";
  wp.od_ptr.assign(nod + 1, 0);
  for (int u = 0; u < nod; ++u) wp.od_ptr[u + 1] = wp.od_ptr[u] + od_cnt[u];
  for (int& sl : wp.grp_slot)
    if (sl >= 0) sl = wp.od_ptr[part_od[sl]] + part_rank[sl];
  return "";
}

std::string build_pg_structure(PGStructure& st, int N, int E, const int* from, const int* to, const uint8_t* fixed) {
  st = PGStructure();
  if (N <= 0 || E < 0) return "invalid pose graph size";
  st.N = N;
  st.E = E;
  // free poses with an edge are the active ones (sparse_optimizer.cpp:217-285)
  std::vector<uint8_t> used(N, 0);
  for (int k = 0; k < E; ++k) {
    if (from[k] < 0 || from[k] >= N || to[k] < 0 || to[k] >= N) return "edge vertex index out of range";
    if (from[k] == to[k]) return "edge joins a pose to itself";
    used[from[k]] = used[to[k]] = 1;
  }
  st.H.assign(N, -1);
  for (int i = 0; i < N; ++i)
    if (!(fixed && fixed[i]) && used[i]) st.H[i] = st.Nf++;
  if (st.Nf == 0) return "no free pose has an edge";
  for (int k = 0; k < E; ++k)
    if (st.H[from[k]] >= 0 || st.H[to[k]] >= 0) st.act.push_back(k);
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

std::vector<int> row_heads(const std::vector<int>& F_ptr, const std::vector<int>& F_col) {
  const int nb = static_cast<int>(F_ptr.size()) - 1;
  std::vector<int> h(static_cast<size_t>(std::max(nb, 1)) * PCG_HEAD_STRIDE, 0);
  for (int r = 0; r < nb; ++r) {
    int* row = h.data() + static_cast<size_t>(r) * PCG_HEAD_STRIDE;
    const int p0 = F_ptr[r], len = F_ptr[r + 1] - p0;
    row[0] = p0;
    row[1] = len;
    for (int t = 0; t < PCG_HEAD_COLS; ++t) row[2 + t] = t < len ? F_col[p0 + t] : r;
  }
  return h;
}

}  // namespace g2ohip
