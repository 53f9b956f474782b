// g2o text-format and BAL input/output for the path's vertex and edge types
// (host code only).
//
// The reference reads a file line by line through a stringstream and a type
// factory (OptimizableGraph::load, g2o/core/optimizable_graph.cpp:344-521).
// Here the file is read whole, cut at line boundaries into one piece per host
// thread, and every piece is tokenised into flat records in parallel. The
// records are then applied in line order on one thread, so the rules that
// depend on order (a FIX or an edge may only name vertices defined on earlier
// lines, the first of two vertices with one id wins) hold exactly as in the
// reference, whatever the number of threads.
#include "../../include/g2o_hip.h"

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <set>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

struct g2ohip_graph {
  // vertex id -> (kind, index in that kind's arrays)
  std::unordered_map<int, std::pair<int, int>> vmap;
  std::vector<int> cam_id, pt_id, pose_id;
  std::vector<double> cam, pt, pose;  // cam: world-to-camera [t q]; pose: Isometry3 [t q]
  std::vector<uint8_t> cam_fix, pt_fix, pose_fix;
  // projection edges (indices into the camera / point arrays); pe_par: the
  // CameraParameters of an EDGE_PROJECT_XYZ2UV edge, -1 for EDGE_SE3_PROJECT_XYZ
  std::vector<int> pe_cam, pe_pt, pe_par;
  // PARAMS_CAMERAPARAMETERS: id -> index; focal length, principal point, baseline
  std::unordered_map<int, int> pmap;
  std::vector<int> par_id;
  std::vector<double> par;
  std::vector<double> pe_meas, pe_omega;  // 2, 3 (upper triangle)
  // pose edges
  std::vector<int> se_from, se_to;
  std::vector<double> se_meas, se_info;  // 7, 36
  // every edge in insertion order: (kind, index)
  std::vector<std::pair<int, int>> edge_order;
  long long lines = 0, unknown_lines = 0, skipped = 0;
  // BAL part
  std::vector<double> bal_cam, bal_pt, bal_meas;
  std::vector<int> bal_cam_idx, bal_pt_idx;
  std::string log;
  mutable std::string err;  // last failure, also of calls on a const graph (save)
  long long log_count = 0;
};

namespace {

enum Kind {
  K_NONE = 0, K_CAM = 1, K_PT = 2, K_POSE = 3, K_PROJ = 4, K_PEDGE = 5, K_FIX = 6, K_PARAM = 7, K_PROJ2 = 8,
  K_UNKNOWN = 9
};

const char* const kTag[] = {"",
                            "VERTEX_SE3:EXPMAP",
                            "VERTEX_TRACKXYZ",
                            "VERTEX_SE3:QUAT",
                            "EDGE_SE3_PROJECT_XYZ:EXPMAP",
                            "EDGE_SE3:QUAT",
                            "FIX",
                            "PARAMS_CAMERAPARAMETERS",   // parameter_cameraparameters.cpp:43-55
                            "EDGE_PROJECT_XYZ2UV:EXPMAP",  // edge_project_xyz2uv.cpp:37-47
                            ""};
// values after the ids
const int kVals[] = {0, 7, 3, 7, 5, 28, 0, 4, 5, 0};
const int kMaxLog = 1000;

int fail(const g2ohip_graph* g, int code, const std::string& msg) {
  g->err = msg;
  return code;
}

// status of an exception caught at the ABI: no exception leaves an entry point
int fail_caught(const g2ohip_graph* g, bool oom) noexcept {
  const int code = oom ? G2OHIP_ERR_OOM : G2OHIP_ERR_INVALID;
  if (!g) return code;
  try {
    g->err = oom ? "out of host memory" : "unexpected host exception";
  } catch (...) {
    g->err.clear();
  }
  return code;
}

void note(g2ohip_graph* g, const std::string& msg) {
  if (g->log_count++ < kMaxLog) {
    g->log += msg;
    g->log += '\n';
  } else if (g->log_count == kMaxLog + 1) {
    g->log += "(further reports suppressed)\n";
  }
}

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// One parsed line. Values live in the piece's flat arrays from `off`.
struct Rec {
  long long line;  // line number within the piece (0-based)
  int kind;
  int bad;         // too few or malformed values
  int a, b;        // vertex / parameter id, or the edge's two vertex ids
  int c;           // EDGE_PROJECT_XYZ2UV: its parameter id
  long long off;   // into vals (or ints for FIX, or names for K_UNKNOWN)
  int n;
};

struct Piece {
  std::vector<Rec> recs;
  std::vector<double> vals;
  std::vector<int> ints;
  std::vector<std::string> names;
  long long lines = 0;
};

inline const char* skip_ws(const char* p, const char* e) {
  while (p < e && is_space(*p)) ++p;
  return p;
}

// std::istream >> accepts a leading '+', std::from_chars does not.
inline bool parse_double(const char*& p, const char* e, double& v) {
  p = skip_ws(p, e);
  if (p < e && *p == '+') ++p;
  auto r = std::from_chars(p, e, v);
  if (r.ec != std::errc() && r.ec != std::errc::result_out_of_range) return false;
  if (r.ptr < e && !is_space(*r.ptr)) return false;
  p = r.ptr;
  return true;
}

inline bool parse_int(const char*& p, const char* e, int& v) {
  p = skip_ws(p, e);
  if (p < e && *p == '+') ++p;
  auto r = std::from_chars(p, e, v);
  if (r.ec != std::errc()) return false;
  if (r.ptr < e && !is_space(*r.ptr)) return false;
  p = r.ptr;
  return true;
}

int tag_kind(const char* t, size_t n) {
  for (int k = 1; k < K_UNKNOWN; ++k)
    if (std::strlen(kTag[k]) == n && std::memcmp(kTag[k], t, n) == 0) return k;
  return K_UNKNOWN;
}

void parse_piece(const char* b, const char* e, Piece& out) {
  const char* p = b;
  long long line = 0;
  while (p < e) {
    const char* le = static_cast<const char*>(std::memchr(p, '\n', e - p));
    if (!le) le = e;
    const char* q = skip_ws(p, le);
    if (q < le && *q != '#') {
      const char* t0 = q;
      while (q < le && !is_space(*q)) ++q;
      Rec r{line, tag_kind(t0, q - t0), 0, -1, -1, -1, 0, 0};
      if (r.kind == K_UNKNOWN) {
        r.off = (long long)out.names.size();
        out.names.emplace_back(t0, q - t0);
      } else if (r.kind == K_FIX) {
        r.off = (long long)out.ints.size();
        int id;
        // "while (currentLine >> id)": stops at the first token that is not an integer
        while (parse_int(q, le, id)) out.ints.push_back(id);
        r.n = (int)(out.ints.size() - r.off);
      } else {
        const bool edge = r.kind == K_PROJ || r.kind == K_PEDGE || r.kind == K_PROJ2;
        if (!parse_int(q, le, r.a) || (edge && !parse_int(q, le, r.b))) r.bad = 1;
        // readParamIds (base_edge.h:187-194) precedes the measurement
        if (!r.bad && r.kind == K_PROJ2 && !parse_int(q, le, r.c)) r.bad = 1;
        r.off = (long long)out.vals.size();
        r.n = kVals[r.kind];
        out.vals.resize(r.off + r.n, 0.0);
        for (int i = 0; i < r.n && !r.bad; ++i)
          if (!parse_double(q, le, out.vals[r.off + i])) r.bad = 1;
      }
      out.recs.push_back(r);
    }
    ++line;
    p = le + 1;
  }
  out.lines = line;
}

bool read_file(const char* filename, std::vector<char>& buf, std::string& err) {
  FILE* f = std::fopen(filename, "rb");
  if (!f) {
    err = std::string("Unable to open file ") + filename;
    return false;
  }
  std::fseek(f, 0, SEEK_END);
  long long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (n < 0) {
    std::fclose(f);
    err = std::string("Unable to read file ") + filename;
    return false;
  }
  buf.resize((size_t)n);
  size_t got = n ? std::fread(buf.data(), 1, (size_t)n, f) : 0;
  std::fclose(f);
  if ((long long)got != n) {
    err = std::string("Unable to read file ") + filename;
    return false;
  }
  return true;
}

int pick_threads(int nthreads, size_t bytes) {
  int n = nthreads > 0 ? nthreads : (int)std::thread::hardware_concurrency();
  n = std::max(1, std::min(n, 16));
  // pieces of at least 1 MiB
  return (int)std::max<size_t>(1, std::min<size_t>((size_t)n, bytes >> 20));
}

// Cut [0, n) into `parts` pieces that end just after a separator.
template <typename IsSep>
std::vector<size_t> cut(const std::vector<char>& buf, int parts, IsSep sep) {
  std::vector<size_t> c{0};
  const size_t n = buf.size();
  for (int i = 1; i < parts; ++i) {
    size_t k = std::max(c.back(), n * (size_t)i / (size_t)parts);
    while (k < n && !sep(buf[k])) ++k;
    c.push_back(std::min(n, k + (k < n ? 1 : 0)));
  }
  c.push_back(n);
  return c;
}

// SE3Quat::normalizeRotation (se3quat.h:251-256): w >= 0, then unit length.
inline void normalize_rotation(double* q) {
  if (q[3] < 0)
    for (int i = 0; i < 4; ++i) q[i] = -q[i];
  const double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  for (int i = 0; i < 4; ++i) q[i] /= n;
}

// q * v for a unit quaternion q = [x y z w] (the cross-product form
// v + 2w (u x v) + 2 u x (u x v)).
inline void rotate(const double* q, const double* v, double* o) {
  double uv[3] = {q[1] * v[2] - q[2] * v[1], q[2] * v[0] - q[0] * v[2], q[0] * v[1] - q[1] * v[0]};
  for (double& x : uv) x += x;
  const double c[3] = {q[1] * uv[2] - q[2] * uv[1], q[2] * uv[0] - q[0] * uv[2], q[0] * uv[1] - q[1] * uv[0]};
  for (int i = 0; i < 3; ++i) o[i] = v[i] + q[3] * uv[i] + c[i];
}

// SE3Quat::inverse (se3quat.h:114-119): [t q] -> [-(q^-1 t), q^-1].
inline void se3_inverse(const double* in, double* out) {
  const double qc[4] = {-in[3], -in[4], -in[5], in[6]};
  const double mt[3] = {-in[0], -in[1], -in[2]};
  rotate(qc, mt, out);
  for (int i = 0; i < 4; ++i) out[3 + i] = qc[i];
}

// Quaternion -> rotation matrix -> quaternion, as toVectorQT
// (isometry3d_mappings.cpp:111-121) writes an Isometry3: the rotation matrix of
// the quaternion (which assumes unit length), converted back by the
// trace / largest-diagonal method (Shepperd), then normalised. This fixes the
// sign the reference prints: w > 0 when the trace is positive, else the
// largest-diagonal component.
void quat_via_matrix(const double* q, double* o) {
  const double x = q[0], y = q[1], z = q[2], w = q[3];
  const double tx = 2 * x, ty = 2 * y, tz = 2 * z;
  const double twx = tx * w, twy = ty * w, twz = tz * w;
  const double txx = tx * x, txy = ty * x, txz = tz * x;
  const double tyy = ty * y, tyz = tz * y, tzz = tz * z;
  const double m[3][3] = {{1 - (tyy + tzz), txy - twz, txz + twy},
                          {txy + twz, 1 - (txx + tzz), tyz - twx},
                          {txz - twy, tyz + twx, 1 - (txx + tyy)}};
  double r[4];  // x y z w
  const double tr = m[0][0] + m[1][1] + m[2][2];
  if (tr > 0) {
    double t = std::sqrt(tr + 1.0);
    r[3] = 0.5 * t;
    t = 0.5 / t;
    r[0] = (m[2][1] - m[1][2]) * t;
    r[1] = (m[0][2] - m[2][0]) * t;
    r[2] = (m[1][0] - m[0][1]) * t;
  } else {
    int i = 0;
    if (m[1][1] > m[0][0]) i = 1;
    if (m[2][2] > m[i][i]) i = 2;
    const int j = (i + 1) % 3, k = (j + 1) % 3;
    double t = std::sqrt(m[i][i] - m[j][j] - m[k][k] + 1.0);
    r[i] = 0.5 * t;
    t = 0.5 / t;
    r[3] = (m[k][j] - m[j][k]) * t;
    r[j] = (m[j][i] + m[i][j]) * t;
    r[k] = (m[k][i] + m[i][k]) * t;
  }
  const double n = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + r[3] * r[3]);
  for (int i = 0; i < 4; ++i) o[i] = r[i] / n;
}

std::vector<int> order_by_id(const std::vector<int>& ids) {
  std::vector<int> o(ids.size());
  for (size_t i = 0; i < o.size(); ++i) o[i] = (int)i;
  std::sort(o.begin(), o.end(), [&](int a, int b) { return ids[a] < ids[b]; });
  return o;
}

std::vector<int> rank_of(const std::vector<int>& order) {
  std::vector<int> r(order.size());
  for (size_t i = 0; i < order.size(); ++i) r[order[i]] = (int)i;
  return r;
}

int next_free_id(const g2ohip_graph* g) {
  int m = -1;
  for (const auto& kv : g->vmap) m = std::max(m, kv.first);
  return m + 1;
}

void apply_piece(g2ohip_graph* g, const Piece& pc, long long line0, std::set<std::string>& warned) {
  char buf[256];
  for (const Rec& r : pc.recs) {
    const long long line = line0 + r.line + 1;  // 1-based, as the reference counts
    const double* v = pc.vals.data() + r.off;
    switch (r.kind) {
      case K_UNKNOWN: {
        ++g->unknown_lines;
        const std::string& t = pc.names[r.off];
        if (warned.insert(t).second) note(g, "Unknown type " + t + " (not one of the path's types)");
        break;
      }
      case K_FIX:
        for (int i = 0; i < r.n; ++i) {
          const int id = pc.ints[r.off + i];
          auto it = g->vmap.find(id);
          if (it == g->vmap.end()) {
            std::snprintf(buf, sizeof buf, "Unable to fix vertex with id %d. Not found in the graph.", id);
            note(g, buf);
            continue;
          }
          const int k = it->second.first, j = it->second.second;
          (k == K_CAM ? g->cam_fix : k == K_PT ? g->pt_fix : g->pose_fix)[j] = 1;
        }
        break;
      case K_PARAM: {
        if (r.bad) {
          ++g->skipped;
          std::snprintf(buf, sizeof buf, "reading data %d for parameter %s at line %lld", r.a, kTag[K_PARAM], line);
          note(g, buf);
          break;
        }
        if (g->pmap.count(r.a)) {
          ++g->skipped;
          std::snprintf(buf, sizeof buf, "Parameter of type: %s id: %d already defined at line %lld", kTag[K_PARAM],
                        r.a, line);
          note(g, buf);
          break;
        }
        g->pmap[r.a] = (int)g->par_id.size();
        g->par_id.push_back(r.a);
        g->par.insert(g->par.end(), v, v + 4);
        break;
      }
      case K_CAM:
      case K_PT:
      case K_POSE: {
        if (r.bad) {
          ++g->skipped;
          std::snprintf(buf, sizeof buf, "Error reading vertex %s %d at line %lld", kTag[r.kind], r.a, line);
          note(g, buf);
          break;
        }
        if (r.a < 0) {
          ++g->skipped;
          std::snprintf(buf, sizeof buf, "FATAL, a vertex with (negative) ID %d cannot be inserted into the graph",
                        r.a);
          note(g, buf);
          break;
        }
        if (g->vmap.count(r.a)) {
          ++g->skipped;
          std::snprintf(buf, sizeof buf, "a vertex with ID %d has already been registered with this graph", r.a);
          note(g, buf);
          break;
        }
        if (r.kind == K_CAM) {
          // vertex_se3_expmap.cpp:35-40: SE3Quat(est).inverse()
          double f[7];
          std::memcpy(f, v, sizeof f);
          normalize_rotation(f + 3);
          double c[7];
          se3_inverse(f, c);
          g->vmap[r.a] = {K_CAM, (int)g->cam_id.size()};
          g->cam_id.push_back(r.a);
          g->cam.insert(g->cam.end(), c, c + 7);
          g->cam_fix.push_back(0);
        } else if (r.kind == K_PT) {
          g->vmap[r.a] = {K_PT, (int)g->pt_id.size()};
          g->pt_id.push_back(r.a);
          g->pt.insert(g->pt.end(), v, v + 3);
          g->pt_fix.push_back(0);
        } else {
          g->vmap[r.a] = {K_POSE, (int)g->pose_id.size()};
          g->pose_id.push_back(r.a);
          g->pose.insert(g->pose.end(), v, v + 7);
          g->pose_fix.push_back(0);
        }
        break;
      }
      case K_PROJ:
      case K_PROJ2:
      case K_PEDGE: {
        auto ia = g->vmap.find(r.a), ib = g->vmap.find(r.b);
        if (!r.bad && (ia == g->vmap.end() || ib == g->vmap.end())) {
          ++g->skipped;
          std::snprintf(buf, sizeof buf, "Unable to find vertices for edge %s at line %lld IDs: %d %d", kTag[r.kind],
                        line, r.a, r.b);
          note(g, buf);
          break;
        }
        const bool proj = r.kind == K_PROJ || r.kind == K_PROJ2;
        // an XYZ2UV edge resolves its CameraParameters when it is added
        // (optimizable_graph.cpp:232-236): the parameter must come first
        auto ip = g->pmap.find(r.c);
        const bool types_ok = !r.bad &&
                              (proj ? (ia->second.first == K_PT && ib->second.first == K_CAM)
                                    : (ia->second.first == K_POSE && ib->second.first == K_POSE && r.a != r.b)) &&
                              (r.kind != K_PROJ2 || ip != g->pmap.end());
        if (!types_ok) {
          ++g->skipped;
          std::snprintf(buf, sizeof buf, "Unable to add edge %s at line %lld IDs: %d %d", kTag[r.kind], line, r.a,
                        r.b);
          note(g, buf);
          break;
        }
        if (proj) {
          g->edge_order.push_back({K_PROJ, (int)g->pe_cam.size()});
          g->pe_pt.push_back(ia->second.second);
          g->pe_cam.push_back(ib->second.second);
          g->pe_par.push_back(r.kind == K_PROJ2 ? ip->second : -1);
          g->pe_meas.insert(g->pe_meas.end(), v, v + 2);
          g->pe_omega.insert(g->pe_omega.end(), v + 2, v + 5);
        } else {
          g->edge_order.push_back({K_PEDGE, (int)g->se_from.size()});
          g->se_from.push_back(ia->second.second);
          g->se_to.push_back(ib->second.second);
          // edge_se3.cpp:43-53: the quaternion is normalised
          double m[7];
          std::memcpy(m, v, sizeof m);
          const double n = std::sqrt(m[3] * m[3] + m[4] * m[4] + m[5] * m[5] + m[6] * m[6]);
          for (int i = 3; i < 7; ++i) m[i] /= n;
          g->se_meas.insert(g->se_meas.end(), m, m + 7);
          // base_edge.h:173-180: upper triangle, mirrored
          double I[36];
          int t = 7;
          for (int i = 0; i < 6; ++i)
            for (int j = i; j < 6; ++j, ++t) I[i * 6 + j] = I[j * 6 + i] = v[t];
          g->se_info.insert(g->se_info.end(), I, I + 36);
        }
        break;
      }
      default:
        break;
    }
  }
}

template <typename F>
void run_threads(int n, F f) {
  // an exception may not leave a worker thread (std::terminate): each worker
  // records a failed allocation, and the caller sees it as std::bad_alloc
  // after every thread has joined
  std::vector<char> failed(n, 0);
  auto guarded = [&](int i) {
    try {
      f(i);
    } catch (const std::bad_alloc&) {
      failed[i] = 1;
    }
  };
  std::vector<std::thread> th;
  for (int i = 1; i < n; ++i) th.emplace_back(guarded, i);
  guarded(0);
  for (auto& t : th) t.join();
  for (char x : failed)
    if (x) throw std::bad_alloc();
}

// Each value as "%.{prec}g" followed by a space: what writeVector
// (io_helper.h:35-39) prints through an ostream at that precision.
void put_vals(std::string& s, const double* v, int n, int prec) {
  char b[40];
  for (int i = 0; i < n; ++i) {
    auto r = std::to_chars(b, b + sizeof b - 1, v[i], std::chars_format::general, prec);
    *r.ptr++ = ' ';
    s.append(b, r.ptr);
  }
}

}  // namespace

extern "C" {

int g2ohip_graph_create(g2ohip_graph** out) {
  if (!out) return G2OHIP_ERR_INVALID;
  *out = new (std::nothrow) g2ohip_graph();
  return *out ? G2OHIP_OK : G2OHIP_ERR_OOM;
}

void g2ohip_graph_destroy(g2ohip_graph* g) { delete g; }

const char* g2ohip_graph_last_error(const g2ohip_graph* g) { return g ? g->err.c_str() : "null graph"; }

const char* g2ohip_graph_log(const g2ohip_graph* g) { return g ? g->log.c_str() : ""; }

int g2ohip_graph_load(g2ohip_graph* g, const char* filename, int nthreads) try {
  if (!g || !filename) return G2OHIP_ERR_INVALID;
  std::vector<char> buf;
  if (!read_file(filename, buf, g->err)) return G2OHIP_ERR_IO;
  const int nt = pick_threads(nthreads, buf.size());
  const std::vector<size_t> c = cut(buf, nt, [](char ch) { return ch == '\n'; });
  std::vector<Piece> pieces(nt);
  try {
    run_threads(nt, [&](int i) { parse_piece(buf.data() + c[i], buf.data() + c[i + 1], pieces[i]); });
  } catch (const std::bad_alloc&) {
    return fail(g, G2OHIP_ERR_OOM, "out of host memory while parsing");
  }
  buf.clear();
  buf.shrink_to_fit();
  std::set<std::string> warned;
  long long line0 = g->lines;
  try {
    for (int i = 0; i < nt; ++i) {
      apply_piece(g, pieces[i], line0, warned);
      line0 += pieces[i].lines;
      std::vector<Rec>().swap(pieces[i].recs);
      std::vector<double>().swap(pieces[i].vals);
    }
  } catch (const std::bad_alloc&) {
    g->lines = line0;
    return fail(g, G2OHIP_ERR_OOM, "out of host memory while loading (the graph holds the lines before it)");
  }
  g->lines = line0;
  return G2OHIP_OK;
} catch (const std::bad_alloc&) {
  return fail_caught(g, true);
} catch (...) {
  return fail_caught(g, false);
}

int g2ohip_graph_info(const g2ohip_graph* g, long long* info) try {
  if (!g || !info) return G2OHIP_ERR_INVALID;
  long long nfix = 0;
  for (auto f : g->cam_fix) nfix += f;
  for (auto f : g->pt_fix) nfix += f;
  for (auto f : g->pose_fix) nfix += f;
  long long n2uv = 0;
  for (int p : g->pe_par) n2uv += p >= 0;
  const long long v[12] = {(long long)g->cam_id.size(), (long long)g->pt_id.size(), (long long)g->pe_cam.size(),
                           (long long)g->pose_id.size(), (long long)g->se_from.size(), nfix, g->lines,
                           g->unknown_lines, g->skipped, (long long)g->bal_cam_idx.size(), n2uv,
                           (long long)g->par_id.size()};
  std::memcpy(info, v, sizeof v);
  return G2OHIP_OK;
} catch (const std::bad_alloc&) {
  return fail_caught(g, true);
} catch (...) {
  return fail_caught(g, false);
}

int g2ohip_graph_get_ba(const g2ohip_graph* g, int* cam_id, double* cams, uint8_t* cam_fixed, int* pt_id,
                        double* pts, uint8_t* pt_fixed, int* edge_cam, int* edge_pt, double* meas, double* omega) try {
  if (!g) return G2OHIP_ERR_INVALID;
  const std::vector<int> oc = order_by_id(g->cam_id), op = order_by_id(g->pt_id);
  for (size_t i = 0; i < oc.size(); ++i) {
    if (cam_id) cam_id[i] = g->cam_id[oc[i]];
    if (cams) std::memcpy(cams + 7 * i, &g->cam[7 * (size_t)oc[i]], 7 * sizeof(double));
    if (cam_fixed) cam_fixed[i] = g->cam_fix[oc[i]];
  }
  for (size_t i = 0; i < op.size(); ++i) {
    if (pt_id) pt_id[i] = g->pt_id[op[i]];
    if (pts) std::memcpy(pts + 3 * i, &g->pt[3 * (size_t)op[i]], 3 * sizeof(double));
    if (pt_fixed) pt_fixed[i] = g->pt_fix[op[i]];
  }
  const std::vector<int> rc = rank_of(oc), rp = rank_of(op);
  const size_t E = g->pe_cam.size();
  for (size_t e = 0; e < E; ++e) {
    if (edge_cam) edge_cam[e] = rc[g->pe_cam[e]];
    if (edge_pt) edge_pt[e] = rp[g->pe_pt[e]];
  }
  if (meas && E) std::memcpy(meas, g->pe_meas.data(), 2 * E * sizeof(double));
  if (omega && E) std::memcpy(omega, g->pe_omega.data(), 3 * E * sizeof(double));
  return G2OHIP_OK;
} catch (const std::bad_alloc&) {
  return fail_caught(g, true);
} catch (...) {
  return fail_caught(g, false);
}

int g2ohip_graph_get_ba_params(const g2ohip_graph* g, int* param_id, double* params, int* edge_param) try {
  if (!g) return G2OHIP_ERR_INVALID;
  const std::vector<int> o = order_by_id(g->par_id);
  const std::vector<int> rk = rank_of(o);
  for (size_t i = 0; i < o.size(); ++i) {
    if (param_id) param_id[i] = g->par_id[o[i]];
    if (params) std::memcpy(params + 4 * i, &g->par[4 * (size_t)o[i]], 4 * sizeof(double));
  }
  if (edge_param)
    for (size_t e = 0; e < g->pe_par.size(); ++e) edge_param[e] = g->pe_par[e] < 0 ? -1 : rk[g->pe_par[e]];
  return G2OHIP_OK;
} catch (const std::bad_alloc&) {
  return fail_caught(g, true);
} catch (...) {
  return fail_caught(g, false);
}

int g2ohip_graph_get_pg(const g2ohip_graph* g, int* pose_id, double* poses, uint8_t* fixed, int* edge_from,
                        int* edge_to, double* meas, double* info) try {
  if (!g) return G2OHIP_ERR_INVALID;
  const std::vector<int> o = order_by_id(g->pose_id);
  for (size_t i = 0; i < o.size(); ++i) {
    if (pose_id) pose_id[i] = g->pose_id[o[i]];
    if (poses) std::memcpy(poses + 7 * i, &g->pose[7 * (size_t)o[i]], 7 * sizeof(double));
    if (fixed) fixed[i] = g->pose_fix[o[i]];
  }
  const std::vector<int> r = rank_of(o);
  const size_t E = g->se_from.size();
  for (size_t e = 0; e < E; ++e) {
    if (edge_from) edge_from[e] = r[g->se_from[e]];
    if (edge_to) edge_to[e] = r[g->se_to[e]];
  }
  if (meas && E) std::memcpy(meas, g->se_meas.data(), 7 * E * sizeof(double));
  if (info && E) std::memcpy(info, g->se_info.data(), 36 * E * sizeof(double));
  return G2OHIP_OK;
} catch (const std::bad_alloc&) {
  return fail_caught(g, true);
} catch (...) {
  return fail_caught(g, false);
}

int g2ohip_graph_set_ba_state(g2ohip_graph* g, const double* cams, const double* pts) try {
  if (!g) return G2OHIP_ERR_INVALID;
  if (cams) {
    const std::vector<int> o = order_by_id(g->cam_id);
    for (size_t i = 0; i < o.size(); ++i) {
      double* c = &g->cam[7 * (size_t)o[i]];
      std::memcpy(c, cams + 7 * i, 7 * sizeof(double));
      normalize_rotation(c + 3);
    }
  }
  if (pts) {
    const std::vector<int> o = order_by_id(g->pt_id);
    for (size_t i = 0; i < o.size(); ++i) std::memcpy(&g->pt[3 * (size_t)o[i]], pts + 3 * i, 3 * sizeof(double));
  }
  return G2OHIP_OK;
} catch (const std::bad_alloc&) {
  return fail_caught(g, true);
} catch (...) {
  return fail_caught(g, false);
}

int g2ohip_graph_set_pg_state(g2ohip_graph* g, const double* poses) try {
  if (!g) return G2OHIP_ERR_INVALID;
  if (poses) {
    const std::vector<int> o = order_by_id(g->pose_id);
    for (size_t i = 0; i < o.size(); ++i) std::memcpy(&g->pose[7 * (size_t)o[i]], poses + 7 * i, 7 * sizeof(double));
  }
  return G2OHIP_OK;
} catch (const std::bad_alloc&) {
  return fail_caught(g, true);
} catch (...) {
  return fail_caught(g, false);
}

int g2ohip_graph_add_ba(g2ohip_graph* g, int num_cams, int num_pts, int num_edges, const int* cam_id,
                        const double* cams, const uint8_t* cam_fixed, const int* pt_id, const double* pts,
                        const uint8_t* pt_fixed, const int* edge_cam, const int* edge_pt, const double* meas,
                        const double* omega) try {
  if (!g) return G2OHIP_ERR_INVALID;
  if (num_cams < 0 || num_pts < 0 || num_edges < 0 || (num_cams && !cams) || (num_pts && !pts) ||
      (num_edges && (!edge_cam || !edge_pt || !meas)))
    return fail(g, G2OHIP_ERR_INVALID, "graph_add_ba: negative size or missing array");
  const int base = next_free_id(g);
  std::vector<int> ci(num_cams), pi(num_pts);
  std::set<int> seen;
  for (int i = 0; i < num_cams; ++i) ci[i] = cam_id ? cam_id[i] : base + i;
  for (int i = 0; i < num_pts; ++i) pi[i] = pt_id ? pt_id[i] : base + num_cams + i;
  for (int id : ci)
    if (id < 0 || g->vmap.count(id) || !seen.insert(id).second)
      return fail(g, G2OHIP_ERR_INVALID, "graph_add_ba: negative or duplicate vertex id " + std::to_string(id));
  for (int id : pi)
    if (id < 0 || g->vmap.count(id) || !seen.insert(id).second)
      return fail(g, G2OHIP_ERR_INVALID, "graph_add_ba: negative or duplicate vertex id " + std::to_string(id));
  for (int e = 0; e < num_edges; ++e)
    if (edge_cam[e] < 0 || edge_cam[e] >= num_cams || edge_pt[e] < 0 || edge_pt[e] >= num_pts)
      return fail(g, G2OHIP_ERR_INVALID, "graph_add_ba: edge " + std::to_string(e) + " vertex index out of range");
  const int c0 = (int)g->cam_id.size(), p0 = (int)g->pt_id.size();
  for (int i = 0; i < num_cams; ++i) {
    g->vmap[ci[i]] = {K_CAM, c0 + i};
    g->cam_id.push_back(ci[i]);
    double c[7];
    std::memcpy(c, cams + 7 * (size_t)i, sizeof c);
    normalize_rotation(c + 3);
    g->cam.insert(g->cam.end(), c, c + 7);
    g->cam_fix.push_back(cam_fixed ? (cam_fixed[i] != 0) : 0);
  }
  for (int i = 0; i < num_pts; ++i) {
    g->vmap[pi[i]] = {K_PT, p0 + i};
    g->pt_id.push_back(pi[i]);
    g->pt.insert(g->pt.end(), pts + 3 * (size_t)i, pts + 3 * (size_t)i + 3);
    g->pt_fix.push_back(pt_fixed ? (pt_fixed[i] != 0) : 0);
  }
  for (int e = 0; e < num_edges; ++e) {
    g->edge_order.push_back({K_PROJ, (int)g->pe_cam.size()});
    g->pe_cam.push_back(c0 + edge_cam[e]);
    g->pe_pt.push_back(p0 + edge_pt[e]);
    g->pe_par.push_back(-1);
    g->pe_meas.insert(g->pe_meas.end(), meas + 2 * (size_t)e, meas + 2 * (size_t)e + 2);
    const double I[3] = {1.0, 0.0, 1.0};
    const double* o = omega ? omega + 3 * (size_t)e : I;
    g->pe_omega.insert(g->pe_omega.end(), o, o + 3);
  }
  return G2OHIP_OK;
} catch (const std::bad_alloc&) {
  return fail_caught(g, true);
} catch (...) {
  return fail_caught(g, false);
}

int g2ohip_graph_add_pg(g2ohip_graph* g, int num_poses, int num_edges, const int* pose_id, const double* poses,
                        const uint8_t* fixed, const int* edge_from, const int* edge_to, const double* meas,
                        const double* info) try {
  if (!g) return G2OHIP_ERR_INVALID;
  if (num_poses < 0 || num_edges < 0 || (num_poses && !poses) || (num_edges && (!edge_from || !edge_to || !meas)))
    return fail(g, G2OHIP_ERR_INVALID, "graph_add_pg: negative size or missing array");
  const int base = next_free_id(g);
  std::vector<int> ids(num_poses);
  std::set<int> seen;
  for (int i = 0; i < num_poses; ++i) {
    ids[i] = pose_id ? pose_id[i] : base + i;
    if (ids[i] < 0 || g->vmap.count(ids[i]) || !seen.insert(ids[i]).second)
      return fail(g, G2OHIP_ERR_INVALID, "graph_add_pg: negative or duplicate vertex id " + std::to_string(ids[i]));
  }
  for (int e = 0; e < num_edges; ++e)
    if (edge_from[e] < 0 || edge_from[e] >= num_poses || edge_to[e] < 0 || edge_to[e] >= num_poses ||
        edge_from[e] == edge_to[e])
      return fail(g, G2OHIP_ERR_INVALID, "graph_add_pg: edge " + std::to_string(e) + " has bad vertex indices");
  const int n0 = (int)g->pose_id.size();
  for (int i = 0; i < num_poses; ++i) {
    g->vmap[ids[i]] = {K_POSE, n0 + i};
    g->pose_id.push_back(ids[i]);
    g->pose.insert(g->pose.end(), poses + 7 * (size_t)i, poses + 7 * (size_t)i + 7);
    g->pose_fix.push_back(fixed ? (fixed[i] != 0) : 0);
  }
  for (int e = 0; e < num_edges; ++e) {
    g->edge_order.push_back({K_PEDGE, (int)g->se_from.size()});
    g->se_from.push_back(n0 + edge_from[e]);
    g->se_to.push_back(n0 + edge_to[e]);
    g->se_meas.insert(g->se_meas.end(), meas + 7 * (size_t)e, meas + 7 * (size_t)e + 7);
    double I[36] = {0};
    for (int i = 0; i < 6; ++i) I[7 * i] = 1.0;
    const double* o = info ? info + 36 * (size_t)e : I;
    g->se_info.insert(g->se_info.end(), o, o + 36);
  }
  return G2OHIP_OK;
} catch (const std::bad_alloc&) {
  return fail_caught(g, true);
} catch (...) {
  return fail_caught(g, false);
}

int g2ohip_graph_save(const g2ohip_graph* g, const char* filename, int precision) try {
  if (!g || !filename) return G2OHIP_ERR_INVALID;
  if (precision < 1 || precision > 17) return fail(g, G2OHIP_ERR_INVALID, "graph_save: precision must be 1..17");
  // vertices of the edges, sorted by id (optimizable_graph.cpp:541-554)
  std::vector<std::pair<int, std::pair<int, int>>> vs;
  {
    std::vector<uint8_t> uc(g->cam_id.size()), up(g->pt_id.size()), uq(g->pose_id.size());
    for (size_t e = 0; e < g->pe_cam.size(); ++e) uc[g->pe_cam[e]] = up[g->pe_pt[e]] = 1;
    for (size_t e = 0; e < g->se_from.size(); ++e) uq[g->se_from[e]] = uq[g->se_to[e]] = 1;
    for (size_t i = 0; i < uc.size(); ++i)
      if (uc[i]) vs.push_back({g->cam_id[i], {K_CAM, (int)i}});
    for (size_t i = 0; i < up.size(); ++i)
      if (up[i]) vs.push_back({g->pt_id[i], {K_PT, (int)i}});
    for (size_t i = 0; i < uq.size(); ++i)
      if (uq[i]) vs.push_back({g->pose_id[i], {K_POSE, (int)i}});
    std::sort(vs.begin(), vs.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
  }
  // the parameters first, by id (optimizable_graph.cpp:538-539)
  std::string head;
  {
    const std::vector<int> po = order_by_id(g->par_id);
    char b[64];
    for (int i : po) {
      std::snprintf(b, sizeof b, "%s %d ", kTag[K_PARAM], g->par_id[i]);
      head += b;
      put_vals(head, &g->par[4 * (size_t)i], 4, precision);
      head += '\n';
    }
  }
  // item k < nv is vertex vs[k], else edge edge_order[k - nv]
  const size_t nv = vs.size(), ni = nv + g->edge_order.size();
  auto format = [&](size_t k, std::string& s) {
    char b[64];
    if (k < nv) {
      const auto& v = vs[k];
      const int kind = v.second.first, i = v.second.second;
      std::snprintf(b, sizeof b, "%s %d ", kTag[kind], v.first);
      s += b;
      uint8_t fixed = 0;
      if (kind == K_CAM) {
        // vertex_se3_expmap.cpp:42-44: estimate().inverse().toVector()
        double o[7];
        se3_inverse(&g->cam[7 * (size_t)i], o);
        put_vals(s, o, 7, precision);
        fixed = g->cam_fix[i];
      } else if (kind == K_PT) {
        put_vals(s, &g->pt[3 * (size_t)i], 3, precision);
        fixed = g->pt_fix[i];
      } else {
        const double* p = &g->pose[7 * (size_t)i];
        double o[7] = {p[0], p[1], p[2]};
        quat_via_matrix(p + 3, o + 3);
        put_vals(s, o, 7, precision);
        fixed = g->pose_fix[i];
      }
      s += '\n';
      if (fixed) {
        std::snprintf(b, sizeof b, "FIX %d\n", v.first);
        s += b;
      }
      return;
    }
    const auto& eo = g->edge_order[k - nv];
    const int i = eo.second;
    if (eo.first == K_PROJ) {
      const int pi = g->pe_par[i];
      if (pi < 0)
        std::snprintf(b, sizeof b, "%s %d %d ", kTag[K_PROJ], g->pt_id[g->pe_pt[i]], g->cam_id[g->pe_cam[i]]);
      else
        std::snprintf(b, sizeof b, "%s %d %d %d ", kTag[K_PROJ2], g->pt_id[g->pe_pt[i]], g->cam_id[g->pe_cam[i]],
                      g->par_id[pi]);
      s += b;
      put_vals(s, &g->pe_meas[2 * (size_t)i], 2, precision);
      put_vals(s, &g->pe_omega[3 * (size_t)i], 3, precision);
    } else {
      std::snprintf(b, sizeof b, "%s %d %d ", kTag[K_PEDGE], g->pose_id[g->se_from[i]], g->pose_id[g->se_to[i]]);
      s += b;
      const double* m = &g->se_meas[7 * (size_t)i];
      double o[7] = {m[0], m[1], m[2]};
      quat_via_matrix(m + 3, o + 3);
      put_vals(s, o, 7, precision);
      const double* I = &g->se_info[36 * (size_t)i];
      for (int r = 0; r < 6; ++r) put_vals(s, I + 6 * r + r, 6 - r, precision);
    }
    s += '\n';
  };
  FILE* f = std::fopen(filename, "wb");
  if (!f) return fail(g, G2OHIP_ERR_IO, std::string("Unable to open file ") + filename + " for writing");
  // Format in blocks of items, one block per thread at a time, and write the
  // blocks in order.
  const int nt = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
  const size_t blk = 1 << 16;
  bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size();
  std::vector<std::string> out(nt);
  try {
    for (size_t k0 = 0; k0 < ni && ok; k0 += blk * nt) {
      run_threads(nt, [&](int t) {
        out[t].clear();
        const size_t a = std::min(ni, k0 + blk * t), b = std::min(ni, a + blk);
        for (size_t k = a; k < b; ++k) format(k, out[t]);
      });
      for (int t = 0; t < nt && ok; ++t) ok = std::fwrite(out[t].data(), 1, out[t].size(), f) == out[t].size();
    }
  } catch (const std::bad_alloc&) {
    std::fclose(f);
    return fail(g, G2OHIP_ERR_OOM, "out of host memory while writing");
  }
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(g, G2OHIP_ERR_IO, std::string("Unable to write file ") + filename);
  return G2OHIP_OK;
} catch (const std::bad_alloc&) {
  return fail_caught(g, true);
} catch (...) {
  return fail_caught(g, false);
}

int g2ohip_graph_load_bal(g2ohip_graph* g, const char* filename, int nthreads) try {
  if (!g || !filename) return G2OHIP_ERR_INVALID;
  std::vector<char> buf;
  if (!read_file(filename, buf, g->err)) return G2OHIP_ERR_IO;
  // The format is a flat stream of numbers; every piece tokenises its share.
  const int nt = pick_threads(nthreads, buf.size());
  auto ws = [](char ch) { return ch == '\n' || is_space(ch); };
  const std::vector<size_t> c = cut(buf, nt, ws);
  std::vector<std::vector<double>> tok(nt);
  std::vector<int> bad(nt, 0);
  try {
    run_threads(nt, [&](int i) {
      const char* p = buf.data() + c[i];
      const char* e = buf.data() + c[i + 1];
      tok[i].reserve((e - p) / 8);
      while (true) {
        while (p < e && ws(*p)) ++p;
        if (p >= e) break;
        const char* t = p;
        while (t < e && !ws(*t)) ++t;
        double v;
        const char* q = p;
        if (!parse_double(q, t, v) || q != t) {
          bad[i] = 1;
          break;
        }
        tok[i].push_back(v);
        p = t;
      }
    });
  } catch (const std::bad_alloc&) {
    return fail(g, G2OHIP_ERR_OOM, "out of host memory while parsing");
  }
  std::vector<double> all;
  for (int i = 0; i < nt; ++i)
    if (bad[i]) return fail(g, G2OHIP_ERR_INVALID, std::string("malformed number in BAL file ") + filename);
  try {
    for (int i = 0; i < nt; ++i) {
      all.insert(all.end(), tok[i].begin(), tok[i].end());
      std::vector<double>().swap(tok[i]);
    }
  } catch (const std::bad_alloc&) {
    return fail(g, G2OHIP_ERR_OOM, "out of host memory while parsing");
  }
  auto integral = [](double v) { return v == std::floor(v) && v >= 0 && v < 2147483648.0; };
  if (all.size() < 3 || !integral(all[0]) || !integral(all[1]) || !integral(all[2]))
    return fail(g, G2OHIP_ERR_INVALID, "BAL header must be three non-negative integers");
  const long long C = (long long)all[0], P = (long long)all[1], E = (long long)all[2];
  const long long need = 3 + 4 * E + 9 * C + 3 * P;
  if ((long long)all.size() < need)
    return fail(g, G2OHIP_ERR_INVALID,
                "BAL file truncated: " + std::to_string(all.size()) + " numbers, " + std::to_string(need) + " needed");
  std::vector<int> ec(E), ep(E);
  std::vector<double> meas(2 * E);
  for (long long e = 0; e < E; ++e) {
    const double* r = &all[3 + 4 * e];
    // bal_example.cpp:297-303 asserts both indices are in range
    if (!integral(r[0]) || !integral(r[1]) || r[0] >= C || r[1] >= P)
      return fail(g, G2OHIP_ERR_INVALID, "BAL observation " + std::to_string(e) + " has an index out of range");
    ec[e] = (int)r[0];
    ep[e] = (int)r[1];
    meas[2 * e] = r[2];
    meas[2 * e + 1] = r[3];
  }
  g->bal_cam.assign(all.begin() + 3 + 4 * E, all.begin() + 3 + 4 * E + 9 * C);
  g->bal_pt.assign(all.begin() + 3 + 4 * E + 9 * C, all.begin() + need);
  g->bal_cam_idx.swap(ec);
  g->bal_pt_idx.swap(ep);
  g->bal_meas.swap(meas);
  return G2OHIP_OK;
} catch (const std::bad_alloc&) {
  return fail_caught(g, true);
} catch (...) {
  return fail_caught(g, false);
}

int g2ohip_graph_get_bal(const g2ohip_graph* g, long long* info, double* cams9, double* pts, int* edge_cam,
                         int* edge_pt, double* meas) try {
  if (!g) return G2OHIP_ERR_INVALID;
  const size_t C = g->bal_cam.size() / 9, P = g->bal_pt.size() / 3, E = g->bal_cam_idx.size();
  if (info) {
    info[0] = (long long)C;
    info[1] = (long long)P;
    info[2] = (long long)E;
  }
  if (cams9 && C) std::memcpy(cams9, g->bal_cam.data(), 9 * C * sizeof(double));
  if (pts && P) std::memcpy(pts, g->bal_pt.data(), 3 * P * sizeof(double));
  if (edge_cam && E) std::memcpy(edge_cam, g->bal_cam_idx.data(), E * sizeof(int));
  if (edge_pt && E) std::memcpy(edge_pt, g->bal_pt_idx.data(), E * sizeof(int));
  if (meas && E) std::memcpy(meas, g->bal_meas.data(), 2 * E * sizeof(double));
  return G2OHIP_OK;
} catch (const std::bad_alloc&) {
  return fail_caught(g, true);
} catch (...) {
  return fail_caught(g, false);
}

}  // extern "C"
