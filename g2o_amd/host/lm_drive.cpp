// Runs Levenberg-Marquardt on a bundle-adjustment problem through
// BlockSolverHIP in two ways and prints one line per iteration:
//   solver   - the host loop of OptimizationAlgorithmLevenberg::solve
//              (optimization_algorithm_levenberg.cpp:60-153) written against
//              the g2o::Solver calls (buildSystem, setLambda, solve,
//              restoreDiagonal, x) and the optimizer's state calls, as g2o
//              drives a solver plugin;
//   optimize - the library's own device-resident loop (g2ohip_optimize).
// Usage: lm_drive <problem.bin> <solver|optimize> <iterations> <device>
// problem.bin: int32 C, P, E, precision, refine_steps; f64 cams[C*7],
// intr[C*4], pts[P*3]; int32 edge_cam[E], edge_pt[E]; f64 meas[E*2],
// weight[E]; u8 cam_fixed[C], pt_fixed[P]; f64 huber.
//   or: lm_drive <problem.bal> <solver|optimize|gn> <iterations> <device>
// (BAL: 9-DoF cameras, BlockSolver<9,3>, FP64)
//   or: lm_drive <graph.g2o> <solver|optimize|gn|dogleg> <iterations> <device>
//                [fx fy cx cy] [out.g2o]
// (no intrinsics when every projection edge is an EDGE_PROJECT_XYZ2UV naming
// its PARAMS_CAMERAPARAMETERS)
// reads a g2o file, runs FP64 Levenberg-Marquardt (solver: the host loop;
// optimize: on the device), Gauss-Newton or Dogleg and, as g2o's
// command-line tool does with -o, saves the optimised graph. A bundle-
// adjustment file needs the intrinsics (the format holds none); a file of
// VERTEX_SE3:QUAT / EDGE_SE3:QUAT runs the pose-graph solver.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "g2o_hip.hpp"

namespace {

template <typename T>
bool read_vec(FILE* f, std::vector<T>& v, size_t n) {
  v.resize(n);
  return n == 0 || std::fread(v.data(), sizeof(T), n, f) == n;
}

// OptimizationAlgorithmLevenberg::computeLambdaInit (optimization_algorithm_
// levenberg.cpp:155-167) as g2o runs it: the largest diagonal entry over the
// vertex Hessians the block solver mapped, times tau, or the user's
// initialLambda.
template <typename Solver>
double lambda_init_from_vertices(Solver& s) {
  constexpr int D = Solver::kCamDim;
  if (s.options().user_lambda_init > 0) return s.options().user_lambda_init;
  double maxDiagonal = 0;
  for (int c = 0; c < s.numCameras(); ++c)
    if (const double* H = s.vertexHessian(true, c))
      for (int j = 0; j < D; ++j) maxDiagonal = std::max(std::fabs(H[(D + 1) * j]), maxDiagonal);
  for (int p = 0; p < s.numPoints(); ++p)
    if (const double* H = s.vertexHessian(false, p))
      for (int j = 0; j < 3; ++j) maxDiagonal = std::max(std::fabs(H[4 * j]), maxDiagonal);
  return s.options().tau * maxDiagonal;
}

template <typename Solver>
int run_solver_loop(Solver& s, int iterations) {
  s.init();
  double lambda = 0, ni = 2;
  for (int i = 0; i < iterations; ++i) {
    double currentChi = s.computeActiveErrors();
    const double chiBefore = currentChi;
    if (!s.buildSystem()) return 1;
    if (i == 0) {
      // the host loop takes lambda from the vertex Hessians, as g2o does; the
      // device's own computeLambdaInit must give the same value
      lambda = lambda_init_from_vertices(s);
      const double dev = s.computeLambdaInit();
      std::fprintf(stderr, "lambda_init vertex_hessians %.17g device %.17g\n", lambda, dev);
      if (!(lambda == dev)) return 1;
      ni = 2;
    }
    double rho = 0;
    int q = 0, pcg = 0;
    do {
      s.push();
      s.setLambda(lambda, true);
      const bool ok = s.solve();
      pcg += s.iterationsLinearSolver();
      s.restoreDiagonal();
      double tempChi = std::numeric_limits<double>::max();
      if (ok) {
        s.update();
        tempChi = s.computeActiveErrors();
      }
      rho = currentChi - tempChi;
      const double scale = ok ? s.computeScale() + 1e-3 : 1.0;
      rho /= scale;
      if (rho > 0 && std::isfinite(tempChi) && ok) {
        double alpha = 1. - std::pow((2 * rho - 1), 3);
        alpha = std::min(alpha, 2. / 3.);
        lambda *= std::max(1. / 3., alpha);
        ni = 2;
        currentChi = tempChi;
        s.discardTop();
      } else {
        lambda *= ni;
        ni *= 2;
        s.pop();
        if (!std::isfinite(lambda)) break;
      }
      q++;
    } while (rho < 0 && q < s.options().max_trials);
    std::printf("%d %.17g %.17g %.17g %d %d\n", i, chiBefore, currentChi, lambda, q, pcg);
    if (q == s.options().max_trials || rho == 0 || !std::isfinite(lambda)) break;
  }
  return 0;
}

void print_stats(const std::vector<g2ohip_iter_stats>& st) {
  for (size_t i = 0; i < st.size(); ++i)
    std::printf("%zu %.17g %.17g %.17g %d %d\n", i, st[i].chi2_before, st[i].chi2, st[i].lambda,
                st[i].levenberg_iterations, st[i].iterations_linear_solver);
}

int algorithm_of(const std::string& mode) {
  return mode == "gn" ? G2OHIP_GAUSS_NEWTON : mode == "dogleg" ? G2OHIP_DOGLEG : G2OHIP_LEVENBERG;
}

template <typename Solver>
int run(Solver& s, const char* mode, int iters) {
  if (std::string(mode) == "solver") {
    if (run_solver_loop(s, iters)) {
      std::fprintf(stderr, "solver loop failed: %s\n", s.lastError().c_str());
      return 1;
    }
  } else {
    std::vector<g2ohip_iter_stats> st;
    if (!s.setAlgorithm(algorithm_of(mode)) || s.optimize(iters, &st) < 0) {
      std::fprintf(stderr, "optimize failed: %s\n", s.lastError().c_str());
      return 1;
    }
    print_stats(st);
  }
  return 0;
}

int run_pose_graph_file(g2o_hip::GraphIO& io, int argc, char** argv) {
  g2o_hip::GraphIO::PoseGraph pg;
  if (!io.getPoseGraph(pg)) {
    std::fprintf(stderr, "load failed: %s\n", io.lastError().c_str());
    return 1;
  }
  g2o_hip::PoseGraphSolverHIP s(std::atoi(argv[4]));
  const int N = static_cast<int>(pg.poseId.size()), E = static_cast<int>(pg.edgeFrom.size());
  std::vector<g2ohip_iter_stats> st;
  if (!s.setGraph(N, E, pg.edgeFrom.data(), pg.edgeTo.data(), pg.fixed.data()) ||
      !s.setMeasurements(pg.meas.data(), pg.information.data()) || !s.setEstimates(pg.poses.data()) ||
      !s.setAlgorithm(algorithm_of(argv[2])) || s.optimize(std::atoi(argv[3]), &st) < 0) {
    std::fprintf(stderr, "pose graph failed: %s\n", s.lastError().c_str());
    return 1;
  }
  print_stats(st);
  if (argc > 5) {
    if (!s.getEstimates(pg.poses.data()) || !io.setPoseEstimates(pg.poses.data()) || !io.save(argv[5], 17)) {
      std::fprintf(stderr, "save failed: %s %s\n", s.lastError().c_str(), io.lastError().c_str());
      return 1;
    }
  }
  return 0;
}

// A BAL file (bal_example.cpp:265-335) through BALSolverHIP: the host loop
// (solver) or the device loop (optimize, gn), FP64, nothing fixed and the
// identity information, as bal_example.cpp sets up its graph.
int run_bal_file(int argc, char** argv) {
  g2o_hip::GraphIO io;
  if (!io.loadBAL(argv[1])) {
    std::fprintf(stderr, "cannot read %s: %s\n", argv[1], io.lastError().c_str());
    return 1;
  }
  g2o_hip::GraphIO::BAL bal;
  if (!io.getBAL(bal)) {
    std::fprintf(stderr, "BAL arrays: %s\n", io.lastError().c_str());
    return 1;
  }
  const int C = static_cast<int>(bal.cams.size() / 9), P = static_cast<int>(bal.points.size() / 3);
  const int E = static_cast<int>(bal.edgeCam.size());
  g2o_hip::BALSolverHIP s(std::atoi(argv[4]));
  if (!s.setGraph(C, P, E, bal.edgeCam.data(), bal.edgePoint.data()) || !s.setMeasurements(bal.meas.data()) ||
      !s.setEstimates(bal.cams.data(), bal.points.data())) {
    std::fprintf(stderr, "setup failed: %s\n", s.lastError().c_str());
    return 1;
  }
  return run(s, argv[2], std::atoi(argv[3]));
}

int run_g2o_file(int argc, char** argv) {
  g2o_hip::GraphIO io;
  if (!io.load(argv[1])) {
    std::fprintf(stderr, "load failed: %s\n", io.lastError().c_str());
    return 1;
  }
  const std::string log = io.log();
  if (!log.empty()) std::fprintf(stderr, "%s", log.c_str());
  const std::vector<long long> info = io.info();
  if (info[2] == 0 && info[4] > 0) return run_pose_graph_file(io, argc, argv);
  // a file whose projection edges all name camera parameters carries its intrinsics
  const bool from_params = info[2] > 0 && info[10] == info[2];
  if (!from_params && argc < 9) {
    std::fprintf(stderr, "usage: %s graph.g2o solver|optimize|gn|dogleg iterations device fx fy cx cy [out.g2o]\n",
                 argv[0]);
    return 2;
  }
  const int out_arg = from_params ? 5 : 9;
  g2o_hip::GraphIO::BA ba;
  if (!io.getBA(ba)) {
    std::fprintf(stderr, "load failed: %s\n", io.lastError().c_str());
    return 1;
  }
  const int C = (int)ba.camId.size(), P = (int)ba.pointId.size(), E = (int)ba.edgeCam.size();
  std::vector<double> w(E), intr(4 * static_cast<size_t>(C));
  // scalar information (O = w I) as edge weights; otherwise the full 2x2
  // matrices through setInformation with unit weights
  bool scalar = true;
  for (int e = 0; e < E; ++e) {
    const double* o = &ba.omega[3 * static_cast<size_t>(e)];
    if (o[1] != 0.0 || o[0] != o[2]) scalar = false;
    w[e] = o[0];
  }
  if (!scalar) std::fill(w.begin(), w.end(), 1.0);
  if (from_params) {
    std::vector<int> ids, edgeParam;
    std::vector<double> params;
    if (!io.getCameraParameters(ids, params, edgeParam)) return 1;
    // EdgeProjectXYZ2UV: fx = fy = f, principal point (cx, cy). g2o resolves
    // the parameters per edge; the solver holds them per camera, so a camera
    // seen through two different parameter sets cannot be represented
    std::vector<int> camParam(C, -1);
    for (int e = 0; e < E; ++e) {
      int& cp = camParam[ba.edgeCam[e]];
      if (cp >= 0 && cp != edgeParam[e]) {
        const double* a = &params[4 * static_cast<size_t>(cp)];
        const double* b = &params[4 * static_cast<size_t>(edgeParam[e])];
        if (a[0] != b[0] || a[1] != b[1] || a[2] != b[2]) {
          std::fprintf(stderr, "camera %d is observed through two different camera parameter sets\n",
                       ba.camId[ba.edgeCam[e]]);
          return 1;
        }
      }
      cp = edgeParam[e];
      const double* p = &params[4 * static_cast<size_t>(edgeParam[e])];
      double* in = &intr[4 * static_cast<size_t>(ba.edgeCam[e])];
      in[0] = in[1] = p[0];
      in[2] = p[1];
      in[3] = p[2];
    }
  } else {
    for (int c = 0; c < C; ++c)
      for (int k = 0; k < 4; ++k) intr[4 * static_cast<size_t>(c) + k] = std::atof(argv[5 + k]);
  }
  g2o_hip::BlockSolverHIP s(std::atoi(argv[4]));
  if (!s.setGraph(C, P, E, ba.edgeCam.data(), ba.edgePoint.data(), ba.camFixed.data(), ba.pointFixed.data()) ||
      !s.setMeasurements(ba.meas.data(), w.data(), intr.data(), 0.0) ||
      (!scalar && !s.setInformation(ba.omega.data())) ||
      !s.setEstimates(ba.cams.data(), ba.points.data())) {
    std::fprintf(stderr, "setup failed: %s\n", s.lastError().c_str());
    return 1;
  }
  if (run(s, argv[2], std::atoi(argv[3]))) return 1;
  if (argc > out_arg) {
    if (!s.getEstimates(ba.cams.data(), ba.points.data()) || !io.setBAEstimates(ba.cams.data(), ba.points.data()) ||
        !io.save(argv[out_arg], 17)) {
      std::fprintf(stderr, "save failed: %s %s\n", s.lastError().c_str(), io.lastError().c_str());
      return 1;
    }
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s problem.bin solver|optimize iterations device\n", argv[0]);
    return 2;
  }
  const std::string in(argv[1]);
  if (in.size() > 4 && in.compare(in.size() - 4, 4, ".g2o") == 0) return run_g2o_file(argc, argv);
  if (in.size() > 4 && in.compare(in.size() - 4, 4, ".bal") == 0) return run_bal_file(argc, argv);
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int hdr[5];
  if (std::fread(hdr, sizeof(int), 5, f) != 5) return 2;
  const int C = hdr[0], P = hdr[1], E = hdr[2];
  std::vector<double> cams, intr, pts, meas, w;
  std::vector<int> ec, ep;
  std::vector<uint8_t> cf, pf;
  double huber = 0;
  bool ok = read_vec(f, cams, 7 * static_cast<size_t>(C)) && read_vec(f, intr, 4 * static_cast<size_t>(C)) &&
            read_vec(f, pts, 3 * static_cast<size_t>(P)) && read_vec(f, ec, E) && read_vec(f, ep, E) &&
            read_vec(f, meas, 2 * static_cast<size_t>(E)) && read_vec(f, w, E) && read_vec(f, cf, C) &&
            read_vec(f, pf, P) && std::fread(&huber, sizeof(double), 1, f) == 1;
  std::fclose(f);
  if (!ok) return 2;
  g2o_hip::BlockSolverHIP s(std::atoi(argv[4]));
  g2ohip_options opt = s.options();
  opt.precision = hdr[3];
  opt.refine_steps = hdr[4];
  if (!s.setOptions(opt) || !s.setGraph(C, P, E, ec.data(), ep.data(), cf.data(), pf.data()) ||
      !s.setMeasurements(meas.data(), w.data(), intr.data(), huber) || !s.setEstimates(cams.data(), pts.data())) {
    std::fprintf(stderr, "setup failed: %s\n", s.lastError().c_str());
    return 1;
  }
  return run(s, argv[2], std::atoi(argv[3]));
}
