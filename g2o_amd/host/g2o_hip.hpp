// C++ host classes over the C ABI of include/g2o_hip.h, with the method names
// and call contracts of the g2o interfaces they stand in for:
//   BlockSolverHIP   - g2o::Solver (g2o/core/solver.h:44-153) as implemented
//                      by BlockSolver<BlockSolverTraits<6,3>>
//                      (g2o/core/block_solver.h:124, .hpp:110-606), plus the
//                      SparseOptimizer / OptimizationAlgorithmLevenberg calls
//                      that read or write vertex state
//                      (sparse_optimizer.cpp:78-130, 457-468, 547-567;
//                      optimization_algorithm_levenberg.cpp:155-176);
//   LinearSolverHIP  - g2o::LinearSolver<MatrixType>::solve on a host
//                      symmetric block matrix (linear_solver.h:47-145) with
//                      LinearSolverPCG semantics (linear_solver_pcg.h:45-50);
//   BALSolverHIP     - the same Solver contract for BlockSolver<
//                      BlockSolverTraits<9,3>> over the BAL example's types;
//   PoseGraphSolverHIP - the same Solver contract for VertexSE3 / EdgeSE3
//                      graphs without Schur;
//   GraphIO          - OptimizableGraph::load / save (optimizable_graph.cpp:
//                      344-567) for the path's vertex and edge types, and the
//                      BAL parser of bal_example.cpp:265-335.
// As in g2o, calls return bool and never throw; lastError() gives the reason.
// The classes hold no Eigen or g2o types, so they build without either; the
// adapter that derives from g2o::Solver and forwards to them is shown in the
// integration notes.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/g2o_hip.h"

namespace g2o_hip {

class DeviceContext {
 public:
  explicit DeviceContext(int device) { _status = g2ohip_create(device, &_ctx); }
  ~DeviceContext() { if (_ctx) g2ohip_destroy(_ctx); }
  DeviceContext(const DeviceContext&) = delete;
  DeviceContext& operator=(const DeviceContext&) = delete;
  g2ohip_ctx* get() const { return _ctx; }
  bool ok(int status) {
    _status = status;
    return status == G2OHIP_OK;
  }
  int status() const { return _status; }
  std::string lastError() const { return _ctx ? g2ohip_last_error(_ctx) : "no context"; }

 private:
  g2ohip_ctx* _ctx = nullptr;
  int _status = G2OHIP_OK;
};

class BlockSolverHIP {
 public:
  static constexpr int kCamDim = 6;
  explicit BlockSolverHIP(int device = 0) : _dev(device) {
    g2ohip_default_options(&_opt);
    if (_dev.get()) _dev.ok(g2ohip_set_options(_dev.get(), &_opt));
  }

  // ---- graph data (SparseOptimizer::initializeOptimization + the vertices' and edges' data)
  bool setGraph(int numCams, int numPoints, int numEdges, const int* edgeCam, const int* edgePoint,
                const uint8_t* camFixed, const uint8_t* pointFixed) {
    if (!_dev.ok(g2ohip_ba_set_structure(_dev.get(), numCams, numPoints, numEdges, edgeCam, edgePoint, camFixed,
                                         pointFixed)))
      return false;
    long long info[13];
    if (!_dev.ok(g2ohip_ba_structure_info(_dev.get(), info))) return false;
    _size = static_cast<size_t>(6 * info[0] + 3 * info[1]);
    // hessianIndex of each vertex (sparse_optimizer.cpp:175-200): the free
    // vertices that have an edge, in input order
    std::vector<uint8_t> camUsed(numCams, 0), pointUsed(numPoints, 0);
    for (int k = 0; k < numEdges; ++k) camUsed[edgeCam[k]] = pointUsed[edgePoint[k]] = 1;
    _camIndex.assign(numCams, -1);
    _pointIndex.assign(numPoints, -1);
    int cf = 0, pf = 0;
    for (int c = 0; c < numCams; ++c)
      if (camUsed[c] && !(camFixed && camFixed[c])) _camIndex[c] = cf++;
    for (int p = 0; p < numPoints; ++p)
      if (pointUsed[p] && !(pointFixed && pointFixed[p])) _pointIndex[p] = pf++;
    _hpp.assign(36 * static_cast<size_t>(cf), 0.0);
    _hll.assign(9 * static_cast<size_t>(pf), 0.0);
    _hessianStale = true;
    return true;
  }
  // BaseEdge::setInformation for every edge: omega[E*3] = O00 O01 O11
  bool setInformation(const double* omega) { return _dev.ok(g2ohip_ba_set_information(_dev.get(), omega)); }
  bool setMeasurements(const double* meas, const double* weight, const double* intrinsics, double huberDelta) {
    return _dev.ok(g2ohip_ba_set_measurements(_dev.get(), meas, weight, intrinsics, huberDelta));
  }
  bool setEstimates(const double* cams, const double* points) {
    return _dev.ok(g2ohip_ba_set_state(_dev.get(), cams, points));
  }
  bool getEstimates(double* cams, double* points) { return _dev.ok(g2ohip_ba_get_state(_dev.get(), cams, points)); }
  bool setOptions(const g2ohip_options& opt) {
    _opt = opt;
    return _dev.ok(g2ohip_set_options(_dev.get(), &_opt));
  }
  const g2ohip_options& options() const { return _opt; }
  // SparseOptimizer::setAlgorithm: G2OHIP_LEVENBERG or G2OHIP_GAUSS_NEWTON, used by optimize()
  bool setAlgorithm(int algorithm) { return _dev.ok(g2ohip_set_algorithm(_dev.get(), algorithm)); }
  // Solver::computeMarginals: blocks (rows[k], cols[k]) of Hpp^-1 into blocks[36 * n]
  bool computeMarginals(int n, const int* rows, const int* cols, double* blocks, double tol = 0.0) {
    return _dev.ok(g2ohip_compute_marginals(_dev.get(), n, rows, cols, blocks, tol));
  }

  // ---- g2o::Solver
  // init(): the linear solver forgets its PCG residual memory (LinearSolverPCG::init)
  bool init(bool /*online*/ = false) { return _dev.ok(g2ohip_set_options(_dev.get(), &_opt)); }
  // the block structure is built once by setGraph
  bool buildStructure(bool /*zeroBlocks*/ = false) { return _size > 0; }
  bool buildSystem() {
    _hessianStale = true;
    return _dev.ok(g2ohip_build_system(_dev.get()));
  }
  bool solve() {
    int ok = 0;
    return _dev.ok(g2ohip_solve(_dev.get(), &ok)) && ok != 0;
  }
  // BlockSolverBase::multiplyHessian: the pose rows of dest = Hpp src (both
  // of vectorSize() doubles; the landmark rows of dest are left as they are)
  bool multiplyHessian(double* dest, const double* src) {
    return _dev.ok(g2ohip_multiply_hessian(_dev.get(), src, dest));
  }
  // Vertex Hessians. g2o's BlockSolver::buildStructure maps each vertex's
  // _hessian onto its diagonal block (block_solver.hpp:144-161), which
  // OptimizationAlgorithmLevenberg::computeLambdaInit reads
  // (optimization_algorithm_levenberg.cpp:155-167). Here the blocks live in a
  // host mirror with a stable address per vertex from setGraph on
  // (vertexHessianMemory: row-major 6x6 for a camera, 3x3 for a point, nullptr
  // for a vertex without a block - fixed, or without edges), filled from the
  // last buildSystem by syncVertexHessians; vertexHessian syncs on demand.
  double* vertexHessianMemory(bool camera, int index) {
    const std::vector<int>& idx = camera ? _camIndex : _pointIndex;
    if (index < 0 || index >= static_cast<int>(idx.size()) || idx[index] < 0) return nullptr;
    return camera ? &_hpp[36 * static_cast<size_t>(idx[index])] : &_hll[9 * static_cast<size_t>(idx[index])];
  }
  bool syncVertexHessians() {
    if (!_hessianStale) return true;
    if (!_dev.ok(g2ohip_ba_get_hessian(_dev.get(), _hpp.data(), _hll.data(), nullptr))) return false;
    _hessianStale = false;
    return true;
  }
  const double* vertexHessian(bool camera, int index) {
    double* m = vertexHessianMemory(camera, index);
    return m && syncVertexHessians() ? m : nullptr;
  }
  int numCameras() const { return static_cast<int>(_camIndex.size()); }
  int numPoints() const { return static_cast<int>(_pointIndex.size()); }
  bool setLambda(double lambda, bool /*backup*/ = false) { return _dev.ok(g2ohip_set_lambda(_dev.get(), lambda)); }
  void restoreDiagonal() { _dev.ok(g2ohip_restore_diagonal(_dev.get())); }
  bool schur() const { return true; }
  size_t vectorSize() const { return _size; }
  // host copies of the update and right-hand side (Solver::x(), Solver::b())
  const double* x() {
    _x.resize(_size);
    return _dev.ok(g2ohip_get_x(_dev.get(), _x.data())) ? _x.data() : nullptr;
  }
  const double* b() {
    _b.resize(_size);
    return _dev.ok(g2ohip_get_b(_dev.get(), _b.data())) ? _b.data() : nullptr;
  }

  // ---- state-side calls the LM algorithm makes, executed on the device
  double computeActiveErrors() {  // computeActiveErrors + activeRobustChi2
    double chi = -1;
    return _dev.ok(g2ohip_compute_active_errors(_dev.get(), &chi)) ? chi : -1;
  }
  bool update() { return _dev.ok(g2ohip_update(_dev.get())); }  // SparseOptimizer::update(solver->x())
  bool push() { return _dev.ok(g2ohip_push(_dev.get())); }
  bool pop() { return _dev.ok(g2ohip_pop(_dev.get())); }
  bool discardTop() { return _dev.ok(g2ohip_discard_top(_dev.get())); }
  double computeLambdaInit() {
    double l = -1;
    return _dev.ok(g2ohip_lambda_init(_dev.get(), &l)) ? l : -1;
  }
  double computeScale() {
    double s = 0;
    return _dev.ok(g2ohip_compute_scale(_dev.get(), &s)) ? s : 0;
  }
  int iterationsLinearSolver() {
    double info[4];
    return _dev.ok(g2ohip_last_solve_info(_dev.get(), info)) ? static_cast<int>(info[0]) : -1;
  }
  // the whole SparseOptimizer::optimize loop on the device
  int optimize(int iterations, std::vector<g2ohip_iter_stats>* stats = nullptr) {
    std::vector<g2ohip_iter_stats> s(iterations > 0 ? iterations : 1);
    int done = 0;
    if (!_dev.ok(g2ohip_optimize(_dev.get(), iterations, s.data(), &done))) return -1;
    if (stats) stats->assign(s.begin(), s.begin() + done);
    return done;
  }
  std::string lastError() const { return _dev.lastError(); }
  g2ohip_ctx* context() const { return _dev.get(); }

 private:
  DeviceContext _dev;
  g2ohip_options _opt{};
  size_t _size = 0;
  std::vector<double> _x, _b;
  std::vector<int> _camIndex, _pointIndex;  // hessianIndex per vertex, -1 = none
  std::vector<double> _hpp, _hll;           // vertex Hessian mirror
  bool _hessianStale = true;
};

// The Solver contract of BlockSolver<BlockSolverTraits<9, 3>> over g2o's BAL
// types (g2o/examples/bal/bal_example.cpp:70-207, 227-341): VertexCameraBAL
// (9 values, additive update), VertexPointBAL, EdgeObservationBAL; FP64,
// explicit Schur, LinearSolverPCG. Same calls as BlockSolverHIP; the vertex
// Hessian mirror holds 9x9 camera blocks.
class BALSolverHIP {
 public:
  static constexpr int kCamDim = 9;
  explicit BALSolverHIP(int device = 0) : _dev(device) {
    g2ohip_default_options(&_opt);
    if (_dev.get()) _dev.ok(g2ohip_set_options(_dev.get(), &_opt));
  }
  bool setGraph(int numCams, int numPoints, int numEdges, const int* edgeCam, const int* edgePoint,
                const uint8_t* camFixed = nullptr, const uint8_t* pointFixed = nullptr) {
    if (!_dev.ok(g2ohip_bal_set_structure(_dev.get(), numCams, numPoints, numEdges, edgeCam, edgePoint, camFixed,
                                          pointFixed)))
      return false;
    long long info[7];
    if (!_dev.ok(g2ohip_bal_structure_info(_dev.get(), info))) return false;
    _size = static_cast<size_t>(9 * info[0] + 3 * info[1]);
    std::vector<uint8_t> camUsed(numCams, 0), pointUsed(numPoints, 0);
    for (int k = 0; k < numEdges; ++k) camUsed[edgeCam[k]] = pointUsed[edgePoint[k]] = 1;
    _camIndex.assign(numCams, -1);
    _pointIndex.assign(numPoints, -1);
    int cf = 0, pf = 0;
    for (int c = 0; c < numCams; ++c)
      if (camUsed[c] && !(camFixed && camFixed[c])) _camIndex[c] = cf++;
    for (int q = 0; q < numPoints; ++q)
      if (pointUsed[q] && !(pointFixed && pointFixed[q])) _pointIndex[q] = pf++;
    _hpp.assign(81 * static_cast<size_t>(cf), 0.0);
    _hll.assign(9 * static_cast<size_t>(pf), 0.0);
    _hessianStale = true;
    return true;
  }
  // meas[E*2]; weight[E] (NULL = identity information); Huber delta (0 = none)
  bool setMeasurements(const double* meas, const double* weight = nullptr, double huberDelta = 0.0) {
    return _dev.ok(g2ohip_bal_set_measurements(_dev.get(), meas, weight, huberDelta));
  }
  bool setEstimates(const double* cams9, const double* points) {
    return _dev.ok(g2ohip_bal_set_state(_dev.get(), cams9, points));
  }
  bool getEstimates(double* cams9, double* points) {
    return _dev.ok(g2ohip_bal_get_state(_dev.get(), cams9, points));
  }
  bool setOptions(const g2ohip_options& opt) {
    _opt = opt;
    return _dev.ok(g2ohip_set_options(_dev.get(), &_opt));
  }
  const g2ohip_options& options() const { return _opt; }
  bool setAlgorithm(int algorithm) { return _dev.ok(g2ohip_set_algorithm(_dev.get(), algorithm)); }

  bool init(bool /*online*/ = false) { return _dev.ok(g2ohip_set_options(_dev.get(), &_opt)); }
  bool buildStructure(bool /*zeroBlocks*/ = false) { return _size > 0; }
  bool buildSystem() {
    _hessianStale = true;
    return _dev.ok(g2ohip_build_system(_dev.get()));
  }
  bool solve() {
    int ok = 0;
    return _dev.ok(g2ohip_solve(_dev.get(), &ok)) && ok != 0;
  }
  // vertex Hessians as BlockSolverHIP: row-major 9x9 for a camera, 3x3 for a point
  double* vertexHessianMemory(bool camera, int index) {
    const std::vector<int>& idx = camera ? _camIndex : _pointIndex;
    if (index < 0 || index >= static_cast<int>(idx.size()) || idx[index] < 0) return nullptr;
    return camera ? &_hpp[81 * static_cast<size_t>(idx[index])] : &_hll[9 * static_cast<size_t>(idx[index])];
  }
  bool syncVertexHessians() {
    if (!_hessianStale) return true;
    if (!_dev.ok(g2ohip_bal_get_hessian(_dev.get(), _hpp.data(), _hll.data(), nullptr))) return false;
    _hessianStale = false;
    return true;
  }
  const double* vertexHessian(bool camera, int index) {
    double* m = vertexHessianMemory(camera, index);
    return m && syncVertexHessians() ? m : nullptr;
  }
  int numCameras() const { return static_cast<int>(_camIndex.size()); }
  int numPoints() const { return static_cast<int>(_pointIndex.size()); }
  bool setLambda(double lambda, bool /*backup*/ = false) { return _dev.ok(g2ohip_set_lambda(_dev.get(), lambda)); }
  void restoreDiagonal() { _dev.ok(g2ohip_restore_diagonal(_dev.get())); }
  bool schur() const { return true; }
  size_t vectorSize() const { return _size; }
  const double* x() {
    _x.resize(_size);
    return _dev.ok(g2ohip_get_x(_dev.get(), _x.data())) ? _x.data() : nullptr;
  }
  const double* b() {
    _b.resize(_size);
    return _dev.ok(g2ohip_get_b(_dev.get(), _b.data())) ? _b.data() : nullptr;
  }
  double computeActiveErrors() {
    double chi = -1;
    return _dev.ok(g2ohip_compute_active_errors(_dev.get(), &chi)) ? chi : -1;
  }
  bool update() { return _dev.ok(g2ohip_update(_dev.get())); }
  bool push() { return _dev.ok(g2ohip_push(_dev.get())); }
  bool pop() { return _dev.ok(g2ohip_pop(_dev.get())); }
  bool discardTop() { return _dev.ok(g2ohip_discard_top(_dev.get())); }
  double computeLambdaInit() {
    double l = -1;
    return _dev.ok(g2ohip_lambda_init(_dev.get(), &l)) ? l : -1;
  }
  double computeScale() {
    double s = 0;
    return _dev.ok(g2ohip_compute_scale(_dev.get(), &s)) ? s : 0;
  }
  int iterationsLinearSolver() {
    double info[4];
    return _dev.ok(g2ohip_last_solve_info(_dev.get(), info)) ? static_cast<int>(info[0]) : -1;
  }
  int optimize(int iterations, std::vector<g2ohip_iter_stats>* stats = nullptr) {
    std::vector<g2ohip_iter_stats> s(iterations > 0 ? iterations : 1);
    int done = 0;
    if (!_dev.ok(g2ohip_optimize(_dev.get(), iterations, s.data(), &done))) return -1;
    if (stats) stats->assign(s.begin(), s.begin() + done);
    return done;
  }
  std::string lastError() const { return _dev.lastError(); }
  g2ohip_ctx* context() const { return _dev.get(); }

 private:
  DeviceContext _dev;
  g2ohip_options _opt{};
  size_t _size = 0;
  std::vector<double> _x, _b;
  std::vector<int> _camIndex, _pointIndex;
  std::vector<double> _hpp, _hll;
  bool _hessianStale = true;
};

// LinearSolver<Matrix{3,6}d>::solve for a host matrix given by its upper
// blocks (the unit-test path: upload, then the device PCG).
class LinearSolverHIP {
 public:
  LinearSolverHIP(int device, int blockDim) : _dev(device), _bd(blockDim) {}
  bool init() {
    _residual = -1.0;  // LinearSolverPCG::init
    return true;
  }
  void setTolerance(double t) { _tol = t; }
  void setMaxIterations(int m) { _maxIter = m; }
  bool solve(int numBlocks, const std::vector<int>& rows, const std::vector<int>& cols,
             const std::vector<double>& blocks, double* x, const double* b) {
    return _dev.ok(g2ohip_linear_solve_pcg(_dev.get(), numBlocks, _bd, static_cast<int>(rows.size()), rows.data(),
                                           cols.data(), blocks.data(), b, x, _tol, _maxIter, &_residual, &_iters));
  }
  int iterations() const { return _iters; }
  std::string lastError() const { return _dev.lastError(); }

 private:
  DeviceContext _dev;
  int _bd;
  double _tol = 1e-6;
  int _maxIter = -1;
  double _residual = -1.0;
  int _iters = 0;
};

// Solver contract for a VertexSE3 / EdgeSE3 pose graph (6x6 blocks, no Schur).
class PoseGraphSolverHIP {
 public:
  explicit PoseGraphSolverHIP(int device = 0) : _dev(device) {
    g2ohip_default_options(&_opt);
    _opt.pcg_tolerance = 1e-8;
    if (_dev.get()) _dev.ok(g2ohip_set_options(_dev.get(), &_opt));
  }
  bool setGraph(int numPoses, int numEdges, const int* from, const int* to, const uint8_t* fixed) {
    if (!_dev.ok(g2ohip_pg_set_structure(_dev.get(), numPoses, numEdges, from, to, fixed))) return false;
    long long info[4];
    if (!_dev.ok(g2ohip_pg_structure_info(_dev.get(), info))) return false;
    _size = static_cast<size_t>(6 * info[0]);
    return true;
  }
  bool setMeasurements(const double* meas7, const double* info36) {
    return _dev.ok(g2ohip_pg_set_measurements(_dev.get(), meas7, info36));
  }
  bool setEstimates(const double* poses7) { return _dev.ok(g2ohip_pg_set_state(_dev.get(), poses7)); }
  bool getEstimates(double* poses7) { return _dev.ok(g2ohip_pg_get_state(_dev.get(), poses7)); }
  // SparseOptimizer::setAlgorithm: G2OHIP_LEVENBERG or G2OHIP_GAUSS_NEWTON, used by optimize()
  bool setAlgorithm(int algorithm) { return _dev.ok(g2ohip_set_algorithm(_dev.get(), algorithm)); }
  // Solver::computeMarginals: blocks (rows[k], cols[k]) of Hpp^-1 into blocks[36 * n]
  bool computeMarginals(int n, const int* rows, const int* cols, double* blocks, double tol = 0.0) {
    return _dev.ok(g2ohip_compute_marginals(_dev.get(), n, rows, cols, blocks, tol));
  }
  bool setOptions(const g2ohip_options& opt) {
    _opt = opt;
    return _dev.ok(g2ohip_set_options(_dev.get(), &_opt));
  }
  bool buildSystem() { return _dev.ok(g2ohip_build_system(_dev.get())); }
  bool solve() {
    int ok = 0;
    return _dev.ok(g2ohip_solve(_dev.get(), &ok)) && ok != 0;
  }
  bool setLambda(double lambda, bool = false) { return _dev.ok(g2ohip_set_lambda(_dev.get(), lambda)); }
  void restoreDiagonal() { _dev.ok(g2ohip_restore_diagonal(_dev.get())); }
  // BlockSolverBase::multiplyHessian: dest = H src (the whole Hessian of the pose graph)
  bool multiplyHessian(double* dest, const double* src) {
    return _dev.ok(g2ohip_multiply_hessian(_dev.get(), src, dest));
  }
  size_t vectorSize() const { return _size; }
  int optimize(int iterations, std::vector<g2ohip_iter_stats>* stats = nullptr) {
    std::vector<g2ohip_iter_stats> s(iterations > 0 ? iterations : 1);
    int done = 0;
    if (!_dev.ok(g2ohip_optimize(_dev.get(), iterations, s.data(), &done))) return -1;
    if (stats) stats->assign(s.begin(), s.begin() + done);
    return done;
  }
  std::string lastError() const { return _dev.lastError(); }

 private:
  DeviceContext _dev;
  g2ohip_options _opt{};
  size_t _size = 0;
};

// g2o text files and BAL files through the library's native graph container.
// Arrays come out in the layout BlockSolverHIP / PoseGraphSolverHIP take:
// vertices of one type sorted by id, edges in file order.
class GraphIO {
 public:
  struct BA {
    std::vector<int> camId, pointId, edgeCam, edgePoint;
    std::vector<double> cams, points, meas, omega;  // omega: O00 O01 O11 per edge
    std::vector<uint8_t> camFixed, pointFixed;
  };
  struct BAL {  // bal_example.cpp's arrays: cameras [C*9], points [P*3], observations in file order
    std::vector<double> cams, points, meas;
    std::vector<int> edgeCam, edgePoint;
  };
  struct PoseGraph {
    std::vector<int> poseId, edgeFrom, edgeTo;
    std::vector<double> poses, meas, information;
    std::vector<uint8_t> fixed;
  };

  GraphIO() { _status = g2ohip_graph_create(&_g); }
  ~GraphIO() { g2ohip_graph_destroy(_g); }
  GraphIO(const GraphIO&) = delete;
  GraphIO& operator=(const GraphIO&) = delete;

  bool load(const char* filename, int nthreads = 0) { return ok(g2ohip_graph_load(_g, filename, nthreads)); }
  bool loadBAL(const char* filename, int nthreads = 0) { return ok(g2ohip_graph_load_bal(_g, filename, nthreads)); }
  // precision 6 prints as g2o's default ostream; 17 round-trips exactly
  bool save(const char* filename, int precision = 6) { return ok(g2ohip_graph_save(_g, filename, precision)); }
  std::string log() const { return _g ? g2ohip_graph_log(_g) : ""; }
  std::string lastError() const { return _g ? g2ohip_graph_last_error(_g) : "no graph"; }
  int status() const { return _status; }

  // [cameras, points, projection edges, poses, pose edges, fixed, lines,
  //  unknown-tag lines, skipped elements, BAL observations, XYZ2UV edges,
  //  camera parameters]
  std::vector<long long> info() const {
    std::vector<long long> v(12, 0);
    if (_g) g2ohip_graph_info(_g, v.data());
    return v;
  }

  bool getBA(BA& o) {
    const std::vector<long long> n = info();
    const size_t C = n[0], P = n[1], E = n[2];
    o.camId.resize(C); o.cams.resize(7 * C); o.camFixed.resize(C);
    o.pointId.resize(P); o.points.resize(3 * P); o.pointFixed.resize(P);
    o.edgeCam.resize(E); o.edgePoint.resize(E); o.meas.resize(2 * E); o.omega.resize(3 * E);
    return ok(g2ohip_graph_get_ba(_g, o.camId.data(), o.cams.data(), o.camFixed.data(), o.pointId.data(),
                                  o.points.data(), o.pointFixed.data(), o.edgeCam.data(), o.edgePoint.data(),
                                  o.meas.data(), o.omega.data()));
  }
  bool getBAL(BAL& o) {
    long long n[3] = {0, 0, 0};
    if (!ok(g2ohip_graph_get_bal(_g, n, nullptr, nullptr, nullptr, nullptr, nullptr))) return false;
    o.cams.resize(9 * n[0]); o.points.resize(3 * n[1]);
    o.edgeCam.resize(n[2]); o.edgePoint.resize(n[2]); o.meas.resize(2 * n[2]);
    return ok(g2ohip_graph_get_bal(_g, nullptr, o.cams.data(), o.points.data(), o.edgeCam.data(),
                                   o.edgePoint.data(), o.meas.data()));
  }
  bool getPoseGraph(PoseGraph& o) {
    const std::vector<long long> n = info();
    const size_t N = n[3], E = n[4];
    o.poseId.resize(N); o.poses.resize(7 * N); o.fixed.resize(N);
    o.edgeFrom.resize(E); o.edgeTo.resize(E); o.meas.resize(7 * E); o.information.resize(36 * E);
    return ok(g2ohip_graph_get_pg(_g, o.poseId.data(), o.poses.data(), o.fixed.data(), o.edgeFrom.data(),
                                  o.edgeTo.data(), o.meas.data(), o.information.data()));
  }
  // camera parameters sorted by id (f cx cy baseline) and each projection
  // edge's parameter index, -1 for EDGE_SE3_PROJECT_XYZ
  bool getCameraParameters(std::vector<int>& ids, std::vector<double>& params, std::vector<int>& edgeParam) {
    const std::vector<long long> n = info();
    ids.resize(n[11]);
    params.resize(4 * n[11]);
    edgeParam.resize(n[2]);
    return ok(g2ohip_graph_get_ba_params(_g, ids.data(), params.data(), edgeParam.data()));
  }
  bool setBAEstimates(const double* cams, const double* points) {
    return ok(g2ohip_graph_set_ba_state(_g, cams, points));
  }
  bool setPoseEstimates(const double* poses) { return ok(g2ohip_graph_set_pg_state(_g, poses)); }

 private:
  bool ok(int status) {
    _status = status;
    return status == G2OHIP_OK;
  }
  g2ohip_graph* _g = nullptr;
  int _status = G2OHIP_OK;
};

}  // namespace g2o_hip
