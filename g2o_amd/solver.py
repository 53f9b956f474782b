"""Python host mirror of g2o's SparseOptimizer + OptimizationAlgorithmLevenberg
+ BlockSolver_6_3 API for bundle adjustment, backed by libg2o_hip.so.

Method names follow the reference (g2o/core/sparse_optimizer.h,
g2o/core/solver.h, g2o/core/optimization_algorithm_levenberg.h):
``initialize_optimization`` / ``optimize`` / ``compute_active_errors`` /
``active_robust_chi2`` on the optimizer, ``build_system`` / ``set_lambda`` /
``solve`` / ``restore_diagonal`` / ``x`` / ``b`` on the solver. Errors raise
:class:`G2oHipError` carrying the C-ABI status.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib as L


class G2oHipError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"g2o_hip status {status}: {msg}")
        self.status = status


def _dp(a):
    return a.ctypes.data_as(L._dp)


def _ip(a):
    return a.ctypes.data_as(L._ip)


def _up(a):
    return a.ctypes.data_as(L._up)


def default_options() -> L.Options:
    o = L.Options()
    L.lib().g2ohip_default_options(ct.byref(o))
    return o


class Context:
    """One device context (``device=-1``: host-only, structure checks only)."""

    def __init__(self, device: int = 0):
        self._lib = L.lib()
        h = ct.c_void_p()
        st = self._lib.g2ohip_create(device, ct.byref(h))
        if st != L.OK:
            raise G2oHipError(st, f"cannot create a context on device {device}")
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            self._lib.g2ohip_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def check(self, st: int):
        if st != L.OK:
            raise G2oHipError(st, self._lib.g2ohip_last_error(self.h).decode())

    def call(self, name: str, *args):
        self.check(getattr(self._lib, name)(self.h, *args))

    def set_options(self, **kw):
        o = L.Options()
        self.check(self._lib.g2ohip_get_options(self.h, ct.byref(o)))
        for k, v in kw.items():
            setattr(o, k, v)
        self.check(self._lib.g2ohip_set_options(self.h, ct.byref(o)))
        return o


TRIAL_FIELDS = ["lambda", "chi2_current", "chi2_trial", "scale", "rho", "accepted", "pcg_iterations",
                "refine_steps", "fp64_fallback", "pcg_residual_mem"]


def _lm_trials(ctx):
    """The LM trials of the last optimize(): float64[n, 10] with the columns
    TRIAL_FIELDS."""
    n = ct.c_int()
    ctx.call("g2ohip_lm_trials", None, 0, ct.byref(n))
    out = np.zeros((max(n.value, 1), len(TRIAL_FIELDS)))
    ctx.call("g2ohip_lm_trials", _dp(out), n.value, ct.byref(n))
    return out[:n.value]


def _marginals(ctx, blocks, tol):
    """SparseOptimizer::computeMarginals for the (row, col) pairs in blocks
    (free-vertex indices); returns float64[n, 6, 6]."""
    bl = np.asarray(blocks, dtype=np.int32).reshape(-1, 2)
    r = np.ascontiguousarray(bl[:, 0])
    c = np.ascontiguousarray(bl[:, 1])
    out = np.zeros((bl.shape[0], 6, 6))
    ctx.call("g2ohip_compute_marginals", int(bl.shape[0]), _ip(r), _ip(c), _dp(out), float(tol))
    return out


class BundleAdjuster:
    """SparseOptimizer with OptimizationAlgorithmLevenberg and a BlockSolverHIP
    for one bundle-adjustment problem (``g2o_amd.datasets.BAProblem``)."""

    def __init__(self, prob, device: int = 0, precision: int = L.FP64, schur_mode: int = L.SCHUR_EXPLICIT,
                 pcg_tolerance: float = 1e-6, pcg_max_iter: int = -1, refine_steps: int = 4, max_trials: int = 10):
        self.ctx = Context(device)
        self.prob = prob
        self.set_options(precision=precision, schur_mode=schur_mode, pcg_tolerance=pcg_tolerance,
                         pcg_max_iter=pcg_max_iter, refine_steps=refine_steps, max_trials=max_trials)
        self.C, self.P, self.E = prob.num_cams, prob.num_pts, prob.num_edges
        self._ec = np.ascontiguousarray(prob.edge_cam, dtype=np.int32)
        self._ep = np.ascontiguousarray(prob.edge_pt, dtype=np.int32)
        self._cf = np.ascontiguousarray(prob.cam_fixed, dtype=np.uint8)
        self._pf = np.ascontiguousarray(prob.pt_fixed, dtype=np.uint8)
        self.ctx.call("g2ohip_ba_set_structure", self.C, self.P, self.E, _ip(self._ec), _ip(self._ep), _up(self._cf),
                      _up(self._pf))
        info = self.structure_info()
        self.Cf, self.Pf = int(info[0]), int(info[1])
        if device >= 0:
            self.set_measurements(prob.meas, prob.weight, prob.intr, prob.huber)
            if getattr(prob, "omega", None) is not None:
                self.set_information(prob.omega)
            self.set_state(prob.cams, prob.pts)

    # ---- setup
    def set_options(self, **kw):
        return self.ctx.set_options(**kw)

    def set_algorithm(self, algorithm: int):
        """SparseOptimizer::setAlgorithm: L.LEVENBERG or L.GAUSS_NEWTON."""
        self.ctx.call("g2ohip_set_algorithm", int(algorithm))

    def compute_marginals(self, blocks, tol: float = 0.0) -> np.ndarray:
        """Blocks of Hpp^-1 (see g2ohip_compute_marginals)."""
        return _marginals(self.ctx, blocks, tol)

    def structure_info(self) -> np.ndarray:
        info = (ct.c_longlong * 13)()
        self.ctx.check(self.ctx._lib.g2ohip_ba_structure_info(self.ctx.h, info))
        return np.array(info[:], dtype=np.int64)

    def set_measurements(self, meas, weight, intr, huber):
        m = np.ascontiguousarray(meas, dtype=np.float64)
        w = np.ascontiguousarray(weight, dtype=np.float64)
        k = np.ascontiguousarray(intr, dtype=np.float64)
        self.ctx.call("g2ohip_ba_set_measurements", _dp(m), _dp(w), _dp(k), float(huber))

    def set_information(self, omega):
        """BaseEdge::setInformation for every edge: omega[E, 3] = the upper
        triangle [O00, O01, O11] of each 2x2 information matrix (None: identity)."""
        if omega is None:
            self.ctx.call("g2ohip_ba_set_information", None)
            return
        o = np.ascontiguousarray(omega, dtype=np.float64).reshape(-1, 3)
        if o.shape[0] != self.E:
            raise ValueError(f"omega needs {self.E} rows of 3")
        self.ctx.call("g2ohip_ba_set_information", _dp(o))

    def set_state(self, cams, pts):
        c = np.ascontiguousarray(cams, dtype=np.float64)
        p = np.ascontiguousarray(pts, dtype=np.float64)
        self.ctx.call("g2ohip_ba_set_state", _dp(c), _dp(p))

    def state(self):
        c = np.zeros((self.C, 7))
        p = np.zeros((self.P, 3))
        self.ctx.call("g2ohip_ba_get_state", _dp(c), _dp(p))
        return c, p

    # ---- SparseOptimizer
    def compute_active_errors(self) -> float:
        """computeActiveErrors + activeRobustChi2; returns the robust chi2."""
        v = ct.c_double()
        self.ctx.call("g2ohip_compute_active_errors", ct.byref(v))
        return v.value

    def errors(self) -> np.ndarray:
        e = np.zeros((self.E, 2))
        self.ctx.call("g2ohip_get_errors", _dp(e))
        return e

    def optimize(self, iterations: int):
        stats = (L.IterStats * max(iterations, 1))()
        done = ct.c_int()
        self.ctx.call("g2ohip_optimize", iterations, stats, ct.byref(done))
        out = []
        for s in stats[:done.value]:
            out.append({f: getattr(s, f) for f, _ in L.IterStats._fields_})
        return out

    def update(self):
        self.ctx.call("g2ohip_update")

    def push(self):
        self.ctx.call("g2ohip_push")

    def pop(self):
        self.ctx.call("g2ohip_pop")

    def discard_top(self):
        self.ctx.call("g2ohip_discard_top")

    # ---- Solver (BlockSolverHIP)
    def build_system(self):
        self.ctx.call("g2ohip_build_system")

    def set_lambda(self, lam: float):
        self.ctx.call("g2ohip_set_lambda", float(lam))

    def restore_diagonal(self):
        self.ctx.call("g2ohip_restore_diagonal")

    def solve(self) -> bool:
        ok = ct.c_int()
        self.ctx.call("g2ohip_solve", ct.byref(ok))
        return bool(ok.value)

    def x(self) -> np.ndarray:
        v = np.zeros(6 * self.Cf + 3 * self.Pf)
        self.ctx.call("g2ohip_get_x", _dp(v))
        return v

    def b(self) -> np.ndarray:
        v = np.zeros(6 * self.Cf + 3 * self.Pf)
        self.ctx.call("g2ohip_get_b", _dp(v))
        return v

    def multiply_hessian(self, src) -> np.ndarray:
        """BlockSolverBase::multiplyHessian: the pose rows of Hpp src (the
        landmark rows of the result are zero)."""
        s = np.ascontiguousarray(src, dtype=np.float64)
        if s.size != 6 * self.Cf + 3 * self.Pf:
            raise ValueError("src must hold 6 Cf + 3 Pf values")
        v = np.zeros_like(s)
        self.ctx.call("g2ohip_multiply_hessian", _dp(s), _dp(v))
        return v

    def lambda_init(self) -> float:
        v = ct.c_double()
        self.ctx.call("g2ohip_lambda_init", ct.byref(v))
        return v.value

    def compute_scale(self) -> float:
        v = ct.c_double()
        self.ctx.call("g2ohip_compute_scale", ct.byref(v))
        return v.value

    def hessian(self):
        Hpp = np.zeros((self.Cf, 6, 6))
        Hll = np.zeros((self.Pf, 3, 3))
        Hpl = np.zeros((self.E, 6, 3))
        self.ctx.call("g2ohip_ba_get_hessian", _dp(Hpp), _dp(Hll), _dp(Hpl))
        return Hpp, Hll, Hpl

    def schur(self, nschur: int):
        rows = np.zeros(nschur, dtype=np.int32)
        cols = np.zeros(nschur, dtype=np.int32)
        blocks = np.zeros((nschur, 6, 6))
        bs = np.zeros(6 * self.Cf)
        self.ctx.call("g2ohip_ba_get_schur", _ip(rows), _ip(cols), _dp(blocks), _dp(bs))
        return rows, cols, blocks, bs

    def set_profiling(self, on: bool = True):
        self.ctx.call("g2ohip_set_profiling", int(on))

    def replay_pcg(self, reps: int = 200):
        """(us per SpMV launch, us per PCG iteration) on the last solve's
        operator, back-to-back launches between HIP events (g2ohip_replay_pcg)."""
        a, b = ct.c_double(), ct.c_double()
        self.ctx.call("g2ohip_replay_pcg", int(reps), ct.byref(a), ct.byref(b))
        return a.value, b.value

    def profile(self) -> dict:
        """{kernel name: (total ms, launches)} since set_profiling(True)."""
        out = {}
        for k, name in enumerate(L.KERNEL_NAMES):
            t = ct.c_double()
            n = ct.c_int()
            self.ctx.call("g2ohip_get_profile", k, ct.byref(t), ct.byref(n))
            out[name] = (t.value, n.value)
        return out

    # ---- exact Hessian: stationary point and implicit differentiation
    def newton(self, steps: int = 3, tol: float = -1.0, max_step: float = 1.0) -> bool:
        """Damped exact-Hessian Newton steps with a trust region of max_step
        per component (drive an LM solution to the stationary point). Returns
        False if a step found no acceptable damping."""
        ok = ct.c_int()
        self.ctx.call("g2ohip_ba_newton", int(steps), float(tol), float(max_step), ct.byref(ok))
        return bool(ok.value)

    def adjoint(self, dL_dcam, dL_dpt=None, tol: float = -1.0, max_iter: int = -1):
        """Implicit-function gradient of a loss of the converged state.
        dL_dcam [C,6] in the left exp-map coordinates [omega, upsilon], dL_dpt [P,3].
        Returns (dL_dmeas [E,2], dL_dweight [E], PCG iterations)."""
        dc = np.ascontiguousarray(dL_dcam, dtype=np.float64).reshape(self.C, 6)
        dp = None if dL_dpt is None else np.ascontiguousarray(dL_dpt, dtype=np.float64).reshape(self.P, 3)
        dz = np.zeros((self.E, 2))
        dw = np.zeros(self.E)
        it = ct.c_int()
        self.ctx.call("g2ohip_ba_adjoint", _dp(dc), None if dp is None else _dp(dp), _dp(dz), _dp(dw), float(tol),
                      int(max_iter), ct.byref(it))
        return dz, dw, it.value

    def last_solve_info(self):
        """[PCG iterations, refinement corrections, PCG residual memory, 1 if FP64 fallback]"""
        v = np.zeros(4)
        self.ctx.call("g2ohip_last_solve_info", _dp(v))
        return v

    def lm_trials(self) -> np.ndarray:
        """Per LM trial of the last optimize(): columns TRIAL_FIELDS."""
        return _lm_trials(self.ctx)


class PoseGraphOptimizer:
    """SparseOptimizer with OptimizationAlgorithmLevenberg over VertexSE3 /
    EdgeSE3 (``g2o_amd.datasets.PoseGraphProblem``), 6x6 blocks without Schur,
    block-Jacobi PCG in FP64."""

    def __init__(self, prob, device: int = 0, pcg_tolerance: float = 1e-8, pcg_max_iter: int = -1,
                 max_trials: int = 10):
        self.ctx = Context(device)
        self.prob = prob
        self.ctx.set_options(precision=L.FP64, pcg_tolerance=pcg_tolerance, pcg_max_iter=pcg_max_iter,
                             max_trials=max_trials)
        self.N, self.E = prob.num_poses, prob.num_edges
        self._fr = np.ascontiguousarray(prob.edge_from, dtype=np.int32)
        self._to = np.ascontiguousarray(prob.edge_to, dtype=np.int32)
        self._fx = np.ascontiguousarray(prob.fixed, dtype=np.uint8)
        self.ctx.call("g2ohip_pg_set_structure", self.N, self.E, _ip(self._fr), _ip(self._to), _up(self._fx))
        info = self.structure_info()
        self.Nf, self.nblocks = int(info[0]), int(info[2])
        if device >= 0:
            self.set_measurements(prob.meas, prob.info)
            self.set_state(prob.poses)

    def set_options(self, **kw):
        return self.ctx.set_options(**kw)

    def set_algorithm(self, algorithm: int):
        """SparseOptimizer::setAlgorithm: L.LEVENBERG or L.GAUSS_NEWTON."""
        self.ctx.call("g2ohip_set_algorithm", int(algorithm))

    def compute_marginals(self, blocks, tol: float = 0.0) -> np.ndarray:
        """Blocks of Hpp^-1 (see g2ohip_compute_marginals)."""
        return _marginals(self.ctx, blocks, tol)

    def structure_info(self) -> np.ndarray:
        info = (ct.c_longlong * 4)()
        self.ctx.check(self.ctx._lib.g2ohip_pg_structure_info(self.ctx.h, info))
        return np.array(info[:], dtype=np.int64)

    def set_measurements(self, meas, info):
        m = np.ascontiguousarray(meas, dtype=np.float64)
        i = np.ascontiguousarray(info, dtype=np.float64)
        self.ctx.call("g2ohip_pg_set_measurements", _dp(m), _dp(i))

    def set_state(self, poses):
        p = np.ascontiguousarray(poses, dtype=np.float64)
        self.ctx.call("g2ohip_pg_set_state", _dp(p))

    def state(self) -> np.ndarray:
        p = np.zeros((self.N, 7))
        self.ctx.call("g2ohip_pg_get_state", _dp(p))
        return p

    def compute_active_errors(self) -> float:
        v = ct.c_double()
        self.ctx.call("g2ohip_compute_active_errors", ct.byref(v))
        return v.value

    def errors(self) -> np.ndarray:
        e = np.zeros((self.E, 6))
        self.ctx.call("g2ohip_get_errors", _dp(e))
        return e

    def build_system(self):
        self.ctx.call("g2ohip_build_system")

    def set_lambda(self, lam: float):
        self.ctx.call("g2ohip_set_lambda", float(lam))

    def solve(self) -> bool:
        ok = ct.c_int()
        self.ctx.call("g2ohip_solve", ct.byref(ok))
        return bool(ok.value)

    def x(self) -> np.ndarray:
        v = np.zeros(6 * self.Nf)
        self.ctx.call("g2ohip_get_x", _dp(v))
        return v

    def multiply_hessian(self, src) -> np.ndarray:
        """BlockSolverBase::multiplyHessian: H src with the undamped Hessian
        of the last build_system."""
        s = np.ascontiguousarray(src, dtype=np.float64)
        if s.size != 6 * self.Nf:
            raise ValueError("src must hold 6 Nf values")
        v = np.zeros_like(s)
        self.ctx.call("g2ohip_multiply_hessian", _dp(s), _dp(v))
        return v

    def lambda_init(self) -> float:
        v = ct.c_double()
        self.ctx.call("g2ohip_lambda_init", ct.byref(v))
        return v.value

    def system(self):
        """(rows, cols, blocks [n,6,6], b) of the last build_system, undamped upper blocks."""
        rows = np.zeros(self.nblocks, dtype=np.int32)
        cols = np.zeros(self.nblocks, dtype=np.int32)
        blocks = np.zeros((self.nblocks, 6, 6))
        b = np.zeros(6 * self.Nf)
        self.ctx.call("g2ohip_pg_get_system", _ip(rows), _ip(cols), _dp(blocks), _dp(b))
        return rows, cols, blocks, b

    def optimize(self, iterations: int):
        stats = (L.IterStats * max(iterations, 1))()
        done = ct.c_int()
        self.ctx.call("g2ohip_optimize", iterations, stats, ct.byref(done))
        return [{f: getattr(s, f) for f, _ in L.IterStats._fields_} for s in stats[:done.value]]

    def last_solve_info(self):
        """[PCG iterations, 0, PCG residual memory, 0] of the last solve."""
        v = np.zeros(4)
        self.ctx.call("g2ohip_last_solve_info", _dp(v))
        return v

    def lm_trials(self) -> np.ndarray:
        """Per LM trial of the last optimize(): columns TRIAL_FIELDS."""
        return _lm_trials(self.ctx)

    def set_profiling(self, on: bool = True):
        self.ctx.call("g2ohip_set_profiling", int(on))

    def profile(self) -> dict:
        out = {}
        for k, name in enumerate(L.KERNEL_NAMES):
            t = ct.c_double()
            n = ct.c_int()
            self.ctx.call("g2ohip_get_profile", k, ct.byref(t), ct.byref(n))
            out[name] = (t.value, n.value)
        return out


def linear_solve_pcg(ctx: Context, nb: int, bd: int, rows, cols, blocks, b, tol=1e-6, max_iter=-1, residual=-1.0):
    """LinearSolverPCG::solve on a host block matrix (upper blocks).
    Returns (x, iterations, residual memory)."""
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    blocks = np.ascontiguousarray(blocks, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(nb * bd)
    r = ct.c_double(residual)
    it = ct.c_int()
    ctx.call("g2ohip_linear_solve_pcg", nb, bd, len(rows), _ip(rows), _ip(cols), _dp(blocks), _dp(b), _dp(x),
             float(tol), int(max_iter), ct.byref(r), ct.byref(it))
    return x, it.value, r.value


def full_block_rows(nb: int, bd: int, rows, cols, blocks):
    """Upper blocks (row <= col) -> full block rows of the symmetric matrix
    (both triangles, columns sorted within a row): (F_ptr, F_col, F_val,
    F_diag), the layout g2ohip_linear_solve_pcg_device takes."""
    rows = np.asarray(rows, dtype=np.int64)
    cols = np.asarray(cols, dtype=np.int64)
    blocks = np.asarray(blocks, dtype=np.float64).reshape(-1, bd, bd)
    off = rows != cols
    r = np.concatenate([rows, cols[off]])
    c = np.concatenate([cols, rows[off]])
    v = np.concatenate([blocks, blocks[off].transpose(0, 2, 1)])
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    F_ptr = np.zeros(nb + 1, dtype=np.int32)
    np.add.at(F_ptr, r + 1, 1)
    F_ptr = np.cumsum(F_ptr).astype(np.int32)
    diag = np.flatnonzero(r == c)
    if diag.size != nb:
        raise ValueError("every block row needs its diagonal block")
    F_diag = diag.astype(np.int32)
    return F_ptr, c.astype(np.int32), np.ascontiguousarray(v), F_diag


def linear_solve_pcg_device(ctx: Context, nb: int, bd: int, F_ptr, F_col, F_val, F_diag, b, x, tol=1e-6,
                            max_iter=-1, residual=-1.0):
    """LinearSolverPCG::solve on a device matrix, allocating nothing
    (g2ohip_linear_solve_pcg_device). F_ptr / F_col / F_diag (int32),
    F_val / b / x (float64) are device addresses (ints) or objects with a
    data_ptr() (e.g. contiguous CUDA tensors of those dtypes); x is
    overwritten. Returns (iterations, residual memory)."""
    def addr(a):
        return int(a.data_ptr()) if hasattr(a, "data_ptr") else int(a)
    r = ct.c_double(residual)
    it = ct.c_int()
    ctx.call("g2ohip_linear_solve_pcg_device", nb, bd, addr(F_ptr), addr(F_col), addr(F_val), addr(F_diag),
             addr(b), addr(x), float(tol), int(max_iter), ct.byref(r), ct.byref(it))
    return it.value, r.value


class BALAdjuster:
    """SparseOptimizer with OptimizationAlgorithmLevenberg and
    BlockSolver<BlockSolverTraits<9, 3>> + LinearSolverPCG over g2o's BAL
    types (g2o/examples/bal/bal_example.cpp:70-207, 227-341): VertexCameraBAL
    [w, t, f, k1, k2] with additive updates, VertexPointBAL, EdgeObservationBAL
    (``g2o_amd.datasets.BALProblem``). FP64, explicit Schur."""

    def __init__(self, prob, device: int = 0, pcg_tolerance: float = 1e-6, pcg_max_iter: int = -1,
                 max_trials: int = 10):
        self.ctx = Context(device)
        self.prob = prob
        self.ctx.set_options(precision=L.FP64, schur_mode=L.SCHUR_EXPLICIT, pcg_tolerance=pcg_tolerance,
                             pcg_max_iter=pcg_max_iter, max_trials=max_trials)
        self.C, self.P, self.E = prob.num_cams, prob.num_pts, prob.num_edges
        self._ec = np.ascontiguousarray(prob.edge_cam, dtype=np.int32)
        self._ep = np.ascontiguousarray(prob.edge_pt, dtype=np.int32)
        cf = prob.cam_fixed if prob.cam_fixed is not None else np.zeros(self.C, np.uint8)
        pf = prob.pt_fixed if prob.pt_fixed is not None else np.zeros(self.P, np.uint8)
        self._cf = np.ascontiguousarray(cf, dtype=np.uint8)
        self._pf = np.ascontiguousarray(pf, dtype=np.uint8)
        self.ctx.call("g2ohip_bal_set_structure", self.C, self.P, self.E, _ip(self._ec), _ip(self._ep),
                      _up(self._cf), _up(self._pf))
        info = self.structure_info()
        self.Cf, self.Pf, self.nschur = int(info[0]), int(info[1]), int(info[3])
        if device >= 0:
            self.set_measurements(prob.meas, prob.weight, prob.huber)
            self.set_state(prob.cams, prob.pts)

    def set_options(self, **kw):
        return self.ctx.set_options(**kw)

    def set_algorithm(self, algorithm: int):
        """SparseOptimizer::setAlgorithm: L.LEVENBERG or L.GAUSS_NEWTON."""
        self.ctx.call("g2ohip_set_algorithm", int(algorithm))

    def structure_info(self) -> np.ndarray:
        """[free cameras, free points, active edges, Hschur upper blocks, blocks
        in both triangles, off-diagonal pair products, Hpl blocks]"""
        info = (ct.c_longlong * 7)()
        self.ctx.check(self.ctx._lib.g2ohip_bal_structure_info(self.ctx.h, info))
        return np.array(info[:], dtype=np.int64)

    def set_measurements(self, meas, weight=None, huber: float = 0.0):
        m = np.ascontiguousarray(meas, dtype=np.float64)
        w = None if weight is None else np.ascontiguousarray(weight, dtype=np.float64)
        self._meas, self._w = m, w
        self.ctx.call("g2ohip_bal_set_measurements", _dp(m), None if w is None else _dp(w), float(huber or 0.0))

    def set_state(self, cams, pts):
        c = np.ascontiguousarray(cams, dtype=np.float64)
        p = np.ascontiguousarray(pts, dtype=np.float64)
        self.ctx.call("g2ohip_bal_set_state", _dp(c), _dp(p))

    def state(self):
        c = np.zeros((self.C, 9))
        p = np.zeros((self.P, 3))
        self.ctx.call("g2ohip_bal_get_state", _dp(c), _dp(p))
        return c, p

    def compute_active_errors(self) -> float:
        v = ct.c_double()
        self.ctx.call("g2ohip_compute_active_errors", ct.byref(v))
        return v.value

    def errors(self) -> np.ndarray:
        e = np.zeros((self.E, 2))
        self.ctx.call("g2ohip_get_errors", _dp(e))
        return e

    def optimize(self, iterations: int):
        stats = (L.IterStats * max(iterations, 1))()
        done = ct.c_int()
        self.ctx.call("g2ohip_optimize", iterations, stats, ct.byref(done))
        return [{f: getattr(s, f) for f, _ in L.IterStats._fields_} for s in stats[:done.value]]

    def build_system(self):
        self.ctx.call("g2ohip_build_system")

    def set_lambda(self, lam: float):
        self.ctx.call("g2ohip_set_lambda", float(lam))

    def solve(self) -> bool:
        ok = ct.c_int()
        self.ctx.call("g2ohip_solve", ct.byref(ok))
        return bool(ok.value)

    def x(self) -> np.ndarray:
        v = np.zeros(9 * self.Cf + 3 * self.Pf)
        self.ctx.call("g2ohip_get_x", _dp(v))
        return v

    def b(self) -> np.ndarray:
        v = np.zeros(9 * self.Cf + 3 * self.Pf)
        self.ctx.call("g2ohip_get_b", _dp(v))
        return v

    def update(self):
        self.ctx.call("g2ohip_update")

    def push(self):
        self.ctx.call("g2ohip_push")

    def pop(self):
        self.ctx.call("g2ohip_pop")

    def discard_top(self):
        self.ctx.call("g2ohip_discard_top")

    def lambda_init(self) -> float:
        v = ct.c_double()
        self.ctx.call("g2ohip_lambda_init", ct.byref(v))
        return v.value

    def compute_scale(self) -> float:
        v = ct.c_double()
        self.ctx.call("g2ohip_compute_scale", ct.byref(v))
        return v.value

    def hessian(self):
        Hpp = np.zeros((self.Cf, 9, 9))
        Hll = np.zeros((self.Pf, 3, 3))
        Hpl = np.zeros((self.E, 9, 3))
        self.ctx.call("g2ohip_bal_get_hessian", _dp(Hpp), _dp(Hll), _dp(Hpl))
        return Hpp, Hll, Hpl

    def schur(self):
        rows = np.zeros(self.nschur, dtype=np.int32)
        cols = np.zeros(self.nschur, dtype=np.int32)
        blocks = np.zeros((self.nschur, 9, 9))
        bs = np.zeros(9 * self.Cf)
        self.ctx.call("g2ohip_bal_get_schur", _ip(rows), _ip(cols), _dp(blocks), _dp(bs))
        return rows, cols, blocks, bs

    def last_solve_info(self):
        v = np.zeros(4)
        self.ctx.call("g2ohip_last_solve_info", _dp(v))
        return v

    def lm_trials(self) -> np.ndarray:
        """Per LM trial of the last optimize(): columns TRIAL_FIELDS."""
        return _lm_trials(self.ctx)

    def set_profiling(self, on: bool = True):
        self.ctx.call("g2ohip_set_profiling", int(on))

    def profile(self) -> dict:
        out = {}
        for k, name in enumerate(L.KERNEL_NAMES):
            t = ct.c_double()
            n = ct.c_int()
            self.ctx.call("g2ohip_get_profile", k, ct.byref(t), ct.byref(n))
            out[name] = (t.value, n.value)
        return out
