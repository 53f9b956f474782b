"""ctypes wrapper of the FP64 CPU oracle (oracle/g2o_oracle.cpp).

Test infrastructure only: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "libg2o_oracle.so")

_dp = ct.POINTER(ct.c_double)
_ip = ct.POINTER(ct.c_int)
_up = ct.POINTER(ct.c_ubyte)


def build(force: bool = False) -> str:
    """Compile the oracle with make (g++ -fopenmp)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ct.CDLL(LIB_PATH)
        L.orc_ba_new.restype = ct.c_void_p
        L.orc_ba_new.argtypes = [ct.c_int, ct.c_int, ct.c_int, _dp, _dp, _dp, _ip, _ip, _dp, _dp, _up, _up,
                                 ct.c_double]
        L.orc_ba_free.argtypes = [ct.c_void_p]
        L.orc_ba_dims.argtypes = [ct.c_void_p, _ip, _ip, _ip]
        L.orc_ba_get_state.argtypes = [ct.c_void_p, _dp, _dp]
        L.orc_ba_set_state.argtypes = [ct.c_void_p, _dp, _dp]
        L.orc_ba_chi2.restype = ct.c_double
        L.orc_ba_chi2.argtypes = [ct.c_void_p]
        L.orc_ba_linearize.restype = ct.c_double
        L.orc_ba_linearize.argtypes = [ct.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_ba_schur.restype = ct.c_int
        L.orc_ba_schur.argtypes = [ct.c_void_p, ct.c_double, _ip, _ip, _dp, _dp]
        L.orc_ba_solve.restype = ct.c_int
        L.orc_ba_solve.argtypes = [ct.c_void_p, ct.c_double, ct.c_int, ct.c_double, ct.c_int, _dp, _ip]
        L.orc_ba_optimize.restype = ct.c_int
        L.orc_ba_optimize.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_double, ct.c_int, _dp, _dp]
        L.orc_ba_gauss_newton.restype = ct.c_int
        L.orc_ba_gauss_newton.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_double, ct.c_int, _dp]
        L.orc_se3_exp.argtypes = [_dp, _dp]
        L.orc_se3_log.argtypes = [_dp, _dp]
        L.orc_se3_mul.argtypes = [_dp, _dp, _dp]
        L.orc_quat_from_R.argtypes = [_dp, _dp]
        L.orc_quat_to_R.argtypes = [_dp, _dp]
        L.orc_huber.argtypes = [ct.c_double, ct.c_double, _dp]
        L.orc_project_edge.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_pcg_blocks.restype = ct.c_int
        L.orc_pcg_blocks.argtypes = [ct.c_int, ct.c_int, ct.c_int, _ip, _ip, _dp, _dp, _dp, ct.c_double, ct.c_int,
                                     _dp]
        L.orc_chol_solve.restype = ct.c_int
        L.orc_chol_solve.argtypes = [ct.c_int, _dp, _dp, _dp]
        L.orc_ba_set_meas.argtypes = [ct.c_void_p, _dp, _dp]
        L.orc_ba_set_info.argtypes = [ct.c_void_p, _dp]
        L.orc_ba_set_user_lambda.argtypes = [ct.c_void_p, ct.c_double]
        L.orc_ba_trials.restype = ct.c_int
        L.orc_ba_trials.argtypes = [ct.c_void_p, _dp, ct.c_int]
        L.orc_pg_trials.restype = ct.c_int
        L.orc_pg_trials.argtypes = [ct.c_void_p, _dp, ct.c_int]
        L.orc_ba_adjoint.restype = ct.c_int
        L.orc_ba_adjoint.argtypes = [ct.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.orc_ba_newton.restype = ct.c_int
        L.orc_ba_newton.argtypes = [ct.c_void_p, ct.c_int, ct.c_double]
        L.orc_ba_exact_hessian.argtypes = [ct.c_void_p, _dp, _dp, _dp]
        L.orc_edge_energy_derivs.restype = ct.c_double
        L.orc_edge_energy_derivs.argtypes = [_dp, _dp, _dp, ct.c_double, _dp, ct.c_double, _dp, _dp]
        L.orc_pg_new.restype = ct.c_void_p
        L.orc_pg_new.argtypes = [ct.c_int, ct.c_int, _dp, _ip, _ip, _dp, _dp, _up]
        L.orc_pg_free.argtypes = [ct.c_void_p]
        L.orc_pg_nfree.restype = ct.c_int
        L.orc_pg_nfree.argtypes = [ct.c_void_p]
        L.orc_pg_nblocks.restype = ct.c_int
        L.orc_pg_nblocks.argtypes = [ct.c_void_p]
        L.orc_pg_get_state.argtypes = [ct.c_void_p, _dp]
        L.orc_pg_set_state.argtypes = [ct.c_void_p, _dp]
        L.orc_pg_chi2.restype = ct.c_double
        L.orc_pg_chi2.argtypes = [ct.c_void_p]
        L.orc_pg_edge.argtypes = [ct.c_void_p, ct.c_int, _dp, _dp]
        L.orc_pg_linearize.restype = ct.c_double
        L.orc_pg_linearize.argtypes = [ct.c_void_p, _ip, _ip, _dp, _dp]
        L.orc_pg_optimize.restype = ct.c_int
        L.orc_pg_optimize.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_double, ct.c_int, _dp, _dp]
        L.orc_ba_dogleg.restype = ct.c_int
        L.orc_ba_dogleg.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_double, ct.c_int, _dp]
        L.orc_pg_dogleg.restype = ct.c_int
        L.orc_pg_dogleg.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_double, ct.c_int, _dp]
        L.orc_pg_gauss_newton.restype = ct.c_int
        L.orc_pg_gauss_newton.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_double, ct.c_int, _dp]
        L.orc_bal_new.restype = ct.c_void_p
        L.orc_bal_new.argtypes = [ct.c_int, ct.c_int, ct.c_int, _dp, _dp, _ip, _ip, _dp, _dp, _up, _up, ct.c_double]
        L.orc_bal_free.argtypes = [ct.c_void_p]
        L.orc_bal_dims.argtypes = [ct.c_void_p, _ip, _ip, _ip]
        L.orc_bal_get_state.argtypes = [ct.c_void_p, _dp, _dp]
        L.orc_bal_set_state.argtypes = [ct.c_void_p, _dp, _dp]
        L.orc_bal_chi2.restype = ct.c_double
        L.orc_bal_chi2.argtypes = [ct.c_void_p, _dp]
        L.orc_bal_edge.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_bal_linearize.restype = ct.c_double
        L.orc_bal_linearize.argtypes = [ct.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_bal_schur.restype = ct.c_int
        L.orc_bal_schur.argtypes = [ct.c_void_p, ct.c_double, _ip, _ip, _dp, _dp]
        L.orc_bal_solve.restype = ct.c_int
        L.orc_bal_solve.argtypes = [ct.c_void_p, ct.c_double, ct.c_int, ct.c_double, ct.c_int, _dp, _ip]
        L.orc_bal_optimize.restype = ct.c_int
        L.orc_bal_optimize.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_double, ct.c_int, _dp, _dp]
        L.orc_bal_trials.restype = ct.c_int
        L.orc_bal_trials.argtypes = [ct.c_void_p, _dp, ct.c_int]
        L.orc_num_threads.restype = ct.c_int
        L.orc_set_num_threads.argtypes = [ct.c_int]
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def _u(a):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    return a, a.ctypes.data_as(_up)


def _trials(fn, h):
    n = fn(h, None, 0)
    out = np.zeros((max(n, 1), 6))
    fn(h, out.ctypes.data_as(_dp), n)
    return out[:n]


class OracleBA:
    """FP64 CPU restatement of g2o's BlockSolver_6_3 + LM on one BA problem."""

    def __init__(self, prob):
        L = lib()
        keep = []
        args = []
        for arr, conv in ((prob.cams, _d), (prob.intr, _d), (prob.pts, _d), (prob.edge_cam, _i),
                          (prob.edge_pt, _i), (prob.meas, _d), (prob.weight, _d), (prob.cam_fixed, _u),
                          (prob.pt_fixed, _u)):
            a, p = conv(arr)
            keep.append(a)
            args.append(p)
        self.C, self.P, self.E = prob.num_cams, prob.num_pts, prob.num_edges
        self.h = L.orc_ba_new(self.C, self.P, self.E, *args, float(prob.huber))
        cf, pf, ns = ct.c_int(), ct.c_int(), ct.c_int()
        L.orc_ba_dims(self.h, ct.byref(cf), ct.byref(pf), ct.byref(ns))
        self.Cf, self.Pf, self.nschur = cf.value, pf.value, ns.value
        if getattr(prob, "omega", None) is not None:
            self.set_info(prob.omega)

    def set_info(self, omega):
        """Per-edge information [O00, O01, O11] (E x 3); None: identity."""
        if omega is None:
            lib().orc_ba_set_info(self.h, None)
            return
        o, po = _d(np.asarray(omega, dtype=np.float64).reshape(-1, 3))
        lib().orc_ba_set_info(self.h, po)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_ba_free(self.h)
            self.h = None

    def state(self):
        c = np.zeros((self.C, 7))
        p = np.zeros((self.P, 3))
        lib().orc_ba_get_state(self.h, _d(c)[1], _d(p)[1])
        return c, p

    def set_state(self, cams, pts):
        c, pc = _d(cams)
        p, pp = _d(pts)
        lib().orc_ba_set_state(self.h, pc, pp)

    def chi2(self) -> float:
        return lib().orc_ba_chi2(self.h)

    def linearize(self):
        Hpp = np.zeros((self.Cf, 6, 6))
        Hll = np.zeros((self.Pf, 3, 3))
        Hpl = np.zeros((self.E, 6, 3))
        b = np.zeros(6 * self.Cf + 3 * self.Pf)
        chi = lib().orc_ba_linearize(self.h, Hpp.ctypes.data_as(_dp), Hll.ctypes.data_as(_dp),
                                     Hpl.ctypes.data_as(_dp), b.ctypes.data_as(_dp))
        return chi, Hpp, Hll, Hpl, b

    def schur(self, lam: float):
        rows = np.zeros(self.nschur, dtype=np.int32)
        cols = np.zeros(self.nschur, dtype=np.int32)
        blocks = np.zeros((self.nschur, 6, 6))
        bs = np.zeros(6 * self.Cf)
        st = lib().orc_ba_schur(self.h, lam, rows.ctypes.data_as(_ip), cols.ctypes.data_as(_ip),
                                blocks.ctypes.data_as(_dp), bs.ctypes.data_as(_dp))
        return st, rows, cols, blocks, bs

    def solve(self, lam: float, solver: int = 0, tol: float = 1e-6, maxiter: int = -1):
        dx = np.zeros(6 * self.Cf + 3 * self.Pf)
        it = ct.c_int()
        st = lib().orc_ba_solve(self.h, lam, solver, tol, maxiter, dx.ctypes.data_as(_dp), ct.byref(it))
        return st, dx, it.value

    def optimize(self, iters: int, solver: int = 1, tol: float = 1e-6, maxiter: int = -1):
        trace = np.zeros((iters, 6))
        times = np.zeros(6)
        n = lib().orc_ba_optimize(self.h, iters, solver, tol, maxiter, trace.ctypes.data_as(_dp),
                                  times.ctypes.data_as(_dp))
        return n, trace[:n], times

    def gauss_newton(self, iters: int, solver: int = 1, tol: float = 1e-6, maxiter: int = -1):
        """Rows: chi2 before, chi2 after, pcg iterations, status (1 = Fail)."""
        trace = np.zeros((iters, 4))
        n = lib().orc_ba_gauss_newton(self.h, iters, solver, tol, maxiter, trace.ctypes.data_as(_dp))
        return n, trace[:n]

    def dogleg(self, iters: int, solver: int = 1, tol: float = 1e-6, maxiter: int = -1):
        """Rows: chi2 before, chi2 after, delta, tries, pcg iterations, status."""
        trace = np.zeros((iters, 6))
        n = lib().orc_ba_dogleg(self.h, iters, solver, tol, maxiter, trace.ctypes.data_as(_dp))
        return n, trace[:n]

    def set_user_lambda(self, lam: float):
        lib().orc_ba_set_user_lambda(self.h, float(lam))

    def trials(self):
        """LM trials of the last optimize(): rows of (lambda used, currentChi,
        tempChi, rho divisor, rho, accepted)."""
        return _trials(lib().orc_ba_trials, self.h)

    def set_meas(self, meas, weight):
        m, pm = _d(meas)
        w, pw = _d(weight)
        lib().orc_ba_set_meas(self.h, pm, pw)

    def adjoint(self, dL_dcam, dL_dpt):
        """Implicit-function gradient of a loss of the converged state with
        respect to the measurements and information weights.
        Returns (status, dL_dmeas [E,2], dL_dw [E], lambda [6Cf+3Pf])."""
        dc, pdc = _d(dL_dcam)
        dp_, pdp = _d(dL_dpt)
        dz = np.zeros((self.E, 2))
        dw = np.zeros(self.E)
        lam = np.zeros(6 * self.Cf + 3 * self.Pf)
        st = lib().orc_ba_adjoint(self.h, pdc, pdp, dz.ctypes.data_as(_dp), dw.ctypes.data_as(_dp),
                                  lam.ctypes.data_as(_dp))
        return st, dz, dw, lam

    def newton(self, steps: int, max_step: float = 1.0) -> int:
        """Damped exact-Hessian Newton steps with a trust region (0 = ok, 2 =
        no acceptable step)."""
        return lib().orc_ba_newton(self.h, steps, float(max_step))

    def exact_hessian(self):
        Hpp = np.zeros((self.Cf, 6, 6))
        Hll = np.zeros((self.Pf, 3, 3))
        Hpl = np.zeros((self.E, 6, 3))
        lib().orc_ba_exact_hessian(self.h, Hpp.ctypes.data_as(_dp), Hll.ctypes.data_as(_dp),
                                   Hpl.ctypes.data_as(_dp))
        return Hpp, Hll, Hpl


def bal_edge(cam9, X, z, jacobians: bool = True):
    """EdgeObservationBAL's error (prediction - z) and, by dual numbers, its
    Jacobians: returns e[2], Jc[2,9], Jp[2,3] (Jacobians None if not asked)."""
    c, pc = _d(cam9)
    x, px = _d(X)
    zz, pz = _d(z)
    e = np.zeros(2)
    Jc = np.zeros((2, 9)) if jacobians else None
    Jp = np.zeros((2, 3)) if jacobians else None
    lib().orc_bal_edge(pc, px, pz, e.ctypes.data_as(_dp), None if Jc is None else Jc.ctypes.data_as(_dp),
                       None if Jp is None else Jp.ctypes.data_as(_dp))
    return e, Jc, Jp


class OracleBAL:
    """FP64 CPU restatement of g2o's BAL example: VertexCameraBAL (9 values,
    additive update), EdgeObservationBAL, BlockSolver<9, 3> with Schur and LM."""

    def __init__(self, prob):
        L = lib()
        C, P, E = prob.cams.shape[0], prob.pts.shape[0], prob.edge_cam.shape[0]
        w = np.ones(E) if getattr(prob, "weight", None) is None else prob.weight
        cf = np.zeros(C, np.uint8) if getattr(prob, "cam_fixed", None) is None else prob.cam_fixed
        pf = np.zeros(P, np.uint8) if getattr(prob, "pt_fixed", None) is None else prob.pt_fixed
        keep, args = [], []
        for arr, conv in ((prob.cams, _d), (prob.pts, _d), (prob.edge_cam, _i), (prob.edge_pt, _i),
                          (prob.meas, _d), (w, _d), (cf, _u), (pf, _u)):
            a, p = conv(arr)
            keep.append(a)
            args.append(p)
        self.C, self.P, self.E = C, P, E
        self.h = L.orc_bal_new(C, P, E, *args, float(getattr(prob, "huber", 0.0) or 0.0))
        cf_, pf_, ns = ct.c_int(), ct.c_int(), ct.c_int()
        L.orc_bal_dims(self.h, ct.byref(cf_), ct.byref(pf_), ct.byref(ns))
        self.Cf, self.Pf, self.nschur = cf_.value, pf_.value, ns.value

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_bal_free(self.h)
            self.h = None

    def state(self):
        c = np.zeros((self.C, 9))
        p = np.zeros((self.P, 3))
        lib().orc_bal_get_state(self.h, _d(c)[1], _d(p)[1])
        return c, p

    def set_state(self, cams, pts):
        c, pc = _d(cams)
        p, pp = _d(pts)
        lib().orc_bal_set_state(self.h, pc, pp)

    def chi2(self) -> float:
        return lib().orc_bal_chi2(self.h, None)

    def errors(self):
        e = np.zeros((self.E, 2))
        lib().orc_bal_chi2(self.h, e.ctypes.data_as(_dp))
        return e

    def linearize(self):
        Hpp = np.zeros((self.Cf, 9, 9))
        Hll = np.zeros((self.Pf, 3, 3))
        Hpl = np.zeros((self.E, 9, 3))
        b = np.zeros(9 * self.Cf + 3 * self.Pf)
        chi = lib().orc_bal_linearize(self.h, Hpp.ctypes.data_as(_dp), Hll.ctypes.data_as(_dp),
                                      Hpl.ctypes.data_as(_dp), b.ctypes.data_as(_dp))
        return chi, Hpp, Hll, Hpl, b

    def schur(self, lam: float):
        rows = np.zeros(self.nschur, dtype=np.int32)
        cols = np.zeros(self.nschur, dtype=np.int32)
        blocks = np.zeros((self.nschur, 9, 9))
        bs = np.zeros(9 * self.Cf)
        st = lib().orc_bal_schur(self.h, lam, rows.ctypes.data_as(_ip), cols.ctypes.data_as(_ip),
                                 blocks.ctypes.data_as(_dp), bs.ctypes.data_as(_dp))
        return st, rows, cols, blocks, bs

    def solve(self, lam: float, solver: int = 0, tol: float = 1e-6, maxiter: int = -1):
        dx = np.zeros(9 * self.Cf + 3 * self.Pf)
        it = ct.c_int()
        st = lib().orc_bal_solve(self.h, lam, solver, tol, maxiter, dx.ctypes.data_as(_dp), ct.byref(it))
        return st, dx, it.value

    def optimize(self, iters: int, solver: int = 1, tol: float = 1e-6, maxiter: int = -1):
        trace = np.zeros((iters, 6))
        times = np.zeros(6)
        n = lib().orc_bal_optimize(self.h, iters, solver, tol, maxiter, trace.ctypes.data_as(_dp),
                                   times.ctypes.data_as(_dp))
        return n, trace[:n], times

    def trials(self):
        return _trials(lib().orc_bal_trials, self.h)


class OraclePG:
    """FP64 CPU restatement of g2o's VertexSE3 / EdgeSE3 pose graph with LM
    over 6x6 blocks (no Schur)."""

    def __init__(self, prob):
        L = lib()
        self._keep = [np.ascontiguousarray(a, dtype=d) for a, d in (
            (prob.poses, np.float64), (prob.edge_from, np.int32), (prob.edge_to, np.int32),
            (prob.meas, np.float64), (prob.info, np.float64), (prob.fixed, np.uint8))]
        po, fr, to, me, inf, fx = self._keep
        self.N, self.E = prob.num_poses, prob.num_edges
        self.h = L.orc_pg_new(self.N, self.E, po.ctypes.data_as(_dp), fr.ctypes.data_as(_ip), to.ctypes.data_as(_ip),
                              me.ctypes.data_as(_dp), inf.ctypes.data_as(_dp), fx.ctypes.data_as(_up))
        self.Nf = L.orc_pg_nfree(self.h)
        self.nblocks = L.orc_pg_nblocks(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_pg_free(self.h)
            self.h = None

    def state(self):
        out = np.zeros((self.N, 7))
        lib().orc_pg_get_state(self.h, out.ctypes.data_as(_dp))
        return out

    def set_state(self, poses):
        a, p = _d(poses)
        lib().orc_pg_set_state(self.h, p)

    def chi2(self) -> float:
        return lib().orc_pg_chi2(self.h)

    def edge(self, k: int):
        e = np.zeros(6)
        J = np.zeros((6, 12))
        lib().orc_pg_edge(self.h, int(k), e.ctypes.data_as(_dp), J.ctypes.data_as(_dp))
        return e, J

    def linearize(self):
        rows = np.zeros(self.nblocks, dtype=np.int32)
        cols = np.zeros(self.nblocks, dtype=np.int32)
        blocks = np.zeros((self.nblocks, 6, 6))
        b = np.zeros(6 * self.Nf)
        chi = lib().orc_pg_linearize(self.h, rows.ctypes.data_as(_ip), cols.ctypes.data_as(_ip),
                                     blocks.ctypes.data_as(_dp), b.ctypes.data_as(_dp))
        return chi, rows, cols, blocks, b

    def optimize(self, iters: int, solver: int = 1, tol: float = 1e-8, maxiter: int = -1):
        trace = np.zeros((iters, 6))
        times = np.zeros(6)
        n = lib().orc_pg_optimize(self.h, iters, solver, tol, maxiter, trace.ctypes.data_as(_dp),
                                  times.ctypes.data_as(_dp))
        return n, trace[:n], times

    def trials(self):
        """LM trials of the last optimize() (columns as OracleBA.trials)."""
        return _trials(lib().orc_pg_trials, self.h)

    def gauss_newton(self, iters: int, solver: int = 1, tol: float = 1e-8, maxiter: int = -1):
        """Rows: chi2 before, chi2 after, pcg iterations, status (1 = Fail)."""
        trace = np.zeros((iters, 4))
        n = lib().orc_pg_gauss_newton(self.h, iters, solver, tol, maxiter, trace.ctypes.data_as(_dp))
        return n, trace[:n]

    def dogleg(self, iters: int, solver: int = 1, tol: float = 1e-8, maxiter: int = -1):
        """Rows: chi2 before, chi2 after, delta, tries, pcg iterations, status."""
        trace = np.zeros((iters, 6))
        n = lib().orc_pg_dogleg(self.h, iters, solver, tol, maxiter, trace.ctypes.data_as(_dp))
        return n, trace[:n]


def edge_energy_derivs(cam7, X, z, w, intr4, huber_delta):
    """0.5 rho of one projection edge, its gradient [12] and Hessian [12,12] in
    (omega, upsilon, dX, dz, dw) at zero (hyper-dual evaluation)."""
    g = np.zeros(12)
    H = np.zeros((12, 12))
    E = lib().orc_edge_energy_derivs(_d(cam7)[1], _d(X)[1], _d(z)[1], float(w), _d(intr4)[1], float(huber_delta),
                                     g.ctypes.data_as(_dp), H.ctypes.data_as(_dp))
    return E, g, H


def se3_exp(upd):
    out = np.zeros(7)
    lib().orc_se3_exp(_d(upd)[1], out.ctypes.data_as(_dp))
    return out


def se3_log(p7):
    out = np.zeros(6)
    lib().orc_se3_log(_d(p7)[1], out.ctypes.data_as(_dp))
    return out


def se3_mul(a7, b7):
    out = np.zeros(7)
    lib().orc_se3_mul(_d(a7)[1], _d(b7)[1], out.ctypes.data_as(_dp))
    return out


def quat_from_R(R):
    out = np.zeros(4)
    lib().orc_quat_from_R(_d(R)[1], out.ctypes.data_as(_dp))
    return out


def quat_to_R(q):
    out = np.zeros((3, 3))
    lib().orc_quat_to_R(_d(q)[1], out.ctypes.data_as(_dp))
    return out


def huber(e, delta):
    out = np.zeros(3)
    lib().orc_huber(float(e), float(delta), out.ctypes.data_as(_dp))
    return out


def project_edge(cam7, intr4, X, z):
    e = np.zeros(2)
    Jp = np.zeros((2, 3))
    Jc = np.zeros((2, 6))
    lib().orc_project_edge(_d(cam7)[1], _d(intr4)[1], _d(X)[1], _d(z)[1], e.ctypes.data_as(_dp),
                           Jp.ctypes.data_as(_dp), Jc.ctypes.data_as(_dp))
    return e, Jp, Jc


def pcg_blocks(nb, bd, rows, cols, blocks, b, tol=1e-6, maxiter=-1, residual=-1.0):
    """Returns (x, iterations, residual memory)."""
    x = np.zeros(nb * bd)
    r = ct.c_double(residual)
    rr, rp = _i(rows)
    cc, cp = _i(cols)
    bb, bp = _d(blocks)
    b2, b2p = _d(b)
    it = lib().orc_pcg_blocks(nb, bd, len(rows), rp, cp, bp, b2p, x.ctypes.data_as(_dp), tol, maxiter,
                              ct.byref(r))
    return x, it, r.value


def chol_solve(A, b):
    A, ap = _d(A)
    b, bp = _d(b)
    x = np.zeros(b.size)
    st = lib().orc_chol_solve(b.size, ap, bp, x.ctypes.data_as(_dp))
    return st, x
