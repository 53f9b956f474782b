"""g2o text-format and BAL files, read and written by libg2o_hip.so's native
graph container (include/g2o_hip.h, ``g2ohip_graph_*``).

``G2oGraph`` mirrors the part of g2o's ``OptimizableGraph`` that loads and
saves graphs (g2o/core/optimizable_graph.cpp:344-567) for the path's types:
VERTEX_SE3:EXPMAP cameras, VERTEX_TRACKXYZ points and
EDGE_SE3_PROJECT_XYZ:EXPMAP projections (bundle adjustment), VERTEX_SE3:QUAT
poses and EDGE_SE3:QUAT edges (3D pose graphs). Its arrays come out in the
layout ``BundleAdjuster`` / ``PoseGraphOptimizer`` take, so a file runs
through the GPU path unchanged:

    g = G2oGraph.load("ba.g2o")
    prob = g.to_ba_problem(intr)          # intrinsics are not in the format
    ba = BundleAdjuster(prob); ba.optimize(10)
    g.set_ba_state(*ba.state()); g.save("ba_opt.g2o")

``load_bal`` reads the BAL format as g2o's BAL example parses it
(g2o/examples/bal/bal_example.cpp:265-335); ``bal_to_ba`` maps its 9-parameter
cameras onto the path's 6-DoF pinhole model.
"""
from __future__ import annotations

import ctypes as ct
import os
import numpy as np

from . import _lib as L
from .datasets import BALProblem, BAProblem, PoseGraphProblem


class G2oIOError(RuntimeError):
    pass


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(ct.POINTER(t))


class G2oGraph:
    """Native graph container; see ``g2ohip_graph_*`` in include/g2o_hip.h."""

    def __init__(self):
        self._lib = L.lib()
        h = ct.c_void_p()
        st = self._lib.g2ohip_graph_create(ct.byref(h))
        if st != L.OK:
            raise G2oIOError(f"g2ohip_graph_create failed ({st})")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.g2ohip_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, st: int):
        if st != L.OK:
            msg = self._lib.g2ohip_graph_last_error(self._h).decode()
            raise G2oIOError(f"status {st}: {msg}")

    # ------------------------------------------------------------------ files
    @classmethod
    def load(cls, path: str, nthreads: int = 0) -> "G2oGraph":
        g = cls()
        g.append(path, nthreads)
        return g

    def append(self, path: str, nthreads: int = 0):
        """OptimizableGraph::load: adds the file's elements to this graph."""
        self._check(self._lib.g2ohip_graph_load(self._h, os.fsencode(path), int(nthreads)))

    def save(self, path: str, precision: int = 17):
        """OptimizableGraph::save. precision=6 prints as g2o's default
        ostream does; 17 round-trips every double exactly."""
        self._check(self._lib.g2ohip_graph_save(self._h, os.fsencode(path), int(precision)))

    @property
    def log(self) -> str:
        return self._lib.g2ohip_graph_log(self._h).decode()

    def info(self) -> dict:
        v = (ct.c_longlong * 12)()
        self._check(self._lib.g2ohip_graph_info(self._h, v))
        keys = ["cameras", "points", "projection_edges", "poses", "pose_edges", "fixed", "lines", "unknown_lines",
                "skipped", "bal_observations", "xyz2uv_edges", "camera_parameters"]
        return dict(zip(keys, (int(x) for x in v)))

    # ----------------------------------------------------- bundle adjustment
    def ba_arrays(self) -> dict:
        n = self.info()
        C, P, E = n["cameras"], n["points"], n["projection_edges"]
        out = dict(cam_id=np.empty(C, np.int32), cams=np.empty((C, 7)), cam_fixed=np.empty(C, np.uint8),
                   pt_id=np.empty(P, np.int32), pts=np.empty((P, 3)), pt_fixed=np.empty(P, np.uint8),
                   edge_cam=np.empty(E, np.int32), edge_pt=np.empty(E, np.int32), meas=np.empty((E, 2)),
                   omega=np.empty((E, 3)))
        o = out
        self._check(self._lib.g2ohip_graph_get_ba(
            self._h, _ptr(o["cam_id"], ct.c_int), _ptr(o["cams"], ct.c_double), _ptr(o["cam_fixed"], ct.c_ubyte),
            _ptr(o["pt_id"], ct.c_int), _ptr(o["pts"], ct.c_double), _ptr(o["pt_fixed"], ct.c_ubyte),
            _ptr(o["edge_cam"], ct.c_int), _ptr(o["edge_pt"], ct.c_int), _ptr(o["meas"], ct.c_double),
            _ptr(o["omega"], ct.c_double)))
        return out

    def ba_params(self) -> dict:
        """PARAMS_CAMERAPARAMETERS sorted by id (f cx cy baseline) and each
        projection edge's parameter index (-1 for EDGE_SE3_PROJECT_XYZ)."""
        n = self.info()
        o = dict(param_id=np.empty(n["camera_parameters"], np.int32), params=np.empty((n["camera_parameters"], 4)),
                 edge_param=np.empty(n["projection_edges"], np.int32))
        self._check(self._lib.g2ohip_graph_get_ba_params(self._h, _ptr(o["param_id"], ct.c_int),
                                                         _ptr(o["params"], ct.c_double),
                                                         _ptr(o["edge_param"], ct.c_int)))
        return o

    def camera_intrinsics(self) -> np.ndarray:
        """Per-camera fx fy cx cy from the CameraParameters of the cameras'
        EDGE_PROJECT_XYZ2UV edges (fx = fy = f, cam_map in
        parameter_cameraparameters.cpp:59-65). Every camera needs such edges,
        all with one parameter set, since the path keeps intrinsics per camera."""
        a = self.ba_arrays()
        pr = self.ba_params()
        C = a["cams"].shape[0]
        intr = np.full((C, 4), np.nan)
        for e in range(a["edge_cam"].shape[0]):
            k = pr["edge_param"][e]
            if k < 0:
                raise G2oIOError(f"edge {e} is an EDGE_SE3_PROJECT_XYZ: its intrinsics are not in the file")
            f, cx, cy, _ = pr["params"][k]
            row = np.array([f, f, cx, cy])
            c = a["edge_cam"][e]
            if np.isnan(intr[c, 0]):
                intr[c] = row
            elif not np.array_equal(intr[c], row):
                raise G2oIOError(f"camera {c} is observed through two different camera parameters")
        if np.isnan(intr).any():
            raise G2oIOError("a camera has no projection edge to take its intrinsics from")
        return intr

    def to_ba_problem(self, intr=None, huber: float = 0.0, name: str = "") -> BAProblem:
        """The graph's bundle-adjustment part as a BAProblem. ``intr`` is
        fx fy cx cy, one row for every camera or a single row for all; None
        takes them from the file's camera parameters (EDGE_PROJECT_XYZ2UV
        edges), since EDGE_SE3_PROJECT_XYZ keeps its intrinsics out of the
        format. Information matrices that are all w * I become the edge
        weights; otherwise the weights are 1 and ``omega`` holds each edge's
        full 2x2 matrix (O00 O01 O11)."""
        if intr is None:
            intr = self.camera_intrinsics()
        a = self.ba_arrays()
        om = a["omega"]
        scalar = not (om.size and (np.any(om[:, 1] != 0.0) or np.any(om[:, 0] != om[:, 2])))
        C = a["cams"].shape[0]
        intr = np.asarray(intr, np.float64).reshape(-1, 4)
        if intr.shape[0] == 1:
            intr = np.repeat(intr, C, axis=0)
        if intr.shape[0] != C:
            raise G2oIOError(f"intrinsics for {intr.shape[0]} cameras, graph has {C}")
        prob = BAProblem(a["cams"], intr.copy(), a["pts"], a["edge_cam"], a["edge_pt"], a["meas"],
                         om[:, 0].copy() if scalar else np.ones(om.shape[0]), a["cam_fixed"], a["pt_fixed"],
                         float(huber), name=name, info=dict(cam_id=a["cam_id"], pt_id=a["pt_id"]))
        if not scalar:
            prob.omega = om.copy()
        return prob

    def set_ba_state(self, cams, pts):
        cams = None if cams is None else np.ascontiguousarray(cams, np.float64)
        pts = None if pts is None else np.ascontiguousarray(pts, np.float64)
        n = self.info()
        if cams is not None and cams.shape != (n["cameras"], 7):
            raise G2oIOError(f"cams shape {cams.shape}, graph has {n['cameras']} cameras")
        if pts is not None and pts.shape != (n["points"], 3):
            raise G2oIOError(f"pts shape {pts.shape}, graph has {n['points']} points")
        self._check(self._lib.g2ohip_graph_set_ba_state(self._h, _ptr(cams, ct.c_double), _ptr(pts, ct.c_double)))

    def add_ba(self, prob: BAProblem, cam_id=None, pt_id=None):
        """Adds a BAProblem's vertices and edges; each edge's information is
        weight * omega (omega = I when the problem has none)."""
        C, P, E = prob.num_cams, prob.num_pts, prob.num_edges
        cams = np.ascontiguousarray(prob.cams, np.float64)
        pts = np.ascontiguousarray(prob.pts, np.float64)
        ec = np.ascontiguousarray(prob.edge_cam, np.int32)
        ep = np.ascontiguousarray(prob.edge_pt, np.int32)
        meas = np.ascontiguousarray(prob.meas, np.float64)
        w = np.ones(E) if prob.weight is None else np.asarray(prob.weight, np.float64)
        o = getattr(prob, "omega", None)
        base = np.tile([1.0, 0.0, 1.0], (E, 1)) if o is None else np.asarray(o, np.float64).reshape(E, 3)
        om = np.ascontiguousarray(base * w[:, None])
        cf = np.ascontiguousarray(prob.cam_fixed, np.uint8)
        pf = np.ascontiguousarray(prob.pt_fixed, np.uint8)
        ci = None if cam_id is None else np.ascontiguousarray(cam_id, np.int32)
        pi = None if pt_id is None else np.ascontiguousarray(pt_id, np.int32)
        self._check(self._lib.g2ohip_graph_add_ba(
            self._h, C, P, E, _ptr(ci, ct.c_int), _ptr(cams, ct.c_double), _ptr(cf, ct.c_ubyte),
            _ptr(pi, ct.c_int), _ptr(pts, ct.c_double), _ptr(pf, ct.c_ubyte), _ptr(ec, ct.c_int),
            _ptr(ep, ct.c_int), _ptr(meas, ct.c_double), _ptr(om, ct.c_double)))

    # ------------------------------------------------------------ pose graph
    def pg_arrays(self) -> dict:
        n = self.info()
        N, E = n["poses"], n["pose_edges"]
        o = dict(pose_id=np.empty(N, np.int32), poses=np.empty((N, 7)), fixed=np.empty(N, np.uint8),
                 edge_from=np.empty(E, np.int32), edge_to=np.empty(E, np.int32), meas=np.empty((E, 7)),
                 info=np.empty((E, 6, 6)))
        self._check(self._lib.g2ohip_graph_get_pg(
            self._h, _ptr(o["pose_id"], ct.c_int), _ptr(o["poses"], ct.c_double), _ptr(o["fixed"], ct.c_ubyte),
            _ptr(o["edge_from"], ct.c_int), _ptr(o["edge_to"], ct.c_int), _ptr(o["meas"], ct.c_double),
            _ptr(o["info"], ct.c_double)))
        return o

    def to_pose_graph(self, name: str = "") -> PoseGraphProblem:
        a = self.pg_arrays()
        return PoseGraphProblem(a["poses"], a["edge_from"], a["edge_to"], a["meas"], a["info"], a["fixed"],
                                name=name, info_dict=dict(pose_id=a["pose_id"]))

    def set_pg_state(self, poses):
        poses = np.ascontiguousarray(poses, np.float64)
        if poses.shape != (self.info()["poses"], 7):
            raise G2oIOError(f"poses shape {poses.shape} does not match the graph")
        self._check(self._lib.g2ohip_graph_set_pg_state(self._h, _ptr(poses, ct.c_double)))

    def add_pg(self, prob: PoseGraphProblem, pose_id=None):
        poses = np.ascontiguousarray(prob.poses, np.float64)
        ef = np.ascontiguousarray(prob.edge_from, np.int32)
        et = np.ascontiguousarray(prob.edge_to, np.int32)
        meas = np.ascontiguousarray(prob.meas, np.float64)
        info = np.ascontiguousarray(prob.info, np.float64)
        fx = np.ascontiguousarray(prob.fixed, np.uint8)
        pi = None if pose_id is None else np.ascontiguousarray(pose_id, np.int32)
        self._check(self._lib.g2ohip_graph_add_pg(
            self._h, prob.num_poses, prob.num_edges, _ptr(pi, ct.c_int), _ptr(poses, ct.c_double),
            _ptr(fx, ct.c_ubyte), _ptr(ef, ct.c_int), _ptr(et, ct.c_int), _ptr(meas, ct.c_double),
            _ptr(info, ct.c_double)))

    # -------------------------------------------------------------------- BAL
    def append_bal(self, path: str, nthreads: int = 0):
        self._check(self._lib.g2ohip_graph_load_bal(self._h, os.fsencode(path), int(nthreads)))

    def bal_arrays(self) -> "BALProblem":
        n = (ct.c_longlong * 3)()
        self._check(self._lib.g2ohip_graph_get_bal(self._h, n, None, None, None, None, None))
        C, P, E = (int(x) for x in n)
        b = BALProblem(np.empty((C, 9)), np.empty((P, 3)), np.empty(E, np.int32), np.empty(E, np.int32),
                       np.empty((E, 2)))
        self._check(self._lib.g2ohip_graph_get_bal(self._h, None, _ptr(b.cams, ct.c_double), _ptr(b.pts, ct.c_double),
                                                   _ptr(b.edge_cam, ct.c_int), _ptr(b.edge_pt, ct.c_int),
                                                   _ptr(b.meas, ct.c_double)))
        return b


def save_ba_problem(path: str, prob: BAProblem, precision: int = 17, cam_id=None, pt_id=None):
    g = G2oGraph()
    g.add_ba(prob, cam_id, pt_id)
    g.save(path, precision)


def save_pose_graph(path: str, prob: PoseGraphProblem, precision: int = 17, pose_id=None):
    g = G2oGraph()
    g.add_pg(prob, pose_id)
    g.save(path, precision)


# ------------------------------------------------------------------------ BAL
def load_bal(path: str, nthreads: int = 0) -> BALProblem:
    g = G2oGraph()
    g.append_bal(path, nthreads)
    return g.bal_arrays()


def save_bal(path: str, bal: BALProblem):
    """Writes the BAL text layout (header, observations, cameras, points)."""
    with open(path, "w") as f:
        f.write(f"{bal.cams.shape[0]} {bal.pts.shape[0]} {bal.edge_cam.shape[0]}\n")
        for c, p, m in zip(bal.edge_cam, bal.edge_pt, bal.meas):
            f.write(f"{int(c)} {int(p)} {float(m[0])!r} {float(m[1])!r}\n")
        for v in np.concatenate([bal.cams.reshape(-1), bal.pts.reshape(-1)]):
            f.write(f"{float(v)!r}\n")


def angle_axis_to_quat(w: np.ndarray) -> np.ndarray:
    """[N,3] angle-axis -> [N,4] unit quaternions [x y z w] with w >= 0."""
    w = np.asarray(w, np.float64).reshape(-1, 3)
    th = np.linalg.norm(w, axis=1)
    half = 0.5 * th
    # sin(th/2)/th, with its series below 1e-4 (error < 1e-18)
    s = np.where(th > 1e-4, np.sin(half) / np.where(th > 0, th, 1.0), 0.5 - th * th / 48.0)
    q = np.concatenate([w * s[:, None], np.cos(half)[:, None]], axis=1)
    q[q[:, 3] < 0] *= -1.0
    return q


def bal_to_ba(bal: BALProblem, undistort: bool = False, name: str = "bal") -> BAProblem:
    """Maps BAL cameras onto EdgeSE3ProjectXYZ's pinhole model.

    BAL projects u = f r(p) p with p = -P.xy / P.z and r(p) = 1 + k1|p|^2 +
    k2|p|^4 (bal_example.cpp:130-205). With k1 = k2 = 0 this is the pinhole
    u = fx P.x / P.z + cx with fx = fy = -f, cx = cy = 0, and the mapping is
    exact: every edge's error is the BAL error with its sign flipped, so chi2
    is equal. With distortion, ``undistort=True`` removes it from the
    measurements instead (r solved per observation by Newton's method on
    s r(s) = |u| / |f|): the residuals are then measured in undistorted pixels,
    an approximation of the BAL model that the 6-DoF path can represent.
    Otherwise a camera with distortion raises."""
    C = bal.cams.shape[0]
    k = bal.cams[:, 7:9]
    f = bal.cams[:, 6]
    meas = np.array(bal.meas, np.float64)
    if np.any(k != 0.0):
        if not undistort:
            raise G2oIOError("BAL cameras with radial distortion (k1, k2 != 0): pass undistort=True to remove it "
                             "from the measurements")
        kc = k[bal.edge_cam]
        fc = np.abs(f[bal.edge_cam])
        d = np.linalg.norm(meas, axis=1) / np.where(fc > 0, fc, 1.0)
        s = d.copy()
        for _ in range(30):
            s2 = s * s
            g = s * (1 + kc[:, 0] * s2 + kc[:, 1] * s2 * s2) - d
            dg = 1 + 3 * kc[:, 0] * s2 + 5 * kc[:, 1] * s2 * s2
            s = s - g / np.where(dg != 0, dg, 1.0)
        s2 = s * s
        r = 1 + kc[:, 0] * s2 + kc[:, 1] * s2 * s2
        meas = meas / r[:, None]
    cams = np.concatenate([bal.cams[:, 3:6], angle_axis_to_quat(bal.cams[:, 0:3])], axis=1)
    intr = np.stack([-f, -f, np.zeros(C), np.zeros(C)], axis=1)
    E = bal.edge_cam.shape[0]
    return BAProblem(np.ascontiguousarray(cams), intr, np.array(bal.pts, np.float64),
                     np.asarray(bal.edge_cam, np.int32).copy(), np.asarray(bal.edge_pt, np.int32).copy(), meas,
                     np.ones(E), np.zeros(C, np.uint8), np.zeros(bal.pts.shape[0], np.uint8), 0.0, name=name)


def bal_chi2(bal: BALProblem) -> float:
    """chi2 of the BAL model (identity information), the error of
    EdgeObservationBAL (bal_example.cpp:150-205) in numpy."""
    c = bal.cams[bal.edge_cam]
    X = bal.pts[bal.edge_pt]
    w = c[:, 0:3]
    th = np.linalg.norm(w, axis=1)
    safe = np.where(th > 0, th, 1.0)
    v = w / safe[:, None]
    cth, sth = np.cos(th)[:, None], np.sin(th)[:, None]
    rot = X * cth + np.cross(v, X) * sth + v * np.sum(v * X, axis=1, keepdims=True) * (1 - cth)
    rot = np.where(th[:, None] > 0, rot, X + np.cross(w, X))
    P = rot + c[:, 3:6]
    p = -P[:, 0:2] / P[:, 2:3]
    r2 = np.sum(p * p, axis=1)
    r = 1 + c[:, 7] * r2 + c[:, 8] * r2 * r2
    pred = c[:, 6:7] * r[:, None] * p
    e = pred - bal.meas
    return float(np.sum(e * e))
