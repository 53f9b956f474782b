"""Seeded synthetic bundle-adjustment and pose-graph problems.

Problems are plain numpy arrays in the layout the C-ABI takes:

* ``cams``  float64[C, 7]  world-to-camera SE3 as [tx ty tz qx qy qz qw]
  (the ``SE3Quat::toVector`` order, g2o/types/slam3d/se3quat.h);
* ``intr``  float64[C, 4]  fx fy cx cy per camera (EdgeSE3ProjectXYZ keeps
  them per edge, g2o/types/sba/edge_project_xyz.h:53; here they are per camera);
* ``pts``   float64[P, 3];
* ``edge_cam``/``edge_pt`` int32[E], ``meas`` float64[E, 2], ``weight`` float64[E]
  (information matrix ``weight * I2``);
* ``cam_fixed``/``pt_fixed`` uint8 masks, ``huber`` the Huber delta (0 = none).

The generators follow the shapes of g2o's own demos (ba_demo.cpp:128-227 for the
small BA case, create_sphere.cpp:42-241 for the pose graph) but are this
package's own, seeded through ``numpy.random.default_rng``.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

IMG_W, IMG_H = 640.0, 480.0


@dataclass
class BAProblem:
    cams: np.ndarray
    intr: np.ndarray
    pts: np.ndarray
    edge_cam: np.ndarray
    edge_pt: np.ndarray
    meas: np.ndarray
    weight: np.ndarray
    cam_fixed: np.ndarray
    pt_fixed: np.ndarray
    huber: float
    name: str = ""
    gt_cams: np.ndarray | None = None
    gt_pts: np.ndarray | None = None
    info: dict = field(default_factory=dict)
    # per-edge 2x2 information as its upper triangle [O00, O01, O11] (E x 3);
    # None: identity (the information is weight * Omega)
    omega: np.ndarray | None = None

    @property
    def num_cams(self) -> int:
        return int(self.cams.shape[0])

    @property
    def num_pts(self) -> int:
        return int(self.pts.shape[0])

    @property
    def num_edges(self) -> int:
        return int(self.edge_cam.shape[0])

    def copy(self) -> "BAProblem":
        return BAProblem(**{k: (v.copy() if isinstance(v, np.ndarray) else v)
                            for k, v in self.__dict__.items()})


# ----------------------------------------------------------------- rotations
def quat_to_rot(q: np.ndarray) -> np.ndarray:
    """Rotation matrices of unit quaternions [x, y, z, w] (..., 4) -> (..., 3, 3)."""
    x, y, z, w = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    R = np.empty(q.shape[:-1] + (3, 3))
    R[..., 0, 0] = 1 - 2 * (y * y + z * z)
    R[..., 0, 1] = 2 * (x * y - z * w)
    R[..., 0, 2] = 2 * (x * z + y * w)
    R[..., 1, 0] = 2 * (x * y + z * w)
    R[..., 1, 1] = 1 - 2 * (x * x + z * z)
    R[..., 1, 2] = 2 * (y * z - x * w)
    R[..., 2, 0] = 2 * (x * z - y * w)
    R[..., 2, 1] = 2 * (y * z + x * w)
    R[..., 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def rot_to_quat(R: np.ndarray) -> np.ndarray:
    """Unit quaternions [x, y, z, w] with w >= 0 of rotation matrices (..., 3, 3)."""
    R = np.asarray(R, dtype=np.float64)
    shp = R.shape[:-2]
    R = R.reshape(-1, 3, 3)
    q = np.empty((R.shape[0], 4))
    tr = R[:, 0, 0] + R[:, 1, 1] + R[:, 2, 2]
    for n in range(R.shape[0]):
        m = R[n]
        if tr[n] > 0:
            s = np.sqrt(tr[n] + 1.0) * 2
            q[n] = [(m[2, 1] - m[1, 2]) / s, (m[0, 2] - m[2, 0]) / s, (m[1, 0] - m[0, 1]) / s, 0.25 * s]
        else:
            i = int(np.argmax([m[0, 0], m[1, 1], m[2, 2]]))
            j, k = (i + 1) % 3, (i + 2) % 3
            s = np.sqrt(1.0 + m[i, i] - m[j, j] - m[k, k]) * 2
            v = np.empty(3)
            v[i] = 0.25 * s
            v[j] = (m[j, i] + m[i, j]) / s
            v[k] = (m[k, i] + m[i, k]) / s
            q[n] = [v[0], v[1], v[2], (m[k, j] - m[j, k]) / s]
    q *= np.where(q[:, 3:4] < 0, -1.0, 1.0)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return q.reshape(shp + (4,))


def axis_angle_to_rot(w: np.ndarray) -> np.ndarray:
    """Rodrigues' formula, (..., 3) -> (..., 3, 3)."""
    w = np.asarray(w, dtype=np.float64)
    th = np.linalg.norm(w, axis=-1)[..., None, None]
    K = np.zeros(w.shape[:-1] + (3, 3))
    K[..., 0, 1], K[..., 0, 2] = -w[..., 2], w[..., 1]
    K[..., 1, 0], K[..., 1, 2] = w[..., 2], -w[..., 0]
    K[..., 2, 0], K[..., 2, 1] = -w[..., 1], w[..., 0]
    small = th < 1e-12
    ths = np.where(small, 1.0, th)
    a = np.where(small, 1.0, np.sin(ths) / ths)
    b = np.where(small, 0.5, (1 - np.cos(ths)) / (ths * ths))
    return np.eye(3) + a * K + b * (K @ K)


def pose_from_Rt(R: np.ndarray, t: np.ndarray) -> np.ndarray:
    return np.concatenate([t, rot_to_quat(R)], axis=-1)


def pose_to_Rt(p: np.ndarray):
    return quat_to_rot(p[..., 3:7]), p[..., 0:3]


def look_rotation(fwd: np.ndarray, up_hint=np.array([0.0, 0.0, 1.0])) -> np.ndarray:
    """World-to-camera rotations for cameras looking along ``fwd`` (x right,
    y down, z forward)."""
    z = fwd / np.linalg.norm(fwd, axis=-1, keepdims=True)
    x = np.cross(z, up_hint)
    x /= np.linalg.norm(x, axis=-1, keepdims=True)
    y = np.cross(z, x)
    Rwc = np.stack([x, y, z], axis=-1)  # columns are the camera axes in world
    return np.swapaxes(Rwc, -1, -2)


def project(cams: np.ndarray, intr: np.ndarray, pts: np.ndarray, ec: np.ndarray, ep: np.ndarray):
    R, t = pose_to_Rt(cams)
    Xc = np.einsum("eij,ej->ei", R[ec], pts[ep]) + t[ec]
    u = Xc[:, 0] / Xc[:, 2] * intr[ec, 0] + intr[ec, 2]
    v = Xc[:, 1] / Xc[:, 2] * intr[ec, 1] + intr[ec, 3]
    return np.stack([u, v], axis=1), Xc[:, 2]


def perturb_poses(rng, cams: np.ndarray, sig_t: float, sig_r: float) -> np.ndarray:
    """Left-multiplied perturbation exp(delta) * T with delta ~ N(0, diag)."""
    R, t = pose_to_Rt(cams)
    dR = axis_angle_to_rot(rng.normal(0, sig_r, size=(cams.shape[0], 3)))
    dt = rng.normal(0, sig_t, size=(cams.shape[0], 3))
    R2 = dR @ R
    t2 = np.einsum("cij,cj->ci", dR, t) + dt
    return pose_from_Rt(R2, t2)


# ------------------------------------------------------------------- BA cases
def make_small_ba(seed: int = 1000, num_cams: int = 16, num_pts: int = 500, huber: float = 0.0) -> BAProblem:
    """Small BA: cameras on a circle of radius 15 looking inward, points uniform
    in a 10 x 10 x 5 box, fx = fy = 500, (cx, cy) = (320, 240), 640 x 480 image,
    in-image and in-front projections kept, pixel noise N(0, 1); poses perturbed
    by sigma_t = 0.1, sigma_r = 0.02 rad, points by sigma = 0.2; the first two
    cameras fixed (the ba_demo.cpp:154-160 gauge)."""
    rng = np.random.default_rng(seed)
    ang = np.arange(num_cams) * 2 * np.pi / num_cams
    centers = np.stack([15 * np.cos(ang), 15 * np.sin(ang), np.zeros(num_cams)], axis=1)
    R = look_rotation(-centers)
    t = -np.einsum("cij,cj->ci", R, centers)
    gt_cams = pose_from_Rt(R, t)
    intr = np.tile([500.0, 500.0, 320.0, 240.0], (num_cams, 1))
    gt_pts = rng.uniform([-5, -5, -2.5], [5, 5, 2.5], size=(num_pts, 3))
    ec, ep = np.meshgrid(np.arange(num_cams), np.arange(num_pts), indexing="ij")
    ec, ep = ec.ravel().astype(np.int32), ep.ravel().astype(np.int32)
    uv, depth = project(gt_cams, intr, gt_pts, ec, ep)
    keep = (depth > 0) & (uv[:, 0] >= 0) & (uv[:, 1] >= 0) & (uv[:, 0] < IMG_W) & (uv[:, 1] < IMG_H)
    # points seen by fewer than two cameras are dropped (ba_demo.cpp:186)
    cnt = np.bincount(ep[keep], minlength=num_pts)
    keep &= cnt[ep] >= 2
    used = np.flatnonzero(cnt >= 2)
    remap = -np.ones(num_pts, dtype=np.int64)
    remap[used] = np.arange(used.size)
    ec, ep, uv = ec[keep], remap[ep[keep]].astype(np.int32), uv[keep]
    gt_pts = gt_pts[used]
    meas = uv + rng.normal(0, 1.0, size=uv.shape)
    cams = perturb_poses(rng, gt_cams, 0.1, 0.02)
    cam_fixed = np.zeros(num_cams, dtype=np.uint8)
    cam_fixed[:2] = 1
    cams[:2] = gt_cams[:2]
    pts = gt_pts + rng.normal(0, 0.2, size=gt_pts.shape)
    return BAProblem(cams, intr, pts, ec.astype(np.int32), ep, meas, np.ones(ec.size), cam_fixed,
                     np.zeros(gt_pts.shape[0], dtype=np.uint8), huber, name="small_ba", gt_cams=gt_cams,
                     gt_pts=gt_pts)


def make_walk_ba(seed: int = 1002, num_cams: int = 1500, num_pts: int = 500_000, mean_track: float = 6.0,
                 outlier_ratio: float = 0.05, huber: float = 2.0, weights: tuple | None = None,
                 step: float = 2.0, sig_t: float = 0.05, sig_r: float = 0.002, sig_p: float = 0.1,
                 shuffle: bool = True) -> BAProblem:
    """Large BA along a camera random walk (``step`` metres per camera, 3 km
    for 1500 cameras). Per-camera fx ~ U(400, 800) = fy, principal point at the
    centre of a 640 x 480 image. Each point is placed ahead of the last camera
    of a track of consecutive cameras whose length is 2 + Geometric with mean
    ``mean_track``; projections behind a camera or outside the image are
    dropped (points left with < 2 are dropped). Pixel noise N(0, 1);
    ``outlier_ratio`` of the observations are replaced by uniform image points.
    Initial poses are perturbed by (sig_t, sig_r), points by sig_p; the first
    two cameras are fixed (gauge incl. scale). ``weights=(lo, hi)`` draws
    per-edge information weights U(lo, hi), else 1."""
    rng = np.random.default_rng(seed)
    C = num_cams
    heading = np.cumsum(rng.normal(0, 0.02, size=C)) + 0.3 * np.sin(np.arange(C) / 150.0)
    pos = np.zeros((C, 3))
    pos[1:, 0] = np.cumsum(step * np.cos(heading[:-1]))
    pos[1:, 1] = np.cumsum(step * np.sin(heading[:-1]))
    pos[:, 2] = 1.5 + 0.3 * np.sin(np.arange(C) / 40.0)
    yaw = heading + rng.normal(0, 0.03, size=C)
    pitch = rng.normal(0, 0.02, size=C)
    fwd = np.stack([np.cos(yaw) * np.cos(pitch), np.sin(yaw) * np.cos(pitch), np.sin(pitch)], axis=1)
    R = look_rotation(fwd)
    t = -np.einsum("cij,cj->ci", R, pos)
    gt_cams = pose_from_Rt(R, t)
    fx = rng.uniform(400, 800, size=C)
    intr = np.stack([fx, fx, np.full(C, IMG_W / 2), np.full(C, IMG_H / 2)], axis=1)

    P = num_pts
    p_geo = 1.0 / (mean_track - 1.0)
    n = np.minimum(1 + rng.geometric(p_geo, size=P), C)
    last = rng.integers(0, C, size=P)
    last = np.maximum(last, n - 1)
    # point ahead of the last camera: depth U(4, 25), inside 80 % of the view
    depth = rng.uniform(4.0, 25.0, size=P)
    xn = rng.uniform(-0.8, 0.8, size=P) * (IMG_W / 2) / fx[last]
    yn = rng.uniform(-0.8, 0.8, size=P) * (IMG_H / 2) / fx[last]
    Xc = np.stack([xn * depth, yn * depth, depth], axis=1)
    Rl, tl = R[last], t[last]
    gt_pts = np.einsum("pji,pj->pi", Rl, Xc - tl)  # R^T (Xc - t)
    # observations: cameras last-n+1 .. last
    offs = np.repeat(np.cumsum(n) - n, n)
    k = np.arange(n.sum()) - offs
    ep = np.repeat(np.arange(P), n)
    ec = (np.repeat(last - n + 1, n) + k)
    uv, d = project(gt_cams, intr, gt_pts, ec, ep)
    keep = (d > 0.1) & (uv[:, 0] >= 0) & (uv[:, 1] >= 0) & (uv[:, 0] < IMG_W) & (uv[:, 1] < IMG_H)
    cnt = np.bincount(ep[keep], minlength=P)
    keep &= cnt[ep] >= 2
    used = np.flatnonzero(cnt >= 2)
    remap = -np.ones(P, dtype=np.int64)
    if shuffle:
        perm = rng.permutation(used.size)
    else:
        perm = np.arange(used.size)
    remap[used] = perm
    gt_pts_new = np.empty((used.size, 3))
    gt_pts_new[perm] = gt_pts[used]
    ec, ep, uv = ec[keep], remap[ep[keep]], uv[keep]
    E = ec.size
    meas = uv + rng.normal(0, 1.0, size=uv.shape)
    out = rng.random(E) < outlier_ratio
    meas[out] = np.stack([rng.uniform(0, IMG_W, out.sum()), rng.uniform(0, IMG_H, out.sum())], axis=1)
    # observations ordered camera-major (as BAL files list them)
    order = np.lexsort((ep, ec))
    ec, ep, meas = ec[order], ep[order], meas[order]
    w = np.ones(E) if weights is None else rng.uniform(weights[0], weights[1], size=E)
    cams = perturb_poses(rng, gt_cams, sig_t, sig_r)
    cam_fixed = np.zeros(C, dtype=np.uint8)
    cam_fixed[:2] = 1
    cams[:2] = gt_cams[:2]
    pts = gt_pts_new + rng.normal(0, sig_p, size=gt_pts_new.shape)
    info = {"outliers": int(out.sum()), "mean_track": float(E / used.size)}
    return BAProblem(cams, intr, pts, ec.astype(np.int32), ep.astype(np.int32), meas, w, cam_fixed,
                     np.zeros(used.size, dtype=np.uint8), huber, name=f"walk_ba_{C}c_{used.size}p",
                     gt_cams=gt_cams, gt_pts=gt_pts_new, info=info)


def truncated_zipf(rng, a: float, lo: int, hi: int, size: int) -> np.ndarray:
    """Samples of P(n) proportional to n^-a on the integers lo..hi (the Zipf
    law restricted to that range, by inverting its cumulative table)."""
    n = np.arange(lo, hi + 1)
    cdf = np.cumsum(n.astype(np.float64) ** -a)
    cdf /= cdf[-1]
    return n[np.searchsorted(cdf, rng.random(size), side="right").clip(0, n.size - 1)]


def make_zipf_ba(seed: int = 1003, num_cams: int = 4000, num_pts: int = 1_200_000, zipf: float = 1.8,
                 max_track: int = 300, min_track: int = 2, step: float = 0.1, near_ratio: float = 0.02,
                 outlier_ratio: float = 0.0, huber: float = 2.0, sig_t: float = 0.05, sig_r: float = 0.002,
                 sig_p: float = 0.1, shuffle: bool = True) -> BAProblem:
    """Uneven-structure BA: cameras ``step`` metres apart along a gently
    curving path (0.4 km for 4000), each looking to the left of its heading
    (fx ~ U(400, 800) = fy, principal point at the centre of 640 x 480). Track
    lengths are min_track - 1 + k with k following the Zipf law k^-``zipf`` on
    1..max_track - min_track + 1 (numpy's zipf(1.8) + 1 kept within 2..300:
    mean ~6.3, so ~7.5M observations for 1.2M points);
    each track is a run of consecutive cameras, so the co-visibility of the
    reduced camera system is banded. A point is placed beyond the middle of
    its run at a lateral distance that keeps it inside 80 % of every run
    camera's view: at least the run's half-length over the narrowest tangent,
    else log-uniform on 2..60 m; ``near_ratio`` of the points sit 0.3..1 m from
    their cameras instead (short tracks there), so Hll conditioning spans
    many decades. Projections behind a camera or outside the image are
    dropped (points left with < 2 are dropped). Pixel noise N(0, 1);
    ``outlier_ratio`` of the observations replaced by uniform image points.
    Initial poses perturbed by (sig_t, sig_r), points by sig_p; the first two
    cameras fixed."""
    rng = np.random.default_rng(seed)
    C = num_cams
    heading = np.cumsum(rng.normal(0, 0.002, size=C)) + 0.2 * np.sin(np.arange(C) / 700.0)
    pos = np.zeros((C, 3))
    pos[1:, 0] = np.cumsum(step * np.cos(heading[:-1]))
    pos[1:, 1] = np.cumsum(step * np.sin(heading[:-1]))
    pos[:, 2] = 1.5 + 0.05 * np.sin(np.arange(C) / 90.0)
    yaw = heading + np.pi / 2 + rng.normal(0, 0.01, size=C)
    pitch = rng.normal(0, 0.01, size=C)
    fwd = np.stack([np.cos(yaw) * np.cos(pitch), np.sin(yaw) * np.cos(pitch), np.sin(pitch)], axis=1)
    R = look_rotation(fwd)
    t = -np.einsum("cij,cj->ci", R, pos)
    gt_cams = pose_from_Rt(R, t)
    fx = rng.uniform(400, 800, size=C)
    intr = np.stack([fx, fx, np.full(C, IMG_W / 2), np.full(C, IMG_H / 2)], axis=1)

    P = num_pts
    near = rng.random(P) < near_ratio
    n = min_track - 1 + truncated_zipf(rng, zipf, 1, max_track - min_track + 1, P)
    n = np.where(near, np.minimum(n, 4), n)
    n = np.minimum(n, C)
    first = (rng.random(P) * (C - n + 1)).astype(np.int64)
    mid = first + (n - 1) // 2
    # narrowest half-view tangent (80 % of the half width) over the run
    tan_min = 0.8 * (IMG_W / 2) / _run_max(fx, first, n)
    half = 0.5 * step * (n - 1)
    lateral = np.maximum(1.2 * half / tan_min, np.exp(rng.uniform(np.log(2.0), np.log(60.0), size=P)))
    lateral = np.where(near, rng.uniform(0.3, 1.0, size=P), lateral)
    height = rng.uniform(-0.5, 0.5, size=P) * lateral * (IMG_H / IMG_W)
    gt_pts = pos[mid] + fwd[mid] * lateral[:, None] + np.array([0.0, 0.0, 1.0]) * height[:, None]
    offs = np.repeat(np.cumsum(n) - n, n)
    k = np.arange(n.sum()) - offs
    ep = np.repeat(np.arange(P), n)
    ec = np.repeat(first, n) + k
    uv, d = project(gt_cams, intr, gt_pts, ec, ep)
    keep = (d > 0.1) & (uv[:, 0] >= 0) & (uv[:, 1] >= 0) & (uv[:, 0] < IMG_W) & (uv[:, 1] < IMG_H)
    cnt = np.bincount(ep[keep], minlength=P)
    keep &= cnt[ep] >= 2
    used = np.flatnonzero(cnt >= 2)
    remap = -np.ones(P, dtype=np.int64)
    perm = rng.permutation(used.size) if shuffle else np.arange(used.size)
    remap[used] = perm
    gt_pts_new = np.empty((used.size, 3))
    gt_pts_new[perm] = gt_pts[used]
    ec, ep, uv = ec[keep], remap[ep[keep]], uv[keep]
    E = ec.size
    meas = uv + rng.normal(0, 1.0, size=uv.shape)
    out = rng.random(E) < outlier_ratio
    meas[out] = np.stack([rng.uniform(0, IMG_W, out.sum()), rng.uniform(0, IMG_H, out.sum())], axis=1)
    order = np.lexsort((ep, ec))
    ec, ep, meas = ec[order], ep[order], meas[order]
    cams = perturb_poses(rng, gt_cams, sig_t, sig_r)
    cam_fixed = np.zeros(C, dtype=np.uint8)
    cam_fixed[:2] = 1
    cams[:2] = gt_cams[:2]
    pts = gt_pts_new + rng.normal(0, sig_p, size=gt_pts_new.shape)
    tracks = np.bincount(ep, minlength=used.size)
    info = {"outliers": int(out.sum()), "mean_track": float(E / max(used.size, 1)),
            "max_track": int(tracks.max()) if tracks.size else 0}
    return BAProblem(cams, intr, pts, ec.astype(np.int32), ep.astype(np.int32), meas, np.ones(E), cam_fixed,
                     np.zeros(used.size, dtype=np.uint8), huber, name=f"zipf_ba_{C}c_{used.size}p",
                     gt_cams=gt_cams, gt_pts=gt_pts_new, info=info)


def _run_max(v: np.ndarray, first: np.ndarray, n: np.ndarray) -> np.ndarray:
    """max(v[first[i] : first[i] + n[i]]) for every i (sparse table)."""
    levels = [v]
    span = 1
    while 2 * span <= int(n.max()):
        prev = levels[-1]
        nxt = prev.copy()
        nxt[:-span] = np.maximum(prev[:-span], prev[span:])
        levels.append(nxt)
        span *= 2
    lvl = np.floor(np.log2(n)).astype(np.int64)
    out = np.empty(first.size)
    for L in range(len(levels)):
        m = lvl == L
        if m.any():
            w = 1 << L
            out[m] = np.maximum(levels[L][first[m]], levels[L][first[m] + n[m] - w])
    return out


def make_metric_ba(seed: int = 1002) -> BAProblem:
    """The headline workload: 1,500 cameras on a 3 km walk, 500k points,
    ~3M observations, Huber delta 2, 5 % outliers."""
    return make_walk_ba(seed=seed, num_cams=1500, num_pts=500_000)


def make_diff_ba(seed: int = 1004) -> BAProblem:
    """Differentiable-BA workload: 200 cameras, 50k points, ~300k observations,
    per-edge weights U(0.5, 2)."""
    return make_walk_ba(seed=seed, num_cams=200, num_pts=50_000, weights=(0.5, 2.0))


def select_edges(prob: BAProblem, keep: np.ndarray, min_obs: int = 2) -> BAProblem:
    """Sub-problem with the edges where ``keep`` is true, as after g2o's
    outlier rejection (edges set to level 1 and left out of
    initializeOptimization(0)); points left with fewer than ``min_obs`` edges
    are dropped with their edges. info['edge_index'] / info['pt_index'] map
    back to the original edges / points."""
    keep = np.asarray(keep, dtype=bool).copy()
    while True:
        cnt = np.bincount(prob.edge_pt[keep], minlength=prob.num_pts)
        drop = keep & (cnt[prob.edge_pt] < min_obs)
        if not drop.any():
            break
        keep &= ~drop
    used = np.flatnonzero(np.bincount(prob.edge_pt[keep], minlength=prob.num_pts) > 0)
    remap = -np.ones(prob.num_pts, dtype=np.int64)
    remap[used] = np.arange(used.size)
    eidx = np.flatnonzero(keep)
    info = dict(prob.info)
    info.update(edge_index=eidx, pt_index=used)
    return BAProblem(prob.cams.copy(), prob.intr.copy(), prob.pts[used].copy(), prob.edge_cam[eidx].copy(),
                     remap[prob.edge_pt[eidx]].astype(np.int32), prob.meas[eidx].copy(), prob.weight[eidx].copy(),
                     prob.cam_fixed.copy(), prob.pt_fixed[used].copy(), prob.huber, name=prob.name + "_sel",
                     gt_cams=prob.gt_cams, gt_pts=None if prob.gt_pts is None else prob.gt_pts[used], info=info)


# ------------------------------------------------------------ 3D pose graph
@dataclass
class PoseGraphProblem:
    """EdgeSE3 / VertexSE3 pose graph. poses [N,7] = Isometry3 [tx ty tz qx qy
    qz qw] (x_world = R x + t); meas [E,7] relative pose from -> to; info
    [E,6,6] information matrices; fixed [N] 0/1."""
    poses: np.ndarray
    edge_from: np.ndarray
    edge_to: np.ndarray
    meas: np.ndarray
    info: np.ndarray
    fixed: np.ndarray
    name: str = ""
    gt_poses: np.ndarray | None = None
    info_dict: dict = field(default_factory=dict)

    @property
    def num_poses(self) -> int:
        return int(self.poses.shape[0])

    @property
    def num_edges(self) -> int:
        return int(self.edge_from.shape[0])


def _quat_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    ax, ay, az, aw = np.moveaxis(a, -1, 0)
    bx, by, bz, bw = np.moveaxis(b, -1, 0)
    return np.stack([aw * bx + ax * bw + ay * bz - az * by, aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw, aw * bw - ax * bx - ay * by - az * bz], axis=-1)


def _relative(pa: np.ndarray, pb: np.ndarray) -> np.ndarray:
    """pa^-1 pb for Isometry3 poses [.., 7]."""
    Ra = quat_to_rot(pa[..., 3:])
    t = np.einsum("...ji,...j->...i", Ra, pb[..., :3] - pa[..., :3])
    qa_inv = pa[..., 3:] * np.array([-1, -1, -1, 1.0])
    q = _quat_mul(qa_inv, pb[..., 3:])
    q *= np.where(q[..., 3:] < 0, -1.0, 1.0)
    return np.concatenate([t, q / np.linalg.norm(q, axis=-1, keepdims=True)], axis=-1)


def _compose(pa: np.ndarray, pb: np.ndarray) -> np.ndarray:
    """pa pb for Isometry3 poses [.., 7]."""
    Ra = quat_to_rot(pa[..., 3:])
    t = pa[..., :3] + np.einsum("...ij,...j->...i", Ra, pb[..., :3])
    q = _quat_mul(pa[..., 3:], pb[..., 3:])
    return np.concatenate([t, q / np.linalg.norm(q, axis=-1, keepdims=True)], axis=-1)


def make_sphere_pose_graph(seed: int = 1001, nodes_per_level: int = 1000, laps: int = 10, radius: float = 5.0,
                           sigma_t: float = 0.01, sigma_r: float = 0.005) -> PoseGraphProblem:
    """Sphere trajectory as g2o's create_sphere example builds it
    (examples/sphere/create_sphere.cpp): ``laps`` laps of ``nodes_per_level``
    poses, azimuth -pi + 2 pi n / nodes_per_level, elevation rising from
    -pi/2 to pi/2 over the whole run, position R (radius, 0, 0). Odometry
    edges (i-1, i) and, from each pose of lap f-1, edges to the poses n-1,
    n, n+1 of lap f (the adjacent laps are 2 r sin(pi / (2 laps)) apart: 1.56
    m at r = 5). Measurements are the true relative poses with noise: the
    translation plus N(0, sigma_t^2), the rotation right-multiplied by the
    normalised quaternion (1 - |v|, v), v ~ N(0, sigma_r^2). Information =
    diag(1/sigma_t^2 x3, 1/sigma_r^2 x3). The initial estimate chains the
    noisy odometry from the first pose, which is fixed."""
    rng = np.random.default_rng(seed)
    N = nodes_per_level * laps
    ids = np.arange(N)
    n = ids % nodes_per_level
    az = -np.pi + 2 * n * np.pi / nodes_per_level
    el = -0.5 * np.pi + (ids + 1) * np.pi / N
    # rot = Rz(az) Ry(el)
    cz, sz, cy, sy = np.cos(az), np.sin(az), np.cos(el), np.sin(el)
    R = np.zeros((N, 3, 3))
    R[:, 0, 0] = cz * cy; R[:, 0, 1] = -sz; R[:, 0, 2] = cz * sy
    R[:, 1, 0] = sz * cy; R[:, 1, 1] = cz; R[:, 1, 2] = sz * sy
    R[:, 2, 0] = -sy; R[:, 2, 1] = 0; R[:, 2, 2] = cy
    t = R[:, :, 0] * radius
    gt = np.concatenate([t, rot_to_quat(R)], axis=1)
    gt[:, 3:] *= np.where(gt[:, 6:7] < 0, -1.0, 1.0)
    fr, to = [np.arange(N - 1)], [np.arange(1, N)]
    for f in range(1, laps):
        for d in (-1, 0, 1):
            nn = np.arange(nodes_per_level)
            if f == laps - 1 and d == 1:
                continue
            src = (f - 1) * nodes_per_level + nn
            dst = f * nodes_per_level + nn + d
            ok = (dst >= 0) & (dst < N)
            fr.append(src[ok])
            to.append(dst[ok])
    fr = np.concatenate(fr).astype(np.int32)
    to = np.concatenate(to).astype(np.int32)
    E = fr.size
    meas = _relative(gt[fr], gt[to])
    v = rng.normal(0, sigma_r, size=(E, 3))
    qw = np.maximum(1.0 - np.linalg.norm(v, axis=1), 0.0)
    dq = np.concatenate([v, qw[:, None]], axis=1)
    dq /= np.linalg.norm(dq, axis=1, keepdims=True)
    meas[:, 3:] = _quat_mul(meas[:, 3:], dq)
    meas[:, :3] += rng.normal(0, sigma_t, size=(E, 3))
    info = np.zeros((E, 6, 6))
    info[:, [0, 1, 2], [0, 1, 2]] = 1.0 / sigma_t ** 2
    info[:, [3, 4, 5], [3, 4, 5]] = 1.0 / sigma_r ** 2
    poses = np.empty_like(gt)
    poses[0] = gt[0]
    for i in range(1, N):  # odometry edges are the first N - 1
        poses[i] = _compose(poses[i - 1], meas[i - 1])
    fixed = np.zeros(N, dtype=np.uint8)
    fixed[0] = 1
    return PoseGraphProblem(poses, fr, to, meas, info, fixed, name=f"sphere_{laps}x{nodes_per_level}",
                            gt_poses=gt, info_dict={"radius": radius})


# ------------------------------------------------------------------------ BAL
@dataclass
class BALProblem:
    """BAL data (g2o/examples/bal/bal_example.cpp:70-207): cams [C, 9] =
    angle-axis rotation, translation, f, k1, k2 (world to camera, P = R X + t,
    the camera looking down its -z axis); pts [P, 3]; observations in file
    order, meas [E, 2] pixels relative to the image centre. Optional, for the
    solver (the BAL file format has none of them): weight [E] (information
    weight * I2, None = 1), cam_fixed / pt_fixed 0/1 masks (None = all free,
    as bal_example.cpp leaves them), huber (RobustKernelHuber delta, 0 = none).
    gt_cams / gt_pts: the generator's ground truth, if any."""
    cams: np.ndarray
    pts: np.ndarray
    edge_cam: np.ndarray
    edge_pt: np.ndarray
    meas: np.ndarray
    weight: np.ndarray | None = None
    cam_fixed: np.ndarray | None = None
    pt_fixed: np.ndarray | None = None
    huber: float = 0.0
    name: str = ""
    gt_cams: np.ndarray | None = None
    gt_pts: np.ndarray | None = None

    @property
    def num_cams(self) -> int:
        return int(self.cams.shape[0])

    @property
    def num_pts(self) -> int:
        return int(self.pts.shape[0])

    @property
    def num_edges(self) -> int:
        return int(self.edge_cam.shape[0])

    def copy(self) -> "BALProblem":
        return BALProblem(**{k: (v.copy() if isinstance(v, np.ndarray) else v)
                             for k, v in self.__dict__.items()})


def make_bal(seed: int = 11, num_cams: int = 8, num_pts: int = 300, k=(0.0, 0.0), noise: float = 0.5) -> BALProblem:
    """Seeded BAL problem: cameras near the origin looking down their -z axis
    at a point cloud 12 units away, focal lengths 500 + 10 c, radial
    distortion k, Gaussian pixel noise; each point seen by each camera with
    probability 0.7."""
    rng = np.random.default_rng(seed)
    pts = rng.uniform(-3, 3, size=(num_pts, 3))
    cams, obs = [], []
    for c in range(num_cams):
        w = rng.normal(size=3) * 0.3
        R = axis_angle_to_rot(w)
        t = np.array([0.0, 0.0, -12.0]) + rng.normal(size=3) * 0.2
        f = 500.0 + 10 * c
        cams.append(np.r_[w, t, f, k[0], k[1]])
        P = pts @ R.T + t
        p = -P[:, :2] / P[:, 2:3]
        r2 = np.sum(p * p, axis=1)
        r = 1 + k[0] * r2 + k[1] * r2 * r2
        u = f * r[:, None] * p + rng.normal(size=(num_pts, 2)) * noise
        seen = rng.random(num_pts) < 0.7
        for j in np.flatnonzero(seen):
            obs.append((c, j, u[j, 0], u[j, 1]))
    obs = np.array(obs)
    return BALProblem(np.array(cams), pts, obs[:, 0].astype(np.int32), obs[:, 1].astype(np.int32),
                      obs[:, 2:4].copy())


def rot_to_angle_axis(R: np.ndarray) -> np.ndarray:
    """Angle-axis vectors (..., 3) of rotation matrices (..., 3, 3), through
    the unit quaternion (robust up to angles of pi)."""
    q = rot_to_quat(R).reshape(-1, 4)
    s = np.linalg.norm(q[:, :3], axis=1)
    ang = 2.0 * np.arctan2(s, q[:, 3])
    scale = np.where(s > 1e-12, ang / np.where(s > 1e-12, s, 1.0), 2.0)
    return (q[:, :3] * scale[:, None]).reshape(np.shape(R)[:-2] + (3,))


def bal_project(cams9: np.ndarray, pts: np.ndarray, ec: np.ndarray, ep: np.ndarray) -> np.ndarray:
    """BAL projection f r(p) p, p = -P.xy / P.z, P = R(w) X + t, per observation."""
    c = cams9[ec]
    P = np.einsum("eij,ej->ei", axis_angle_to_rot(c[:, 0:3]), pts[ep]) + c[:, 3:6]
    p = -P[:, 0:2] / P[:, 2:3]
    r2 = np.sum(p * p, axis=1)
    r = 1 + c[:, 7] * r2 + c[:, 8] * r2 * r2
    return c[:, 6:7] * r[:, None] * p


def make_walk_bal(seed: int = 1005, num_cams: int = 1500, num_pts: int = 500_000, mean_track: float = 6.0,
                  k_sigma=(0.05, 0.01), noise: float = 1.0, sig_f: float = 0.005) -> BALProblem:
    """BAL-model counterpart of make_walk_ba: the same random-walk cameras,
    tracks and initial perturbations (no outliers, no robust kernel, nothing
    fixed, as bal_example.cpp runs), each camera turned to look down its -z
    axis (a half turn about x) and given a focal length f = fx and radial
    distortion k1 ~ N(0, k_sigma[0]), k2 ~ N(0, k_sigma[1]). Measurements are
    the distorted projections of the ground truth plus N(0, noise) pixels,
    relative to the image centre. The initial cameras start from f (1 +
    N(0, sig_f)) and k1 = k2 = 0."""
    ba = make_walk_ba(seed=seed, num_cams=num_cams, num_pts=num_pts, mean_track=mean_track, outlier_ratio=0.0,
                      huber=0.0)
    rng = np.random.default_rng(seed + 7)
    C = ba.num_cams
    flip = np.diag([1.0, -1.0, -1.0])

    def to_bal(cams7, f, k):
        R = flip @ quat_to_rot(cams7[:, 3:7])
        t = cams7[:, 0:3] @ flip.T
        return np.concatenate([rot_to_angle_axis(R), t, f[:, None], k], axis=1)

    f = ba.intr[:, 0].copy()
    k = np.stack([rng.normal(0, k_sigma[0], C), rng.normal(0, k_sigma[1], C)], axis=1)
    gt = to_bal(ba.gt_cams, f, k)
    meas = bal_project(gt, ba.gt_pts, ba.edge_cam, ba.edge_pt) + rng.normal(0, noise, size=(ba.num_edges, 2))
    init = to_bal(ba.cams, f * (1 + rng.normal(0, sig_f, C)), np.zeros((C, 2)))
    return BALProblem(init, ba.pts.copy(), ba.edge_cam.copy(), ba.edge_pt.copy(), meas,
                      name=f"walk_bal_{C}c_{ba.num_pts}p", gt_cams=gt, gt_pts=ba.gt_pts.copy())


def random_information(num_edges: int, seed: int, lo: float = 0.3, hi: float = 3.0) -> np.ndarray:
    """Per-edge positive-definite 2x2 information matrices as their upper
    triangles [O00, O01, O11] (num_edges x 3): eigenvalues U(lo, hi) and a
    uniformly random orientation (anisotropic pixel noise)."""
    rng = np.random.default_rng(seed)
    th = rng.uniform(0, np.pi, num_edges)
    l1, l2 = rng.uniform(lo, hi, num_edges), rng.uniform(lo, hi, num_edges)
    c, s = np.cos(th), np.sin(th)
    return np.stack([l1 * c * c + l2 * s * s, (l1 - l2) * c * s, l1 * s * s + l2 * c * c], 1)
