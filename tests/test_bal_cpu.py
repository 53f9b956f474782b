"""Pinning the oracle's BAL path (no GPU): g2o's BAL example
(g2o/examples/bal/bal_example.cpp:70-207), VertexCameraBAL with 9 values
[w, t, f, k1, k2] and additive updates, EdgeObservationBAL, BlockSolver<9, 3>.

Each check is independent of the formula it checks:
* the rotation against scipy's rotation-vector map, and the first-order
  branch at w = 0 (bal_example.cpp:163-180);
* without distortion, the chi2 of the BAL model equals the 6-DoF oracle's
  chi2 of the same cameras mapped onto EdgeSE3ProjectXYZ (fx = fy = -f),
  whose projection is pinned separately;
* the dual-number Jacobians (the reference differentiates automatically,
  auto_differentiation.h:155-302) against central differences of the error
  through the additive update, also next to w = 0;
* the whole quadratic form against a dense numeric Jacobian, with Huber
  weights and fixed vertices;
* Schur elimination + back-substitution against a dense solve of the damped
  system, PCG against Cholesky;
* LM reaches the statistical chi2 of a sigma = 1 problem.
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from g2o_amd import datasets, graphio


@pytest.fixture(scope="module")
def O(oracle_lib):
    return oracle_lib


def _bal_error_scipy(cam, X, z):
    P = Rotation.from_rotvec(cam[0:3]).apply(X) + cam[3:6]
    p = -P[0:2] / P[2]
    r2 = p @ p
    return cam[6] * (1 + cam[7] * r2 + cam[8] * r2 * r2) * p - z


def test_bal_error_matches_scipy_rotation(O):
    rng = np.random.default_rng(3)
    for _ in range(200):
        cam = np.r_[rng.normal(size=3) * rng.choice([1e-3, 0.3, 2.0]), rng.normal(size=2), -8 + rng.normal(),
                    rng.uniform(300, 900), rng.normal() * 0.1, rng.normal() * 0.02]
        X = rng.normal(size=3)
        z = rng.normal(size=2) * 50
        e, _, _ = O.bal_edge(cam, X, z, jacobians=False)
        assert np.allclose(e, _bal_error_scipy(cam, X, z), rtol=1e-12, atol=1e-9)


def test_bal_first_order_branch_at_zero_rotation(O):
    # w = 0 takes the reference's first-order branch p = X + w x X = X; the
    # Jacobian in w there is -[X]_x propagated, continuous with w != 0
    cam = np.r_[0.0, 0.0, 0.0, 0.3, -0.2, -9.0, 600.0, 0.05, -0.01]
    X = np.array([0.4, -0.7, 1.1])
    z = np.array([10.0, -20.0])
    e0, Jc0, Jp0 = O.bal_edge(cam, X, z)
    assert np.allclose(e0, _bal_error_scipy(cam, X, z), rtol=1e-14)
    cam2 = cam.copy()
    cam2[0:3] = [1e-7, -2e-7, 1.5e-7]
    _, Jc1, Jp1 = O.bal_edge(cam2, X, z)
    # the two states differ by |w| ~ 3e-7, so the Jacobians by ~ |J| 3e-7
    assert np.allclose(Jc0, Jc1, rtol=1e-5, atol=1e-4)
    assert np.allclose(Jp0, Jp1, rtol=1e-5, atol=1e-4)


def test_bal_chi2_equals_pinhole_oracle_without_distortion(O):
    bal = datasets.make_bal(seed=5, num_cams=6, num_pts=150, k=(0.0, 0.0))
    ba = graphio.bal_to_ba(bal)
    chi_bal = O.OracleBAL(bal).chi2()
    chi_ba = O.OracleBA(ba).chi2()
    assert chi_bal == pytest.approx(chi_ba, rel=1e-11)


def test_bal_jacobians_match_central_differences(O):
    rng = np.random.default_rng(7)
    worst = 0.0
    for trial in range(300):
        scale = [1e-6, 1e-2, 0.5, 2.5][trial % 4]
        cam = np.r_[rng.normal(size=3) * scale, rng.normal(size=2), -10 + rng.normal(),
                    rng.uniform(300, 900), rng.normal() * 0.1, rng.normal() * 0.03]
        X = rng.normal(size=3) * 2
        z = rng.normal(size=2) * 30
        _, Jc, Jp = O.bal_edge(cam, X, z)
        h = 1e-6
        num = np.zeros((2, 12))
        for j in range(12):
            for sgn in (1.0, -1.0):
                c, x = cam.copy(), X.copy()
                if j < 9:
                    c[j] += sgn * h
                else:
                    x[j - 9] += sgn * h
                e, _, _ = O.bal_edge(c, x, z, jacobians=False)
                num[:, j] += sgn * e / (2 * h)
        ana = np.concatenate([Jc, Jp], axis=1)
        err = np.abs(ana - num).max() / max(1.0, np.abs(ana).max())
        worst = max(worst, err)
    assert worst < 1e-6, worst


def _small_problem(huber=0.0, fixed=True):
    bal = datasets.make_bal(seed=21, num_cams=6, num_pts=80, k=(0.08, -0.02), noise=0.8)
    rng = np.random.default_rng(2)
    bal.cams = bal.cams + np.c_[rng.normal(0, 0.01, (6, 3)), rng.normal(0, 0.05, (6, 3)),
                                rng.normal(0, 2.0, 6), rng.normal(0, 0.01, 6), rng.normal(0, 0.002, 6)]
    bal.pts = bal.pts + rng.normal(0, 0.05, bal.pts.shape)
    bal.weight = rng.uniform(0.5, 2.0, bal.num_edges)
    bal.huber = huber
    if fixed:
        bal.cam_fixed = np.zeros(bal.num_cams, np.uint8)
        bal.cam_fixed[0] = 1
        bal.pt_fixed = np.zeros(bal.num_pts, np.uint8)
        bal.pt_fixed[[3, 11]] = 1
    return bal


@pytest.mark.parametrize("huber", [0.0, 1.5])
def test_bal_quadratic_form_vs_dense_numeric_jacobian(O, huber):
    bal = _small_problem(huber)
    ora = O.OracleBAL(bal)
    chi, Hpp, Hll, Hpl, b = ora.linearize()
    cf = bal.cam_fixed.astype(bool)
    pf = bal.pt_fixed.astype(bool)
    camH = -np.ones(bal.num_cams, int)
    camH[~cf] = np.arange((~cf).sum())
    ptH = -np.ones(bal.num_pts, int)
    ptH[~pf] = np.arange((~pf).sum())
    Cf, Pf = int((~cf).sum()), int((~pf).sum())
    n = 9 * Cf + 3 * Pf
    E = bal.num_edges
    J = np.zeros((2 * E, n))
    e0 = np.zeros(2 * E)
    h = 1e-6
    for k in range(E):
        c, p = bal.edge_cam[k], bal.edge_pt[k]
        e, _, _ = O.bal_edge(bal.cams[c], bal.pts[p], bal.meas[k], jacobians=False)
        e0[2 * k:2 * k + 2] = e
        cols = []
        if camH[c] >= 0:
            cols += [("c", j, 9 * camH[c] + j) for j in range(9)]
        if ptH[p] >= 0:
            cols += [("p", j, 9 * Cf + 3 * ptH[p] + j) for j in range(3)]
        for kind, j, col in cols:
            d = np.zeros(2)
            for sgn in (1.0, -1.0):
                cc, xx = bal.cams[c].copy(), bal.pts[p].copy()
                if kind == "c":
                    cc[j] += sgn * h
                else:
                    xx[j] += sgn * h
                ee, _, _ = O.bal_edge(cc, xx, bal.meas[k], jacobians=False)
                d += sgn * ee / (2 * h)
            J[2 * k:2 * k + 2, col] = d
    s = bal.weight * (e0.reshape(-1, 2) ** 2).sum(1)
    rho1 = np.ones(E) if huber == 0 else np.where(s <= huber ** 2, 1.0, huber / np.sqrt(s))
    W = np.repeat(bal.weight * rho1, 2)
    H = J.T @ (W[:, None] * J)
    g = -J.T @ (W * e0)
    scale = np.abs(H).max()
    for c in range(Cf):
        assert np.allclose(Hpp[c], H[9 * c:9 * c + 9, 9 * c:9 * c + 9], atol=1e-6 * scale)
    for l in range(Pf):
        o = 9 * Cf + 3 * l
        assert np.allclose(Hll[l], H[o:o + 3, o:o + 3], atol=1e-6 * scale)
    for k in range(E):
        ci, li = camH[bal.edge_cam[k]], ptH[bal.edge_pt[k]]
        if ci < 0 or li < 0:
            assert not Hpl[k].any()
            continue
        o = 9 * Cf + 3 * li
        # one edge per (camera, point) in this problem: the block is the edge's
        assert np.allclose(Hpl[k], H[9 * ci:9 * ci + 9, o:o + 3], atol=1e-6 * scale)
    assert np.allclose(b, g, atol=1e-6 * np.abs(g).max())


@pytest.mark.parametrize("lam", [1e-2, 1.0, 100.0])
def test_bal_schur_solution_equals_dense_solve(O, lam):
    bal = _small_problem(1.5)
    ora = O.OracleBAL(bal)
    _, Hpp, Hll, Hpl, b = ora.linearize()
    Cf, Pf = Hpp.shape[0], Hll.shape[0]
    n = 9 * Cf + 3 * Pf
    H = np.zeros((n, n))
    for c in range(Cf):
        H[9 * c:9 * c + 9, 9 * c:9 * c + 9] = Hpp[c]
    for l in range(Pf):
        o = 9 * Cf + 3 * l
        H[o:o + 3, o:o + 3] = Hll[l]
    cf = bal.cam_fixed.astype(bool)
    pf = bal.pt_fixed.astype(bool)
    camH = -np.ones(bal.num_cams, int)
    camH[~cf] = np.arange(Cf)
    ptH = -np.ones(bal.num_pts, int)
    ptH[~pf] = np.arange(Pf)
    for k in range(bal.num_edges):
        ci, li = camH[bal.edge_cam[k]], ptH[bal.edge_pt[k]]
        if ci < 0 or li < 0:
            continue
        o = 9 * Cf + 3 * li
        H[9 * ci:9 * ci + 9, o:o + 3] = Hpl[k]
        H[o:o + 3, 9 * ci:9 * ci + 9] = Hpl[k].T
    x_ref = np.linalg.solve(H + lam * np.eye(n), b)
    st, x, _ = ora.solve(lam, solver=0)
    assert st == 0
    assert np.linalg.norm(x - x_ref) <= 1e-9 * np.linalg.norm(x_ref)
    st, xp, it = ora.solve(lam, solver=1, tol=1e-14)
    assert st == 0 and it > 0
    assert np.linalg.norm(xp - x_ref) <= 1e-6 * np.linalg.norm(x_ref)


def test_bal_lm_reaches_statistical_chi2(O):
    bal = datasets.make_walk_bal(seed=77, num_cams=30, num_pts=1500)
    ora = O.OracleBAL(bal)
    n, tr, _ = ora.optimize(20, solver=0)
    ndof = 2 * bal.num_edges - (9 * bal.num_cams + 3 * bal.num_pts)
    assert n == 20 or tr[-1, 5] == 1
    assert np.all(np.diff(tr[:, 1]) <= 0)
    assert tr[-1, 1] < 1.1 * ndof, (tr[-1, 1], ndof)
    # the PCG-solved LM follows the Cholesky-solved one
    ora2 = O.OracleBAL(bal)
    n2, tr2, _ = ora2.optimize(20, solver=1, tol=1e-12)
    assert abs(tr2[-1, 1] - tr[-1, 1]) <= 1e-6 * tr[-1, 1]


def test_bal_structure_on_the_host(O):
    """The C ABI's BAL structure (host-only context): free vertices, the
    Hschur pattern of the oracle, one pair product per two observations of a
    free point by free cameras; a camera seeing a point twice is rejected."""
    import g2o_amd
    from g2o_amd.solver import G2oHipError
    bal = _small_problem(0.0)
    ba = g2o_amd.BALAdjuster(bal, device=-1)
    info = ba.structure_info()
    ora = O.OracleBAL(bal)
    assert (info[0], info[1], info[3]) == (ora.Cf, ora.Pf, ora.nschur)
    cf, pf = bal.cam_fixed.astype(bool), bal.pt_fixed.astype(bool)
    keep = ~cf[bal.edge_cam] & ~pf[bal.edge_pt]
    n = np.bincount(bal.edge_pt[keep], minlength=bal.num_pts)
    assert info[5] == int((n * (n - 1) // 2).sum())
    assert info[6] == int(keep.sum())
    dup = bal.copy()
    k = int(np.flatnonzero(keep)[0])  # an edge between a free camera and a free point
    dup.edge_cam = np.r_[dup.edge_cam, dup.edge_cam[k]].astype(np.int32)
    dup.edge_pt = np.r_[dup.edge_pt, dup.edge_pt[k]].astype(np.int32)
    dup.meas = np.r_[dup.meas, dup.meas[k:k + 1]]
    with pytest.raises(G2oHipError) as ex:
        g2o_amd.BALAdjuster(dup, device=-1)
    assert ex.value.status == g2o_amd._lib.ERR_UNSUPPORTED
