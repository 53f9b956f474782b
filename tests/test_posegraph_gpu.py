"""3D pose graph (EdgeSE3 / VertexSE3, 6x6 blocks, PCG) on the GPU against
the FP64 CPU oracle: errors, Hessian blocks, LM traces, the reference's
two-vertex tests, and the 10k-pose sphere graph."""
import time

import numpy as np
import pytest

import g2o_amd
from g2o_amd import datasets

pytestmark = pytest.mark.gpu


def _oracle(prob):
    from oracle import oracle
    return oracle.OraclePG(prob)


def test_errors_and_system_match_oracle():
    prob = datasets.make_sphere_pose_graph(seed=3, nodes_per_level=40, laps=4)
    ora = _oracle(prob)
    pgo = g2o_amd.PoseGraphOptimizer(prob, device=0)
    chi_o, rows_o, cols_o, blocks_o, b_o = ora.linearize()
    chi = pgo.compute_active_errors()
    assert abs(chi - chi_o) <= 1e-11 * chi_o
    pgo.build_system()
    rows, cols, blocks, b = pgo.system()
    key_o = {(r, c): k for k, (r, c) in enumerate(zip(rows_o, cols_o))}
    assert len(key_o) == len(rows)
    perm = np.array([key_o[(r, c)] for r, c in zip(rows, cols)])
    scale = np.abs(blocks_o).max()
    assert np.abs(blocks - blocks_o[perm]).max() <= 1e-11 * scale
    assert np.abs(b - b_o).max() <= 1e-11 * np.abs(b_o).max()
    e = pgo.errors()
    e_o = np.stack([ora.edge(k)[0] for k in range(prob.num_edges)])
    assert np.abs(e - e_o).max() <= 1e-12


def test_damped_solve_matches_dense_solve():
    """One setLambda + solve with a tight PCG tolerance equals the dense FP64
    solution of (H + lambda I) x = b built from the oracle's blocks."""
    prob = datasets.make_sphere_pose_graph(seed=4, nodes_per_level=50, laps=4)
    ora = _oracle(prob)
    _, rows, cols, blocks, b = ora.linearize()
    pgo = g2o_amd.PoseGraphOptimizer(prob, device=0, pcg_tolerance=1e-24)
    pgo.build_system()
    lam = pgo.lambda_init()
    diag = np.array([np.abs(np.diag(blocks[k])).max() for k in range(len(rows)) if rows[k] == cols[k]])
    assert abs(lam - 1e-5 * diag.max()) <= 1e-12 * lam
    pgo.set_lambda(lam)
    assert pgo.solve()
    n = 6 * pgo.Nf
    A = np.zeros((n, n))
    for r, c, B in zip(rows, cols, blocks):
        A[6 * r:6 * r + 6, 6 * c:6 * c + 6] = B
        A[6 * c:6 * c + 6, 6 * r:6 * r + 6] = B.T
    A += lam * np.eye(n)
    x_ref = np.linalg.solve(A, b)
    x = pgo.x()
    assert np.linalg.norm(x - x_ref) <= 1e-8 * np.linalg.norm(x_ref)


def test_lm_trace_matches_oracle():
    """Same LM control as the oracle. The PCG stops at tolerance 1e-8 on an
    ill-conditioned chain, so host and device steps differ in the last
    digits (tolerance 1e-8 on r.M^-1 r leaves a relative residual ~1e-4)
    and the LM paths drift apart (per-iteration chi2 within 1e-4, the first
    solve's iteration count equal); the final chi2 agrees to 1e-6."""
    prob = datasets.make_sphere_pose_graph(seed=4, nodes_per_level=50, laps=4)
    ora = _oracle(prob)
    n, trace, _ = ora.optimize(15, solver=1, tol=1e-8)
    pgo = g2o_amd.PoseGraphOptimizer(prob, device=0, pcg_tolerance=1e-8)
    stats = pgo.optimize(15)
    assert len(stats) == n
    dev = max(abs(s["chi2"] - t[1]) / t[1] for s, t in zip(stats, trace))
    print(f"\nmax relative chi2 deviation per iteration {dev:.2e}", flush=True)
    assert abs(stats[0]["iterations_linear_solver"] - int(trace[0, 4])) <= 1
    assert abs(stats[0]["lambda_"] - trace[0, 2]) <= 1e-10 * trace[0, 2]
    for s, t in zip(stats, trace):
        assert abs(s["chi2"] - t[1]) <= 1e-4 * t[1]
        assert s["levenberg_iterations"] == int(t[3])
    assert abs(stats[-1]["chi2"] - trace[-1, 1]) <= 1e-6 * trace[-1, 1]


def test_lm_with_rejected_trials_matches_oracle():
    """A 60-pose sphere whose initial poses are scrambled (1.5 m, 1.5 rad):
    the LM rejects trials with rho well below zero, up to four in a row in
    one iteration. The device takes the same decisions as the oracle: the
    same trials per iteration and the same lambda after each (nu doubling on
    rejection, reset on acceptance, levenberg.cpp:125-141)."""
    pg = datasets.make_sphere_pose_graph(seed=1001, nodes_per_level=20, laps=3)
    rng = np.random.default_rng(23)
    P = datasets.perturb_poses(rng, pg.poses, 1.5, 1.5)
    P[0] = pg.poses[0]
    pg.poses = P
    # PCG run to 1e-20 on r.M^-1 r: the oracle's steps then agree with its
    # dense Cholesky solve to ~1e-9, so the two paths do not drift apart
    ora = _oracle(pg)
    n, trace, _ = ora.optimize(12, solver=1, tol=1e-20)
    assert trace[:, 3].max() >= 3
    pgo = g2o_amd.PoseGraphOptimizer(pg, device=0, pcg_tolerance=1e-20)
    stats = pgo.optimize(12)
    assert len(stats) == n
    for s, t in zip(stats, trace):
        assert s["levenberg_iterations"] == int(t[3])
        assert abs(s["lambda_"] - t[2]) <= 1e-6 * t[2]
        assert abs(s["chi2"] - t[1]) <= 1e-5 * t[1]
    # trial by trial: lambda used, chi2 before and after, rho, the decision
    tg, to = pgo.lm_trials(), ora.trials()
    assert tg.shape[0] == to.shape[0] and (to[:, 5] == 0).sum() >= 3
    assert np.array_equal(tg[:, 5], to[:, 5])
    assert np.allclose(tg[:, 0], to[:, 0], rtol=1e-6)
    assert np.allclose(tg[:, 1], to[:, 1], rtol=1e-5)
    assert np.allclose(tg[:, 4], to[:, 4], rtol=1e-4, atol=1e-6)


@pytest.mark.parametrize("which", ["translation", "rotation"])
def test_two_vertex_lm_like_reference(which):
    """optimization_slam3d.cpp:62-138 on the device."""
    I7 = np.array([0, 0, 0, 0, 0, 0, 1.0])
    p2 = I7.copy()
    if which == "translation":
        p2[:3] = 10.0
    else:
        a = np.deg2rad(2)
        p2[3:] = np.concatenate([np.sin(a / 2) * np.ones(3) / np.sqrt(3), [np.cos(a / 2)]])
    prob = datasets.PoseGraphProblem(np.stack([I7, p2]), np.array([0], np.int32), np.array([1], np.int32),
                                     I7[None], np.eye(6)[None], np.array([1, 0], np.uint8))
    pgo = g2o_amd.PoseGraphOptimizer(prob, device=0, pcg_tolerance=1e-6)
    assert pgo.compute_active_errors() > 0
    stats = pgo.optimize(100)
    assert len(stats) > 0
    assert pgo.compute_active_errors() < 1e-6
    st = pgo.state()
    assert np.abs(st[1, :3]).max() < 1e-3
    assert np.abs(np.diag(datasets.quat_to_rot(st[1, 3:])) - 1).max() < 1e-3


def test_sphere_10k_poses_matches_oracle():
    """10k poses on 10 laps, ~36k edges, PCG tolerance 1e-8, 20 LM
    iterations from the odometry chain; first pose fixed."""
    prob = datasets.make_sphere_pose_graph()
    pgo = g2o_amd.PoseGraphOptimizer(prob, device=0, pcg_tolerance=1e-8)
    pgo.optimize(1)
    pgo.set_state(prob.poses)
    t = time.perf_counter()
    stats = pgo.optimize(20)
    dt = time.perf_counter() - t
    ora = _oracle(prob)
    t = time.perf_counter()
    n, trace, _ = ora.optimize(20, solver=1, tol=1e-8)
    dto = time.perf_counter() - t
    print(f"\nsphere 10k: GPU {len(stats)} LM iterations in {dt * 1e3:.1f} ms "
          f"({len(stats) / dt:.1f} it/s), oracle {n} in {dto:.2f} s; chi2 {stats[-1]['chi2']:.6e} vs "
          f"{trace[-1, 1]:.6e}; PCG iterations {[s['iterations_linear_solver'] for s in stats]}", flush=True)
    assert len(stats) == n
    assert abs(stats[-1]["chi2"] - trace[-1, 1]) <= 1e-6 * trace[-1, 1]
    gt = prob.gt_poses
    err = np.linalg.norm(pgo.state()[:, :3] - gt[:, :3], axis=1)
    assert err.max() < 0.5


def test_multiply_hessian_is_the_hessian_product():
    """BlockSolverBase::multiplyHessian on the pose graph: H src with the whole
    undamped Hessian of the last buildSystem (both triangles)."""
    prob = datasets.make_sphere_pose_graph(seed=3, nodes_per_level=40, laps=4)
    pgo = g2o_amd.PoseGraphOptimizer(prob, device=0)
    pgo.build_system()
    rows, cols, blocks, _ = pgo.system()
    n = 6 * pgo.Nf
    H = np.zeros((n, n))
    for r, c, B in zip(rows, cols, blocks):
        H[6 * r:6 * r + 6, 6 * c:6 * c + 6] = B
        if r != c:
            H[6 * c:6 * c + 6, 6 * r:6 * r + 6] = B.T
    src = np.random.default_rng(5).normal(size=n)
    out = pgo.multiply_hessian(src)
    ref = H @ src
    assert np.max(np.abs(out - ref)) <= 1e-12 * np.max(np.abs(ref))
