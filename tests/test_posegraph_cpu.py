"""Pins the pose-graph oracle (VertexSE3 / EdgeSE3, 6x6 blocks, no Schur):
closed-form errors and Jacobians at known configurations, Jacobians against
central differences through VertexSE3's right-multiplied increment
(vertex_se3.h:99-108, g2o's own numeric scheme in base_fixed_sized_edge.hpp:
152-193), the reference's two-vertex LM tests (optimization_slam3d.cpp:62-138)
and LM on a small sphere graph."""
import numpy as np
import pytest

from g2o_amd import datasets
from oracle import oracle


def incr(pose, d):
    """pose * fromVectorMQT(d) (isometry3d_mappings.cpp:120-125)."""
    v = d[3:]
    w2 = 1.0 - v @ v
    q = np.array([0, 0, 0, 1.0]) if w2 < 0 else np.array([v[0], v[1], v[2], np.sqrt(w2)])
    return datasets._compose(pose, np.concatenate([d[:3], q]))


def two_node(pi, pj, z, info=None):
    prob = datasets.PoseGraphProblem(np.stack([pi, pj]), np.array([0], np.int32), np.array([1], np.int32),
                                     z[None], (np.eye(6) if info is None else info)[None],
                                     np.zeros(2, np.uint8))
    return oracle.OraclePG(prob)


def random_pose(rng, scale=3.0):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    q *= np.sign(q[3])
    return np.concatenate([rng.normal(0, scale, 3), q])


def test_identity_error_and_jacobians():
    I7 = np.array([0, 0, 0, 0, 0, 0, 1.0])
    e, J = two_node(I7, I7, I7).edge(0)
    assert np.abs(e).max() == 0
    np.testing.assert_allclose(J[:, :6], -np.eye(6), atol=1e-15)
    np.testing.assert_allclose(J[:, 6:], np.eye(6), atol=1e-15)


def test_error_is_relative_pose_residual():
    rng = np.random.default_rng(1)
    for _ in range(20):
        pi, pj = random_pose(rng), random_pose(rng)
        z = datasets._relative(pi, pj)
        e, _ = two_node(pi, pj, z).edge(0)
        assert np.abs(e).max() < 1e-12
        # a measurement off by a known increment d gives toVectorMQT(d^-1)
        d = np.concatenate([rng.normal(0, 0.1, 3), rng.normal(0, 0.05, 3)])
        e, _ = two_node(pi, pj, incr(z, d)).edge(0)
        dinv = datasets._relative(incr(np.array([0, 0, 0, 0, 0, 0, 1.0]), d), np.array([0, 0, 0, 0, 0, 0, 1.0]))
        np.testing.assert_allclose(e, np.concatenate([dinv[:3], dinv[3:6]]), atol=1e-12)


def test_jacobians_match_central_differences():
    rng = np.random.default_rng(2)
    h = 1e-6
    worst = 0.0
    for trial in range(300):
        pi, pj = random_pose(rng), random_pose(rng)
        z = incr(datasets._relative(pi, pj), np.concatenate([rng.normal(0, 0.3, 3), rng.normal(0, 0.2, 3)]))
        ora = two_node(pi, pj, z)
        _, J = ora.edge(0)
        Jn = np.zeros((6, 12))
        for v, base in ((0, pi), (1, pj)):
            for k in range(6):
                d = np.zeros(6)
                d[k] = h
                st = [pi, pj]
                st[v] = incr(base, d)
                ora.set_state(np.stack(st))
                ep, _ = ora.edge(0)
                st[v] = incr(base, -d)
                ora.set_state(np.stack(st))
                em, _ = ora.edge(0)
                Jn[:, 6 * v + k] = (ep - em) / (2 * h)
        worst = max(worst, np.abs(J - Jn).max())
    assert worst < 1e-6, worst


@pytest.mark.parametrize("which", ["translation", "rotation"])
def test_two_vertex_lm_like_reference(which):
    I7 = np.array([0, 0, 0, 0, 0, 0, 1.0])
    p2 = I7.copy()
    if which == "translation":
        p2[:3] = 10.0
    else:
        ax = np.ones(3) / np.sqrt(3)
        a = np.deg2rad(2)
        p2[3:] = np.concatenate([np.sin(a / 2) * ax, [np.cos(a / 2)]])
    prob = datasets.PoseGraphProblem(np.stack([I7, p2]), np.array([0], np.int32), np.array([1], np.int32),
                                     I7[None], np.eye(6)[None], np.array([1, 0], np.uint8))
    for solver in (0, 1):
        ora = oracle.OraclePG(prob)
        assert ora.chi2() > 0
        n, trace, _ = ora.optimize(100, solver=solver)
        assert n > 0
        assert ora.chi2() < 1e-6
        st = ora.state()
        assert np.abs(st[1, :3]).max() < 1e-3
        R = datasets.quat_to_rot(st[1, 3:])
        assert np.abs(np.diag(R) - 1).max() < 1e-3


def test_sphere_lm_converges_and_pcg_matches_cholesky():
    prob = datasets.make_sphere_pose_graph(seed=3, nodes_per_level=40, laps=4)
    res = {}
    for solver in (0, 1):
        ora = oracle.OraclePG(prob)
        chi0 = ora.chi2()
        n, trace, _ = ora.optimize(20, solver=solver, tol=1e-8)
        assert np.all(np.diff(trace[:, 1]) <= 1e-9 * trace[0, 0])
        res[solver] = (chi0, ora.chi2(), ora.state())
    chi0, chi_c, st_c = res[0]
    _, chi_p, st_p = res[1]
    # chi2 at the optimum ~ dof = 6 (E - N + 1)
    dof = 6 * (prob.num_edges - prob.num_poses + 1)
    assert chi_c < 2 * dof < chi0
    assert abs(chi_p - chi_c) <= 1e-6 * chi_c
    assert np.abs(st_p[:, :3] - st_c[:, :3]).max() < 1e-4
