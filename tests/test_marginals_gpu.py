"""SparseOptimizer::computeMarginals on the device (g2ohip_compute_marginals)
against the oracle's linearisation inverted by numpy.

Bundle adjustment: the reference inverts Hpp (block_solver.hpp:479-490), which
is block diagonal for projection edges: diagonal blocks equal numpy's inverse
of the oracle's Hpp blocks to 1e-10 relative, other blocks are zero. Pose
graph: blocks of the inverse of the whole Hessian, from PCG solves at
tolerance 1e-24, equal to numpy's dense inverse to 1e-8 of the block norm.
The requested pattern is g2o.cpp:620-649's: (i, i) and (i-1, i).
"""
import numpy as np
import pytest

from g2o_amd import datasets

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _pattern(n):
    out = []
    for i in range(n):
        out.append((i, i))
        if i > 0:
            out.append((i - 1, i))
    return out


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_ba_marginals_invert_hpp(precision, oracle_lib):
    import g2o_amd
    prob = datasets.make_small_ba(seed=1000)
    prec = g2o_amd.FP64 if precision == "fp64" else g2o_amd.FP32
    ba = g2o_amd.BundleAdjuster(prob, device=0, precision=prec)
    with pytest.raises(g2o_amd.G2oHipError):
        ba.compute_marginals([(0, 0)])  # nothing linearised yet
    ba.optimize(4)
    ba.build_system()
    pat = _pattern(ba.Cf)
    M = ba.compute_marginals(pat)
    ora = oracle_lib.OracleBA(prob)
    ora.set_state(*ba.state())
    _, Hpp, _, _, _ = ora.linearize()
    for (r, c), blk in zip(pat, M):
        if r == c:
            want = np.linalg.inv(Hpp[r])
            assert np.max(np.abs(blk - want)) <= 1e-10 * np.max(np.abs(want)), (r, c)
            assert np.allclose(blk, blk.T, rtol=0, atol=1e-14 * np.max(np.abs(want)))
        else:
            assert not np.any(blk), (r, c)
    with pytest.raises(g2o_amd.G2oHipError):
        ba.compute_marginals([(0, ba.Cf)])


def test_pose_graph_marginals_match_dense_inverse(oracle_lib):
    import g2o_amd
    pg = datasets.make_sphere_pose_graph(nodes_per_level=30, laps=3)
    opt = g2o_amd.PoseGraphOptimizer(pg, device=0)
    opt.optimize(3)
    opt.build_system()
    Nf = opt.Nf
    _say(f"pose graph marginals: N={pg.num_poses} free={Nf}")
    rng = np.random.default_rng(0)
    cols = sorted(set([1, 2, 3, Nf - 1] + list(rng.integers(1, Nf, 6))))
    pat = []
    for c in cols:
        pat += [(c, c), (c - 1, c)]
    pat.append((0, Nf - 1))  # far from the diagonal
    M = opt.compute_marginals(pat, tol=1e-24)
    ora = oracle_lib.OraclePG(pg)
    ora.set_state(opt.state())
    _, rows, cbl, blocks, _ = ora.linearize()
    H = np.zeros((6 * Nf, 6 * Nf))
    for r, c, b in zip(rows, cbl, blocks):
        H[6 * r:6 * r + 6, 6 * c:6 * c + 6] = b
        if r != c:
            H[6 * c:6 * c + 6, 6 * r:6 * r + 6] = b.T
    Hi = np.linalg.inv(H)
    for (r, c), blk in zip(pat, M):
        want = Hi[6 * r:6 * r + 6, 6 * c:6 * c + 6]
        assert np.max(np.abs(blk - want)) <= 1e-8 * np.max(np.abs(Hi[6 * c:6 * c + 6, 6 * c:6 * c + 6])), (r, c)
    # the optimiser continues normally afterwards (the damping is restored)
    st = opt.optimize(2)
    assert st[-1]["chi2"] <= st[0]["chi2_before"]
