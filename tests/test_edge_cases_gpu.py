"""Ragged and boundary inputs through the C ABI against the oracle.

* mixed: fixed points, inactive edges (fixed camera and fixed point), points
  with a single observation, and a free camera and a free point without edges
  (not active, sparse_optimizer.cpp:217-285): chi2 and Hessian blocks 1e-11,
  LM trace 1e-8 with the same trial and PCG counts, FP64 and FP32;
* the per-point limit (400 free cameras): a point seen by 400 free cameras
  matches the oracle, one more is rejected with G2OHIP_ERR_UNSUPPORTED before
  any device work;
* no edge at all: rejected (the reference's optimize returns -1, "0 vertices
  to optimize", sparse_optimizer.cpp:393-399).
"""
import numpy as np
import pytest

from g2o_amd import datasets

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _mixed():
    p = datasets.make_walk_ba(seed=11, num_cams=40, num_pts=3000)
    rng = np.random.default_rng(4)
    # drop edges at random so that some points keep a single observation
    keep = rng.random(p.num_edges) > 0.35
    p = datasets.select_edges(p, keep, min_obs=1)
    p.pt_fixed = p.pt_fixed.copy()
    p.pt_fixed[::13] = 1
    p.cam_fixed = p.cam_fixed.copy()
    p.cam_fixed[5] = 1
    # a free camera and a free point that no edge uses
    p = datasets.BAProblem(np.vstack([p.cams, p.cams[-1:]]), np.vstack([p.intr, p.intr[-1:]]),
                           np.vstack([p.pts, p.pts[-1:] + 1.0]), p.edge_cam, p.edge_pt, p.meas, p.weight,
                           np.r_[p.cam_fixed, 0].astype(np.uint8), np.r_[p.pt_fixed, 0].astype(np.uint8), p.huber)
    return p


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_mixed_fixed_inactive_single_and_isolated(precision, oracle_lib):
    import g2o_amd
    p = _mixed()
    deg = np.bincount(p.edge_pt, minlength=p.num_pts)
    both = (p.cam_fixed[p.edge_cam] == 1) & (p.pt_fixed[p.edge_pt] == 1)
    _say(f"mixed: C={p.num_cams} P={p.num_pts} E={p.num_edges} single-observation points={int((deg == 1).sum())} "
         f"inactive edges={int(both.sum())}")
    assert (deg == 1).sum() > 0 and both.sum() > 0
    prec = g2o_amd.FP64 if precision == "fp64" else g2o_amd.FP32
    ba = g2o_amd.BundleAdjuster(p, device=0, precision=prec)
    ora = oracle_lib.OracleBA(p)
    assert (ba.Cf, ba.Pf) == (ora.Cf, ora.Pf)
    if precision == "fp64":
        chi_o, Hpp_o, Hll_o, _, b_o = ora.linearize()
        assert abs(ba.compute_active_errors() - chi_o) <= 1e-11 * chi_o
        ba.build_system()
        Hpp, Hll, _ = ba.hessian()
        assert np.max(np.abs(Hpp - Hpp_o)) <= 1e-11 * np.max(np.abs(Hpp_o))
        assert np.max(np.abs(Hll - Hll_o)) <= 1e-11 * np.max(np.abs(Hll_o))
        assert np.max(np.abs(ba.b() - b_o)) <= 1e-11 * np.max(np.abs(b_o))
    n, tr, _ = ora.optimize(8, solver=1, tol=1e-6)
    st = ba.optimize(8)
    _say(f"gpu {[s['chi2'] for s in st]}")
    _say(f"cpu {list(tr[:, 1])}")
    assert len(st) == n
    if precision == "fp64":
        for s, t in zip(st, tr):
            assert abs(s["chi2"] - t[1]) <= 1e-8 * t[1]
            assert s["levenberg_iterations"] == int(t[3])
            assert s["iterations_linear_solver"] == int(t[4])
    else:
        assert abs(st[-1]["chi2"] - tr[-1, 1]) <= 1e-6 * tr[-1, 1]
    # the isolated vertices keep their estimates
    cams, pts = ba.state()
    assert np.array_equal(cams[-1], p.cams[-1]) and np.array_equal(pts[-1], p.pts[-1])


def test_point_seen_by_the_most_free_cameras(oracle_lib):
    """A point may be observed by up to 400 free cameras (the LDS camera
    accumulators of the chunk kernels): 402 cameras on a circle, two of them
    fixed, every point seen by all of them, match the oracle; one more camera
    is rejected as unsupported."""
    import g2o_amd
    p = datasets.make_small_ba(seed=3, num_cams=402, num_pts=8)
    assert np.bincount(p.edge_pt).max() == 402
    ba = g2o_amd.BundleAdjuster(p, device=0, precision=g2o_amd.FP64)
    info = ba.structure_info()
    _say(f"free cameras {info[0]}, max observations per point {info[8]}, chunks {info[4]}")
    ora = oracle_lib.OracleBA(p)
    chi_o = ora.linearize()[0]
    assert abs(ba.compute_active_errors() - chi_o) <= 1e-11 * chi_o
    n, tr, _ = ora.optimize(3, solver=1, tol=1e-6)
    st = ba.optimize(3)
    assert len(st) == n
    for s, t in zip(st, tr):
        assert abs(s["chi2"] - t[1]) <= 1e-8 * t[1]
    q = datasets.make_small_ba(seed=3, num_cams=403, num_pts=8)
    with pytest.raises(g2o_amd.G2oHipError) as ei:
        g2o_amd.BundleAdjuster(q, device=0)
    assert ei.value.status == g2o_amd._lib.ERR_UNSUPPORTED


def test_no_edges_is_rejected():
    import g2o_amd
    p = datasets.make_small_ba(seed=1)
    empty = datasets.BAProblem(p.cams, p.intr, p.pts, p.edge_cam[:0], p.edge_pt[:0], p.meas[:0], p.weight[:0],
                               p.cam_fixed, p.pt_fixed, 0.0)
    with pytest.raises(g2o_amd.G2oHipError) as ei:
        g2o_amd.BundleAdjuster(empty, device=0)
    assert ei.value.status == g2o_amd._lib.ERR_UNSUPPORTED or ei.value.status == g2o_amd._lib.ERR_INVALID
