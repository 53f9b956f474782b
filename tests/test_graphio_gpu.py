"""Files through the GPU path: a g2o or BAL file, loaded by the native graph
container, runs through the C ABI exactly as the in-memory problem it was
written from, and the optimised state saved back reloads unchanged.

Tolerances: the loaded cameras differ from the written ones by the rounding
of the file's camera-to-world inversion (<= 1e-14 absolute), so chi2 and LM
traces agree to 1e-10 relative; the BAL chi2 (k1 = k2 = 0) matches the numpy
BAL model to 1e-11.
"""
import numpy as np
import pytest

from g2o_amd import datasets
from g2o_amd import graphio as gio

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_g2o_file_runs_like_the_problem_it_holds(tmp_path, oracle_lib):
    import g2o_amd
    p = datasets.make_small_ba(seed=1000)
    path = str(tmp_path / "small.g2o")
    gio.save_ba_problem(path, p)
    g = gio.G2oGraph.load(path)
    q = g.to_ba_problem(p.intr)
    _say(f"g2o file: C={q.num_cams} P={q.num_pts} E={q.num_edges}")
    a = g2o_amd.BundleAdjuster(q, device=0, precision=g2o_amd.FP64)
    b = g2o_amd.BundleAdjuster(p, device=0, precision=g2o_amd.FP64)
    ca, cb = a.compute_active_errors(), b.compute_active_errors()
    assert abs(ca - cb) <= 1e-10 * cb
    sa, sb = a.optimize(6), b.optimize(6)
    assert len(sa) == len(sb)
    for x, y in zip(sa, sb):
        assert abs(x["chi2"] - y["chi2"]) <= 1e-10 * y["chi2"]
        assert x["levenberg_iterations"] == y["levenberg_iterations"]
    # against the oracle's LM on the loaded problem
    n_o, trace, _ = oracle_lib.OracleBA(q).optimize(6, solver=1, tol=1e-6)
    assert n_o == len(sa)
    for x, t in zip(sa, trace):
        assert abs(x["chi2"] - t[1]) <= 1e-8 * t[1]
    # save the optimised state and read it back
    cams, pts = a.state()
    g.set_ba_state(cams, pts)
    out = str(tmp_path / "small_opt.g2o")
    g.save(out)
    r = gio.G2oGraph.load(out).to_ba_problem(p.intr)
    assert np.max(np.abs(r.cams - cams)) < 1e-13
    assert np.array_equal(r.pts, pts)
    c = g2o_amd.BundleAdjuster(r, device=0, precision=g2o_amd.FP64)
    assert abs(c.compute_active_errors() - sa[-1]["chi2"]) <= 1e-10 * sa[-1]["chi2"]


def test_pose_graph_file(tmp_path):
    import g2o_amd
    pg = datasets.make_sphere_pose_graph(nodes_per_level=100, laps=5)
    path = str(tmp_path / "sphere.g2o")
    gio.save_pose_graph(path, pg)
    r = gio.G2oGraph.load(path).to_pose_graph()
    _say(f"pose graph file: N={r.num_poses} E={r.num_edges}")
    a = g2o_amd.PoseGraphOptimizer(r, device=0)
    b = g2o_amd.PoseGraphOptimizer(pg, device=0)
    sa, sb = a.optimize(5), b.optimize(5)
    assert len(sa) == len(sb)
    for x, y in zip(sa, sb):
        assert abs(x["chi2"] - y["chi2"]) <= 1e-9 * y["chi2"]


def test_bal_file_chi2_and_lm(tmp_path, oracle_lib):
    import g2o_amd
    bal = datasets.make_bal(21, num_cams=10, num_pts=400)
    path = str(tmp_path / "p.txt")
    gio.save_bal(path, bal)
    prob = gio.bal_to_ba(gio.load_bal(path))
    _say(f"BAL file: C={prob.num_cams} P={prob.num_pts} E={prob.num_edges}")
    ba = g2o_amd.BundleAdjuster(prob, device=0, precision=g2o_amd.FP64)
    want = gio.bal_chi2(bal)
    assert abs(ba.compute_active_errors() - want) <= 1e-11 * want
    prob.cam_fixed[0] = 1  # gauge
    ba = g2o_amd.BundleAdjuster(prob, device=0, precision=g2o_amd.FP64)
    s = ba.optimize(5)
    n_o, trace, _ = oracle_lib.OracleBA(prob).optimize(5, solver=1, tol=1e-6)
    assert n_o == len(s)
    for x, t in zip(s, trace):
        assert abs(x["chi2"] - t[1]) <= 1e-8 * t[1]
    assert s[-1]["chi2"] < want


def test_xyz2uv_file_runs_like_the_plain_edges(tmp_path, oracle_lib):
    """A ba_demo-style file (PARAMS_CAMERAPARAMETERS + EDGE_PROJECT_XYZ2UV)
    takes its intrinsics from the file and runs as EDGE_SE3_PROJECT_XYZ edges
    with fx = fy = f."""
    import g2o_amd
    p = datasets.make_small_ba(seed=1000)
    plain = tmp_path / "plain.g2o"
    gio.save_ba_problem(str(plain), p)
    lines = ["PARAMS_CAMERAPARAMETERS 0 500 320 240 0.5"]
    for ln in plain.read_text().splitlines():
        if ln.startswith("EDGE_SE3_PROJECT_XYZ:EXPMAP "):
            t = ln.split(" ", 3)
            ln = f"EDGE_PROJECT_XYZ2UV:EXPMAP {t[1]} {t[2]} 0 {t[3]}"
        lines.append(ln)
    path = tmp_path / "xyz2uv.g2o"
    path.write_text("\n".join(lines) + "\n")
    q = gio.G2oGraph.load(str(path)).to_ba_problem()
    a = g2o_amd.BundleAdjuster(q, device=0, precision=g2o_amd.FP64)
    b = g2o_amd.BundleAdjuster(p, device=0, precision=g2o_amd.FP64)
    sa, sb = a.optimize(4), b.optimize(4)
    assert len(sa) == len(sb)
    for x, y in zip(sa, sb):
        assert abs(x["chi2"] - y["chi2"]) <= 1e-10 * y["chi2"]
