"""The C++ host classes (g2o_amd/host/g2o_hip.hpp): the reference's LM loop
written against the g2o::Solver calls of BlockSolverHIP (what a g2o plugin
adapter would drive) must reproduce the library's own device-resident
optimize, iteration by iteration."""
import os
import subprocess

import numpy as np
import pytest

import g2o_amd
from g2o_amd import datasets

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "g2o_amd", "host", "lm_drive")


def write_problem(path, prob, precision, refine):
    with open(path, "wb") as f:
        np.array([prob.num_cams, prob.num_pts, prob.num_edges, precision, refine], np.int32).tofile(f)
        for a, dt in ((prob.cams, np.float64), (prob.intr, np.float64), (prob.pts, np.float64),
                      (prob.edge_cam, np.int32), (prob.edge_pt, np.int32), (prob.meas, np.float64),
                      (prob.weight, np.float64), (prob.cam_fixed, np.uint8), (prob.pt_fixed, np.uint8)):
            np.ascontiguousarray(a, dtype=dt).tofile(f)
        np.array([prob.huber], np.float64).tofile(f)


def run(path, mode, iters):
    out = subprocess.run([DRIVER, path, mode, str(iters), "0"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    if mode == "solver":
        # the host loop took lambda_0 from the mapped vertex Hessians, as g2o's
        # computeLambdaInit does, and it equals the device's
        _lambda_line(out.stderr)
    return np.array([[float(v) for v in line.split()] for line in out.stdout.strip().splitlines()])


def _lambda_line(stderr):
    line = next(ln for ln in stderr.splitlines() if ln.startswith("lambda_init vertex_hessians"))
    v = line.split()
    assert float(v[2]) == float(v[4]) and float(v[2]) > 0
    return float(v[2])


@pytest.mark.parametrize("precision,refine", [(g2o_amd.FP64, 0), (g2o_amd.FP32, 1)])
def test_solver_contract_loop_matches_device_optimize(tmp_path, precision, refine):
    assert os.path.exists(DRIVER), "build first (make -C g2o_amd/csrc)"
    prob = datasets.make_walk_ba(seed=21, num_cams=40, num_pts=4000)
    path = str(tmp_path / "p.bin")
    write_problem(path, prob, precision, refine)
    a = run(path, "solver", 8)
    b = run(path, "optimize", 8)
    assert a.shape == b.shape and a.shape[0] >= 5
    # chi2 before / after and lambda: same decisions, values equal up to the
    # summation order of the two chi2 passes (FP64); the FP32 Schur sums its
    # pair products with float LDS atomics, so two FP32 runs differ in the
    # last float digits and PCG counts may differ by an iteration or two
    fp32 = precision == g2o_amd.FP32
    rtol = 1e-5 if fp32 else 1e-10
    np.testing.assert_allclose(a[:, 1:4], b[:, 1:4], rtol=rtol)
    np.testing.assert_array_equal(a[:, 4], b[:, 4])
    assert np.abs(a[:, 5] - b[:, 5]).max() <= (3 if fp32 else 0)
    # and the Python mirror of the same library
    ba = g2o_amd.BundleAdjuster(prob, device=0, precision=precision, refine_steps=refine)
    st = ba.optimize(8)
    np.testing.assert_allclose([s["chi2"] for s in st], b[:, 2], rtol=1e-5 if fp32 else 1e-10)


def test_g2o_file_through_the_cpp_driver(tmp_path):
    """load (GraphIO) -> BlockSolverHIP LM -> save, as g2o's command-line tool
    does with a file: the trace equals the Python mirror's on the same file and
    the saved graph holds the optimised state."""
    from g2o_amd import graphio as gio
    prob = datasets.make_small_ba(seed=1000)
    fx, fy, cx, cy = (float(v) for v in prob.intr[0])
    assert np.all(prob.intr == prob.intr[0])
    src = str(tmp_path / "small.g2o")
    dst = str(tmp_path / "small_opt.g2o")
    gio.save_ba_problem(src, prob)
    out = subprocess.run([DRIVER, src, "optimize", "6", "0", repr(fx), repr(fy), repr(cx), repr(cy), dst],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    a = np.array([[float(v) for v in line.split()] for line in out.stdout.strip().splitlines()])
    loaded = gio.G2oGraph.load(src).to_ba_problem(prob.intr)
    ba = g2o_amd.BundleAdjuster(loaded, device=0, precision=g2o_amd.FP64)
    st = ba.optimize(6)
    np.testing.assert_allclose(a[:, 2], [s["chi2"] for s in st], rtol=1e-10)
    cams, pts = ba.state()
    saved = gio.G2oGraph.load(dst).to_ba_problem(prob.intr)
    assert np.max(np.abs(saved.cams - cams)) < 1e-12
    assert np.max(np.abs(saved.pts - pts)) < 1e-9 * np.max(np.abs(pts))


@pytest.mark.parametrize("mode,algorithm", [("optimize", g2o_amd.LEVENBERG), ("gn", g2o_amd.GAUSS_NEWTON),
                                            ("dogleg", g2o_amd.DOGLEG)])
def test_pose_graph_file_through_the_cpp_driver(tmp_path, mode, algorithm):
    """A VERTEX_SE3:QUAT / EDGE_SE3:QUAT file through GraphIO and
    PoseGraphSolverHIP with each algorithm: the trace equals the Python
    mirror's on the same file, and the saved graph holds the final poses."""
    from g2o_amd import graphio as gio
    pg = datasets.make_sphere_pose_graph(nodes_per_level=60, laps=4)
    src = str(tmp_path / "sphere.g2o")
    dst = str(tmp_path / "sphere_opt.g2o")
    gio.save_pose_graph(src, pg)
    out = subprocess.run([DRIVER, src, mode, "5", "0", dst], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    a = np.array([[float(v) for v in line.split()] for line in out.stdout.strip().splitlines()])
    loaded = gio.G2oGraph.load(src).to_pose_graph()
    opt = g2o_amd.PoseGraphOptimizer(loaded, device=0)  # PCG tolerance 1e-8, as PoseGraphSolverHIP
    opt.set_algorithm(algorithm)
    st = opt.optimize(5)
    assert a.shape[0] == len(st)
    np.testing.assert_allclose(a[:, 2], [s["chi2"] for s in st], rtol=1e-10)
    saved = gio.G2oGraph.load(dst).to_pose_graph()
    poses = opt.state()
    assert np.max(np.abs(saved.poses[:, :3] - poses[:, :3])) < 1e-9 * max(1.0, np.max(np.abs(poses[:, :3])))


def test_full_information_file_through_the_solver_contract_loop(tmp_path):
    """A file whose projection edges carry full 2x2 information matrices
    (EdgeSE3ProjectXYZ::read, edge_project_xyz.cpp:34-42) through GraphIO and
    the host LM loop over BlockSolverHIP: lambda_0 from the vertex Hessians
    mapped after buildSystem, and the trace equal to the Python mirror's on the
    same problem."""
    from g2o_amd import graphio as gio
    prob = datasets.make_small_ba(seed=1000)
    prob.omega = datasets.random_information(prob.num_edges, seed=9)
    fx, fy, cx, cy = (float(v) for v in prob.intr[0])
    src = str(tmp_path / "info.g2o")
    gio.save_ba_problem(src, prob)
    out = subprocess.run([DRIVER, src, "solver", "6", "0", repr(fx), repr(fy), repr(cx), repr(cy)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lam0 = _lambda_line(out.stderr)
    a = np.array([[float(v) for v in line.split()] for line in out.stdout.strip().splitlines()])
    ba = g2o_amd.BundleAdjuster(gio.G2oGraph.load(src).to_ba_problem(prob.intr), device=0,
                                precision=g2o_amd.FP64)
    ba.build_system()
    # another context: its linearisation's LDS sums may round the diagonal
    # differently in the last bit
    assert ba.lambda_init() == pytest.approx(lam0, rel=1e-14)
    ba.set_state(prob.cams, prob.pts)
    st = ba.optimize(6)
    np.testing.assert_allclose(a[:, 2], [s["chi2"] for s in st], rtol=1e-10)
    np.testing.assert_array_equal(a[:, 4], [s["levenberg_iterations"] for s in st])


def test_xyz2uv_file_through_the_cpp_driver(tmp_path):
    """A ba_demo-style file carries its intrinsics: the driver needs none."""
    from g2o_amd import graphio as gio
    prob = datasets.make_small_ba(seed=1000)
    plain = tmp_path / "plain.g2o"
    gio.save_ba_problem(str(plain), prob)
    lines = ["PARAMS_CAMERAPARAMETERS 0 500 320 240 0.5"]
    for ln in plain.read_text().splitlines():
        if ln.startswith("EDGE_SE3_PROJECT_XYZ:EXPMAP "):
            t = ln.split(" ", 3)
            ln = f"EDGE_PROJECT_XYZ2UV:EXPMAP {t[1]} {t[2]} 0 {t[3]}"
        lines.append(ln)
    src = tmp_path / "xyz2uv.g2o"
    src.write_text("\n".join(lines) + "\n")
    dst = str(tmp_path / "xyz2uv_opt.g2o")
    out = subprocess.run([DRIVER, str(src), "optimize", "5", "0", dst], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    a = np.array([[float(v) for v in line.split()] for line in out.stdout.strip().splitlines()])
    ba = g2o_amd.BundleAdjuster(prob, device=0, precision=g2o_amd.FP64)
    st = ba.optimize(5)
    np.testing.assert_allclose(a[:, 2], [s["chi2"] for s in st], rtol=1e-10)
    saved = gio.G2oGraph.load(dst)
    assert saved.info()["xyz2uv_edges"] == prob.num_edges


def test_bal_file_through_the_cpp_solver_contract(tmp_path, oracle_lib):
    """A BAL file through BALSolverHIP (BlockSolver<9,3>): the reference's LM
    loop on the Solver calls, with lambda_0 from the mapped 9x9 / 3x3 vertex
    Hessians, equals the device loop and the oracle, iteration by
    iteration."""
    from g2o_amd import graphio
    assert os.path.exists(DRIVER), "build first (make -C g2o_amd/csrc)"
    prob = datasets.make_walk_bal(seed=47, num_cams=30, num_pts=1500)
    path = str(tmp_path / "walk.bal")
    graphio.save_bal(path, prob)
    a = run(path, "solver", 8)
    b = run(path, "optimize", 8)
    assert a.shape == b.shape and a.shape[0] == 8
    assert np.allclose(a[:, 1:4], b[:, 1:4], rtol=1e-9)
    assert np.array_equal(a[:, 4], b[:, 4])
    ora = oracle_lib.OracleBAL(prob)
    n, trace, _ = ora.optimize(8, solver=1, tol=1e-6)
    assert n == 8
    assert np.allclose(b[:, 2], trace[:, 1], rtol=1e-8)
    assert np.array_equal(b[:, 4], trace[:, 3])
