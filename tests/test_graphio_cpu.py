"""g2o text-format and BAL IO (libg2o_hip.so's host-side graph container) on
the CPU.

Pinned by:
* the number format: a C++ program writes the same values through std::ostream
  (what writeVector, g2o/core/io_helper.h:35-39, uses) at precision 6 and 17,
  and the saved file must match it byte for byte;
* the load rules of OptimizableGraph::load / addVertex / addEdge
  (optimizable_graph.cpp:179-246, 344-521) on a file that exercises each one;
* closed forms for the pose conversions: VERTEX_SE3:EXPMAP holds the inverse of
  the world-to-camera estimate (vertex_se3_expmap.cpp:35-44), checked with
  scipy rotation matrices; EDGE_SE3:QUAT normalises its quaternion
  (edge_se3.cpp:43-53);
* round trips at precision 17, independence from the parser thread count;
* BAL: chi2 of the BAL model (numpy restatement of EdgeObservationBAL,
  bal_example.cpp:130-205) equals the oracle's EdgeSE3ProjectXYZ chi2 of the
  mapped problem when k1 = k2 = 0, and the undistorted mapping of noise-free
  data has zero chi2.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from g2o_amd import datasets
from g2o_amd import graphio as gio
from oracle.oracle import OracleBA

VALUES = [0.1, 1.5, -0.0, 1e-5, 2.5e-7, 123456789.0, 100000.0, 1e6, 1e300, 5e-324, np.pi, -2.0 / 3.0, 7.0, 1e-4,
          0.00012345678, 9.9999995, 99999.95, 1.0 / 3.0]


def _ostream_lines(tmp_path, values, precision):
    """Each value as an ostream prints it (default float format), "v " per
    value, the way writeVector does."""
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = tmp_path / "fmt.cpp"
    src.write_text(
        "#include <iostream>\n#include <iomanip>\n#include <cstdlib>\n"
        "int main(int argc, char** argv) {\n"
        "  std::cout << std::setprecision(std::atoi(argv[1]));\n"
        "  for (int i = 2; i < argc; ++i) std::cout << std::strtod(argv[i], nullptr) << ' ';\n"
        "  std::cout << '\\n';\n}\n")
    exe = tmp_path / "fmt"
    subprocess.run(["g++", "-O1", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe), str(precision)] + [repr(float(v)) for v in values], check=True,
                         capture_output=True, text=True).stdout
    return out


@pytest.mark.parametrize("precision", [6, 17])
def test_number_format_matches_ostream(tmp_path, precision):
    vals = np.array(VALUES, np.float64)
    n = len(vals) // 3 * 3
    pts = vals[:n].reshape(-1, 3)
    P = pts.shape[0]
    # one camera observing every point, so that every point is saved
    prob = datasets.BAProblem(np.array([[0, 0, 0, 0, 0, 0, 1.0]]), np.array([[1.0, 1, 0, 0]]), pts,
                              np.zeros(P, np.int32), np.arange(P, dtype=np.int32), np.zeros((P, 2)), np.ones(P),
                              np.zeros(1, np.uint8), np.zeros(P, np.uint8), 0.0)
    path = tmp_path / "pts.g2o"
    gio.save_ba_problem(str(path), prob, precision=precision)
    lines = [ln for ln in path.read_text().splitlines() if ln.startswith("VERTEX_TRACKXYZ")]
    got = "".join(ln.split(" ", 2)[2] for ln in lines)  # each value is followed by a space
    want = _ostream_lines(tmp_path, pts.reshape(-1), precision).rstrip("\n")
    assert got == want


RULES = """# header comment

VERTEX_SE3:EXPMAP 0 1 2 3 0 0 0 1
FIX 1
VERTEX_SE3:EXPMAP 1 0.5 -1 2 0.1 0.2 0.3 -0.9
FIX 0 7
VERTEX_TRACKXYZ 5 1 2 10
VERTEX_TRACKXYZ 5 9 9 9
VERTEX_TRACKXYZ -3 1 1 1
VERTEX_TRACKXYZ 6 1 2
EDGE_SE3_PROJECT_XYZ:EXPMAP 5 0 100 200 1 0 1
EDGE_SE3_PROJECT_XYZ:EXPMAP 5 1 110 210 2 0 2 trailing tokens are ignored
EDGE_SE3_PROJECT_XYZ:EXPMAP 0 5 1 2 1 0 1
EDGE_SE3_PROJECT_XYZ:EXPMAP 5 9 1 2 1 0 1
EDGE_SE3_PROJECT_XYZ:EXPMAP 5 1 1 2 1 0
VERTEX_XY 3 1 2
VERTEX_XY 4 1 2
VERTEX_SE3:QUAT 10 0 0 0 0 0 0 1
VERTEX_SE3:QUAT 11 1 0 0 0 0 0 1
EDGE_SE3:QUAT 10 11 1 0 0 0 0 0 2 2 0 0 0 0 0 1 0 0 0 0 1 0 0 0 1 0 0 1 0 1
EDGE_SE3:QUAT 10 10 1 0 0 0 0 0 1 1 0 0 0 0 0 1 0 0 0 0 1 0 0 0 1 0 0 1 0 1
"""


def test_load_rules(tmp_path):
    path = tmp_path / "rules.g2o"
    path.write_text(RULES)
    g = gio.G2oGraph.load(str(path))
    n = g.info()
    assert n["lines"] == RULES.count("\n")
    assert (n["cameras"], n["points"], n["projection_edges"], n["poses"], n["pose_edges"]) == (2, 1, 2, 2, 1)
    # "FIX 1" precedes vertex 1: not found, so only vertex 0 is fixed
    assert n["fixed"] == 1
    assert n["unknown_lines"] == 2
    # duplicate 5, negative -3, short 6, wrong vertex order, missing vertex 9,
    # short edge, edge from a vertex to itself
    assert n["skipped"] == 7
    log = g.log.splitlines()
    assert "Unable to fix vertex with id 1. Not found in the graph." in log
    assert "Unable to fix vertex with id 7. Not found in the graph." in log
    assert "a vertex with ID 5 has already been registered with this graph" in log
    assert "FATAL, a vertex with (negative) ID -3 cannot be inserted into the graph" in log
    assert "Error reading vertex VERTEX_TRACKXYZ 6 at line 10" in log
    assert "Unable to add edge EDGE_SE3_PROJECT_XYZ:EXPMAP at line 13 IDs: 0 5" in log
    assert "Unable to find vertices for edge EDGE_SE3_PROJECT_XYZ:EXPMAP at line 14 IDs: 5 9" in log
    assert "Unable to add edge EDGE_SE3_PROJECT_XYZ:EXPMAP at line 15 IDs: 5 1" in log
    assert "Unable to add edge EDGE_SE3:QUAT at line 21 IDs: 10 10" in log
    assert sum(1 for ln in log if ln.startswith("Unknown type VERTEX_XY")) == 1
    a = g.ba_arrays()
    assert a["cam_id"].tolist() == [0, 1] and a["cam_fixed"].tolist() == [1, 0]
    assert a["pt_id"].tolist() == [5] and a["pts"].tolist() == [[1.0, 2.0, 10.0]]  # the first vertex 5 wins
    assert a["edge_cam"].tolist() == [0, 1] and a["edge_pt"].tolist() == [0, 0]
    assert a["meas"].tolist() == [[100.0, 200.0], [110.0, 210.0]]
    assert a["omega"].tolist() == [[1.0, 0.0, 1.0], [2.0, 0.0, 2.0]]
    pg = g.pg_arrays()
    assert pg["pose_id"].tolist() == [10, 11]
    assert pg["edge_from"].tolist() == [0] and pg["edge_to"].tolist() == [1]
    # the quaternion (0 0 0 2) is normalised; the information upper triangle is mirrored
    assert pg["meas"][0].tolist() == [1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0]
    I = np.eye(6)
    I[0, 0] = 2.0
    assert np.array_equal(pg["info"][0], I)


def test_camera_vertex_holds_the_inverse_pose(tmp_path):
    rng = np.random.default_rng(5)
    q = rng.normal(size=(20, 4))
    q[::3, 3] = -np.abs(q[::3, 3])  # some with w < 0
    t = rng.normal(size=(20, 3)) * 10
    path = tmp_path / "cams.g2o"
    with open(path, "w") as f:
        for i in range(20):
            f.write("VERTEX_SE3:EXPMAP %d %s\n" % (i, " ".join(repr(float(x)) for x in np.r_[t[i], q[i]])))
    a = gio.G2oGraph.load(str(path)).ba_arrays()
    for i in range(20):
        R_c2w = Rotation.from_quat(q[i]).as_matrix()  # scipy normalises
        R_w2c = R_c2w.T
        got = a["cams"][i]
        assert got[6] >= 0.0
        assert abs(np.linalg.norm(got[3:]) - 1.0) < 1e-15
        assert np.allclose(Rotation.from_quat(got[3:]).as_matrix(), R_w2c, atol=1e-14)
        assert np.allclose(got[:3], -R_w2c @ t[i], rtol=1e-14, atol=1e-13)


def test_ba_round_trip(tmp_path):
    p = datasets.make_small_ba(seed=1000)
    path = str(tmp_path / "ba.g2o")
    gio.save_ba_problem(path, p)
    g = gio.G2oGraph.load(path)
    q = g.to_ba_problem(p.intr)
    C = p.num_cams
    assert q.info["cam_id"].tolist() == list(range(C))
    assert q.info["pt_id"].tolist() == list(range(C, C + p.num_pts))
    assert np.allclose(q.cams, p.cams, rtol=0, atol=1e-14)
    for k in ("pts", "edge_cam", "edge_pt", "meas", "weight", "cam_fixed", "pt_fixed"):
        assert np.array_equal(getattr(q, k), getattr(p, k)), k
    # the FIX lines follow the fixed vertices
    lines = open(path).read().splitlines()
    for c in np.flatnonzero(p.cam_fixed):
        i = lines.index(next(ln for ln in lines if ln.startswith(f"VERTEX_SE3:EXPMAP {c} ")))
        assert lines[i + 1] == f"FIX {c}"
    # save after a state change: set_ba_state / save / load
    cams2 = p.cams.copy()
    cams2[:, :3] += 0.25
    g.set_ba_state(cams2, p.pts * 2)
    g.save(path)
    r = gio.G2oGraph.load(path).to_ba_problem(p.intr)
    assert np.allclose(r.cams, cams2, rtol=0, atol=1e-13)
    assert np.array_equal(r.pts, p.pts * 2)


def test_full_information_round_trip(tmp_path):
    """EdgeSE3ProjectXYZ::read/write take the upper triangle of the 2x2
    information (edge_project_xyz.cpp:34-42): a problem with per-edge full
    information saves weight * omega and loads it back as omega (weights 1)."""
    p = datasets.make_small_ba(seed=1000)
    p.omega = datasets.random_information(p.num_edges, seed=3)
    path = str(tmp_path / "ba_info.g2o")
    gio.save_ba_problem(path, p)
    q = gio.G2oGraph.load(path).to_ba_problem(p.intr)
    assert np.array_equal(q.weight, np.ones(p.num_edges))
    np.testing.assert_allclose(q.omega, p.omega * p.weight[:, None], rtol=1e-15)
    # scalar information stays a weight
    p2 = datasets.make_small_ba(seed=1000)
    p2.weight = np.linspace(0.5, 2.0, p2.num_edges)
    gio.save_ba_problem(path, p2)
    q2 = gio.G2oGraph.load(path).to_ba_problem(p2.intr)
    assert q2.omega is None and np.array_equal(q2.weight, p2.weight)


def test_cpp_driver_rejects_a_camera_with_two_parameter_sets(tmp_path):
    """The solver holds intrinsics per camera while g2o resolves
    EDGE_PROJECT_XYZ2UV's camera parameters per edge: a camera observed through
    two different parameter sets is refused before anything runs."""
    driver = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "g2o_amd", "host",
                          "lm_drive")
    if not os.path.exists(driver):
        pytest.skip("driver not built")
    p = datasets.make_small_ba(seed=1000)
    plain = tmp_path / "plain.g2o"
    gio.save_ba_problem(str(plain), p)
    lines = ["PARAMS_CAMERAPARAMETERS 0 500 320 240 0.5", "PARAMS_CAMERAPARAMETERS 1 510 320 240 0.5"]
    n = 0
    for ln in plain.read_text().splitlines():
        if ln.startswith("EDGE_SE3_PROJECT_XYZ:EXPMAP "):
            t = ln.split(" ", 3)
            ln = f"EDGE_PROJECT_XYZ2UV:EXPMAP {t[1]} {t[2]} {n % 2} {t[3]}"
            n += 1
        lines.append(ln)
    src = tmp_path / "two_params.g2o"
    src.write_text("\n".join(lines) + "\n")
    out = subprocess.run([driver, str(src), "optimize", "2", "0"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 1
    assert "two different camera parameter sets" in out.stderr


def test_pose_graph_round_trip(tmp_path):
    pg = datasets.make_sphere_pose_graph(nodes_per_level=40, laps=3)
    path = str(tmp_path / "pg.g2o")
    gio.save_pose_graph(path, pg)
    r = gio.G2oGraph.load(path).to_pose_graph()
    assert np.array_equal(r.poses[:, :3], pg.poses[:, :3])
    R0 = Rotation.from_quat(pg.poses[:, 3:]).as_matrix()
    R1 = Rotation.from_quat(r.poses[:, 3:]).as_matrix()
    assert np.max(np.abs(R0 - R1)) < 1e-14
    for k in ("edge_from", "edge_to", "info", "fixed"):
        assert np.array_equal(getattr(r, k), getattr(pg, k)), k
    M0 = Rotation.from_quat(pg.meas[:, 3:]).as_matrix()
    M1 = Rotation.from_quat(r.meas[:, 3:]).as_matrix()
    assert np.max(np.abs(M0 - M1)) < 1e-14
    assert np.array_equal(r.meas[:, :3], pg.meas[:, :3])


def test_saved_quaternion_sign_follows_the_rotation_matrix(tmp_path):
    """toVectorQT (isometry3d_mappings.cpp:111-121) prints the quaternion of the
    rotation matrix: w > 0 if trace(R) > 0, else the component of the largest
    diagonal entry is positive."""
    rng = np.random.default_rng(3)
    q = rng.normal(size=(400, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    poses = np.concatenate([rng.normal(size=(400, 3)), q], axis=1)
    prob = datasets.PoseGraphProblem(poses, np.arange(399, dtype=np.int32), np.arange(1, 400, dtype=np.int32),
                                     np.tile([0, 0, 0, 0, 0, 0, 1.0], (399, 1)), np.tile(np.eye(6), (399, 1, 1)),
                                     np.zeros(400, np.uint8))
    path = str(tmp_path / "sign.g2o")
    gio.save_pose_graph(path, prob)
    r = gio.G2oGraph.load(path).to_pose_graph()
    for i in range(400):
        R = Rotation.from_quat(q[i]).as_matrix()
        s = r.poses[i, 3:]
        assert min(np.max(np.abs(s - q[i])), np.max(np.abs(s + q[i]))) < 1e-15
        if np.trace(R) > 0:
            assert s[3] > 0
        else:
            assert s[int(np.argmax(np.diag(R)))] > 0


def test_parse_is_independent_of_thread_count(tmp_path):
    p = datasets.make_walk_ba(seed=7, num_cams=60, num_pts=30000)
    path = str(tmp_path / "walk.g2o")
    gio.save_ba_problem(path, p)
    with open(path, "a") as f:  # a few rejected lines at the end of the file
        f.write("VERTEX_TRACKXYZ 3 0 0 0\nEDGE_SE3_PROJECT_XYZ:EXPMAP 99999999 0 1 1 1 0 1\nFOO 1\n")
    assert os.path.getsize(path) > 4 << 20  # several parser pieces
    a = gio.G2oGraph.load(path, nthreads=1)
    b = gio.G2oGraph.load(path, nthreads=8)
    assert a.info() == b.info() and a.log == b.log
    x, y = a.ba_arrays(), b.ba_arrays()
    for k in x:
        assert np.array_equal(x[k], y[k]), k
    assert a.info()["skipped"] == 2 and a.info()["unknown_lines"] == 1


def test_errors(tmp_path):
    with pytest.raises(gio.G2oIOError, match="Unable to open"):
        gio.G2oGraph.load(str(tmp_path / "missing.g2o"))
    g = gio.G2oGraph()
    with pytest.raises(gio.G2oIOError, match="precision"):
        g.save(str(tmp_path / "x.g2o"), precision=0)
    path = tmp_path / "aniso.g2o"
    path.write_text("VERTEX_SE3:EXPMAP 0 0 0 0 0 0 0 1\nVERTEX_TRACKXYZ 1 0 0 5\n"
                    "EDGE_SE3_PROJECT_XYZ:EXPMAP 1 0 1 2 1 0.5 1\n")
    # an anisotropic information matrix is kept as the edge's full Omega
    aniso = gio.G2oGraph.load(str(path)).to_ba_problem([500, 500, 320, 240])
    assert aniso.omega.tolist() == [[1.0, 0.5, 1.0]] and aniso.weight.tolist() == [1.0]
    p = datasets.make_small_ba(seed=1000)
    g = gio.G2oGraph()
    g.add_ba(p)
    with pytest.raises(gio.G2oIOError, match="duplicate"):
        g.add_ba(p, cam_id=np.arange(p.num_cams))
    assert g.info()["cameras"] == p.num_cams  # the failed call added nothing
    bad = tmp_path / "bad.bal"
    bad.write_text("1 1 2\n0 0 1 2\n0 1 1 2\n" + "0\n" * 12)
    with pytest.raises(gio.G2oIOError, match="out of range"):
        gio.load_bal(str(bad))
    bad.write_text("1 1 1\n0 0 1 2\n" + "0\n" * 11)
    with pytest.raises(gio.G2oIOError, match="truncated"):
        gio.load_bal(str(bad))


def test_bal_load_and_exact_mapping(tmp_path):
    bal = datasets.make_bal(11)
    path = str(tmp_path / "p.txt")
    gio.save_bal(path, bal)
    b = gio.load_bal(path)
    for k in ("cams", "pts", "edge_cam", "edge_pt", "meas"):
        assert np.array_equal(getattr(b, k), getattr(bal, k)), k
    # angle-axis -> quaternion
    q = gio.angle_axis_to_quat(bal.cams[:, :3])
    assert np.allclose(Rotation.from_quat(q).as_matrix(), Rotation.from_rotvec(bal.cams[:, :3]).as_matrix(),
                       atol=1e-15)
    prob = gio.bal_to_ba(b)
    want = gio.bal_chi2(b)
    got = OracleBA(prob).chi2()
    assert want > 0 and abs(got - want) <= 1e-12 * want


def test_bal_undistorted_mapping_of_exact_data(tmp_path):
    bal = datasets.make_bal(12, k=(-0.3, 0.08), noise=0.0)
    path = str(tmp_path / "d.txt")
    gio.save_bal(path, bal)
    b = gio.load_bal(path)
    assert gio.bal_chi2(b) < 1e-18 * len(b.meas)
    with pytest.raises(gio.G2oIOError, match="distortion"):
        gio.bal_to_ba(b)
    prob = gio.bal_to_ba(b, undistort=True)
    # every undistorted measurement is the pinhole projection of its point
    assert OracleBA(prob).chi2() < 1e-16 * len(b.meas) * 500.0 ** 2


def test_whitespace_signs_and_exponents(tmp_path):
    """CRLF line ends, tabs, leading '+' and exponents parse as an ostream
    reads them (std::istream >> double accepts all of these)."""
    path = tmp_path / "ws.g2o"
    path.write_bytes(b"VERTEX_TRACKXYZ\t3\t+1.5e0  -2E-1 3.\r\n"
                     b"  VERTEX_SE3:EXPMAP 7 0 0 0 0 0 0 +1\r\n"
                     b"EDGE_SE3_PROJECT_XYZ:EXPMAP 3 7 1e2 2.5e+2 4 0 4\r\n"
                     b"\r\n# trailing comment\r\n")
    g = gio.G2oGraph.load(str(path))
    n = g.info()
    assert (n["points"], n["cameras"], n["projection_edges"], n["skipped"], n["lines"]) == (1, 1, 1, 0, 5)
    a = g.ba_arrays()
    assert a["pts"].tolist() == [[1.5, -0.2, 3.0]]
    assert a["meas"].tolist() == [[100.0, 250.0]] and a["omega"].tolist() == [[4.0, 0.0, 4.0]]


def xyz2uv_text(prob, tmp_path, f=500.0, cx=320.0, cy=240.0):
    """The problem as g2o's ba_demo writes it: one PARAMS_CAMERAPARAMETERS and
    EDGE_PROJECT_XYZ2UV:EXPMAP edges naming it (edge_project_xyz2uv.cpp:37-47)."""
    path = tmp_path / "plain.g2o"
    gio.save_ba_problem(str(path), prob)
    out = [f"PARAMS_CAMERAPARAMETERS 0 {f!r} {cx!r} {cy!r} 0.5"]
    for ln in path.read_text().splitlines():
        if ln.startswith("EDGE_SE3_PROJECT_XYZ:EXPMAP "):
            t = ln.split(" ", 3)
            ln = f"EDGE_PROJECT_XYZ2UV:EXPMAP {t[1]} {t[2]} 0 {t[3]}"
        out.append(ln)
    return "\n".join(out) + "\n"


def test_camera_parameters_and_xyz2uv_edges(tmp_path):
    """EdgeProjectXYZ2UV is EdgeSE3ProjectXYZ with fx = fy = f and (cx, cy)
    taken from its CameraParameters (cam_map, parameter_cameraparameters.cpp:
    59-65; the Jacobians of edge_project_xyz2uv.cpp:58-97 are those of
    edge_project_xyz.cpp:57-93 with fx = fy = f), so a ba_demo-style file runs
    on the path with the intrinsics the file holds."""
    p = datasets.make_small_ba(seed=1000)
    path = tmp_path / "xyz2uv.g2o"
    path.write_text(xyz2uv_text(p, tmp_path))
    g = gio.G2oGraph.load(str(path))
    n = g.info()
    assert n["xyz2uv_edges"] == n["projection_edges"] == p.num_edges and n["camera_parameters"] == 1
    assert n["skipped"] == 0
    q = g.to_ba_problem()
    assert np.array_equal(q.intr, np.tile([500.0, 500.0, 320.0, 240.0], (p.num_cams, 1)))
    for k in ("pts", "edge_cam", "edge_pt", "meas", "weight", "cam_fixed"):
        assert np.array_equal(getattr(q, k), getattr(p, k)), k
    assert abs(OracleBA(q).chi2() - OracleBA(p).chi2()) <= 1e-12 * OracleBA(p).chi2()
    # saved back: the parameter first, the edges keep their tag and parameter id
    out = tmp_path / "xyz2uv_out.g2o"
    g.save(str(out), precision=17)
    lines = out.read_text().splitlines()
    assert lines[0] == "PARAMS_CAMERAPARAMETERS 0 500 320 240 0.5 "
    assert sum(ln.startswith("EDGE_PROJECT_XYZ2UV:EXPMAP ") for ln in lines) == p.num_edges
    r = gio.G2oGraph.load(str(out)).to_ba_problem()
    assert np.array_equal(r.intr, q.intr) and np.array_equal(r.meas, q.meas)


def test_camera_parameter_rules(tmp_path):
    path = tmp_path / "prules.g2o"
    path.write_text("VERTEX_SE3:EXPMAP 0 0 0 0 0 0 0 1\n"
                    "VERTEX_TRACKXYZ 1 0 0 5\n"
                    "EDGE_PROJECT_XYZ2UV:EXPMAP 1 0 3 1 2 1 0 1\n"   # parameter 3 not defined yet
                    "PARAMS_CAMERAPARAMETERS 3 400 300 200 0.1\n"
                    "PARAMS_CAMERAPARAMETERS 3 999 0 0 0\n"         # duplicate id
                    "EDGE_PROJECT_XYZ2UV:EXPMAP 1 0 3 1 2 1 0 1\n")
    g = gio.G2oGraph.load(str(path))
    n = g.info()
    assert (n["projection_edges"], n["camera_parameters"], n["skipped"]) == (1, 1, 2)
    log = g.log.splitlines()
    assert "Unable to add edge EDGE_PROJECT_XYZ2UV:EXPMAP at line 3 IDs: 1 0" in log
    assert "Parameter of type: PARAMS_CAMERAPARAMETERS id: 3 already defined at line 5" in log
    assert g.to_ba_problem().intr.tolist() == [[400.0, 400.0, 300.0, 200.0]]
