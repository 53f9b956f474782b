"""The one-launch PCG iteration (k_pcg_fused, the default for systems of up
to 2,048 block rows) against the two-launch iteration (direction + product,
then update), both on the device and selected per context through
G2OHIP_PCG_FUSED at creation. In exact arithmetic the two are the same
block-Jacobi PCG (linear_solver_pcg.hpp:112-155); in floating point the
one-launch form takes r.M^-1 r from a recurrence instead of a direct sum, so
the stop iteration may move by one and the solutions agree to the PCG
tolerance, not bit for bit."""
import numpy as np
import pytest

from g2o_amd import datasets

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _pair(monkeypatch, make):
    out = []
    for flag in ("1", "0"):
        monkeypatch.setenv("G2OHIP_PCG_FUSED", flag)
        out.append(make())
    monkeypatch.delenv("G2OHIP_PCG_FUSED")
    return out


# lambda a thousandth of the LM's first one: the walk systems then take tens
# of PCG iterations instead of a handful
LAM = 1e-3


def _solve(opt, lam_scale=1.0):
    opt.build_system()
    lam = lam_scale * opt.lambda_init()
    opt.set_lambda(lam)
    assert opt.solve()
    return opt.x(), int(opt.last_solve_info()[0])


def _compare(xa, ia, xb, ib, x_ref, count_slack=1):
    """Iteration counts within count_slack; the two solutions no further
    apart than the PCG tolerance leaves either from the exact step."""
    ea = np.linalg.norm(xa - x_ref) / np.linalg.norm(x_ref)
    eb = np.linalg.norm(xb - x_ref) / np.linalg.norm(x_ref)
    dab = np.linalg.norm(xa - xb) / np.linalg.norm(x_ref)
    _say(f"PCG iterations one-launch {ia}, two-launch {ib}; error vs exact step {ea:.2e} / {eb:.2e}, "
         f"apart {dab:.2e}")
    assert ia >= 8 and abs(ia - ib) <= count_slack
    assert ea <= 3.0 * eb + 1e-9
    assert dab <= 3.0 * eb + 1e-9


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-6), ("fp64", 1e-12), ("fp32", 1e-6)])
def test_ba_one_launch_equals_two_launch(monkeypatch, precision, tol):
    import g2o_amd
    prob = datasets.make_walk_ba(seed=7, num_cams=60, num_pts=6000)
    prec = g2o_amd.FP64 if precision == "fp64" else g2o_amd.FP32
    _say(f"BA walk 60 cams, {precision}, tol {tol}")
    a, b = _pair(monkeypatch, lambda: g2o_amd.BundleAdjuster(prob, device=0, precision=prec, pcg_tolerance=tol))
    xa, ia = _solve(a, LAM)
    xb, ib = _solve(b, LAM)
    x_ref, _ = _solve(g2o_amd.BundleAdjuster(prob, device=0, pcg_tolerance=1e-24), LAM)
    # the FP32 count sums the PCG runs of each refinement step
    _compare(xa, ia, xb, ib, x_ref, 1 if precision == "fp64" else 2 + ib // 50)


def test_ba_one_launch_lm_trace(monkeypatch):
    """Ten LM iterations of the FP32 path: equal trial counts, chi2 per
    iteration within 1e-6."""
    import g2o_amd
    prob = datasets.make_walk_ba(seed=8, num_cams=40, num_pts=3000)
    a, b = _pair(monkeypatch, lambda: g2o_amd.BundleAdjuster(prob, device=0, precision=g2o_amd.FP32))
    sa, sb = a.optimize(10), b.optimize(10)
    assert len(sa) == len(sb)
    for x, y in zip(sa, sb):
        assert x["levenberg_iterations"] == y["levenberg_iterations"]
        assert abs(x["chi2"] - y["chi2"]) <= 1e-6 * y["chi2"]


@pytest.mark.parametrize("tol", [1e-6, 1e-12])
def test_bal_one_launch_equals_two_launch(monkeypatch, tol):
    import g2o_amd
    prob = datasets.make_walk_bal(seed=41, num_cams=60, num_pts=4000)
    a, b = _pair(monkeypatch, lambda: g2o_amd.BALAdjuster(prob, device=0, pcg_tolerance=tol))
    xa, ia = _solve(a, LAM)
    xb, ib = _solve(b, LAM)
    x_ref, _ = _solve(g2o_amd.BALAdjuster(prob, device=0, pcg_tolerance=1e-24), LAM)
    _compare(xa, ia, xb, ib, x_ref)


@pytest.mark.parametrize("tol", [1e-8, 1e-16])
def test_posegraph_one_launch_equals_two_launch(monkeypatch, tol):
    import g2o_amd
    prob = datasets.make_sphere_pose_graph(seed=4, nodes_per_level=50, laps=4)
    a, b = _pair(monkeypatch, lambda: g2o_amd.PoseGraphOptimizer(prob, device=0, pcg_tolerance=tol))
    xa, ia = _solve(a)
    xb, ib = _solve(b)
    x_ref, _ = _solve(g2o_amd.PoseGraphOptimizer(prob, device=0, pcg_tolerance=1e-28))
    _compare(xa, ia, xb, ib, x_ref)


@pytest.mark.parametrize("nf", [2047, 2048, 2049])
def test_posegraph_at_the_one_launch_size_limit(monkeypatch, nf):
    """The one-launch iteration serves up to 2,048 block rows (256 workgroups
    of 8 rows, whose partials each lane of the next launch loads four at a
    time); at 2,049 both settings run the two-launch iteration and agree bit
    for bit."""
    import g2o_amd
    prob = datasets.make_sphere_pose_graph(seed=6, nodes_per_level=70, laps=30)
    assert prob.num_poses == 2100
    fixed = np.zeros(prob.num_poses, np.uint8)
    fixed[:prob.num_poses - nf] = 1
    prob.fixed = fixed
    a, b = _pair(monkeypatch, lambda: g2o_amd.PoseGraphOptimizer(prob, device=0, pcg_tolerance=1e-10))
    assert a.Nf == nf
    xa, ia = _solve(a)
    xb, ib = _solve(b)
    if nf > 2048:
        _say(f"Nf {nf}: PCG iterations {ia} / {ib}")
        assert ia == ib and np.array_equal(xa, xb)
        return
    x_ref, _ = _solve(g2o_amd.PoseGraphOptimizer(prob, device=0, pcg_tolerance=1e-28))
    _compare(xa, ia, xb, ib, x_ref)


@pytest.mark.parametrize("kind", ["ba", "bal"])
def test_one_free_camera(monkeypatch, kind, oracle_lib):
    """One free camera: one workgroup, and block-Jacobi PCG on a single
    block is exact after one iteration, as in the oracle."""
    import g2o_amd
    if kind == "ba":
        prob = datasets.make_small_ba(seed=12, num_cams=2, num_pts=60)
        prob.cam_fixed = np.array([1, 0], np.uint8)
        ora = oracle_lib.OracleBA(prob)
        mk = lambda: g2o_amd.BundleAdjuster(prob, device=0, pcg_tolerance=1e-10)
    else:
        prob = datasets.make_bal(seed=12, num_cams=2, num_pts=60, k=(0.05, -0.01))
        prob.cam_fixed = np.array([1, 0], np.uint8)
        ora = oracle_lib.OracleBAL(prob)
        mk = lambda: g2o_amd.BALAdjuster(prob, device=0, pcg_tolerance=1e-10)
    a, b = _pair(monkeypatch, mk)
    a.build_system()
    lam = a.lambda_init()
    xs = []
    for opt in (a, b):
        if opt is b:
            opt.build_system()
        opt.set_lambda(lam)
        assert opt.solve()
        xs.append((opt.x(), int(opt.last_solve_info()[0])))
    ora.linearize()
    st, x_o, _ = ora.solve(lam, solver=0)
    st1, _, it_o = ora.solve(lam, solver=1, tol=1e-10)
    assert st == 0 and st1 == 0
    _say(f"{kind}: PCG iterations one-launch {xs[0][1]}, two-launch {xs[1][1]}, oracle {it_o}")
    for x, it in xs:
        assert it == it_o == 1
        assert np.linalg.norm(x - x_o) <= 1e-9 * np.linalg.norm(x_o)
