"""GPU parity of the BAL path (9-DoF cameras, BlockSolver<9, 3>,
bal_example.cpp) against the FP64 CPU oracle, through the C ABI.

Tolerances (FP64 throughout; only the summation order and the closed-form
against dual-number Jacobians differ from the oracle): chi2 and errors
relative 1e-11, Hessian blocks and b relative 1e-10 of their largest entry,
Schur blocks 1e-9, the solve 1e-8 against the oracle's dense Cholesky, LM
chi2 traces 1e-8 per iteration with equal trial counts.
"""
import numpy as np
import pytest

from g2o_amd import datasets

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _perturbed_bal(seed=21, cams=6, pts=80, k=(0.08, -0.02), huber=0.0, fixed=True, weights=True):
    bal = datasets.make_bal(seed=seed, num_cams=cams, num_pts=pts, k=k, noise=0.8)
    rng = np.random.default_rng(seed + 1)
    bal.cams = bal.cams + np.c_[rng.normal(0, 0.01, (cams, 3)), rng.normal(0, 0.05, (cams, 3)),
                                rng.normal(0, 2.0, cams), rng.normal(0, 0.01, cams), rng.normal(0, 0.002, cams)]
    bal.pts = bal.pts + rng.normal(0, 0.05, bal.pts.shape)
    if weights:
        bal.weight = rng.uniform(0.5, 2.0, bal.num_edges)
    bal.huber = huber
    if huber > 0:
        bal.meas[::23] += 30.0  # outliers on the Huber branch
    if fixed:
        bal.cam_fixed = np.zeros(cams, np.uint8)
        bal.cam_fixed[0] = 1
        bal.pt_fixed = np.zeros(pts, np.uint8)
        bal.pt_fixed[[3, 11]] = 1
    return bal


@pytest.fixture(scope="module")
def small():
    return _perturbed_bal()


@pytest.fixture(scope="module")
def small_huber():
    return _perturbed_bal(seed=31, huber=2.0)


@pytest.fixture(scope="module")
def walk():
    return datasets.make_walk_bal(seed=41, num_cams=60, num_pts=4000)


@pytest.mark.parametrize("case", ["small", "small_huber", "walk"])
def test_bal_chi2_and_hessian(case, request, oracle_lib):
    import g2o_amd
    prob = request.getfixturevalue(case)
    _say(f"case {case}: C={prob.num_cams} P={prob.num_pts} E={prob.num_edges}")
    ba = g2o_amd.BALAdjuster(prob, device=0)
    ora = oracle_lib.OracleBAL(prob)
    chi_o, Hpp_o, Hll_o, Hpl_o, b_o = ora.linearize()
    chi_g = ba.compute_active_errors()
    assert abs(chi_g - chi_o) <= 1e-11 * abs(chi_o)
    assert _rel(ba.errors(), ora.errors()) < 1e-11
    ba.build_system()
    Hpp, Hll, Hpl = ba.hessian()
    assert _rel(Hpp, Hpp_o) < 1e-10
    assert _rel(Hll, Hll_o) < 1e-10
    assert _rel(Hpl, Hpl_o) < 1e-10
    assert _rel(ba.b(), b_o) < 1e-10


@pytest.mark.parametrize("case", ["small", "small_huber", "walk"])
def test_bal_schur_and_solve(case, request, oracle_lib):
    import g2o_amd
    prob = request.getfixturevalue(case)
    _say(f"case {case}: C={prob.num_cams} P={prob.num_pts} E={prob.num_edges}")
    ba = g2o_amd.BALAdjuster(prob, device=0, pcg_tolerance=1e-20)
    ora = oracle_lib.OracleBAL(prob)
    ora.linearize()
    ba.build_system()
    lam = 5.0 * ba.lambda_init()  # 5e-5 of the largest diagonal entry
    st, rows_o, cols_o, blocks_o, bs_o = ora.schur(lam)
    assert st == 0
    ba.set_lambda(lam)
    assert ba.solve()
    rows, cols, blocks, bs = ba.schur()
    assert np.array_equal(rows, rows_o) and np.array_equal(cols, cols_o)
    assert _rel(blocks, blocks_o) < 1e-9
    assert _rel(bs, bs_o) < 1e-9
    st, x_o, _ = ora.solve(lam, solver=0)
    assert st == 0
    x = ba.x()
    assert np.linalg.norm(x - x_o) <= 1e-8 * np.linalg.norm(x_o)


def test_bal_pcg_iterations_follow_the_reference_rule(walk, oracle_lib):
    import g2o_amd
    ba = g2o_amd.BALAdjuster(walk, device=0, pcg_tolerance=1e-6)
    ora = oracle_lib.OracleBAL(walk)
    ora.linearize()
    ba.build_system()
    lam = ba.lambda_init()
    ba.set_lambda(lam)
    assert ba.solve()
    it_g = int(ba.last_solve_info()[0])
    st, x_o, it_o = ora.solve(lam, solver=1, tol=1e-6)
    assert st == 0
    _say(f"PCG iterations: device {it_g}, oracle {it_o}")
    assert abs(it_g - it_o) <= 1
    assert np.linalg.norm(ba.x() - x_o) <= 1e-5 * np.linalg.norm(x_o)


@pytest.mark.parametrize("case", ["small_huber", "walk"])
def test_bal_lm_trace_matches_oracle(case, request, oracle_lib):
    import g2o_amd
    prob = request.getfixturevalue(case)
    iters = 10
    ba = g2o_amd.BALAdjuster(prob, device=0, pcg_tolerance=1e-10)
    stats = ba.optimize(iters)
    ora = oracle_lib.OracleBAL(prob)
    n, trace, _ = ora.optimize(iters, solver=1, tol=1e-10)
    _say(f"case {case}: chi2 {trace[0, 0]:.6e} -> device {stats[-1]['chi2']:.9e}, oracle {trace[n - 1, 1]:.9e}")
    assert len(stats) == n
    for s, t in zip(stats, trace):
        assert abs(s["chi2"] - t[1]) <= 1e-8 * t[1], (s["chi2"], t[1])
        assert s["levenberg_iterations"] == int(t[3])
    cams, pts = ba.state()
    cams_o, pts_o = ora.state()
    assert _rel(cams, cams_o) < 1e-6
    # the problem converges: the final chi2 is near its statistical value
    if case == "walk":
        ndof = 2 * prob.num_edges - (9 * prob.num_cams + 3 * prob.num_pts)
        assert stats[-1]["chi2"] < 1.2 * ndof


def test_bal_gauss_newton_matches_lm_step_at_zero_lambda(oracle_lib):
    import g2o_amd
    # every free point seen by at least two cameras: the undamped system is
    # positive definite
    small = _perturbed_bal(cams=8, pts=200)
    ba = g2o_amd.BALAdjuster(small, device=0, pcg_tolerance=1e-20)
    ba.build_system()
    ba.set_lambda(0.0)
    assert ba.solve()
    x0 = ba.x()
    ora = oracle_lib.OracleBAL(small)
    ora.linearize()
    st, x_o, _ = ora.solve(0.0, solver=0)
    assert st == 0
    assert np.linalg.norm(x0 - x_o) <= 1e-7 * np.linalg.norm(x_o)
    ba2 = g2o_amd.BALAdjuster(small, device=0, pcg_tolerance=1e-20)
    ba2.set_algorithm(g2o_amd.GAUSS_NEWTON)
    st2 = ba2.optimize(1)
    assert st2[0]["status"] == 0
    cams, _ = ba2.state()
    cf = small.cam_fixed.astype(bool)
    assert np.allclose(cams[~cf], small.cams[~cf] + x0[:9 * (~cf).sum()].reshape(-1, 9), rtol=0, atol=1e-9)


def test_bal_file_round_trip_runs_like_the_problem(tmp_path, oracle_lib):
    import g2o_amd
    from g2o_amd import graphio
    prob = datasets.make_walk_bal(seed=43, num_cams=20, num_pts=800)
    path = str(tmp_path / "walk.bal")
    graphio.save_bal(path, prob)
    loaded = graphio.load_bal(path)
    a = g2o_amd.BALAdjuster(prob, device=0)
    b = g2o_amd.BALAdjuster(loaded, device=0)
    sa, sb = a.optimize(5), b.optimize(5)
    assert len(sa) == len(sb)
    for x, y in zip(sa, sb):
        assert abs(x["chi2"] - y["chi2"]) <= 1e-9 * x["chi2"]


def test_bal_rejections(small):
    import g2o_amd
    from g2o_amd.solver import G2oHipError
    dup = small.copy()
    free = (small.cam_fixed[small.edge_cam] == 0) & (small.pt_fixed[small.edge_pt] == 0)
    k = int(np.flatnonzero(free)[0])  # an edge between a free camera and a free point
    dup.edge_cam = np.r_[dup.edge_cam, dup.edge_cam[k]].astype(np.int32)
    dup.edge_pt = np.r_[dup.edge_pt, dup.edge_pt[k]].astype(np.int32)
    dup.meas = np.r_[dup.meas, dup.meas[k:k + 1]]
    dup.weight = np.r_[dup.weight, 1.0]
    with pytest.raises(G2oHipError) as ex:
        g2o_amd.BALAdjuster(dup, device=0)
    assert ex.value.status == g2o_amd._lib.ERR_UNSUPPORTED
    ba = g2o_amd.BALAdjuster(small, device=0)
    ba.set_options(precision=g2o_amd.FP32)
    with pytest.raises(G2oHipError) as ex:
        ba.build_system()
    assert ex.value.status == g2o_amd._lib.ERR_UNSUPPORTED
