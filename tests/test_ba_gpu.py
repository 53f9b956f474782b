"""GPU parity of the bundle-adjustment path against the FP64 CPU oracle.

Every call goes through the C ABI (libg2o_hip.so). Tolerances:
* FP64 path: blocks, gradients and chi2 relative 1e-11 (only the summation
  order differs from the oracle);
* FP32 path: Hessian blocks relative 2e-5 of the block norm; final LM chi2
  relative 1e-6 of the FP64 result.
"""
import numpy as np
import pytest

from g2o_amd import datasets

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def small():
    return datasets.make_small_ba(seed=1000)


@pytest.fixture(scope="module")
def small_huber():
    p = datasets.make_small_ba(seed=1001)
    p.huber = 2.0
    # a few gross outliers so that the Huber branch is taken
    p.meas[::37] += 25.0
    return p


@pytest.fixture(scope="module")
def walk():
    return datasets.make_walk_ba(seed=7, num_cams=60, num_pts=6000)


@pytest.fixture(scope="module")
def small_info(small_huber):
    """Full 2x2 information per edge (EdgeSE3ProjectXYZ::setInformation)."""
    p = small_huber.copy()
    p.omega = datasets.random_information(p.num_edges, 21)
    return p


@pytest.fixture(scope="module")
def walk_info(walk):
    p = walk.copy()
    p.omega = datasets.random_information(p.num_edges, 22)
    return p


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _mk(prob, **kw):
    import g2o_amd
    return g2o_amd.BundleAdjuster(prob, device=0, **kw)


@pytest.mark.parametrize("case", ["small", "small_huber", "walk", "small_info", "walk_info"])
def test_chi2_and_hessian_fp64(case, request, oracle_lib):
    prob = request.getfixturevalue(case)
    _say(f"case {case}: C={prob.num_cams} P={prob.num_pts} E={prob.num_edges}")
    import g2o_amd
    ba = _mk(prob, precision=g2o_amd.FP64)
    ora = oracle_lib.OracleBA(prob)
    chi_o, Hpp_o, Hll_o, Hpl_o, b_o = ora.linearize()
    chi_g = ba.compute_active_errors()
    assert abs(chi_g - chi_o) <= 1e-11 * abs(chi_o)
    ba.build_system()
    Hpp, Hll, Hpl = ba.hessian()
    b = ba.b()
    assert _rel(Hpp, Hpp_o) < 1e-11
    assert _rel(Hll, Hll_o) < 1e-11
    assert _rel(Hpl, Hpl_o) < 1e-11
    assert _rel(b, b_o) < 1e-11
    # errors per edge
    e = ba.errors()
    assert e.shape == (prob.num_edges, 2)


@pytest.mark.parametrize("case", ["small", "walk"])
def test_hessian_fp32(case, request, oracle_lib):
    prob = request.getfixturevalue(case)
    import g2o_amd
    ba = _mk(prob, precision=g2o_amd.FP32)
    ora = oracle_lib.OracleBA(prob)
    chi_o, Hpp_o, Hll_o, Hpl_o, b_o = ora.linearize()
    ba.build_system()
    Hpp, Hll, Hpl = ba.hessian()
    for a, o in ((Hpp, Hpp_o), (Hll, Hll_o), (Hpl, Hpl_o)):
        scale = np.maximum(np.linalg.norm(o.reshape(o.shape[0], -1), axis=1), 1e-300)
        err = np.linalg.norm((a - o).reshape(o.shape[0], -1), axis=1) / scale
        assert err.max() < 2e-5
    assert _rel(ba.b(), b_o) < 2e-5


@pytest.mark.parametrize("case", ["small", "small_huber", "walk", "small_info"])
def test_schur_fp64(case, request, oracle_lib):
    prob = request.getfixturevalue(case)
    import g2o_amd
    ba = _mk(prob, precision=g2o_amd.FP64, pcg_tolerance=1e-30, pcg_max_iter=400)
    ora = oracle_lib.OracleBA(prob)
    ora.linearize()
    lam = 3.7
    st, rows_o, cols_o, blocks_o, bs_o = ora.schur(lam)
    assert st == 0
    ba.build_system()
    ba.set_lambda(lam)
    assert ba.solve()
    rows, cols, blocks, bs = ba.schur(ora.nschur)
    assert np.array_equal(rows, rows_o) and np.array_equal(cols, cols_o)
    assert _rel(blocks, blocks_o) < 1e-11
    assert _rel(bs, bs_o) < 1e-11
    # the PCG run to convergence gives the dense Cholesky solution
    st, dx_o, _ = ora.solve(lam, solver=0)
    assert st == 0
    assert _rel(ba.x(), dx_o) < 1e-8


@pytest.mark.parametrize("case", ["small", "walk"])
def test_pcg_iterations_match_reference_rule(case, request, oracle_lib):
    prob = request.getfixturevalue(case)
    import g2o_amd
    ba = _mk(prob, precision=g2o_amd.FP64)
    ora = oracle_lib.OracleBA(prob)
    ora.linearize()
    ba.build_system()
    lam = 1.0
    st, dx_o, it_o = ora.solve(lam, solver=1, tol=1e-6)
    ba.set_lambda(lam)
    assert ba.solve()
    info = ba.last_solve_info()
    _say(f"pcg iterations gpu {info[0]} oracle {it_o}")
    assert int(info[0]) == it_o
    assert _rel(ba.x(), dx_o) < 1e-9


@pytest.mark.parametrize("case", ["small", "small_huber", "walk", "walk_info"])
def test_lm_trace_fp64(case, request, oracle_lib):
    prob = request.getfixturevalue(case)
    import g2o_amd
    ba = _mk(prob, precision=g2o_amd.FP64)
    ora = oracle_lib.OracleBA(prob)
    n_o, trace, _ = ora.optimize(8, solver=1, tol=1e-6)
    stats = ba.optimize(8)
    _say(f"{case}: gpu chi2 {[s['chi2'] for s in stats]}")
    _say(f"{case}: cpu chi2 {list(trace[:, 1])}")
    assert len(stats) == n_o
    # PCG counts: equal, or one apart where the stop test dn <= d0 compares two
    # values within rounding of each other (the device's sums round
    # differently from the oracle's, and the linearisation's LDS sums
    # differently from run to run); the first solve's count is exact
    assert stats[0]["iterations_linear_solver"] == int(trace[0, 4])
    off = 0
    for s, t in zip(stats, trace):
        assert abs(s["chi2"] - t[1]) <= 1e-8 * t[1]
        assert s["levenberg_iterations"] == int(t[3])
        assert abs(s["iterations_linear_solver"] - int(t[4])) <= 1
        off += s["iterations_linear_solver"] != int(t[4])
    assert off <= 1
    c, p = ba.state()
    co, po = ora.state()
    # inexact PCG (tol 1e-6) lets rounding differences grow slowly over 8 iterations
    assert _rel(p, po) < 1e-6


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_host_mapped_and_synchronised_paths_agree(precision, walk, monkeypatch):
    """The PCG's progress and final scalars and the LM trial scalars reach
    the host through host-mapped memory; with G2OHIP_HOST_MAPPED=0 the host
    copies them after a stream synchronisation instead. Both run the same
    kernels, so the LM traces agree to the run-to-run rounding of the LDS sums
    (and take the same trials)."""
    import g2o_amd
    prec = g2o_amd.FP64 if precision == "fp64" else g2o_amd.FP32
    runs = []
    for env in ("1", "0"):
        monkeypatch.setenv("G2OHIP_HOST_MAPPED", env)
        ba = _mk(walk, precision=prec, refine_steps=0 if precision == "fp64" else 1)
        runs.append((ba.optimize(6), ba.state()))
        del ba
    (s1, (c1, p1)), (s0, (c0, p0)) = runs
    _say(f"{precision}: mapped {[s['chi2'] for s in s1]}")
    _say(f"{precision}: synced {[s['chi2'] for s in s0]}")
    assert len(s1) == len(s0)
    for a, b in zip(s1, s0):
        assert a["levenberg_iterations"] == b["levenberg_iterations"]
        assert abs(a["iterations_linear_solver"] - b["iterations_linear_solver"]) <= 1
        assert abs(a["chi2"] - b["chi2"]) <= 1e-8 * b["chi2"]
    assert _rel(p1, p0) < 1e-6


def test_fp32_falls_back_to_fp64_when_refinement_stalls(walk):
    """At a tolerance the FP32-rounded reduced system cannot reach (1e-18 of
    the initial r.M^-1 r: the FP64 residual of an FP32 solution is about
    (6e-8)^2 of it), the FP64 residual misses the threshold by more than 10x,
    so the solve is redone in FP64 at the same state and the rest of
    optimize() stays FP64: the trace is then the FP64 path's."""
    import g2o_amd
    kw = dict(pcg_tolerance=1e-18, pcg_max_iter=300)
    b32 = _mk(walk, precision=g2o_amd.FP32, refine_steps=1, **kw)
    s32 = b32.optimize(3)
    b64 = _mk(walk, precision=g2o_amd.FP64, refine_steps=0, **kw)
    s64 = b64.optimize(3)
    _say(f"fallback flags {[s['fp64_fallback'] for s in s32]}")
    _say(f"fp32 {[s['chi2'] for s in s32]}")
    _say(f"fp64 {[s['chi2'] for s in s64]}")
    assert s32[0]["fp64_fallback"] == 1
    assert len(s32) == len(s64)
    for a, b in zip(s32, s64):
        assert a["levenberg_iterations"] == b["levenberg_iterations"]
        assert abs(a["chi2"] - b["chi2"]) <= 1e-8 * b["chi2"]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_lm_trace_with_rejected_trials(precision, oracle_lib):
    """A start far from the optimum makes a step overshoot, so a trial is
    rejected and popped: the next trial must linearise at the restored
    state (regression: the device camera table once kept the rejected one)."""
    import g2o_amd
    prob = datasets.make_walk_ba(seed=8, num_cams=60, num_pts=6000, sig_t=0.3, sig_r=0.01, sig_p=0.5)
    prec = g2o_amd.FP64 if precision == "fp64" else g2o_amd.FP32
    ba = _mk(prob, precision=prec, refine_steps=0 if precision == "fp64" else 1)
    ora = oracle_lib.OracleBA(prob)
    n_o, trace, _ = ora.optimize(8, solver=1, tol=1e-6)
    stats = ba.optimize(8)
    _say(f"{precision}: trials gpu {[s['levenberg_iterations'] for s in stats]} cpu {list(trace[:, 3].astype(int))}")
    assert trace[:, 3].sum() > n_o  # some trials were rejected
    assert len(stats) == n_o
    for s, t in zip(stats, trace):
        assert s["levenberg_iterations"] == int(t[3])
        assert abs(s["chi2"] - t[1]) <= (1e-8 if precision == "fp64" else 1e-6) * t[1]


@pytest.mark.parametrize("refine", [0, 1, 2])
def test_lm_fp32_refine(refine, walk, oracle_lib):
    import g2o_amd
    ba64 = _mk(walk, precision=g2o_amd.FP64)
    s64 = ba64.optimize(8)
    ba32 = _mk(walk, precision=g2o_amd.FP32, refine_steps=refine)
    s32 = ba32.optimize(8)
    _say(f"refine {refine}: fp64 {s64[-1]['chi2']:.10e} fp32 {s32[-1]['chi2']:.10e}")
    rel = abs(s32[-1]["chi2"] - s64[-1]["chi2"]) / s64[-1]["chi2"]
    assert rel < 1e-6


def test_linear_solver_pcg_fixture():
    """LinearSolverPCG on the reference's 36x36 block fixture
    (unit_test/solver/sparse_system_helper.cpp), solved twice as in
    linear_solver_test.cpp:89-101, to 1e-6."""
    import json
    import os
    import g2o_amd
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "sparse_system_36.json")))
    ctx = g2o_amd.Context(0)
    rows, cols = np.array(fx["rows"]), np.array(fx["cols"])
    blocks = np.array(fx["blocks"]).reshape(-1, 3, 3)
    b = np.array(fx["b"])
    x_ref = np.array(fx["x"])
    resid = -1.0
    for _ in range(2):
        x, it, resid = g2o_amd.linear_solve_pcg(ctx, 12, 3, rows, cols, blocks, b, tol=1e-12, max_iter=200,
                                                residual=-1.0)
        assert np.linalg.norm(x - x_ref) <= 1e-6 * np.linalg.norm(x_ref) + 1e-7


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_implicit_matches_explicit(prec, walk, oracle_lib):
    """Implicit (matrix-free) Schur inside the PCG applies the same reduced
    operator as the explicit Hschur: run to a tight tolerance both give the
    step of the oracle's dense Cholesky solve of (H + lambda I) x = b."""
    import g2o_amd
    p = g2o_amd.FP64 if prec == "fp64" else g2o_amd.FP32
    tol = 1e-24 if prec == "fp64" else 1e-12
    ora = oracle_lib.OracleBA(walk)
    ora.linearize()
    st, dx_o, _ = ora.solve(2.5, solver=0)
    assert st == 0
    xs = []
    for mode in (g2o_amd.SCHUR_EXPLICIT, g2o_amd.SCHUR_IMPLICIT):
        ba = _mk(walk, precision=p, schur_mode=mode, pcg_tolerance=tol, pcg_max_iter=600, refine_steps=0)
        ba.build_system()
        ba.set_lambda(2.5)
        assert ba.solve()
        xs.append(ba.x())
    bound = 1e-9 if prec == "fp64" else 2e-4
    _say(f"implicit vs explicit {prec}: {_rel(xs[1], xs[0]):.3e}; vs dense: {_rel(xs[0], dx_o):.3e} "
         f"{_rel(xs[1], dx_o):.3e}")
    assert _rel(xs[1], xs[0]) < bound
    dense_bound = 1e-8 if prec == "fp64" else 2e-4
    assert _rel(xs[0], dx_o) < dense_bound
    assert _rel(xs[1], dx_o) < dense_bound


def test_profile_counts(walk):
    import g2o_amd
    ba = _mk(walk, precision=g2o_amd.FP32)
    ba.set_profiling(True)
    ba.optimize(2)
    prof = ba.profile()
    _say(f"profile {prof}")
    assert prof["linearize"][1] >= 1 and prof["schur"][1] >= 1 and prof["pcg"][1] >= 1
    assert all(v[0] >= 0 for v in prof.values())


@pytest.mark.parametrize("case", ["small", "walk"])
def test_multiply_hessian_is_the_pose_block_product(case, request):
    """BlockSolverBase::multiplyHessian (block_solver.h:149-151): the pose rows
    of Hpp src with the undamped pose blocks of the last buildSystem (block
    diagonal for bundle adjustment); the landmark rows stay zero."""
    prob = request.getfixturevalue(case)
    import g2o_amd
    ba = _mk(prob, precision=g2o_amd.FP64)
    ba.build_system()
    ba.set_lambda(123.0)  # the damping is not part of the product
    Hpp, _, _ = ba.hessian()
    rng = np.random.default_rng(4)
    src = rng.normal(size=6 * ba.Cf + 3 * ba.Pf)
    out = ba.multiply_hessian(src)
    ref = np.einsum("cij,cj->ci", Hpp, src[:6 * ba.Cf].reshape(-1, 6)).ravel()
    assert _rel(out[:6 * ba.Cf], ref) < 1e-14
    assert np.all(out[6 * ba.Cf:] == 0)
