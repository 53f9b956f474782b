"""Exact-Hessian path on the GPU (g2ohip_ba_newton, g2ohip_ba_adjoint) and
the torch.autograd wrapper, against the oracle (hyper-dual exact Hessian,
Cholesky) and against central finite differences of re-converged solves."""
import os
import sys

import numpy as np
import pytest

import g2o_amd
from g2o_amd import datasets

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_adjoint_cpu import centre_loss  # noqa: E402

pytestmark = pytest.mark.gpu


def _small_problem():
    prob = datasets.make_small_ba(seed=5, num_cams=8, num_pts=80, huber=1.5)
    prob.weight = np.random.default_rng(1).uniform(0.5, 2.0, prob.num_edges)
    return prob


def _oracle_converged(prob):
    from oracle import oracle
    ora = oracle.OracleBA(prob)
    ora.optimize(30, solver=0)
    assert ora.newton(4) == 0
    return ora


def test_exact_hessian_blocks_match_oracle():
    prob = _small_problem()
    ora = _oracle_converged(prob)
    cams, pts = ora.state()
    # a state away from the optimum as well, where the second-order terms differ most
    for c, p in ((cams, pts), (prob.cams, prob.pts)):
        ora.set_state(c, p)
        Hpp_o, Hll_o, Hpl_o = ora.exact_hessian()
        ba = g2o_amd.BundleAdjuster(prob, device=0)
        ba.set_state(c, p)
        try:
            ba.adjoint(np.zeros((prob.num_cams, 6)))
        except g2o_amd.G2oHipError as ex:  # away from the optimum H may be indefinite
            assert ex.status == g2o_amd._lib.ERR_NOT_PD
        Hpp, Hll, Hpl = ba.hessian()
        for a, b in ((Hpp, Hpp_o), (Hll, Hll_o), (Hpl, Hpl_o)):
            assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()


def test_newton_reaches_the_oracle_stationary_point():
    prob = _small_problem()
    ora = _oracle_converged(prob)
    cams_o, pts_o = ora.state()
    ba = g2o_amd.BundleAdjuster(prob, device=0)
    ba.optimize(30)
    assert ba.newton(4)
    cams, pts = ba.state()
    assert np.abs(cams[:, :3] - cams_o[:, :3]).max() <= 1e-9
    assert np.abs(np.abs(np.einsum("ci,ci->c", cams[:, 3:], cams_o[:, 3:])) - 1).max() <= 1e-12
    assert np.abs(pts - pts_o).max() <= 1e-9


def test_adjoint_matches_oracle():
    prob = _small_problem()
    ora = _oracle_converged(prob)
    cams, pts = ora.state()
    L, gcam = centre_loss(cams, prob.gt_cams)
    gpt = np.random.default_rng(2).normal(size=(prob.num_pts, 3)) * 1e-2
    st, dz_o, dw_o, _ = ora.adjoint(gcam, gpt)
    assert st == 0
    ba = g2o_amd.BundleAdjuster(prob, device=0)
    ba.set_state(cams, pts)
    dz, dw, it = ba.adjoint(gcam, gpt)
    assert it > 0
    assert np.linalg.norm(dz - dz_o) <= 1e-8 * np.linalg.norm(dz_o)
    assert np.linalg.norm(dw - dw_o) <= 1e-8 * np.linalg.norm(dw_o)


def test_autograd_function_matches_adjoint():
    import torch
    from g2o_amd import diff
    prob = _small_problem()
    ba = g2o_amd.BundleAdjuster(prob, device=0)
    meas = torch.tensor(prob.meas, dtype=torch.float32, requires_grad=True)
    weight = torch.tensor(prob.weight, dtype=torch.float32, requires_grad=True)
    centres, points = diff.solve(meas, weight, ba, prob.cams, prob.pts, 30, 4)
    gt = torch.tensor(diff.camera_centres(prob.gt_cams), dtype=torch.float32)
    loss = ((centres - gt) ** 2).sum()
    loss.backward()
    assert meas.grad.dtype == torch.float32 and weight.grad.dtype == torch.float32
    # same gradient from the C-ABI directly (measurements rounded to fp32 as the tensor holds them)
    prob32 = prob.copy()
    prob32.meas = meas.detach().double().numpy()
    prob32.weight = weight.detach().double().numpy()
    ba2 = g2o_amd.BundleAdjuster(prob32, device=0)
    ba2.optimize(30)
    assert ba2.newton(4)
    cams, _ = ba2.state()
    c = diff.camera_centres(cams)
    g = 2 * (c - diff.camera_centres(prob.gt_cams)).astype(np.float32).astype(np.float64)
    dz, dw, _ = ba2.adjoint(diff.centres_grad_to_tangent(cams, g))
    assert np.linalg.norm(meas.grad.double().numpy() - dz) <= 1e-5 * np.linalg.norm(dz)
    assert np.linalg.norm(weight.grad.double().numpy() - dw) <= 1e-5 * np.linalg.norm(dw)


def test_differentiable_ba_200_cams_finite_differences():
    """200 cameras, ~50k points, ~300k observations (Huber 2, 5 % outliers,
    weights U(0.5, 2)), solved by FP32 + FP64-refinement LM, conditioned for
    differentiation (outlier edges and ill-conditioned landmarks dropped) and
    polished by exact-Hessian Newton; loss = sum |C - C_gt|^2. The adjoint
    gradient is checked against central differences (1e-3 px) on 200 random
    measurement coordinates and 20 weights, and against the FP64 CPU
    oracle's implicit gradient."""
    import time
    from g2o_amd import diff
    prob = datasets.make_diff_ba()
    t = time.perf_counter()
    sub, ba = diff.prepare(prob)
    cams0, pts0 = ba.state()
    L0, gcam = centre_loss(cams0, prob.gt_cams)
    dz, dw, it = ba.adjoint(gcam)
    print(f"\nkept E={sub.num_edges} of {prob.num_edges}, P={sub.num_pts}; prepare {time.perf_counter() - t:.1f} s; "
          f"adjoint PCG iterations {it}, L={L0:.6e}", flush=True)
    rng = np.random.default_rng(4)
    ks = rng.choice(sub.num_edges, size=200, replace=False)
    js = rng.integers(0, 2, size=200)

    def loss_at(meas, weight):
        ba.set_measurements(meas, weight, sub.intr, sub.huber)
        ba.set_state(cams0, pts0)
        assert ba.newton(2)
        return centre_loss(ba.state()[0], prob.gt_cams)[0]

    t = time.perf_counter()
    h = 1e-3
    fd = np.zeros(200)
    for i, (k, j) in enumerate(zip(ks, js)):
        vals = []
        for sgn in (1.0, -1.0):
            m = sub.meas.copy()
            m[k, j] += sgn * h
            vals.append(loss_at(m, sub.weight))
        fd[i] = (vals[0] - vals[1]) / (2 * h)
    an = dz[ks, js]
    rel = np.linalg.norm(an - fd) / np.linalg.norm(fd)
    kw = ks[:20]
    fdw = np.zeros(20)
    for i, k in enumerate(kw):
        hw = 1e-4 * sub.weight[k]
        vals = []
        for sgn in (1.0, -1.0):
            w = sub.weight.copy()
            w[k] += sgn * hw
            vals.append(loss_at(sub.meas, w))
        fdw[i] = (vals[0] - vals[1]) / (2 * hw)
    relw = np.linalg.norm(dw[kw] - fdw) / np.linalg.norm(fdw)
    print(f"FD check ({time.perf_counter() - t:.1f} s): rel err meas {rel:.3e} (200 coords), "
          f"weights {relw:.3e} (20)", flush=True)
    assert rel <= 1e-4 and relw <= 1e-4
    # FP64 CPU implicit-gradient oracle at the same state
    from oracle import oracle
    ora = oracle.OracleBA(sub)
    ora.set_state(cams0, pts0)
    st, dz_o, dw_o, _ = ora.adjoint(gcam, np.zeros((sub.num_pts, 3)))
    assert st == 0
    relo = np.linalg.norm(dz - dz_o) / np.linalg.norm(dz_o)
    relwo = np.linalg.norm(dw - dw_o) / np.linalg.norm(dw_o)
    print(f"vs oracle: rel err meas {relo:.3e}, weights {relwo:.3e}", flush=True)
    assert relo <= 1e-6 and relwo <= 1e-6


def test_dropped_observations_have_zero_gradient():
    """The differentiable solve conditions the problem first: outlier edges
    (robust chi2 above the 95 % bound) and landmarks left ill-conditioned are
    dropped before the implicit differentiation, as g2o workflows move outlier
    edges to level 1. The full pipeline's loss therefore does not depend on a
    dropped observation while the selection stays the same: central
    differences through the whole pipeline (the robust LM on every edge,
    which sets the start state, then the solve of the selection) on a sample
    of dropped measurements give zero, next to the kept edges' adjoint
    gradient. The perturbed runs take the unperturbed run's selection: edges
    near the chi2 bound can change sides between runs through rounding alone
    (the LM's FP32 sums are not bit-reproducible). Prints the kept
    fraction."""
    import time
    from g2o_amd import diff
    prob = datasets.make_diff_ba()
    t = time.perf_counter()
    sub, ba = diff.prepare(prob)
    cams0, _ = ba.state()
    L0, gcam = centre_loss(cams0, prob.gt_cams)
    dz, _, _ = ba.adjoint(gcam)
    kept = sub.info["edge_index"]
    dropped = np.setdiff1d(np.arange(prob.num_edges), kept)
    print(f"\nkept {kept.size} of {prob.num_edges} observations ({kept.size / prob.num_edges:.1%}), "
          f"{sub.num_pts} of {prob.num_pts} points; prepare {time.perf_counter() - t:.1f} s", flush=True)
    assert 0.5 < kept.size / prob.num_edges < 1.0
    rng = np.random.default_rng(8)
    sample = rng.choice(dropped, size=3, replace=False)
    h = 1e-3
    fd = []
    for k in sample:
        vals = []
        for sgn in (1.0, -1.0):
            p = prob.copy()
            p.meas = prob.meas.copy()
            p.meas[k, 0] += sgn * h
            s2, b2 = diff.prepare(p, keep_edges=kept)
            assert np.array_equal(s2.info["edge_index"], kept)  # the same selection
            vals.append(centre_loss(b2.state()[0], prob.gt_cams)[0])
        fd.append((vals[0] - vals[1]) / (2 * h))
    scale = float(np.median(np.abs(dz[:, 0])))
    print(f"FD of the loss on dropped measurements {np.abs(fd).max():.2e}; median |dL/dz| of kept ones {scale:.2e}",
          flush=True)
    assert np.abs(fd).max() <= 1e-3 * scale
