"""Differentiable bundle adjustment: a torch.autograd.Function around the
HIP solver.

Forward: LM to convergence (any precision the BundleAdjuster is set to), then
undamped exact-Hessian Newton steps in FP64 so the state is a stationary point
of F = sum_e rho(w_e |e_e|^2). Outputs the camera centres C = -R^T t [C,3]
and the points [P,3].

Backward: one implicit-function solve at the optimum (g2ohip_ba_adjoint):
H lambda = dL/dx with the exact Hessian, dL/dz = -(dG/dz)^T lambda,
dL/dw = -(dG/dw)^T lambda. The gradients come back in the dtype and on the
device of the measurement / weight tensors.

There is no counterpart in g2o, which stops at the solution.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .datasets import BAProblem, quat_to_rot, select_edges
from .solver import BundleAdjuster, G2oHipError


def camera_centres(cams7: np.ndarray) -> np.ndarray:
    """C = -R^T t for world-to-camera poses [tx ty tz qx qy qz qw]."""
    R = quat_to_rot(cams7[:, 3:7])
    return -np.einsum("cji,cj->ci", R, cams7[:, :3])


def centres_grad_to_tangent(cams7: np.ndarray, g_centres: np.ndarray) -> np.ndarray:
    """Pull dL/dC back to the left exp-map update [omega, upsilon] of each
    camera: C(exp(d) T) = C - R^T upsilon + O(|d|^2), so dL/domega = 0 and
    dL/dupsilon = -R dL/dC."""
    R = quat_to_rot(cams7[:, 3:7])
    out = np.zeros((cams7.shape[0], 6))
    out[:, 3:] = -np.einsum("cij,cj->ci", R, g_centres)
    return out


class BundleAdjustmentFunction(torch.autograd.Function):
    """(meas [E,2], weight [E]) -> (centres [C,3], points [P,3])."""

    @staticmethod
    def forward(ctx, meas, weight, ba: BundleAdjuster, init_cams, init_pts, lm_iterations: int,
                newton_steps: int):
        m = meas.detach().to("cpu", torch.float64).numpy()
        w = weight.detach().to("cpu", torch.float64).numpy()
        prob = ba.prob
        ba.set_measurements(m, w, prob.intr, prob.huber)
        ba.set_state(init_cams, init_pts)
        ba.optimize(lm_iterations)
        if newton_steps > 0 and not ba.newton(newton_steps):
            raise RuntimeError("exact Hessian not positive definite: the LM solution is not a strict minimum")
        cams, pts = ba.state()
        ctx.ba = ba
        ctx.cams = cams
        ctx.meta = (meas.dtype, meas.device, weight.dtype, weight.device)
        centres = torch.from_numpy(camera_centres(cams)).to(meas.device, meas.dtype)
        points = torch.from_numpy(pts).to(meas.device, meas.dtype)
        return centres, points

    @staticmethod
    def backward(ctx, g_centres, g_points):
        mdt, mdev, wdt, wdev = ctx.meta
        cams = ctx.cams
        gc = np.zeros((cams.shape[0], 3)) if g_centres is None else g_centres.detach().to("cpu", torch.float64).numpy()
        gp = None if g_points is None else g_points.detach().to("cpu", torch.float64).numpy()
        dz, dw, _ = ctx.ba.adjoint(centres_grad_to_tangent(cams, gc), gp)
        return (torch.from_numpy(dz).to(mdev, mdt), torch.from_numpy(dw).to(wdev, wdt), None, None, None, None, None)


def solve(meas: torch.Tensor, weight: torch.Tensor, ba: BundleAdjuster, init_cams: np.ndarray, init_pts: np.ndarray,
          lm_iterations: int = 30, newton_steps: int = 3):
    """Differentiable solve; see BundleAdjustmentFunction."""
    return BundleAdjustmentFunction.apply(meas, weight, ba, init_cams, init_pts, lm_iterations, newton_steps)


def prepare(prob: BAProblem, device: int = 0, precision: int = _lib.FP32, refine_steps: int = 1,
            lm_iterations: int = 30, chi2_max: float | None = None, min_obs: int = 3, max_cond: float = 1e6,
            newton_steps: int = 12, polish_iterations: int = 100, polish_tolerance: float = 1e-12,
            keep_edges: np.ndarray | None = None):
    """Conditions a robust BA problem for implicit differentiation and solves it.

    The implicit function theorem needs a strict stationary point with a
    well-conditioned exact Hessian. Under a Huber kernel, uniform outliers and
    landmarks seen by two cameras whose residuals both lie in the kernel's
    linear range have indefinite or near-singular exact point blocks. So:
    robust LM; edges with w |e|^2 > chi2_max (default 5.991 delta^2, the 95 %
    chi-square bound scaled to the kernel) are dropped, as g2o workflows set
    outlier edges to level 1; landmarks left with < min_obs edges are dropped;
    LM again; landmarks whose exact 3x3 block has condition number > max_cond
    are dropped; LM again (polish_iterations at PCG tolerance
    polish_tolerance, so that Newton starts inside its quadratic region) and
    exact-Hessian Newton steps until the gradient reaches its rounding floor
    (info['gradient_max']).

    keep_edges (indices into the input's edges, e.g. an earlier call's
    info['edge_index']): the selection is taken as given instead of from the
    robust chi2 and the conditioning test; the robust LM on every edge still
    runs first and provides the start state.

    Returns (sub_problem, BundleAdjuster at the stationary point). The
    sub-problem's info['edge_index'] maps its edges to the input's."""
    ba = BundleAdjuster(prob, device=device, precision=precision, refine_steps=refine_steps)
    ba.optimize(lm_iterations)
    cams, pts = ba.state()
    e = ba.errors()
    chi = prob.weight * (e * e).sum(1)
    if chi2_max is None:
        chi2_max = 5.991 * prob.huber ** 2 if prob.huber > 0 else np.inf
    if keep_edges is not None:
        mask = np.zeros(prob.num_edges, dtype=bool)
        mask[np.asarray(keep_edges)] = True
    else:
        mask = chi <= chi2_max
    sub = select_edges(prob, mask, min_obs)
    sub.cams, sub.pts = cams, pts[sub.info["pt_index"]]
    ba = BundleAdjuster(sub, device=device, precision=precision, refine_steps=refine_steps,
                        pcg_tolerance=polish_tolerance)
    ba.optimize(polish_iterations)
    try:
        ba.adjoint(np.zeros((sub.num_cams, 6)))  # evaluates the exact blocks at the state
    except G2oHipError as ex:  # blocks are there even where the reduced system is not yet PD
        if ex.status != _lib.ERR_NOT_PD:
            raise
    _, Hll, _ = ba.hessian()
    ev = np.linalg.eigvalsh(Hll)
    free = np.flatnonzero(sub.pt_fixed == 0)
    weak = np.zeros(sub.num_pts, dtype=bool)
    weak[free] = ev[:, 0] <= ev[:, 2] / max_cond
    if weak.any() and keep_edges is None:
        cams, pts = ba.state()
        eidx = sub.info["edge_index"]
        sub2 = select_edges(sub, ~weak[sub.edge_pt], min_obs)
        sub2.info["edge_index"] = eidx[sub2.info["edge_index"]]
        sub2.info["pt_index"] = sub.info["pt_index"][sub2.info["pt_index"]]
        sub2.cams, sub2.pts = cams, pts[np.searchsorted(sub.info["pt_index"], sub2.info["pt_index"])]
        sub = sub2
        ba = BundleAdjuster(sub, device=device, precision=precision, refine_steps=refine_steps,
                            pcg_tolerance=polish_tolerance)
        ba.optimize(polish_iterations)
    g = _newton_to_stationary(ba, newton_steps)
    sub.info["gradient_max"] = g
    return sub, ba


def _newton_to_stationary(ba: BundleAdjuster, max_steps: int, gtol: float = 1e-6) -> float:
    """Exact-Hessian Newton steps until max |grad| stops falling below 1e-3 of
    its previous value's rounding floor: at most max_steps steps, early exit
    at max |grad| <= gtol * (1 + initial max |grad|). Returns the final max
    |grad|; raises if no step is acceptable or the gradient stays large.
    The gradient is read from an FP64 linearisation (an FP32 one carries
    rounding far above the stationary point's)."""
    prec = ba.ctx.set_options().precision
    ba.set_options(precision=_lib.FP64)
    try:
        return _newton_loop(ba, max_steps, gtol)
    finally:
        ba.set_options(precision=prec)


def _newton_loop(ba: BundleAdjuster, max_steps: int, gtol: float) -> float:
    ba.build_system()
    g0 = g = float(np.abs(ba.b()).max())
    for _ in range(max_steps):
        if g <= gtol * (1.0 + g0):
            break
        if not ba.newton(1):
            raise RuntimeError("exact-Hessian Newton found no descent step")
        ba.build_system()
        g_new = float(np.abs(ba.b()).max())
        if g_new > 0.5 * g and g_new <= 1e-4 * (1.0 + g0):
            g = g_new
            break  # at the rounding floor
        g = g_new
    if g > 1e-3 * (1.0 + g0):
        raise RuntimeError(f"exact-Hessian Newton did not reach a stationary point (max |grad| {g:.3e})")
    return g
