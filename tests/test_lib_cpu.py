"""C-ABI library checks that need no GPU: it loads, exports every symbol
include/g2o_hip.h declares, and the host-side structure builder (device -1
context) produces the reference's index mapping and Hschur pattern."""
import os
import re

import numpy as np
import pytest

import g2o_amd
from g2o_amd import datasets

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "g2o_hip.h")).read()
    return sorted(set(re.findall(r"\b(g2ohip_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = g2o_amd.lib()
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    assert sorted(g2o_amd._lib.EXPORTS) == names


def test_default_options_follow_reference():
    o = g2o_amd.default_options()
    assert o.pcg_tolerance == 1e-6 and o.pcg_max_iter == -1  # linear_solver_pcg.h:45-50
    assert o.max_trials == 10 and o.tau == 1e-5              # levenberg.cpp:42-56
    assert o.precision == g2o_amd.FP64 and o.schur_mode == g2o_amd.SCHUR_EXPLICIT


def test_host_only_context_rejects_device_calls():
    ctx = g2o_amd.Context(-1)
    with pytest.raises(g2o_amd.G2oHipError):
        ctx.call("g2ohip_build_system")


def _pairs_numpy(prob):
    free = np.where(prob.cam_fixed == 0, np.cumsum(prob.cam_fixed == 0) - 1, -1)
    by_pt = {}
    for c, p in zip(prob.edge_cam, prob.edge_pt):
        if free[c] >= 0 and prob.pt_fixed[p] == 0:
            by_pt.setdefault(int(p), []).append(int(free[c]))
    pairs = 0
    Cf = int((prob.cam_fixed == 0).sum())
    keys = {(i, i) for i in range(Cf)}
    for cams in by_pt.values():
        cams = sorted(cams)
        for a in range(len(cams)):
            for b in range(a, len(cams)):
                keys.add((cams[a], cams[b]))
                pairs += cams[a] != cams[b]
    return pairs, len(keys)


@pytest.mark.parametrize("seed,cams,pts", [(3, 40, 3000), (9, 120, 8000)])
def test_structure_matches_reference_pattern(seed, cams, pts):
    prob = datasets.make_walk_ba(seed=seed, num_cams=cams, num_pts=pts)
    ba = g2o_amd.BundleAdjuster(prob, device=-1)
    info = ba.structure_info()
    Cf = int((prob.cam_fixed == 0).sum())
    assert info[0] == Cf and info[1] == prob.num_pts and info[2] == prob.num_edges
    pairs, nS = _pairs_numpy(prob)
    assert info[6] == pairs      # off-diagonal Schur pair products
    assert info[3] == nS         # upper Hschur blocks (block_solver.hpp:237-266)
    assert info[9] == 2 * nS - Cf  # both triangles
    # every off-diagonal block has at least one partial sum; windows hold at
    # most 512 observations each
    assert info[7] >= nS - Cf
    nq = int(((prob.cam_fixed[prob.edge_cam] == 0) & (prob.pt_fixed[prob.edge_pt] == 0)).sum())
    assert (nq + 511) // 512 <= info[5] <= nq


def test_long_tracks_take_the_camera_row_schur():
    """Tracks longer than 64 cameras making most of the pair products (Zipf
    law, tracks up to 300): the off-diagonal Schur blocks are formed by
    camera rows, so the structure has no point windows and no partial sums;
    a walk with geometric tracks keeps the window plan."""
    z = datasets.make_zipf_ba(seed=5, num_cams=300, num_pts=20000)
    info = g2o_amd.BundleAdjuster(z, device=-1).structure_info()
    pairs, nS = _pairs_numpy(z)
    assert info[5] == 0 and info[7] == 0
    assert info[6] == pairs and info[3] == nS
    w = datasets.make_walk_ba(seed=3, num_cams=40, num_pts=3000)
    assert g2o_amd.BundleAdjuster(w, device=-1).structure_info()[5] > 0


def test_structure_rejects_bad_indices():
    prob = datasets.make_small_ba(seed=1)
    prob.edge_cam = prob.edge_cam.copy()
    prob.edge_cam[5] = prob.num_cams + 3
    with pytest.raises(g2o_amd.G2oHipError) as ei:
        g2o_amd.BundleAdjuster(prob, device=-1)
    assert ei.value.status == g2o_amd._lib.ERR_INVALID


def test_fixed_points_and_cameras_are_excluded():
    prob = datasets.make_walk_ba(seed=5, num_cams=30, num_pts=2000)
    prob.pt_fixed = prob.pt_fixed.copy()
    prob.pt_fixed[::10] = 1
    ba = g2o_amd.BundleAdjuster(prob, device=-1)
    info = ba.structure_info()
    assert info[1] == prob.num_pts - int(prob.pt_fixed.sum())
    # edges with both ends fixed are inactive
    both = (prob.cam_fixed[prob.edge_cam] == 1) & (prob.pt_fixed[prob.edge_pt] == 1)
    assert info[2] == prob.num_edges - int(both.sum())


def with_isolated_vertices(prob):
    """The problem plus one free camera and one free point that no edge uses."""
    return datasets.BAProblem(np.vstack([prob.cams, prob.cams[-1:]]), np.vstack([prob.intr, prob.intr[-1:]]),
                              np.vstack([prob.pts, prob.pts[-1:]]), prob.edge_cam, prob.edge_pt, prob.meas,
                              prob.weight, np.r_[prob.cam_fixed, 0].astype(np.uint8),
                              np.r_[prob.pt_fixed, 0].astype(np.uint8), prob.huber)


def test_vertices_without_edges_are_not_active():
    """initializeOptimization(level) activates a vertex only through an
    active edge (sparse_optimizer.cpp:217-285), so a free vertex without edges
    gets no Hessian index; with no edge at all there is nothing to optimise."""
    prob = datasets.make_small_ba(seed=1)
    base = g2o_amd.BundleAdjuster(prob, device=-1).structure_info()
    iso = g2o_amd.BundleAdjuster(with_isolated_vertices(prob), device=-1).structure_info()
    assert iso[0] == base[0] and iso[1] == base[1] and iso[2] == base[2]
    empty = datasets.BAProblem(prob.cams, prob.intr, prob.pts, prob.edge_cam[:0], prob.edge_pt[:0],
                               prob.meas[:0], prob.weight[:0], prob.cam_fixed, prob.pt_fixed, 0.0)
    with pytest.raises(g2o_amd.G2oHipError):
        g2o_amd.BundleAdjuster(empty, device=-1)
    pg = datasets.make_sphere_pose_graph(nodes_per_level=20, laps=2)
    pgi = datasets.PoseGraphProblem(np.vstack([pg.poses, pg.poses[-1:]]), pg.edge_from, pg.edge_to, pg.meas,
                                    pg.info, np.r_[pg.fixed, 0].astype(np.uint8))
    a = g2o_amd.PoseGraphOptimizer(pg, device=-1).structure_info()
    b = g2o_amd.PoseGraphOptimizer(pgi, device=-1).structure_info()
    assert np.array_equal(a, b)


def test_pose_graph_structure_matches_reference_pattern():
    prob = datasets.make_sphere_pose_graph(seed=3, nodes_per_level=30, laps=4)
    pgo = g2o_amd.PoseGraphOptimizer(prob, device=-1)
    info = pgo.structure_info()
    free = prob.fixed == 0
    H = np.where(free, np.cumsum(free) - 1, -1)
    hi, hj = H[prob.edge_from], H[prob.edge_to]
    both = (hi >= 0) & (hj >= 0)
    pairs = {(min(a, b), max(a, b)) for a, b in zip(hi[both], hj[both])}
    assert info[0] == free.sum()
    assert info[1] == int(((hi >= 0) | (hj >= 0)).sum())
    assert info[2] == free.sum() + len(pairs)
    assert info[3] == free.sum() + 2 * len(pairs)


def test_pose_graph_structure_rejects_bad_edges():
    prob = datasets.make_sphere_pose_graph(seed=3, nodes_per_level=10, laps=2)
    prob.edge_to = prob.edge_to.copy()
    prob.edge_to[3] = prob.edge_from[3]
    with pytest.raises(g2o_amd.G2oHipError) as ei:
        g2o_amd.PoseGraphOptimizer(prob, device=-1)
    assert ei.value.status == g2o_amd._lib.ERR_INVALID


_OOM_CHILD = r"""
import ctypes as ct, resource, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import g2o_amd
L = g2o_amd.lib()
ctx = g2o_amd.Context(-1)
E = 20_000_000
ec = np.zeros(E, dtype=np.int32)
ep = np.arange(E, dtype=np.int32)
# cap the address space a little above what the process maps now: the
# structure builder's host vectors for 20M edges cannot fit
with open("/proc/self/status") as f:
    vm = [int(l.split()[1]) for l in f if l.startswith("VmSize:")][0] * 1024
resource.setrlimit(resource.RLIMIT_AS, (vm + (64 << 20), vm + (64 << 20)))
st = L.g2ohip_ba_set_structure(ctx.h, 1, E, E, ec.ctypes.data_as(ct.POINTER(ct.c_int)),
                               ep.ctypes.data_as(ct.POINTER(ct.c_int)), None, None)
print("status", st, L.g2ohip_last_error(ctx.h).decode(), flush=True)
"""


def test_host_allocation_failure_returns_oom_status():
    """A std::bad_alloc inside an entry point (here the structure builder's host
    vectors) comes back as G2OHIP_ERR_OOM, not as a terminated process: no
    exception crosses the C ABI."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, "-c", _OOM_CHILD, ROOT], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    line = [l for l in r.stdout.splitlines() if l.startswith("status")][0]
    assert line.split()[1] == str(g2o_amd._lib.ERR_OOM), line
