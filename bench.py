#!/usr/bin/env python
"""Benchmark: LM iterations/s of bundle adjustment on one MI355X.

Workload: 1,500 cameras on a 3 km random walk, ~500k points, ~3M observations
(Huber delta 2, 5 % outliers), seeded synthetic data (g2o_amd.datasets). A step
is one Levenberg-Marquardt iteration (linearisation + inner trials, each a
Schur complement + block-Jacobi PCG + back-substitution + update + chi2), i.e.
one OptimizationAlgorithmLevenberg::solve. `value` is the FP32+FP64-refinement
path; the plain FP64 GPU path runs the same iterations from the same start and
its rate and the relative final-chi2 difference are reported in `config`.

Prints one JSON line. Only N=1 is supported (one GPU).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "LM iter/s at 1.5k-cam/500k-pt BA, FP32+refine vs FP64, rel chi2 diff ≤1e-6"
PEAK_HBM = 8.0e12


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(kernel, info, prec_bytes, lin_cam_rows):
    """Compulsory HBM bytes of one launch (every array touched once).
    info = structure_info(): [Cf, Pf, E, nS, nlc, windows, od pairs, od partials, maxslots, nF, lin rows, npos, nq]."""
    Cf, Pf, E, nF = int(info[0]), int(info[1]), int(info[2]), int(info[9])
    pairs, slots, nq = int(info[6]), int(info[7]), int(info[12])
    hb = 18 * prec_bytes  # one 6x3 Hpl block
    if kernel == "linearize":
        # slot: cam 4 + pos 4 + meas 16 + w 8 + Hpl write; point: ids 12 + X 24 + Hll 48 + b_l 24;
        # chunk camera partials 27 f64
        return E * (32 + hb) + Pf * 108 + lin_cam_rows * 216
    if kernel == "schur_windows":
        # observation: Hpl block + q index 4 + point index 4; point: Hll 48 (6 unique f64);
        # pair list 4 per pair product; partial: 36 values written + slot index 4
        return nq * (hb + 8) + Pf * 48 + pairs * 4 + slots * (36 * prec_bytes + 4)
    if kernel == "pcg_spmv":
        # matrix blocks (both triangles) + column indices + row pointers; vectors s, d_old read and d, q
        # written once (6 Cf doubles each); per-row dot partials
        return nF * (36 * prec_bytes + 4) + (Cf + 1) * 4 + 4 * 6 * Cf * 8 + Cf * 8
    if kernel == "refine":
        # slot: cam 4 + pos 4 + meas 16 + w 8; point: ids 12 + X 24 + x_l write 24; chunk camera partials 6 f64
        return E * 32 + Pf * 60 + lin_cam_rows * 48
    if kernel == "chi2":
        return E * 32 + Pf * 24
    return None


KERNEL_SYMBOL = {"schur_windows": "k_schur_win<float, 512>", "pcg_spmv": "k_pcg_dir_spmv<float, 6, false>",
                 "linearize": "k_linearize<float>", "refine": "k_refine", "chi2": "k_chi2"}


def measured_traffic(kernel):
    """HBM bytes per launch of the FP32-path kernel from the committed rocprofv3
    counter summary of this workload (FETCH_SIZE + WRITE_SIZE, each calibrated
    against a plain copy), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_ba_1500cam_500kpt.md")
    key = f"traffic_bytes_per_launch[{KERNEL_SYMBOL.get(kernel)}]:"
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key):
                    return int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return None


def run_gpu(prob, precision, refine, steps, warmup):
    """Times `steps` LM iterations (profiling off), then repeats them from the
    same start with per-kernel HIP-event timing for the breakdown."""
    import torch
    import g2o_amd
    ba = g2o_amd.BundleAdjuster(prob, device=0, precision=precision, refine_steps=refine)
    if warmup > 0:
        ba.optimize(warmup)
        ba.set_state(prob.cams, prob.pts)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = ba.optimize(steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ba.set_state(prob.cams, prob.pts)
    ba.set_profiling(True)
    pstats = ba.optimize(steps)
    prof = ba.profile()
    ba.set_profiling(False)
    # full PCG iterations of the profiled run (the SpMV slot also holds the
    # cheap launches that only find the solve finished)
    prof["pcg_spmv_iterations"] = sum(st["iterations_linear_solver"] for st in pstats)
    return ba, stats, t1 - t0, prof


def cpu_baseline(prob, iters):
    """The oracle's LM run from the same start: (baseline record, its chi2
    after each iteration)."""
    from oracle import oracle
    n_threads = oracle.lib().orc_num_threads()
    t = time.perf_counter()
    ora = oracle.OracleBA(prob)
    setup = time.perf_counter() - t
    t = time.perf_counter()
    n, trace, times = ora.optimize(iters, solver=1, tol=1e-6)
    dt = time.perf_counter() - t
    return ({"value": n / dt, "unit": "LM iter/s", "cores": int(n_threads), "kind": "port",
             "sample": f"{n} LM iterations (FP64 explicit Schur + block-Jacobi PCG, tol 1e-6) of the same "
                       f"workload from the same start on the host; structure setup {setup:.1f} s excluded; "
                       f"chi2 after: {trace[-1][1]:.6e}"},
            [float(c) for c in trace[:n, 1]])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--refine", type=int, default=4)
    ap.add_argument("--cpu-iters", type=int, default=None, help="oracle LM iterations (default: --steps)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cams", type=int, default=1500)
    ap.add_argument("--points", type=int, default=500_000)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 plain device copies (out = -in, fp32) of MIB MiB each: the known-size "
                         "reference for the FETCH_SIZE / WRITE_SIZE counters, then exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(5):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": 4 * n, "bytes_written": 4 * n, "launches": 5}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the solver path is single-GPU (no multi-GPU run in this build)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np  # noqa: F401
    import g2o_amd
    from g2o_amd import datasets

    t = time.perf_counter()
    prob = datasets.make_walk_ba(seed=1002, num_cams=args.cams, num_pts=args.points)
    log(f"workload: C={prob.num_cams} P={prob.num_pts} E={prob.num_edges} ({time.perf_counter() - t:.1f} s)")

    ba32, s32, dt32, prof32 = run_gpu(prob, g2o_amd.FP32, args.refine, args.steps, args.warmup)
    log(f"fp32+refine: {len(s32)} iters in {dt32 * 1e3:.1f} ms; chi2 {s32[-1]['chi2']:.9e}")
    log(f"profile fp32: {prof32}")
    ba64, s64, dt64, prof64 = run_gpu(prob, g2o_amd.FP64, 0, args.steps, args.warmup)
    log(f"fp64: {len(s64)} iters in {dt64 * 1e3:.1f} ms; chi2 {s64[-1]['chi2']:.9e}")
    log(f"profile fp64: {prof64}")
    n32 = len(s32)
    value = n32 / dt32
    rel = abs(s32[-1]["chi2"] - s64[-1]["chi2"]) / s64[-1]["chi2"]

    info = ba32.structure_info()
    lin_cam_rows = int(info[10])
    # per-kernel rooflines of the FP32 path. `roofline` is the Schur window
    # kernel: the largest launch (a third of the bytes of a trial), whose
    # HIP-event duration is its kernel time. The PCG iteration kernel has the
    # largest total time, but its launches are ~12 us, so the events around
    # each one also time its launch latency (its rocprof kernel duration is in
    # the kernel-trace summary).
    roofs = {}
    for k in ("pcg_spmv", "schur_windows", "linearize", "refine", "chi2"):
        tot, cnt = prof32[k]
        if k == "pcg_spmv":
            cnt = prof32["pcg_spmv_iterations"]
        if not cnt:
            continue
        avg_s = tot / cnt / 1e3
        bytes_ = algorithmic_bytes(k, info, 4, lin_cam_rows)
        ach = bytes_ / avg_s
        roofs[k] = {"bound": "hbm", "kernel": KERNEL_SYMBOL[k], "achieved": round(ach / 1e9, 1),
                    "peak": PEAK_HBM / 1e9, "unit": "GB/s", "frac": round(ach / PEAK_HBM, 4),
                    "traffic": measured_traffic(k), "bytes_per_launch": int(bytes_),
                    "avg_launch_us": round(avg_s * 1e6, 2), "launches": cnt, "total_ms": round(tot, 3)}
    roof = roofs.get("schur_windows")
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "LM iter/s",
        "n_gpus": 1,
        "steps": n32,
        "warmup": args.warmup,
        "ms_per_step": round(dt32 / n32 * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 refinement, chi2, gain ratio, exp-map update)",
        "data": "synthetic",
        "config": {
            "workload": f"BA random walk {prob.num_cams} cams / {prob.num_pts} pts / {prob.num_edges} obs, "
                        "Huber 2, 5% outliers, seed 1002",
            "path": f"explicit Schur, FP32 blocks + FP64 refinement x{args.refine}",
            "fp64_path_iter_per_s": round(len(s64) / dt64, 3),
            "fp64_path_ms_per_step": round(dt64 / len(s64) * 1e3, 3),
            "speedup_vs_fp64_path": round(value / (len(s64) / dt64), 3),
            "rel_final_chi2_diff_vs_fp64": rel,
            "final_chi2_fp32_refine": s32[-1]["chi2"],
            "final_chi2_fp64": s64[-1]["chi2"],
            "lm_trials": [s["levenberg_iterations"] for s in s32],
            "pcg_iterations": [s["iterations_linear_solver"] for s in s32],
            "kernel_ms_fp32": {k: round(v[0], 3) for k, v in prof32.items() if k in g2o_amd.KERNEL_NAMES},
            "kernel_ms_fp64": {k: round(v[0], 3) for k, v in prof64.items() if k in g2o_amd.KERNEL_NAMES},
            "hschur_blocks": int(info[3]),
            "schur_windows": int(info[5]),
            "schur_offdiag_pairs": int(info[6]),
            "schur_partials": int(info[7]),
        },
        "roofline": roof,
        "roofline_by_kernel": roofs,
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"], chi_o = cpu_baseline(prob, args.cpu_iters or args.steps)
        # the same LM iterations on the oracle: chi2 error of the FP32+refine
        # path against the FP64 CPU oracle, per iteration
        k = min(len(chi_o), len(s32))
        if k:
            out["config"]["rel_chi2_diff_vs_oracle_per_iter"] = [abs(s32[i]["chi2"] - chi_o[i]) / chi_o[i]
                                                                 for i in range(k)]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
