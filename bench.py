"""Benchmark: cell-stage updates/s on 2D Sedov 640^2, FP64 vs FP32-bulk at stated L1 error.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-full-run]
    python bench.py --workload sedov1d     (the 1D program's case: 800 cells to T = 1e-3)

One step = one SSP-RK3 step (three stages: WENO fluxes, RK update with the
positivity check, sweep when flagged) over the whole 640^2 domain, inputs
resident in HBM.  The timed region is K steps from t = 0 (the program's initial
data and initial bc), or the complete run to T = 1 (28 650 steps) when K reaches
it; `config.workload` names which.  The W warmup steps run on the same solver
before it (they capture its 16-step graph), then the initial state is uploaded
again.  Each precision mode is timed separately; `value` is the plain FP64 path.

Unless --no-full-run, every mode is also run from t = 0 to T = 1 (the accuracy
target: positivity held, the sweep statistics, the relative L1 density
difference from FP64-exact, which is bit-identical to the reference's FP64
program) and reported under `full_run_to_T1`.

cpu_baseline: the reference program itself (ex4_sedov_2d built as a one-rank
copy by oracle/Makefile, `oracle/_ref/bin/ex4_640w`: 640^2 from t = 0 to
T = 1.2e-5, three steps, timed by its own MPI_Wtime) on one host core; the C
oracle's single step is reported beside it as `port`.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-stage updates/s, 2D Sedov 640²: FP64 vs FP32-bulk at stated L1 error"
N = 640
T_END = 1.0
STEPS_TO_T1 = 28650         # SSP-RK3 steps of Sedov 640^2 (cfl 0.5) from t = 0 to T = 1
# Algorithmic flops per cell, counted from the reference formulas (+ - * / sqrt = 1,
# FMA = 2, structural zeros of the eigenvectors removed): one direction of wenofd
# (primitives, split, Roe average, eigenvectors, 2 x 5-point projections, 8 WENO
# reconstructions, back-projection) and the per-cell stage work (scheme, RK
# combination, findalpha, checkpositivity).
FLOPS_PER_CELL_DIR = 906.5
FLOPS_PER_CELL_UPDATE = 77.0
FLOPS_PER_CELL_STAGE = 2 * FLOPS_PER_CELL_DIR + FLOPS_PER_CELL_UPDATE
FP64_VALU_PEAK = 78.6e12
N_CU = 256

# The 1D program (ex2_sedov_1d): 800 cells on [-2, 2], cfl 0.6, ambient E 1e-12,
# swp_num 2, T = 1e-3 (2617 steps).  Algorithmic flops per cell and stage, counted
# the same way: split fluxes 26, local eigenvectors 25, 2 x 5-point projections of
# 3 fields 150, 6 WENO reconstructions 6 x 67, back-projection 30, scheme 12, RK 15,
# findalpha and checkpositivity 13.
N1D = 800
T1D = 1e-3
STEPS_1D = 2617
FLOPS_PER_CELL_STAGE_1D = 673.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=32)
    ap.add_argument("--no-full-run", action="store_true", help="skip the runs to T = 1")
    ap.add_argument("--modes", default="fp64,fp32_bulk,fp32_full,fp64_exact")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--workload", default="sedov2d", choices=["sedov2d", "sedov1d"])
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="run one plain device copy of this size first (calibrates HBM counters under rocprofv3)")
    return ap.parse_args()


def _fresh(s, u0):
    """The program's start: initial data, then its bc (ex4_sedov_2d/main.f90:1186)."""
    s.upload(u0)
    s.bc()
    s.sync()


def run_mode(S, prec: str, steps: int, warmup: int):
    cfg = S.sedov_config(N, prec)
    u0 = S.init_sedov(cfg)
    s = S.Solver(cfg)
    _fresh(s, u0)
    s.advance(0.0, T_END, max_steps=warmup)  # also captures the solver's step graph (first advance)
    _fresh(s, u0)
    st0 = s.stats()
    t0 = time.perf_counter()
    t, k = s.advance(0.0, T_END, max_steps=steps)
    s.sync()
    wall = time.perf_counter() - t0
    rho = s.download()[3:-3, 3:-3, 0].copy()
    st1 = s.stats()
    xms = s.time_kernel("xflux", 20)
    yms = s.time_kernel("ystage", 20)
    s.close()
    window = {key: st1[key] - st0[key] for key in ("sweep_count", "total_sweep", "total_neg", "total_pos",
                                                     "den_count", "pre_count")}
    return dict(steps=k, t_reached=t, seconds=wall, value=N * N * 3 * k / wall, ms_per_step=1e3 * wall / max(k, 1),
                xflux_kernel_ms=xms, ystage_kernel_ms=yms, window_stats=window), rho


def run_full(S, prec: str):
    """From t = 0 to T = 1 in one call (graph captured beforehand): rate, statistics,
    final density, smallest density and pressure."""
    cfg = S.sedov_config(N, prec)
    u0 = S.init_sedov(cfg)
    s = S.Solver(cfg)
    _fresh(s, u0)
    s.advance(0.0, T_END, max_steps=16)  # also captures the solver's step graph (first advance)
    _fresh(s, u0)
    z = s.stats()
    t0 = time.perf_counter()
    t, k = s.advance(0.0, T_END)
    s.sync()
    wall = time.perf_counter() - t0
    u = s.download()[3:-3, 3:-3]
    st = {key: v - z[key] if key in ("sweep_count", "den_count", "pre_count", "total_sweep", "total_neg",
                                     "total_pos", "stage_sweeps", "steps") else v for key, v in s.stats().items()}
    s.close()
    pres = 0.4 * (u[..., 3] - 0.5 * u[..., 1] ** 2 / u[..., 0] - 0.5 * u[..., 2] ** 2 / u[..., 0])
    return dict(steps=k, t_reached=t, seconds=wall, value=N * N * 3 * k / wall, stats=st,
                min_rho=float(u[..., 0].min()), min_p=float(pres.min())), u[..., 0].copy()


def cpu_baseline_reference():
    """The reference program (ex4_sedov_2d, one-rank build copy, oracle/_ref/bin/ex4_640w):
    640^2 from t = 0 to 1.2e-5, timed by its own MPI_Wtime around the time loop."""
    exe = os.path.join(ROOT, "oracle", "_ref", "bin", "ex4_640w")
    if not os.path.exists(exe):
        return None
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run(["bash", "-c", f"ulimit -s unlimited; exec {exe} {d}"], cwd=d, capture_output=True,
                           text=True, timeout=600)
    line = [l for l in r.stdout.splitlines() if l.strip().startswith("CPU Time")]
    if r.returncode != 0 or not line:
        return None
    secs, steps = float(line[0].split()[2]), int(line[0].split()[3])
    return {"value": N * N * 3 * steps / secs, "unit": "cell-stage updates/s", "cores": 1, "kind": "reference",
            "sample": f"{steps} SSP-RK3 steps of Sedov 640^2 from t=0 to 1.2e-5 by the reference program "
                      f"ex4_sedov_2d (one-rank build copy, flang -O3 -fdefault-real-8), {secs:.1f} s by its own "
                      f"MPI_Wtime timer, single thread"}


def cpu_baseline(steps: int):
    from oracle import oracle as O  # the checker, timed as the CPU baseline
    oc = O.sedov_config(N)
    o = O.Oracle(oc)
    o.set_state(O.init_sedov(oc))
    o.bc()
    t0 = time.perf_counter()
    rc, t, k = o.run(0.0, T_END, max_steps=steps)
    wall = time.perf_counter() - t0
    return {"value": N * N * 3 * k / wall, "unit": "cell-stage updates/s", "cores": 1, "kind": "port",
            "sample": f"{k} SSP-RK3 step(s) of Sedov 640^2 from t=0 by the C oracle (FP64 with binary128 sites "
                      f"as the reference build), single thread, {wall:.1f} s"}


def cpu_baselines(steps: int):
    port = cpu_baseline(steps)
    ref = cpu_baseline_reference()
    if ref is None:
        return port
    ref["port"] = port
    return ref


def traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter runs."""
    p = os.path.join(ROOT, "profiles", "stage_traffic.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get(kernel, {}).get("bytes_per_launch")
    return None


def run_line(S, prec: str, warmup: int):
    cfg = S.sedov1d_config(N1D, prec)
    u0 = S.init_sedov1d(cfg)
    with S.Solver1D(cfg) as s:
        if warmup:
            s.upload(u0)
            s.advance(0.0, T1D, max_steps=warmup)
        s.upload(u0)
        t0 = time.perf_counter()
        t, k = s.advance(0.0, T1D)
        wall = time.perf_counter() - t0
        ms = s.last_ms()
        rho = s.download()[3:-3, 0].copy()
        stats = s.stats()
    return dict(steps=k, t_end=t, seconds=wall, kernel_ms=ms, value=N1D * 3 * k / wall,
                ms_per_step=1e3 * wall / max(k, 1), stats=stats), rho


def cpu_baseline_line(steps: int):
    from oracle import oracle as O  # the checker, timed as the CPU baseline
    oc = O.sedov1d_config(N1D)
    o = O.Oracle1D(oc)
    o.set_state(O.init_sedov1d(oc))
    t0 = time.perf_counter()
    t, k = o.run(0.0, T1D, steps)
    wall = time.perf_counter() - t0
    return {"value": N1D * 3 * k / wall, "unit": "cell-stage updates/s", "cores": 1, "kind": "port",
            "sample": f"{k} SSP-RK3 steps of the 1D case from t=0 by the C oracle (FP64 with binary128 sites as "
                      f"the reference build), single thread, {wall:.1f} s"}


def main_line(a, S):
    modes = a.modes.split(",")
    res, rho = {}, {}
    for m in modes:
        res[m], rho[m] = run_line(S, m, min(a.warmup, 50))
    err = {}
    if "fp64_exact" in rho:
        for m in modes:
            if m != "fp64_exact":
                err[m] = float(np.abs(rho[m] - rho["fp64_exact"]).sum() / np.abs(rho["fp64_exact"]).sum())
    head = res.get("fp64") or res[modes[0]]
    ach = FLOPS_PER_CELL_STAGE_1D * (N1D + 6) * 3 * head["steps"] / (head["kernel_ms"] * 1e-3)
    out = {
        "metric": "cell-stage updates/s, 1D Sedov 800 cells to T=1e-3",
        "value": head["value"], "unit": "cell-stage updates/s", "n_gpus": 1, "steps": head["steps"],
        "warmup": min(a.warmup, 50), "ms_per_step": head["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "sedov1d_800_to_T1e-3", "cells": N1D, "cfl": 0.6, "swp_num": 2, "t_end": T1D,
                   "precision": "fp64", "parallelism": "one workgroup, whole time loop in one launch"},
        "roofline": {"bound": "valu", "achieved": ach / 1e12, "peak": FP64_VALU_PEAK / 1e12, "unit": "TFLOP/s",
                     "frac": ach / FP64_VALU_PEAK, "frac_of_one_cu": ach / (FP64_VALU_PEAK / N_CU),
                     "traffic": None, "kernel": "k_run1d<fp64> (every stage of every step)",
                     "kernel_ms": head["kernel_ms"]},
        "modes": res,
        "l1_density_rel_vs_fp64_exact": err,
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_line(a.cpu_steps if a.cpu_steps > 1 else 600)
    print(json.dumps(out))


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not provided (see DESIGN.md, "
                                   "sharding)", "n_gpus": a.gpus}))
        sys.exit(2)
    import sweeping_fd_weno_amd as S

    if a.workload == "sedov1d":
        main_line(a, S)
        return

    if a.copy_probe_mb:
        import torch

        x = torch.ones(a.copy_probe_mb * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
    modes = a.modes.split(",")
    res, rho = {}, {}
    for m in modes:
        res[m], rho[m] = run_mode(S, m, a.steps, a.warmup)
    err = {}
    if "fp64_exact" in rho:
        ref = rho["fp64_exact"]
        for m in modes:
            if m != "fp64_exact":
                err[m] = float(np.abs(rho[m] - ref).sum() / np.abs(ref).sum())
    full, frho = {}, {}
    if not a.no_full_run:
        for m in modes:
            full[m], frho[m] = run_full(S, m)
        if "fp64_exact" in frho:
            ref = frho["fp64_exact"]
            for m in modes:
                full[m]["l1_density_rel_vs_fp64_exact"] = float(np.abs(frho[m] - ref).sum() / np.abs(ref).sum())
        for m in ("fp32_bulk", "fp32_full"):
            if m in full and "fp64" in full:
                full[m + "_speedup_vs_fp64"] = full[m]["value"] / full["fp64"]["value"]
    head = res.get("fp64") or res[modes[0]]
    # dominant kernel: the stage kernel (y-pass of wenofd + per-cell stage work)
    ach_y = (FLOPS_PER_CELL_DIR + FLOPS_PER_CELL_UPDATE) * N * N / (head["ystage_kernel_ms"] * 1e-3)
    ach_x = FLOPS_PER_CELL_DIR * N * N / (head["xflux_kernel_ms"] * 1e-3)
    ach_run = FLOPS_PER_CELL_STAGE * N * N * 3 * head["steps"] / head["seconds"]
    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "cell-stage updates/s",
        "n_gpus": 1,
        "steps": head["steps"],
        "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": ("sedov2d_640x640_to_T1" if not S.time_continue(head["t_reached"], T_END)
                                else f"sedov2d_640x640_steps0-{head['steps']}"),
                   "grid": [N, N], "cfl": 0.5, "t_start": 0.0, "t_reached": head["t_reached"],
                   "precision": "fp64", "parallelism": "single GPU, whole domain"},
        "roofline": {"bound": "valu", "achieved": ach_y / 1e12, "peak": FP64_VALU_PEAK / 1e12, "unit": "TFLOP/s",
                     "frac": ach_y / FP64_VALU_PEAK, "traffic": traffic("k_ystage<fp64>"),
                     "kernel": "k_ystage<fp64> (y-pass of wenofd + scheme/RK/positivity/bc/alpha)",
                     "kernel_ms": head["ystage_kernel_ms"],
                     "algorithmic_flops_per_launch": (FLOPS_PER_CELL_DIR + FLOPS_PER_CELL_UPDATE) * N * N,
                     "other": {"k_xflux<fp64>": {"achieved": ach_x / 1e12, "frac": ach_x / FP64_VALU_PEAK,
                                                 "kernel_ms": head["xflux_kernel_ms"],
                                                 "traffic": traffic("k_xflux<fp64>")},
                               "whole_run": {"achieved": ach_run / 1e12, "frac": ach_run / FP64_VALU_PEAK}}},
        "modes": res,
        "l1_density_rel_vs_fp64_exact": err,
    }
    for m in ("fp32_bulk", "fp32_full"):
        if m in res and "fp64" in res:
            out[m + "_speedup_vs_fp64"] = res[m]["value"] / res["fp64"]["value"]
    if full:
        out["full_run_to_T1"] = full
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baselines(a.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
