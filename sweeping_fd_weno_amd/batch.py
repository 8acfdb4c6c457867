"""Batched ensemble: many independent 2D solves of uneven sizes advanced together.

    python -m sweeping_fd_weno_amd.batch [--steps 200 --precision fp32_bulk --reference fp64_exact]

The ensemble is drawn with splitmix64 from seed 20261203, the families in turn: the
Sedov corners (ex4_sedov_2d's problem on [0, 1.3]^2) take a corner energy E0 ~ U[0.1,
1.0] in place of 0.244816, the low-density vortices (ex3_vortex_accuracy_2d's on
[0, 10]^2) a strength mag ~ U[5, 10.0828], the shock diffractions past a backward step
(ex6_shock_diffraction's problem on [0, 13] x [0, 11]) a Mach number ~ U[3, 6].  The
j-th instance of a family has the grid SIZES[j % 8] (Sedov, vortex) or 13T x 11T with
T = STEP_TOTALS[j % 8] (step).

All instances run in one swf_batch (one launch of each stage kernel per RK stage
for the Sedov and vortex instances, one for the step instances, whose wenofd
differs), each with its own time step and sweep, to its own final time (Sedov T = 1,
vortex T = 0.01, step T = 2.3) or `steps` steps.  The tile schedule is picked by
swf_batch_tune.  Reported: aggregate cell-stage updates/s (sum over instances of
nx*ny*3*steps over the wall time, sweeps included) and, against a batch in the
reference precision, each instance's relative L1 density difference.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import _lib as L
from .problems import init_sedov, init_step, init_vortex

SEED = 20261203
SIZES = [(90, 90), (180, 180), (260, 220), (360, 360), (520, 440), (640, 640), (780, 660), (1040, 880)]
STEP_TOTALS = [8, 16, 24, 32, 40, 48, 64, 80]  # 104 x 88 .. 1040 x 880
MASK = (1 << 64) - 1


def splitmix64(state: int):
    """One splitmix64 step: (next state, 64-bit output)."""
    state = (state + 0x9E3779B97F4A7C15) & MASK
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return state, z ^ (z >> 31)


def uniform(state: int, lo: float, hi: float):
    state, z = splitmix64(state)
    return state, lo + (hi - lo) * ((z >> 11) * (1.0 / (1 << 53)))


def ensemble(precision: str, n_sedov: int = 16, n_vortex: int = 16, n_step: int = 16, seed: int = SEED):
    """[(label, cfg, initial state, tfinal)] in instance order (Sedov, vortex, step in
    turn while each family has instances left; each family takes its sizes in turn)."""
    state = seed
    out = []
    left = [n_sedov, n_vortex, n_step]
    done = [0, 0, 0]
    k = 0
    while any(left):
        fam = k % 3
        k += 1
        if not left[fam]:
            continue
        j = done[fam]
        left[fam] -= 1
        done[fam] += 1
        if fam == 0:
            nx, ny = SIZES[j % len(SIZES)]
            state, e0 = uniform(state, 0.1, 1.0)
            cfg = L._config(nx=nx, ny=ny, xlo=0.0, ylo=0.0, dx=1.3 / nx, dy=1.3 / ny, gamma=1.4, cfl=0.5, eps=1e-13,
                            eps2=1e-8, delta=1e-12, bc=L.BC_SEDOV, use_limiter=2, swp_num=20000, pbar_check=1,
                            fail_stop=1, precision=L._prec(precision), device=0)
            out.append((f"sedov{nx}x{ny}_E{e0:.3f}", cfg, init_sedov(cfg, e_corner_num=e0), 1.0))
        elif fam == 1:
            nx, ny = SIZES[j % len(SIZES)]
            state, mag = uniform(state, 5.0, 10.0828)
            cfg = L._config(nx=nx, ny=ny, xlo=0.0, ylo=0.0, dx=10.0 / nx, dy=10.0 / ny, gamma=1.4, cfl=0.5,
                            eps=1e-13, eps2=1e-8, delta=1e-12, bc=L.BC_PERIODIC, use_limiter=2, swp_num=20000,
                            alpha_in_scheme=1, roe_eps_double=1, ydens_plain=1, eps2_quad_literal=1, findalpha_ex3=1,
                            dt_rule_dx53=1, poscheck_le=1, precision=L._prec(precision), device=0)
            out.append((f"vortex{nx}x{ny}_mag{mag:.3f}", cfg, init_vortex(cfg, mag=mag), 0.01))
        else:
            total = STEP_TOTALS[j % len(STEP_TOTALS)]
            state, mach = uniform(state, 3.0, 6.0)
            cfg = L.step_config(total, precision, mach=mach)
            out.append((f"step{cfg.nx}x{cfg.ny}_M{mach:.3f}", cfg, init_step(cfg), 2.3))
    return out


def run_batch(precision: str, steps: int, tune: bool = True, n_sedov: int = 16, n_vortex: int = 16,
              n_step: int = 16) -> dict:
    ens = ensemble(precision, n_sedov, n_vortex, n_step)
    bt = L.Batch([c for _, c, _, _ in ens])
    for i, (_, c, u, _) in enumerate(ens):
        bt.instance(i).upload(u)
        if c.bc != L.BC_STEP:  # the step program applies no bc before its first step
            bt.instance(i).bc()
    sched, ms = bt.tune(10) if tune else (1, None)
    for i in range(bt.n):
        bt.instance(i).sync()
    t0 = time.perf_counter()
    t, n, st = bt.advance(np.zeros(bt.n), np.array([tf for *_, tf in ens]), max_steps=steps)
    el = time.perf_counter() - t0
    cells = np.array([c.nx * c.ny for _, c, _, _ in ens], dtype=np.float64)
    rows = []
    for i, (label, c, _, tf) in enumerate(ens):
        s = bt.instance(i).stats()
        rows.append({"label": label, "nx": c.nx, "ny": c.ny, "steps": int(n[i]), "t": float(t[i]),
                     "status": int(st[i]), "sweeps": s["sweep_count"],
                     "rho": bt.instance(i).download()[3:-3, 3:-3, 0].copy()})
    bt.close()
    return {"precision": precision, "instances": len(ens), "cells": int(cells.sum()), "seconds": el,
            "cell_stage_updates_per_s": float((cells * 3 * n).sum() / el), "schedule": sched, "schedule_ms": ms,
            "rows": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--precision", default="fp32_bulk")
    ap.add_argument("--reference", default="fp64_exact", help="precision of the comparison batch ('' for none)")
    a = ap.parse_args()
    r = run_batch(a.precision, a.steps)
    ref = run_batch(a.reference, a.steps) if a.reference else None
    worst = 0.0
    for i, row in enumerate(r["rows"]):
        if ref is not None and ref["rows"][i]["steps"] == row["steps"]:
            d = ref["rows"][i]["rho"]
            row["l1_vs_reference"] = float(np.abs(row["rho"] - d).sum() / np.abs(d).sum())
            worst = max(worst, row["l1_vs_reference"])
    for row in r["rows"]:
        print(json.dumps({k: v for k, v in row.items() if k != "rho"}), flush=True)
    summary = {k: v for k, v in r.items() if k != "rows"}
    if ref is not None:
        summary["reference"] = {"precision": ref["precision"], "seconds": ref["seconds"],
                                "cell_stage_updates_per_s": ref["cell_stage_updates_per_s"]}
        summary["worst_l1_vs_reference"] = worst
    print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    main()
