"""Shock diffraction past a backward-facing step, the physics of ex6_shock_diffraction on
one domain: [0, 13] x [0, 11] with 13*total x 11*total cells, a step of total x 6*total
cells at the lower left, a Mach 5.09 shock entering from the left, cfl 0.9, T = 2.3:

    python -m sweeping_fd_weno_amd.step [--total 80 --mach 5.09 --tfinal 2.3]

Runs each precision mode to tfinal on the GPU from the program's initial data (no bc
before the first step, as there), reports cell-stage updates/s, the sweep statistics,
the smallest density and pressure of the final state, and the relative L1 / Linf
density difference from the FP64-exact run.  `--perturb` adds the exact run from an
initial state perturbed by a relative 1e-15, the problem's own sensitivity at tfinal.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import _lib as L
from .problems import init_step


def run_step(total: int, precision: str, mach: float = 5.09, tfinal: float = 2.3, chunk: int = 200,
             perturb: float = 0.0, max_steps: int = 1 << 40, progress=None) -> dict:
    cfg = L.step_config(total, precision, mach=mach)
    s = L.Solver(cfg)
    u = init_step(cfg)
    if perturb:
        u = u * (1.0 + perturb * np.random.default_rng(12345).standard_normal(u.shape))
    s.upload(u)
    s.sync()
    t, k, t0, rc = 0.0, 0, time.perf_counter(), 0
    while k < max_steps:
        try:
            t, kk = s.advance(t, tfinal, max_steps=min(chunk, max_steps - k))
        except L.SwfError as e:  # the program's stops: P(u_bar) <= 0, SWEEP FAIL, the pressure-sum check
            rc = e.code
            break
        k += kk
        if progress:
            progress(precision, k, t, time.perf_counter() - t0)
        if kk < chunk:
            break
    s.sync()
    el = time.perf_counter() - t0
    v = s.download()[3:-3, 3:-3]
    rho = v[:, :, 0].copy()
    p = (cfg.gamma - 1.0) * (v[:, :, 3] - 0.5 * (v[:, :, 1] ** 2 + v[:, :, 2] ** 2) / v[:, :, 0])
    st = s.stats()
    s.close()
    nx, ny = cfg.nx, cfg.ny
    return {"precision": precision, "total": total, "nx": nx, "ny": ny, "mach": mach, "rc": rc, "steps": k, "t": t,
            "seconds": el, "cell_stage_updates_per_s": nx * ny * 3 * k / el, "stats": st,
            "min_rho": float(rho.min()), "min_p": float(p.min()), "rho": rho}


def step_series(total=80, mach=5.09, tfinal=2.3, precisions=("fp64_exact", "fp64", "fp32_bulk"), perturb=True,
                progress=None):
    rows = []
    ref = None
    for prec in precisions:
        rows.append(run_step(total, prec, mach, tfinal, progress=progress))
        if prec == "fp64_exact":
            ref = rows[-1]["rho"]
    if perturb:
        r = run_step(total, "fp64_exact", mach, tfinal, perturb=1e-15, progress=progress)
        r["precision"] = "fp64_exact_perturbed_1e-15"
        rows.append(r)
    if ref is not None:
        for r in rows:
            if r["precision"] != "fp64_exact":
                d = np.abs(r["rho"] - ref)
                r["L1_vs_exact"] = float(d.sum() / np.abs(ref).sum())
                r["Linf_vs_exact"] = float(d.max() / np.abs(ref).max())
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--total", type=int, default=80)
    ap.add_argument("--mach", type=float, default=5.09)
    ap.add_argument("--tfinal", type=float, default=2.3)
    ap.add_argument("--precisions", default="fp64_exact,fp64,fp32_bulk")
    ap.add_argument("--no-perturb", action="store_true")
    a = ap.parse_args()

    def progress(prec, k, t, el):
        print(f"# {prec} step {k} t={t:.5g} {el:.1f}s", flush=True)

    rows = step_series(a.total, a.mach, a.tfinal, tuple(a.precisions.split(",")), not a.no_perturb, progress)
    for r in rows:
        r.pop("rho", None)
        print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
