"""Mach 2000 jet, full domain (inflow for |y| < 0.05 at x = 0, zero-gradient at
the other faces), ex5_mach2000_jet's physics on [0, 1] x [-0.25, 0.25]:

    python -m sweeping_fd_weno_amd.jet [--nx 800 --ny 400 --cfl 0.5 --tfinal 0.001]

Runs each precision mode to tfinal on the GPU from the reference's initial data
(ambient rho = 0.5, p = 0.4127; ex5_mach2000_jet/main.f90:743-768, then `bc`),
reports cell-stage updates/s, the sweep statistics and the relative L1 / Linf
density difference from the FP64-exact run (bit-identical to the reference's
FP64 arithmetic).  `--perturb` adds the exact run from an initial state
perturbed by a relative 1e-15: its difference from the exact run is the
problem's own sensitivity at tfinal, the floor against which the FP64 and
FP32-bulk differences are read.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import _lib as L
from .problems import init_jet


def run_jet(nx: int, ny: int, precision: str, cfl: float = 0.5, tfinal: float = 0.001, chunk: int = 200,
            perturb: float = 0.0, progress=None) -> dict:
    cfg = L.jet_config(nx, ny, precision, cfl=cfl, half=False)
    s = L.Solver(cfg)
    u = init_jet(cfg)
    if perturb:
        u = u * (1.0 + perturb * np.random.default_rng(12345).standard_normal(u.shape))
    s.upload(u)
    s.bc()
    s.sync()
    t, k, t0 = 0.0, 0, time.perf_counter()
    while True:
        t, kk = s.advance(t, tfinal, max_steps=chunk)
        k += kk
        if progress:
            progress(precision, k, t, time.perf_counter() - t0)
        if kk < chunk:
            break
    s.sync()
    el = time.perf_counter() - t0
    rho = s.download()[3:-3, 3:-3, 0].copy()
    st = s.stats()
    s.close()
    return {"precision": precision, "nx": nx, "ny": ny, "steps": k, "t": t, "seconds": el,
            "cell_stage_updates_per_s": nx * ny * 3 * k / el, "stats": st, "rho": rho}


def jet_series(nx=800, ny=400, cfl=0.5, tfinal=0.001, precisions=("fp64_exact", "fp64", "fp32_bulk"),
               perturb=True, progress=None):
    rows = []
    ref = None
    for prec in precisions:
        rows.append(run_jet(nx, ny, prec, cfl, tfinal, progress=progress))
        if prec == "fp64_exact":
            ref = rows[-1]["rho"]
    if perturb:
        r = run_jet(nx, ny, "fp64_exact", cfl, tfinal, perturb=1e-15, progress=progress)
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
    ap.add_argument("--nx", type=int, default=800)
    ap.add_argument("--ny", type=int, default=400)
    ap.add_argument("--cfl", type=float, default=0.5)
    ap.add_argument("--tfinal", type=float, default=0.001)
    ap.add_argument("--precisions", default="fp64_exact,fp64,fp32_bulk")
    ap.add_argument("--no-perturb", action="store_true")
    a = ap.parse_args()

    def progress(prec, k, t, el):
        print(f"  {prec} step {k} t {t:.4e} {el:.1f}s", flush=True)

    rows = jet_series(a.nx, a.ny, a.cfl, a.tfinal, tuple(a.precisions.split(",")), not a.no_perturb, progress)
    for r in rows:
        print(json.dumps({k: v for k, v in r.items() if k != "rho"}), flush=True)


if __name__ == "__main__":
    main()
