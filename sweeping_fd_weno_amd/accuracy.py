"""Accuracy series of the low-density isentropic vortex (ex3_vortex_accuracy_2d).

    python -m sweeping_fd_weno_amd.accuracy [--grids 45,90,180,360] [--precisions fp64_exact,fp64,fp32_bulk]

For each grid and precision mode: run the vortex (mag = 10.0828, minimum density
~7.8e-15, minimum pressure ~1.7e-20) to T = 0.01 with the reference's time step
(dx^(5/3) cfl / alpha, cfl 0.5) on the GPU, form ex3's error norms against the
exact solution (ex3_vortex_accuracy_2d/main.f90:725-747: mean |rho - rho_exact| and
max |rho - rho_exact| over the interior) and their orders between successive
grids, and for the FP64 / FP32-bulk modes the L1 / Linf density difference from
the FP64-exact run (bit-identical to the reference's FP64 run).  The reference's
own program, output files included, is fortran/swf_vortex.
"""
from __future__ import annotations

import argparse
import json
import math

import numpy as np

from . import _lib as L
from .problems import _vortex_point, init_vortex

MAG = 10.0828
T_END = 0.01


def exact_density(cfg, t: float, mag: float = MAG) -> np.ndarray:
    """Exact density at the interior cell centres, indexed [j-1, i-1]."""
    out = np.empty((cfg.ny, cfg.nx))
    for j in range(1, cfg.ny + 1):
        y = cfg.ylo + j * cfg.dy - cfg.dy / 2.0
        for i in range(1, cfg.nx + 1):
            x = cfg.xlo + i * cfg.dx - cfg.dx / 2.0
            out[j - 1, i - 1] = _vortex_point(cfg.gamma, mag, x, y, t)[0]
    return out


def vortex_errors(rho: np.ndarray, cfg, t: float, mag: float = MAG):
    """ex3 `output`: er1 = sum over i (outer), j (inner) of |rho_exact - rho|, / nx / ny;
    eri = the maximum.  rho is the interior density indexed [j-1, i-1]."""
    d = np.abs(exact_density(cfg, t, mag) - rho)
    er1 = float(np.cumsum(d.T.ravel())[-1]) / cfg.nx / cfg.ny  # the reference's summation order
    return er1, float(d.max())


def run_vortex(n: int, precision: str, tfinal: float = T_END, cfl: float = 0.5) -> dict:
    cfg = L.vortex_config(n, precision, cfl=cfl)
    s = L.Solver(cfg)
    s.upload(init_vortex(cfg))
    s.bc()
    t, k = s.advance(0.0, tfinal)
    rho = s.download()[3:-3, 3:-3, 0].copy()
    stats = s.stats()
    s.close()
    er1, eri = vortex_errors(rho, cfg, t)
    return {"n": n, "precision": precision, "steps": k, "t": t, "L1": er1, "Linf": eri, "stats": stats,
            "rho": rho}


def vortex_series(grids=(45, 90, 180, 360), precisions=("fp64_exact", "fp64", "fp32_bulk"), cfl: float = 0.5):
    """Rows {n, precision, steps, t, L1, L1_order, Linf, Linf_order, [L1_vs_exact, Linf_vs_exact]}."""
    rows = []
    exact_rho = {}
    for prec in precisions:
        prev = None
        for n in grids:
            r = run_vortex(n, prec, cfl=cfl)
            r["L1_order"] = math.log(prev["L1"] / r["L1"]) / math.log(2.0) if prev else 0.0
            r["Linf_order"] = math.log(prev["Linf"] / r["Linf"]) / math.log(2.0) if prev else 0.0
            if prec == "fp64_exact":
                exact_rho[n] = r["rho"]
            elif n in exact_rho:
                d = np.abs(r["rho"] - exact_rho[n])
                r["L1_vs_exact"] = float(d.sum() / np.abs(exact_rho[n]).sum())
                r["Linf_vs_exact"] = float(d.max() / np.abs(exact_rho[n]).max())
            prev = r
            rows.append(r)
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--grids", default="45,90,180,360")
    ap.add_argument("--precisions", default="fp64_exact,fp64,fp32_bulk")
    ap.add_argument("--cfl", type=float, default=0.5)
    ap.add_argument("--json", action="store_true", help="one JSON line per row")
    a = ap.parse_args()
    rows = vortex_series(tuple(int(g) for g in a.grids.split(",")), tuple(a.precisions.split(",")), a.cfl)
    for r in rows:
        r = {k: v for k, v in r.items() if k != "rho"}
        if a.json:
            print(json.dumps(r), flush=True)
        else:
            extra = ""
            if "L1_vs_exact" in r:
                extra = f"  vs exact: L1 {r['L1_vs_exact']:.3e} Linf {r['Linf_vs_exact']:.3e}"
            print(f"{r['precision']:10s} {r['n']:4d} steps {r['steps']:3d}  L1 {r['L1']:.6e} ({r['L1_order']:5.2f})"
                  f"  Linf {r['Linf']:.6e} ({r['Linf_order']:5.2f})  sweeps {r['stats']['total_sweep']}{extra}",
                  flush=True)


if __name__ == "__main__":
    main()
