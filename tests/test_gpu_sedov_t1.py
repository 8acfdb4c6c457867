"""The benchmark's accuracy target: 2D Sedov 640^2 on [0, 1.3]^2 to T = 1 (28 650 steps)
in the three precision modes, from the program's initial data and initial bc
(ex4_sedov_2d/main.f90:1121-1187).

- Positivity held throughout: every mode reaches T = 1 with no stop (the program stops
  on P(u_bar) <= 0 and on SWEEP FAIL, main.f90:1447-1454,1747-1752; both are enabled),
  and the final state has positive density and pressure.
- The sweep statistics (triggers, negative pairs, positive neighbours, limiter counts)
  are identical in every mode.
- Relative L1 density difference from the FP64-exact run (bit-identical to the
  reference's FP64 program; see test_gpu_reference_runs.py for the 640^2 window against
  the reference itself): at most 1e-12 for FP64 and 1e-6 for FP32-bulk and FP32-full.
"""
import numpy as np
import pytest

import sweeping_fd_weno_amd as S

pytestmark = pytest.mark.gpu
STATS = ("sweep_count", "den_count", "pre_count", "total_sweep", "max_sweep", "total_neg", "max_neg", "total_pos",
         "max_pos")


def _run_to_t1(prec):
    cfg = S.sedov_config(640, prec)
    with S.Solver(cfg) as s:
        s.upload(S.init_sedov(cfg))
        s.bc()
        t, k = 0.0, 0
        while S.time_continue(t, 1.0):
            t, kk = s.advance(t, 1.0, max_steps=4096)
            k += kk
            print(f"  {prec}: {k} steps, t = {t:.6f}", flush=True)
            if kk == 0:
                break
        u = s.download()
        st = s.stats()
    return t, k, u, st


def test_sedov640_to_t1_all_modes():
    runs = {p: _run_to_t1(p) for p in ("fp64_exact", "fp64", "fp32_bulk", "fp32_full")}
    ex = runs["fp64_exact"]
    assert not S.time_continue(ex[0], 1.0) and ex[1] == 28650
    rho_ex = ex[2][3:-3, 3:-3, 0]
    for p, (t, k, u, st) in runs.items():
        v = u[3:-3, 3:-3]
        pres = 0.4 * (v[..., 3] - 0.5 * v[..., 1] ** 2 / v[..., 0] - 0.5 * v[..., 2] ** 2 / v[..., 0])
        assert np.isfinite(v).all() and v[..., 0].min() > 0 and pres.min() > 0, p
        assert not S.time_continue(t, 1.0), p
        for key in STATS:
            assert st[key] == ex[3][key], (p, key)
        l1 = float(np.abs(v[..., 0] - rho_ex).sum() / np.abs(rho_ex).sum())
        print(f"{p}: {k} steps, t = {t!r}, L1 rho vs exact {l1:.3e}, stats {st}")
        if p == "fp64":
            assert l1 <= 1e-12
        elif p in ("fp32_bulk", "fp32_full"):
            assert l1 <= 1e-6
    assert ex[3]["sweep_count"] > 0
