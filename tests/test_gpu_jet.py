"""The Mach 2000 jet program (ex5_mach2000_jet/main.f90) on the GPU against the
reference itself, built with its grid PARAMETERs set to 160 x 40 and run to T = 0.001
(612 steps; the sweep runs at 1 687 stages with 120 653 negative-neighbour pairs):
tests/golden/ref_ex5_jet160x40.json.

- FP64-exact: the log density / pressure / p-over-rho fields the program writes,
  bitwise (logarithms by the C library, as the program's), and every printed statistic.
- fortran/swf_jet, the ex5 program over the C-ABI, writes den.dat / pre.dat / m1.dat
  byte for byte and prints the same text (progress percentages, statistics).
- The same two at the program's own 800 x 200 grid to T = 5e-5
  (tests/golden/ref_ex5_jet800x200.json).
"""
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

import sweeping_fd_weno_amd as S
from test_oracle_golden import jet_log_fields

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "ref_ex5_jet160x40.json")
DRIVER = os.path.join(ROOT, "sweeping_fd_weno_amd", "fortran", "swf_jet")


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def test_jet160x40_exact_matches_reference_run():
    g = json.load(open(GOLD))["full"]
    cfg = S.jet_config(160, 40, "fp64_exact")
    with S.Solver(cfg) as s:
        s.upload(S.init_jet(cfg))
        s.bc()
        t, k = s.advance(0.0, 0.001)
        u = s.download()
        st = s.stats()
    print("steps", k, "sweeps", st["sweep_count"], flush=True)
    for name, v in jet_log_fields(u, 160, 40).items():
        assert _digest(v) == g["sha256"][name], name
    assert st["total_sweep"] == g["stats"]["total sweep"] and st["max_sweep"] == g["stats"]["max sweep"]
    assert st["total_neg"] == g["stats"]["total negative pairs"] and st["max_neg"] == g["stats"]["max negative pairs"]
    assert st["total_pos"] == g["stats"]["total positive neighbors"]
    assert st["max_pos"] == g["stats"]["max positive neighbors"]
    assert st["den_count"] == g["stats"]["den_count"] and st["pre_count"] == g["stats"]["pre_count"]
    assert k // 20 * 20 == g["last_progress_step"]


def test_jet_driver_reproduces_reference_output_files(tmp_path):
    g = json.load(open(GOLD))["full"]
    res = subprocess.run([DRIVER, "160", "40"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    for fn, digest in g["file_sha256"].items():
        assert hashlib.sha256((tmp_path / fn).read_bytes()).hexdigest() == digest, fn
    assert hashlib.sha256(res.stdout.encode()).hexdigest() == g["stdout_sha256"]


GOLD_SHIP = os.path.join(ROOT, "tests", "golden", "ref_ex5_jet800x200.json")


def test_jet_shipped_grid_exact_matches_reference_run():
    """The program's own 800 x 200 grid (cfl 0.25), to T = 5e-5 (time_terminal set in a
    build copy): fields bitwise and every statistic."""
    g = json.load(open(GOLD_SHIP))["ship"]
    cfg = S.jet_config(800, 200, "fp64_exact")
    with S.Solver(cfg) as s:
        s.upload(S.init_jet(cfg))
        s.bc()
        t, k = s.advance(0.0, 5e-5)
        u = s.download()
        st = s.stats()
    print("steps", k, "sweeps", st["sweep_count"], flush=True)
    for name, v in jet_log_fields(u, 800, 200).items():
        assert _digest(v) == g["sha256"][name], name
    for key, name in (("total sweep", "total_sweep"), ("max sweep", "max_sweep"), ("total negative pairs", "total_neg"),
                      ("max negative pairs", "max_neg"), ("total positive neighbors", "total_pos"),
                      ("max positive neighbors", "max_pos"), ("den_count", "den_count"), ("pre_count", "pre_count")):
        assert st[name] == g["stats"][key], key
    assert k // 20 * 20 == g["last_progress_step"]


def test_jet_driver_shipped_grid_output_files(tmp_path):
    """fortran/swf_jet at the program's grid, T = 5e-5: den.dat / pre.dat / m1.dat and the
    printout byte for byte."""
    g = json.load(open(GOLD_SHIP))["ship"]
    res = subprocess.run([DRIVER, "800", "200", "0", "0.00005"], cwd=tmp_path, capture_output=True, text=True,
                         timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    for fn, digest in g["file_sha256"].items():
        assert hashlib.sha256((tmp_path / fn).read_bytes()).hexdigest() == digest, fn
    assert hashlib.sha256(res.stdout.encode()).hexdigest() == g["stdout_sha256"]


def test_jet800x400_fp32_bulk_to_t1e_3():
    """The jet row at its stated size: full domain [0, 1] x [-0.25, 0.25], 800 x 400, cfl
    0.5, FP32-bulk to T = 0.001, against FP64-exact on the same grid.  The run holds
    positivity with no stop (ex5 prints and continues on a failed sweep; here a failure
    would show as a non-positive density), reaches T, and sweeps at nearly every stage.
    At this resolution the jet is chaotic: a relative 1e-15 perturbation of the initial
    state moves the exact run's density by 1.2e-2 in relative L1 at T (measured), so
    FP32-bulk is held to that scale (5e-2), not to rounding."""
    from sweeping_fd_weno_amd import jet as J

    print("jet 800x400 fp64_exact and fp32_bulk to T=1e-3", flush=True)
    ex = J.run_jet(800, 400, "fp64_exact")
    fb = J.run_jet(800, 400, "fp32_bulk")
    for r in (ex, fb):
        assert not S.time_continue(r["t"], 0.001), r["precision"]
        assert np.isfinite(r["rho"]).all() and r["rho"].min() > 0, r["precision"]
        assert r["stats"]["sweep_count"] > r["steps"], r["precision"]
    l1 = float(np.abs(fb["rho"] - ex["rho"]).sum() / np.abs(ex["rho"]).sum())
    print(f"steps {ex['steps']} / {fb['steps']}, sweeps {ex['stats']['sweep_count']} / {fb['stats']['sweep_count']}, "
          f"L1 {l1:.3e}, {fb['cell_stage_updates_per_s']:.3e} cell-stage/s")
    assert abs(fb["steps"] - ex["steps"]) <= 0.02 * ex["steps"]
    assert l1 <= 5e-2
