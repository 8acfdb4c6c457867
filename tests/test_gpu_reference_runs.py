"""The GPU exact mode and the Fortran ex4 driver against runs of the reference's own
MPI programs, built as one-rank copies (oracle/Makefile: `use MPI` -> `include
'mpif.h'`, cut_x = cut_y = 1; fixtures from tests/golden/make_golden.py ex4 ex6).

- ex4_sedov_2d (the benchmark program): 64^2 to T = 0.002, 128^2 to T = 0.01 and the
  benchmark grid 640^2 to T = 1.2e-5.  FP64-exact from the program's initial data (and
  its initial bc, ex4_sedov_2d/main.f90:1186): den/pre/m1/m2 bitwise, step count and
  every printed statistic equal.
- fortran/swf_driver (ex4's parameter surface and time loop over the iso_c_binding
  module): den.dat / pre.dat / m1.dat / m2.dat byte for byte, and its printout equal to
  the program's apart from the wall-clock `CPU Time` line.
- ex6_shock_diffraction: total = 8 (104 x 88) to the shipped T = 2.3 (150 steps, 190
  sweeps, 57 density sweeps) and total = 16 (208 x 176) to T = 0.5: den/pre as the
  program writes them (the step blanked) bitwise, step count and statistics equal.
"""
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

import sweeping_fd_weno_amd as S

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
DRIVER = os.path.join(ROOT, "sweeping_fd_weno_amd", "fortran", "swf_driver")
EX4 = {"n64": (64, 0.002), "n128": (128, 0.01), "n640w": (640, 1.2e-5)}
EX6 = {"total8": (8, 2.3), "total16": (16, 0.5)}


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def _stats_as_printed(s):
    return {"total sweep": s["total_sweep"], "max sweep": s["max_sweep"], "total negative pairs": s["total_neg"],
            "max negative pairs": s["max_neg"], "total positive neighbors": s["total_pos"],
            "max positive neighbors": s["max_pos"], "den_count": s["den_count"], "pre_count": s["pre_count"]}


def _columns(u, gamma=1.4, blank_total=0):
    """Value columns of the programs' field files, (j outer, i inner) over the interior."""
    U = u[3:-3, 3:-3, :]
    rho, m1, m2, e = (U[:, :, c].copy() for c in range(4))
    pre = (gamma - 1.0) * (e - 0.5 * m1 ** 2 / rho - 0.5 * m2 ** 2 / rho)
    if blank_total:
        jj, ii = np.meshgrid(np.arange(1, U.shape[0] + 1), np.arange(1, U.shape[1] + 1), indexing="ij")
        blank = (ii < blank_total) & (jj < 6 * blank_total)
        rho[blank] = 0.0
        pre[blank] = 0.0
    return {"den": rho.reshape(-1), "pre": pre.reshape(-1), "m1": m1.reshape(-1), "m2": m2.reshape(-1)}


@pytest.mark.parametrize("case", list(EX4))
def test_sedov_exact_matches_ex4_reference_run(case):
    g = json.load(open(os.path.join(GOLD, "ref_ex4_sedov.json")))[case]
    n, T = EX4[case]
    print(f"ex4 {case}: {n}^2 to T={T} fp64_exact", flush=True)
    cfg = S.sedov_config(n, "fp64_exact")
    with S.Solver(cfg) as s:
        s.upload(S.init_sedov(cfg))
        s.bc()
        t, k = s.advance(0.0, T)
        u = s.download()
        st = s.stats()
    assert k == g["steps"]
    cols = _columns(u)
    for nm in ("den", "pre", "m1", "m2"):
        assert _digest(cols[nm]) == g["sha256"][nm], nm
    assert _stats_as_printed(st) == g["stats"]


@pytest.mark.parametrize("case", ["n64", "n128"])
def test_sedov_driver_reproduces_ex4_output_files(case, tmp_path):
    g = json.load(open(os.path.join(GOLD, "ref_ex4_sedov.json")))[case]
    n, T = EX4[case]
    print(f"swf_driver {case}", flush=True)
    res = subprocess.run([DRIVER, str(n), str(n), str(T), "0", str(tmp_path)], capture_output=True, text=True,
                         timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    for fn, want in g["file_sha256"].items():
        assert hashlib.sha256((tmp_path / fn).read_bytes()).hexdigest() == want, fn
    lines = res.stdout.splitlines()
    timer = [l for l in lines if l.strip().startswith("CPU Time")]
    assert len(timer) == 1 and int(timer[0].split()[3]) == g["steps"]
    assert "\n".join(l for l in lines if not l.strip().startswith("CPU Time")) + "\n" == g["printout_minus_timer"]


@pytest.mark.parametrize("case", list(EX6))
def test_step_exact_matches_ex6_reference_run(case):
    g = json.load(open(os.path.join(GOLD, "ref_ex6_step.json")))[case]
    total, T = EX6[case]
    print(f"ex6 {case}: total={total} to T={T} fp64_exact", flush=True)
    cfg = S.step_config(total, "fp64_exact")
    with S.Solver(cfg) as s:
        s.upload(S.init_step(cfg))  # ex6 has no bc after initdata
        t, k = s.advance(0.0, T)
        u = s.download()
        st = s.stats()
    assert k == g["steps"]
    cols = _columns(u, blank_total=total)
    for nm in ("den", "pre"):
        assert _digest(cols[nm]) == g["sha256"][nm], nm
    assert _stats_as_printed(st) == g["stats"]
