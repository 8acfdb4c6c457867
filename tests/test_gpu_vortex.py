"""Low-density vortex accuracy (ex3_vortex_accuracy_2d) on the GPU.

- fortran/swf_vortex, the ex3 program over the C-ABI, reproduces the reference
  run's own output files byte for byte: err_table.dat (chained from fort.7 =
  1.0 1.0), fort.7, the printed norms, den/pre/m1/m2/den2/pre2.dat.
- The accuracy series 45^2 .. 360^2 (config of the north-star vortex row): the
  exact mode's 180^2 norms equal the reference's; FP64 stays within 1e-12 and
  FP32-bulk within 1e-6 of the exact run in relative L1 density (tolerances
  written here; FP32-full, with the projections and reconstructions in FP32 too, also
  within 1e-6); the order between 180^2 and 360^2 exceeds 4 in every mode.
"""
import hashlib
import json
import os
import subprocess

import pytest

from sweeping_fd_weno_amd import accuracy as A

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "ref_ex3_vortex180.json")
DRIVER = os.path.join(ROOT, "sweeping_fd_weno_amd", "fortran", "swf_vortex")


def test_vortex_driver_reproduces_reference_output_files(tmp_path):
    g = json.load(open(GOLD))
    (tmp_path / "fort.7").write_text("%13.5e%13.5e\n" % (1.0, 1.0))
    (tmp_path / "err_table.dat").write_text("")
    res = subprocess.run([DRIVER, "180", "0"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    assert (tmp_path / "err_table.dat").read_text() == g["err_table_dat"]
    assert (tmp_path / "fort.7").read_text() == g["fort7"]
    assert "\n".join(res.stdout.splitlines()[:4]) == g["stdout_head"]
    for fn, digest in g["file_sha256"].items():
        assert hashlib.sha256((tmp_path / fn).read_bytes()).hexdigest() == digest, fn
    line = [l for l in res.stdout.splitlines() if "total sweep" in l][0]
    assert int(line.split("=")[1]) == g["stats"]["total sweep"]


def test_vortex_driver_flux_limiter_reproduces_reference_output_files(tmp_path):
    """The same program with use_limiter = 1 (4th argument), against the reference
    built with that switch: output files, printed norms and limiter counts."""
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "ref_ex3_l1_vortex180.json")))
    (tmp_path / "fort.7").write_text("%13.5e%13.5e\n" % (1.0, 1.0))
    (tmp_path / "err_table.dat").write_text("")
    res = subprocess.run([DRIVER, "180", "0", "0", "1"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    assert (tmp_path / "err_table.dat").read_text() == g["err_table_dat"]
    assert (tmp_path / "fort.7").read_text() == g["fort7"]
    assert "\n".join(res.stdout.splitlines()[:4]) == g["stdout_head"]
    for fn, digest in g["file_sha256"].items():
        assert hashlib.sha256((tmp_path / fn).read_bytes()).hexdigest() == digest, fn
    line = [l for l in res.stdout.splitlines() if "# of limiter" in l][0]
    assert [int(v) for v in line.split()[-2:]] == [g["den_count"], g["pre_count"]]


def test_vortex_accuracy_series():
    g = json.load(open(GOLD))
    rows = A.vortex_series((45, 90, 180, 360), ("fp64_exact", "fp64", "fp32_bulk", "fp32_full"))
    for r in rows:
        print(r["precision"], r["n"], r["steps"], "L1 %.6e (%.2f) Linf %.6e (%.2f)" %
              (r["L1"], r["L1_order"], r["Linf"], r["Linf_order"]), r.get("L1_vs_exact"), r.get("Linf_vs_exact"))
    by = {(r["precision"], r["n"]): r for r in rows}
    ex = by[("fp64_exact", 180)]
    assert ex["steps"] == 9
    # the reference's printed norms (its exact solution evaluated by this package in binary64)
    assert abs(ex["L1"] - g["L1"]) <= 1e-12 * g["L1"] and abs(ex["Linf"] - g["Linf"]) <= 1e-12 * g["Linf"]
    for n in (45, 90, 180, 360):
        assert by[("fp64", n)]["L1_vs_exact"] <= 1e-12
        assert by[("fp32_bulk", n)]["L1_vs_exact"] <= 1e-6
        assert by[("fp32_full", n)]["L1_vs_exact"] <= 1e-6
    for prec in ("fp64_exact", "fp64", "fp32_bulk", "fp32_full"):
        assert by[(prec, 360)]["L1_order"] > 4.0


def test_vortex_accuracy_series_at_cfl_0_6():
    """The vortex row's stated CFL 0.6 (the reference's dx^(5/3) time-step rule with cfl
    0.6 in place of its 0.5): FP64 within 1e-12 and FP32-bulk within 1e-6 of the exact
    run in relative L1 density, fifth order between 180^2 and 360^2 in every mode, and the
    exact mode bitwise to the oracle on 45^2 and 90^2."""
    import numpy as np

    from oracle import oracle as O

    rows = A.vortex_series((45, 90, 180, 360), ("fp64_exact", "fp64", "fp32_bulk"), cfl=0.6)
    by = {(r["precision"], r["n"]): r for r in rows}
    for r in rows:
        print(r["precision"], r["n"], r["steps"], "L1 %.6e (%.2f)" % (r["L1"], r["L1_order"]), r.get("L1_vs_exact"))
    for n in (45, 90, 180, 360):
        assert by[("fp64", n)]["L1_vs_exact"] <= 1e-12
        assert by[("fp32_bulk", n)]["L1_vs_exact"] <= 1e-6
    for prec in ("fp64_exact", "fp64", "fp32_bulk"):
        assert by[(prec, 360)]["L1_order"] > 4.0
    for n in (45, 90):
        oc = O.vortex_config(n, cfl=0.6)
        o = O.Oracle(oc)
        o.set_state(O.init_vortex(oc))
        o.bc()
        rc, t, k = o.run(0.0, A.T_END)
        assert rc == 0 and k == by[("fp64_exact", n)]["steps"]
        assert np.array_equal(o.state()[3:-3, 3:-3, 0], by[("fp64_exact", n)]["rho"])
