"""The first-difference WENO forms used by the FP64 and FP32-bulk modes
(sweeping_fd_weno_amd/csrc/swf_weno.h), compiled for the host, against the
reference's weno_right / weno_left (weno5.f90:86-146) evaluated in binary128.
The error is measured relative to the size of the correction max|a(k)-a(0)|:
a consistent form stays at the rounding level of its own precision for any
offset a(0), for smooth data, jumps, 40-decade spreads and near-vacuum values.
Also pins the closed form R*L = I + dH e_E (L1 + L4) of the reference's
eigenvectors (cm carries +1e-15), which the fast modes use for the centre value."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# max |form - reference| / max|a(k) - a(0)|, written in the test
TOL = {"d64": 1e-15, "d32": 5e-7, "rl": 1e-20}


@pytest.fixture(scope="module")
def weno_check(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("wenod") / "weno_diff_check")
    subprocess.run(["g++", "-O2", "-std=gnu++17", "-fext-numeric-literals", "-ffp-contract=off",
                    "-I", os.path.join(ROOT, "sweeping_fd_weno_amd", "csrc"),
                    os.path.join(ROOT, "tests", "weno_diff_check.cpp"), "-o", exe, "-lquadmath", "-lm"], check=True)
    return exe


def test_difference_forms_match_binary128_reference(weno_check):
    out = subprocess.run([weno_check, "60000"], check=True, capture_output=True, text=True).stdout
    rows = [l.split() for l in out.splitlines() if l.strip()]
    assert len(rows) == 25
    for form, side, kind, n, err in rows:
        assert int(n) > 0
        assert float(err) <= TOL[form], f"{form} {side} {kind}: {err}"
