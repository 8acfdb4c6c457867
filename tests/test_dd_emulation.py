"""The library's emulation of the reference's binary128 expression sites
(sweeping_fd_weno_amd/csrc/swf_dd.h + swf_q128.h), compiled for the host, against
__float128 evaluation of the same expressions on adversarial inputs (gaussian,
near-equal groups, 60-decade exponent spread, small dyadic values): every site
must round to the identical binary64 result, and so must the exact soft-binary128
path on its own.  The WENO weight site is also checked with 1d-12 (the step
program's weno5.f90) on smoothness indicators scaled near that constant."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def dd_check(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("dd") / "dd_check")
    subprocess.run(["g++", "-O2", "-std=gnu++17", "-fext-numeric-literals", "-ffp-contract=off",
                    "-I", os.path.join(ROOT, "sweeping_fd_weno_amd", "csrc"),
                    os.path.join(ROOT, "tests", "dd_check.cpp"), "-o", exe, "-lquadmath", "-lm"], check=True)
    return exe


def test_all_binary128_sites_round_identically(dd_check):
    out = subprocess.run([dd_check, "400000"], check=True, capture_output=True, text=True).stdout
    rows = [l.split() for l in out.splitlines() if l and not l.startswith("by input")]
    assert len(rows) == 10  # alpha12: the step program's 1d-12 in the weights
    for name, bad, total in rows:
        assert int(total) > 0
        assert int(bad) == 0, f"{name}: {bad} of {total} differ from binary128"
