"""The detonation program's exp on the device (exp_glibc, csrc/swf_ex7.h) against the host
C library's exp, which the reference build of ex7_detonation calls (`exp@GLIBC_2.29`,
ex7_detonation/main.f90:770): the restated algorithm, compiled for the host, is bitwise
equal to the library on seeded arguments over six ranges and special values; and its
table of 2^(j/128) = H (1 + T) is re-derived here from 80-digit decimal arithmetic (H and
T correctly rounded)."""
import os
import struct
import subprocess
from decimal import Decimal, getcontext
from fractions import Fraction

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("ex7exp") / "ex7_exp_check")
    subprocess.run(["g++", "-O2", "-std=gnu++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "tests", "emu"),
                    "-I", os.path.join(ROOT, "sweeping_fd_weno_amd", "csrc"), "-Wno-unknown-pragmas",
                    "-Wno-attributes", os.path.join(ROOT, "tests", "ex7_exp_check.cpp"), "-o", out, "-lm"],
                   check=True)
    return out


def _bits(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def test_table_is_the_correctly_rounded_powers_of_two(exe):
    words = [int(w, 16) for w in subprocess.run([exe, "table"], check=True, capture_output=True,
                                                  text=True).stdout.split()]
    getcontext().prec = 80
    ln2 = Decimal(2).ln()
    for j in range(128):
        e = Fraction((Decimal(j) / 128 * ln2).exp())
        h = float(e)  # correctly rounded
        t = float(e / Fraction(h) - 1)
        assert words[2 * j] == _bits(t), j
        assert words[2 * j + 1] == (_bits(h) - (j << 45)) % (1 << 64), j


def test_exp_bitwise_equal_to_the_host_library(exe):
    out = subprocess.run([exe, "300000"], check=True, capture_output=True, text=True).stdout
    for line in out.strip().splitlines():
        lo, hi, bad, n = line.split()
        assert int(bad) == 0, line
