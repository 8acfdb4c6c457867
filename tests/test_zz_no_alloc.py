"""swf_weno_points_noalloc, the entry point that allocates no device memory (its
staging lives in one array of the library per GPU), against swf_weno_points: bitwise
equal results in every precision mode, on two calls in a row, the second with fewer
stencils (the array keeps the first call's contents), and the reference module's own
values in the exact mode.  Sizes above the maximum are rejected before any HIP call."""
import ctypes
import os

import numpy as np
import pytest

import sweeping_fd_weno_amd as S
from sweeping_fd_weno_amd import _lib as L

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _cases():
    g = np.load(os.path.join(GOLD, "ref_weno.npz"))
    rng = np.random.default_rng(20261205)
    big = np.concatenate([g["a"]] * 9 + [rng.standard_normal((L.WENO_NOALLOC_MAX - 9 * len(g["a"]), 5))])
    small = rng.standard_normal((1237, 5)) * 10.0 ** rng.integers(-8, 8, (1237, 1))
    return g, big, small


@pytest.mark.gpu
@pytest.mark.parametrize("prec,eps12", [("fp64_exact", False), ("fp64", False), ("fp32_bulk", False),
                                        ("fp64_exact", True)])
def test_noalloc_equals_allocating_entry_point_twice(prec, eps12):
    g, big, small = _cases()
    assert len(big) == L.WENO_NOALLOC_MAX
    for a in (big, small):
        print(f"noalloc {prec} eps12={eps12} n={len(a)}", flush=True)
        r0, l0 = S.weno_points(a, prec, eps12=eps12)
        r1, l1 = L.weno_points_noalloc(a, prec, eps12=eps12)
        assert np.array_equal(r0.view(np.int64), r1.view(np.int64))
        assert np.array_equal(l0.view(np.int64), l1.view(np.int64))
    if prec == "fp64_exact" and not eps12:
        r, l = L.weno_points_noalloc(g["a"], prec)
        assert np.array_equal(r.view(np.int64), g["right"].view(np.int64))
        assert np.array_equal(l.view(np.int64), g["left"].view(np.int64))


def test_noalloc_rejects_more_than_its_maximum():
    lib = S.lib()
    n = L.WENO_NOALLOC_MAX + 1
    a = np.zeros((n, 5))
    r = np.zeros(n)
    P = ctypes.POINTER(ctypes.c_double)
    rc = lib.swf_weno_points_noalloc(a.ctypes.data_as(P), r.ctypes.data_as(P), r.ctypes.data_as(P), n, 0, 0)
    assert rc == -1  # SWF_E_INVALID, before any HIP call
    lib.swf_last_error.restype = ctypes.c_char_p
    assert b"exceed the maximum 65536" in lib.swf_last_error()
