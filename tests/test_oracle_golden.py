"""The oracle against the reference's own outputs (fixtures from tests/golden/make_golden.py).

- ref_weno.npz: ex4_sedov_2d/weno5.f90 weno_right/weno_left, compiled unmodified with
  flang -O3 -fdefault-real-8, on seeded adversarial stencils -> bitwise.
- ref_ex3_vortex180.json: ex3_vortex_accuracy_2d as shipped (180^2 periodic vortex,
  T = 0.01, sweeping limiter on) -> final rho/m1/m2/p bitwise (SHA-256) and the
  printed sweep statistics and error norms exactly.
- ref_ex3_l1_vortex180.json: the same program built with its switch use_limiter = 1
  (the Zhang-Shu flux limiter in wenofd instead of the sweeping post-process) ->
  final fields bitwise, the printed limiter counts and error norms exactly.
- ref_ex5_jet160x40.json: ex5_mach2000_jet on a 160 x 40 grid to T = 1.5e-4 (90 steps,
  124 sweeps, 401 negative-neighbour pairs) -> the log density / pressure / p-over-rho
  fields it writes bitwise (logarithms by the C library, as the program's), the printed
  statistics exactly.
- ref_ex2_sedov1d.json: ex2_sedov_1d (the 1D program, 800 points) as shipped, at the
  1D benchmark settings (cfl 0.6, ambient E 1e-12, swp_num 2) and as a 40-step cfl 1.5
  run with 16 sweeps -> den/pre/vel bitwise, the sweep statistics exactly, the step
  count through the program's progress lines and its min_rho line.
"""
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_weno_matches_reference_bitwise():
    g = np.load(os.path.join(GOLD, "ref_weno.npz"))
    r = O.weno_right(g["a"])
    l = O.weno_left(g["a"])
    assert np.array_equal(r.view(np.int64), g["right"].view(np.int64))
    assert np.array_equal(l.view(np.int64), g["left"].view(np.int64))


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


@pytest.mark.slow
def test_ex3_vortex_run_matches_reference_bitwise():
    g = json.load(open(os.path.join(GOLD, "ref_ex3_vortex180.json")))
    cfg = O.vortex_config(180)
    o = O.Oracle(cfg)
    o.set_state(O.init_vortex(cfg))
    o.bc()
    rc, t, steps = o.run(0.0, 0.01)
    assert rc == 0 and steps == 9
    U = o.state()[3:-3, 3:-3, :]
    rho, m1, m2, e = (U[:, :, c].reshape(-1) for c in range(4))
    g1 = 1.4 - 1.0
    p = g1 * (e - 0.5 * m1 ** 2 / rho - 0.5 * m2 ** 2 / rho)
    got = {"rho": rho, "m1": m1, "m2": m2, "p": p}
    for k in got:
        assert _digest(got[k]) == g["sha256"][k], k
    s = o.stats()
    st = g["stats"]
    assert s["total_sweep"] == st["total sweep"] and s["max_sweep"] == st["max sweep"]
    assert s["total_neg"] == st["total negative pairs"] and s["max_neg"] == st["max negative pairs"]
    assert s["total_pos"] == st["total positive neighbors"] and s["max_pos"] == st["max positive neighbors"]
    assert s["den_count"] == st["den_count"] and s["pre_count"] == st["pre_count"]
    # error norms exactly as ex3 `output` forms them (:725-741)
    ex = O.vortex_exact(cfg, t)[3:-3, 3:-3, 0]
    er1 = 0.0
    eri = 0.0
    R = U[:, :, 0]
    for i in range(180):
        for j in range(180):
            d = abs(ex[j, i] - R[j, i])
            er1 = er1 + d
            eri = max(eri, d)
    er1 = er1 / 180 / 180
    assert er1 == g["L1"] and eri == g["Linf"]


@pytest.mark.slow
def test_ex3_vortex_flux_limiter_run_matches_reference_bitwise():
    g = json.load(open(os.path.join(GOLD, "ref_ex3_l1_vortex180.json")))
    cfg = O.vortex_config(180, use_limiter=1)
    o = O.Oracle(cfg)
    o.set_state(O.init_vortex(cfg))
    o.bc()
    rc, t, steps = o.run(0.0, 0.01)
    assert rc == 0 and steps == 9
    U = o.state()[3:-3, 3:-3, :]
    rho, m1, m2, e = (U[:, :, c].reshape(-1) for c in range(4))
    p = (1.4 - 1.0) * (e - 0.5 * m1 ** 2 / rho - 0.5 * m2 ** 2 / rho)
    got = {"rho": rho, "m1": m1, "m2": m2, "p": p}
    for k in got:
        assert _digest(got[k]) == g["sha256"][k], k
    s = o.stats()
    assert (s["den_count"], s["pre_count"]) == (g["den_count"], g["pre_count"]) == (20, 416)
    assert s["sweep_count"] == 0
    ex = O.vortex_exact(cfg, t)[3:-3, 3:-3, 0]
    d = np.abs(ex - U[:, :, 0])
    er1 = float(np.cumsum(d.T.ravel())[-1]) / 180 / 180  # i outer, j inner as in `output`
    assert er1 == g["L1"] and float(d.max()) == g["Linf"]


_EX2 = {
    # key: (cfl, swp_num, shipped ambient energy, max steps)
    "shipped": (1.2, 50, True, None),
    "cfg1": (0.6, 2, False, None),
    "stress": (1.5, 50, True, 40),
}


def _stats1d_as_printed(st):
    """The program's printout: maxval over an empty track prints as -huge-1 (int32)."""
    none = st["sweep_count"] == 0
    return {
        "total sweep": st["total_sweep"], "max sweep": -2147483648 if none else st["max_sweep"],
        "total negative pairs": st["total_neg"], "max negative pairs": -2147483648 if none else st["max_neg"],
        "total positive neighbors": st["total_pos"],
        "max positive neighbors": -2147483648 if none else st["max_pos"],
        "den_count": st["den_count"], "pre_count": st["pre_count"],
    }


@pytest.mark.slow
@pytest.mark.parametrize("case", list(_EX2))
def test_ex2_1d_run_matches_reference_bitwise(case):
    g = json.load(open(os.path.join(GOLD, "ref_ex2_sedov1d.json")))[case]
    cfl, swp, shipped, max_steps = _EX2[case]
    cfg = O.sedov1d_config(800, cfl=cfl, swp_num=swp)
    o = O.Oracle1D(cfg)
    o.set_state(O.init_sedov1d(cfg, shipped_energy=shipped))
    t, k = o.run(0.0, 0.001, max_steps if max_steps else 1 << 40)
    u = o.state()
    inner = u[3:3 + 800]
    fields = {"den": inner[:, 0], "pre": O.pressure1d(u)[3:3 + 800], "vel": inner[:, 1] / inner[:, 0]}
    for name, v in fields.items():
        assert _digest(v) == g["sha256"][name], name
    assert _stats1d_as_printed(o.stats()) == g["stats"]
    if max_steps:
        assert k == max_steps
        assert float(u[:, 0].min()) == g["min_rho"]
    else:
        assert k // 50 * 50 == g["last_progress_step"]  # progress printed every 50 steps
    q, neg, pos = o.track()
    assert int(q.sum()) == g["stats"]["total sweep"] and int(pos.sum()) == g["stats"]["total positive neighbors"]


def test_flang_dot_product_sums_left_to_right():
    """ex2's sweep mean state is dot_product(rho_t, weights) (ex2_sedov_1d/main.f90:577-579);
    the 1D oracle sums u*w left to right from 0.  oracle/ref_harness/dot_probe.f90, built
    with ex2's flags, prints dot_product on data where the order matters."""
    exe = os.path.join(os.path.dirname(O.__file__), "_ref", "dot_probe")
    if not os.path.exists(exe):
        if not os.path.exists("/opt/rocm/lib/llvm/bin/flang"):
            pytest.skip("flang not available")
        subprocess.run(["make", "-s", "-C", os.path.dirname(O.__file__), exe], check=True)
    dot_hex, loop_hex = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    x = np.array([1.0 + (i % 7) * 0.37 for i in range(1, 807)])
    x[0], x[399] = 1e17, -1e17
    w = 1 / (800.0 + 6)
    s = 0.0
    for v in x:
        s = s + v * w
    exact = float(sum(__import__("fractions").Fraction(float(v * w)) for v in x))
    assert exact != s  # the data distinguish the orders
    want = "%016X" % np.float64(s).view(np.uint64)
    assert dot_hex == want and loop_hex == want


def _libm_log(a):
    import ctypes
    libm = ctypes.CDLL("libm.so.6")
    libm.log.restype = ctypes.c_double
    libm.log.argtypes = [ctypes.c_double]
    return np.array([libm.log(float(v)) for v in np.ravel(a)])


def jet_log_fields(u, nx, ny, gamma=5.0 / 3.0):
    """The unmirrored half of ex5's den.dat / pre.dat / m1.dat (ex5_mach2000_jet/main.f90:834-847)
    from a state u (ny+6, nx+6, 4): (j outer, i inner) log fields."""
    v = u[3:3 + ny, 3:3 + nx]
    r, m, n, e = v[:, :, 0], v[:, :, 1], v[:, :, 2], v[:, :, 3]
    p = (gamma - 1.0) * ((e - (0.5 * (m * m)) / r) - (0.5 * (n * n)) / r)
    return {"log_rho": _libm_log(np.abs(r)), "log_p": _libm_log(np.abs(p)), "log_p_over_rho": _libm_log(np.abs(p / r))}


@pytest.mark.slow
def test_ex5_jet_160x40_short_run_matches_reference_bitwise():
    g = json.load(open(os.path.join(GOLD, "ref_ex5_jet160x40.json")))["short"]
    cfg = O.jet_config(160, 40)
    o = O.Oracle(cfg)
    o.set_state(O.init_jet(cfg))
    o.bc()
    rc, t, k = o.run(0.0, 0.00015)
    assert rc == 0
    for name, v in jet_log_fields(o.state(), 160, 40).items():
        assert _digest(v) == g["sha256"][name], name
    st = o.stats()
    assert st["total_sweep"] == g["stats"]["total sweep"] and st["total_neg"] == g["stats"]["total negative pairs"]
    assert st["total_pos"] == g["stats"]["total positive neighbors"] and st["pre_count"] == g["stats"]["pre_count"]
    assert st["max_sweep"] == g["stats"]["max sweep"] and st["max_neg"] == g["stats"]["max negative pairs"]
    assert k // 20 * 20 == g["last_progress_step"]
