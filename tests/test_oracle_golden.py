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
- ref_ex4_sedov.json: ex4_sedov_2d (the benchmark program) as one-rank build copies
  (oracle/Makefile: `use MPI` -> `include 'mpif.h'`, cut_x = cut_y = 1): 64^2 to
  T = 0.002 (4 steps, 4 sweeps), 128^2 to T = 0.01 (37 steps, 3 sweeps, 4 negative
  pairs, 90 positive neighbours) and the benchmark grid 640^2 to T = 1.2e-5 (3 steps,
  2 sweeps, 7 negative pairs) -> den/pre/m1/m2 bitwise, the statistics and step counts
  exactly.  This pins ex4's own sites: the binary128 `1d-15` of the Roe weight, sound
  speed and y-pass reciprocal (main.f90:234,240,426,455,461), findalpha's eps2 form
  (:652-683), alpha applied in wenofd (:414-415), the Sedov bc (:916-994) and the
  initial bc of initdata (:1186).  Planted in the oracle, each of those mistakes fails
  the 64^2 fixture (test_ex4_fixture_detects_planted_mistakes).  Two sites no reachable
  state tells apart stay pinned by reading: checkpositivity's strict `<` (:779-801)
  and the REAL(8) `1d-8` literal of fflux, whose alternatives differ only for values
  exactly at those constants.
- ref_ex6_step.json: ex6_shock_diffraction as one-rank build copies at total = 8
  (104 x 88) to the shipped T = 2.3 (150 steps, 190 sweeps, 57 density sweeps, 62
  negative pairs) and total = 16 (208 x 176) to T = 0.5 -> den/pre bitwise (the step
  blanked as the program writes it), the statistics exactly.  This pins the per-line
  alpha (:217-219,449-451), the step walls inside wenofd (:192-202,416-431), its bc
  with the `i = 1` loop (:896-943), the `<=` positivity test, the 1d-12 weights and the
  pressure-sum check's no-stop path.

The long runs are computed concurrently, one process per case, the first time a test
asks for them.
"""
import concurrent.futures as cf
import hashlib
import json
import multiprocessing as mp
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


def _ex3_case(limiter: int) -> dict:
    cfg = O.vortex_config(180, use_limiter=limiter)
    o = O.Oracle(cfg)
    o.set_state(O.init_vortex(cfg))
    o.bc()
    rc, t, steps = o.run(0.0, 0.01)
    U = o.state()[3:-3, 3:-3, :]
    rho, m1, m2, e = (U[:, :, c].reshape(-1) for c in range(4))
    g1 = 1.4 - 1.0
    p = g1 * (e - 0.5 * m1 ** 2 / rho - 0.5 * m2 ** 2 / rho)
    ex = O.vortex_exact(cfg, t)[3:-3, 3:-3, 0]
    R = U[:, :, 0]
    if limiter == 2:
        # error norms exactly as ex3 `output` forms them (:725-741): i outer, j inner
        er1, eri = 0.0, 0.0
        for i in range(180):
            for j in range(180):
                d = abs(ex[j, i] - R[j, i])
                er1 = er1 + d
                eri = max(eri, d)
        er1 = er1 / 180 / 180
    else:
        d = np.abs(ex - R)
        er1 = float(np.cumsum(d.T.ravel())[-1]) / 180 / 180
        eri = float(d.max())
    return {"rc": rc, "steps": steps, "sha": {k: _digest(v) for k, v in (("rho", rho), ("m1", m1), ("m2", m2), ("p", p))},
            "stats": o.stats(), "L1": er1, "Linf": eri}


_EX2 = {
    # key: (cfl, swp_num, shipped ambient energy, max steps)
    "shipped": (1.2, 50, True, None),
    "cfg1": (0.6, 2, False, None),
    "stress": (1.5, 50, True, 40),
}


def _ex2_case(case: str) -> dict:
    cfl, swp, shipped, max_steps = _EX2[case]
    cfg = O.sedov1d_config(800, cfl=cfl, swp_num=swp)
    o = O.Oracle1D(cfg)
    o.set_state(O.init_sedov1d(cfg, shipped_energy=shipped))
    t, k = o.run(0.0, 0.001, max_steps if max_steps else 1 << 40)
    u = o.state()
    inner = u[3:3 + 800]
    fields = {"den": inner[:, 0], "pre": O.pressure1d(u)[3:3 + 800], "vel": inner[:, 1] / inner[:, 0]}
    q, neg, pos = o.track()
    return {"steps": k, "sha": {n: _digest(v) for n, v in fields.items()}, "stats": o.stats(),
            "min_rho": float(u[:, 0].min()), "track_q": int(q.sum()), "track_pos": int(pos.sum())}


def _ex5_short_case() -> dict:
    cfg = O.jet_config(160, 40)
    o = O.Oracle(cfg)
    o.set_state(O.init_jet(cfg))
    o.bc()
    rc, t, k = o.run(0.0, 0.00015)
    return {"rc": rc, "steps": k, "sha": {n: _digest(v) for n, v in jet_log_fields(o.state(), 160, 40).items()},
            "stats": o.stats()}


_EX4 = {"n64": (64, 0.002), "n128": (128, 0.01), "n640w": (640, 1.2e-5)}


def sedov_columns(u: np.ndarray, gamma: float = 1.4) -> dict:
    """The value columns of ex4's den/pre/m1/m2.dat (ex4_sedov_2d/main.f90:1304-1316) from a
    state (ny+6, nx+6, 4): (j outer, i inner) over the interior."""
    U = u[3:-3, 3:-3, :]
    rho, m1, m2, e = (U[:, :, c].reshape(-1) for c in range(4))
    pre = (gamma - 1.0) * (e - 0.5 * m1 ** 2 / rho - 0.5 * m2 ** 2 / rho)
    return {"den": rho, "pre": pre, "m1": m1, "m2": m2}


def _ex4_case(key: str, **planted) -> dict:
    n, T = _EX4[key]
    oc = O.sedov_config(n)
    for k, v in planted.items():
        setattr(oc, k, v)
    o = O.Oracle(oc)
    o.set_state(O.init_sedov(oc))
    o.bc()  # initdata ends with `call bc` (ex4_sedov_2d/main.f90:1186)
    rc, t, k = o.run(0.0, T)
    return {"rc": rc, "steps": k, "t": t, "sha": {nm: _digest(v) for nm, v in sedov_columns(o.state()).items()},
            "stats": o.stats()}


_EX6 = {"total8": (8, 2.3), "total16": (16, 0.5)}


def step_columns(u: np.ndarray, total: int, gamma: float = 1.4) -> dict:
    """ex6's den/pre.dat value columns (ex6_shock_diffraction/main.f90:1218-1230): (j outer,
    i inner), 0 where i < total and j < 6*total."""
    U = u[3:-3, 3:-3, :]
    rho, m1, m2, e = (U[:, :, c].copy() for c in range(4))
    pre = (gamma - 1.0) * (e - 0.5 * m1 ** 2 / rho - 0.5 * m2 ** 2 / rho)
    jj, ii = np.meshgrid(np.arange(1, U.shape[0] + 1), np.arange(1, U.shape[1] + 1), indexing="ij")
    blank = (ii < total) & (jj < 6 * total)
    rho[blank] = 0.0
    pre[blank] = 0.0
    return {"den": rho.reshape(-1), "pre": pre.reshape(-1)}


def _ex6_case(key: str) -> dict:
    total, T = _EX6[key]
    oc = O.step_config(total)
    u0 = O.init_step(oc)
    o = O.Oracle(oc)
    o.set_state(u0)  # no bc after initdata in ex6
    rc, t, k = o.run(0.0, T)
    return {"rc": rc, "steps": k, "t": t, "sha": {nm: _digest(v) for nm, v in step_columns(o.state(), total).items()},
            "stats": o.stats()}


# planted single-site mistakes (oracle switches that are another program's semantics)
_PLANTED = {
    "roe_eps_double": "Roe weight / sound speed `+1d-15` in REAL(8) instead of binary128 (:234,240)",
    "ydens_plain": "y-pass `1./den` instead of binary128 `1./(den+1d-15)` (:426)",
    "findalpha_ex3": "findalpha without the eps2 velocity cut-off (:652-683)",
    "alpha_in_scheme": "alpha applied in scheme instead of wenofd (:414-415)",
}

_CASES = {
    "ex3": (_ex3_case, (2,)), "ex3_l1": (_ex3_case, (1,)),
    **{"ex2_" + c: (_ex2_case, (c,)) for c in _EX2},
    "ex5_short": (_ex5_short_case, ()),
    **{"ex4_" + c: (_ex4_case, (c,)) for c in _EX4},
    **{"ex6_" + c: (_ex6_case, (c,)) for c in _EX6},
}


def _call(name):
    fn, args = _CASES[name]
    return fn(*args)


def _planted(sw):
    return _ex4_case("n64", **{sw: 1})


@pytest.fixture(scope="module")
def runs():
    """Every long oracle run of this module, computed concurrently (one process each)."""
    workers = max(1, min(8, os.cpu_count() or 1))
    with cf.ProcessPoolExecutor(workers, mp_context=mp.get_context("fork")) as ex:
        fut = {n: ex.submit(_call, n) for n in sorted(_CASES, key=lambda n: n.startswith("ex6"), reverse=True)}
        fut.update({"planted_" + sw: ex.submit(_planted, sw) for sw in _PLANTED})
        return {n: f.result() for n, f in fut.items()}


@pytest.mark.slow
def test_ex3_vortex_run_matches_reference_bitwise(runs):
    g = json.load(open(os.path.join(GOLD, "ref_ex3_vortex180.json")))
    r = runs["ex3"]
    assert r["rc"] == 0 and r["steps"] == 9
    for k in ("rho", "m1", "m2", "p"):
        assert r["sha"][k] == g["sha256"][k], k
    s = r["stats"]
    st = g["stats"]
    assert s["total_sweep"] == st["total sweep"] and s["max_sweep"] == st["max sweep"]
    assert s["total_neg"] == st["total negative pairs"] and s["max_neg"] == st["max negative pairs"]
    assert s["total_pos"] == st["total positive neighbors"] and s["max_pos"] == st["max positive neighbors"]
    assert s["den_count"] == st["den_count"] and s["pre_count"] == st["pre_count"]
    assert r["L1"] == g["L1"] and r["Linf"] == g["Linf"]


@pytest.mark.slow
def test_ex3_vortex_flux_limiter_run_matches_reference_bitwise(runs):
    g = json.load(open(os.path.join(GOLD, "ref_ex3_l1_vortex180.json")))
    r = runs["ex3_l1"]
    assert r["rc"] == 0 and r["steps"] == 9
    for k in ("rho", "m1", "m2", "p"):
        assert r["sha"][k] == g["sha256"][k], k
    s = r["stats"]
    assert (s["den_count"], s["pre_count"]) == (g["den_count"], g["pre_count"]) == (20, 416)
    assert s["sweep_count"] == 0
    assert r["L1"] == g["L1"] and r["Linf"] == g["Linf"]


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
def test_ex2_1d_run_matches_reference_bitwise(case, runs):
    g = json.load(open(os.path.join(GOLD, "ref_ex2_sedov1d.json")))[case]
    max_steps = _EX2[case][3]
    r = runs["ex2_" + case]
    for name in ("den", "pre", "vel"):
        assert r["sha"][name] == g["sha256"][name], name
    assert _stats1d_as_printed(r["stats"]) == g["stats"]
    if max_steps:
        assert r["steps"] == max_steps
        assert r["min_rho"] == g["min_rho"]
    else:
        assert r["steps"] // 50 * 50 == g["last_progress_step"]  # progress printed every 50 steps
    assert r["track_q"] == g["stats"]["total sweep"] and r["track_pos"] == g["stats"]["total positive neighbors"]


def _stats2d_equal(s, st):
    assert s["total_sweep"] == st["total sweep"] and s["max_sweep"] == st["max sweep"]
    assert s["total_neg"] == st["total negative pairs"] and s["max_neg"] == st["max negative pairs"]
    assert s["total_pos"] == st["total positive neighbors"] and s["max_pos"] == st["max positive neighbors"]
    assert s["den_count"] == st["den_count"] and s["pre_count"] == st["pre_count"]


@pytest.mark.slow
@pytest.mark.parametrize("case", list(_EX4))
def test_ex4_sedov_run_matches_reference_bitwise(case, runs):
    g = json.load(open(os.path.join(GOLD, "ref_ex4_sedov.json")))[case]
    r = runs["ex4_" + case]
    assert r["rc"] == 0 and r["steps"] == g["steps"]
    for nm in ("den", "pre", "m1", "m2"):
        assert r["sha"][nm] == g["sha256"][nm], nm
    _stats2d_equal(r["stats"], g["stats"])


@pytest.mark.slow
def test_ex4_fixture_detects_planted_mistakes(runs):
    """Each single-site mistake planted in the oracle changes the 64^2 fields, so the ex4
    fixture pins those sites."""
    g = json.load(open(os.path.join(GOLD, "ref_ex4_sedov.json")))["n64"]
    assert runs["ex4_n64"]["sha"]["den"] == g["sha256"]["den"]
    for sw, what in _PLANTED.items():
        assert runs["planted_" + sw]["sha"]["den"] != g["sha256"]["den"], what


@pytest.mark.slow
@pytest.mark.parametrize("case", list(_EX6))
def test_ex6_step_run_matches_reference_bitwise(case, runs):
    g = json.load(open(os.path.join(GOLD, "ref_ex6_step.json")))[case]
    r = runs["ex6_" + case]
    assert r["rc"] == 0 and r["steps"] == g["steps"]
    for nm in ("den", "pre"):
        assert r["sha"][nm] == g["sha256"][nm], nm
    _stats2d_equal(r["stats"], g["stats"])


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
def test_ex5_jet_160x40_short_run_matches_reference_bitwise(runs):
    g = json.load(open(os.path.join(GOLD, "ref_ex5_jet160x40.json")))["short"]
    r = runs["ex5_short"]
    assert r["rc"] == 0
    for name, v in r["sha"].items():
        assert v == g["sha256"][name], name
    st = r["stats"]
    assert st["total_sweep"] == g["stats"]["total sweep"] and st["total_neg"] == g["stats"]["total negative pairs"]
    assert st["total_pos"] == g["stats"]["total positive neighbors"] and st["pre_count"] == g["stats"]["pre_count"]
    assert st["max_sweep"] == g["stats"]["max sweep"] and st["max_neg"] == g["stats"]["max negative pairs"]
    assert r["steps"] // 20 * 20 == g["last_progress_step"]
