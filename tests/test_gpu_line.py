"""The 1D program (ex2_sedov_1d/main.f90) on the GPU through the C-ABI (swf1d_*).

- FP64-exact against the reference's own outputs (tests/golden/ref_ex2_sedov1d.json,
  three builds of the program): den/pre/vel of the 800 interior cells bitwise, the
  printed sweep statistics exactly, the step count through the progress lines.
- FP64-exact against the oracle on the 40-step cfl 1.5 run: the whole state including
  the ghost cells bitwise and the per-sweep tracks.
- The fast modes on the 1D benchmark case (cfl 0.6, ambient E 1e-12, swp_num 2, T = 1e-3)
  against FP64-exact: relative L1 of the density, tolerance in the test.
- fortran/swf_sedov1d, the ex2 program over the C-ABI, writes den.dat, pre.dat and
  vel.dat byte for byte as the reference builds do, and prints the same norms and
  statistics.
- Argument checks of swf1d_create.
"""
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

import sweeping_fd_weno_amd as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sweeping_fd_weno_amd",
                      "fortran", "swf_sedov1d")
DRIVER_ARGS = {"shipped": [], "cfg1": ["0.6", "2", "1"], "stress": ["1.5", "50", "0", "40"]}

# key: (cfl, swp_num, shipped ambient energy, max steps)
CASES = {
    "shipped": (1.2, 50, True, None),
    "cfg1": (0.6, 2, False, None),
    "stress": (1.5, 50, True, 40),
}


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def _run(case, precision="fp64_exact"):
    cfl, swp, shipped, max_steps = CASES[case]
    cfg = S.sedov1d_config(800, precision, cfl=cfl, swp_num=swp)
    with S.Solver1D(cfg) as s:
        s.upload(S.init_sedov1d(cfg, shipped_energy=shipped))
        t, k = s.advance(0.0, 0.001, max_steps if max_steps else 1 << 40)
        return s.download(), t, k, s.stats(), s.track()


@pytest.mark.parametrize("case", list(CASES))
def test_line_exact_matches_reference_program_bitwise(case):
    g = json.load(open(os.path.join(GOLD, "ref_ex2_sedov1d.json")))[case]
    print("case", case, flush=True)
    u, t, k, st, _ = _run(case)
    inner = u[3:803]
    pre = (1.4 - 1.0) * (inner[:, 2] - (0.5 * (inner[:, 1] * inner[:, 1])) / inner[:, 0])  # pre.dat (:481)
    assert _digest(inner[:, 0]) == g["sha256"]["den"]
    assert _digest(pre) == g["sha256"]["pre"]
    assert _digest(inner[:, 1] / inner[:, 0]) == g["sha256"]["vel"]
    none = st["sweep_count"] == 0
    printed = {
        "total sweep": st["total_sweep"], "max sweep": -2147483648 if none else st["max_sweep"],
        "total negative pairs": st["total_neg"], "max negative pairs": -2147483648 if none else st["max_neg"],
        "total positive neighbors": st["total_pos"],
        "max positive neighbors": -2147483648 if none else st["max_pos"],
        "den_count": st["den_count"], "pre_count": st["pre_count"],
    }
    assert printed == g["stats"]
    if CASES[case][3]:
        assert k == CASES[case][3] and float(u[:, 0].min()) == g["min_rho"]
    else:
        assert k // 50 * 50 == g["last_progress_step"]


def test_line_exact_matches_oracle_state_and_tracks():
    cfl, swp, shipped, max_steps = CASES["stress"]
    u, t, k, st, (q, ng, ps) = _run("stress")
    cfg = O.sedov1d_config(800, cfl=cfl, swp_num=swp)
    o = O.Oracle1D(cfg)
    o.set_state(O.init_sedov1d(cfg, shipped_energy=shipped))
    to, ko = o.run(0.0, 0.001, max_steps)
    assert (t, k) == (to, ko)
    assert np.array_equal(u.view(np.int64), o.state().view(np.int64))
    qo, ngo, pso = o.track()
    assert np.array_equal(q, qo) and np.array_equal(ng, ngo) and np.array_equal(ps, pso)
    assert st["stage_sweeps"] == st["sweep_count"] == 16


# relative L1 of the density against FP64-exact after the full run (2617 steps)
FAST_TOL = {"fp64": 1e-12, "fp32_bulk": 1e-6}


@pytest.mark.parametrize("precision", list(FAST_TOL))
def test_line_fast_modes_close_to_exact(precision):
    ue, te, ke, _, _ = _run("cfg1")
    u, t, k, st, _ = _run("cfg1", precision)
    l1 = np.abs(u[:, 0] - ue[:, 0]).sum() / np.abs(ue[:, 0]).sum()
    print(precision, "steps", k, ke, "l1", l1, flush=True)
    assert abs(k - ke) <= 1 and l1 <= FAST_TOL[precision]


def test_line_create_rejects_bad_arguments():
    with pytest.raises(S.SwfError):
        S.Solver1D(S.sedov1d_config(2043))          # n + 6 > 2048 cells
    with pytest.raises(S.SwfError):
        S.Solver1D(S.sedov1d_config(800, use_limiter=1))
    with pytest.raises(S.SwfError):
        S.Solver1D(S.sedov1d_config(800, precision=3))


@pytest.mark.parametrize("case", list(CASES))
def test_line_driver_reproduces_reference_output_files(case, tmp_path):
    g = json.load(open(os.path.join(GOLD, "ref_ex2_sedov1d.json")))[case]
    print("case", case, flush=True)
    res = subprocess.run([DRIVER] + DRIVER_ARGS[case], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    for fn, digest in g["file_sha256"].items():
        assert hashlib.sha256((tmp_path / fn).read_bytes()).hexdigest() == digest, fn
    assert hashlib.sha256(res.stdout.encode()).hexdigest() == g["stdout_sha256"]  # the whole printout
    lines = res.stdout.splitlines()
    at = [i for i, l in enumerate(lines) if l.strip().startswith("L1 =")][0]
    assert [float(v) for v in lines[at + 1].split()] == [g["L1"], g["Linf"]]
    stats = {}
    for l in lines:
        t = l.strip()
        if "=" in t and not t.startswith("L1") and not t.startswith("avg"):
            k, v = t.split("=", 1)
            stats[k.strip()] = int(v)
        if t.startswith("# of limiter"):
            stats["den_count"], stats["pre_count"] = (int(v) for v in t.split()[-2:])
    stats.pop("L1order", None)
    assert stats == g["stats"]
    progress = [int(l.split()[0]) for l in lines if l.strip().endswith("%")]
    assert (progress[-1] if progress else 0) == g["last_progress_step"]
