"""The library's device code run on the CPU (tests/emu: the same sources built
with g++ against an emulated HIP runtime, fibers for threads) against the oracle.

This runs the stage kernels, the sweep and the launch sequence of swf_advance
(graph capture included) without a GPU.  The emulation interleaves a block's
threads only at barriers and runs blocks one after another, so it checks the
arithmetic and control flow, not races; the GPU tests check the real device.

- Sedov 32^2, 10 steps, 4 sweeps: fields bitwise, statistics equal.
- Full-domain jet 64x32 at CFL 0.5, 30 steps, 40 sweeps (q up to 2): bitwise.
- Half-domain jet 80x20 at CFL 0.25, 40 steps: bitwise.
- Vortex 90^2 with the Zhang-Shu flux limiter: bitwise, limiter counts equal.
- swf_sweep on crafted states with 14, 3000 and 12000 cells of negative
  pressure (the last more than one ordered list of flagged positions holds, so
  its passes go chunk by chunk), and with a chain of 100, 300 or 6000 consecutive
  flagged snake positions (walks replayed from their speculative record, and one
  longer than the record's 4096 steps, walked again as it writes): fields
  bitwise, statistics equal.
- FP64 / FP32-bulk full jet 64x32 with sweeps (u_bar from per-tile sums).
- A batch of three instances against each run alone: bitwise.
- After upload of the ambient jet and swf_bc, the first dt sees the Mach 2000
  inflow ghosts (alpha ~ 800) in every precision mode.
- The step diffraction program, 52 x 44 cells for 40 steps from its initial data
  (50 sweeps, 15 density sweeps, 13 negative-neighbour adjustments): bitwise.
- The detonation program (ex7) at 40 x 40 to T = 0.2 (17 steps, 3 sweeps): den / pre /
  rate bitwise to the reference's one-rank run, statistics equal.
- The 1D program (one workgroup, whole time loop): 60 cells for 12 steps and 800
  cells for 40 steps at cfl 1.5 (16 sweeps, 5273 limited pairs) and 30 steps of the
  1D benchmark case: the whole state bitwise, statistics and tracks equal.
"""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(ROOT, "tests", "emu")


# every scenario of this module: run concurrently, one process each, when the module starts
SCENARIOS = ["sedov_sweeps", "jet_full_sweeps", "jet_half", "vortex_flux_limiter", "crafted_sweep_24_14",
             "crafted_sweep_128_3000", "crafted_sweep_160_12000", "crafted_sweep_64_20_100",
             "crafted_sweep_64_20_300", "crafted_sweep_96_20_6000", "crafted_psum_48_400_1", "crafted_psum_48_400_2", "jet_full_sweeps_fp64",
             "jet_full_sweeps_fp32_bulk", "batch_fp64_exact", "batch_fp32_bulk", "first_dt_fp64_exact",
             "first_dt_fp64", "first_dt_fp32_bulk", "line_small", "line_stress_fp64_exact", "line_cfg1",
             "step_fp64_exact_4_40", "ex7_8", "sweep7_crafted"]
_RESULTS = {}


@pytest.fixture(scope="module")
def emu_lib():
    subprocess.run(["make", "-s", "-C", EMU], check=True, capture_output=True)
    lib = os.path.join(EMU, "_build", "libswf_emu.so")
    assert os.path.exists(lib)
    # the longest first, at most as many at once as the host has cores (8 at most)
    order = sorted(SCENARIOS, key=lambda n: not n.startswith(("batch", "step", "jet_full", "vortex")))
    width = max(1, min(8, os.cpu_count() or 1))
    running = []
    for name in order + [None] * width:
        while running and (name is None or len(running) >= width):
            nm, p = running.pop(0)
            out, err = p.communicate(timeout=900)
            _RESULTS[nm] = (p.returncode, out, err)
        if name is not None:
            running.append((name, subprocess.Popen([sys.executable, os.path.join(EMU, "scenarios.py"), name],
                                                   stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
    return lib


def scenario(name):
    rc, out, err = _RESULTS[name]
    assert rc == 0, out + err
    return json.loads(out.strip().splitlines()[-1])


@pytest.mark.parametrize("name", ["sedov_sweeps", "jet_full_sweeps", "jet_half", "vortex_flux_limiter",
                                  "crafted_sweep_24_14", "crafted_sweep_128_3000", "crafted_sweep_160_12000",
                                  "crafted_sweep_64_20_100", "crafted_sweep_64_20_300", "crafted_sweep_96_20_6000"])
def test_emulated_kernels_match_oracle(emu_lib, name):
    r = scenario(name)
    assert r["rc"] == 0
    assert r["steps"][0] == r["steps"][1] and r["time"][0] == r["time"][1]
    assert r["differ"] == 0
    for key, (o, e) in r["stats"].items():
        assert o == e, key
    if name == "jet_full_sweeps":
        assert r["stats"]["sweep_count"][0] >= 30
    if name == "vortex_flux_limiter":
        assert r["stats"]["pre_count"][0] > 0


@pytest.mark.parametrize("seed", [1, 2])
def test_emulated_pressure_sum_check_agrees_with_oracle(emu_lib, seed):
    """Crafted sweeps with the step program's pressure-sum check on: the device raises
    SWF_E_PSUM only for genuine walks (speculative walks that are discarded do not
    count), exactly when the oracle's serial sweep stops.  Pressure is concave in the
    conserved variables, so a pair adjustment does not lower the pair's total pressure
    beyond rounding, and on these states neither side stops."""
    r = scenario(f"crafted_psum_48_400_{seed}")
    assert r["rc"][0] == r["rc"][1]


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-7), ("fp32_bulk", 1e-4)])
def test_emulated_fast_modes_jet_sweeps(emu_lib, prec, tol):
    """FP64 / FP32-bulk full jet 64x32, 25 steps with sweeps at every stage from
    step 18 (u_bar from the stage kernel's per-tile sums): the same number of
    sweep passes as the oracle's exact run, and relative L1 density within the
    stated bound.  The coarse jet amplifies rounding quickly (FP32-bulk: 1.3e-6
    at 25 steps, 2e-3 at 30), so the bounds are loose; the GPU tests hold the
    200x100 jet to 1e-8 / 1e-4."""
    r = scenario("jet_full_sweeps_" + prec)
    assert r["rc"] == 0 and r["steps"] == [25, 25]
    assert r["stats"]["sweep_count"][0] == r["stats"]["sweep_count"][1] >= 20
    assert r["stats"]["total_sweep"][0] == r["stats"]["total_sweep"][1]
    assert r["l1"] < tol


@pytest.mark.parametrize("prec", ["fp64_exact", "fp32_bulk"])
def test_emulated_batch_matches_single_runs(emu_lib, prec):
    """A batch of Sedov 40^2, vortex 36^2, full jet 48x24 and a Mach 4 step
    diffraction 52x44 (the vortex ends after one step while the others go on): every
    instance bitwise equal to its run alone, with the same steps, time and sweeps."""
    r = scenario("batch_" + prec)
    assert r["rc"] == 0 and r["steps"] == [25, 1, 25, 25]
    assert r["differ"] == 0 and r["time_equal"]
    for o, e in r["sweeps"]:
        assert o == e
    assert r["sweeps"][2][0] > 10


@pytest.mark.parametrize("prec", ["fp64_exact", "fp64", "fp32_bulk"])
def test_emulated_first_dt_after_bc(emu_lib, prec):
    r = scenario("first_dt_" + prec)
    dto, dte = r["dt"]
    assert r["rc"] == 0 and dto < 0.05 * 0.001
    assert abs(dte - dto) <= 1e-12 * dto


@pytest.mark.parametrize("name", ["line_small", "line_stress_fp64_exact", "line_cfg1"])
def test_emulated_line_matches_oracle(emu_lib, name):
    r = scenario(name)
    assert r["steps"][0] == r["steps"][1] and r["time"][0] == r["time"][1]
    assert r["differ"] == 0 and r["track_equal"]
    for key, (o, e) in r["stats"].items():
        assert o == e, key
    if name == "line_stress_fp64_exact":
        assert r["stats"]["sweep_count"][0] == 16 and r["stats"]["total_pos"][0] == 5273


def test_emulated_step_program_matches_oracle(emu_lib):
    r = scenario("step_fp64_exact_4_40")
    assert r["rc"] == [0, 0] and r["steps"] == [40, 40] and r["time"][0] == r["time"][1]
    assert r["differ"] == 0
    for key, (o, e) in r["stats"].items():
        assert o == e, key
    assert r["stats"]["den_count"][0] > 0 and r["stats"]["total_neg"][0] > 0


def test_emulated_detonation_matches_reference_run(emu_lib):
    import json as _json
    gold = _json.load(open(os.path.join(ROOT, "tests", "golden", "ref_ex7_detonation.json")))["total8"]
    r = scenario("ex7_8")
    assert r["code"] == 0 and r["steps"] == gold["steps"]
    assert r["sha256"] == gold["sha256"]
    st = gold["stats"]
    assert (r["stats"]["total_sweep"], r["stats"]["total_neg"], r["stats"]["total_pos"]) == (
        st["total sweep"], st["total negative pairs"], st["total positive neighbors"])
    assert (r["stats"]["den_count"], r["stats"]["pre_count"]) == (st["den_count"], st["pre_count"])


def test_emulated_detonation_sweep_equals_oracle_on_crafted_states(emu_lib):
    """swf7_sweep (the workgroup-parallel sweep) against the oracle's serial restatement
    of ex7's sweep_lim on the states of tests/sweep7_cases.py: bitwise, the same stop,
    counts and track entry."""
    r = scenario("sweep7_crafted")
    codes = set()
    for c in r["cases"]:
        assert c["code"][0] == c["code"][1], c["case"]
        assert c["counts"][0] == c["counts"][1], c["case"]
        assert c["track"][0] == c["track"][1], c["case"]
        assert c["differ"] == 0 and c["changed"] > 0, c["case"]
        codes.add(c["code"][1])
    assert codes == {0, 3, 4}
