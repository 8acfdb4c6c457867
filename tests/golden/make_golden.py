"""Writes the golden fixtures under tests/golden/ from the reference programs.

Runs in the build container only (it needs /root/reference and flang): the
reference is compiled unmodified by ``oracle/Makefile`` (target ``ref``) into
oracle/_ref/, run there, and only its OUTPUTS are stored here:

  ref_weno.npz            weno_right / weno_left (ex4_sedov_2d/weno5.f90:86-146) on
                          seeded adversarial 5-point stencils, via oracle/ref_harness.
  ref_weno6.npz           the same reconstructions from ex6_shock_diffraction/weno5.f90
                          (1d-12 in the weights) via oracle/_ref/weno_probe6, on those
                          stencils plus small-amplitude ones where beta ~ 1d-12.
  ref_ex3_vortex180.json  ex3_vortex_accuracy_2d as shipped (180^2, T = 0.01, sweep on):
                          SHA-256 of the final rho/m1/m2/p fields, a strided sample,
                          the printed L1/Linf errors and the sweep statistics.
  ref_ex5_jet160x40.json  ex5_mach2000_jet with its grid PARAMETERs set to 160 x 40, to
                          T = 0.001 (612 steps, 1 687 sweeps) and to T = 1.5e-4: SHA-256
                          of den.dat / pre.dat / m1.dat and of the printed text, the log
                          fields of the unmirrored half, the statistics.
  ref_ex5_jet.json        ex5_mach2000_jet as shipped (800x200, T = 0.001), when its
                          (hour-long) run has finished: log-field hashes + statistics.

  ref_ex2_sedov1d.json    ex2_sedov_1d (800 points, T = 0.001) in three builds: as shipped
                          (cfl 1.2), the 1D benchmark case (cfl 0.6, ambient E 1e-12,
                          swp_num 2) and a 40-step cfl 1.5 run stopped by the program's
                          max_iter branch: SHA-256 of den/pre/vel, samples, statistics.

  ref_ex4_sedov.json      ex4_sedov_2d as a one-rank build copy (oracle/Makefile: `use MPI`
                          replaced by `include 'mpif.h'`, cut_x = cut_y = 1, grid and
                          time_terminal PARAMETERs): 64^2 to T = 0.002, 128^2 to T = 0.01
                          and the benchmark grid 640^2 to T = 1.2e-5 (3 steps, 2 sweeps):
                          SHA-256 of the den/pre/m1/m2 columns and files, samples, the
                          statistics, the step count and the printout minus `CPU Time`.
  ref_ex6_step.json       ex6_shock_diffraction as one-rank build copies: total = 8
                          (104 x 88) to the shipped T = 2.3 (150 steps, 190 sweeps, 57
                          density sweeps) and total = 16 (208 x 176) to T = 0.5: the same
                          records for den.dat / pre.dat.

  ref_ex7_detonation.json ex7_detonation as one-rank build copies (the same edits, and mpirank
                          initialised to 0, which the program tests before MPI_Init): total = 8
                          (40 x 40, 17 steps, 3 sweeps) and total = 16 (80 x 80, 33 steps, 8
                          sweeps) to T = 0.2: the same records for den.dat / pre.dat / rate.dat,
                          and the 40 x 40 fields in full.
  ref_ex7_detonation400.json  the shipped grid (400 x 400, T = 0.6) to the program's own
                          stop: progress lines, STOP text, run time on one core.
  ref_ex5_jet800x200.json ex5_mach2000_jet at its shipped grid to T = 5e-5 (time_terminal
                          set in a build copy): the records of ref_ex5_jet160x40.json.

Usage:  python tests/golden/make_golden.py [weno] [weno6] [ex3] [ex3_l1] [ex2] [ex5s] [ex5] [ex5ship] [ex4]
        [ex6] [ex7] [ex7ship]
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
GOLD = os.path.join(ROOT, "tests", "golden")
REF = os.path.join(ROOT, "oracle", "_ref")


def weno_stencils(seed: int = 20261203, n: int = 1000) -> np.ndarray:
    """Seeded adversarial stencils a(-2..2)."""
    rng = np.random.default_rng(seed)
    parts = [rng.standard_normal((n, 5))]
    b = rng.standard_normal((n, 1)) * 10.0 ** rng.integers(-20, 20, (n, 1))
    parts.append(b + b * rng.standard_normal((n, 5)) * 10.0 ** rng.integers(-17, -8, (n, 1)))
    parts.append(rng.standard_normal((n, 5)) * 10.0 ** rng.integers(-30, 30, (n, 5)))
    x = np.arange(-2, 3)
    c = rng.standard_normal((n, 3))
    parts.append(c[:, :1] + c[:, 1:2] * x + c[:, 2:3] * x ** 2)
    jump = rng.standard_normal((n, 5))
    jump[:, 3:] += 1e6
    parts.append(jump)
    parts.append(np.round(rng.standard_normal((n, 5)) * 8) / 8)
    parts.append(np.zeros((8, 5)))
    return np.ascontiguousarray(np.concatenate(parts))


def make_weno() -> None:
    a = weno_stencils()
    fin, fout = "/tmp/swf_wp_in.bin", "/tmp/swf_wp_out.bin"
    a.tofile(fin)
    subprocess.run([os.path.join(REF, "weno_probe"), fin, fout], check=True)
    r = np.fromfile(fout).reshape(-1, 2)
    np.savez_compressed(os.path.join(GOLD, "ref_weno.npz"), a=a, right=r[:, 0], left=r[:, 1])


def make_weno6() -> None:
    rng = np.random.default_rng(20261204)
    small = [rng.standard_normal((1000, 5)) * 10.0 ** rng.integers(-9, -3, (1000, 1))]
    x = np.arange(-2, 3)
    c = rng.standard_normal((1000, 3)) * 10.0 ** rng.integers(-8, -4, (1000, 1))
    small.append(c[:, :1] + c[:, 1:2] * x + c[:, 2:3] * x ** 2)
    a = np.ascontiguousarray(np.concatenate([weno_stencils()] + small))
    fin, fout = "/tmp/swf_wp6_in.bin", "/tmp/swf_wp6_out.bin"
    a.tofile(fin)
    subprocess.run([os.path.join(REF, "weno_probe6"), fin, fout], check=True)
    r = np.fromfile(fout).reshape(-1, 2)
    np.savez_compressed(os.path.join(GOLD, "ref_weno6.npz"), a=a, right=r[:, 0], left=r[:, 1])


def _tokens(fn: str, ncol: int) -> np.ndarray:
    return np.array(open(fn).read().split(), dtype=np.float64).reshape(-1, ncol)


def field_digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def _parse_stats(text: str) -> dict:
    out = {}
    for line in text.splitlines():
        s = line.strip()
        if "=" in s and not s.startswith("L1"):
            k, v = s.split("=", 1)
            k = k.strip()
            if k in ("total sweep", "max sweep", "total negative pairs", "max negative pairs",
                     "total positive neighbors", "max positive neighbors"):
                out[k] = int(v)
        if s.startswith("# of limiter"):
            d, p = s.split()[-2:]
            out["den_count"], out["pre_count"] = int(d), int(p)
    return out


def _run(prog: str, rundir: str) -> str:
    os.makedirs(rundir, exist_ok=True)
    with open(os.path.join(rundir, "fort.7"), "w") as f:
        f.write("%13.5e%13.5e\n" % (1.0, 1.0))
    open(os.path.join(rundir, "err_table.dat"), "w").close()
    res = subprocess.run(["bash", "-c", "ulimit -s unlimited; exec " + prog], cwd=rundir, check=True,
                         capture_output=True, text=True)
    with open(os.path.join(rundir, "stdout.txt"), "w") as f:
        f.write(res.stdout)
    return res.stdout


def make_ex3() -> None:
    rundir = os.path.join(REF, "run_ex3")
    text = _run(os.path.join(REF, "ex3", "a"), rundir)
    den = _tokens(os.path.join(rundir, "den.dat"), 5)
    m1 = _tokens(os.path.join(rundir, "m1.dat"), 3)
    m2 = _tokens(os.path.join(rundir, "m2.dat"), 3)
    pre = _tokens(os.path.join(rundir, "pre.dat"), 3)
    lines = text.splitlines()
    l1, linf = (float(v) for v in lines[1].split())
    fields = {"rho": den[:, 2], "m1": m1[:, 2], "m2": m2[:, 2], "p": pre[:, 2]}
    out = {
        "program": "ex3_vortex_accuracy_2d (unmodified), nx=ny=180, T=0.01, cfl=0.5, mag=10.0828",
        "layout": "fields in (j outer, i inner) order over the 180x180 interior",
        "sha256": {k: field_digest(v) for k, v in fields.items()},
        "sample_stride": 37,
        "sample": {k: v[::37].tolist() for k, v in fields.items()},
        "exact_rho_sample": den[::37, 3].tolist(),
        "L1": l1, "Linf": linf,
        "stats": _parse_stats(text),
        # the program's own output files (fort.7 / err_table.dat chained from fort.7 = 1.0 1.0)
        "err_table_dat": open(os.path.join(rundir, "err_table.dat")).read(),
        "fort7": open(os.path.join(rundir, "fort.7")).read(),
        "file_sha256": {fn: hashlib.sha256(open(os.path.join(rundir, fn), "rb").read()).hexdigest()
                        for fn in ("den.dat", "pre.dat", "m1.dat", "m2.dat", "den2.dat", "pre2.dat")},
        "stdout_head": "\n".join(lines[:4]),
    }
    with open(os.path.join(GOLD, "ref_ex3_vortex180.json"), "w") as f:
        json.dump(out, f, indent=1)


def make_ex3_l1() -> None:
    """ex3 with use_limiter = 1 (oracle/Makefile target ex3_l1): the Zhang-Shu flux
    limiter instead of the sweep; pins the limiter's fields and den/pre counts."""
    rundir = os.path.join(REF, "run_ex3_l1")
    text = _run(os.path.join(REF, "ex3_l1", "a"), rundir)
    den = _tokens(os.path.join(rundir, "den.dat"), 5)
    m1 = _tokens(os.path.join(rundir, "m1.dat"), 3)
    m2 = _tokens(os.path.join(rundir, "m2.dat"), 3)
    pre = _tokens(os.path.join(rundir, "pre.dat"), 3)
    lines = text.splitlines()
    l1, linf = (float(v) for v in lines[1].split())
    fields = {"rho": den[:, 2], "m1": m1[:, 2], "m2": m2[:, 2], "p": pre[:, 2]}
    limiter = [l for l in lines if l.strip().startswith("# of limiter")][0].split()[-2:]
    out = {
        "program": "ex3_vortex_accuracy_2d with use_limiter = 1, nx=ny=180, T=0.01, cfl=0.5, mag=10.0828",
        "layout": "fields in (j outer, i inner) order over the 180x180 interior",
        "sha256": {k: field_digest(v) for k, v in fields.items()},
        "sample_stride": 37,
        "sample": {k: v[::37].tolist() for k, v in fields.items()},
        "L1": l1, "Linf": linf,
        "den_count": int(limiter[0]), "pre_count": int(limiter[1]),
        "err_table_dat": open(os.path.join(rundir, "err_table.dat")).read(),
        "fort7": open(os.path.join(rundir, "fort.7")).read(),
        "file_sha256": {fn: hashlib.sha256(open(os.path.join(rundir, fn), "rb").read()).hexdigest()
                        for fn in ("den.dat", "pre.dat", "m1.dat", "m2.dat", "den2.dat", "pre2.dat")},
        "stdout_head": "\n".join(lines[:4]),
        "limiter_line": [l for l in lines if l.strip().startswith("# of limiter")][0],
    }
    with open(os.path.join(GOLD, "ref_ex3_l1_vortex180.json"), "w") as f:
        json.dump(out, f, indent=1)


def make_ex2() -> None:
    """The three ex2 builds of oracle/Makefile (ex2, ex2_cfg1, ex2_stress)."""
    cases = {
        "shipped": ("ex2", "source as shipped (n=800, cfl=1.2, swp_num=50, E_amb=1d-9/gamma1, T=0.001), "
                           "built with -O2 -fdefault-real-8 (FP64)"),
        "cfg1": ("ex2_cfg1", "cfl=0.6, swp_num=2, E_amb=1d-12, otherwise as shipped"),
        "stress": ("ex2_stress", "cfl=1.5, max_iter=40, otherwise as shipped (stops after 40 steps)"),
    }
    out = {"layout": "den/pre/vel columns of den.dat/pre.dat/vel.dat over the 800 interior cells",
           "sample_stride": 23}
    for key, (build, desc) in cases.items():
        rundir = os.path.join(REF, "run_" + build)
        text = _run(os.path.join(REF, build, "a"), rundir)
        fields = {nm: _tokens(os.path.join(rundir, nm + ".dat"), 2)[:, 1] for nm in ("den", "pre", "vel")}
        lines = text.splitlines()
        steps = [int(l.split()[0]) for l in lines if l.strip().endswith("%")]
        at = [i for i, l in enumerate(lines) if l.strip().startswith("L1 =")][0]
        l1, linf = (float(v) for v in lines[at + 1].split())
        rec = {
            "program": "ex2_sedov_1d, " + desc,
            "sha256": {k: field_digest(v) for k, v in fields.items()},
            "sample": {k: v[::23].tolist() for k, v in fields.items()},
            "L1": l1, "Linf": linf,
            "stats": _parse_stats(text),
            "last_progress_step": steps[-1] if steps else 0,
            "file_sha256": {fn: hashlib.sha256(open(os.path.join(rundir, fn), "rb").read()).hexdigest()
                            for fn in ("den.dat", "pre.dat", "vel.dat")},
            "stdout_sha256": hashlib.sha256(text.encode()).hexdigest(),
        }
        mr = [l for l in lines if l.strip().startswith("min_rho:")]
        if mr:
            rec["min_rho"] = float(mr[0].split()[-1])
        out[key] = rec
    with open(os.path.join(GOLD, "ref_ex2_sedov1d.json"), "w") as f:
        json.dump(out, f, indent=1)


def _ex5_entry(build: str, nx: int, ny: int, program: str, stride: int) -> dict:
    """Runs ex5 build copy `build` (grid nx x ny) and forms its fixture entry."""
    rundir = os.path.join(REF, "run_" + build)
    text = _run(os.path.join(REF, build, "a"), rundir)
    fields = {}
    for name, fn, skip in (("log_rho", "den.dat", 2), ("log_p", "pre.dat", 1), ("log_p_over_rho", "m1.dat", 1)):
        body = open(os.path.join(rundir, fn)).read().split("\n", skip)[skip]
        t = np.array(body.split(), dtype=np.float64).reshape(-1, 3)
        fields[name] = t[ny * nx:, 2]
    steps = [int(l.split()[0]) for l in text.splitlines() if l.strip().endswith("%")]
    return {
        "program": program,
        "sha256": {k: field_digest(v) for k, v in fields.items()},
        "sample": {k: v[::stride].tolist() for k, v in fields.items()},
        "stats": _parse_stats(text),
        "last_progress_step": steps[-1] if steps else 0,
        "file_sha256": {fn: hashlib.sha256(open(os.path.join(rundir, fn), "rb").read()).hexdigest()
                        for fn in ("den.dat", "pre.dat", "m1.dat")},
        "stdout_sha256": hashlib.sha256(text.encode()).hexdigest(),
    }


def make_ex5_small() -> None:
    """oracle/Makefile targets ex5_small (T = 0.001) and ex5_short (T = 1.5e-4), 160 x 40."""
    out = {"layout": "log fields of the unmirrored half (the file's second half), (j outer, i inner), "
                     "160 x 40 interior"}
    for key, build, desc in (("full", "ex5_small", "T = 0.001"), ("short", "ex5_short", "T = 1.5e-4")):
        out[key] = _ex5_entry(build, 160, 40, f"ex5_mach2000_jet with nx = 160, ny = 40 (PARAMETER line), {desc}, "
                                              "otherwise as shipped", 97)
    with open(os.path.join(GOLD, "ref_ex5_jet160x40.json"), "w") as f:
        json.dump(out, f, indent=1)


def make_ex5_ship() -> None:
    """oracle/Makefile target ex5_ship: the shipped 800 x 200 grid to T = 5e-5 (about half
    an hour on one core)."""
    out = {"layout": "log fields of the unmirrored half (the file's second half), (j outer, i inner), "
                     "800 x 200 interior",
           "ship": _ex5_entry("ex5_ship", 800, 200, "ex5_mach2000_jet as shipped (800 x 200, cfl 0.25) with "
                                                    "time_terminal = 0.00005", 997)}
    with open(os.path.join(GOLD, "ref_ex5_jet800x200.json"), "w") as f:
        json.dump(out, f, indent=1)


def _run_mpi1(build: str) -> tuple:
    """Runs a one-rank build copy with its output directory as the first argument (as the
    shipped programs take it); returns (run dir, printout)."""
    rundir = os.path.join(REF, "run_" + build)
    os.makedirs(rundir, exist_ok=True)
    res = subprocess.run(["bash", "-c", "ulimit -s unlimited; exec " + os.path.join(REF, build, "a") + " " + rundir],
                         cwd=rundir, check=True, capture_output=True, text=True)
    with open(os.path.join(rundir, "stdout.txt"), "w") as f:
        f.write(res.stdout)
    return rundir, res.stdout


def _printout_minus_timer(text: str) -> str:
    """The printout without the `CPU Time` line (wall-clock dependent); its step count and
    final time are kept apart."""
    return "\n".join(l for l in text.splitlines() if not l.strip().startswith("CPU Time")) + "\n"


def _timer_line(text: str) -> tuple:
    l = [l for l in text.splitlines() if l.strip().startswith("CPU Time")][0].split()
    return float(l[2]), int(l[3]), l[4]


def make_ex4() -> None:
    cases = {
        "n64": ("ex4_64", 64, 0.002, "total_x = total_y = 64, time_terminal = 0.002"),
        "n128": ("ex4_128", 128, 0.01, "total_x = total_y = 128, time_terminal = 0.01"),
        "n640w": ("ex4_640w", 640, 1.2e-5, "total_x = total_y = 640 (as shipped), time_terminal = 1.2e-5"),
    }
    out = {"layout": "value columns of den.dat / pre.dat / m1.dat / m2.dat: (j outer, i inner) over the "
                     "interior; pre = gamma1*(E - 0.5*m1**2/rho - 0.5*m2**2/rho) as the program writes it",
           "sample_stride": 97}
    for key, (build, n, T, desc) in cases.items():
        rundir, text = _run_mpi1(build)
        fields = {}
        for nm in ("den", "pre", "m1", "m2"):
            fields[nm] = _tokens_skip(os.path.join(rundir, nm + ".dat"), 3, 2 if nm == "den" else 1)[:, 2]
        secs, steps, tfin = _timer_line(text)
        out[key] = {
            "program": "ex4_sedov_2d, one-rank build copy (use MPI -> include 'mpif.h', cut_x = cut_y = 1), " + desc,
            "n": n, "t_end": T, "steps": steps, "t_printed": tfin, "reference_seconds_1core": secs,
            "sha256": {k: field_digest(v) for k, v in fields.items()},
            "sample": {k: v[::97].tolist() for k, v in fields.items()},
            "stats": _parse_stats(text),
            "file_sha256": {fn: hashlib.sha256(open(os.path.join(rundir, fn), "rb").read()).hexdigest()
                            for fn in ("den.dat", "pre.dat", "m1.dat", "m2.dat")},
            "printout_minus_timer": _printout_minus_timer(text),
        }
    with open(os.path.join(GOLD, "ref_ex4_sedov.json"), "w") as f:
        json.dump(out, f, indent=1)


def make_ex6() -> None:
    cases = {
        "total8": ("ex6_8", 8, 2.3, "total = 8 (104 x 88), time_terminal = 2.3 (as shipped)"),
        "total16": ("ex6_16", 16, 0.5, "total = 16 (208 x 176), time_terminal = 0.5"),
    }
    out = {"layout": "value columns of den.dat / pre.dat: (j outer, i inner) over the interior, 0 where "
                     "i < total and j < 6*total (the program blanks the step)", "sample_stride": 53}
    for key, (build, total, T, desc) in cases.items():
        rundir, text = _run_mpi1(build)
        fields = {nm: _tokens_skip(os.path.join(rundir, nm + ".dat"), 3, 1)[:, 2] for nm in ("den", "pre")}
        secs, steps, tfin = _timer_line(text)
        out[key] = {
            "program": "ex6_shock_diffraction, one-rank build copy (use MPI -> include 'mpif.h', cut_x = cut_y = 1), "
                       + desc,
            "total": total, "t_end": T, "steps": steps, "t_printed": tfin, "reference_seconds_1core": secs,
            "sha256": {k: field_digest(v) for k, v in fields.items()},
            "sample": {k: v[::53].tolist() for k, v in fields.items()},
            "stats": _parse_stats(text),
            "file_sha256": {fn: hashlib.sha256(open(os.path.join(rundir, fn), "rb").read()).hexdigest()
                            for fn in ("den.dat", "pre.dat")},
            "printout_minus_timer": _printout_minus_timer(text),
        }
    with open(os.path.join(GOLD, "ref_ex6_step.json"), "w") as f:
        json.dump(out, f, indent=1)


def make_ex7() -> None:
    cases = {
        "total8": ("ex7_8", 8, 0.2, "total = 8 (40 x 40), time_terminal = 0.2"),
        "total16": ("ex7_16", 16, 0.2, "total = 16 (80 x 80), time_terminal = 0.2"),
    }
    out = {"layout": "value columns of den.dat / pre.dat / rate.dat: (j outer, i inner) over the interior, "
                     "with the program's own values 100 / 300 / 20 where i <= total and j < step*total; "
                     "pre = gamma1*(E - 0.5*m1**2/rho - 0.5*m2**2/rho - q_const*rhoY), rate = rhoY/rho",
           "sample_stride": 41}
    for key, (build, total, T, desc) in cases.items():
        rundir, text = _run_mpi1(build)
        fields = {nm: _tokens_skip(os.path.join(rundir, nm + ".dat"), 3, 1)[:, 2] for nm in ("den", "pre", "rate")}
        secs, steps, tfin = _timer_line(text)
        out[key] = {
            "program": "ex7_detonation, one-rank build copy (use MPI -> include 'mpif.h', cut_x = cut_y = 1, "
                       "mpirank initialised to 0), " + desc,
            "total": total, "t_end": T, "steps": steps, "t_printed": tfin, "reference_seconds_1core": secs,
            "sha256": {k: field_digest(v) for k, v in fields.items()},
            "sample": {k: v[::41].tolist() for k, v in fields.items()},
            "stats": _parse_stats(text),
            "file_sha256": {fn: hashlib.sha256(open(os.path.join(rundir, fn), "rb").read()).hexdigest()
                            for fn in ("den.dat", "pre.dat", "rate.dat")},
            "printout_minus_timer": _printout_minus_timer(text),
        }
        if total == 8:
            out[key]["fields"] = {k: v.tolist() for k, v in fields.items()}
    with open(os.path.join(GOLD, "ref_ex7_detonation.json"), "w") as f:
        json.dump(out, f, indent=1)


def make_ex7_ship() -> None:
    """ex7 at its shipped grid (total = 80, 400 x 400; oracle/Makefile target ex7_80) from
    t = 0 to the program's own stop: its progress lines (every 50 steps), the STOP text
    and the run time on one core (about 25 minutes)."""
    import time as _t
    rundir = os.path.join(REF, "run_ex7_80")
    os.makedirs(rundir, exist_ok=True)
    t0 = _t.perf_counter()
    res = subprocess.run(["bash", "-c", "ulimit -s unlimited; exec " + os.path.join(REF, "ex7_80", "a") + " " + rundir],
                         cwd=rundir, capture_output=True, text=True)
    secs = _t.perf_counter() - t0
    text = res.stdout + res.stderr
    progress = [[int(l.split()[0]), l.split()[1]] for l in text.splitlines() if l.strip().endswith("%")]
    stop = [l.strip() for l in text.splitlines() if "STOP" in l]
    out = {"program": "ex7_detonation, one-rank build copy (use MPI -> include 'mpif.h', cut_x = cut_y = 1, "
                      "mpirank initialised to 0), total = 80 (400 x 400) and time_terminal = 0.6 as shipped",
           "progress": progress, "stop": stop, "returncode": res.returncode,
           "reference_seconds_1core": round(secs, 1)}
    with open(os.path.join(GOLD, "ref_ex7_detonation400.json"), "w") as f:
        json.dump(out, f, indent=1)


def _tokens_skip(fn: str, ncol: int, skip: int) -> np.ndarray:
    body = open(fn).read().split("\n", skip)[skip]
    return np.array(body.split(), dtype=np.float64).reshape(-1, ncol)


def make_ex5() -> None:
    rundir = os.path.join(REF, "run_ex5")
    if not os.path.exists(os.path.join(rundir, "DONE")):
        print("ex5 run not finished; start it with: cd oracle/_ref/run_ex5 && ../ex5/a")
        return
    text = open(os.path.join(rundir, "stdout.txt")).read()
    nx, ny = 800, 200
    fields = {}
    for name, fn in (("log_rho", "den.dat"), ("log_p", "pre.dat"), ("log_p_over_rho", "m1.dat")):
        lines = open(os.path.join(rundir, fn)).read().split("\n", 3)
        body = lines[3] if fn == "den.dat" else open(os.path.join(rundir, fn)).read().split("\n", 1)[1]
        t = np.array(body.split(), dtype=np.float64).reshape(-1, 3)
        fields[name] = t[ny * nx:, 2]  # second (unmirrored) half: j = 1..ny, i = 1..nx
    out = {
        "program": "ex5_mach2000_jet (unmodified), 800x200 half domain, T=0.001, cfl=0.25",
        "layout": "log fields in (j outer, i inner) order over the 800x200 interior",
        "sha256": {k: field_digest(v) for k, v in fields.items()},
        "sample_stride": 101,
        "sample": {k: v[::101].tolist() for k, v in fields.items()},
        "stats": _parse_stats(text),
    }
    with open(os.path.join(GOLD, "ref_ex5_jet.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    what = sys.argv[1:] or ["weno", "weno6", "ex3", "ex3_l1", "ex2", "ex5s", "ex5"]
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    if "weno" in what:
        make_weno()
    if "weno6" in what:
        make_weno6()
    if "ex3" in what:
        make_ex3()
    if "ex3_l1" in what:
        make_ex3_l1()
    if "ex2" in what:
        make_ex2()
    if "ex5s" in what:
        make_ex5_small()
    if "ex5" in what:
        make_ex5()
    if "ex5ship" in what:
        make_ex5_ship()
    if "ex4" in what:
        make_ex4()
    if "ex6" in what:
        make_ex6()
    if "ex7" in what:
        make_ex7()
    if "ex7ship" in what:
        make_ex7_ship()
