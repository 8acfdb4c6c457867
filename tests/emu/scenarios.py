"""Scenarios of the CPU emulation build (tests/emu/_build/libswf_emu.so) against
the oracle, one per process:  python tests/emu/scenarios.py <name>  prints one
JSON line.  The emulated library is the package's own device code compiled for
the host; it is loaded here only, never by the package."""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import sweeping_fd_weno_amd._lib as L  # noqa: E402

L.library_path = os.path.join(ROOT, "tests", "emu", "_build", "libswf_emu.so")
import sweeping_fd_weno_amd as S  # noqa: E402
from oracle import oracle as O  # noqa: E402

STATS = ("sweep_count", "den_count", "pre_count", "total_sweep", "max_sweep", "total_neg", "max_neg", "total_pos",
         "max_pos")


def ocfg(cfg):
    d = cfg.as_dict()
    d.pop("precision")
    d.pop("device")
    return O.make_config(**d)


def run_both(cfg, u, tfinal, nsteps):
    """Oracle and emulated library from the same post-bc state; returns a dict."""
    o = O.Oracle(ocfg(cfg))
    o.set_state(u)
    o.bc()
    u = o.state()
    rc, to, no = o.run(0.0, tfinal, max_steps=nsteps)
    s = S.Solver(cfg)
    s.upload(u)
    ts, ns = s.advance(0.0, tfinal, max_steps=nsteps)
    a, b = s.download(), o.state()
    so, se = o.stats(), s.stats()
    return {"rc": rc, "steps": [no, ns], "time": [to, ts],
            "differ": int((a.view(np.int64) != b.view(np.int64)).sum()),
            "stats": {k: [so[k], se[k]] for k in STATS}}


def fast_vs_oracle(cfg, u, tfinal, nsteps):
    """A fast mode against the oracle's exact run: relative L1 density difference
    and both runs' sweep statistics."""
    o = O.Oracle(ocfg(cfg))
    o.set_state(u)
    o.bc()
    u = o.state()
    rc, to, no = o.run(0.0, tfinal, max_steps=nsteps)
    s = S.Solver(cfg)
    s.upload(u)
    ts, ns = s.advance(0.0, tfinal, max_steps=nsteps)
    a, b = s.download()[3:-3, 3:-3, 0], o.state()[3:-3, 3:-3, 0]
    so, se = o.stats(), s.stats()
    return {"rc": rc, "steps": [no, ns], "l1": float(np.abs(a - b).sum() / np.abs(b).sum()),
            "stats": {k: [so[k], se[k]] for k in STATS}}


def first_dt(cfg):
    """dt of the first step after upload of the initial state and swf_bc."""
    o = O.Oracle(ocfg(cfg))
    o.set_state(S.init_jet(cfg))
    o.bc()
    rc, dto = o.step(0.0, 0.001)
    s = S.Solver(cfg)
    s.upload(S.init_jet(cfg))
    s.bc()
    return {"rc": rc, "dt": [dto, s.step(0.0, 0.001)]}


def crafted_sweep(n, nbad, seed, chain=0, psum=0):
    """swf_sweep on a state with nbad cells of slightly negative pressure (runs of
    consecutive flagged cells included) and three near-vacuum cells, against the
    oracle's sweep: fields bitwise and statistics equal."""
    cfg = S.sedov_config(n, "fp64_exact")
    cfg.psum_check = psum  # the step program's pressure-sum check on these walks
    rng = np.random.default_rng(seed)
    u = np.zeros(S.state_shape(cfg))
    u[:, :, 0] = 1.0 + 0.1 * rng.random(u.shape[:2])
    u[:, :, 1] = 0.05 * rng.standard_normal(u.shape[:2])
    u[:, :, 2] = 0.05 * rng.standard_normal(u.shape[:2])
    u[:, :, 3] = 2.5 + rng.random(u.shape[:2])
    placed = 0
    while placed < nbad:
        i, j = int(rng.integers(1, n + 1)), int(rng.integers(1, n + 1))
        for q in range(int(rng.integers(1, 5))):
            ii = min(i + q, n)
            u[j + 2, ii + 2, 3] = 0.5 * (u[j + 2, ii + 2, 1] ** 2 + u[j + 2, ii + 2, 2] ** 2) / u[j + 2, ii + 2, 0] \
                - 1e-3 * rng.random()
            placed += 1
    for pos in range(n * n // 3, n * n // 3 + chain):  # consecutive snake-order positions
        j = pos // n + 1
        r = pos - (j - 1) * n
        i = r + 1 if j % 2 == 1 else n - r
        u[j + 2, i + 2, 3] = 0.5 * (u[j + 2, i + 2, 1] ** 2 + u[j + 2, i + 2, 2] ** 2) / u[j + 2, i + 2, 0] - 1e-4
    for _ in range(3):
        i, j = int(rng.integers(1, n + 1)), int(rng.integers(1, n + 1))
        u[j + 2, i + 2, 0] = 1e-14
        u[j + 2, i + 2, 1:3] = 0.0
        u[j + 2, i + 2, 3] = 1e-15
    o = O.Oracle(ocfg(cfg))
    o.set_state(u)
    rc = o.sweep()
    s = S.Solver(cfg)
    s.upload(u)
    try:
        s.sweep()
        rce = 0
    except S.SwfError as e:
        rce = e.code
    if psum:
        return {"rc": [rc, rce]}
    a, b = s.download()[3:-3, 3:-3], o.state()[3:-3, 3:-3]
    so, se = o.stats(), s.stats()
    return {"rc": rc, "steps": [0, 0], "time": [0.0, 0.0], "differ": int((a.view(np.int64) != b.view(np.int64)).sum()),
            "stats": {k: [so[k], se[k]] for k in STATS}}


def batch_vs_single(prec, nsteps):
    """A batch of a Sedov, a vortex, a full jet and a step diffraction instance against
    each run alone (same precision): every field bitwise, the same time and steps.  The
    step instance starts from its initial data without bc, as its program does."""
    cfgs = [S.sedov_config(40, prec), S.vortex_config(36, prec), S.jet_config(48, 24, prec, cfl=0.5, half=False),
            S.step_config(4, prec, mach=4.0)]
    inits = [S.init_sedov(cfgs[0]), S.init_vortex(cfgs[1]), S.init_jet(cfgs[2]), S.init_step(cfgs[3])]
    tf = np.array([1.0, 0.01, 0.001, 2.3])
    n = len(cfgs)
    bt = S.Batch(cfgs)
    for i in range(n):
        bt.instance(i).upload(inits[i])
        if i < 3:
            bt.instance(i).bc()
    tb, nb, stb = bt.advance(np.zeros(n), tf, max_steps=nsteps)
    out = {"rc": int(stb.max()), "steps": nb.tolist(), "differ": 0, "time_equal": True}
    for i in range(n):
        s = S.Solver(cfgs[i])
        s.upload(inits[i])
        if i < 3:
            s.bc()
        ts, ns = s.advance(0.0, tf[i], max_steps=nsteps)
        a, b = bt.instance(i).download(), s.download()
        out["differ"] += int((a.view(np.int64) != b.view(np.int64)).sum())
        out["time_equal"] = bool(out["time_equal"] and ts == tb[i] and ns == nb[i])
        out.setdefault("sweeps", []).append([s.stats()["sweep_count"], bt.instance(i).stats()["sweep_count"]])
        s.close()
    bt.close()
    return out


def step_run(prec, total, nsteps, mach=5.09):
    """The step program (no bc before the first step, as there): oracle (exact) and
    the emulated library from the same initial state."""
    cfg = S.step_config(total, prec, mach=mach)
    u = S.init_step(cfg)
    o = O.Oracle(ocfg(cfg))
    o.set_state(u)
    rc, to, no = o.run(0.0, 2.3, max_steps=nsteps)
    s = S.Solver(cfg)
    s.upload(u)
    try:
        ts, ns = s.advance(0.0, 2.3, max_steps=nsteps)
        rce = 0
    except S.SwfError as e:
        ts, ns, rce = None, None, e.code
    a, b = s.download(), o.state()
    so, se = o.stats(), s.stats()
    return {"rc": [rc, rce], "steps": [no, ns], "time": [to, ts],
            "differ": int((a.view(np.int64) != b.view(np.int64)).sum()),
            "l1": float(np.abs(a[3:-3, 3:-3, 0] - b[3:-3, 3:-3, 0]).sum() / np.abs(b[3:-3, 3:-3, 0]).sum()),
            "stats": {k: [so[k], se[k]] for k in STATS}}


def line_run(prec, n, cfl, swp, shipped, nsteps):
    """The 1D program: oracle (exact) and the emulated library from the same state."""
    cfg = S.sedov1d_config(n, prec, cfl=cfl, swp_num=swp)
    u = S.init_sedov1d(cfg, shipped_energy=shipped)
    o = O.Oracle1D(O.sedov1d_config(n, cfl=cfl, swp_num=swp))
    o.set_state(u)
    to, no = o.run(0.0, 0.001, nsteps)
    s = S.Solver1D(cfg)
    s.upload(u)
    ts, ns = s.advance(0.0, 0.001, nsteps)
    a, b = s.download(), o.state()
    so, se = o.stats(), s.stats()
    q, ng, ps = s.track()
    qo, ngo, pso = o.track()
    return {"steps": [no, ns], "time": [to, ts],
            "differ": int((a.view(np.int64) != b.view(np.int64)).sum()),
            "l1_rel": float(np.abs(a[:, 0] - b[:, 0]).sum() / np.abs(b[:, 0]).sum()),
            "stats": {k: [so[k], se[k]] for k in STATS},
            "track_equal": bool(np.array_equal(q, qo) and np.array_equal(ng, ngo) and np.array_equal(ps, pso))}


def ex7_sweeps():
    """swf7_sweep on the crafted states of tests/sweep7_cases.py against the oracle's
    serial restatement of the program's sweep: per case the stop code, the counts and
    the number of differing state values."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from sweep7_cases import CASES, crafted7
    from sweeping_fd_weno_amd import detonation as D
    res = []
    for total, kind, seed in CASES:
        u = crafted7(total, kind, seed)
        ref = O.sweep7(u, total)
        s = D.Solver7(D.det_config(total))
        s.upload(u)
        code = s.sweep()
        got = s.download()
        st = s.stats()
        trk = s.track()
        s.close()
        res.append({"case": [total, kind, seed], "code": [code, ref["code"]],
                    "counts": [[st["sweep_count"], st["den_count"], st["pre_count"]],
                               [ref["sweep_count"], ref["den_count"], ref["pre_count"]]],
                    "track": [trk[-1].tolist() if len(trk) else [],
                              [ref["q"], ref["neg"], ref["pos"]] if ref["code"] == 0 else [0, 0, 0]],
                    "differ": int((got != ref["u"]).sum()), "changed": int((ref["u"] != u).sum())})
    return {"cases": res}


def main(name):
    if name == "sedov_sweeps":
        cfg = S.sedov_config(32, "fp64_exact")
        out = run_both(cfg, S.init_sedov(cfg), 1.0, 10)
    elif name.startswith("ex7_"):
        # the detonation program against its reference run (fixture): fields and statistics
        from sweeping_fd_weno_amd import detonation as D
        import hashlib
        total = int(name[len("ex7_"):])
        r = D.run(total, 0.2)
        out = {"steps": r["steps"], "code": r["code"], "stats": r["stats"],
               "sha256": {k: hashlib.sha256(np.ascontiguousarray(v, dtype="<f8").tobytes()).hexdigest()
                          for k, v in r["fields"].items()}}
    elif name == "sweep7_crafted":
        out = ex7_sweeps()
    elif name == "jet_full_sweeps":
        cfg = S.jet_config(64, 32, "fp64_exact", cfl=0.5, half=False)
        out = run_both(cfg, S.init_jet(cfg), 0.001, 30)
    elif name.startswith("jet_full_sweeps_"):
        cfg = S.jet_config(64, 32, name[len("jet_full_sweeps_"):], cfl=0.5, half=False)
        out = fast_vs_oracle(cfg, S.init_jet(cfg), 0.001, 25)
    elif name == "jet_half":
        cfg = S.jet_config(80, 20, "fp64_exact", cfl=0.25, half=True)
        out = run_both(cfg, S.init_jet(cfg), 0.001, 40)
    elif name == "vortex_flux_limiter":
        cfg = S.vortex_config(90, "fp64_exact", use_limiter=1)
        out = run_both(cfg, S.init_vortex(cfg), 0.01, 100)
    elif name.startswith("crafted_psum_"):
        n, nbad, seed = (int(v) for v in name[len("crafted_psum_"):].split("_"))
        out = crafted_sweep(n, nbad, seed, 0, 1)
    elif name.startswith("crafted_sweep_"):
        n, nbad, *rest = (int(v) for v in name[len("crafted_sweep_"):].split("_"))
        out = crafted_sweep(n, nbad, 7, rest[0] if rest else 0)
    elif name.startswith("batch_"):
        out = batch_vs_single(name[len("batch_"):], 25)
    elif name.startswith("first_dt_"):
        out = first_dt(S.jet_config(40, 20, name[len("first_dt_"):], cfl=0.5, half=False))
    elif name.startswith("step_"):
        prec, total, nsteps = name[len("step_"):].rsplit("_", 2)
        out = step_run(prec, int(total), int(nsteps))
    elif name.startswith("line_stress_"):
        out = line_run(name[len("line_stress_"):], 800, 1.5, 50, True, 40)
    elif name == "line_cfg1":
        out = line_run("fp64_exact", 800, 0.6, 2, False, 30)
    elif name == "line_small":
        out = line_run("fp64_exact", 60, 1.5, 50, True, 12)
    else:
        raise SystemExit("unknown scenario " + name)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main(sys.argv[1])
