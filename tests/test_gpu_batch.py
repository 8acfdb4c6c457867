"""Batches (swf_batch_*) on the GPU: independent instances advanced together by one
launch of each stage kernel per RK stage.  Every instance must equal its run alone
bit for bit, under every tile schedule, and the autotuner must leave the states
untouched.  The instances include two shock diffractions past a step (the step
program's wenofd has batch kernels of its own; no bc before their first step)."""
import numpy as np
import pytest

import sweeping_fd_weno_amd as S

pytestmark = pytest.mark.gpu


def _instances(prec):
    cfgs = [S.sedov_config(64, prec), S.vortex_config(45, prec), S.jet_config(80, 40, prec, cfl=0.5, half=False),
            S.sedov_config(37, prec), S.vortex_config(90, prec), S.step_config(8, prec, mach=3.5),
            S.step_config(16, prec, mach=5.5)]
    inits = [S.init_sedov(cfgs[0]), S.init_vortex(cfgs[1]), S.init_jet(cfgs[2]),
             S.init_sedov(cfgs[3], e_corner_num=0.7), S.init_vortex(cfgs[4], mag=7.5), S.init_step(cfgs[5]),
             S.init_step(cfgs[6])]
    tf = np.array([1.0, 0.01, 0.001, 1.0, 0.01, 2.3, 2.3])
    return cfgs, inits, tf


def _singles(cfgs, inits, tf, nsteps):
    out = []
    for c, u, t in zip(cfgs, inits, tf):
        s = S.Solver(c)
        s.upload(u)
        if c.bc != S.BC_STEP:
            s.bc()
        ts, ns = s.advance(0.0, t, max_steps=nsteps)
        out.append((ts, ns, s.download(), s.stats()))
        s.close()
    return out


def _batch(cfgs, inits, tf, nsteps, schedule=None, tune=False):
    bt = S.Batch(cfgs)
    for i, u in enumerate(inits):
        bt.instance(i).upload(u)
        if cfgs[i].bc != S.BC_STEP:
            bt.instance(i).bc()
    info = None
    if tune:
        info = bt.tune(5)
    elif schedule is not None:
        bt.set_schedule(schedule)
    t, n, st = bt.advance(np.zeros(len(cfgs)), tf, max_steps=nsteps)
    res = [(t[i], n[i], bt.instance(i).download(), bt.instance(i).stats()) for i in range(len(cfgs))]
    bt.close()
    return res, st, info


@pytest.mark.parametrize("prec", ["fp64_exact", "fp64", "fp32_bulk"])
def test_batch_equals_single_runs(prec):
    cfgs, inits, tf = _instances(prec)
    single = _singles(cfgs, inits, tf, 40)
    res, st, _ = _batch(cfgs, inits, tf, 40)
    assert (st == 0).all()
    for i, (a, b) in enumerate(zip(res, single)):
        assert a[0] == b[0] and a[1] == b[1], i
        assert np.array_equal(a[2].view(np.int64), b[2].view(np.int64)), i
        for key in ("sweep_count", "total_sweep", "total_neg", "total_pos", "den_count", "pre_count"):
            assert a[3][key] == b[3][key], (i, key)
    assert single[2][3]["sweep_count"] > 0  # the jet sweeps within 40 steps


@pytest.mark.parametrize("schedule", [0, 2])
def test_batch_schedule_does_not_change_results(schedule):
    cfgs, inits, tf = _instances("fp32_bulk")
    base, _, _ = _batch(cfgs, inits, tf, 30, schedule=1)
    other, _, _ = _batch(cfgs, inits, tf, 30, schedule=schedule)
    for a, b in zip(base, other):
        assert a[1] == b[1] and np.array_equal(a[2].view(np.int64), b[2].view(np.int64))


def test_batch_tune_leaves_states_untouched():
    cfgs, inits, tf = _instances("fp64")
    single = _singles(cfgs, inits, tf, 30)
    res, st, (pick, ms) = _batch(cfgs, inits, tf, 30, tune=True)
    assert pick in (0, 1, 2) and all(m > 0 for m in ms)
    for a, b in zip(res, single):
        assert a[1] == b[1] and np.array_equal(a[2].view(np.int64), b[2].view(np.int64))


def test_batch_fp32_full_rejects_jets_and_matches_single_runs():
    """SWF_FP32_FULL batches: Sedov, vortex and step instances equal their runs alone bit
    for bit (the step instance runs the FP32-bulk form in both); a jet is rejected (its
    FP32-full run would use the FP32-bulk kernels, which the other instances do not)."""
    cfgs, inits, tf = _instances("fp32_full")
    with pytest.raises(S.SwfError):
        S.Batch(cfgs)
    keep = [i for i, c in enumerate(cfgs) if c.bc not in (S.BC_JET_HALF, S.BC_JET_FULL)]
    cfgs, inits, tf = [cfgs[i] for i in keep], [inits[i] for i in keep], tf[keep]
    single = _singles(cfgs, inits, tf, 30)
    res, st, _ = _batch(cfgs, inits, tf, 30)
    assert (st == 0).all()
    for i, (a, b) in enumerate(zip(res, single)):
        assert a[0] == b[0] and a[1] == b[1], i
        assert np.array_equal(a[2].view(np.int64), b[2].view(np.int64)), i
