"""The batched ensemble of the benchmark's batching row: 48 independent solves of uneven
sizes (90^2 .. 1040 x 880; 16 Sedov corners, 16 low-density vortices, 16 shock
diffractions past a step, drawn with splitmix64 from seed 20261203 by
sweeping_fd_weno_amd.batch) advanced in one batch.

- FP64-exact, 20 steps: every instance bit for bit equal to its run alone, with equal
  statistics; the first instance of each family (Sedov 90^2, vortex 90^2, step 104 x 88)
  bitwise to the oracle.
- FP32-bulk and FP32-full, 10 steps: each instance's relative L1 density difference
  from the FP64-exact batch within 1e-6 (the step problem is chaotic beyond ~10 steps on its
  coarse grids: a 1e-15 perturbation grows to 1.4e-4 after 20 steps).
- FP32-bulk and FP32-full, 200 steps (the ensemble run of `python -m sweeping_fd_weno_amd.batch`): no
  stop, every instance finite with positive density and pressure.
"""
import numpy as np
import pytest

import sweeping_fd_weno_amd as S
from sweeping_fd_weno_amd import batch as B
from oracle import oracle as O

pytestmark = pytest.mark.gpu

STATS = ("sweep_count", "den_count", "pre_count", "total_sweep", "max_sweep", "total_neg", "max_neg", "total_pos",
         "max_pos")


def _run_batch(ens, steps):
    bt = S.Batch([c for _, c, _, _ in ens])
    for i, (_, c, u, _) in enumerate(ens):
        bt.instance(i).upload(u)
        if c.bc != S.BC_STEP:  # the step program applies no bc before its first step
            bt.instance(i).bc()
    t, n, st = bt.advance(np.zeros(bt.n), np.array([tf for *_, tf in ens]), max_steps=steps)
    out = [(float(t[i]), int(n[i]), bt.instance(i).download(), bt.instance(i).stats()) for i in range(bt.n)]
    bt.close()
    return out, st


def _single(label, c, u, tf, steps):
    s = S.Solver(c)
    s.upload(u)
    if c.bc != S.BC_STEP:
        s.bc()
    t, n = s.advance(0.0, tf, max_steps=steps)
    r = (t, n, s.download(), s.stats())
    s.close()
    return r


def _ocfg(cfg):
    d = cfg.as_dict()
    d.pop("precision")
    d.pop("device")
    return O.make_config(**d)


def test_ensemble48_exact_equals_single_runs_and_oracle():
    ens = B.ensemble("fp64_exact")
    assert len(ens) == 48
    print("ensemble48 fp64_exact, 20 steps", flush=True)
    res, st = _run_batch(ens, 20)
    assert (st == 0).all()
    for i, (label, c, u, tf) in enumerate(ens):
        one = _single(label, c, u, tf, 20)
        a = res[i]
        assert a[1] == one[1] and a[0] == one[0], label
        assert a[1] == 20 or not S.time_continue(a[0], tf), label  # 20 steps or its final time
        assert np.array_equal(a[2].view(np.int64), one[2].view(np.int64)), label
        for k in STATS:
            assert a[3][k] == one[3][k], (label, k)
    for fam in ("sedov", "vortex", "step"):
        i = next(i for i, e in enumerate(ens) if e[0].startswith(fam))
        label, c, u, tf = ens[i]
        print("oracle", label, flush=True)
        o = O.Oracle(_ocfg(c))
        o.set_state(u)
        if c.bc != S.BC_STEP:
            o.bc()
        rc, t, k = o.run(0.0, tf, max_steps=20)
        assert rc == 0 and k == res[i][1] and t == res[i][0], label
        assert np.array_equal(o.state().view(np.int64), res[i][2].view(np.int64)), label
        so = o.stats()
        for key in STATS:
            assert so[key] == res[i][3][key], (label, key)


@pytest.mark.parametrize("prec", ["fp32_bulk", "fp32_full"])
def test_ensemble48_fp32_close_to_exact(prec):
    steps = 10
    ex, _ = _run_batch(B.ensemble("fp64_exact"), steps)
    ens = B.ensemble(prec)
    fb, st = _run_batch(ens, steps)
    assert (st == 0).all()
    worst = {}
    for i, (label, *_rest) in enumerate(ens):
        a = ex[i][2][3:-3, 3:-3, 0]
        b = fb[i][2][3:-3, 3:-3, 0]
        assert ex[i][1] == fb[i][1], label
        l1 = float(np.abs(a - b).sum() / np.abs(a).sum())
        fam = label[:4]
        worst[fam] = max(worst.get(fam, 0.0), l1)
        assert l1 <= 1e-6, (label, l1)
    print("worst relative L1 density per family:", worst)


@pytest.mark.parametrize("prec", ["fp32_bulk", "fp32_full"])
def test_ensemble48_fp32_200_steps_positive(prec):
    ens = B.ensemble(prec)
    res, st = _run_batch(ens, 200)
    assert (st == 0).all()
    for i, (label, c, _, _) in enumerate(ens):
        t, n, u, s = res[i]
        v = u[3:-3, 3:-3]
        assert np.isfinite(v).all(), label
        p = (c.gamma - 1.0) * (v[..., 3] - 0.5 * (v[..., 1] ** 2 + v[..., 2] ** 2) / v[..., 0])
        assert v[..., 0].min() > 0 and p.min() > 0, label
        assert n == 200 or t >= ens[i][3] - 1e-10, label
