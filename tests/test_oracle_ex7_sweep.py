"""The oracle's restatement of the detonation program's sweep (oracle/swf_oracle_ex7.c,
ex7_detonation/main.f90:1399-1806) against what the mathematics fixes, on the crafted
states of tests/sweep7_cases.py:
- an adjusted pair keeps its sum (v_p_t = v_p + (v_i - v_i_t), weights equal), so
  every component's total over the grid changes by rounding only; the density sweep
  moves rho - 1d-12 to the next cell, so the density total is kept too;
- a sweep that ends without a stop leaves every cell with density and pressure
  (without the reaction term) above eps, the program's convergence test;
- a state with no flagged cell is left unchanged, with the statistics of one pass;
- nod_lim's interpolation parameter: a flagged cell next to a positive one lands on
  pressure eps (to rounding) when the pressure is linear along the segment."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from oracle import oracle as O  # noqa: E402
from sweep7_cases import CASES, GAMMA, Q, crafted7  # noqa: E402


def _interior(u, n):
    return u[3:n + 3, 3:n + 3, :]


def test_sweep_keeps_component_totals():
    for total, kind, seed in CASES:
        n = 5 * total
        u = crafted7(total, kind, seed)
        r = O.sweep7(u, total)
        a, b = _interior(u, n), _interior(r["u"], n)
        scale = np.abs(a).sum(axis=(0, 1)) + 1.0
        assert np.all(np.abs(b.sum(axis=(0, 1)) - a.sum(axis=(0, 1))) <= 1e-12 * scale), (kind, seed)
        # ghosts untouched
        assert np.array_equal(np.delete(np.delete(r["u"], range(3, n + 3), 0), range(3, n + 3), 1),
                              np.delete(np.delete(u, range(3, n + 3), 0), range(3, n + 3), 1))


def test_converged_sweep_leaves_positive_state():
    for total, kind, seed in CASES:
        n = 5 * total
        r = O.sweep7(crafted7(total, kind, seed), total)
        if r["code"] != 0:
            continue
        v = _interior(r["u"], n)
        rho, m1, m2, e = v[..., 0], v[..., 1], v[..., 2], v[..., 3]
        p = (GAMMA - 1.0) * (e - 0.5 * (m1 * m1 + m2 * m2) / rho)
        assert rho.min() > 1e-13 and p.min() > 1e-13, (kind, seed)


def test_unflagged_state_is_unchanged():
    u = crafted7(4, "runs", 11)
    n = 20
    for j in range(-2, n + 4):
        for i in range(-2, n + 4):
            r, m1, m2, ry = u[j + 2, i + 2, 0], u[j + 2, i + 2, 1], u[j + 2, i + 2, 2], u[j + 2, i + 2, 4]
            u[j + 2, i + 2, 0] = max(r, 0.5)
            u[j + 2, i + 2, 3] = 1.0 / (GAMMA - 1.0) + 0.5 * (m1 * m1 + m2 * m2) / u[j + 2, i + 2, 0] + Q * ry
    r = O.sweep7(u, 4)
    assert r["code"] == 0 and np.array_equal(r["u"], u)
    assert (r["sweep_count"], r["den_count"], r["pre_count"], r["q"], r["neg"], r["pos"]) == (1, 0, 1, 1, 0, 0)


def test_single_flagged_cell_lands_on_eps():
    """One cell of pressure -1e-3 with m = 0 and rho*Y = 0 in a uniform gas at rest:
    pressure is linear in E along the segment, so t = (p1 - eps)/(p1 - p2) gives
    pressure eps at the cell (to rounding), and its order-I neighbour takes the rest."""
    total, n = 4, 20
    u = np.zeros((n + 6, n + 6, 5))
    u[..., 0] = 1.0
    u[..., 3] = 1.0 / (GAMMA - 1.0)
    i, j = 7, 3  # odd row: order-I position (3-1)*20 + 7 = 47, forward neighbour (8, 3)
    u[j + 2, i + 2, 3] = -1e-3 / (GAMMA - 1.0)
    r = O.sweep7(u, total)
    assert r["code"] == 0 and (r["neg"], r["pos"]) == (0, 1)
    v = r["u"]
    p_cell = (GAMMA - 1.0) * v[j + 2, i + 2, 3]
    assert abs(p_cell - 1e-13) < 1e-15
    e_tot = v[3:n + 3, 3:n + 3, 3].sum()
    assert abs(e_tot - u[3:n + 3, 3:n + 3, 3].sum()) < 1e-12
    changed = np.argwhere(np.any(v != u, axis=2))
    assert sorted(map(tuple, changed.tolist())) == [(j + 2, i + 2), (j + 2, i + 3)]
