"""Crafted states for the detonation program's sweep (ex7_detonation/main.f90:1399-1806),
shared by the emulated and GPU checks of swf7_sweep against the oracle's serial
restatement (oracle/swf_oracle_ex7.c).

A state is u of shape (n+6, n+6, 5) ([j][i][c], n = 5 total) with the program's
variables (rho, m1, m2, E, rho*Y) and gamma 1.2, q_const 50:
- `runs`: cells of slightly negative pressure in runs along the order-I snake
  (chains of flagged positions) and a few near-vacuum cells (rho < 1d-12: the
  density sweep);
- `deep`: pairs of strongly negative cells, so nod_lim takes its negative-neighbour
  branch and the order-I passes leave cells flagged (order-II passes follow);
- `corner`: the first cells of both snakes strongly negative, so an order-II forward
  pass finds position 1 flagged (the program would read sweepI_i(0));
- `loud`: momenta 1e4 times larger and runs of cells with pressure just below zero:
  the kinetic energy cancels in the pressure to about 1e-10, more than eps, so an
  adjustment can lower the rounded pressure sum and the order-I check stops the sweep.
"""
import numpy as np

GAMMA, Q = 1.2, 50.0


def _snake_I(n, p):
    q = p - 1
    j = q // n + 1
    r = q - (j - 1) * n
    return (r + 1 if j % 2 == 1 else n - r), j


def _set_pressure(u, i, j, p):
    r, m1, m2, ry = u[j + 2, i + 2, 0], u[j + 2, i + 2, 1], u[j + 2, i + 2, 2], u[j + 2, i + 2, 4]
    u[j + 2, i + 2, 3] = p / (GAMMA - 1.0) + 0.5 * (m1 * m1 + m2 * m2) / r + Q * ry


def crafted7(total: int, kind: str, seed: int) -> np.ndarray:
    n = 5 * total
    rng = np.random.default_rng(seed)
    u = np.zeros((n + 6, n + 6, 5))
    rho = 1.0 + 0.5 * rng.random((n + 6, n + 6))
    mag = 1e3 if kind == "loud" else 0.1
    u[:, :, 0] = rho
    u[:, :, 1] = mag * rng.standard_normal((n + 6, n + 6)) * rho
    u[:, :, 2] = mag * rng.standard_normal((n + 6, n + 6)) * rho
    u[:, :, 4] = rho * rng.random((n + 6, n + 6))
    for j in range(-2, n + 4):
        for i in range(-2, n + 4):
            _set_pressure(u, i, j, 0.5 + rng.random())
    N = n * n
    if kind == "runs":
        for _ in range(max(2, N // 40)):
            p0 = int(rng.integers(1, N + 1))
            for p in range(p0, min(N, p0 + int(rng.integers(1, 6))) + 1):
                i, j = _snake_I(n, p)
                _set_pressure(u, i, j, -1e-3 * rng.random())
        for _ in range(3):
            i, j = (int(v) for v in rng.integers(1, n + 1, 2))
            u[j + 2, i + 2, :] = (1e-14, 0.0, 0.0, 0.0, 0.0)
            _set_pressure(u, i, j, 1e-3)
    elif kind == "deep":
        for _ in range(max(2, N // 60)):
            p0 = int(rng.integers(1, N))
            for p in (p0, p0 + 1):
                i, j = _snake_I(n, p)
                _set_pressure(u, i, j, -5.0 * (0.5 + rng.random()))
    elif kind == "corner":
        for p in (1, 2, 3):
            i, j = _snake_I(n, p)
            _set_pressure(u, i, j, -20.0)
        for (i, j) in ((1, 2), (1, 3)):
            _set_pressure(u, i, j, -20.0)
    elif kind == "loud":
        for _ in range(N // 10):
            p0 = int(rng.integers(1, N + 1))
            for p in range(p0, min(N, p0 + 3) + 1):
                i, j = _snake_I(n, p)
                _set_pressure(u, i, j, -1e-9 * rng.random())
    else:
        raise ValueError(kind)
    return u


CASES = [(4, "runs", 1), (4, "runs", 2), (8, "runs", 3), (4, "deep", 4), (8, "deep", 5), (16, "deep", 6),
         (4, "corner", 7), (4, "loud", 8), (8, "loud", 9)]
