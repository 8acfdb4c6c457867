"""Initial data of the reference problems (driver side, host FP64).

Arrays are numpy (ny+6, nx+6, 4) views of the Fortran u(4,-2:nx+3,-2:ny+3).  The
expressions keep the reference's evaluation order; transcendental calls go to the C
library through `math` (as the reference's libm calls do), so the vortex data are
bit-identical to the reference's.
"""
from __future__ import annotations

import math

import numpy as np

from ._lib import state_shape


def init_sedov(cfg, e_corner_num: float = 0.244816, e_amb: float = 1e-12) -> np.ndarray:
    """ex4_sedov_2d/main.f90:1121-1189 at cut_x = cut_y = 1 (ghosts by bc afterwards).

    rho = 1, m = 0, E = 1d-12 everywhere; the cell containing the corner (0, 0) gets
    E = 0.244816/(dx*dy)."""
    u = np.zeros(state_shape(cfg))
    u[:, :, 0] = 1.0
    u[:, :, 3] = e_amb
    dx, dy = cfg.dx, cfg.dy
    for j in range(1, cfg.ny + 1):
        gy = cfg.ylo + dy * (j - 0.5)
        if not abs(gy - 0.0) <= dy / 2.0:
            continue
        for i in range(1, cfg.nx + 1):
            gx = cfg.xlo + dx * (i - 0.5)
            if abs(gx - 0.0) <= dx / 2.0:
                u[j + 2, i + 2, 3] = e_corner_num / (dx * dy)
    _sedov_ghosts(u, cfg)
    return u


def _sedov_ghosts(u: np.ndarray, cfg) -> None:
    """Reflective x = 0 and y = 0, zero-gradient at x = xm and y = ym (ex4 `bc`)."""
    nx, ny = cfg.nx, cfg.ny
    for i in range(1, 4):                      # x = 0 reflection, all rows
        u[:, 2 + 1 - i, :] = u[:, 2 + i, :]
        u[:, 2 + 1 - i, 1] = -u[:, 2 + i, 1]
    for j in range(1, 4):                      # y = 0 reflection, all columns
        u[2 + 1 - j, :, :] = u[2 + j, :, :]
        u[2 + 1 - j, :, 2] = -u[2 + j, :, 2]
    for ig in range(1, 4):
        u[:, 2 + nx + ig, :] = u[:, 2 + nx, :]
    for ig in range(1, 4):
        u[2 + ny + ig, :, :] = u[2 + ny, :, :]


def _vortex_point(gamma: float, mag: float, x: float, y: float, t: float):
    """ex3_vortex_accuracy_2d/main.f90:699-718 `exact`."""
    g1 = gamma - 1.0
    pi = 3.14159265358979323846264338327950
    r = math.sqrt(((x - 5.0) - t) * ((x - 5.0) - t) + ((y - 5.0) - t) * ((y - 5.0) - t))
    e1 = math.pow(1.0 - ((((g1 * (mag * mag)) / 8.0) / gamma) / (pi * pi)) * math.exp(1.0 - r * r), 1.0 / g1)
    e2 = e1 * (1.0 - (((mag / 2.0) / pi) * math.exp((1.0 - r * r) / 2.0)) * (y - 5.0))
    e3 = e1 * (1.0 + (((mag / 2.0) / pi) * math.exp((1.0 - r * r) / 2.0)) * (x - 5.0))
    e4 = (math.pow(e1, gamma) / g1 + (0.5 * (e2 * e2)) / e1) + (0.5 * (e3 * e3)) / e1
    return e1, e2, e3, e4


def vortex_exact(cfg, t: float, mag: float = 10.0828) -> np.ndarray:
    """Exact vortex state at time t at every cell centre, ghosts included."""
    u = np.zeros(state_shape(cfg))
    for j in range(-2, cfg.ny + 4):
        y = cfg.ylo + j * cfg.dy - cfg.dy / 2.0
        for i in range(-2, cfg.nx + 4):
            x = cfg.xlo + i * cfg.dx - cfg.dx / 2.0
            u[j + 2, i + 2, :] = _vortex_point(cfg.gamma, mag, x, y, t)
    return u


def init_vortex(cfg, mag: float = 10.0828) -> np.ndarray:
    """ex3 `initdata`: exact solution at t = 0 everywhere, then the periodic bc."""
    u = vortex_exact(cfg, 0.0, mag)
    nx, ny = cfg.nx, cfg.ny
    for i in range(1, 4):
        u[:, 2 + 1 - i, :] = u[:, 2 + nx + 1 - i, :]
        u[:, 2 + nx + i, :] = u[:, 2 + i, :]
    for j in range(1, 4):
        u[2 + 1 - j, :, :] = u[2 + ny + 1 - j, :, :]
        u[2 + ny + j, :, :] = u[2 + j, :, :]
    return u


def init_jet(cfg) -> np.ndarray:
    """ex5_mach2000_jet/main.f90:743-768: ambient rho = 0.5, p = 0.4127 (ghosts by bc)."""
    u = np.zeros(state_shape(cfg))
    g1 = cfg.gamma - 1.0
    u[:, :, 0] = cfg.amb_rho
    u[:, :, 3] = cfg.amb_p / g1
    return u


def init_step(cfg) -> np.ndarray:
    """initdata of the step program (ex6_shock_diffraction/main.f90:1111-1131) on one
    domain: gas at rest (gamma, 0, 0, 1/gamma1) and the inflow state (cfg.in_rho, in_u,
    in_p) where i <= total/2 and j >= 6*total, ghosts included.  The program applies no
    bc before its first step."""
    T = cfg.step_total
    g1 = cfg.gamma - 1.0
    u = np.zeros(state_shape(cfg))
    u[:, :, 0] = cfg.gamma
    u[:, :, 3] = 1.0 / g1
    dle, ule, ple = cfg.in_rho, cfg.in_u, cfg.in_p
    jj = np.arange(-2, cfg.ny + 4)[:, None]
    ii = np.arange(-2, cfg.nx + 4)[None, :]
    shocked = (ii <= T // 2) & (jj >= 6 * T)
    u[shocked, 0] = dle
    u[shocked, 1] = ule * dle
    u[shocked, 2] = 0.0
    u[shocked, 3] = ple / g1 + (0.5 * dle) * (ule * ule)
    return u


def _round_bits(x, p: int):
    """x (a positive Fraction) rounded to p significant bits, ties to even."""
    from fractions import Fraction
    e = x.numerator.bit_length() - x.denominator.bit_length()
    if Fraction(2) ** e > x:
        e -= 1
    scale = Fraction(2) ** (p - 1 - e)
    m = x * scale
    fl = m.numerator // m.denominator
    rem = m - fl
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and fl % 2 == 1):
        fl += 1
    return Fraction(fl) / scale


def init_sedov1d(cfg, shipped_energy: bool = False, e_amb: float = 1e-12) -> np.ndarray:
    """initdata of the 1D program (ex2_sedov_1d/main.f90:398-425): rho = 1, m = 0,
    E = e_amb everywhere (ghosts included), and E = 3200000.0/dx in cell n/2+1.

    shipped_energy: the program's own ambient E = 1d-9/gamma1 (:423), the binary128
    literal 1d-9 divided in binary128 by gamma1 = gamma - 1 (binary64), then rounded
    to binary64."""
    from fractions import Fraction
    n = cfg.n
    u = np.zeros((n + 6, 3))
    u[:, 0] = 1.0
    if shipped_energy:
        g1 = cfg.gamma - 1.0
        q = _round_bits(_round_bits(Fraction(1, 10 ** 9), 113) / Fraction(g1), 113)
        u[:, 2] = float(q)
    else:
        u[:, 2] = e_amb
    u[n // 2 + 1 + 2, 2] = 3200000.0 / cfg.dx
    return u
