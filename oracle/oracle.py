"""ctypes binding of the CPU oracle (liboracle.so, built from oracle/swf_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` as the checker.  The product package never
imports this module.

The oracle restates ex4_sedov_2d/main.f90 (Sedov), ex3_vortex_accuracy_2d/main.f90
(vortex), ex5_mach2000_jet/main.f90 (jet) and ex2_sedov_1d/main.f90 (the 1D program)
of the reference; see swf_oracle.h.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

BC_SEDOV, BC_PERIODIC, BC_JET_HALF, BC_JET_FULL, BC_STEP = 0, 1, 2, 3, 4


class SwoConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32),
        ("xlo", C.c_double), ("ylo", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
        ("gamma", C.c_double), ("cfl", C.c_double),
        ("eps", C.c_double), ("eps2", C.c_double), ("delta", C.c_double),
        ("bc", C.c_int32), ("use_limiter", C.c_int32), ("swp_num", C.c_int32),
        ("alpha_in_scheme", C.c_int32), ("roe_eps_double", C.c_int32), ("ydens_plain", C.c_int32),
        ("eps2_quad_literal", C.c_int32), ("findalpha_ex3", C.c_int32), ("dt_rule_dx53", C.c_int32),
        ("poscheck_le", C.c_int32), ("success_ge", C.c_int32), ("pbar_check", C.c_int32),
        ("fail_stop", C.c_int32),
        ("jet_rho", C.c_double), ("jet_u", C.c_double), ("amb_rho", C.c_double), ("amb_p", C.c_double),
        ("jet_halfwidth", C.c_double),
        ("step_total", C.c_int32), ("alpha_lines", C.c_int32), ("weno_eps12", C.c_int32), ("psum_check", C.c_int32),
        ("in_rho", C.c_double), ("in_u", C.c_double), ("in_p", C.c_double),
    ]


class SwoStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "sweep_count", "den_count", "pre_count", "total_sweep", "max_sweep", "total_neg", "max_neg",
        "total_pos", "max_pos", "stage_sweeps")]


class Swo1dConfig(C.Structure):
    _fields_ = [("n", C.c_int32), ("xlo", C.c_double), ("dx", C.c_double), ("gamma", C.c_double),
                ("cfl", C.c_double), ("eps", C.c_double), ("use_limiter", C.c_int32), ("swp_num", C.c_int32)]


class Swo1dStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "sweep_count", "den_count", "pre_count", "total_sweep", "max_sweep", "total_neg", "max_neg",
        "total_pos", "max_pos", "steps")]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        dp = P(C.c_double)
        L.swo_create.restype = C.c_void_p
        L.swo_create.argtypes = [P(SwoConfig)]
        L.swo_destroy.argtypes = [C.c_void_p]
        L.swo_set_state.argtypes = [C.c_void_p, dp]
        L.swo_get_state.argtypes = [C.c_void_p, dp]
        L.swo_step.argtypes = [C.c_void_p, C.c_double, C.c_double, dp]
        L.swo_step.restype = C.c_int
        L.swo_bc.argtypes = [C.c_void_p]
        L.swo_alpha_max.argtypes = [C.c_void_p]
        L.swo_alpha_max.restype = C.c_double
        L.swo_rhs.argtypes = [C.c_void_p, C.c_double, dp]
        L.swo_check_positivity.argtypes = [C.c_void_p]
        L.swo_check_positivity.restype = C.c_int
        L.swo_sweep.argtypes = [C.c_void_p]
        L.swo_sweep.restype = C.c_int
        L.swo_get_stats.argtypes = [C.c_void_p, P(SwoStats)]
        L.swo_track.argtypes = [C.c_void_p, P(C.c_int64), P(C.c_int64), P(C.c_int64), C.c_int64]
        L.swo_track.restype = C.c_int64
        L.swo_time_continue.argtypes = [C.c_double, C.c_double]
        L.swo_time_continue.restype = C.c_int
        L.swo_weno_right_n.argtypes = [dp, dp, C.c_int64]
        L.swo_weno_left_n.argtypes = [dp, dp, C.c_int64]
        L.swo_nod_lim.argtypes = [dp, dp, dp, C.c_double, C.c_double, dp]
        L.swo_nod_lim.restype = C.c_int
        L.swo_init_sedov.argtypes = [P(SwoConfig), C.c_double, C.c_double, dp]
        L.swo_init_vortex.argtypes = [P(SwoConfig), C.c_double, C.c_double, dp]
        L.swo_vortex_exact.argtypes = [P(SwoConfig), C.c_double, C.c_double, dp]
        L.swo_init_jet.argtypes = [P(SwoConfig), dp]
        L.swo_init_step.argtypes = [P(SwoConfig), C.c_double, dp]
        L.swo_weno_eps12_n.argtypes = [dp, dp, dp, C.c_int64]
        L.swo1d_create.restype = C.c_void_p
        L.swo1d_create.argtypes = [P(Swo1dConfig)]
        L.swo1d_destroy.argtypes = [C.c_void_p]
        L.swo1d_set_state.argtypes = [C.c_void_p, dp]
        L.swo1d_get_state.argtypes = [C.c_void_p, dp]
        L.swo1d_get_stats.argtypes = [C.c_void_p, P(Swo1dStats)]
        L.swo1d_track.restype = C.c_int64
        L.swo1d_track.argtypes = [C.c_void_p, P(C.c_int64), P(C.c_int64), P(C.c_int64), C.c_int64]
        L.swo1d_init_sedov.argtypes = [P(Swo1dConfig), C.c_int, C.c_double, dp]
        L.swo1d_run.argtypes = [C.c_void_p, P(C.c_double), C.c_double, C.c_int64, P(C.c_int64)]
        L.swo7_sweep.argtypes = [dp, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_int32, P(C.c_int64),
                                 P(C.c_int64)]
        L.swo7_sweep.restype = C.c_int
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_config(**kw) -> SwoConfig:
    c = SwoConfig()
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def state_shape(cfg) -> tuple:
    """numpy shape (ny+6, nx+6, 4) of the Fortran u(4,-2:nx+3,-2:ny+3) array."""
    return (cfg.ny + 6, cfg.nx + 6, 4)


def weno_right(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 5)
    out = np.empty(a.shape[0])
    lib().swo_weno_right_n(_dp(a), _dp(out), a.shape[0])
    return out


def weno_left(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 5)
    out = np.empty(a.shape[0])
    lib().swo_weno_left_n(_dp(a), _dp(out), a.shape[0])
    return out


def nod_lim(v1, v2, vbar, gamma1, eps):
    v1, v2, vb = (np.ascontiguousarray(x, dtype=np.float64) for x in (v1, v2, vbar))
    out = np.empty(4)
    pos = lib().swo_nod_lim(_dp(v1), _dp(v2), _dp(vb), gamma1, eps, _dp(out))
    return out, bool(pos)


def time_continue(t: float, tfinal: float) -> bool:
    return bool(lib().swo_time_continue(t, tfinal))


class Oracle:
    """One solver context of the oracle (reference `rk` and its callees)."""

    def __init__(self, cfg: SwoConfig):
        self.cfg = cfg
        self._p = lib().swo_create(C.byref(cfg))

    def __del__(self):
        if getattr(self, "_p", None):
            lib().swo_destroy(self._p)
            self._p = None

    def set_state(self, u: np.ndarray) -> None:
        u = np.ascontiguousarray(u, dtype=np.float64)
        assert u.shape == state_shape(self.cfg)
        lib().swo_set_state(self._p, _dp(u))

    def state(self) -> np.ndarray:
        u = np.empty(state_shape(self.cfg))
        lib().swo_get_state(self._p, _dp(u))
        return u

    def bc(self) -> None:
        lib().swo_bc(self._p)

    def alpha_max(self) -> float:
        return lib().swo_alpha_max(self._p)

    def rhs(self, alphamax: float) -> np.ndarray:
        r = np.empty(state_shape(self.cfg))
        lib().swo_rhs(self._p, alphamax, _dp(r))
        return r

    def check_positivity(self) -> int:
        return lib().swo_check_positivity(self._p)

    def sweep(self) -> int:
        return lib().swo_sweep(self._p)

    def step(self, t: float, tfinal: float):
        dt = C.c_double(0.0)
        rc = lib().swo_step(self._p, t, tfinal, C.byref(dt))
        return rc, dt.value

    def stats(self) -> dict:
        s = SwoStats()
        lib().swo_get_stats(self._p, C.byref(s))
        return {n: getattr(s, n) for n, _ in SwoStats._fields_}

    def track(self):
        n = lib().swo_track(self._p, None, None, None, 0)
        q = np.zeros(max(n, 1), np.int64)
        ng = np.zeros_like(q)
        ps = np.zeros_like(q)
        P = C.POINTER(C.c_int64)
        lib().swo_track(self._p, q.ctypes.data_as(P), ng.ctypes.data_as(P), ps.ctypes.data_as(P), n)
        return q[:n], ng[:n], ps[:n]

    def run(self, t: float, tfinal: float, max_steps: int | None = None):
        """Reference time loop `DO WHILE (time <= time_terminal - 1d-10)`."""
        steps = 0
        while time_continue(t, tfinal):
            if max_steps is not None and steps >= max_steps:
                break
            rc, dt = self.step(t, tfinal)
            if rc:
                return rc, t, steps
            t = t + dt
            steps += 1
        return 0, t, steps


# ---------------------------------------------------------------------------
# configurations of the reference programs (parameter surfaces)
# ---------------------------------------------------------------------------

def sedov_config(n: int = 640, cfl: float = 0.5, swp_num: int = 20000, xm: float = 1.3) -> SwoConfig:
    """ex4_sedov_2d/main.f90:15-20,50 at cut_x = cut_y = 1 (dx = xm/nx as :1154-1157)."""
    return make_config(nx=n, ny=n, xlo=0.0, ylo=0.0, dx=(xm - 0.0) / n, dy=(xm - 0.0) / n, gamma=1.4,
                       cfl=cfl, eps=1e-13, eps2=1e-8, delta=1e-12, bc=BC_SEDOV, use_limiter=2,
                       swp_num=swp_num, pbar_check=1, fail_stop=1)


def vortex_config(n: int = 180, cfl: float = 0.5, swp_num: int = 20000, use_limiter: int = 2) -> SwoConfig:
    """ex3_vortex_accuracy_2d/main.f90:5-7,570-572,671-697."""
    return make_config(nx=n, ny=n, xlo=0.0, ylo=0.0, dx=(10.0 - 0.0) / n, dy=(10.0 - 0.0) / n, gamma=1.4,
                       cfl=cfl, eps=1e-13, eps2=1e-8, delta=1e-12, bc=BC_PERIODIC, use_limiter=use_limiter,
                       swp_num=swp_num, alpha_in_scheme=1, roe_eps_double=1, ydens_plain=1,
                       eps2_quad_literal=1, findalpha_ex3=1, dt_rule_dx53=1, poscheck_le=1)


def jet_config(nx: int = 800, ny: int = 200, cfl: float = 0.25, half: bool = True,
               swp_num: int = 20000) -> SwoConfig:
    """ex5_mach2000_jet/main.f90:4-6,597,686-768 (half domain), or the full [-0.25,0.25] domain."""
    ylo = 0.0 if half else -0.25
    yhi = 0.25
    return make_config(nx=nx, ny=ny, xlo=0.0, ylo=ylo, dx=(1.0 - 0.0) / nx, dy=(yhi - ylo) / ny,
                       gamma=5.0 / 3.0, cfl=cfl, eps=1e-13, eps2=1e-8, delta=1e-12,
                       bc=BC_JET_HALF if half else BC_JET_FULL, use_limiter=2, swp_num=swp_num,
                       alpha_in_scheme=1, roe_eps_double=1, ydens_plain=1, eps2_quad_literal=1,
                       success_ge=1, jet_rho=5.0, jet_u=800.0, amb_rho=0.5, amb_p=0.4127,
                       jet_halfwidth=0.05)


def step_config(total: int = 80, cfl: float = 0.9, swp_num: int = 20000) -> SwoConfig:
    """ex6_shock_diffraction/main.f90:16-20,1089-1101,695 on one domain (cut_x = cut_y = 1):
    [0, 13] x [0, 11] with 13*total x 11*total cells, a step of total x 6*total cells."""
    nx, ny = 13 * total, 11 * total
    return make_config(nx=nx, ny=ny, xlo=0.0, ylo=0.0, dx=13.0 / nx, dy=11.0 / ny, gamma=1.4, cfl=cfl, eps=1e-13,
                       eps2=1e-8, delta=1e-12, bc=BC_STEP, use_limiter=2, swp_num=swp_num, poscheck_le=1,
                       pbar_check=1, fail_stop=1, step_total=total, alpha_lines=1, weno_eps12=1, psum_check=1)


def init_step(cfg: SwoConfig, mach: float = 5.09) -> np.ndarray:
    """ex6 initdata; also sets cfg.in_rho/in_u/in_p (the inflow state of the step bc)."""
    u = np.zeros(state_shape(cfg))
    lib().swo_init_step(C.byref(cfg), mach, _dp(u))
    return u


def weno_eps12(a: np.ndarray):
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = a.shape[0]
    r, l = np.empty(n), np.empty(n)
    lib().swo_weno_eps12_n(_dp(a), _dp(r), _dp(l), n)
    return r, l


def init_sedov(cfg: SwoConfig, e_corner_num: float = 0.244816, e_amb: float = 1e-12) -> np.ndarray:
    u = np.zeros(state_shape(cfg))
    lib().swo_init_sedov(C.byref(cfg), e_corner_num, e_amb, _dp(u))
    return u


def init_vortex(cfg: SwoConfig, mag: float = 10.0828) -> np.ndarray:
    u = np.zeros(state_shape(cfg))
    lib().swo_init_vortex(C.byref(cfg), mag, 0.0, _dp(u))
    return u


def vortex_exact(cfg: SwoConfig, t: float, mag: float = 10.0828) -> np.ndarray:
    u = np.zeros(state_shape(cfg))
    lib().swo_vortex_exact(C.byref(cfg), mag, t, _dp(u))
    return u


def init_jet(cfg: SwoConfig) -> np.ndarray:
    u = np.zeros(state_shape(cfg))
    lib().swo_init_jet(C.byref(cfg), _dp(u))
    return u


# ---------------------------------------------------------------------------
# the 1D program (ex2_sedov_1d/main.f90)
# ---------------------------------------------------------------------------

def sedov1d_config(n: int = 800, cfl: float = 0.6, swp_num: int = 2, use_limiter: int = 2) -> Swo1dConfig:
    """ex2_sedov_1d/main.f90:4-7,399-405: [-2,2], dx = 4/n, gamma 1.4, eps 1d-13.

    The shipped program has cfl 1.2 and swp_num 50; the 1D benchmark case sets 0.6 and 2.
    """
    return Swo1dConfig(n=n, xlo=-2.0, dx=(2.0 - (-2.0)) / n, gamma=1.4, cfl=cfl, eps=1e-13,
                       use_limiter=use_limiter, swp_num=swp_num)


def state1d_shape(cfg: Swo1dConfig) -> tuple:
    return (cfg.n + 6, 3)


def init_sedov1d(cfg: Swo1dConfig, shipped_energy: bool = False, e_amb: float = 1e-12) -> np.ndarray:
    """initdata (:398-420); the shipped program's ambient E is 1d-9/gamma1 (shipped_energy)."""
    u = np.zeros(state1d_shape(cfg))
    lib().swo1d_init_sedov(C.byref(cfg), int(shipped_energy), e_amb, _dp(u))
    return u


class Oracle1D:
    """One context of the 1D program: state u(3, -2:n+3) stored as rows (n+6, 3)."""

    def __init__(self, cfg: Swo1dConfig):
        self.cfg = cfg
        self._p = lib().swo1d_create(C.byref(cfg))

    def __del__(self):
        if getattr(self, "_p", None):
            lib().swo1d_destroy(self._p)
            self._p = None

    def set_state(self, u: np.ndarray) -> None:
        u = np.ascontiguousarray(u, dtype=np.float64)
        assert u.shape == state1d_shape(self.cfg)
        lib().swo1d_set_state(self._p, _dp(u))

    def state(self) -> np.ndarray:
        u = np.empty(state1d_shape(self.cfg))
        lib().swo1d_get_state(self._p, _dp(u))
        return u

    def stats(self) -> dict:
        s = Swo1dStats()
        lib().swo1d_get_stats(self._p, C.byref(s))
        return {n: getattr(s, n) for n, _ in Swo1dStats._fields_}

    def track(self):
        n = lib().swo1d_track(self._p, None, None, None, 0)
        q = np.zeros(max(n, 1), np.int64)
        ng = np.zeros_like(q)
        ps = np.zeros_like(q)
        P = C.POINTER(C.c_int64)
        lib().swo1d_track(self._p, q.ctypes.data_as(P), ng.ctypes.data_as(P), ps.ctypes.data_as(P), n)
        return q[:n], ng[:n], ps[:n]

    def run(self, t: float, tfinal: float, max_steps: int = 1 << 40):
        """Reference loop `DO WHILE (time <= time_terminal - 1d-10)` (:38-45)."""
        tt = C.c_double(t)
        k = C.c_int64(0)
        lib().swo1d_run(self._p, C.byref(tt), tfinal, max_steps, C.byref(k))
        return tt.value, k.value


def pressure1d(u: np.ndarray, gamma: float = 1.4) -> np.ndarray:
    """The pre.dat column: gamma1*(E - 0.5*m**2/rho) (:481)."""
    g1 = gamma - 1.0
    return g1 * (u[:, 2] - (0.5 * (u[:, 1] * u[:, 1])) / u[:, 0])


def sweep7(u: np.ndarray, total: int, gamma: float = 1.2, q_const: float = 50.0, eps: float = 1e-13,
           swp_num: int = 20000) -> dict:
    """The detonation program's sweep_lim on a copy of u (shape (n+6, n+6, 5), [j][i][c],
    n = 5 total): the swept state, the stop code and the counts."""
    a = np.ascontiguousarray(u, dtype=np.float64).copy()
    cnt = (C.c_int64 * 3)()
    trk = (C.c_int64 * 3)()
    rc = lib().swo7_sweep(_dp(a), total, gamma, q_const, eps, swp_num, cnt, trk)
    return {"u": a, "code": rc, "sweep_count": cnt[0], "den_count": cnt[1], "pre_count": cnt[2],
            "q": trk[0], "neg": trk[1], "pos": trk[2]}
