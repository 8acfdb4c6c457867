"""ctypes binding of libswf.so (include/swf.h) and the solver object.

The parameter constructors restate the PARAMETER blocks of the reference programs:
  sedov_config  ex4_sedov_2d/main.f90:15-20,50,687 at cut_x = cut_y = 1
  vortex_config ex3_vortex_accuracy_2d/main.f90:5-7,570-572,671-697
  jet_config    ex5_mach2000_jet/main.f90:4-6,597,686-768
  step_config   ex6_shock_diffraction/main.f90:16-20,695,1089-1109 on one domain
  sedov1d_config ex2_sedov_1d/main.f90:4-7,398-405 (the 1D program)
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(_HERE, "libswf.so")

PRECISIONS = {"fp64_exact": 0, "fp64": 1, "fp32_bulk": 2, "fp32_full": 3}
BC_SEDOV, BC_PERIODIC, BC_JET_HALF, BC_JET_FULL, BC_STEP = 0, 1, 2, 3, 4

SWF_OK, SWF_E_PBAR, SWF_E_SWEEP_FAIL, SWF_E_PSUM = 0, 1, 2, 3


class SwfConfig(C.Structure):
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
        ("precision", C.c_int32), ("device", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class SwfStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "sweep_count", "den_count", "pre_count", "total_sweep", "max_sweep", "total_neg", "max_neg",
        "total_pos", "max_pos", "stage_sweeps", "steps")]


class Swf1dConfig(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("dx", C.c_double), ("gamma", C.c_double), ("cfl", C.c_double), ("eps", C.c_double),
        ("use_limiter", C.c_int32), ("swp_num", C.c_int32), ("precision", C.c_int32), ("device", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


class SwfError(RuntimeError):
    """A failed library call; `code` is the swf.h return code."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


_lib = None


def lib():
    """The loaded libswf.so.  Raises if the HIP library has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(library_path):
            raise ImportError(f"{library_path} is missing: build it with `python -m sweeping_fd_weno_amd.build`")
        L = C.CDLL(library_path)
        P = C.POINTER
        dp = P(C.c_double)
        vp = C.c_void_p
        L.swf_create.argtypes = [P(SwfConfig), P(vp)]
        L.swf_destroy.argtypes = [vp]
        L.swf_destroy.restype = None
        L.swf_upload.argtypes = [vp, dp]
        L.swf_download.argtypes = [vp, dp]
        L.swf_step.argtypes = [vp, C.c_double, C.c_double, dp]
        L.swf_advance.argtypes = [vp, dp, C.c_double, C.c_int64, P(C.c_int64)]
        L.swf_bc.argtypes = [vp]
        L.swf_alpha_max.argtypes = [vp, dp]
        L.swf_rhs.argtypes = [vp, C.c_double, dp]
        L.swf_check_positivity.argtypes = [vp, P(C.c_int32)]
        L.swf_sweep.argtypes = [vp]
        L.swf_stats.argtypes = [vp, P(SwfStats)]
        L.swf_track.argtypes = [vp, P(C.c_int64), P(C.c_int64), P(C.c_int64), C.c_int64]
        L.swf_track.restype = C.c_int64
        L.swf_weno_points.argtypes = [dp, dp, dp, C.c_int64, C.c_int32, C.c_int32]
        L.swf_weno_points_noalloc.argtypes = [dp, dp, dp, C.c_int64, C.c_int32, C.c_int32]
        L.swf_time_continue.argtypes = [C.c_double, C.c_double]
        L.swf_sync.argtypes = [vp]
        L.swf_time_kernel.argtypes = [vp, C.c_int32, C.c_int32, dp]
        L.swf_batch_create.argtypes = [P(SwfConfig), C.c_int32, P(vp)]
        L.swf_batch_destroy.argtypes = [vp]
        L.swf_batch_destroy.restype = None
        L.swf_batch_instance.argtypes = [vp, C.c_int32]
        L.swf_batch_instance.restype = vp
        L.swf_batch_advance.argtypes = [vp, dp, dp, C.c_int64, P(C.c_int64), P(C.c_int32)]
        L.swf_batch_set_schedule.argtypes = [vp, C.c_int32]
        L.swf_batch_tune.argtypes = [vp, C.c_int32, P(C.c_int32), P(C.c_float)]
        L.swf1d_create.argtypes = [P(Swf1dConfig), P(vp)]
        L.swf1d_destroy.argtypes = [vp]
        L.swf1d_destroy.restype = None
        L.swf1d_upload.argtypes = [vp, dp]
        L.swf1d_download.argtypes = [vp, dp]
        L.swf1d_advance.argtypes = [vp, dp, C.c_double, C.c_int64, P(C.c_int64)]
        L.swf1d_stats.argtypes = [vp, P(SwfStats)]
        L.swf1d_track.argtypes = [vp, P(C.c_int64), P(C.c_int64), P(C.c_int64), C.c_int64]
        L.swf1d_track.restype = C.c_int64
        L.swf1d_last_ms.argtypes = [vp, dp]
        L.swf_strerror.argtypes = [C.c_int]
        L.swf_strerror.restype = C.c_char_p
        L.swf_last_error.argtypes = []
        L.swf_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _check(rc: int, what: str) -> int:
    if rc != SWF_OK:
        L = lib()
        msg = f"{what}: {L.swf_strerror(rc).decode()}"
        detail = L.swf_last_error().decode()
        if detail:
            msg += f" ({detail})"
        raise SwfError(rc, msg)
    return rc


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _prec(p) -> int:
    return PRECISIONS[p] if isinstance(p, str) else int(p)


def state_shape(cfg: SwfConfig) -> tuple:
    """numpy shape (ny+6, nx+6, 4) of the Fortran array u(4,-2:nx+3,-2:ny+3)."""
    return (cfg.ny + 6, cfg.nx + 6, 4)


def _config(**kw) -> SwfConfig:
    c = SwfConfig()
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def sedov_config(n: int = 640, precision="fp64_exact", cfl: float = 0.5, swp_num: int = 20000,
                 xm: float = 1.3, device: int = 0, use_limiter: int = 2) -> SwfConfig:
    """2D Sedov blast on [0, xm]^2 (ex4_sedov_2d, cut_x = cut_y = 1: dx = xm/nx).
    use_limiter: 2 = sweeping, 1 = Zhang-Shu flux limiter, 0 = none."""
    return _config(nx=n, ny=n, xlo=0.0, ylo=0.0, dx=(xm - 0.0) / n, dy=(xm - 0.0) / n, gamma=1.4, cfl=cfl,
                   eps=1e-13, eps2=1e-8, delta=1e-12, bc=BC_SEDOV, use_limiter=use_limiter, swp_num=swp_num,
                   pbar_check=1, fail_stop=1, precision=_prec(precision), device=device)


def vortex_config(n: int = 180, precision="fp64_exact", cfl: float = 0.5, swp_num: int = 20000,
                  device: int = 0, use_limiter: int = 2) -> SwfConfig:
    """Low-density isentropic vortex on the periodic [0, 10]^2 (ex3_vortex_accuracy_2d)."""
    return _config(nx=n, ny=n, xlo=0.0, ylo=0.0, dx=(10.0 - 0.0) / n, dy=(10.0 - 0.0) / n, gamma=1.4,
                   cfl=cfl, eps=1e-13, eps2=1e-8, delta=1e-12, bc=BC_PERIODIC, use_limiter=use_limiter,
                   swp_num=swp_num, alpha_in_scheme=1, roe_eps_double=1, ydens_plain=1, eps2_quad_literal=1,
                   findalpha_ex3=1, dt_rule_dx53=1, poscheck_le=1, precision=_prec(precision), device=device)


def jet_config(nx: int = 800, ny: int = 200, precision="fp64_exact", cfl: float = 0.25, half: bool = True,
               swp_num: int = 20000, device: int = 0) -> SwfConfig:
    """Mach 2000 jet (ex5_mach2000_jet): half domain [0,1]x[0,0.25] with reflective y = 0,
    or the full domain [0,1]x[-0.25,0.25]."""
    ylo = 0.0 if half else -0.25
    return _config(nx=nx, ny=ny, xlo=0.0, ylo=ylo, dx=(1.0 - 0.0) / nx, dy=(0.25 - ylo) / ny,
                   gamma=5.0 / 3.0, cfl=cfl, eps=1e-13, eps2=1e-8, delta=1e-12,
                   bc=BC_JET_HALF if half else BC_JET_FULL, use_limiter=2, swp_num=swp_num,
                   alpha_in_scheme=1, roe_eps_double=1, ydens_plain=1, eps2_quad_literal=1, success_ge=1,
                   jet_rho=5.0, jet_u=800.0, amb_rho=0.5, amb_p=0.4127, jet_halfwidth=0.05,
                   precision=_prec(precision), device=device)


def step_config(total: int = 80, precision="fp64_exact", mach: float = 5.09, cfl: float = 0.9,
                swp_num: int = 20000) -> SwfConfig:
    """ex6_shock_diffraction/main.f90 on one domain (cut_x = cut_y = 1): [0, 13] x [0, 11] with
    13*total x 11*total cells, a step of total x 6*total cells at the lower left, a Mach
    `mach` shock (5.09 in the program) entering from the left; per-line alphas, 1d-12 in the
    WENO weights, the sweep's pressure-sum check, cfl 0.9.  The inflow state is the
    Rankine-Hugoniot state of `step_inflow`."""
    nx, ny = 13 * total, 11 * total
    dle, ule, ple = step_inflow(mach, 1.4)
    return _config(nx=nx, ny=ny, xlo=0.0, ylo=0.0, dx=13.0 / nx, dy=11.0 / ny, gamma=1.4, cfl=cfl, eps=1e-13,
                   eps2=1e-8, delta=1e-12, bc=BC_STEP, use_limiter=2, swp_num=swp_num, poscheck_le=1, pbar_check=1,
                   fail_stop=1, step_total=total, alpha_lines=1, weno_eps12=1, psum_check=1, in_rho=dle, in_u=ule,
                   in_p=ple, precision=_prec(precision), device=0)


def step_inflow(mach: float, gamma: float = 1.4):
    """The post-shock state (dle, ule, ple) of ex6_shock_diffraction/main.f90:1098-1109:
    a shock of Mach `mach` into gas with rho = gamma, u = 0, p = 1, each expression in the
    program's order (binary64)."""
    ss, dri, pri = mach, gamma, 1.0
    ach2 = (((0.0 - ss) * (0.0 - ss)) / gamma / pri) * dri
    ple = pri * (((gamma * ach2) / (gamma + 1.0)) * 2.0 - (gamma - 1.0) / (gamma + 1.0))
    dle = (((dri * (gamma + 1.0)) / 2.0) * ach2) / (((gamma - 1.0) / 2.0) * ach2 + 1.0)
    ule = ((0.0 - ss) / dle) * dri + ss
    return dle, ule, ple


def time_continue(t: float, tfinal: float) -> bool:
    """The reference loop test `time <= time_terminal - 1d-10` (binary128)."""
    return bool(lib().swf_time_continue(t, tfinal))


def weno_points(a: np.ndarray, precision="fp64_exact", device: int = 0, eps12: bool = False):
    """weno_right / weno_left of n stencils a(-2..2) computed on the device; eps12: the
    weights of the step program's weno5.f90 (1d-12 in place of 1d-6)."""
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 5)
    r = np.empty(a.shape[0])
    l = np.empty(a.shape[0])
    flags = _prec(precision) | (0x10 if eps12 else 0)
    _check(lib().swf_weno_points(_dp(a), _dp(r), _dp(l), a.shape[0], flags, device), "swf_weno_points")
    return r, l


WENO_NOALLOC_MAX = 65536  # SWF_WENO_NOALLOC_MAX


def weno_points_noalloc(a: np.ndarray, precision="fp64_exact", device: int = 0, eps12: bool = False):
    """weno_points through the entry point that allocates no device memory (at most
    WENO_NOALLOC_MAX stencils per call)."""
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 5)
    r = np.empty(a.shape[0])
    l = np.empty(a.shape[0])
    flags = _prec(precision) | (0x10 if eps12 else 0)
    _check(lib().swf_weno_points_noalloc(_dp(a), _dp(r), _dp(l), a.shape[0], flags, device),
           "swf_weno_points_noalloc")
    return r, l


class Solver:
    """One device solver context (the reference's global state and `rk` call tree)."""

    def __init__(self, cfg: SwfConfig, _borrowed=None):
        self.cfg = cfg
        self._owned = _borrowed is None
        if _borrowed is not None:  # an instance of a Batch: the batch owns it
            self._h = C.c_void_p(_borrowed)
            return
        h = C.c_void_p()
        _check(lib().swf_create(C.byref(cfg), C.byref(h)), "swf_create")
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            if self._owned:
                lib().swf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # state transfer --------------------------------------------------------
    def upload(self, u: np.ndarray) -> None:
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.shape != state_shape(self.cfg):
            raise ValueError(f"state shape {u.shape} != {state_shape(self.cfg)}")
        _check(lib().swf_upload(self._h, _dp(u)), "swf_upload")

    def download(self) -> np.ndarray:
        u = np.empty(state_shape(self.cfg))
        _check(lib().swf_download(self._h, _dp(u)), "swf_download")
        return u

    # stepping ----------------------------------------------------------------
    def step(self, t: float, tfinal: float) -> float:
        dt = C.c_double(0.0)
        _check(lib().swf_step(self._h, t, tfinal, C.byref(dt)), "swf_step")
        return dt.value

    def advance(self, t: float, tfinal: float, max_steps: int = 1 << 40):
        """Reference time loop: steps while time <= tfinal - 1d-10.  Returns (time, steps)."""
        tt = C.c_double(t)
        n = C.c_int64(0)
        _check(lib().swf_advance(self._h, C.byref(tt), tfinal, max_steps, C.byref(n)), "swf_advance")
        return tt.value, n.value

    # fine-grained operators ----------------------------------------------
    def bc(self) -> None:
        _check(lib().swf_bc(self._h), "swf_bc")

    def alpha_max(self) -> float:
        a = C.c_double(0.0)
        _check(lib().swf_alpha_max(self._h, C.byref(a)), "swf_alpha_max")
        return a.value

    def rhs(self, alphamax: float) -> np.ndarray:
        r = np.empty(state_shape(self.cfg))
        _check(lib().swf_rhs(self._h, alphamax, _dp(r)), "swf_rhs")
        return r

    def check_positivity(self) -> int:
        p = C.c_int32(0)
        _check(lib().swf_check_positivity(self._h, C.byref(p)), "swf_check_positivity")
        return p.value

    def sweep(self) -> None:
        _check(lib().swf_sweep(self._h), "swf_sweep")

    def time_kernel(self, which: str = "xflux", nreps: int = 20) -> float:
        """Average ms of one launch of a stage kernel (HIP events on the solver's stream):
        "xflux" (x-pass of wenofd) or "ystage" (y-pass fused with the RK update)."""
        ms = C.c_double(0.0)
        _check(lib().swf_time_kernel(self._h, {"xflux": 0, "ystage": 1}[which], nreps, C.byref(ms)),
               "swf_time_kernel")
        return ms.value

    def sync(self) -> None:
        _check(lib().swf_sync(self._h), "swf_sync")

    # statistics --------------------------------------------------------------
    def stats(self) -> dict:
        s = SwfStats()
        _check(lib().swf_stats(self._h, C.byref(s)), "swf_stats")
        return {n: getattr(s, n) for n, _ in SwfStats._fields_}

    def track(self):
        n = lib().swf_track(self._h, None, None, None, 0)
        if n < 0:
            _check(int(n), "swf_track")
        q = np.zeros(max(n, 1), np.int64)
        ng = np.zeros_like(q)
        ps = np.zeros_like(q)
        P = C.POINTER(C.c_int64)
        lib().swf_track(self._h, q.ctypes.data_as(P), ng.ctypes.data_as(P), ps.ctypes.data_as(P), n)
        return q[:n], ng[:n], ps[:n]


class Batch:
    """Independent instances advanced together (swf_batch_*): one launch of each
    stage kernel per RK stage over the tiles of all instances, one sweep workgroup
    per instance.  ``instance(i)`` is instance i's Solver (upload, bc, download,
    stats); ``advance`` runs every instance's own time loop."""

    def __init__(self, cfgs):
        self.cfgs = list(cfgs)
        n = len(self.cfgs)
        arr = (SwfConfig * n)(*self.cfgs)
        h = C.c_void_p()
        _check(lib().swf_batch_create(arr, n, C.byref(h)), "swf_batch_create")
        self._h = h
        self.n = n
        self._inst = [Solver(self.cfgs[i], _borrowed=lib().swf_batch_instance(h, i)) for i in range(n)]

    def instance(self, i: int) -> "Solver":
        return self._inst[i]

    def advance(self, times, tfinals, max_steps: int = 1 << 40):
        """Returns (times, steps, statuses) as numpy arrays of n."""
        t = np.ascontiguousarray(times, dtype=np.float64).copy()
        tf = np.ascontiguousarray(tfinals, dtype=np.float64)
        steps = np.zeros(self.n, dtype=np.int64)
        st = np.zeros(self.n, dtype=np.int32)
        rc = lib().swf_batch_advance(self._h, _dp(t), _dp(tf), max_steps,
                                     steps.ctypes.data_as(C.POINTER(C.c_int64)),
                                     st.ctypes.data_as(C.POINTER(C.c_int32)))
        if rc < 0:
            _check(rc, "swf_batch_advance")
        return t, steps, st

    def set_schedule(self, schedule: int) -> None:
        _check(lib().swf_batch_set_schedule(self._h, schedule), "swf_batch_set_schedule")

    def tune(self, reps: int = 10):
        """Times the stage kernels under each schedule, keeps the fastest: (schedule, ms[3])."""
        pick = C.c_int32(0)
        ms = (C.c_float * 3)()
        _check(lib().swf_batch_tune(self._h, reps, C.byref(pick), ms), "swf_batch_tune")
        return pick.value, list(ms)

    def close(self) -> None:
        if getattr(self, "_h", None):
            for s in self._inst:
                s._h = None
            lib().swf_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# the 1D program (ex2_sedov_1d/main.f90)
# ---------------------------------------------------------------------------

def sedov1d_config(n: int = 800, precision="fp64_exact", cfl: float = 0.6, swp_num: int = 2,
                   use_limiter: int = 2, device: int = 0) -> Swf1dConfig:
    """ex2_sedov_1d/main.f90:4-7,399-405: [-2, 2] with dx = 4/n, gamma 1.4, eps 1d-13.

    Defaults are the 1D benchmark case (cfl 0.6, swp_num 2); the program as shipped has
    cfl 1.2 and swp_num 50."""
    c = Swf1dConfig()
    c.n, c.dx, c.gamma, c.cfl, c.eps = n, (2.0 - (-2.0)) / n, 1.4, cfl, 1e-13
    c.use_limiter, c.swp_num, c.precision, c.device = use_limiter, swp_num, _prec(precision), device
    return c


def state1d_shape(cfg: Swf1dConfig) -> tuple:
    """numpy shape (n+6, 3) of the Fortran array u(3, -2:n+3)."""
    return (cfg.n + 6, 3)


class Solver1D:
    """One device context of the 1D program: state, `rk3` time loop and statistics."""

    def __init__(self, cfg: Swf1dConfig):
        self.cfg = cfg
        h = C.c_void_p()
        _check(lib().swf1d_create(C.byref(cfg), C.byref(h)), "swf1d_create")
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().swf1d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, u: np.ndarray) -> None:
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.shape != state1d_shape(self.cfg):
            raise ValueError(f"state shape {u.shape} != {state1d_shape(self.cfg)}")
        _check(lib().swf1d_upload(self._h, _dp(u)), "swf1d_upload")

    def download(self) -> np.ndarray:
        u = np.empty(state1d_shape(self.cfg))
        _check(lib().swf1d_download(self._h, _dp(u)), "swf1d_download")
        return u

    def advance(self, t: float, tfinal: float, max_steps: int = 1 << 40):
        """The program's time loop: steps while time <= tfinal - 1d-10.  Returns (time, steps)."""
        tt = C.c_double(t)
        n = C.c_int64(0)
        _check(lib().swf1d_advance(self._h, C.byref(tt), tfinal, max_steps, C.byref(n)), "swf1d_advance")
        return tt.value, n.value

    def last_ms(self) -> float:
        """Device time (HIP events) of the last advance."""
        ms = C.c_double(0.0)
        _check(lib().swf1d_last_ms(self._h, C.byref(ms)), "swf1d_last_ms")
        return ms.value

    def stats(self) -> dict:
        s = SwfStats()
        _check(lib().swf1d_stats(self._h, C.byref(s)), "swf1d_stats")
        return {n: getattr(s, n) for n, _ in SwfStats._fields_}

    def track(self):
        n = lib().swf1d_track(self._h, None, None, None, 0)
        if n < 0:
            _check(int(n), "swf1d_track")
        q = np.zeros(max(n, 1), np.int64)
        ng = np.zeros_like(q)
        ps = np.zeros_like(q)
        P = C.POINTER(C.c_int64)
        lib().swf1d_track(self._h, q.ctypes.data_as(P), ng.ctypes.data_as(P), ps.ctypes.data_as(P), n)
        return q[:n], ng[:n], ps[:n]
