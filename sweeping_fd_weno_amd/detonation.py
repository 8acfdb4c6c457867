"""The detonation program (ex7_detonation/main.f90) on one domain over the swf7_* C-ABI.

    python -m sweeping_fd_weno_amd.detonation --total 8 --t-end 0.2

The parameter block (:11-18, :814) and initdata (:1196-1260) are restated here; the time
loop, wenofd, scheme with chemsource, rk, checkpositivity, the sweep and bc run on the
GPU in FP64-exact (bit-identical to the reference's one-rank run).  `output_fields`
forms den.dat / pre.dat / rate.dat's value columns as the program's output (:1337-1358)
writes them.
"""
from __future__ import annotations

import argparse
import ctypes as C
import time as _time

import numpy as np

from . import _lib as L


class Swf7Config(C.Structure):
    _fields_ = [
        ("total", C.c_int32), ("step", C.c_int32), ("dx", C.c_double), ("dy", C.c_double),
        ("gamma", C.c_double), ("q_const", C.c_double), ("t_const", C.c_double), ("k_const", C.c_double),
        ("eps", C.c_double), ("eps2", C.c_double), ("cfl", C.c_double),
        ("swp_num", C.c_int32), ("device", C.c_int32),
    ]


def det_config(total: int = 80, device: int = 0, swp_num: int = 20000) -> Swf7Config:
    """PARAMETERs of ex7 at cut_x = cut_y = 1: nx = ny = 5 total on [0, 5]^2 (dx = xlen/nx,
    :1228-1235), gamma 1.2, q 50, t 50, k 2566.4, eps 1d-13, eps2 1d-8, cfl 0.88, step 2."""
    n = 5 * total
    xlen = 5.0 / 1 * (0 + 1) - 5.0 / 1 * 0
    return Swf7Config(total=total, step=2, dx=xlen / n, dy=xlen / n, gamma=1.2, q_const=50.0, t_const=50.0,
                      k_const=2566.4, eps=1e-13, eps2=1e-8, cfl=0.88, swp_num=swp_num, device=device)


def state_shape(cfg: Swf7Config) -> tuple:
    n = 5 * cfg.total
    return (n + 6, n + 6, 5)  # u(5, -2:n+3, -2:n+3) as a C array [j][i][c]


def init_detonation(cfg: Swf7Config) -> np.ndarray:
    """initdata before its `call bc`: u = 0, then cells 0..n+1 left of x = 0.5 the burnt
    gas (11, 11*6.18, 0, 970, 11) and the others the unburnt gas at rest (1, 0, 0, 55, 1)."""
    n = 5 * cfg.total
    u = np.zeros(state_shape(cfg))
    dx = cfg.dx
    for i in range(0, n + 2):
        x = (0.0 + i * dx) - dx / 2.0
        v = (11.0, 11.0 * 6.18, 0.0, 970.0, 11.0) if x < 0.5 else (1.0, 0.0, 0.0, 55.0, 1.0)
        u[2:n + 4, i + 2, :] = v
    return u


def output_fields(cfg: Swf7Config, u: np.ndarray) -> dict:
    """den / pre / rate value columns of the program's files, (j outer, i inner) over the
    interior, with 100 / 300 / 20 where i <= total and j < step*total."""
    n = 5 * cfg.total
    g1 = cfg.gamma - 1.0
    inner = u[3:n + 3, 3:n + 3, :]
    r, m, nn, e, y = (inner[..., c] for c in range(5))
    pres = g1 * (((e - (0.5 * (m * m)) / r) - (0.5 * (nn * nn)) / r) - cfg.q_const * y)
    rate = y / r
    den = r.copy()
    jj, ii = np.meshgrid(np.arange(1, n + 1), np.arange(1, n + 1), indexing="ij")
    blank = (ii <= cfg.total) & (jj < cfg.step * cfg.total)
    den[blank], pres[blank], rate[blank] = 100.0, 300.0, 20.0
    return {"den": den.ravel(), "pre": pres.ravel(), "rate": rate.ravel()}


class Solver7:
    """One detonation domain on the GPU (swf7_* in include/swf.h)."""

    def __init__(self, cfg: Swf7Config):
        self.cfg = cfg
        lib = L.lib()
        P, vp, dp = C.POINTER, C.c_void_p, C.POINTER(C.c_double)
        lib.swf7_create.argtypes = [P(Swf7Config), P(vp)]
        lib.swf7_destroy.argtypes = [vp]
        lib.swf7_destroy.restype = None
        lib.swf7_upload.argtypes = [vp, dp]
        lib.swf7_download.argtypes = [vp, dp]
        lib.swf7_bc.argtypes = [vp]
        lib.swf7_advance.argtypes = [vp, dp, C.c_double, C.c_int64, P(C.c_int64)]
        lib.swf7_stats.argtypes = [vp, P(L.SwfStats)]
        lib.swf7_sweep.argtypes = [vp]
        lib.swf7_track.argtypes = [vp, P(C.c_int64), P(C.c_int64), P(C.c_int64), C.c_int64]
        lib.swf7_track.restype = C.c_int64
        self._lib = lib
        h = vp()
        L._check(lib.swf7_create(C.byref(cfg), C.byref(h)), "swf7_create")
        self._h = h

    def close(self) -> None:
        if self._h:
            self._lib.swf7_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, u: np.ndarray) -> None:
        a = np.ascontiguousarray(u, dtype=np.float64)
        assert a.shape == state_shape(self.cfg)
        L._check(self._lib.swf7_upload(self._h, a.ctypes.data_as(C.POINTER(C.c_double))), "swf7_upload")

    def download(self) -> np.ndarray:
        a = np.empty(state_shape(self.cfg))
        L._check(self._lib.swf7_download(self._h, a.ctypes.data_as(C.POINTER(C.c_double))), "swf7_download")
        return a

    def bc(self) -> None:
        L._check(self._lib.swf7_bc(self._h), "swf7_bc")

    def advance(self, t: float, t_end: float, max_steps: int = 1 << 40) -> tuple:
        """Steps while time <= t_end - 1d-10 (binary128); returns (time, steps, code) with
        code the swf.h return code of the reference's stop paths (0 when none)."""
        tt = C.c_double(t)
        n = C.c_int64(0)
        rc = self._lib.swf7_advance(self._h, C.byref(tt), t_end, max_steps, C.byref(n))
        if rc < 0:
            L._check(rc, "swf7_advance")
        return tt.value, n.value, rc

    def sweep(self) -> int:
        """sweep_lim on the current state; returns 0 or the stop's swf.h code."""
        rc = self._lib.swf7_sweep(self._h)
        if rc < 0:
            L._check(rc, "swf7_sweep")
        return rc

    def track(self) -> np.ndarray:
        """(q, negative pairs, positive neighbours) of each sweep so far."""
        n = self.stats()["sweep_count"]
        q, ng, ps = (np.zeros(max(n, 1), dtype=np.int64) for _ in range(3))
        P = C.POINTER(C.c_int64)
        m = self._lib.swf7_track(self._h, q.ctypes.data_as(P), ng.ctypes.data_as(P), ps.ctypes.data_as(P), n)
        return np.stack([q[:m], ng[:m], ps[:m]], axis=1)

    def stats(self) -> dict:
        s = L.SwfStats()
        L._check(self._lib.swf7_stats(self._h, C.byref(s)), "swf7_stats")
        return {n: getattr(s, n) for n, _ in s._fields_}


def run(total: int, t_end: float, device: int = 0) -> dict:
    cfg = det_config(total, device=device)
    s = Solver7(cfg)
    s.upload(init_detonation(cfg))
    s.bc()
    t0 = _time.perf_counter()
    t, steps, rc = s.advance(0.0, t_end)
    secs = _time.perf_counter() - t0
    u = s.download()
    out = {"total": total, "n": 5 * total, "t": t, "steps": steps, "code": rc, "seconds": secs,
           "stats": s.stats(), "fields": output_fields(cfg, u)}
    s.close()
    return out


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--total", type=int, default=8)
    ap.add_argument("--t-end", type=float, default=0.2)
    a = ap.parse_args()
    r = run(a.total, a.t_end)
    print(f"{r['n']}x{r['n']} to t = {r['t']:.6g}: {r['steps']} steps in {r['seconds']:.3f} s, code {r['code']}")
    print({k: v for k, v in r["stats"].items() if k != "stage_sweeps"})


if __name__ == "__main__":
    main()
