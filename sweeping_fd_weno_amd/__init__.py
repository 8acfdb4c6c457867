"""MI355X-native sweeping FD-WENO hot path.

Host-side mirror of the reference's time-stepping interface (dcgriffi/Sweeping_FD_WENO,
ex4_sedov_2d/main.f90 `rk` and its callees) over the C-ABI of ``libswf.so``
(include/swf.h).  The compute runs in hand-written HIP kernels for gfx950; there is
no CPU fallback: loading the package without the built library, or creating a solver
without a HIP device, raises.

    from sweeping_fd_weno_amd import Solver, sedov_config
    cfg = sedov_config(640, precision="fp64")
    s = Solver(cfg); s.upload(init_sedov(cfg)); t, n = s.advance(0.0, 1.0)
"""
from ._lib import (SwfConfig, SwfError, Solver, Batch, lib, library_path, sedov_config, vortex_config,
                   jet_config, weno_points, time_continue, state_shape, PRECISIONS, BC_SEDOV,
                   BC_PERIODIC, BC_JET_HALF, BC_JET_FULL, BC_STEP, step_config, step_inflow, Swf1dConfig, Solver1D, sedov1d_config,
                   state1d_shape)
from .problems import init_sedov, init_vortex, init_jet, vortex_exact, init_sedov1d, init_step
from .detonation import Swf7Config, Solver7, det_config, init_detonation

__all__ = [
    "SwfConfig", "SwfError", "Solver", "Batch", "lib", "library_path", "sedov_config", "vortex_config", "jet_config",
    "weno_points", "time_continue", "state_shape", "PRECISIONS", "BC_SEDOV", "BC_PERIODIC", "BC_JET_HALF",
    "BC_JET_FULL", "init_sedov", "init_vortex", "init_jet", "vortex_exact", "Swf1dConfig", "Solver1D",
    "sedov1d_config", "state1d_shape", "init_sedov1d", "BC_STEP", "step_config", "step_inflow", "init_step",
    "Swf7Config", "Solver7", "det_config", "init_detonation",
]
