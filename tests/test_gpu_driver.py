"""The Fortran driver (sweeping_fd_weno_amd/fortran/swf_driver, the reference's
ex4 parameter surface and time loop over the iso_c_binding module) against the
oracle: same density/pressure fields and the same printed sweep statistics."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "sweeping_fd_weno_amd", "fortran", "swf_driver")


def _fields(path, n):
    lines = open(path).read().split("\n", 2)[2] if "den" in path else open(path).read().split("\n", 1)[1]
    return np.array(lines.split(), dtype=np.float64).reshape(n * n, 3)[:, 2]


def test_driver_sedov_matches_oracle(tmp_path):
    n, tend = 48, 0.02
    res = subprocess.run([DRIVER, str(n), str(n), str(tend), "0", str(tmp_path)], capture_output=True, text=True,
                         timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    oc = O.sedov_config(n)
    o = O.Oracle(oc)
    o.set_state(O.init_sedov(oc))
    o.bc()
    rc, t, k = o.run(0.0, tend)
    assert rc == 0
    U = o.state()[3:-3, 3:-3, :]
    rho = U[:, :, 0].reshape(-1)
    g1 = 1.4 - 1.0
    p = g1 * (U[:, :, 3] - 0.5 * U[:, :, 1] ** 2 / U[:, :, 0] - 0.5 * U[:, :, 2] ** 2 / U[:, :, 0]).reshape(-1)
    assert np.array_equal(_fields(str(tmp_path / "den.dat"), n), rho)
    assert np.array_equal(_fields(str(tmp_path / "pre.dat"), n), p)
    out = res.stdout
    st = o.stats()
    assert f"CPU Time" in out and f" {k} " in out.split("CPU Time")[1]
    line = [l for l in out.splitlines() if "total sweep" in l][0]
    assert int(line.split("=")[1]) == st["total_sweep"]
    line = [l for l in out.splitlines() if "# of limiter" in l][0]
    assert [int(v) for v in line.split()[-2:]] == [st["den_count"], st["pre_count"]]
