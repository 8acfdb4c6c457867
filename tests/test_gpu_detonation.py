"""The detonation program (ex7_detonation/main.f90) on the GPU in FP64-exact against its
reference runs (tests/golden/ref_ex7_detonation.json, one-rank build copies of the
program): 40 x 40 (17 steps, 3 sweeps) and 80 x 80 (33 steps, 8 sweeps) to T = 0.2 --
den / pre / rate as the program writes them bitwise (SHA-256), the step count and every
printed statistic."""
import json
import os

import numpy as np
import pytest

from sweeping_fd_weno_amd import detonation as D
from sweeping_fd_weno_amd.build import ROOT

GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "ref_ex7_detonation.json")))


def _sha(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


@pytest.mark.gpu
@pytest.mark.parametrize("key,total", [("total8", 8), ("total16", 16)])
def test_detonation_bitwise_to_reference_run(key, total):
    g = GOLD[key]
    r = D.run(total, g["t_end"])
    assert r["code"] == 0
    assert r["steps"] == g["steps"]
    for nm in ("den", "pre", "rate"):
        assert _sha(r["fields"][nm]) == g["sha256"][nm], nm
    st, s = g["stats"], r["stats"]
    assert (s["total_sweep"], s["max_sweep"]) == (st["total sweep"], st["max sweep"])
    assert (s["total_neg"], s["max_neg"]) == (st["total negative pairs"], st["max negative pairs"])
    assert (s["total_pos"], s["max_pos"]) == (st["total positive neighbors"], st["max positive neighbors"])
    assert (s["den_count"], s["pre_count"]) == (st["den_count"], st["pre_count"])
