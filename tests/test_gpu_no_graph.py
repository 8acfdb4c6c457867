"""SWF_NO_GRAPH=1 (every step launched on the stream, no graph captured or launched) gives
the graph path's results bit for bit: a Sedov run with sweeps and a small batch, each advanced
past several 16-step graph lengths, in a child process with the switch set (the library
reads it once per process) against this process."""
import hashlib
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import hashlib, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import sweeping_fd_weno_amd as S
from sweeping_fd_weno_amd import _lib as L
cfg = S.sedov_config(64, "fp64_exact")
s = S.Solver(cfg)
s.upload(S.init_sedov(cfg))
s.bc()
t, k = s.advance(0.0, 1.0, max_steps=50)
st = s.stats()
h = hashlib.sha1(s.download().tobytes()).hexdigest()
s.close()
cfgs = [S.sedov_config(48, "fp64"), S.vortex_config(45, "fp64")]
bt = L.Batch(cfgs)
for i, c in enumerate(cfgs):
    bt.instance(i).upload(S.init_sedov(c) if i == 0 else S.init_vortex(c))
    bt.instance(i).bc()
tt, n, _ = bt.advance(np.zeros(2), np.array([1.0, 1.0]), max_steps=40)
hb = hashlib.sha1(b"".join(bt.instance(i).download().tobytes() for i in range(2))).hexdigest()
bt.close()
print(h, k, repr(t), st["sweep_count"], st["total_neg"], st["total_pos"], hb, list(n), flush=True)
"""


def _run(env_extra):
    env = dict(os.environ)
    env.pop("SWF_NO_GRAPH", None)
    env.update(env_extra)
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT], capture_output=True, text=True, env=env, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout.strip().splitlines()[-1]


@pytest.mark.gpu
def test_stream_launches_equal_graph_launches():
    with_graph = _run({})
    without = _run({"SWF_NO_GRAPH": "1"})
    print(with_graph)
    assert with_graph == without
