"""Bit-reproducibility across processes (runs on an MI355X).

The same design run, started in two fresh processes, gives identical bits:
compliance history, physical density and displacement. The two processes are
pinned to different K.u z-chunks (the launch-shape choice the autotuner makes
per process), which must not matter: every reduction in the solve has a
fixed order and a fixed partial-sum layout.
"""
import hashlib
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import hashlib, json, sys
import numpy as np
sys.path.insert(0, %r)
import topopt_amd as ta
from topopt_amd import problems
nx, ny, nz = 48, 16, 16
bcs, loads = problems.mbb(nx, ny, nz)
with ta.Problem(nx, ny, nz, bcs, loads) as p:
    out = p.optimize(0.2, 2.0, 5, cg_tol=1e-6)
    c, dc, u, info = p.compliance(out["xphys"], tol=1e-8, want_u=True)
    tile = p.ku_tiles()
h = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
print(json.dumps({"tile": tile, "history": h(out["history"]), "xphys": h(out["xphys"]),
                  "u": h(u), "dc": h(dc), "c": c.hex(), "its": info["iterations"]}))
"""


def _run(tz):
    env = dict(os.environ, TOPOHIP_KU_TILE=str(tz))
    r = subprocess.run([sys.executable, "-c", CHILD % ROOT], env=env, capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_two_processes_give_identical_bits():
    print("repro: two fresh processes, z-chunks 4:7 and 16:28", flush=True)
    a = _run("4:7")
    b = _run("16:28")
    print(a, "\n", b, flush=True)
    assert a["tile"] == [4, 7] and b["tile"] == [16, 28]
    for k in ("history", "xphys", "u", "dc", "c", "its"):
        assert a[k] == b[k], k
