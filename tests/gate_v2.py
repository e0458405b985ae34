"""Gate V2 benchmark problems and gates (core/tests/validation/test_gate_v2.cpp,
fixture core/tests/fixtures/benchmarks.json, copied to tests/golden/): the
projected OC loop under heaviside_continuation_schedule on two coarse meshes,
against reference values computed by an independent exact-solve reference."""
import json
import os

import numpy as np

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "benchmarks.json")
SCHEDULE = [(1.0, 0.2, 50), (2.0, 0.2, 50), (4.0, 0.2, 50), (8.0, 0.2, 50), (16.0, 0.1, 50),
            (32.0, 0.05, 50)]
RMIN, VOLFRAC, CG_TOL = 2.5, 0.3, 1e-9


def fixture():
    return json.load(open(GOLD))["benchmarks"]


def node(nx, ny, a, b, c):
    return (c * (ny + 1) + b) * (nx + 1) + a


def problem(name):
    """(dims, bcs, loads) of test_gate_v2.cpp:300-330."""
    nx, ny, nz = (24, 8, 8) if name == "cantilever_24x8x8" else (48, 8, 8)
    bcs, loads = [], []

    def clamp(a):
        for c in range(nz + 1):
            for b in range(ny + 1):
                for q in range(3):
                    bcs.append((node(nx, ny, a, b, c), q, 0.0))

    if name == "cantilever_24x8x8":
        clamp(0)
        fz = -1.0 / ((ny + 1) * (nz + 1))
        for c in range(nz + 1):
            for b in range(ny + 1):
                loads.append((node(nx, ny, nx, b, c), 2, fz))
    else:
        clamp(0)
        clamp(nx)
        fz = -1.0 / (ny + 1)
        for b in range(ny + 1):
            loads.append((node(nx, ny, nx // 2, b, nz), 2, fz))
    return (nx, ny, nz), bcs, loads


def components(dims, rho, thr=0.5):
    nx, ny, nz = dims
    solid = (rho > thr).reshape(nz, ny, nx)
    label = -np.ones_like(solid, dtype=np.int64)
    n = 0
    for k, j, i in zip(*np.nonzero(solid)):
        if label[k, j, i] >= 0:
            continue
        stack = [(k, j, i)]
        label[k, j, i] = n
        while stack:
            a, b, c = stack.pop()
            for da, db, dc in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                p, q, r = a + da, b + db, c + dc
                if 0 <= p < nz and 0 <= q < ny and 0 <= r < nx and solid[p, q, r] and label[p, q, r] < 0:
                    label[p, q, r] = n
                    stack.append((p, q, r))
        n += 1
    return n


def mnd(rho):
    return float(np.mean(4.0 * rho * (1.0 - rho)))


def y_symmetry(dims, rho):
    nx, ny, nz = dims
    r = rho.reshape(nz, ny, nx)
    return float(np.abs(r - r[:, ::-1, :]).max())


def check_gates(name, out, schedule=SCHEDULE):
    """The gates of test_gate_v2.cpp (benchmarks.json "gates")."""
    g = fixture()[name]["gates"]
    dims, _, _ = problem(name)
    h = out["history"]
    assert out["iterations"] == sum(s[2] for s in schedule) == len(h)
    start = 0
    for beta, move, iters in schedule:  # every OC step within its stage's move limit
        assert np.all(h[start:start + iters, 1] <= move + 1e-9)
        start += iters
    ends = [h[50 * (s + 1) - 1, 0] for s in range(len(schedule))]
    for key, s in (("stage_end_compliance_beta_1", 0), ("stage_end_compliance_beta_8", 3)):
        ref = g[key]
        assert abs(ends[s] - ref["value"]) <= ref["rel_tolerance"] * ref["value"], (key, ends[s])
    if "final_analysis_compliance_band" in g:
        lo, hi = g["final_analysis_compliance_band"]
        assert lo <= out["compliance"] <= hi
    else:
        ref = g["final_analysis_compliance"]
        assert abs(out["compliance"] - ref["value"]) <= ref["rel_tolerance"] * ref["value"]
    assert abs(out["volume_fraction"] - VOLFRAC) <= g["achieved_projected_volume_fraction_abs_tolerance"]
    rho = out["xphys"]
    assert components(dims, rho) == g["connected_components_at_0p5"]
    assert y_symmetry(dims, rho) <= g["y_symmetry_error_max"]
    assert mnd(rho) <= g["discreteness_Mnd_max"]
    return ends
