"""Oracle pins for the MMA updater (CPU only).

mma_update: on a separable problem with a closed-form optimum (identity
filter, f(x) = sum c_e / x_e, sum x_e = V: x_e* = V sqrt(c_e) / sum sqrt(c)) the
iterates converge to x*; the move box and bounds hold on every step; the volume
constraint of the result is met once the asymptotes settle. The plateau stop:
the reference's hand-built curves and verdicts
(tests/golden/mma_plateau_curves.json, from core/tests/unit/test_simp.cpp:1245-1435).
The loop: monotone descent on the cantilever and the volume target.
"""
import json
import os

import numpy as np
import pytest

import oracle
from topopt_amd import problems

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                  "mma_plateau_curves.json")))


def _separable(n_steps, nx=10, ny=6, nz=4, vf=0.3, move=0.2):
    s = oracle.Setup(nx, ny, nz, bcs=[], loads=[])
    N = nx * ny * nz
    c = np.random.default_rng(3).uniform(1.0, 4.0, N)
    x = np.full(N, vf)
    xs = []
    # identity filter: radius 1 gives every neighbour weight max(0, 1 - d) = 0
    for _ in range(n_steps):
        xs.append(x.copy())
        out, lam = oracle.mma_steps(s, 1.0, np.array(xs), np.array([-c / xx ** 2 for xx in xs]),
                                    vf, move)
        x = out[-1]
    return s, c, np.array(xs), x, vf, N


def test_mma_converges_to_the_closed_form_optimum():
    s, c, xs, x, vf, N = _separable(40)
    V = vf * N
    xstar = V * np.sqrt(c) / np.sqrt(c).sum()
    assert xstar.min() > 0.1 and xstar.max() < 1.0
    assert np.abs(x - xstar).max() < 1e-4
    assert abs(x.sum() - V) / V < 1e-6


def test_mma_steps_respect_bounds_and_move():
    s, c, xs, x, vf, N = _separable(8)
    move = 0.2
    outs, _ = oracle.mma_steps(s, 1.0, xs, np.array([-c / xx ** 2 for xx in xs]), vf, move)
    for xin, xout in zip(xs, outs):
        assert np.all(xout >= 1e-3 - 1e-15) and np.all(xout <= 1.0 + 1e-15)
        assert np.all(np.abs(xout - xin) <= move * (1.0 - 1e-3) + 1e-12)


def _curve(name):
    return FIX["curves"]["short3"] if name == "short3" else FIX["curves"][name]


def test_plateau_reference_verdicts():
    for name, pre, w, tol, drop, fw, want in FIX["cases"]:
        cv = _curve(name)
        if pre is not None:
            cv = cv[:pre]
        assert oracle.mma_plateau(cv, w, tol, drop, fw) == want, (name, pre, w, drop, fw)


def test_plateau_first_fire_points():
    for name, spec in FIX["first_fire"].items():
        cv = _curve(name)
        fire = next(n for n in range(1, len(cv) + 1)
                    if oracle.mma_plateau(cv[:n], spec["window"], 1e-3, spec["min_drop"],
                                          spec["flat_windows"]))
        best = min(cv[:fire])
        if "greater_than" in spec:
            assert fire > spec["greater_than"] and best < spec["best_below"], (name, fire)
        else:
            assert fire >= spec["at_least"], (name, fire)


def test_mma_loop_on_the_cantilever():
    nx, ny, nz = 16, 6, 6
    bcs, loads = problems.cantilever(nx, ny, nz)
    s = oracle.Setup(nx, ny, nz, bcs=bcs, loads=loads)
    out = oracle.optimize_mma(s, 0.3, 1.5, 30, cg_tol=1e-8, plateau_window=0)
    h = out["history"]
    assert out["iterations"] == 30
    assert h[-1, 0] < 0.6 * h[0, 0]
    assert abs(out["volume_fraction"] - 0.3) < 2e-3
    ref = oracle.optimize(s, 0.3, 1.5, 30, cg_tol=1e-8)
    # MMA and OC reach designs of similar stiffness on this problem
    assert abs(out["compliance"] - ref["compliance"]) / ref["compliance"] < 0.1


def test_masked_mma_with_an_all_active_mask_is_the_plain_loop():
    """mma_update_masked with an all-Active mask on an all-solid grid
    reproduces mma_update bit for bit (core/src/simp/simp.cpp:2733-2735)."""
    nx, ny, nz = 12, 4, 4
    bcs, loads = problems.cantilever(nx, ny, nz)
    s = oracle.Setup(nx, ny, nz, bcs=bcs, loads=loads)
    a = oracle.optimize_mma(s, 0.3, 1.5, 8, cg_tol=1e-10, plateau_window=0)
    b = oracle.optimize_mma(s, 0.3, 1.5, 8, cg_tol=1e-10, plateau_window=0,
                            mask=np.zeros(nx * ny * nz, np.uint8))
    assert np.array_equal(a["history"], b["history"])
    assert np.array_equal(a["design"], b["design"])


def test_masked_mma_moves_only_active_voxels():
    nx, ny, nz = 12, 4, 4
    bcs, loads = problems.cantilever(nx, ny, nz)
    s = oracle.Setup(nx, ny, nz, bcs=bcs, loads=loads)
    mask = np.zeros((nz, ny, nx), np.uint8)
    mask[:, :, -2:] = oracle.FROZEN_SOLID
    out = oracle.optimize_mma(s, 0.3, 1.5, 6, cg_tol=1e-10, plateau_window=0, mask=mask.ravel())
    m = mask.ravel()
    assert np.all(out["design"][m == oracle.FROZEN_SOLID] == 0.0)   # mma_update_joint's zeros
    assert np.all(out["xphys"][m == oracle.FROZEN_SOLID] == 1.0)    # pinned physical density
    assert abs(out["volume_fraction"] - 0.3) < 5e-3
    h = out["history"]
    assert h[0, 1] == 1.0 and np.all(h[1:, 1] < 1.0)   # the first step zeroes the frozen x
