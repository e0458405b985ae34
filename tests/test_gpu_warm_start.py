"""Warm-started compliance (the PLSM state solve: a loose tolerance seeded with
the previous iteration's displacement, core/src/simp/plsm.cpp:671-689) on the
HIP path vs the oracle's warm-started simp_compliance (runs on an MI355X).

FP64 form: same design, same guess, same tolerance -> the oracle's iteration
count (+-1) and compliance to 1e-9 relative when the counts agree; the warm
solve takes fewer iterations than a cold one at that tolerance and lands within
the loose tolerance of the tight answer.
"""
import numpy as np
import pytest

import oracle
import topopt_amd as ta
from topopt_amd import problems

pytestmark = pytest.mark.gpu

NX, NY, NZ = 24, 8, 8
LOOSE = 1e-4


def setup_module():
    oracle.set_threads(0)


def _designs():
    """A smooth design and the next one (a small step, as a design iteration makes)."""
    k, j, i = np.meshgrid(np.arange(NZ), np.arange(NY), np.arange(NX), indexing="ij")
    x1 = (0.4 + 0.25 * np.sin(i / 3.0) * np.cos(j / 2.0) + 0.1 * np.cos(k)).ravel()
    x2 = np.clip(x1 + 0.03 * np.sin(i / 2.0 + j).ravel(), 0.05, 1.0)
    return x1, x2


def test_warm_compliance_matches_oracle_f64():
    bcs, loads = problems.cantilever(NX, NY, NZ)
    x1, x2 = _designs()
    s = oracle.Setup(NX, NY, NZ, bcs=bcs, loads=loads)
    _, _, u1_ref, _ = oracle.simp_compliance(s, x1, tol=1e-10)
    c_ref, dc_ref, _, info_ref = oracle.simp_compliance(s, x2, tol=LOOSE, guess=u1_ref)
    c_tight, _, _, _ = oracle.simp_compliance(s, x2, tol=1e-10)
    with ta.Problem(NX, NY, NZ, bcs, loads, precision=ta.PREC_F64) as p:
        _, _, u1, _ = p.compliance(x1, tol=1e-10, want_u=True)
        assert np.abs(u1 - u1_ref).max() / np.abs(u1_ref).max() < 1e-8
        c, dc, _, info = p.compliance(x2, tol=LOOSE, guess=u1)
        _, _, _, cold = p.compliance(x2, tol=LOOSE)
    print("warm", info["iterations"], info_ref["iterations"], "cold", cold["iterations"],
          c, c_ref, c_tight, flush=True)
    assert abs(info["iterations"] - info_ref["iterations"]) <= 1
    assert info["iterations"] < cold["iterations"]
    # same stopping iterate: agreement to rounding; one iteration apart: to the
    # size of one loose-tolerance step
    tol_c, tol_dc = (1e-9, 1e-8) if info["iterations"] == info_ref["iterations"] else (1e-5, 1e-4)
    assert abs(c - c_ref) / c_ref < tol_c
    assert np.abs(dc - dc_ref).max() / np.abs(dc_ref).max() < tol_dc
    assert abs(c - c_tight) / c_tight < 10 * LOOSE


def test_warm_start_at_the_solution_stops_after_one_iteration():
    """Seeded with the converged field the solve still runs its first
    iteration (convergence is tested after the update), as the oracle does."""
    bcs, loads = problems.cantilever(NX, NY, NZ)
    x1, _ = _designs()
    s = oracle.Setup(NX, NY, NZ, bcs=bcs, loads=loads)
    _, _, u_ref, _ = oracle.simp_compliance(s, x1, tol=1e-10)
    _, _, _, info_ref = oracle.simp_compliance(s, x1, tol=LOOSE, guess=u_ref)
    with ta.Problem(NX, NY, NZ, bcs, loads, precision=ta.PREC_F64) as p:
        c1, _, u1, _ = p.compliance(x1, tol=1e-10, want_u=True)
        c, _, _, info = p.compliance(x1, tol=LOOSE, guess=u1)
    assert info["iterations"] == info_ref["iterations"] == 1
    assert abs(c - c1) / c1 < 1e-9


def test_warm_mixed_form_within_loose_tolerance():
    bcs, loads = problems.cantilever(NX, NY, NZ)
    x1, x2 = _designs()
    s = oracle.Setup(NX, NY, NZ, bcs=bcs, loads=loads)
    c_tight, _, _, _ = oracle.simp_compliance(s, x2, tol=1e-10)
    with ta.Problem(NX, NY, NZ, bcs, loads) as p:
        _, _, u1, _ = p.compliance(x1, tol=1e-10, want_u=True)
        c, _, _, info = p.compliance(x2, tol=LOOSE, guess=u1)
        _, _, _, cold = p.compliance(x2, tol=LOOSE)
    assert info["iterations"] < cold["iterations"]
    assert abs(c - c_tight) / c_tight < 10 * LOOSE


def test_warm_guess_size_is_checked():
    bcs, loads = problems.cantilever(NX, NY, NZ)
    x1, _ = _designs()
    with ta.Problem(NX, NY, NZ, bcs, loads) as p:
        with pytest.raises(ValueError):
            p.compliance(x1, guess=np.zeros(5))
