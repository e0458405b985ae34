"""The per-run multigrid stagnation latch on the device path.

Reference: core/src/fea/multigrid.cpp:123-150 (policy), 2138-2155 (a latched
solve skips the hierarchy), 2261-2275 (bookkeeping); reset and query
core/include/topopt/fea.hpp:1081-1091. A solve that builds a hierarchy and does
not converge with it is a stagnation; three in a row latch multigrid off for the
thread until the driver resets it; a converging solve clears the count.

Stagnations are forced deterministically with an iteration cap of 2: multigrid
cannot converge in 2 cycles, the Jacobi-CG fallback neither, and the solve
raises SolverNonConvergence after counting the stagnation (the reference counts
it before its fallback throws). A latched solve is checked against the oracle's
Jacobi-CG solve (use_mg=False), the route the reference then takes.
"""
import threading

import numpy as np
import pytest

import oracle
import topopt_amd as ta
from topopt_amd import problems

pytestmark = pytest.mark.gpu

DIMS = (32, 16, 16)


def setup_module():
    oracle.set_threads(0)


def stagnate(p, E):
    with pytest.raises(ta.SolverNonConvergence):
        p.solve(E, tol=1e-8, max_iterations=2)


@pytest.mark.parametrize("precision", [ta.PREC_MIXED, ta.PREC_F64])
def test_three_stagnations_latch_multigrid_off(precision):
    nx, ny, nz = DIMS
    bcs, loads = problems.cantilever(nx, ny, nz)
    E = np.full(nx * ny * nz, 0.4 ** 3)
    ta.reset_mg_stagnation_latch()
    with ta.Problem(nx, ny, nz, bcs, loads, precision=precision) as p:
        for k in range(3):
            assert not ta.mg_stagnation_latched()
            stagnate(p, E)
        assert ta.mg_stagnation_latched()
        # latched: no hierarchy, straight to Jacobi-CG
        u, info = p.solve(E, tol=1e-8)
        print(f"latched solve: its {info['iterations']} hier_built {info['hier_built']} "
              f"mg {info['used_multigrid']}", flush=True)
        assert info["hier_built"] == 0 and info["used_multigrid"] == 0
        assert info["mg_cycles_attempted"] == 0 and info["converged"] == 1
        s = oracle.Setup(nx, ny, nz, bcs=bcs, loads=loads)
        u_ref, info_ref = oracle.solve(s, E, tol=1e-8, use_mg=False)
        assert np.abs(u - u_ref).max() / np.abs(u_ref).max() < 1e-6
        assert abs(info["iterations"] - info_ref["iterations"]) <= 1
        assert ta.mg_stagnation_latched()  # sticky until the driver resets it
        ta.reset_mg_stagnation_latch()
        assert not ta.mg_stagnation_latched()
        _, info = p.solve(E, tol=1e-8)
        assert info["hier_built"] == 1 and info["used_multigrid"] == 1
    ta.reset_mg_stagnation_latch()


def test_a_converging_solve_clears_the_count():
    nx, ny, nz = DIMS
    bcs, loads = problems.cantilever(nx, ny, nz)
    E = np.full(nx * ny * nz, 0.4 ** 3)
    ta.reset_mg_stagnation_latch()
    with ta.Problem(nx, ny, nz, bcs, loads) as p:
        stagnate(p, E)
        stagnate(p, E)
        _, info = p.solve(E, tol=1e-8)  # converges with multigrid: count back to 0
        assert info["used_multigrid"] == 1
        stagnate(p, E)
        stagnate(p, E)
        assert not ta.mg_stagnation_latched()
        stagnate(p, E)
        assert ta.mg_stagnation_latched()
    ta.reset_mg_stagnation_latch()


def test_latch_is_per_thread():
    nx, ny, nz = DIMS
    bcs, loads = problems.cantilever(nx, ny, nz)
    E = np.full(nx * ny * nz, 0.4 ** 3)
    ta.reset_mg_stagnation_latch()
    with ta.Problem(nx, ny, nz, bcs, loads) as p:
        for _ in range(3):
            stagnate(p, E)
        assert ta.mg_stagnation_latched()
        seen = {}

        def other():
            seen["latched"] = ta.mg_stagnation_latched()
            _, info = p.solve(E, tol=1e-8)
            seen["mg"] = info["used_multigrid"]

        t = threading.Thread(target=other)
        t.start()
        t.join()
        assert seen == {"latched": False, "mg": 1}
        assert ta.mg_stagnation_latched()
    ta.reset_mg_stagnation_latch()


def test_rearm_period_retries_multigrid():
    """The measurement-only re-arm period (multigrid.cpp:152-180, 2144-2156,
    2248-2262): 0 keeps the latch; 1 lets every latched solve retry multigrid,
    a stagnating retry re-latches at once (one stagnation, not three) and a
    converging one clears the latch and counts as a carry. The period survives
    the reset, which zeroes the counters."""
    nx, ny, nz = DIMS
    bcs, loads = problems.cantilever(nx, ny, nz)
    E = np.full(nx * ny * nz, 0.4 ** 3)
    ta.reset_mg_stagnation_latch()
    try:
        with ta.Problem(nx, ny, nz, bcs, loads) as p:
            for _ in range(3):
                stagnate(p, E)
            assert ta.mg_stagnation_latched()
            _, info = p.solve(E, tol=1e-8)  # period 0: stays latched, no hierarchy
            assert info["hier_built"] == 0 and ta.mg_rearm_counts() == (0, 0, 0)
            ta.set_mg_rearm_period(1)
            stagnate(p, E)  # re-armed retry stagnates: latched again at once
            assert ta.mg_stagnation_latched() and ta.mg_rearm_counts() == (1, 1, 0)
            _, info = p.solve(E, tol=1e-8)  # re-armed retry converges with multigrid
            print("rearmed solve", info["used_multigrid"], ta.mg_rearm_counts(), flush=True)
            assert info["hier_built"] == 1 and info["used_multigrid"] == 1
            assert not ta.mg_stagnation_latched() and ta.mg_rearm_counts() == (1, 2, 1)
            ta.reset_mg_stagnation_latch()
            assert ta.mg_rearm_counts() == (1, 0, 0)
    finally:
        ta.set_mg_rearm_period(0)
        ta.reset_mg_stagnation_latch()
