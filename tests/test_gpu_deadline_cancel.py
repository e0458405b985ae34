"""Solve deadline and design-loop cancellation through the C ABI (MI355X).

* fea_set_solve_deadline_ms (core/include/topopt/fea.hpp:440-469): an armed
  deadline that has passed stops a solve with SolverDeadlineExceeded (a
  SolverNonConvergence) carrying the iterations and residual reached; a design
  loop ends non-convergent with deadline_exceeded; a deadline far in the future
  leaves the solve bit-identical to a disarmed one. The deadline is the calling
  thread's and sticky, so every test disarms it again.
* SimpOptions::cancel (core/include/topopt/simp.hpp:1505-1521): polled at the
  start of each design iteration; a cancelled run reports cancelled, not
  converged, its iterations so far equal an uncancelled run's, and its
  compliance is that of its returned physical density.
"""
import ctypes as C

import numpy as np
import pytest

import topopt_amd as ta
from topopt_amd import problems

pytestmark = pytest.mark.gpu

GRID = (32, 16, 16)


def _problem():
    bcs, loads = problems.cantilever(*GRID)
    return ta.Problem(*GRID, bcs, loads)


def test_passed_deadline_stops_the_solve():
    E = np.full(GRID[0] * GRID[1] * GRID[2], 0.3 ** 3)
    with _problem() as p:
        u0, info0 = p.solve(E, tol=1e-8)
        try:
            ta.set_solve_deadline_ms(ta.steady_ms() - 1.0)
            with pytest.raises(ta.SolverDeadlineExceeded) as e:
                p.solve(E, tol=1e-8)
            assert isinstance(e.value, ta.SolverNonConvergence)
            assert e.value.iterations < info0["iterations"]
            assert e.value.residual > 1e-8
            out = p.optimize(0.3, 1.5, 4, cg_tol=1e-8)
            assert out["non_convergent"] == 1 and out["deadline_exceeded"] == 1
            assert out["converged"] == 0 and out["iterations"] == 0
            # armed far in the future: never reached, nothing changes
            ta.set_solve_deadline_ms(ta.steady_ms() + 3.6e6)
            u1, info1 = p.solve(E, tol=1e-8)
        finally:
            prev = ta.set_solve_deadline_ms(0.0)
        assert prev > 0.0
        u2, _ = p.solve(E, tol=1e-8)
    assert np.array_equal(u1, u0) and info1["iterations"] == info0["iterations"]
    assert np.array_equal(u2, u0)


def test_cancel_flag_stops_the_loop_at_an_iteration_start():
    with _problem() as p:
        full = p.optimize(0.3, 1.5, 6, cg_tol=1e-8)
        flag = C.c_int(1)
        p.set_cancel_flag(flag)
        out0 = p.optimize(0.3, 1.5, 6, cg_tol=1e-8)
        assert out0["cancelled"] == 1 and out0["converged"] == 0 and out0["iterations"] == 0
        flag.value = 0
        seen = []

        def observe(o):
            seen.append(o.iteration)
            if o.iteration == 3:
                flag.value = 1  # another thread would do the same

        p.set_observer(observe)
        out = p.optimize(0.3, 1.5, 6, cg_tol=1e-8)
        p.set_observer(None)
        p.set_cancel_flag(None)
        c, _, _, _ = p.compliance(out["xphys"], tol=1e-8)
    print("cancel: iterations", out["iterations"], "observed", seen, flush=True)
    assert out["cancelled"] == 1 and out["converged"] == 0
    assert out["iterations"] == 3 and seen == [1, 2, 3]
    assert np.array_equal(out["history"], full["history"][:3])
    assert out["compliance"] == c
    assert full["cancelled"] == 0
