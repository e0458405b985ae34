"""The Chebyshev smoother option against the oracle's restatement of it
(MI355X, through the C ABI). It is this engine's option, default off: the
reference smooths with damped Jacobi only (core/src/fea/multigrid.cpp:72,
1128-1160), so the oracle's Chebyshev V-cycle (oracle.cpp: power_lambda,
cheb_smooth) is the checker and is pinned by its own CPU tests.

* lambda_max(D^-1 A) per smoothed level: FP64 form within 1e-9 of the oracle,
  mixed form (FP32 level operators) within 1e-5;
* degree-2 solves: FP64 form PCG iterations within +-1 of the oracle's, field
  within 1e-6 (max-norm relative, tol 1e-8); mixed form within +2 iterations;
* a 5-iteration MBB design loop with degree 2 (FP64 form) tracks the oracle's
  compliance history to 1e-6;
* two load cases solved together equal each solved alone, bit for bit.
"""
import numpy as np
import pytest

import oracle
import topopt_amd as ta
from topopt_amd import problems

pytestmark = pytest.mark.gpu


def setup_module():
    oracle.set_threads(0)


def teardown_module():
    oracle.set_smoother("jacobi", 1)


def _graded(nx, ny, nz, bcs, loads):
    s = oracle.Setup(nx, ny, nz, bcs=bcs, loads=loads)
    x = oracle.filter(s, 2.0, np.random.default_rng(1).uniform(0.05, 1, nx * ny * nz))
    return s, np.clip(x, 1e-3, 1.0) ** 3


@pytest.mark.parametrize("precision", [ta.PREC_F64, ta.PREC_MIXED])
def test_chebyshev_solve_matches_oracle(precision):
    nx, ny, nz = 60, 20, 20
    bcs, loads = problems.cantilever(nx, ny, nz)
    s, E = _graded(nx, ny, nz, bcs, loads)
    try:
        oracle.set_smoother("chebyshev", 2)
        lam_ref = oracle.cheb_lambdas(s, E)
        u_ref, info_ref = oracle.solve(s, E, tol=1e-8)
    finally:
        oracle.set_smoother("jacobi", 1)
    with ta.Problem(nx, ny, nz, bcs, loads, precision=precision) as p:
        _, info_j = p.solve(E, tol=1e-8)
        p.set_smoother("chebyshev", 2)
        u, info = p.solve(E, tol=1e-8)
        lam = p.smoother_lambdas()
    lam_err = np.abs(lam - lam_ref) / lam_ref
    err = np.abs(u - u_ref).max() / np.abs(u_ref).max()
    print(f"chebyshev(2) prec {precision}: lambdas {lam} (oracle {lam_ref}, rel {lam_err}); "
          f"iterations {info['iterations']} (oracle {info_ref['iterations']}, Jacobi "
          f"{info_j['iterations']}); field err {err:.2e}", flush=True)
    assert len(lam) == len(lam_ref) == 2
    assert np.all(lam_err < (1e-9 if precision == ta.PREC_F64 else 1e-5))
    assert info["used_multigrid"] == info_ref["used_multigrid"] == 1
    assert info["residual"] <= 1e-8 and err < 1e-6
    if precision == ta.PREC_F64:
        assert abs(info["iterations"] - info_ref["iterations"]) <= 1
    else:
        assert info["iterations"] <= info_ref["iterations"] + 2
    assert info_ref["iterations"] < info_j["iterations"]


def test_chebyshev_design_loop_matches_oracle():
    nx, ny, nz = 48, 16, 16
    bcs, loads = problems.mbb(nx, ny, nz)
    s = oracle.Setup(nx, ny, nz, bcs=bcs, loads=loads)
    try:
        oracle.set_smoother("chebyshev", 2)
        ref = oracle.optimize(s, 0.2, 2.0, 5, cg_tol=1e-8)
    finally:
        oracle.set_smoother("jacobi", 1)
    with ta.Problem(nx, ny, nz, bcs, loads, precision=ta.PREC_F64) as p:
        p.set_smoother("chebyshev", 2)
        out = p.optimize(0.2, 2.0, 5, cg_tol=1e-8)
    hc, hr = out["history"][:, 0], ref["history"][:, 0]
    print("chebyshev loop", hc, hr, out["history"][:, 3], ref["history"][:, 3], flush=True)
    assert np.all(np.abs(hc - hr) / hr < 1e-6)
    assert np.all(np.abs(out["history"][:, 3] - ref["history"][:, 3]) <= 1)


def test_chebyshev_block_solve_equals_single_solves():
    nx, ny, nz = 24, 8, 12
    bcs, cases = problems.bridge(nx, ny, nz)
    x = np.random.default_rng(3).uniform(0.2, 1.0, nx * ny * nz)
    with ta.Problem(nx, ny, nz, bcs, cases[:2], precision=ta.PREC_MIXED) as p:
        p.set_smoother("chebyshev", 3)
        c2, dc2, _, _ = p.compliance(x, tol=1e-8)
    parts = []
    for case in cases[:2]:
        with ta.Problem(nx, ny, nz, bcs, case, precision=ta.PREC_MIXED) as p:
            p.set_smoother("chebyshev", 3)
            parts.append(p.compliance(x, tol=1e-8))
    assert c2 == parts[0][0] + parts[1][0]
    assert np.array_equal(dc2, parts[0][1] + parts[1][1])
