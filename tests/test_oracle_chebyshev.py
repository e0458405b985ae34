"""Pins of the oracle's Chebyshev smoother option (CPU only).

The reference has no Chebyshev smoother, so nothing of its own pins this; the
checks come from the mathematics:
* the power-iteration estimate of lambda_max(D^-1 A) lies below the largest
  eigenvalue of the level's generalized problem A v = lambda D v (scipy, on the
  oracle's level operator, itself pinned to the Galerkin product), and close
  enough that no mode is amplified: the smoothing polynomial of the interval
  [0.3, 1.1] * estimate exceeds 1 in magnitude only beyond 1.4 * estimate,
  above the true top of the spectrum;
* the V-cycle it smooths stays a symmetric positive preconditioner: MG-PCG
  with it reaches the same field as the Jacobi-smoothed solve (1e-6) and a
  dense direct solve on a small grid (1e-8);
* on the graded cantilever it needs fewer PCG iterations than damped Jacobi.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from topopt_amd import problems


def setup_module():
    oracle.set_threads(0)


def teardown_module():
    oracle.set_smoother("jacobi", 1)


def _case(nx, ny, nz):
    bcs, loads = problems.cantilever(nx, ny, nz)
    s = oracle.Setup(nx, ny, nz, bcs=bcs, loads=loads)
    x = oracle.filter(s, 2.0, np.random.default_rng(1).uniform(0.05, 1, nx * ny * nz))
    return s, np.clip(x, 1e-3, 1.0) ** 3


def test_power_estimate_brackets_the_top_eigenvalue():
    s, E = _case(40, 20, 20)  # three levels: the fine one and level 1 are smoothed
    try:
        oracle.set_smoother("chebyshev", 2)
        lam = oracle.cheb_lambdas(s, E)
    finally:
        oracle.set_smoother("jacobi", 1)
    assert len(lam) == 2
    n, ptr, col, val = oracle.level_operator(s, E, 1)
    A = sp.csr_matrix((val, col, ptr), shape=(n, n))
    d = A.diagonal()
    top = spla.eigsh(A, k=1, M=sp.diags(d), which="LA", return_eigenvectors=False, tol=1e-10)[0]
    print("level-1 lambda_max", top, "estimate", lam[1], flush=True)
    assert lam[1] <= top * (1 + 1e-9)
    assert top < (0.3 + 1.1) * lam[1]


@pytest.mark.parametrize("degree", [1, 2, 3])
def test_chebyshev_vcycle_preconditions_to_the_same_field(degree):
    s, E = _case(60, 20, 20)
    u_j, info_j = oracle.solve(s, E, tol=1e-10)
    try:
        oracle.set_smoother("chebyshev", degree)
        u_c, info_c = oracle.solve(s, E, tol=1e-10)
    finally:
        oracle.set_smoother("jacobi", 1)
    print(f"degree {degree}: {info_c['iterations']} iterations (Jacobi {info_j['iterations']})",
          flush=True)
    assert info_c["used_multigrid"] == 1 and info_c["converged"] == 1
    assert np.abs(u_c - u_j).max() / np.abs(u_j).max() < 1e-6
    if degree >= 2:
        assert info_c["iterations"] < info_j["iterations"]


def test_chebyshev_solve_matches_direct_solve():
    nx, ny, nz = 8, 4, 4
    bcs, loads = problems.cantilever(nx, ny, nz)
    s = oracle.Setup(nx, ny, nz, bcs=bcs, loads=loads)
    E = np.random.default_rng(2).uniform(0.1, 1.0, nx * ny * nz)
    u_ref, _ = oracle.solve(s, E, tol=1e-13, use_mg=False)
    try:
        oracle.set_smoother("chebyshev", 2)
        u, info = oracle.solve(s, E, tol=1e-12)
    finally:
        oracle.set_smoother("jacobi", 1)
    assert info["used_multigrid"] == 1
    assert np.abs(u - u_ref).max() / np.abs(u_ref).max() < 1e-8
