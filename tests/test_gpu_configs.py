"""The benchmark configurations and the solver's fallback and error routes on
the device, against the FP64 oracle (runs on an MI355X, through the C ABI).

* MBB half-beam 192x64x64 (2.45 M DOF), the headline workload, mixed form:
  the first design iterations against the oracle's (compliance per iteration
  within 1e-6 relative, the target this engine is built to), the displacement
  of the first solve (||u|| and max|du| / max|u| within 1e-6), and the OC
  update fed with the oracle's own iteration-1 design and sensitivities at full
  size, bit-exact (new design, lambda, bisection count, active set).
* MBB 48x16x16 over 20 iterations (the same BC set at a size the oracle runs
  in seconds): compliance history within 1e-5 relative per iteration, final
  density within 1e-4.
* Penalty continuation (1 -> 3 by 0.5, 6 iterations per stage) on a 32x16x16
  cantilever: history within 1e-6 per iteration through every stage change.
* The 512x256x256 cantilever builds the reference rule's six levels with a
  coarsest level of 3888 DOFs (<= 6000).
* The reference's measured count of 15 MG-PCG iterations on the solid 32^3
  cantilever at tol 1e-8 (docs/handoffs/072-geometric-multigrid-solver.md:34).
* An odd fine axis has no hierarchy (coarsen.hpp rule): the solve goes straight
  to the exact Jacobi-CG (core/src/fea/matfree.cpp:841-1076) and takes the
  oracle's iteration count (+-1) to the oracle's field.
* Error routes through the ABI: the void-DOF gate (every free DOF void, a load
  on a void DOF) is TOPOHIP_EUNDERCONSTRAINED; an iteration cap that neither MG
  nor Jacobi-CG can meet is TOPOHIP_ENONCONVERGENCE with the iterations and
  the residual reached (core/src/fea/matfree.cpp:782-814; fea.hpp:418-469).
"""
import os

import numpy as np
import pytest

import oracle
import topopt_amd as ta
from topopt_amd import problems

pytestmark = pytest.mark.gpu


def setup_module():
    oracle.set_threads(min(16, len(os.sched_getaffinity(0))))


def rel(a, b):
    return abs(a - b) / abs(b)


# ---- MBB 192x64x64 (the headline configuration) --------------------------------

C2 = dict(grid=(192, 64, 64), volfrac=0.2, rmin=2.0, tol=1e-6)
C2_ITERS = 3


@pytest.fixture(scope="module")
def c2_oracle():
    nx, ny, nz = C2["grid"]
    bcs, loads = problems.mbb(nx, ny, nz)
    s = oracle.Setup(nx, ny, nz, bcs=bcs, loads=loads)
    x0 = np.full(nx * ny * nz, C2["volfrac"])
    xp0 = oracle.filter(s, C2["rmin"], x0)
    c0, dc0, u0, info0 = oracle.simp_compliance(s, xp0, tol=C2["tol"])
    x1, nb1, lam1 = oracle.oc_update(s, C2["rmin"], x0, dc0, C2["volfrac"])
    run = oracle.optimize(s, C2["volfrac"], C2["rmin"], C2_ITERS, cg_tol=C2["tol"])
    print(f"oracle C2: c0 {c0:.10g} its {info0['iterations']} levels {info0['mg_levels']} "
          f"lambda {lam1!r} bisections {nb1}", flush=True)
    return dict(s=s, bcs=bcs, loads=loads, x0=x0, xp0=xp0, c0=c0, dc0=dc0, u0=u0, info0=info0,
                x1=x1, nb1=nb1, lam1=lam1, run=run)


def test_mbb_192_first_solve_matches_oracle(c2_oracle):
    o = c2_oracle
    nx, ny, nz = C2["grid"]
    with ta.Problem(nx, ny, nz, o["bcs"], o["loads"]) as p:
        c, dc, u, info = p.compliance(o["xp0"], tol=C2["tol"], want_u=True)
    du = np.abs(u - o["u0"]).max() / np.abs(o["u0"]).max()
    nu = rel(np.linalg.norm(u), np.linalg.norm(o["u0"]))
    ddc = np.abs(dc - o["dc0"]).max() / np.abs(o["dc0"]).max()
    print(f"C2 first solve: c {c:.12g} (oracle {o['c0']:.12g}, rel {rel(c, o['c0']):.2e}) "
          f"||u|| rel {nu:.2e} max|du|/max|u| {du:.2e} dc {ddc:.2e} iters {info['iterations']} "
          f"(oracle {o['info0']['iterations']}) levels {info['mg_levels']}", flush=True)
    assert info["used_multigrid"] == 1
    assert info["mg_levels"] == o["info0"]["mg_levels"] == 4
    assert abs(info["iterations"] - o["info0"]["iterations"]) <= 1
    assert rel(c, o["c0"]) < 1e-6
    assert nu < 1e-6 and du < 1e-6
    assert ddc < 1e-5


def test_mbb_192_oc_update_bit_exact_at_full_size(c2_oracle):
    """The oracle's iteration-1 inputs (uniform design, its sensitivities) and
    a random design with the same sensitivities: identical bits out."""
    o = c2_oracle
    nx, ny, nz = C2["grid"]
    n = nx * ny * nz
    with ta.Problem(nx, ny, nz) as p:
        xn, act, nb, lam = p.oc_update(C2["rmin"], o["x0"], o["dc0"], C2["volfrac"])
        gap, _, bound = p.oc_margin()
        print(f"C2 OC: lambda {lam!r} (oracle {o['lam1']!r}) bisections {nb} (oracle {o['nb1']}) "
              f"decision gap {gap:.2e} vs rounding bound {bound:.2e}", flush=True)
        assert nb == o["nb1"] and lam == o["lam1"]
        assert np.array_equal(xn, o["x1"])
        lo = np.maximum(1e-3, o["x0"] - 0.2)
        hi = np.minimum(1.0, o["x0"] + 0.2)
        assert np.array_equal(act == 1, (o["x1"] == lo) & (act != 2))
        assert np.all(xn[act == 2] == hi[act == 2])
        assert gap > bound
        xr = np.random.default_rng(21).uniform(1e-3, 1.0, n)
        xn2, act2, nb2, lam2 = p.oc_update(C2["rmin"], xr, o["dc0"], C2["volfrac"])
        gap2, _, _ = p.oc_margin()
    xo2, nbo2, lamo2 = oracle.oc_update(o["s"], C2["rmin"], xr, o["dc0"], C2["volfrac"])
    print(f"C2 OC random design: lambda {lam2!r} (oracle {lamo2!r}) gap {gap2:.2e}", flush=True)
    assert nb2 == nbo2 and lam2 == lamo2
    assert np.array_equal(xn2, xo2)
    assert np.count_nonzero(act2 == 1) == np.count_nonzero(xo2 == np.maximum(1e-3, xr - 0.2))


def test_mbb_192_design_iterations_match_oracle(c2_oracle):
    o = c2_oracle
    nx, ny, nz = C2["grid"]
    ref = o["run"]
    with ta.Problem(nx, ny, nz, o["bcs"], o["loads"]) as p:
        out = p.optimize(C2["volfrac"], C2["rmin"], C2_ITERS, cg_tol=C2["tol"])
        _, run_gap, bound = p.oc_margin()
    hc, hr = out["history"][:, 0], ref["history"][:, 0]
    errs = np.abs(hc - hr) / hr
    dx = np.abs(out["xphys"] - ref["xphys"]).max()
    print(f"C2 loop: compliance {hc} oracle {hr} rel {errs} final {out['compliance']:.10g} "
          f"(oracle {ref['compliance']:.10g}) density max|d| {dx:.2e} pcg "
          f"{out['history'][:, 3]} oracle {ref['history'][:, 3]} OC gap {run_gap:.2e} "
          f"(bound {bound:.2e})", flush=True)
    assert out["iterations"] == ref["iterations"] == C2_ITERS
    assert np.all(errs < 1e-6)
    assert rel(out["compliance"], ref["compliance"]) < 1e-6
    assert abs(out["volume_fraction"] - ref["volume_fraction"]) < 1e-9
    assert dx < 1e-5
    assert run_gap > bound


# ---- MBB 48x16x16, 20 iterations ------------------------------------------------

def test_mbb_48_twenty_iterations_match_oracle():
    nx, ny, nz = 48, 16, 16
    bcs, loads = problems.mbb(nx, ny, nz)
    s = oracle.Setup(nx, ny, nz, bcs=bcs, loads=loads)
    ref = oracle.optimize(s, 0.2, 2.0, 20, cg_tol=1e-6)
    res = {}
    for prec in (ta.PREC_MIXED, ta.PREC_F64):
        with ta.Problem(nx, ny, nz, bcs, loads, precision=prec) as p:
            res[prec] = p.optimize(0.2, 2.0, 20, cg_tol=1e-6)
    for prec, out in res.items():
        hc, hr = out["history"][:, 0], ref["history"][:, 0]
        err = np.abs(hc - hr) / hr
        dx = np.abs(out["xphys"] - ref["xphys"]).max()
        print(f"MBB 48x16x16 prec {prec}: max rel compliance err {err.max():.2e} (iter "
              f"{int(err.argmax()) + 1}), final {out['compliance']:.10g} vs "
              f"{ref['compliance']:.10g}, density {dx:.2e}", flush=True)
        assert out["iterations"] == 20
        assert np.all(err < 1e-5)
        assert rel(out["compliance"], ref["compliance"]) < 1e-5
        assert dx < 1e-4


# ---- penalty continuation --------------------------------------------------------

def test_penalty_continuation_matches_oracle():
    nx, ny, nz = 32, 16, 16
    bcs, loads = problems.cantilever(nx, ny, nz)
    s = oracle.Setup(nx, ny, nz, bcs=bcs, loads=loads)
    ramp = (1.0, 3.0, 0.5, 6)
    ref = oracle.optimize(s, 0.3, 1.5, 30, cg_tol=1e-8, continuation=ramp)
    with ta.Problem(nx, ny, nz, bcs, loads, precision=ta.PREC_F64) as p:
        out = p.optimize(0.3, 1.5, 30, cg_tol=1e-8, continuation=ramp)
    hc, hr = out["history"][:, 0], ref["history"][:, 0]
    err = np.abs(hc - hr) / hr
    print(f"continuation: max rel err {err.max():.2e}; stage-start compliance device "
          f"{hc[::6]} oracle {hr[::6]}", flush=True)
    # the ramp's jumps are visible: compliance rises when p steps up
    assert hr[6] > hr[5] and hr[12] > hr[11]
    assert np.all(err < 1e-6)
    assert rel(out["compliance"], ref["compliance"]) < 1e-6
    assert np.abs(out["xphys"] - ref["xphys"]).max() < 1e-5


# ---- the six-level plan of the 512x256x256 cantilever --------------------------------

def test_large_cantilever_builds_six_levels():
    nx, ny, nz = 512, 256, 256
    bcs, loads = problems.cantilever(nx, ny, nz)
    # reference rule: halve while every element axis is even and >= 4, stop
    # once the coarsest has <= 6000 active DOFs; node grids 513x257x257 ->
    # ... -> 17x9x9, whose clamped x = 0 face (9 x 9 nodes) leaves 3888 DOFs.
    expect_dofs = 3 * 17 * 9 * 9 - 3 * 9 * 9
    with ta.Problem(nx, ny, nz, bcs, loads) as p:
        c, _, _, info = p.compliance(np.full(nx * ny * nz, 0.1), tol=1e-6)
    print(f"512x256x256: levels {info['mg_levels']} coarse dofs {info['coarse_dofs']} "
          f"iters {info['iterations']} compliance {c:.8g}", flush=True)
    assert info["used_multigrid"] == 1
    assert info["mg_levels"] == 6
    assert info["coarse_dofs"] == expect_dofs <= 6000
    assert info["residual"] <= 1e-6


# ---- the reference's measured iteration count ----------------------------------------

@pytest.mark.parametrize("precision", [ta.PREC_F64, ta.PREC_MIXED])
def test_solid_cantilever_32_takes_the_reference_count(precision):
    n = 32
    bcs, loads = problems.cantilever(n, n, n)
    with ta.Problem(n, n, n, bcs, loads, precision=precision) as p:
        _, info = p.solve(np.ones(n ** 3), tol=1e-8)
    print(f"solid 32^3 prec {precision}: {info['iterations']} iterations, "
          f"{info['mg_levels']} levels", flush=True)
    assert info["used_multigrid"] == 1 and info["mg_levels"] == 3
    assert info["iterations"] == 15


# ---- Jacobi-CG when there is no hierarchy -----------------------------------------------

@pytest.mark.parametrize("precision", [ta.PREC_F64, ta.PREC_MIXED])
def test_odd_axis_goes_straight_to_jacobi_cg(precision):
    nx, ny, nz = 15, 6, 6
    bcs, loads = problems.cantilever(nx, ny, nz)
    rng = np.random.default_rng(4)
    E = rng.uniform(0.2, 1.0, nx * ny * nz) ** 3
    s = oracle.Setup(nx, ny, nz, bcs=bcs, loads=loads)
    assert oracle.hierarchy(s, E)[0] == 0
    u_ref, info_ref = oracle.solve(s, E, tol=1e-10)
    with ta.Problem(nx, ny, nz, bcs, loads, precision=precision) as p:
        u, info = p.solve(E, tol=1e-10)
    err = np.abs(u - u_ref).max() / np.abs(u_ref).max()
    print(f"Jacobi-CG prec {precision}: {info['iterations']} iterations (oracle "
          f"{info_ref['iterations']}), field err {err:.2e}", flush=True)
    assert info_ref["used_multigrid"] == 0 and info["used_multigrid"] == 0
    assert info["mg_levels"] == 0 and info["mg_cycles_attempted"] == 0
    assert info["converged"] == 1 and info["residual"] <= 1e-10
    assert abs(info["iterations"] - info_ref["iterations"]) <= 1
    assert err < 1e-8


# ---- error routes through the ABI ---------------------------------------------------------

def test_void_gate_is_underconstrained():
    nx, ny, nz = 6, 4, 4
    bcs, loads = problems.cantilever(nx, ny, nz)
    # every voxel empty: no free DOF is touched by a solid element
    with pytest.raises(ta.UnderConstrained):
        ta.Problem(nx, ny, nz, bcs, loads, solid=np.zeros(nx * ny * nz, dtype=np.uint8))
    # the loaded tip column empty: the load sits on a void DOF
    solid = np.ones((nz, ny, nx), dtype=np.uint8)
    solid[:, :, nx - 1] = 0
    with pytest.raises(ta.UnderConstrained):
        ta.Problem(nx, ny, nz, bcs, loads, solid=solid.ravel())
    s = oracle.Setup(nx, ny, nz, solid=solid.ravel(), bcs=bcs, loads=loads)
    with pytest.raises(oracle.OracleError) as e:
        oracle.solve(s, np.ones(nx * ny * nz), tol=1e-8)
    assert e.value.code == ta.EUNDERCONSTRAINED


def test_iteration_cap_is_nonconvergence():
    nx, ny, nz = 16, 8, 8
    bcs, loads = problems.cantilever(nx, ny, nz)
    E = np.random.default_rng(2).uniform(0.01, 1.0, nx * ny * nz) ** 3
    with ta.Problem(nx, ny, nz, bcs, loads) as p:
        with pytest.raises(ta.SolverNonConvergence) as e:
            p.solve(E, tol=1e-14, max_iterations=3)
        assert e.value.iterations == 3
        assert 0.0 < e.value.residual < 1.0
        # the handle stays usable after the failure
        _, info = p.solve(E, tol=1e-8)
        assert info["converged"] == 1
    s = oracle.Setup(nx, ny, nz, bcs=bcs, loads=loads)
    with pytest.raises(oracle.OracleError) as eo:
        oracle.solve(s, E, tol=1e-14, max_it=3)
    assert eo.value.code == ta.ENONCONVERGENCE


# ---- serving: one job of each stream grid size against the oracle --------------------------

SERVED = [
    dict(nx=48, ny=16, nz=16, bc_kind=0, volfrac=0.3, rmin=1.5, iterations=5, cg_tol=1e-6),
    dict(nx=96, ny=32, nz=32, bc_kind=1, volfrac=0.25, rmin=2.2, iterations=5, cg_tol=1e-6),
    dict(nx=128, ny=64, nz=32, bc_kind=0, volfrac=0.15, rmin=1.3, iterations=5, cg_tol=1e-6),
    dict(nx=192, ny=64, nz=64, bc_kind=1, volfrac=C2["volfrac"], rmin=C2["rmin"],
         iterations=C2_ITERS, cg_tol=C2["tol"]),
]


def _serve(jobs, max_slots, one_at_a_time):
    out = []
    with ta.Server(budget_bytes=64 << 30, max_slots=max_slots) as srv:
        ids = []
        for j in jobs:
            ids.append(srv.submit(j))
            if one_at_a_time:
                srv.drain()
        srv.drain()
        for j, jid in zip(jobs, ids):
            r = srv.poll(jid)
            assert r["state"] == 2 and r["status"] == 0, r
            out.append((r, srv.density(jid, j["nx"] * j["ny"] * j["nz"])))
    return out


def test_served_jobs_match_oracle_and_are_the_same_alone(c2_oracle):
    """The four grid sizes of the serving stream, served together (their design
    loops overlap on the device), against the oracle's runs of the same jobs
    (the 192x64x64 MBB job is the headline run above); then the same jobs one
    at a time: every result identical to its in-mix result."""
    mix = _serve(SERVED, max_slots=len(SERVED), one_at_a_time=False)
    spans = sorted((r["t_start_ms"], r["t_end_ms"]) for r, _ in mix)
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    for j, (r, x) in zip(SERVED, mix):
        if j["nx"] == 192:
            ref = c2_oracle["run"]
        else:
            bcs, loads = (problems.cantilever if j["bc_kind"] == 0 else problems.mbb)(
                j["nx"], j["ny"], j["nz"])
            s = oracle.Setup(j["nx"], j["ny"], j["nz"], bcs=bcs, loads=loads)
            ref = oracle.optimize(s, j["volfrac"], j["rmin"], j["iterations"], cg_tol=j["cg_tol"])
        dx = np.abs(x - ref["xphys"]).max()
        print(f"served {j['nx']}x{j['ny']}x{j['nz']} bc {j['bc_kind']}: compliance "
              f"{r['compliance']:.10g} oracle {ref['compliance']:.10g} rel "
              f"{rel(r['compliance'], ref['compliance']):.2e} density {dx:.2e}", flush=True)
        assert r["iterations"] == j["iterations"] == ref["iterations"]
        assert rel(r["compliance"], ref["compliance"]) < 1e-6
        assert abs(r["volume_fraction"] - ref["volume_fraction"]) < 1e-9
        assert dx < 1e-5
    alone = _serve(SERVED, max_slots=1, one_at_a_time=True)
    for (rm, xm), (ra, xa) in zip(mix, alone):
        assert rm["compliance"] == ra["compliance"]
        assert np.array_equal(xm, xa)
