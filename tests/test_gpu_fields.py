"""Result fields and run records from the device (runs on an MI355X).

von Mises: the device field against the oracle's fea_von_mises_field on a
solved displacement with void voxels (rel <= 1e-15; the device sums the
reference's terms in its order without contraction), and against the closed
form of a uniform strain. Narrowed arrays: float32 of the oracle field and of
the displacement, exact. Snapshot halves: bit-exact against the oracle's
float_to_half and numpy's IEEE float16 for every finite class. The observer
reports what the design loop's history holds, and its rows make a CSV.
"""
import json

import numpy as np
import pytest

import oracle
import topopt_amd as ta
from topopt_amd import problems

pytestmark = pytest.mark.gpu


def setup_module():
    oracle.set_threads(0)


def test_von_mises_matches_oracle_on_a_solved_field():
    nx, ny, nz = 20, 8, 6
    bcs, loads = problems.cantilever(nx, ny, nz)
    solid = np.ones(nx * ny * nz, np.uint8)
    solid[np.random.default_rng(4).choice(solid.size, 60, replace=False)] = 0
    solid.reshape(nz, ny, nx)[:, :, 0] = 1
    solid.reshape(nz, ny, nx)[:, :, -1] = 1
    with ta.Problem(nx, ny, nz, bcs, loads, h=0.7, solid=solid, precision=ta.PREC_F64) as p:
        E = 210.0
        u, _ = p.solve(np.full(nx * ny * nz, E), tol=1e-10)
        vm = p.von_mises(E)            # resident displacement of the solve
        vm_u = p.von_mises(E, u=u)     # the same field passed in
        vm32, du32 = p.fields_f32(E)
    s = oracle.Setup(nx, ny, nz, h=0.7, solid=solid, bcs=bcs, loads=loads)
    ref = oracle.von_mises_field(s, u, E=E, nu=0.3)
    assert np.array_equal(vm, vm_u)
    assert np.all(vm[solid == 0] == 0) and np.all(vm[solid == 1] > 0)
    rel = np.abs(vm - ref) / np.maximum(np.abs(ref), 1e-300)
    print("max rel", rel.max(), "bitwise mismatches", int(np.sum(vm != ref)), flush=True)
    assert rel.max() <= 1e-15
    same = vm == ref
    assert np.array_equal(vm32[same], ref[same].astype(np.float32))
    assert np.array_equal(du32, u.astype(np.float32))


def test_von_mises_uniform_strain_closed_form():
    nx, ny, nz, h = 5, 4, 3, 2.0
    bcs, loads = problems.cantilever(nx, ny, nz)
    k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    a, g = 1e-3, 5e-4
    u = np.zeros((nz + 1, ny + 1, nx + 1, 3))
    u[..., 0] = a * i * h + g * j * h   # eps_xx = a, gamma_xy = g
    E, nu = 200.0, 0.3
    c = E / ((1 + nu) * (1 - 2 * nu))
    G = E / (2 * (1 + nu))
    sxx, syy, txy = c * (1 - nu) * a, c * nu * a, G * g
    want = np.sqrt((sxx - syy) ** 2 + 3 * txy ** 2)
    with ta.Problem(nx, ny, nz, bcs, loads, h=h) as p:
        vm = p.von_mises(E, u=u.ravel())
    assert np.allclose(vm, want, rtol=1e-13, atol=0)


def test_von_mises_needs_a_displacement():
    bcs, loads = problems.cantilever(4, 2, 2)
    with ta.Problem(4, 2, 2, bcs, loads) as p:
        with pytest.raises(ValueError):
            p.von_mises(1.0)
        with pytest.raises(ValueError):
            p.von_mises(-1.0, u=np.zeros(p.ndof))


def test_density_halves_bit_exact():
    rng = np.random.default_rng(12)
    nx, ny, nz = 64, 32, 33          # 67584 voxels: not a multiple of 4 per thread tail
    n = nx * ny * nz
    vals = np.concatenate([
        rng.uniform(0, 1, n // 2),
        np.exp(rng.uniform(np.log(1e-9), np.log(7e4), n // 4)) * rng.choice([-1, 1], n // 4),
        np.array([0.0, -0.0, 1.0, 0.5, 65504.0, 65519.99, 65520.0, 1e300, -1e300, 2.0 ** -24,
                  2.0 ** -25, 3 * 2.0 ** -26, 6.103515625e-05, 1e-320, np.inf, -np.inf]),
    ])
    vals = np.concatenate([vals, rng.uniform(0, 1, n - vals.size)])
    bcs, loads = problems.cantilever(nx, ny, nz)
    with ta.Problem(nx, ny, nz, bcs, loads) as p:
        got = p.density_f16(vals)
    with np.errstate(over="ignore"):
        f32 = vals.astype(np.float32)
    assert np.array_equal(got, oracle.float_to_half(f32))
    with np.errstate(over="ignore"):
        assert np.array_equal(got, f32.astype(np.float16).view(np.uint16))


def test_observer_rows_match_history_and_make_a_csv(tmp_path):
    nx, ny, nz = 24, 8, 8
    bcs, loads = problems.cantilever(nx, ny, nz)
    rows = []
    with ta.Problem(nx, ny, nz, bcs, loads) as p:
        p.set_observer(lambda o: rows.append(ta.IterationObs.from_buffer_copy(o)))
        out = p.optimize(0.3, 1.5, 6, cg_tol=1e-8)
        snap = p.density_f16()
        p.write_density_snapshot(tmp_path / "snaps", "snap_rung0_iter0006_boundary", 0, 6, True,
                                 origin=(1.0, 2.0, 3.0))
        p.set_observer(None)
        p.optimize(0.3, 1.5, 2, cg_tol=1e-8)   # cleared: no more rows
    h = out["history"]
    assert [r.iteration for r in rows] == list(range(1, 7))
    assert np.array_equal([r.compliance for r in rows], h[:, 0])
    assert np.array_equal([r.change for r in rows], h[:, 1])
    assert np.array_equal([r.volume_fraction for r in rows], h[:, 2])
    assert np.array_equal([r.cg_iterations for r in rows], h[:, 3])
    for r in rows:
        assert r.cg_used_multigrid == 1 and r.cg_hier_built == 1 and r.beta == 0.0
        assert r.fea_solves == 1 and r.matvecs > 0 and r.rss_mb > 0 and r.peak_rss_mb > 0
        assert r.solve_ms > 0 and r.fea_ms > 0 and r.total_ms >= r.solve_ms
        parts = r.filter_ms + r.project_ms + r.solve_ms + r.update_ms + r.analysis_ms + r.observe_ms
        assert abs(parts + r.residual_ms - r.total_ms) < 1e-9
    path = tmp_path / "iterations.csv"
    with ta.CsvWriter(path) as w:
        for r in rows:
            w.append(0, r)
    lines = path.read_text().splitlines()
    assert len(lines) == 7 and lines[0] == ta.iterations_csv_header()
    assert [int(l.split(",")[1]) for l in lines[1:]] == list(range(1, 7))
    # the snapshot: resident physical density, halves as the device formed them
    raw = np.fromfile(tmp_path / "snaps" / "snap_rung0_iter0006_boundary.f16", dtype="<u2")
    assert np.array_equal(raw, snap)
    assert np.array_equal(snap, oracle.float_to_half(out["xphys"].astype(np.float32)))
    js = json.loads((tmp_path / "snaps" / "snap_rung0_iter0006_boundary.json").read_text())
    assert js == {"dims": [nx, ny, nz], "spacing": 1, "origin": [1, 2, 3], "rung": 0, "iter": 6,
                  "boundary": True, "voxels": nx * ny * nz, "dtype": "float16",
                  "order": "x-fastest"}


def test_projected_observer_reports_stage_beta():
    nx, ny, nz = 16, 8, 8
    bcs, loads = problems.cantilever(nx, ny, nz)
    rows = []
    with ta.Problem(nx, ny, nz, bcs, loads) as p:
        p.set_observer(lambda o: rows.append((o.iteration, o.beta)))
        p.optimize_projected(0.3, 1.5, [(1.0, 0.2, 2), (4.0, 0.1, 3)], cg_tol=1e-8)
    assert rows == [(1, 1.0), (2, 1.0), (3, 4.0), (4, 4.0), (5, 4.0)]


def test_run_records_of_a_design(tmp_path):
    nx, ny, nz = 16, 8, 8
    bcs, loads = problems.cantilever(nx, ny, nz)
    with ta.Problem(nx, ny, nz, bcs, loads) as p:
        out = p.optimize(0.3, 1.5, 4, cg_tol=1e-8)
        c, _, u, _ = p.compliance(out["xphys"], tol=1e-8, want_u=True)
        vm32, du32 = p.fields_f32(1.0)
    grid = ta.GridFrame.make(nx, ny, nz, 1.0)
    ta.write_design_file(tmp_path / "design.bin", grid, [{
        "requested_volume_fraction": 0.3, "achieved_volume_fraction": out["volume_fraction"],
        "accepted": 1, "iterations": out["iterations"], "max_von_mises_mpa": float(vm32.max()),
        "density": out["xphys"]}])
    back = ta.read_design_file(tmp_path / "design.bin")["variants"][0]
    assert np.array_equal(back["density"], out["xphys"])
    assert back["fingerprint"] == ta.design_fingerprint(out["xphys"])
    ta.write_fields_file(tmp_path / "fields.bin", grid, [{
        "requested_volume_fraction": 0.3, "von_mises": vm32, "displacement": du32}])
    b = (tmp_path / "fields.bin").read_bytes()
    off = 64 + 8 + 8 + 8 + 24
    assert np.array_equal(np.frombuffer(b, "<f4", vm32.size, off), vm32)
    assert np.array_equal(np.frombuffer(b, "<f4", du32.size, off + 4 * vm32.size),
                          u.astype(np.float32))


def test_snapshot_capture_cadence_and_eviction(tmp_path):
    """The reference's cadence + cap case (core/tests/unit/test_observability.cpp:
    221-244): every 10 iterations, cap 2, a boundary at 26 kept through
    eviction, the oldest per-iteration snapshot removed at 30."""
    bcs, loads = problems.cantilever(3, 3, 3)
    d = tmp_path / "caps"
    with ta.Problem(3, 3, 3, bcs, loads) as p, ta.SnapshotCapture(d, 10, 2) as cap:
        dens = np.full(27, 0.4)
        got = [cap.capture(p, 0, it, False, density=dens) for it in range(1, 26)]
        assert [i + 1 for i, g in enumerate(got) if g] == [10, 20]
        assert cap.capture(p, 0, 26, True, density=dens)
        assert cap.counts() == (3, 0)
        assert cap.capture(p, 0, 30, False, density=dens)
        assert cap.counts() == (4, 1)
    assert not (d / "snap_rung0_iter0010.f16").exists()
    assert not (d / "snap_rung0_iter0010.json").exists()
    for name in ("snap_rung0_iter0026_boundary", "snap_rung0_iter0020", "snap_rung0_iter0030"):
        assert (d / f"{name}.f16").exists() and (d / f"{name}.json").exists()
    raw = np.fromfile(d / "snap_rung0_iter0030.f16", dtype="<u2")
    assert np.array_equal(raw, np.full(27, np.float16(0.4)).view(np.uint16))
