"""Run records (host formats in libtopopt_hip.so), CPU only.

iterations.csv: the reference's golden bytes (tests/golden/iterations_csv_golden.json,
from core/tests/unit/test_observability.cpp:76-192). fields.bin: the checks of
core/tests/unit/test_fields.cpp, with an independent little-endian reader.
design.bin: an independent reader of the layout in
core/include/topopt/design_store.hpp:46-86, the fingerprint against the
oracle's FNV-1a (itself pinned on the published vectors), and the reader's
refusals (version, counts, fingerprint). PROGRESS / VARIANT: the CLI's printf
lines (core/src/cli/run_job.cpp:10225,10256-10259) and the worker's
key=value parse (tools/topopt-worker/topopt_worker.py:231-266).
"""
import json
import os
import struct

import numpy as np
import pytest

import oracle
import topopt_amd as ta

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "iterations_csv_golden.json")


def test_iterations_csv_golden_bytes(tmp_path):
    g = json.load(open(GOLDEN))
    assert ta.iterations_csv_header() == g["header"]
    path = tmp_path / "iterations.csv"
    with ta.CsvWriter(path) as w:
        for r in g["rows"]:
            w.append_at(r["rung"], ta.IterationObs.make(**r["obs"]), r["wall_ms"])
        assert w.rows() == 2
    assert path.read_bytes() == g["expected"].encode()


def test_iterations_csv_append_stamps_wall_clock(tmp_path):
    import time
    path = tmp_path / "it.csv"
    t0 = int(time.time() * 1000)
    with ta.CsvWriter(path) as w:
        w.append(3, ta.IterationObs.make(iteration=7, compliance=1.0))
    t1 = int(time.time() * 1000)
    lines = path.read_text().splitlines()
    assert len(lines) == 2 and lines[0] == ta.iterations_csv_header()
    f = lines[1].split(",")
    assert len(f) == len(lines[0].split(","))
    assert f[0] == "3" and f[1] == "7" and t0 <= int(f[2]) <= t1


def test_iterations_csv_row_buffer_and_open_errors(tmp_path):
    row = ta.iterations_csv_row(0, ta.IterationObs.make(iteration=1), 5)
    assert row.endswith("\n") and row.startswith("0,1,5,")
    with pytest.raises(RuntimeError):
        ta.CsvWriter(tmp_path / "missing_dir" / "x.csv")


class LE:
    def __init__(self, b):
        self.b, self.at = b, 0

    def get(self, fmt):
        v = struct.unpack_from("<" + fmt, self.b, self.at)
        self.at += struct.calcsize("<" + fmt)
        return v[0] if len(v) == 1 else v


def _fields_variant(nvox, nnode, vf, mass, support, seed):
    i = np.arange(nvox, dtype=np.float32)
    k = np.arange(3 * nnode, dtype=np.float32)
    return {"requested_volume_fraction": vf, "mass_grams": mass, "support_volume_voxels": support,
            "von_mises": np.float32(seed) + i, "displacement": -np.float32(seed) - k * np.float32(0.5)}


def test_fields_bin_layout(tmp_path):
    grid = ta.GridFrame.make(3, 2, 2, 2.5, (1.0, -2.0, 0.5))
    nvox, nnode = 12, 36
    vs = [_fields_variant(nvox, nnode, 0.68, 12.5, 7, 100.0),
          _fields_variant(nvox, nnode, 0.38, 6.25, 0, 300.0)]
    path = tmp_path / "fields.bin"
    ta.write_fields_file(path, grid, vs)
    b = path.read_bytes()
    r = LE(b)
    assert r.get("B") == 1
    r.at += 3
    assert r.get("3i") == (3, 2, 2)
    assert r.get("3d") == (1.0, -2.0, 0.5)
    assert r.get("d") == 2.5 and r.get("d") == 2.5 ** 3
    assert r.get("i") == 2
    r.at += 4
    for v, seed in zip(vs, (100.0, 300.0)):
        assert r.get("d") == v["requested_volume_fraction"] and r.get("d") == v["mass_grams"]
        assert r.get("i") == v["support_volume_voxels"]
        r.at += 4
        assert r.get("3q") == (nvox, 0, 3 * nnode)
        vm = np.frombuffer(b, "<f4", nvox, r.at)
        r.at += 4 * nvox
        du = np.frombuffer(b, "<f4", 3 * nnode, r.at)
        r.at += 12 * nnode
        assert np.array_equal(vm, np.float32(seed) + np.arange(nvox, dtype=np.float32))
        assert np.array_equal(du, -np.float32(seed) - np.arange(3 * nnode, dtype=np.float32) * 0.5)
    assert r.at == len(b)
    assert not os.path.exists(str(path) + ".part")


def test_fields_bin_empty_and_missing_field(tmp_path):
    grid = ta.GridFrame.make(3, 2, 2, 2.5)
    ta.write_fields_file(tmp_path / "e.bin", grid, [])
    b = (tmp_path / "e.bin").read_bytes()
    assert len(b) == 4 + 12 + 40 + 8 and struct.unpack_from("<i", b, 56)[0] == 0
    v = _fields_variant(12, 36, 0.5, 1.0, 1, 1.0)
    v["von_mises"] = None
    with pytest.raises(ValueError):
        ta.write_fields_file(tmp_path / "bad.bin", grid, [v])


def _design(nvox, seed, vf):
    rng = np.random.default_rng(seed)
    return {"requested_volume_fraction": vf, "achieved_volume_fraction": vf - 0.001,
            "margin_worst_case": 1.75, "margin_effective": 1.5, "max_von_mises_mpa": 31.25,
            "accepted": 1, "iterations": 80 + seed, "applied_build_dir": (0.0, -1.0, 0.0),
            "build_direction_auto_applied": 1, "export_baked": 0,
            "density": rng.uniform(0, 1, nvox)}


def test_design_bin_layout_and_round_trip(tmp_path):
    grid = ta.GridFrame.make(4, 3, 2, 1.25, (0.5, 0.0, -3.0))
    nvox = 24
    vs = [_design(nvox, 1, 0.5), _design(nvox, 2, 0.3)]
    path = tmp_path / "design.bin"
    ta.write_design_file(path, grid, vs)
    b = path.read_bytes()
    r = LE(b)
    assert r.get("B") == 1 and b[1:4] == b"\0\0\0"
    r.at += 3
    assert r.get("3i") == (4, 3, 2) and r.get("3d") == (0.5, 0.0, -3.0) and r.get("d") == 1.25
    assert r.get("i") == 2
    r.at += 4
    for v in vs:
        assert r.get("5d") == (v["requested_volume_fraction"], v["achieved_volume_fraction"],
                               1.75, 1.5, 31.25)
        assert r.get("2i") == (1, v["iterations"])
        assert r.get("3d") == (0.0, -1.0, 0.0)
        assert r.get("2i") == (1, 0)
        fp = r.get("Q")
        assert r.get("q") == nvox
        dens = np.frombuffer(b, "<f8", nvox, r.at)
        r.at += 8 * nvox
        assert np.array_equal(dens, v["density"])
        assert fp == oracle.fnv1a64(v["density"].astype("<f8").tobytes())
        assert fp == ta.design_fingerprint(v["density"])
    assert r.at == len(b)
    back = ta.read_design_file(path)
    assert (back["grid"].nx, back["grid"].ny, back["grid"].nz) == (4, 3, 2)
    for v, w in zip(vs, back["variants"]):
        assert np.array_equal(w["density"], v["density"])
        assert w["iterations"] == v["iterations"] and w["accepted"] == 1
        assert w["applied_build_dir"] == (0.0, -1.0, 0.0)


def test_design_bin_reader_refusals(tmp_path):
    grid = ta.GridFrame.make(4, 3, 2, 1.0)
    path = tmp_path / "design.bin"
    ta.write_design_file(path, grid, [_design(24, 3, 0.4)])
    good = path.read_bytes()
    header = 4 + 12 + 32 + 8
    dens_at = header + 5 * 8 + 8 + 24 + 8 + 8 + 8
    cases = {
        "version": bytes([2]) + good[1:],
        "fingerprint": good[:dens_at + 5] + bytes([good[dens_at + 5] ^ 1]) + good[dens_at + 6:],
        "truncated": good[:-3],
        "count": good[:dens_at - 8] + struct.pack("<q", 23) + good[dens_at:],
    }
    for name, blob in cases.items():
        p = tmp_path / f"{name}.bin"
        p.write_bytes(blob)
        with pytest.raises(ValueError):
            ta.read_design_file(p)


def test_density_f16_reader(tmp_path):
    v = (np.arange(24) / 24).astype(np.float16)
    p = tmp_path / "s.f16"
    p.write_bytes(v.tobytes())
    assert np.array_equal(ta.read_density_f16(p, 24), v.astype(np.float32))
    with pytest.raises(ValueError):
        ta.read_density_f16(p, 25)


def _worker_kv(line, prefix):
    assert line.startswith(prefix) and line.endswith("\n")
    return dict(t.split("=", 1) for t in line[len(prefix):].split())


def test_progress_and_variant_lines():
    line = ta.format_progress(2, 4, 137)
    assert line == "PROGRESS rung=2 rungs=4 iter=137\n"
    assert _worker_kv(line, "PROGRESS ") == {"rung": "2", "rungs": "4", "iter": "137"}
    line = ta.format_variant(0.3, 0.2998766, 0.31, 1.23456789, "/out/variant_30.stl")
    assert line == ("VARIANT vf=0.300000 achieved=0.299877 printed=0.310000 margin=1.23457 "
                    "accepted=1 mesh=/out/variant_30.stl\n")
    kv = _worker_kv(line, "VARIANT ")
    assert kv["accepted"] == "1" and float(kv["achieved"]) == 0.299877


def test_library_plateau_stop_matches_reference_curves():
    """The device loop's plateau stop (host code in the library) on the
    reference's curves (tests/golden/mma_plateau_curves.json)."""
    fix = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                      "mma_plateau_curves.json")))
    for name, pre, w, tol, drop, fw, want in fix["cases"]:
        cv = fix["curves"][name][: pre] if pre is not None else fix["curves"][name]
        assert ta.mma_plateau(cv, w, tol, drop, fw) == want, (name, pre, w, drop, fw)
    for name, spec in fix["first_fire"].items():
        cv = fix["curves"][name]
        fire = next(n for n in range(1, len(cv) + 1)
                    if ta.mma_plateau(cv[:n], spec["window"], 1e-3, spec["min_drop"],
                                      spec["flat_windows"]))
        got = oracle.mma_plateau(cv[:fire], spec["window"], 1e-3, spec["min_drop"],
                                 spec["flat_windows"])
        assert got and not oracle.mma_plateau(cv[:fire - 1], spec["window"], 1e-3,
                                               spec["min_drop"], spec["flat_windows"])


REF_FIELDS = ("/root/reference/docs/handoffs/evidence/2026-07-26-constrained-smooth-ui/smooth/"
              "s000/fields.bin")


@pytest.mark.skipif(not os.path.exists(REF_FIELDS), reason="reference checkout not mounted")
def test_reference_written_fields_bin_round_trips_byte_identically(tmp_path):
    """A fields.bin the reference itself wrote (v1, 48x32x48, one variant,
    docs/handoffs/evidence/2026-07-26-constrained-smooth-ui/smooth/s000/):
    parsed with the independent reader above and re-emitted by this library's
    writer, the bytes are identical."""
    b = open(REF_FIELDS, "rb").read()
    r = LE(b)
    assert r.get("B") == 1
    r.at += 3
    nx, ny, nz = r.get("3i")
    origin = r.get("3d")
    spacing = r.get("d")
    assert r.get("d") == spacing ** 3
    nvar = r.get("i")
    r.at += 4
    assert (nx, ny, nz) == (48, 32, 48) and nvar == 1
    nvox, nnode = nx * ny * nz, (nx + 1) * (ny + 1) * (nz + 1)
    vs = []
    for _ in range(nvar):
        vf, mass = r.get("d"), r.get("d")
        support = r.get("i")
        r.at += 4
        assert r.get("3q") == (nvox, 0, 3 * nnode)
        vm = np.frombuffer(b, "<f4", nvox, r.at).copy()
        r.at += 4 * nvox
        du = np.frombuffer(b, "<f4", 3 * nnode, r.at).copy()
        r.at += 12 * nnode
        vs.append({"requested_volume_fraction": vf, "mass_grams": mass,
                   "support_volume_voxels": support, "von_mises": vm, "displacement": du})
    assert r.at == len(b)
    grid = ta.GridFrame.make(nx, ny, nz, spacing, origin)
    out = tmp_path / "fields.bin"
    ta.write_fields_file(out, grid, vs)
    assert out.read_bytes() == b


def test_reference_written_design_bin_reads_and_round_trips(tmp_path):
    """A design.bin the reference wrote (tools/topopt-worker/e2e/.../out/
    design.bin, 4x4x4, two variants; committed as tests/golden/
    design_bin_reference_4x4x4.bin): this library's reader accepts it (its
    fingerprints are the reference's FNV-1a with the reference's offset basis)
    and its writer re-emits the same bytes."""
    src = os.path.join(os.path.dirname(__file__), "golden", "design_bin_reference_4x4x4.bin")
    back = ta.read_design_file(src)
    g = back["grid"]
    assert (g.nx, g.ny, g.nz) == (4, 4, 4) and len(back["variants"]) == 2
    for v in back["variants"]:
        assert v["fingerprint"] == ta.design_fingerprint(v["density"])
        assert v["iterations"] == 30 and v["accepted"] == 1
    out = tmp_path / "design.bin"
    ta.write_design_file(out, g, back["variants"])
    assert out.read_bytes() == open(src, "rb").read()


def test_run_info_json_matches_the_reference_record(tmp_path):
    """run_info_json's keys as the reference's test checks them
    (core/tests/unit/test_observability.cpp:270-311): the fingerprint, the
    solver, the observed multigrid outcome (null until observed), the
    Galerkin cache and thread echoes, the ladder, the timestamp; the file
    writer writes the same text; the record is JSON with the reference's key
    order for the keys it carries."""
    import json
    kw = dict(cli_version="topopt_amd", fingerprint="abc123def456", mode="minimize_plastic",
              material="PLA", source_format="step", resolution=96, load_source="loadcase",
              solver="MultigridCG_Matfree", cg_multigrid_observed=1, cg_multigrid=1,
              mg_levels=4, mixed_precision=1, matfree_threads=6, solved_grid_dofs=2446275,
              rung_non_convergent=[0, 1], rung_non_convergent_iteration=[0, 300],
              rung_non_convergent_residual=[0.0, 3.5e-4],
              ladder=[0.68, 0.52, 0.38, 0.26], ladder_mode="reduction",
              created_wall_ms=1234567890123, iteration_csv=1, snapshot_every=10, snapshot_cap=2)
    js = ta.run_info_json(**kw)
    for frag in ['"fingerprint": "abc123def456"', '"solver": "MultigridCG_Matfree"',
                 '"cg_multigrid": true', '"mg_levels": 4', '"galerkin_block_cache": true',
                 '"matfree_threads": 6', '"ladder": [0.68, 0.52, 0.38, 0.26]',
                 '"created_wall_ms": 1234567890123', '"rung_non_convergent": [false, true]',
                 '"rung_non_convergent_residual": [0, 0.00035]']:
        assert frag in js, frag
    pending = ta.run_info_json(**dict(kw, cg_multigrid_observed=0))
    assert '"cg_multigrid": null' in pending and '"mg_levels": null' in pending
    d = json.loads(js)
    order = ["cli_version", "fingerprint", "mode", "material", "source_format", "resolution",
             "load_source", "solver", "cg_multigrid", "mg_levels", "mg_mode",
             "galerkin_block_cache", "mixed_precision", "matfree_threads", "krylov_recycling",
             "krylov_recycle_dim", "warm_start_inherit", "solved_grid_dofs", "projection",
             "rung_non_convergent", "rung_non_convergent_iteration",
             "rung_non_convergent_residual", "ladder", "ladder_mode", "created_wall_ms",
             "iteration_csv", "density_snapshots", "snapshot_every", "snapshot_cap"]
    assert list(d) == order
    assert js.endswith('  "snapshot_cap": 2\n}\n')
    ta.write_run_info(tmp_path / "run_info.json", **kw)
    assert (tmp_path / "run_info.json").read_text() == js
