"""Serving loop on the GPU: each job's result is the same alone and in the mix.

Jobs of uneven grid sizes are submitted together to a server with several
slots, so their design loops overlap on the device; the same jobs are then run
one at a time (max_slots = 1), and each job also runs through the plain
problem handle with the same BCs. Compliance and physical density must be
bit-identical: every reduction has a fixed order and a partial-sum layout that
does not depend on the launch shape or on what else runs on the device.
"""
import numpy as np
import pytest

import topopt_amd as ta
from topopt_amd import problems

pytestmark = pytest.mark.gpu

JOBS = [
    dict(nx=24, ny=8, nz=8, bc_kind=0, volfrac=0.3, rmin=1.5, iterations=6, cg_tol=1e-8),
    dict(nx=32, ny=16, nz=8, bc_kind=1, volfrac=0.25, rmin=1.8, iterations=5, cg_tol=1e-8),
    dict(nx=48, ny=16, nz=16, bc_kind=1, volfrac=0.2, rmin=2.0, iterations=4, cg_tol=1e-8),
    dict(nx=16, ny=8, nz=8, bc_kind=0, volfrac=0.4, rmin=1.2, iterations=7, cg_tol=1e-8),
    dict(nx=40, ny=8, nz=16, bc_kind=0, volfrac=0.15, rmin=2.5, iterations=5, cg_tol=1e-8),
]


def _run(max_slots, one_at_a_time):
    out = []
    with ta.Server(budget_bytes=32 << 30, max_slots=max_slots) as srv:
        ids = []
        for j in JOBS:
            ids.append(srv.submit(j))
            if one_at_a_time:
                srv.drain()
        srv.drain()
        for j, jid in zip(JOBS, ids):
            r = srv.poll(jid)
            assert r["state"] == 2 and r["status"] == 0, r
            assert r["iterations"] == j["iterations"]
            out.append((r, srv.density(jid, j["nx"] * j["ny"] * j["nz"])))
    return out


def test_results_same_alone_and_in_mix():
    print("serving: mix of", len(JOBS), "jobs, then one at a time", flush=True)
    mix = _run(max_slots=len(JOBS), one_at_a_time=False)
    alone = _run(max_slots=1, one_at_a_time=True)
    for (rm, xm), (ra, xa) in zip(mix, alone):
        assert rm["compliance"] == ra["compliance"]
        assert np.array_equal(xm, xa)
    # the mix really overlapped: some job started before another one ended
    spans = sorted((r["t_start_ms"], r["t_end_ms"]) for r, _ in mix)
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    # and the serial run did not
    spans = sorted((r["t_start_ms"], r["t_end_ms"]) for r, _ in alone)
    assert all(spans[i + 1][0] >= spans[i][1] for i in range(len(spans) - 1))


def test_server_job_matches_problem_handle():
    j = JOBS[1]
    bcs, loads = problems.mbb(j["nx"], j["ny"], j["nz"])
    with ta.Problem(j["nx"], j["ny"], j["nz"], bcs, loads) as p:
        ref = p.optimize(j["volfrac"], j["rmin"], j["iterations"], cg_tol=j["cg_tol"])
    with ta.Server(max_slots=2) as srv:
        jid = srv.submit(j)
        srv.drain()
        r = srv.poll(jid)
        x = srv.density(jid, j["nx"] * j["ny"] * j["nz"])
    assert r["compliance"] == ref["compliance"]
    assert np.array_equal(x, ref["xphys"])


def test_admission_respects_slots_and_budget():
    small = dict(nx=16, ny=8, nz=8, bc_kind=0, volfrac=0.3, rmin=1.5, iterations=3)
    fp = ta.Server.footprint(ta.Server.make_job(**small))
    # room for two jobs in the budget, four slots: never more than two run at once
    with ta.Server(budget_bytes=2 * fp + fp // 2, max_slots=4) as srv:
        ids = [srv.submit(small) for _ in range(5)]
        srv.drain()
        res = [srv.poll(i) for i in ids]
    assert all(r["state"] == 2 for r in res)
    events = sorted([(r["t_start_ms"], 1) for r in res] + [(r["t_end_ms"], -1) for r in res])
    live = peak = 0
    for _, d in events:
        live += d
        peak = max(peak, live)
    assert peak <= 2


def test_server_run_records(tmp_path):
    """Per-job run directories: iterations.csv rows per iteration, the CLI's
    PROGRESS / VARIANT lines as the worker parses them
    (tools/topopt-worker/topopt_worker.py:231-266), design.bin holding the
    job's density, fields.bin with the device's von Mises and displacement."""
    jobs = [dict(nx=16, ny=8, nz=8, bc_kind=0, volfrac=0.3, rmin=1.5, iterations=4),
            dict(nx=24, ny=8, nz=8, bc_kind=1, volfrac=0.4, rmin=1.5, iterations=3)]
    with ta.Server(max_slots=2) as srv:
        srv.set_output(tmp_path)
        ids = [srv.submit(j) for j in jobs]
        srv.drain()
        res = [srv.poll(i) for i in ids]
        dens = [srv.density(i, j["nx"] * j["ny"] * j["nz"]) for i, j in zip(ids, jobs)]
    for jid, j, r, x in zip(ids, jobs, res, dens):
        d = tmp_path / f"job_{jid}"
        assert r["state"] == 2
        lines = (d / "iterations.csv").read_text().splitlines()
        assert lines[0] == ta.iterations_csv_header() and len(lines) == 1 + j["iterations"]
        out = (d / "stdout.txt").read_text().splitlines()
        assert out[:-1] == [f"PROGRESS rung=0 rungs=1 iter={k}" for k in range(1, j["iterations"] + 1)]
        kv = dict(t.split("=", 1) for t in out[-1][len("VARIANT "):].split())
        assert out[-1].startswith("VARIANT ") and kv["accepted"] == "1"
        assert float(kv["vf"]) == pytest.approx(j["volfrac"], abs=1e-6)
        assert float(kv["achieved"]) == pytest.approx(r["volume_fraction"], abs=1e-6)
        back = ta.read_design_file(d / "design.bin")
        v = back["variants"][0]
        assert np.array_equal(v["density"], x) and v["iterations"] == j["iterations"]
        b = (d / "fields.bin").read_bytes()
        nv = j["nx"] * j["ny"] * j["nz"]
        nd = 3 * (j["nx"] + 1) * (j["ny"] + 1) * (j["nz"] + 1)
        assert len(b) == 64 + 48 + 4 * nv + 4 * nd
        vm = np.frombuffer(b, "<f4", nv, 64 + 48)
        assert np.all(vm >= 0) and vm.max() == np.float32(v["max_von_mises_mpa"])
        import json
        ri = json.loads((d / "run_info.json").read_text())
        assert ri["solver"] == "Hip" and ri["cg_multigrid"] is True and ri["mg_levels"] >= 2
        assert ri["solved_grid_dofs"] == nd and ri["ladder"] == [pytest.approx(j["volfrac"])]
        assert ri["rung_non_convergent"] == [False] and ri["mixed_precision"] is True
