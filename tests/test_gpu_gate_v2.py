"""Gate V2 on the GPU: the projected OC loop with the reference's beta
continuation schedule (6 stages x 50 iterations) on both benchmark meshes,
checked against every gate of core/tests/validation/test_gate_v2.cpp (the
reference values come from an independent exact-solve computation held in
benchmarks.json), and its first stage against the oracle's trajectory."""
import numpy as np
import pytest

import gate_v2
import oracle
import topopt_amd as ta

pytestmark = pytest.mark.gpu


def setup_module():
    oracle.set_threads(0)


@pytest.mark.parametrize("name", ["cantilever_24x8x8", "clamped_beam_48x8x8"])
def test_gate_v2_on_device(name):
    dims, bcs, loads = gate_v2.problem(name)
    print("gate v2", name, flush=True)
    with ta.Problem(*dims, bcs, loads, precision=ta.PREC_F64) as p:
        out = p.optimize_projected(gate_v2.VOLFRAC, gate_v2.RMIN, gate_v2.SCHEDULE,
                                   cg_tol=gate_v2.CG_TOL)
    ends = gate_v2.check_gates(name, out)
    print(name, "stage ends", np.round(ends, 6), "final", out["compliance"], flush=True)


def test_gate_v2_first_stage_matches_oracle():
    name = "cantilever_24x8x8"
    dims, bcs, loads = gate_v2.problem(name)
    stage = gate_v2.SCHEDULE[:1]
    s = oracle.Setup(*dims, bcs=bcs, loads=loads)
    ref = oracle.optimize_projected(s, gate_v2.VOLFRAC, gate_v2.RMIN, stage, cg_tol=gate_v2.CG_TOL)
    with ta.Problem(*dims, bcs, loads, precision=ta.PREC_F64) as p:
        out = p.optimize_projected(gate_v2.VOLFRAC, gate_v2.RMIN, stage, cg_tol=gate_v2.CG_TOL)
    hc, hr = out["history"], ref["history"]
    assert out["iterations"] == ref["iterations"] == 50
    assert np.max(np.abs(hc[:, 0] - hr[:, 0]) / hr[:, 0]) < 1e-6
    assert np.max(np.abs(hc[:, 2] - hr[:, 2])) < 1e-6
    assert np.abs(out["design"] - ref["design"]).max() < 1e-4
    assert abs(out["compliance"] - ref["compliance"]) / ref["compliance"] < 1e-6
