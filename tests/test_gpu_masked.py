"""The masked (passive-region) design loop on the GPU against the oracle.

The oracle runs simp_optimize with a design mask on the original grid and
derives the analysis grid itself (FrozenVoid removed); the device handle is
built on that analysis grid, as the C ABI requires. Filter and OC are
bit-exact to the oracle on identical inputs, so the design trajectories agree
to the PCG's tolerance, and the pins hold exactly.
"""
import numpy as np
import pytest

import oracle
import topopt_amd as ta
from test_oracle_masked import frozen_box_problem

pytestmark = pytest.mark.gpu

A, FS, FV = oracle.ACTIVE, oracle.FROZEN_SOLID, oracle.FROZEN_VOID


def setup_module():
    oracle.set_threads(0)


def test_masked_filter_bit_exact():
    nx, ny, nz = 13, 7, 9
    rng = np.random.default_rng(21)
    n = nx * ny * nz
    mask = rng.choice([A, A, A, A, FS, FV], size=n).astype(np.uint8)
    solid = (mask != FV).astype(np.uint8)  # the analysis grid
    x = rng.uniform(0, 1, n)
    s = oracle.Setup(nx, ny, nz, solid=solid)
    print("masked filter", (nx, ny, nz), flush=True)
    with ta.Problem(nx, ny, nz, solid=solid) as p:
        for r in (1.2, 1.5, 2.0, 3.0):
            for adj in (False, True):
                assert np.array_equal(p.filter(r, x, adjoint=adj, mask=mask),
                                      oracle.filter(s, r, x, adjoint=adj, mask=mask))


@pytest.mark.parametrize("precision", [ta.PREC_F64, ta.PREC_MIXED])
def test_masked_optimize_matches_oracle(precision):
    nx, ny, nz = 24, 8, 8
    bcs, loads, mask, analysis = frozen_box_problem(nx, ny, nz)
    print("masked optimize", (nx, ny, nz), "precision", precision, flush=True)
    s = oracle.Setup(nx, ny, nz, bcs=bcs, loads=loads)
    ref = oracle.optimize(s, 0.3, 1.5, 8, cg_tol=1e-10, mask=mask)
    with ta.Problem(nx, ny, nz, bcs, loads, solid=analysis, precision=precision) as p:
        out = p.optimize(0.3, 1.5, 8, cg_tol=1e-10, mask=mask)
    hc, hr = out["history"], ref["history"]
    assert out["iterations"] == ref["iterations"] == 8
    tol = 1e-6 if precision == ta.PREC_F64 else 1e-5
    assert np.max(np.abs(hc[:, 0] - hr[:, 0]) / hr[:, 0]) < tol
    assert np.max(np.abs(hc[:, 2] - hr[:, 2])) < 1e-6
    assert abs(out["compliance"] - ref["compliance"]) / ref["compliance"] < tol
    assert abs(out["volume_fraction"] - ref["volume_fraction"]) < 1e-6
    assert np.abs(out["design"] - ref["design"]).max() < 1e-4
    xp, d = out["xphys"], out["design"]
    assert np.all(xp[mask == FS] == 1.0) and np.all(xp[mask == FV] == 0.0)
    assert np.all(d[mask == FS] == 1.0) and np.all(d[mask == FV] == 0.0)


def test_all_active_mask_is_the_plain_loop_on_device():
    nx, ny, nz = 24, 8, 8
    bcs, loads, _, _ = frozen_box_problem(nx, ny, nz)
    with ta.Problem(nx, ny, nz, bcs, loads, precision=ta.PREC_F64) as p:
        plain = p.optimize(0.3, 1.5, 4, cg_tol=1e-8)
        masked = p.optimize(0.3, 1.5, 4, cg_tol=1e-8, mask=np.zeros(nx * ny * nz, np.uint8))
    assert np.array_equal(plain["history"], masked["history"])
    assert np.array_equal(plain["design"], masked["design"])
    assert np.array_equal(plain["xphys"], masked["xphys"])


def test_frozen_void_on_a_solid_voxel_is_rejected():
    nx, ny, nz = 8, 4, 4
    bcs, loads, mask, _ = frozen_box_problem(nx, ny, nz)
    with ta.Problem(nx, ny, nz, bcs, loads) as p:  # full grid, not the analysis grid
        with pytest.raises(ValueError, match="FrozenVoid"):
            p.optimize(0.3, 1.5, 2, mask=mask)


@pytest.mark.parametrize("precision", [ta.PREC_F64, ta.PREC_MIXED])
def test_masked_projected_stages_match_oracle(precision):
    """The masked loop's Heaviside projection stages (simp.cpp:2989-3689 with
    options.projection; oc_update_masked_projected :2801-2855) against the
    oracle: two stages (beta 1 then 4), frozen voxels kept and pinned."""
    nx, ny, nz = 24, 8, 8
    bcs, loads, mask, analysis = frozen_box_problem(nx, ny, nz)
    stages = [(1.0, 0.2, 4), (4.0, 0.1, 4)]
    print("masked projected", (nx, ny, nz), "precision", precision, flush=True)
    s = oracle.Setup(nx, ny, nz, bcs=bcs, loads=loads)
    ref = oracle.optimize_projected(s, 0.3, 1.5, stages, cg_tol=1e-10, mask=mask)
    with ta.Problem(nx, ny, nz, bcs, loads, solid=analysis, precision=precision) as p:
        out = p.optimize_projected(0.3, 1.5, stages, cg_tol=1e-10, mask=mask)
        plain = p.optimize_projected(0.3, 1.5, stages, cg_tol=1e-10,
                                     mask=np.where(mask == FV, FV, A).astype(np.uint8))
    hc, hr = out["history"], ref["history"]
    print("compliance", out["compliance"], ref["compliance"], flush=True)
    assert out["iterations"] == ref["iterations"] == 8
    tol = 1e-6 if precision == ta.PREC_F64 else 1e-5
    assert np.max(np.abs(hc[:, 0] - hr[:, 0]) / hr[:, 0]) < tol
    assert np.max(np.abs(hc[:, 2] - hr[:, 2])) < 1e-6
    assert abs(out["compliance"] - ref["compliance"]) / ref["compliance"] < tol
    assert abs(out["volume_fraction"] - ref["volume_fraction"]) < 1e-6
    assert np.abs(out["design"] - ref["design"]).max() < 1e-4
    xp, d = out["xphys"], out["design"]
    assert np.all(xp[mask == FS] == 1.0) and np.all(xp[mask == FV] == 0.0)
    assert np.all(d[mask == FS] == 1.0) and np.all(d[mask == FV] == 0.0)
    # the FrozenSolid pad is what holds the load: without it the design differs
    assert not np.array_equal(plain["design"], out["design"])
