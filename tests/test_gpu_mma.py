"""The MMA updater on the device vs the oracle's mma_update (runs on an MI355X).

Single steps: the same designs and sensitivities into both, four successive
steps with one state: new designs to 1e-8 (the device sums the dual's terms in
a fixed block order, so a bisection decision taken at a g within rounding of
zero may differ by one bracket width). The loop: 15 plain MMA iterations in the
FP64 form, compliance history to 1e-6 and design to 1e-5; with the reference's
default plateau stop, the same stopping iteration. MMA refuses projection and
masks, as the reference's validate_updater_options does for projection.
"""
import numpy as np
import pytest

import oracle
import topopt_amd as ta
from topopt_amd import problems

pytestmark = pytest.mark.gpu


def setup_module():
    oracle.set_threads(0)


def test_mma_steps_match_oracle():
    nx, ny, nz, vf, r = 24, 8, 8, 0.3, 1.5
    bcs, loads = problems.cantilever(nx, ny, nz)
    s = oracle.Setup(nx, ny, nz, bcs=bcs, loads=loads)
    x = np.full(nx * ny * nz, vf)
    xs, dcs = [], []
    for k in range(4):
        xp = oracle.filter(s, r, x)
        _, dc, _, _ = oracle.simp_compliance(s, xp, tol=1e-10)
        xs.append(x.copy())
        dcs.append(dc)
        outs, lams = oracle.mma_steps(s, r, np.array(xs), np.array(dcs), vf)
        x = outs[-1]
    with ta.Problem(nx, ny, nz, bcs, loads, precision=ta.PREC_F64) as p:
        for k in range(4):
            xn, lam, ch = p.mma_update(r, k + 1, xs[k], dcs[k], vf)
            d = np.abs(xn - outs[k]).max()
            print("step", k + 1, "lambda", lam, lams[k], "max|dx|", d, flush=True)
            assert d < 1e-8
            assert abs(lam - lams[k]) <= 1e-8 * max(1.0, abs(lams[k]))
            assert ch == pytest.approx(np.abs(outs[k] - xs[k]).max(), abs=1e-8)


def test_mma_loop_matches_oracle_f64():
    nx, ny, nz = 24, 8, 8
    bcs, loads = problems.cantilever(nx, ny, nz)
    s = oracle.Setup(nx, ny, nz, bcs=bcs, loads=loads)
    ref = oracle.optimize_mma(s, 0.3, 1.5, 15, cg_tol=1e-10, plateau_window=0)
    with ta.Problem(nx, ny, nz, bcs, loads, precision=ta.PREC_F64) as p:
        out = p.optimize_mma(0.3, 1.5, 15, cg_tol=1e-10, plateau_window=0)
    hc, hr = out["history"][:, 0], ref["history"][:, 0]
    print("gpu", hc, "\noracle", hr, flush=True)
    assert out["iterations"] == ref["iterations"] == 15
    assert np.max(np.abs(hc - hr) / hr) < 1e-6
    assert np.abs(out["xphys"] - ref["xphys"]).max() < 1e-5
    assert abs(out["compliance"] - ref["compliance"]) / ref["compliance"] < 1e-6


def test_mma_plateau_stop_matches_oracle():
    nx, ny, nz = 16, 6, 6
    bcs, loads = problems.cantilever(nx, ny, nz)
    s = oracle.Setup(nx, ny, nz, bcs=bcs, loads=loads)
    ref = oracle.optimize_mma(s, 0.3, 1.5, 100, cg_tol=1e-10)
    rows = []
    with ta.Problem(nx, ny, nz, bcs, loads, precision=ta.PREC_F64) as p:
        p.set_observer(lambda o: rows.append(o.plateau))
        out = p.optimize_mma(0.3, 1.5, 100, cg_tol=1e-10)
    print("iterations", out["iterations"], ref["iterations"], flush=True)
    assert ref["iterations"] < 100
    assert out["iterations"] == ref["iterations"]
    assert rows == [0] * (len(rows) - 1) + [1] and out["converged"] == 1
    assert abs(out["compliance"] - ref["compliance"]) / ref["compliance"] < 1e-5


def test_mma_mixed_form_tracks_fp64():
    nx, ny, nz = 24, 8, 8
    bcs, loads = problems.cantilever(nx, ny, nz)
    with ta.Problem(nx, ny, nz, bcs, loads, precision=ta.PREC_F64) as p:
        a = p.optimize_mma(0.3, 1.5, 15, cg_tol=1e-8, plateau_window=0)
    with ta.Problem(nx, ny, nz, bcs, loads) as p:
        b = p.optimize_mma(0.3, 1.5, 15, cg_tol=1e-8, plateau_window=0)
    assert np.max(np.abs(a["history"][:, 0] - b["history"][:, 0]) / a["history"][:, 0]) < 1e-5


def test_mma_rejects_bad_inputs():
    bcs, loads = problems.cantilever(8, 4, 4)
    with ta.Problem(8, 4, 4, bcs, loads) as p:
        x = np.full(128, 0.3)
        with pytest.raises(ValueError):
            p.mma_update(1.5, 3, x, -np.ones(128), 0.3)   # no state yet for mma_iter > 2
        with pytest.raises(ValueError):
            p.mma_update(1.5, 1, x, -np.ones(128), 1.5)   # volume fraction out of range


def test_masked_mma_matches_oracle_f64():
    nx, ny, nz = 16, 6, 6
    bcs, loads = problems.cantilever(nx, ny, nz)
    mask = np.zeros((nz, ny, nx), np.uint8)
    mask[:, :, -2:] = oracle.FROZEN_SOLID
    mask[: nz // 2, : ny // 2, nx // 2] = oracle.FROZEN_VOID
    m = mask.ravel()
    solid = (m != oracle.FROZEN_VOID).astype(np.uint8)   # the analysis grid
    s = oracle.Setup(nx, ny, nz, bcs=bcs, loads=loads)
    ref = oracle.optimize_mma(s, 0.3, 1.5, 10, cg_tol=1e-10, plateau_window=0, mask=m)
    with ta.Problem(nx, ny, nz, bcs, loads, solid=solid, precision=ta.PREC_F64) as p:
        out = p.optimize_mma(0.3, 1.5, 10, cg_tol=1e-10, plateau_window=0, mask=m)
    hc, hr = out["history"], ref["history"]
    print("gpu", hc[:, 0], "\noracle", hr[:, 0], flush=True)
    assert out["iterations"] == ref["iterations"] == 10
    assert np.max(np.abs(hc[:, 0] - hr[:, 0]) / hr[:, 0]) < 1e-6
    assert np.max(np.abs(hc[:, 1:3] - hr[:, 1:3])) < 1e-6
    assert np.abs(out["design"] - ref["design"]).max() < 1e-5
    assert np.all(out["xphys"][m == oracle.FROZEN_SOLID] == 1.0)
    assert np.all(out["xphys"][m == oracle.FROZEN_VOID] == 0.0)


def test_masked_mma_all_active_equals_plain_on_device():
    nx, ny, nz = 16, 6, 6
    bcs, loads = problems.cantilever(nx, ny, nz)
    with ta.Problem(nx, ny, nz, bcs, loads) as p:
        a = p.optimize_mma(0.3, 1.5, 8, cg_tol=1e-8, plateau_window=0)
        b = p.optimize_mma(0.3, 1.5, 8, cg_tol=1e-8, plateau_window=0,
                           mask=np.zeros(nx * ny * nz, np.uint8))
    assert np.array_equal(a["history"], b["history"])
    assert np.array_equal(a["design"], b["design"])


def test_mma_projection_continuation_matches_oracle():
    """SimpOptions::mma_projection on the device against the oracle: the
    projected MMA step at each beta, the beta advances on plateaus and the
    stop at the capped beta (core/src/simp/simp.cpp:1509-1636, 2461-2491)."""
    nx, ny, nz = 16, 6, 6
    bcs, loads = problems.cantilever(nx, ny, nz)
    s = oracle.Setup(nx, ny, nz, bcs=bcs, loads=loads)
    ref = oracle.optimize_mma(s, 0.3, 1.5, 200, cg_tol=1e-10, plateau_window=5,
                              projection=(1.0, 8.0))
    betas = []
    with ta.Problem(nx, ny, nz, bcs, loads, precision=ta.PREC_F64) as p:
        p.set_observer(lambda o: betas.append(o.beta))
        out = p.optimize_mma(0.3, 1.5, 200, cg_tol=1e-10, plateau_window=5,
                             projection=(1.0, 8.0))
        with pytest.raises(ValueError):
            p.optimize_mma(0.3, 1.5, 5, projection=(4.0, 2.0))
    hc, hr = out["history"], ref["history"]
    print("iterations", out["iterations"], ref["iterations"], "betas",
          sorted(set(betas)), flush=True)
    assert out["iterations"] == ref["iterations"] < 200
    assert sorted(set(betas)) == [1.0, 2.0, 4.0, 8.0] and betas == sorted(betas)
    assert out["converged"] == 1
    assert np.max(np.abs(hc[:, 0] - hr[:, 0]) / hr[:, 0]) < 1e-5
    assert abs(out["compliance"] - ref["compliance"]) / ref["compliance"] < 1e-5
    assert abs(out["volume_fraction"] - ref["volume_fraction"]) < 1e-6
    assert np.abs(out["xphys"] - ref["xphys"]).max() < 1e-4


def test_masked_mma_projection_matches_oracle():
    nx, ny, nz = 16, 6, 6
    bcs, loads = problems.cantilever(nx, ny, nz)
    mask = np.zeros((nz, ny, nx), np.uint8)
    mask[:, :, -2:] = oracle.FROZEN_SOLID
    mask[: nz // 2, : ny // 2, nx // 2] = oracle.FROZEN_VOID
    m = mask.ravel()
    solid = (m != oracle.FROZEN_VOID).astype(np.uint8)
    s = oracle.Setup(nx, ny, nz, bcs=bcs, loads=loads)
    ref = oracle.optimize_mma(s, 0.3, 1.5, 12, cg_tol=1e-10, plateau_window=0, mask=m,
                              projection=(2.0, 8.0))
    with ta.Problem(nx, ny, nz, bcs, loads, solid=solid, precision=ta.PREC_F64) as p:
        out = p.optimize_mma(0.3, 1.5, 12, cg_tol=1e-10, plateau_window=0, mask=m,
                             projection=(2.0, 8.0))
    hc, hr = out["history"], ref["history"]
    assert out["iterations"] == ref["iterations"] == 12
    assert np.max(np.abs(hc[:, 0] - hr[:, 0]) / hr[:, 0]) < 1e-6
    assert np.max(np.abs(hc[:, 1:3] - hr[:, 1:3])) < 1e-6
    assert np.abs(out["design"] - ref["design"]).max() < 1e-5
    assert np.all(out["xphys"][m == oracle.FROZEN_SOLID] == 1.0)
    assert np.all(out["xphys"][m == oracle.FROZEN_VOID] == 0.0)


def test_adaptive_move_matches_oracle():
    """SimpOptions::adaptive_move on the device (oscillation counts over the
    design set, the seed / shrink / grow / clamp rule) against the oracle."""
    nx, ny, nz = 16, 6, 6
    bcs, loads = problems.cantilever(nx, ny, nz)
    ad = (1.2, 0.7, 0.15, 0.30, 0.5, 0.02)
    s = oracle.Setup(nx, ny, nz, bcs=bcs, loads=loads)
    ref = oracle.optimize_mma(s, 0.3, 1.5, 15, cg_tol=1e-10, plateau_window=0, adaptive=ad)
    with ta.Problem(nx, ny, nz, bcs, loads, precision=ta.PREC_F64) as p:
        out = p.optimize_mma(0.3, 1.5, 15, cg_tol=1e-10, plateau_window=0, adaptive=ad)
        with pytest.raises(ValueError):
            p.optimize_mma(0.3, 1.5, 3, adaptive=(1.2, 0.7, 0.15, 0.30, 1.0, 0.02))
    hc, hr = out["history"], ref["history"]
    print("change gpu", hc[:, 1], "\noracle", hr[:, 1], flush=True)
    assert out["iterations"] == ref["iterations"] == 15
    assert np.max(np.abs(hc[:, 0] - hr[:, 0]) / hr[:, 0]) < 1e-6
    assert np.max(np.abs(hc[:, 1:3] - hr[:, 1:3])) < 1e-6
    assert np.abs(out["design"] - ref["design"]).max() < 1e-5
