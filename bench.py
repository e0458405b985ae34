"""Benchmark: voxel-SIMP design iterations/s on one MI355X.

Workload (default): MBB half-beam 192x64x64 hex8 (2.45 M DOF), volfrac 0.2,
rmin 2.0, p = 3, OC move 0.2, MG-PCG tol 1e-6, FP32 4-level V-cycle (FP64
coarsest) preconditioning an FP64 PCG (FP64 A.p). A step is one design iteration
(filter -> hierarchy rebuild + solve -> sensitivity -> OC update); the timed
region is one optimisation run of --steps iterations from the uniform design,
final certification solve included. Inputs are generated on the host before
the timed region and live in HBM throughout.

Also reported on the same JSON line:
  fp64_form    the same run in the plain FP64 GPU form (every apply and the
               whole V-cycle in FP64)
  ku           the fine K.u apply (FP32 and FP64) timed with HIP events over
               back-to-back launches on the problem's stream: GDOF/s and HBM
               fraction of the algorithmic bytes
  roofline     that FP32 K.u kernel against MI355X's 8 TB/s
  cpu_baseline the FP64 oracle (a C++ restatement of the reference CPU path,
               8-colour threaded apply) timed on this host for one design
               iteration of the same workload

Other workloads (--workload): cantilever_60x20x20, bridge_256x64x96 (four load
cases solved as one block PCG, summed compliance), cantilever_512x256x256, and
serving (a Poisson stream of optimisation jobs through the continuous-batching
server: p50/p99 job latency and aggregate design-iters/s; --jobs, --rate).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def measured_traffic(workload):
    """HBM bytes per fine K.u launch from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE, each scaled by what the counter reads for a
    plain copy of known size), or None when no summary covers the workload."""
    path = os.path.join(ROOT, "profiles", f"ku_fine_f32_{workload.split('_', 1)[-1]}_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def ku_bytes(nx, ny, nz, word):
    ndof = 3 * (nx + 1) * (ny + 1) * (nz + 1)
    return 2 * word * ndof + word * nx * ny * nz, ndof


def run_design(ta, wl, precision, steps, warmup, mma=False):
    """One timed optimisation run (OC, or MMA with the plateau stop off so the
    run has exactly `steps` iterations)."""
    nx, ny, nz = wl["grid"]
    with ta.Problem(nx, ny, nz, wl["bcs"], wl["loads"], precision=precision) as p:
        kw = dict(penalty=wl["penalty"], move=wl["move"], cg_tol=wl["tol"],
                  continuation=wl.get("continuation"))
        run = (lambda n: p.optimize_mma(wl["volfrac"], wl["rmin"], n, plateau_window=0, **kw)) \
            if mma else (lambda n: p.optimize(wl["volfrac"], wl["rmin"], n, **kw))
        if warmup > 0:
            run(warmup)
        t0 = time.perf_counter()
        out = run(steps)
        dt = time.perf_counter() - t0
    return out, dt


def ku_measure(ta, nx, ny, nz, reps=100):
    with ta.Problem(nx, ny, nz) as p:
        p.bench_prepare(0x5EED)
        ms32 = p.bench_ku(ta.PREC_MIXED, reps)
        ms64 = p.bench_ku(ta.PREC_F64, reps)
    b32, ndof = ku_bytes(nx, ny, nz, 4)
    b64, _ = ku_bytes(nx, ny, nz, 8)
    res = {}
    for name, ms, b in (("f32", ms32, b32), ("f64", ms64, b64)):
        gbs = b / (ms * 1e-3) / 1e9
        res[name] = {"ms": ms, "gdof_s": ndof / (ms * 1e-3) / 1e9, "GBs": gbs,
                     "hbm_frac": gbs / HBM_PEAK_GBS, "bytes": b,
                     "gflops": 1152.0 * nx * ny * nz / (ms * 1e-3) / 1e9}
    return res


def fields_measure(ta, nx, ny, nz, reps=50):
    """Result-field kernels (von Mises -> f32, displacement -> f32, density ->
    f16) on resident inputs; algorithmic bytes: u once (8 B per DOF) + 4 B per
    voxel out; 8 + 4 B per DOF; 8 + 2 B per voxel."""
    with ta.Problem(nx, ny, nz) as p:
        p.bench_prepare(0x5EED)
        ms = p.bench_fields(reps)
    nv = nx * ny * nz
    nd = 3 * (nx + 1) * (ny + 1) * (nz + 1)
    out = {}
    for name, t, b in (("von_mises_f32", ms[0], 8 * nd + 4 * nv), ("narrow_f32", ms[1], 12 * nd),
                       ("density_f16", ms[2], 10 * nv)):
        gbs = b / (t * 1e-3) / 1e9
        out[name] = {"us": t * 1e3, "bytes": b, "GBs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS}
    return out


def run_serving(ta, njobs, rate, seed, iterations):
    """Poisson arrivals of jobs with uneven grids through the serving loop."""
    rng = np.random.default_rng(seed)
    grids = [(48, 16, 16), (96, 32, 32), (128, 64, 32), (192, 64, 64)]
    jobs = []
    for _ in range(njobs):
        g = grids[rng.integers(len(grids))]
        jobs.append(dict(nx=g[0], ny=g[1], nz=g[2], bc_kind=int(rng.integers(2)),
                         volfrac=float(rng.uniform(0.1, 0.4)), rmin=float(rng.uniform(1.2, 2.5)),
                         iterations=iterations, cg_tol=1e-6))
    gaps = rng.exponential(1.0 / rate, njobs)
    with ta.Server(budget_bytes=200 << 30, max_slots=16) as srv:
        ids = []
        t0 = time.perf_counter()
        t_next = t0
        for j, gap in zip(jobs, gaps):
            t_next += gap
            delay = t_next - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            ids.append(srv.submit(j))
        srv.drain()
        makespan = time.perf_counter() - t0
        res = [srv.poll(i) for i in ids]
    failed = sum(r["state"] != 2 for r in res)
    lat = np.array([r["t_end_ms"] - r["t_submit_ms"] for r in res])
    its = sum(r["iterations"] for r in res)
    return {"jobs": njobs, "failed": failed, "rate_per_s": rate, "seed": seed,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "design_iters": its, "makespan_s": makespan,
            "design_iters_per_s": its / makespan}


def cpu_baseline(wl, threads):
    """One design iteration of the workload on the FP64 oracle (bounded sample)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.set_threads(threads)
    nx, ny, nz = wl["grid"]
    s = oracle.Setup(nx, ny, nz, bcs=wl["bcs"], loads=wl["loads"][0])
    x = np.full(nx * ny * nz, wl["volfrac"])
    t0 = time.perf_counter()
    xp = oracle.filter(s, wl["rmin"], x)
    c, dc, _, info = oracle.simp_compliance(s, xp, penalty=wl["penalty"], tol=wl["tol"])
    oracle.oc_update(s, wl["rmin"], x, dc, wl["volfrac"], wl["move"])
    dt = time.perf_counter() - t0
    return {"value": 1.0 / dt, "unit": "design iters/s", "cores": threads, "kind": "port",
            "sample": f"1 design iteration (filter + MG-PCG solve incl. hierarchy build + "
                      f"sensitivity + OC) of {nx}x{ny}x{nz} from the uniform design, FP64 oracle, "
                      f"{info['iterations']} PCG iterations, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="mbb_192x64x64")
    ap.add_argument("--no-fp64-form", action="store_true")
    ap.add_argument("--no-mma", action="store_true", help="skip the MMA-updater leg")

    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--jobs", type=int, default=128)
    ap.add_argument("--rate", type=float, default=4.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--ku-only", action="store_true",
                    help="only the K.u microbenchmark, after a 256 MiB device copy that "
                         "calibrates the PMC byte counters (for rocprofv3 --pmc passes)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this engine has no multi-GPU path (each job and each "
              "problem runs on a single MI355X); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import topopt_amd as ta
    from topopt_amd import problems

    if a.workload == "serving":
        print(f"# serving: {a.jobs} jobs, Poisson rate {a.rate}/s, seed {a.seed}", file=sys.stderr,
              flush=True)
        sv = run_serving(ta, a.jobs, a.rate, a.seed, iterations=60)
        print(json.dumps({
            "metric": "serving: aggregate design iterations/s (job p50/p99 latency in config)",
            "value": sv["design_iters_per_s"], "unit": "design iters/s", "n_gpus": 1,
            "steps": sv["design_iters"], "warmup": 0,
            "ms_per_step": 1e3 * sv["makespan_s"] / max(1, sv["design_iters"]),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32 V-cycle, f64 operator and PCG",
            "data": "synthetic (Poisson job stream, cantilever/MBB boxes)",
            "config": {"workload": "serving", **sv}}), flush=True)
        return

    wl = problems.workload(a.workload)
    nx, ny, nz = wl["grid"]
    if a.ku_only:
        import torch  # device buffers for the calibration copy only
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        ku = ku_measure(ta, nx, ny, nz, reps=10)
        print(json.dumps({"workload": a.workload, "ku": ku, "calibration_copy_bytes": 256 << 20}),
              flush=True)
        return
    print(f"# workload {a.workload}: warmup {a.warmup}, timed {a.steps} design iterations",
          file=sys.stderr, flush=True)
    out, dt = run_design(ta, wl, ta.PREC_MIXED, a.steps, a.warmup)
    value = out["iterations"] / dt
    print(f"# mixed form: {value:.2f} design iters/s, compliance {out['compliance']:.6g}, "
          f"cg iters {out['cg_iterations_total']}", file=sys.stderr, flush=True)
    fp64 = None
    if not a.no_fp64_form:
        out64, dt64 = run_design(ta, wl, ta.PREC_F64, a.steps, min(a.warmup, 1))
        v64 = out64["iterations"] / dt64
        rel = abs(out["compliance"] - out64["compliance"]) / abs(out64["compliance"])
        fp64 = {"value": v64, "ratio": value / v64, "compliance": out64["compliance"],
                "compliance_rel_diff_vs_mixed": rel, "cg_iterations_total": out64["cg_iterations_total"]}
        print(f"# fp64 form: {v64:.2f} design iters/s", file=sys.stderr, flush=True)
    mma = None
    if not a.no_mma:
        outm, dtm = run_design(ta, wl, ta.PREC_MIXED, a.steps, min(a.warmup, 1), mma=True)
        mma = {"value": outm["iterations"] / dtm, "compliance": outm["compliance"],
               "cg_iterations_total": outm["cg_iterations_total"], "t_oc_ms": outm["t_oc_ms"]}
        print(f"# mma updater: {mma['value']:.2f} design iters/s, compliance "
              f"{outm['compliance']:.6g}", file=sys.stderr, flush=True)
    ku = ku_measure(ta, nx, ny, nz)
    fields = fields_measure(ta, nx, ny, nz)
    print("# fields " + ", ".join(f"{k} {v['us']:.1f} us {v['GBs']:.0f} GB/s"
                                  for k, v in fields.items()), file=sys.stderr, flush=True)
    print(f"# K.u f32 {ku['f32']['ms']*1e3:.1f} us {ku['f32']['GBs']:.0f} GB/s; "
          f"f64 {ku['f64']['ms']*1e3:.1f} us", file=sys.stderr, flush=True)
    cpu = None
    if not a.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(wl, threads)
        print(f"# cpu baseline {cpu['value']:.4f} design iters/s on {threads} threads",
              file=sys.stderr, flush=True)
    line = {
        "metric": f"design iterations/s ({a.workload}, FP32 multigrid V-cycle in FP64 PCG)",
        "value": value,
        "unit": "design iters/s",
        "n_gpus": 1,
        "steps": out["iterations"],
        "warmup": a.warmup,
        "ms_per_step": dt * 1e3 / max(1, out["iterations"]),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 V-cycle, f64 operator and PCG",
        "data": "synthetic (analytic box workload, uniform initial density)",
        "config": {"workload": a.workload, "grid": [nx, ny, nz], "dof": ku["f32"]["bytes"] and
                   3 * (nx + 1) * (ny + 1) * (nz + 1), "volfrac": wl["volfrac"], "rmin": wl["rmin"],
                   "penalty": wl["penalty"], "cg_tol": wl["tol"], "load_cases": len(wl["loads"]),
                   "compliance": out["compliance"], "cg_iterations_total": out["cg_iterations_total"],
                   "t_solve_ms": out["t_solve_ms"], "t_build_ms": out["t_build_ms"],
                   "t_oc_ms": out["t_oc_ms"], "t_filter_ms": out["t_filter_ms"]},
        "fp64_form": fp64,
        "mma_updater": mma,
        "ku": ku,
        "result_fields": fields,
        "roofline": {"bound": "hbm", "achieved": ku["f32"]["GBs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": ku["f32"]["hbm_frac"],
                     "traffic": measured_traffic(a.workload),
                     "kernel": "ku_face<float> fine modal apply (autotuned tile)",
                     "bytes_per_launch": ku["f32"]["bytes"], "us_per_launch": ku["f32"]["ms"] * 1e3},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
