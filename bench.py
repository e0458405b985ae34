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
  chebyshev_smoother  the same run with the Chebyshev(2) smoother option
               instead of the reference's damped Jacobi
  vs_oracle    per-result errors of the first design iteration against the
               FP64 oracle (compliance, ||u||, PCG count, the OC step's new
               density and active set) and the OC step fed with the oracle's
               own inputs (bit-exact or not), with its decision margin
  ku           the fine K.u kernels timed with HIP events over launches that
               rotate through operand sets larger than twice the last-level
               cache: the plain FP32/FP64 apply and the three variants the
               design loop launches (FP64 A.p with the fused <p,Ap>, the FP32
               V-cycle pre-smoothing residual, the FP32 post-smoothing sweep
               with the fused <r,z>), each with its algorithmic bytes
  roofline     the in-loop FP64 A.p kernel (the largest per-launch K.u cost of
               a PCG iteration) against MI355X's 8 TB/s
  cpu_baseline the FP64 oracle (a C++ restatement of the reference CPU path,
               8-colour threaded apply) timed on this host for the first design
               iteration of the same workload, at 1 thread and at the physical
               cores available to this process
  multi_load_bridge  bridge 256x64x96 with four load cases (three deck point
               loads and the self-weight) solved as one block PCG, summed
               compliance, 80 design iterations (mixed form)
  serving      512 optimisation jobs (grids 48x16x16 .. 192x64x64, cantilever /
               MBB, volfrac U(0.1,0.4), rmin U(1.2,2.5), 60 iterations each)
               arriving Poisson through the continuous-batching server: p50/p99
               job latency and aggregate design-iters/s
  large        cantilever 512x256x256 (101.6 M DOF): design iterations/s, device
               memory held, K.u and smoother bandwidth

Other workloads (--workload): cantilever_60x20x20, bridge_256x64x96 (four load
cases solved as one block PCG, summed compliance), cantilever_512x256x256, and
serving.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def ku_bytes(nx, ny, nz, variant):
    """Algorithmic HBM bytes of one fine K.u launch (BC mask bytes counted as 0)."""
    ndof = 3 * (nx + 1) * (ny + 1) * (nz + 1)
    nv = nx * ny * nz
    w = 8 if variant in ("f64", "ap64_dot", "ap64_pdir") else 4
    if variant == "ap64_pdir":
        b = 2 * w * ndof + w * nv + 12 * ndof       # p_old read, Ap written, E; p written, FP32 z read
    elif variant == "res0_f32":
        b = 4 * w * ndof + w * nv                   # b, D^-1 read; y, x0 written; E read
    elif variant == "jac_dot_f32":
        b = 4 * w * ndof + 8 * ndof + w * nv        # x, b, D^-1 read; y written; FP64 r read; E
    else:
        b = 2 * w * ndof + w * nv                   # x read, y written, E read
    return b, ndof


def measured_traffic(workload, variant, tile):
    """HBM bytes per launch from the committed rocprofv3 PMC summary of this
    kernel variant, workload and launch tile (FETCH_SIZE and WRITE_SIZE, each
    scaled by what the counter reads for a plain copy of known size), or None
    when no summary covers them."""
    grid = workload.split("_", 1)[-1]
    path = os.path.join(ROOT, "profiles", f"ku_{variant}_{grid}_traffic.md")
    if not os.path.exists(path):
        return None
    key = "traffic_bytes_per_launch[" + ",".join(str(int(v)) for v in tile) + "]:"
    for line in open(path):
        if line.startswith(key):
            return float(line.split(":", 1)[1])
    return None


# Design-loop kernel categories that run in FP64 in both forms (the coarsest
# factorisation and solve, the outer PCG's vector work, the A.p apply, the
# SIMP kernels and the hierarchy build): with the mixed form's FP32 parts made
# free, the mixed/FP64 ratio could not exceed (measured ratio) / share.
FP64_IN_BOTH = ("coarsest factorisation", "coarsest solve", "PCG vector kernels",
                "fine K.u FP64", "SIMP", "hierarchy build")


def fp64_invariant_share(workload):
    """Share of the design loop's kernel time that is FP64 in both forms, from
    the committed rocprofv3 kernel summary of the mixed form's design loop."""
    path = os.path.join(ROOT, "profiles", f"{workload}_design_loop.md")
    if not os.path.exists(path):
        return None
    share = 0.0
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 4 and cells[3].replace(".", "", 1).isdigit():
            if any(cells[0].startswith(k) for k in FP64_IN_BOTH):
                share += float(cells[3]) / 100.0
    return share or None


def run_design(ta, wl, precision, steps, warmup, mma=False, smoother=None):
    """One timed optimisation run (OC, or MMA with the plateau stop off so the
    run has exactly `steps` iterations); smoother = (kind, degree) or None for
    the reference's damped Jacobi."""
    nx, ny, nz = wl["grid"]
    ta.reset_mg_stagnation_latch()  # a new run (the latch is per thread and sticky)
    with ta.Problem(nx, ny, nz, wl["bcs"], wl["loads"], precision=precision) as p:
        if smoother:
            p.set_smoother(*smoother)
        kw = dict(penalty=wl["penalty"], move=wl["move"], cg_tol=wl["tol"],
                  continuation=wl.get("continuation"))
        run = (lambda n: p.optimize_mma(wl["volfrac"], wl["rmin"], n, plateau_window=0, **kw)) \
            if mma else (lambda n: p.optimize(wl["volfrac"], wl["rmin"], n, **kw))
        if warmup > 0:
            run(warmup)
        t0 = time.perf_counter()
        out = run(steps)
        dt = time.perf_counter() - t0
        out["oc_gap"] = p.oc_margin()
    return out, dt


KU_LOOP = ("ap64_pdir", "res0_f32", "jac_dot_f32")


def variant_tile(tiles, variant):
    """Launch tile (tx, ty, tz) of a K.u variant: FP64 variants run on the
    FP64 applies' z-chunk, FP32 ones on the FP32 applies'."""
    return tiles["f64"] if variant in ("f64", "ap64_dot", "ap64_pdir") else tiles["f32"]


def ku_measure(ta, nx, ny, nz, reps=100, variants=("f32", "f64", "ap64_dot") + KU_LOOP):
    res = {}
    with ta.Problem(nx, ny, nz) as p:
        p.bench_prepare(0x5EED)
        tx, ty, _ = p.ku_tile()
        tz32, tz64 = p.ku_tiles()
        tiles = {"f32": [tx, ty, tz32], "f64": [tx, ty, tz64]}
        for v in variants:
            ms, sets = p.bench_ku_variant(v, reps)
            b, ndof = ku_bytes(nx, ny, nz, v)
            gbs = b / (ms * 1e-3) / 1e9
            res[v] = {"us": ms * 1e3, "GBs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS, "bytes": b,
                      "gdof_s": ndof / (ms * 1e-3) / 1e9, "rotated_sets": sets}
    return res, tiles


def fields_measure(ta, nx, ny, nz, reps=50):
    """Result-field kernels (von Mises -> f32, displacement -> f32, density ->
    f16) over rotated inputs; algorithmic bytes: u once (8 B per DOF) + 4 B per
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


def serving_jobs(njobs, seed, iterations):
    rng = np.random.default_rng(seed)
    grids = [(48, 16, 16), (96, 32, 32), (128, 64, 32), (192, 64, 64)]
    jobs = []
    for _ in range(njobs):
        g = grids[rng.integers(len(grids))]
        jobs.append(dict(nx=g[0], ny=g[1], nz=g[2], bc_kind=int(rng.integers(2)),
                         volfrac=float(rng.uniform(0.1, 0.4)), rmin=float(rng.uniform(1.2, 2.5)),
                         iterations=iterations, cg_tol=1e-6))
    return jobs, rng.exponential(1.0, njobs)


def run_serving(ta, njobs, rate, seed, iterations, slots=16):
    """Poisson arrivals of jobs with uneven grids through the serving loop."""
    jobs, unit_gaps = serving_jobs(njobs, seed, iterations)
    gaps = unit_gaps / rate
    with ta.Server(budget_bytes=200 << 30, max_slots=slots) as srv:
        ids = []
        t0 = time.perf_counter()
        t_next = t0
        for j, gap in zip(jobs, gaps):
            t_next += gap
            delay = t_next - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            ids.append(srv.submit(j))
            if len(ids) % 64 == 0:
                log(f"# serving: {len(ids)} of {njobs} jobs submitted")
        srv.drain()
        makespan = time.perf_counter() - t0
        res = [srv.poll(i) for i in ids]
    failed = sum(r["state"] != 2 for r in res)
    lat = np.array([r["t_end_ms"] - r["t_submit_ms"] for r in res])
    its = sum(r["iterations"] for r in res)
    return {"jobs": njobs, "failed": failed, "rate_per_s": rate, "seed": seed, "slots": slots,
            "iterations_per_job": iterations,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "design_iters": its, "makespan_s": makespan,
            "design_iters_per_s": its / makespan}


def hip_mem_used():
    """Device bytes in use (hipMemGetInfo), or None without the runtime."""
    try:
        hip = C.CDLL("libamdhip64.so")
        free, total = C.c_size_t(0), C.c_size_t(0)
        if hip.hipMemGetInfo(C.byref(free), C.byref(total)) != 0:
            return None
        return total.value - free.value
    except OSError:
        return None


def run_bridge(ta, problems, warmup=1):
    """Multi-load bridge 256x64x96: the four load cases solved as one block PCG
    (every operator apply and V-cycle one launch over the four cases), summed
    compliance, the configuration's 80 design iterations in the mixed form.
    Late in the run the self-weight case's multigrid exhausts its 300-cycle
    budget (high-contrast thin design under a load on every voxel) and that
    case falls back to Jacobi-CG, as the reference's orchestration does; the
    record counts those design iterations."""
    wl = problems.workload("bridge_256x64x96")
    nx, ny, nz = wl["grid"]
    out, dt = run_design(ta, wl, ta.PREC_MIXED, wl["iterations"], warmup)
    v = out["iterations"] / dt
    pcg = out["history"][:, 3].astype(int)
    with ta.Problem(nx, ny, nz, wl["bcs"], wl["loads"]) as p:
        tz32, tz64 = p.ku_tiles()
    return {"workload": "bridge_256x64x96", "load_cases": len(wl["loads"]),
            "dof": 3 * (nx + 1) * (ny + 1) * (nz + 1), "iterations": out["iterations"],
            "design_iters_per_s": v, "ms_per_iteration": 1e3 * dt / out["iterations"],
            "cg_iterations_total": out["cg_iterations_total"], "compliance": out["compliance"],
            "pcg_iterations_first_10_mean": float(pcg[:10].mean()),
            "design_iterations_past_mg_budget": int((pcg > 300).sum()),
            "ku_tile": {"f32": [32, 8, tz32], "f64": [32, 8, tz64]},
            "note": f"{out['iterations']} design iterations from the uniform design after "
                    f"{warmup} warm-up iteration; volfrac 0.15, rmin 2.0, tol 1e-6; PCG "
                    f"iterations per design iteration are the slowest case's"}


def run_large(ta, problems, iterations=3):
    """Cantilever 512x256x256 (101.6 M DOF): a few design iterations, the
    device memory the problem holds, K.u / smoother bandwidth at that size."""
    wl = problems.workload("cantilever_512x256x256")
    nx, ny, nz = wl["grid"]
    ta.reset_mg_stagnation_latch()
    m0 = hip_mem_used()
    with ta.Problem(nx, ny, nz, wl["bcs"], wl["loads"]) as p:
        out = p.optimize(wl["volfrac"], wl["rmin"], 1, cg_tol=wl["tol"],
                         continuation=wl["continuation"])
        t0 = time.perf_counter()
        out = p.optimize(wl["volfrac"], wl["rmin"], iterations, cg_tol=wl["tol"],
                         continuation=wl["continuation"])
        dt = time.perf_counter() - t0
        m1 = hip_mem_used()
    ku, tiles = ku_measure(ta, nx, ny, nz, reps=10, variants=("f32", "f64", "ap64_dot") + KU_LOOP)
    return {"workload": "cantilever_512x256x256", "dof": 3 * (nx + 1) * (ny + 1) * (nz + 1),
            "iterations": out["iterations"], "design_iters_per_s": out["iterations"] / dt,
            "cg_iterations_total": out["cg_iterations_total"], "compliance": out["compliance"],
            "device_bytes_held": None if m0 is None or m1 is None else m1 - m0,
            "ku_tile": tiles, "ku": ku,
            "note": f"{iterations} iterations from the uniform design (continuation stage p = 1) "
                    f"after one warm-up iteration; the full config runs 150"}


def cpu_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    allowed = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in allowed:
        base = f"/sys/devices/system/cpu/cpu{c}/topology/"
        try:
            cores.add((open(base + "physical_package_id").read().strip(),
                       open(base + "core_id").read().strip()))
        except OSError:
            cores.add(("?", str(c)))
    return {"model": model, "logical_cpus_host": os.cpu_count(), "cpus_allowed": len(allowed),
            "physical_cores_allowed": len(cores)}


def oracle_first_iteration(oracle, s, wl, threads):
    oracle.set_threads(threads)
    nx, ny, nz = wl["grid"]
    x = np.full(nx * ny * nz, wl["volfrac"])
    t0 = time.perf_counter()
    xp = oracle.filter(s, wl["rmin"], x)
    c, dc, u, info = oracle.simp_compliance(s, xp, penalty=wl["penalty"], tol=wl["tol"])
    x1, nb, lam = oracle.oc_update(s, wl["rmin"], x, dc, wl["volfrac"], wl["move"])
    dt = time.perf_counter() - t0
    return dict(dt=dt, x0=x, xp=xp, c=c, dc=dc, u=u, info=info, x1=x1, nb=nb, lam=lam)


def cpu_baseline_and_parity(ta, wl, workload, gpu_history):
    """The oracle's first design iteration on this host's cores (1 thread and
    the physical cores this process may use, at most 16), and the GPU's first
    design iteration compared with it result by result."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    nx, ny, nz = wl["grid"]
    s = oracle.Setup(nx, ny, nz, bcs=wl["bcs"], loads=wl["loads"][0])
    ci = cpu_info()
    nthreads = max(1, min(16, ci["physical_cores_allowed"]))
    ref = oracle_first_iteration(oracle, s, wl, nthreads)
    log(f"# cpu oracle first iteration on {nthreads} threads: {ref['dt']:.1f} s")
    ref1 = oracle_first_iteration(oracle, s, wl, 1)
    log(f"# cpu oracle first iteration on 1 thread: {ref1['dt']:.1f} s")
    cpu = {"value": 1.0 / ref["dt"], "unit": "design iters/s", "cores": nthreads, "kind": "port",
           "value_1_thread": 1.0 / ref1["dt"], "cpu": ci,
           "sample": f"the first design iteration (filter + MG-PCG solve incl. hierarchy build + "
                     f"sensitivity + OC) of {workload} from the uniform design on the FP64 oracle, "
                     f"{ref['info']['iterations']} PCG iterations: {ref['dt']:.1f} s on "
                     f"{nthreads} threads, {ref1['dt']:.1f} s on 1 thread (the run's later "
                     f"iterations take about twice the PCG iterations)"}
    # GPU first iteration, result by result
    with ta.Problem(nx, ny, nz, wl["bcs"], wl["loads"]) as p:
        c, dc, u, info = p.compliance(ref["xp"], penalty=wl["penalty"], tol=wl["tol"], want_u=True)
        x1, act, nb, lam = p.oc_update(wl["rmin"], ref["x0"], dc, wl["volfrac"], wl["move"])
        xo, acto, nbo, lamo = p.oc_update(wl["rmin"], ref["x0"], ref["dc"], wl["volfrac"],
                                          wl["move"])
        gap, _, bound = p.oc_margin()
    lo = np.maximum(1e-3, ref["x0"] - wl["move"])
    hi = np.minimum(1.0, ref["x0"] + wl["move"])
    ref_act = np.where(ref["x1"] == lo, 1, np.where(ref["x1"] == hi, 2, 0))
    par = {
        "iteration": 1,
        "compliance": c, "compliance_oracle": ref["c"],
        "compliance_rel_err": abs(c - ref["c"]) / abs(ref["c"]),
        "loop_history0_compliance": float(gpu_history[0, 0]),
        "loop_history0_rel_err": abs(gpu_history[0, 0] - ref["c"]) / abs(ref["c"]),
        "u_norm_rel_err": abs(np.linalg.norm(u) - np.linalg.norm(ref["u"])) / np.linalg.norm(ref["u"]),
        "u_max_rel_err": float(np.abs(u - ref["u"]).max() / np.abs(ref["u"]).max()),
        "dc_max_rel_err": float(np.abs(dc - ref["dc"]).max() / np.abs(ref["dc"]).max()),
        "pcg_iterations": info["iterations"], "pcg_iterations_oracle": ref["info"]["iterations"],
        "loop_history0_pcg_iterations": int(gpu_history[0, 3]),
        "mg_levels": info["mg_levels"], "mg_levels_oracle": ref["info"]["mg_levels"],
        "density_max_abs_err": float(np.abs(x1 - ref["x1"]).max()),
        "active_set_mismatches": int(np.count_nonzero(act != ref_act)),
        "oc_on_oracle_inputs_bit_exact": bool(np.array_equal(xo, ref["x1"]) and nbo == ref["nb"]
                                              and lamo == ref["lam"]),
        "oc_lambda": lamo, "oc_lambda_oracle": ref["lam"], "oc_bisections": nbo,
        "oc_bisections_oracle": ref["nb"],
        "oc_decision_gap_rel": gap, "oc_rounding_bound_rel": bound,
    }
    return cpu, par


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="mbb_192x64x64")
    ap.add_argument("--no-fp64-form", action="store_true")
    ap.add_argument("--no-mma", action="store_true", help="skip the MMA-updater leg")
    ap.add_argument("--no-chebyshev", action="store_true",
                    help="skip the Chebyshev-smoother leg")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serving", action="store_true", help="skip the 512-job serving leg")
    ap.add_argument("--no-large", action="store_true", help="skip the 512x256x256 leg")
    ap.add_argument("--no-bridge", action="store_true",
                    help="skip the four-load-case bridge 256x64x96 leg")
    ap.add_argument("--jobs", type=int, default=512)
    ap.add_argument("--rate", type=float, default=4.0)
    ap.add_argument("--slots", type=int, default=16, help="serving: concurrent jobs admitted")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--ku-only", action="store_true",
                    help="only the K.u microbenchmark, after a 256 MiB device copy that "
                         "calibrates the PMC byte counters (for rocprofv3 --pmc passes)")
    ap.add_argument("--variant", default=None,
                    help="with --ku-only: one K.u variant (f32, f64, ap64_dot, ap64_pdir, "
                         "res0_f32, jac_dot_f32)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this engine has no multi-GPU path (each job and each "
              "problem runs on a single MI355X); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import topopt_amd as ta
    from topopt_amd import problems

    if a.workload == "serving":
        log(f"# serving: {a.jobs} jobs, Poisson rate {a.rate}/s, seed {a.seed}")
        sv = run_serving(ta, a.jobs, a.rate, a.seed, iterations=60, slots=a.slots)
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
        variants = (a.variant,) if a.variant else ("f32", "f64", "ap64_dot") + KU_LOOP
        ku, tiles = ku_measure(ta, nx, ny, nz, reps=10, variants=variants)
        print(json.dumps({"workload": a.workload, "tile": tiles, "ku": ku,
                          "calibration_copy_bytes": 256 << 20}), flush=True)
        return
    log(f"# workload {a.workload}: warmup {a.warmup}, timed {a.steps} design iterations")
    out, dt = run_design(ta, wl, ta.PREC_MIXED, a.steps, a.warmup)
    value = out["iterations"] / dt
    log(f"# mixed form: {value:.2f} design iters/s, compliance {out['compliance']:.6g}, "
        f"cg iters {out['cg_iterations_total']}")
    fp64 = None
    if not a.no_fp64_form:
        out64, dt64 = run_design(ta, wl, ta.PREC_F64, a.steps, min(a.warmup, 1))
        v64 = out64["iterations"] / dt64
        rel = abs(out["compliance"] - out64["compliance"]) / abs(out64["compliance"])
        share = fp64_invariant_share(a.workload)
        fp64 = {"value": v64, "ratio": value / v64, "compliance": out64["compliance"],
                "compliance_rel_diff_vs_mixed": rel,
                "cg_iterations_total": out64["cg_iterations_total"],
                "fp64_in_both_share": share,
                "ratio_bound_if_fp32_parts_were_free": None if not share else value / v64 / share}
        log(f"# fp64 form: {v64:.2f} design iters/s")
    mma = None
    if not a.no_mma:
        outm, dtm = run_design(ta, wl, ta.PREC_MIXED, a.steps, min(a.warmup, 1), mma=True)
        mma = {"value": outm["iterations"] / dtm, "compliance": outm["compliance"],
               "cg_iterations_total": outm["cg_iterations_total"], "t_oc_ms": outm["t_oc_ms"]}
        log(f"# mma updater: {mma['value']:.2f} design iters/s, compliance "
            f"{outm['compliance']:.6g}")
    cheb = None
    if not a.no_chebyshev:
        outc, dtc = run_design(ta, wl, ta.PREC_MIXED, a.steps, min(a.warmup, 1),
                               smoother=("chebyshev", 2))
        cheb = {"degree": 2, "value": outc["iterations"] / dtc,
                "compliance": outc["compliance"],
                "compliance_rel_diff_vs_jacobi": abs(outc["compliance"] - out["compliance"]) /
                abs(out["compliance"]),
                "cg_iterations_total": outc["cg_iterations_total"]}
        log(f"# chebyshev(2) smoother: {cheb['value']:.2f} design iters/s, "
            f"cg iters {cheb['cg_iterations_total']} (Jacobi {out['cg_iterations_total']})")
    ku, tiles = ku_measure(ta, nx, ny, nz)
    log("# K.u " + ", ".join(f"{k} {v['us']:.1f} us {v['GBs']:.0f} GB/s" for k, v in ku.items())
        + f" (tiles {tiles})")
    fields = fields_measure(ta, nx, ny, nz)
    log("# fields " + ", ".join(f"{k} {v['us']:.1f} us {v['GBs']:.0f} GB/s"
                                for k, v in fields.items()))
    cpu = par = None
    if not a.no_cpu:
        cpu, par = cpu_baseline_and_parity(ta, wl, a.workload, out["history"])
        log(f"# vs oracle: compliance rel {par['compliance_rel_err']:.2e}, ||u|| rel "
            f"{par['u_norm_rel_err']:.2e}, OC on oracle inputs bit-exact "
            f"{par['oc_on_oracle_inputs_bit_exact']}")
    serving = large = bridge = None
    if not a.no_bridge and a.workload == "mbb_192x64x64":
        bridge = run_bridge(ta, problems)
        log(f"# bridge 256x64x96 (4 load cases): {bridge['design_iters_per_s']:.2f} design "
            f"iters/s, cg iters {bridge['cg_iterations_total']}, "
            f"{bridge['design_iterations_past_mg_budget']} design iterations past the MG budget")
    if not a.no_serving and a.workload == "mbb_192x64x64":
        log(f"# serving: {a.jobs} jobs, Poisson rate {a.rate}/s, seed {a.seed}")
        serving = run_serving(ta, a.jobs, a.rate, a.seed, iterations=60, slots=a.slots)
        log(f"# serving: {serving['design_iters_per_s']:.1f} design iters/s, p50 "
            f"{serving['p50_ms']:.0f} ms, p99 {serving['p99_ms']:.0f} ms")
    if not a.no_large and a.workload == "mbb_192x64x64":
        large = run_large(ta, problems)
        log(f"# 512x256x256: {large['design_iters_per_s']:.2f} design iters/s, "
            f"K.u f32 {large['ku']['f32']['us']:.0f} us")
    dom = ku["ap64_pdir"]
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
        "config": {"workload": a.workload, "grid": [nx, ny, nz],
                   "dof": 3 * (nx + 1) * (ny + 1) * (nz + 1), "volfrac": wl["volfrac"],
                   "rmin": wl["rmin"], "penalty": wl["penalty"], "cg_tol": wl["tol"],
                   "load_cases": len(wl["loads"]), "compliance": out["compliance"],
                   "cg_iterations_total": out["cg_iterations_total"],
                   "t_solve_ms": out["t_solve_ms"], "t_build_ms": out["t_build_ms"],
                   "t_oc_ms": out["t_oc_ms"], "t_filter_ms": out["t_filter_ms"],
                   "ku_tile": tiles, "oc_decision_gap_min_rel": out["oc_gap"][1],
                   "oc_rounding_bound_rel": out["oc_gap"][2]},
        "fp64_form": fp64,
        "vs_oracle": par,
        "mma_updater": mma,
        "chebyshev_smoother": cheb,
        "ku": ku,
        "result_fields": fields,
        "roofline": {"bound": "hbm", "achieved": dom["GBs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": dom["hbm_frac"],
                     "traffic": measured_traffic(a.workload, "ap64_pdir", tiles["f64"]),
                     "kernel": "ku_face<double> EPI_AX + <p,Ap> partials with the direction "
                               "update p = z + beta p fused in (the PCG's A.p)",
                     "tile": tiles["f64"], "bytes_per_launch": dom["bytes"],
                     "us_per_launch": dom["us"]},
        "roofline_loop": {k: {"frac": ku[k]["hbm_frac"], "GBs": ku[k]["GBs"], "us": ku[k]["us"],
                              "bytes": ku[k]["bytes"],
                              "traffic": measured_traffic(a.workload, k, variant_tile(tiles, k))}
                          for k in KU_LOOP},
        "cpu_baseline": cpu,
        "multi_load_bridge": bridge,
        "serving": serving,
        "large": large,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
