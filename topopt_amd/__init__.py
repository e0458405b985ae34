"""topopt_amd — MI355X (gfx950) engine for the voxel-SIMP inner loop.

Python mirror of the reference core's solver interface for this path, over the
C ABI of libtopopt_hip.so (include/topopt_hip.h). Names, argument meaning and
error behaviour follow the reference:

  fea_matfree_apply        core/include/topopt/fea.hpp:763-766
  fea_solve_mgcg_matfree   core/include/topopt/fea.hpp:707-715
  simp_compliance          core/include/topopt/simp.hpp:247-255
  DensityFilter            core/include/topopt/simp.hpp:387-406
  oc_update                core/include/topopt/simp.hpp:481-485
  simp_optimize            core/include/topopt/simp.hpp:1783

Exceptions map the C status codes onto the reference taxonomy:
ValueError (std::invalid_argument), UnderConstrained (std::runtime_error from
the void gate), SolverNonConvergence (with .iterations / .residual),
DeviceError (a HIP failure). The library must be built (topopt_amd.build);
nothing here falls back to a CPU path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtopopt_hip.so")

PREC_MIXED = 0
PREC_F64 = 1

OK, EINVAL, EUNDERCONSTRAINED, ENONCONVERGENCE, EDEADLINE, EDEVICE, EINTERNAL = range(7)


class UnderConstrained(RuntimeError):
    pass


class SolverNonConvergence(RuntimeError):
    def __init__(self, msg, iterations=0, residual=0.0):
        super().__init__(msg)
        self.iterations = iterations
        self.residual = residual


class SolverDeadlineExceeded(SolverNonConvergence):
    """An armed solve deadline passed before CG reached its tolerance."""


class DeviceError(RuntimeError):
    pass


class CgInfo(C.Structure):
    _fields_ = [
        ("converged", C.c_int), ("iterations", C.c_int), ("used_multigrid", C.c_int),
        ("mg_levels", C.c_int), ("hier_built", C.c_int), ("mg_cycles_attempted", C.c_int),
        ("residual", C.c_double), ("matvecs", C.c_longlong),
        ("t_build_ms", C.c_double), ("t_mg_build_ms", C.c_double), ("t_mg_ms", C.c_double),
        ("t_cg_ms", C.c_double), ("t_total_ms", C.c_double),
        ("residual_replacements", C.c_int), ("coarse_dofs", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Options(C.Structure):
    _fields_ = [
        ("volfrac", C.c_double), ("rmin", C.c_double), ("move", C.c_double),
        ("penalty", C.c_double), ("E0", C.c_double), ("rho_min", C.c_double),
        ("change_tol", C.c_double), ("cg_tol", C.c_double),
        ("max_iterations", C.c_int), ("cg_max_iterations", C.c_int),
        ("pc_p0", C.c_double), ("pc_p1", C.c_double), ("pc_step", C.c_double),
        ("pc_per_stage", C.c_int),
    ]


class MmaOptions(C.Structure):
    _fields_ = [("plateau_window", C.c_int), ("plateau_tol", C.c_double),
                ("plateau_min_drop", C.c_double), ("plateau_flat_windows", C.c_int),
                ("adaptive_move", C.c_int), ("adaptive_move_grow", C.c_double),
                ("adaptive_move_shrink", C.c_double), ("adaptive_move_osc_lo", C.c_double),
                ("adaptive_move_osc_hi", C.c_double), ("adaptive_move_max", C.c_double),
                ("adaptive_move_min", C.c_double)]


def mma_plateau(c, window=10, rel_tol=1e-3, min_drop=0.05, flat_windows=3):
    """mma_objective_plateau on a compliance history."""
    cc = np.ascontiguousarray(c, dtype=np.float64)
    out = C.c_int(0)
    _check(lib().topohip_mma_plateau(_p(cc), cc.size, int(window), C.c_double(rel_tol),
                                     C.c_double(min_drop), int(flat_windows), C.byref(out)))
    return bool(out.value)


class Result(C.Structure):
    _fields_ = [
        ("iterations", C.c_int), ("converged", C.c_int), ("non_convergent", C.c_int),
        ("compliance", C.c_double), ("initial_compliance", C.c_double),
        ("volume_fraction", C.c_double),
        ("t_total_ms", C.c_double), ("t_solve_ms", C.c_double), ("t_build_ms", C.c_double),
        ("t_filter_ms", C.c_double), ("t_sens_ms", C.c_double), ("t_oc_ms", C.c_double),
        ("cg_iterations_total", C.c_longlong),
        ("cancelled", C.c_int), ("deadline_exceeded", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Job(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("bc_kind", C.c_int),
        ("volfrac", C.c_double), ("rmin", C.c_double), ("penalty", C.c_double),
        ("move", C.c_double), ("cg_tol", C.c_double), ("iterations", C.c_int),
        ("precision", C.c_int),
    ]


class JobResult(C.Structure):
    _fields_ = [
        ("state", C.c_int), ("iterations", C.c_int), ("compliance", C.c_double),
        ("volume_fraction", C.c_double), ("t_submit_ms", C.c_double),
        ("t_start_ms", C.c_double), ("t_end_ms", C.c_double), ("status", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Load the in-tree HIP library; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DeviceError(f"{LIB_PATH} is missing: build it with `python -m topopt_amd.build`")
        _lib = C.CDLL(LIB_PATH)
        _lib.topohip_version.restype = C.c_char_p
        _lib.topohip_job_footprint.restype = C.c_size_t
        _lib.topohip_job_footprint.argtypes = [C.POINTER(Job)]
        _lib.topohip_server_create.argtypes = [C.c_size_t, C.c_int, C.POINTER(C.c_void_p)]
    return _lib


def _err():
    buf = C.create_string_buffer(1024)
    lib().topohip_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def _check(code, info=None):
    if code == OK:
        return
    msg = _err()
    if code == EINVAL:
        raise ValueError(msg)
    if code == EUNDERCONSTRAINED:
        raise UnderConstrained(msg)
    if code == ENONCONVERGENCE:
        it = info.iterations if info is not None else 0
        res = info.residual if info is not None else 0.0
        raise SolverNonConvergence(msg, it, res)
    if code == EDEADLINE:
        it = info.iterations if info is not None else 0
        res = info.residual if info is not None else 0.0
        raise SolverDeadlineExceeded(msg, it, res)
    if code == EDEVICE:
        raise DeviceError(msg)
    raise RuntimeError(f"topohip status {code}: {msg}")


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


class Problem:
    """Device-resident problem: grid, material, Dirichlet BCs and load cases.

    bcs / load cases are sequences of (node, component, value) triples, the
    reference's DirichletBC / NodalLoad.
    """

    def __init__(self, nx, ny, nz, bcs=(), loads=None, h=1.0, poisson=0.3, solid=None,
                 precision=PREC_MIXED):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.nvox = self.nx * self.ny * self.nz
        self.ndof = 3 * (self.nx + 1) * (self.ny + 1) * (self.nz + 1)
        self.precision = precision
        bcs = list(bcs)
        bn = np.array([b[0] for b in bcs], dtype=np.int32)
        bc = np.array([b[1] for b in bcs], dtype=np.int32)
        bv = np.array([b[2] for b in bcs], dtype=np.float64)
        sol = None if solid is None else np.ascontiguousarray(solid, dtype=np.uint8).ravel()
        if sol is not None and sol.size != self.nvox:
            raise ValueError("solid tag size != voxel_count")
        self._h = C.c_void_p()
        _check(lib().topohip_problem_create(
            self.nx, self.ny, self.nz, C.c_double(h), C.c_double(poisson), _p(sol), len(bcs),
            _p(bn), _p(bc), _p(bv), int(precision), C.byref(self._h)))
        self.nrhs = 0
        if loads is not None:
            self.set_loads(loads)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().topohip_problem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_loads(self, cases):
        """cases: one list of (node, comp, value) triples, or a list of them."""
        cases = list(cases)
        if cases and isinstance(cases[0], tuple):
            cases = [cases]
        counts = np.array([len(c) for c in cases], dtype=np.int32)
        flat = [t for c in cases for t in c]
        n = np.array([t[0] for t in flat], dtype=np.int32)
        cc = np.array([t[1] for t in flat], dtype=np.int32)
        v = np.array([t[2] for t in flat], dtype=np.float64)
        _check(lib().topohip_set_loads(self._h, len(cases), _p(counts), _p(n), _p(cc), _p(v)))
        self.nrhs = len(cases)

    # -- operators ----------------------------------------------------------
    def apply(self, youngs, u, precision="f32"):
        E = _f64(youngs, self.nvox)
        u = _f64(u, self.ndof)
        y = np.zeros(self.ndof)
        fn = lib().topohip_ku_apply_f32 if precision == "f32" else lib().topohip_ku_apply_f64
        _check(fn(self._h, _p(E), _p(u), _p(y)))
        return y

    def solve(self, youngs, tol=1e-8, max_iterations=0, initial_guess=None):
        E = _f64(youngs, self.nvox)
        g = None if initial_guess is None else _f64(initial_guess)
        if g is not None and g.size != self.ndof:
            g = None  # wrong length: ignored, as the reference's initial_guess
        u = np.zeros(self.ndof)
        info = CgInfo()
        code = lib().topohip_solve(self._h, _p(E), C.c_double(tol), int(max_iterations), _p(g),
                                   _p(u), C.byref(info))
        _check(code, info)
        return u, info.as_dict()

    def compliance(self, xphys, E0=1.0, penalty=3.0, rho_min=1e-3, tol=1e-8, max_iterations=0,
                   want_u=False, guess=None):
        """simp_compliance; guess (full displacement) warm-starts the solve."""
        x = _f64(xphys, self.nvox)
        c = C.c_double(0)
        dc = np.zeros(self.nvox)
        u = np.zeros(self.ndof) if want_u else None
        info = CgInfo()
        if guess is None:
            code = lib().topohip_simp_compliance(self._h, _p(x), C.c_double(E0),
                                                 C.c_double(penalty), C.c_double(rho_min),
                                                 C.c_double(tol), int(max_iterations), C.byref(c),
                                                 _p(dc), _p(u), C.byref(info))
        else:
            g = _f64(guess, self.ndof)
            code = lib().topohip_simp_compliance_warm(self._h, _p(x), C.c_double(E0),
                                                      C.c_double(penalty), C.c_double(rho_min),
                                                      C.c_double(tol), int(max_iterations),
                                                      _p(g), C.byref(c), _p(dc), _p(u),
                                                      C.byref(info))
        _check(code, info)
        return c.value, dc, u, info.as_dict()

    def von_mises(self, youngs=1.0, u=None):
        """fea_von_mises_field: centroid von Mises per solid voxel (0 on empty
        voxels) of u (default: the displacement of the last solve)."""
        uu = None if u is None else _f64(u, self.ndof)
        out = np.zeros(self.nvox)
        _check(lib().topohip_von_mises_field(self._h, C.c_double(youngs), _p(uu), _p(out)))
        return out

    def fields_f32(self, youngs=1.0, u=None):
        """fields.bin's float32 arrays: (von Mises, displacement)."""
        uu = None if u is None else _f64(u, self.ndof)
        vm = np.zeros(self.nvox, dtype=np.float32)
        du = np.zeros(self.ndof, dtype=np.float32)
        _check(lib().topohip_fields_f32(self._h, C.c_double(youngs), _p(uu), _p(vm), _p(du)))
        return vm, du

    def density_f16(self, density=None):
        """Binary16 bits (uint16) of float(density) per voxel (default: the
        physical density of the last optimisation)."""
        d = None if density is None else _f64(density, self.nvox)
        out = np.zeros(self.nvox, dtype=np.uint16)
        _check(lib().topohip_density_f16(self._h, _p(d), _p(out)))
        return out

    def write_density_snapshot(self, directory, basename, rung, iteration, boundary=False,
                               density=None, origin=(0.0, 0.0, 0.0)):
        d = None if density is None else _f64(density, self.nvox)
        o = np.ascontiguousarray(origin, dtype=np.float64)
        _check(lib().topohip_write_density_snapshot(self._h, str(directory).encode(),
                                                    str(basename).encode(), _p(o), _p(d),
                                                    C.c_size_t(rung), int(iteration),
                                                    int(bool(boundary))))

    def set_smoother(self, kind="jacobi", degree=1):
        """V-cycle smoother of later solves: "jacobi" (damped, omega 0.6; the
        reference's) or "chebyshev" of the given degree."""
        _check(lib().topohip_set_smoother(self._h, {"jacobi": 0, "chebyshev": 1}[kind],
                                          int(degree)))

    def smoother_lambdas(self):
        lam = np.zeros(16)
        n = C.c_int(0)
        _check(lib().topohip_smoother_lambdas(self._h, 16, _p(lam), C.byref(n)))
        return lam[: n.value].copy()

    def set_cancel_flag(self, flag):
        """flag: a ctypes.c_int the caller keeps (nonzero cancels the design
        loop at the start of its next iteration), or None to clear."""
        self._cancel = flag
        _check(lib().topohip_set_cancel_flag(self._h, None if flag is None else C.byref(flag)))

    def set_observer(self, fn):
        """fn(IterationObs) after every design iteration (None clears)."""
        from .records import ITERATION_CB
        if fn is None:
            self._observer = None
            _check(lib().topohip_set_iteration_observer(self._h, None, None))
            return
        cb = ITERATION_CB(lambda o, _u: fn(o.contents))
        self._observer = cb  # keep the trampoline alive
        _check(lib().topohip_set_iteration_observer(self._h, cb, None))

    def _mask(self, mask):
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        if m.size != self.nvox:
            raise ValueError("mask size != voxel count")
        return m

    def filter(self, radius, x, adjoint=False, mask=None):
        """filter_density (adjoint False) / filter_sensitivity; with a design
        mask (MaskValue bytes: 0 Active, 1 FrozenSolid, 2 FrozenVoid) the
        Active-only filter of make_density_filter(grid, radius, mask)."""
        x = _f64(x, self.nvox)
        out = np.zeros(self.nvox)
        if mask is None:
            _check(lib().topohip_filter(self._h, C.c_double(radius), _p(x), _p(out),
                                        int(adjoint)))
        else:
            m = self._mask(mask)
            _check(lib().topohip_filter_masked(self._h, C.c_double(radius), _p(m), _p(x),
                                               _p(out), int(adjoint)))
        return out

    def oc_update(self, radius, x, dc, volfrac, move=0.2, rho_min=1e-3):
        x = _f64(x, self.nvox)
        dc = _f64(dc, self.nvox)
        out = np.zeros(self.nvox)
        act = np.zeros(self.nvox, dtype=np.uint8)
        nb = C.c_int(0)
        lam = C.c_double(0)
        _check(lib().topohip_oc_update(self._h, C.c_double(radius), _p(x), _p(dc),
                                       C.c_double(volfrac), C.c_double(move), C.c_double(rho_min),
                                       _p(out), _p(act), C.byref(nb), C.byref(lam)))
        return out, act, nb.value, lam.value

    def oc_margin(self):
        """(last_gap, run_min_gap, bound): the OC search's smallest relative
        |vol - target| (last search; smallest since the last optimize started)
        and the rounding bound of the volume sum."""
        a, b, c = C.c_double(0), C.c_double(0), C.c_double(0)
        _check(lib().topohip_oc_margin(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def mma_update(self, radius, mma_iter, x, dc, volfrac, move=0.2, rho_min=1e-3):
        """mma_update: one MMA step; the handle keeps MmaState between calls
        (restart a run at mma_iter 1). Returns (x_new, lambda, change)."""
        x = _f64(x, self.nvox)
        dc = _f64(dc, self.nvox)
        out = np.zeros(self.nvox)
        lam = C.c_double(0)
        ch = C.c_double(0)
        _check(lib().topohip_mma_update(self._h, C.c_double(radius), int(mma_iter), _p(x), _p(dc),
                                        C.c_double(volfrac), C.c_double(move),
                                        C.c_double(rho_min), _p(out), C.byref(lam), C.byref(ch)))
        return out, lam.value, ch.value

    def optimize_mma(self, volfrac, rmin, iterations, penalty=3.0, move=0.2, E0=1.0, rho_min=1e-3,
                     change_tol=0.0, cg_tol=1e-8, cg_max_iterations=0, continuation=None,
                     plateau_window=10, plateau_tol=1e-3, plateau_min_drop=0.05,
                     plateau_flat_windows=3, mask=None, projection=None, eta=0.5,
                     adaptive=None):
        """simp_optimize with SimpUpdater::MMA (plateau stop); with a design
        mask the masked loop (mma_update_masked). projection = (beta0,
        beta_max) turns on SimpOptions::mma_projection: Heaviside projection
        with beta continuation, iterations the safety cap. adaptive = (grow,
        shrink, osc_lo, osc_hi, max, min) turns on SimpOptions::adaptive_move."""
        o = Options(volfrac, rmin, move, penalty, E0, rho_min, change_tol, cg_tol, int(iterations),
                    int(cg_max_iterations), 0.0, 0.0, 0.0, 0)
        if continuation:
            o.pc_p0, o.pc_p1, o.pc_step, o.pc_per_stage = continuation
        m = MmaOptions(int(plateau_window), plateau_tol, plateau_min_drop, int(plateau_flat_windows))
        if adaptive is not None:
            m.adaptive_move = 1
            (m.adaptive_move_grow, m.adaptive_move_shrink, m.adaptive_move_osc_lo,
             m.adaptive_move_osc_hi, m.adaptive_move_max, m.adaptive_move_min) = adaptive
        design = np.zeros(self.nvox)
        xphys = np.zeros(self.nvox)
        hist = np.zeros(4 * iterations)
        res = Result()
        mk = None if mask is None else self._mask(mask)
        if projection is None:
            _check(lib().topohip_optimize_mma(self._h, C.byref(o), C.byref(m), _p(mk),
                                              _p(design), _p(xphys), _p(hist), C.byref(res)))
        else:
            b0, bmax = projection
            _check(lib().topohip_optimize_mma_projected(
                self._h, C.byref(o), C.byref(m), C.c_double(b0), C.c_double(bmax),
                C.c_double(eta), _p(mk), _p(design), _p(xphys), _p(hist), C.byref(res)))
        n = res.iterations
        out = res.as_dict()
        out.update(design=design, xphys=xphys, history=hist[: 4 * n].reshape(n, 4))
        return out

    def optimize(self, volfrac, rmin, iterations, penalty=3.0, move=0.2, E0=1.0, rho_min=1e-3,
                 change_tol=0.0, cg_tol=1e-8, cg_max_iterations=0, continuation=None, mask=None):
        """The OC design loop; with a design mask the masked loop (the handle
        must be built on the analysis grid: FrozenVoid voxels empty)."""
        o = Options(volfrac, rmin, move, penalty, E0, rho_min, change_tol, cg_tol, int(iterations),
                    int(cg_max_iterations), 0.0, 0.0, 0.0, 0)
        if continuation:
            o.pc_p0, o.pc_p1, o.pc_step, o.pc_per_stage = continuation
        design = np.zeros(self.nvox)
        xphys = np.zeros(self.nvox)
        hist = np.zeros(4 * iterations)
        res = Result()
        if mask is None:
            _check(lib().topohip_optimize(self._h, C.byref(o), _p(design), _p(xphys), _p(hist),
                                          C.byref(res)))
        else:
            m = self._mask(mask)
            _check(lib().topohip_optimize_masked(self._h, _p(m), C.byref(o), _p(design),
                                                 _p(xphys), _p(hist), C.byref(res)))
        n = res.iterations
        out = res.as_dict()
        out.update(design=design, xphys=xphys, history=hist[: 4 * n].reshape(n, 4))
        return out

    def optimize_projected(self, volfrac, rmin, stages, eta=0.5, penalty=3.0, E0=1.0,
                           rho_min=1e-3, change_tol=0.0, cg_tol=1e-8, cg_max_iterations=0,
                           mask=None):
        """The design loop with Heaviside projection stages [(beta, move, iterations), ...];
        with a design mask, the masked loop's projected stages (the handle on the
        analysis grid, as optimize(mask=...))."""
        n_it = sum(int(st[2]) for st in stages)
        o = Options(volfrac, rmin, 0.2, penalty, E0, rho_min, change_tol, cg_tol, max(n_it, 1),
                    int(cg_max_iterations), 0.0, 0.0, 0.0, 0)
        betas = np.array([st[0] for st in stages], dtype=np.float64)
        moves = np.array([st[1] for st in stages], dtype=np.float64)
        iters = np.array([st[2] for st in stages], dtype=np.int32)
        design = np.zeros(self.nvox)
        xphys = np.zeros(self.nvox)
        hist = np.zeros(4 * n_it)
        res = Result()
        if mask is None:
            _check(lib().topohip_optimize_projected(self._h, C.byref(o), len(stages), _p(betas),
                                                    _p(moves), _p(iters), C.c_double(eta),
                                                    _p(design), _p(xphys), _p(hist),
                                                    C.byref(res)))
        else:
            m = self._mask(mask)
            _check(lib().topohip_optimize_masked_projected(
                self._h, _p(m), C.byref(o), len(stages), _p(betas), _p(moves), _p(iters),
                C.c_double(eta), _p(design), _p(xphys), _p(hist), C.byref(res)))
        n = res.iterations
        out = res.as_dict()
        out.update(design=design, xphys=xphys, history=hist[: 4 * n].reshape(n, 4))
        return out

    # -- device-resident benchmarking ----------------------------------------
    def bench_prepare(self, seed=0x5EED):
        _check(lib().topohip_bench_prepare(self._h, C.c_uint64(seed)))

    def bench_ku(self, precision=PREC_MIXED, reps=50):
        ms = C.c_double(0)
        _check(lib().topohip_bench_ku(self._h, int(precision), int(reps), C.byref(ms)))
        return ms.value

    KU_VARIANTS = {"f32": 0, "f64": 1, "ap64_dot": 2, "res0_f32": 3, "jac_dot_f32": 4,
                   "ap64_pdir": 5}

    def bench_ku_variant(self, variant, reps=50):
        """(ms per launch, rotated operand sets) of a fine K.u kernel variant."""
        ms = C.c_double(0)
        ns = C.c_int(0)
        _check(lib().topohip_bench_ku_variant(self._h, int(self.KU_VARIANTS[variant]), int(reps),
                                              C.byref(ms), C.byref(ns)))
        return ms.value, ns.value

    def ku_tile(self):
        """(tx, ty, tz) of the FP64 applies (the PCG's A.p)."""
        t = [C.c_int(0), C.c_int(0), C.c_int(0)]
        _check(lib().topohip_ku_tile(self._h, *[C.byref(v) for v in t]))
        return tuple(v.value for v in t)

    def ku_tiles(self):
        """(z-chunk of the FP32 applies, z-chunk of the FP64 applies)."""
        t = [C.c_int(0), C.c_int(0)]
        _check(lib().topohip_ku_tiles(self._h, *[C.byref(v) for v in t]))
        return tuple(v.value for v in t)

    def bench_fields(self, reps=50):
        """ms per launch: (von Mises -> f32, displacement -> f32, density -> f16)."""
        ms = (C.c_double * 3)()
        _check(lib().topohip_bench_fields(self._h, int(reps), ms))
        return tuple(ms)


class Server:
    """Persistent continuous-batching job loop (topohip_server_*).

    Jobs are dicts (or Job structs) with nx, ny, nz, bc_kind (0 cantilever,
    1 MBB), volfrac, rmin, penalty, move, cg_tol, iterations, precision.
    """

    def __init__(self, budget_bytes=64 << 30, max_slots=8):
        self._s = C.c_void_p()
        code = lib().topohip_server_create(C.c_size_t(budget_bytes), int(max_slots),
                                           C.byref(self._s))
        if code != OK:
            raise ValueError(f"topohip_server_create failed ({code})")

    @staticmethod
    def make_job(nx, ny, nz, bc_kind=0, volfrac=0.3, rmin=1.5, penalty=3.0, move=0.2,
                 cg_tol=1e-6, iterations=60, precision=PREC_MIXED):
        return Job(int(nx), int(ny), int(nz), int(bc_kind), volfrac, rmin, penalty, move, cg_tol,
                   int(iterations), int(precision))

    @staticmethod
    def footprint(job):
        return lib().topohip_job_footprint(C.byref(job))

    def set_output(self, directory):
        """Per-job run records under directory/job_<id>/ (None turns them off)."""
        d = None if directory is None else str(directory).encode()
        _check(lib().topohip_server_set_output(self._s, d))

    def submit(self, job):
        if isinstance(job, dict):
            job = self.make_job(**job)
        jid = C.c_int64()
        code = lib().topohip_server_submit(self._s, C.byref(job), C.byref(jid))
        if code != OK:
            raise ValueError(f"topohip_server_submit rejected the job ({code})")
        return jid.value

    def poll(self, jid):
        r = JobResult()
        _check(lib().topohip_server_poll(self._s, C.c_int64(jid), C.byref(r)))
        out = r.as_dict()
        if out["state"] == 3:
            buf = C.create_string_buffer(1024)
            lib().topohip_server_job_error(self._s, C.c_int64(jid), buf, 1024)
            out["error"] = buf.value.decode(errors="replace")
        return out

    def drain(self):
        _check(lib().topohip_server_drain(self._s))

    def density(self, jid, nvox):
        out = np.zeros(int(nvox))
        _check(lib().topohip_server_fetch_density(self._s, C.c_int64(jid), _p(out)))
        return out

    def close(self):
        if getattr(self, "_s", None) is not None and self._s.value:
            lib().topohip_server_destroy(self._s)
            self._s = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# -- reference-named free functions ------------------------------------------

def fea_matfree_apply(nx, ny, nz, youngs_per_voxel, u, poisson=0.3, h=1.0, solid=None,
                      precision="f32"):
    """y = K(E) u over all solid voxels, no BCs (fea_matfree_apply, graded)."""
    with Problem(nx, ny, nz, h=h, poisson=poisson, solid=solid) as p:
        return p.apply(youngs_per_voxel, u, precision)


def fea_solve_mgcg_matfree(nx, ny, nz, youngs_per_voxel, poisson, bcs, loads, tolerance=1e-8,
                           max_iterations=0, initial_guess=None, h=1.0, solid=None,
                           precision=PREC_MIXED):
    """Multigrid-CG solve; returns (u, CgInfo dict)."""
    with Problem(nx, ny, nz, bcs, loads, h=h, poisson=poisson, solid=solid,
                 precision=precision) as p:
        return p.solve(youngs_per_voxel, tolerance, max_iterations, initial_guess)


def set_solve_deadline_ms(absolute_steady_ms):
    """fea_set_solve_deadline_ms for the calling thread (<= 0 disarms);
    returns the previous value. Times are on steady_ms()'s clock."""
    f = lib().topohip_set_solve_deadline_ms
    f.restype = C.c_double
    return f(C.c_double(absolute_steady_ms))


def steady_ms():
    f = lib().topohip_steady_ms
    f.restype = C.c_double
    return f()


def reset_mg_stagnation_latch():
    """fea_matfree_reset_mg_stagnation_latch for the calling thread: the run's
    driver calls it at run start (the latch is sticky otherwise)."""
    lib().topohip_reset_mg_stagnation_latch()


def mg_stagnation_latched():
    """fea_matfree_mg_stagnation_latched: whether the calling thread's solves
    skip multigrid after 3 consecutive stagnated solves."""
    return bool(lib().topohip_mg_stagnation_latched())


def set_mg_rearm_period(period):
    """fea_matfree_set_mg_rearm_period for the calling thread (measurement only;
    0 = off, the default)."""
    lib().topohip_set_mg_rearm_period(int(period))


def mg_rearm_counts():
    """(period, attempts, carries) of the calling thread's latch re-arm."""
    L = lib()
    L.topohip_mg_rearm_attempts.restype = C.c_longlong
    L.topohip_mg_rearm_carries.restype = C.c_longlong
    return (int(L.topohip_mg_rearm_period()), int(L.topohip_mg_rearm_attempts()),
            int(L.topohip_mg_rearm_carries()))


def version():
    return lib().topohip_version().decode()


from .records import (CsvWriter, DesignVariant, FieldsVariant, GridFrame,  # noqa: E402
                      IterationObs, design_fingerprint, format_progress, format_variant,
                      iterations_csv_header, iterations_csv_row, read_density_f16,
                      read_design_file, write_design_file, write_fields_file, SnapshotCapture,
                      RunInfo, run_info_json, write_run_info)
