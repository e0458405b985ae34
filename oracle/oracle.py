"""ctypes binding of the FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the bench's
CPU-baseline leg, never by the product package. Each function mirrors the
reference routine named in oracle.hpp.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class Problem(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("h", C.c_double),
        ("solid", C.c_void_p),
        ("nbc", C.c_int), ("bc_node", C.c_void_p), ("bc_comp", C.c_void_p), ("bc_val", C.c_void_p),
        ("nload", C.c_int), ("ld_node", C.c_void_p), ("ld_comp", C.c_void_p), ("ld_val", C.c_void_p),
    ]


class Info(C.Structure):
    _fields_ = [
        ("converged", C.c_int), ("iterations", C.c_int), ("used_multigrid", C.c_int),
        ("mg_levels", C.c_int), ("hier_built", C.c_int), ("mg_cycles_attempted", C.c_int),
        ("residual", C.c_double), ("matvecs", C.c_longlong),
        ("t_build_ms", C.c_double), ("t_mg_build_ms", C.c_double), ("t_mg_ms", C.c_double),
        ("t_cg_ms", C.c_double), ("t_total_ms", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class OracleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"oracle status {code}: {msg}")
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        _lib.oracle_last_error.restype = C.c_char_p
        _lib.oracle_fnv1a64.restype = C.c_uint64
        _lib.oracle_fnv1a64.argtypes = [C.c_void_p, C.c_size_t]
        _lib.oracle_fnv1a64_basis.restype = C.c_uint64
        _lib.oracle_fnv1a64_basis.argtypes = [C.c_void_p, C.c_size_t, C.c_uint64]
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _check(code):
    if code != 0:
        raise OracleError(code, lib().oracle_last_error().decode())


def set_threads(n):
    lib().oracle_set_threads(int(n))


class Setup:
    """Grid + BCs + loads, held as numpy arrays so the ctypes pointers stay valid."""

    def __init__(self, nx, ny, nz, h=1.0, solid=None, bcs=(), loads=()):
        self.nx, self.ny, self.nz, self.h = nx, ny, nz, h
        self.solid = None if solid is None else np.ascontiguousarray(solid, dtype=np.uint8).ravel()
        bcs = list(bcs)
        loads = list(loads)
        self.bc_node = np.array([b[0] for b in bcs], dtype=np.int32)
        self.bc_comp = np.array([b[1] for b in bcs], dtype=np.int32)
        self.bc_val = np.array([b[2] for b in bcs], dtype=np.float64)
        self.ld_node = np.array([l[0] for l in loads], dtype=np.int32)
        self.ld_comp = np.array([l[1] for l in loads], dtype=np.int32)
        self.ld_val = np.array([l[2] for l in loads], dtype=np.float64)
        self.p = Problem(nx, ny, nz, h, _ptr(self.solid),
                         len(bcs), _ptr(self.bc_node), _ptr(self.bc_comp), _ptr(self.bc_val),
                         len(loads), _ptr(self.ld_node), _ptr(self.ld_comp), _ptr(self.ld_val))

    @property
    def nvox(self):
        return self.nx * self.ny * self.nz

    @property
    def ndof(self):
        return 3 * (self.nx + 1) * (self.ny + 1) * (self.nz + 1)


def hex8_stiffness(E=1.0, nu=0.3, h=1.0):
    K = np.zeros(576)
    _check(lib().oracle_hex8_stiffness(C.c_double(E), C.c_double(nu), C.c_double(h), _ptr(K)))
    return K.reshape(24, 24)


def apply(s, E, u, nu=0.3):
    E = np.ascontiguousarray(E, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    y = np.zeros(s.ndof)
    _check(lib().oracle_apply(C.byref(s.p), _ptr(E), C.c_double(nu), _ptr(u), _ptr(y)))
    return y


def solve(s, E, nu=0.3, tol=1e-8, max_it=0, use_mg=True, guess=None):
    E = np.ascontiguousarray(E, dtype=np.float64)
    g = None if guess is None else np.ascontiguousarray(guess, dtype=np.float64)
    u = np.zeros(s.ndof)
    info = Info()
    code = lib().oracle_solve(C.byref(s.p), _ptr(E), C.c_double(nu), C.c_double(tol),
                              int(max_it), int(use_mg), _ptr(g), _ptr(u), C.byref(info))
    _check(code)
    return u, info.as_dict()


def hierarchy(s, E, nu=0.3):
    E = np.ascontiguousarray(E, dtype=np.float64)
    lv = C.c_int(0)
    dofs = np.zeros(16, dtype=np.int32)
    _check(lib().oracle_hierarchy(C.byref(s.p), _ptr(E), C.c_double(nu), C.byref(lv), _ptr(dofs)))
    return lv.value, dofs[: lv.value].tolist()


def set_smoother(kind="jacobi", degree=1):
    """V-cycle smoother of every later solve: "jacobi" (the reference's) or
    "chebyshev" of the given degree."""
    _check(lib().oracle_set_smoother({"jacobi": 0, "chebyshev": 1}[kind], int(degree)))


def cheb_lambdas(s, E, nu=0.3):
    E = np.ascontiguousarray(E, dtype=np.float64)
    n = C.c_int(0)
    lam = np.zeros(16)
    _check(lib().oracle_cheb_lambdas(C.byref(s.p), _ptr(E), C.c_double(nu), C.byref(n),
                                     _ptr(lam)))
    return lam[: n.value].copy()


def level_operator(s, E, level, nu=0.3):
    """Hierarchy level `level` (>= 1) operator as (n, indptr, indices, data)
    over the level's active DOFs (node-major, then component)."""
    E = np.ascontiguousarray(E, dtype=np.float64)
    n = C.c_int(0)
    nnz = C.c_int(0)
    _check(lib().oracle_level_csr(C.byref(s.p), _ptr(E), C.c_double(nu), int(level), C.byref(n),
                                  C.byref(nnz), None, None, None))
    ptr = np.zeros(n.value + 1, dtype=np.int32)
    col = np.zeros(max(1, nnz.value), dtype=np.int32)
    val = np.zeros(max(1, nnz.value))
    _check(lib().oracle_level_csr(C.byref(s.p), _ptr(E), C.c_double(nu), int(level), C.byref(n),
                                  C.byref(nnz), _ptr(ptr), _ptr(col), _ptr(val)))
    return n.value, ptr, col[: nnz.value], val[: nnz.value]


def simp_compliance(s, density, E0=1.0, nu=0.3, penalty=3.0, rho_min=1e-3, tol=1e-8,
                    max_it=0, use_mg=True, guess=None):
    d = np.ascontiguousarray(density, dtype=np.float64)
    c = C.c_double(0)
    dc = np.zeros(s.nvox)
    u = np.zeros(s.ndof)
    info = Info()
    if guess is None:
        _check(lib().oracle_simp_compliance(C.byref(s.p), _ptr(d), C.c_double(E0), C.c_double(nu),
                                            C.c_double(penalty), C.c_double(rho_min),
                                            C.c_double(tol), int(max_it), int(use_mg), C.byref(c),
                                            _ptr(dc), _ptr(u), C.byref(info)))
    else:
        g = np.ascontiguousarray(guess, dtype=np.float64)
        _check(lib().oracle_simp_compliance_warm(C.byref(s.p), _ptr(d), C.c_double(E0),
                                                 C.c_double(nu), C.c_double(penalty),
                                                 C.c_double(rho_min), C.c_double(tol), int(max_it),
                                                 _ptr(g), C.byref(c), _ptr(dc), _ptr(u),
                                                 C.byref(info)))
    return c.value, dc, u, info.as_dict()


# Design masks: one MaskValue byte per voxel (0 Active, 1 FrozenSolid,
# 2 FrozenVoid; core/include/topopt/voxel.hpp:42-46). mask=None is the plain path.
ACTIVE, FROZEN_SOLID, FROZEN_VOID = 0, 1, 2


def _mask(s, mask):
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    if m.size != s.nvox:
        raise ValueError("mask size != voxel count")
    return m, m.ctypes.data_as(C.c_void_p)


def filter(s, radius, x, adjoint=False, mask=None):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(s.nvox)
    if mask is None:
        _check(lib().oracle_filter(C.byref(s.p), C.c_double(radius), _ptr(x), _ptr(out),
                                   int(adjoint)))
    else:
        m, mp = _mask(s, mask)
        _check(lib().oracle_filter_masked(C.byref(s.p), C.c_double(radius), mp, _ptr(x), _ptr(out),
                                          int(adjoint)))
    return out


def filter_rows(s, radius, mask=None):
    nnz = np.zeros(s.nvox, dtype=np.int32)
    hs = np.zeros(s.nvox)
    if mask is None:
        _check(lib().oracle_filter_rows(C.byref(s.p), C.c_double(radius), _ptr(nnz), _ptr(hs)))
    else:
        m, mp = _mask(s, mask)
        _check(lib().oracle_filter_rows_masked(C.byref(s.p), C.c_double(radius), mp, _ptr(nnz),
                                               _ptr(hs)))
    return nnz, hs


def oc_update(s, radius, x, dc, volfrac, move=0.2, rho_min=1e-3, mask=None):
    x = np.ascontiguousarray(x, dtype=np.float64)
    dc = np.ascontiguousarray(dc, dtype=np.float64)
    out = np.zeros(s.nvox)
    nb = C.c_int(0)
    lam = C.c_double(0)
    if mask is None:
        _check(lib().oracle_oc_update(C.byref(s.p), C.c_double(radius), _ptr(x), _ptr(dc),
                                      C.c_double(volfrac), C.c_double(move), C.c_double(rho_min),
                                      _ptr(out), C.byref(nb), C.byref(lam)))
    else:
        m, mp = _mask(s, mask)
        _check(lib().oracle_oc_update_masked(C.byref(s.p), C.c_double(radius), mp, _ptr(x),
                                             _ptr(dc), C.c_double(volfrac), C.c_double(move),
                                             C.c_double(rho_min), _ptr(out), C.byref(nb),
                                             C.byref(lam)))
    return out, nb.value, lam.value


def optimize(s, volfrac, radius, iterations, E0=1.0, nu=0.3, penalty=3.0, rho_min=1e-3,
             move=0.2, change_tol=0.0, cg_tol=1e-8, use_mg=True, continuation=None, mask=None):
    design = np.zeros(s.nvox)
    xphys = np.zeros(s.nvox)
    hist = np.zeros(4 * iterations)
    it = C.c_int(0)
    comp = C.c_double(0)
    vf = C.c_double(0)
    p0, p1, step, per = continuation if continuation else (0.0, 0.0, 0.0, 0)
    args = (C.c_double(E0), C.c_double(nu), C.c_double(penalty), C.c_double(rho_min),
            C.c_double(volfrac), C.c_double(radius), C.c_double(move), int(iterations),
            C.c_double(change_tol), C.c_double(cg_tol), int(use_mg), C.c_double(p0),
            C.c_double(p1), C.c_double(step), int(per), _ptr(design), _ptr(xphys), _ptr(hist),
            C.byref(it), C.byref(comp), C.byref(vf))
    if mask is None:
        _check(lib().oracle_optimize(C.byref(s.p), *args))
    else:
        m, mp = _mask(s, mask)
        _check(lib().oracle_optimize_masked(C.byref(s.p), mp, *args))
    n = it.value
    return {"design": design, "xphys": xphys, "history": hist[: 4 * n].reshape(n, 4),
            "iterations": n, "compliance": comp.value, "volume_fraction": vf.value}


# heaviside_continuation_schedule (core/src/simp/simp.cpp:367-373): (beta, move, iterations)
GATE_V2_SCHEDULE = [(1.0, 0.2, 50), (2.0, 0.2, 50), (4.0, 0.2, 50), (8.0, 0.2, 50),
                    (16.0, 0.1, 50), (32.0, 0.05, 50)]


def optimize_projected(s, volfrac, radius, stages, eta=0.5, E0=1.0, nu=0.3, penalty=3.0,
                       rho_min=1e-3, change_tol=0.0, cg_tol=1e-8, use_mg=True, mask=None):
    """The OC loop with projected stages [(beta, move, iterations), ...]; with a
    design mask the masked loop's projected stages (simp.cpp:2989-3689)."""
    n_it = sum(int(st[2]) for st in stages)
    betas = np.array([st[0] for st in stages], dtype=np.float64)
    moves = np.array([st[1] for st in stages], dtype=np.float64)
    iters = np.array([st[2] for st in stages], dtype=np.int32)
    design = np.zeros(s.nvox)
    xphys = np.zeros(s.nvox)
    hist = np.zeros(4 * n_it)
    it = C.c_int(0)
    comp = C.c_double(0)
    vf = C.c_double(0)
    mp = None
    if mask is not None:
        m, mp = _mask(s, mask)
    _check(lib().oracle_optimize_projected(
        C.byref(s.p), mp, C.c_double(E0), C.c_double(nu), C.c_double(penalty), C.c_double(rho_min),
        C.c_double(volfrac), C.c_double(radius), len(stages), _ptr(betas), _ptr(moves),
        iters.ctypes.data_as(C.c_void_p), C.c_double(eta), C.c_double(change_tol),
        C.c_double(cg_tol), int(use_mg), _ptr(design), _ptr(xphys), _ptr(hist), C.byref(it),
        C.byref(comp), C.byref(vf)))
    n = it.value
    return {"design": design, "xphys": xphys, "history": hist[: 4 * n].reshape(n, 4),
            "iterations": n, "compliance": comp.value, "volume_fraction": vf.value}


def self_weight(s, density=1.0, gravity=1.0, direction=(0.0, 0.0, -1.0)):
    f = np.zeros(s.ndof)
    _check(lib().oracle_self_weight(C.byref(s.p), C.c_double(density), C.c_double(gravity),
                                    C.c_double(direction[0]), C.c_double(direction[1]),
                                    C.c_double(direction[2]), _ptr(f)))
    return f


def penalty_at(p0, p1, step, per_stage, iteration):
    out = C.c_double(0)
    _check(lib().oracle_penalty_at(C.c_double(p0), C.c_double(p1), C.c_double(step),
                                   int(per_stage), int(iteration), C.byref(out)))
    return out.value


def heaviside(rt, beta, eta=0.5):
    v = C.c_double(0)
    d = C.c_double(0)
    _check(lib().oracle_heaviside(C.c_double(rt), C.c_double(beta), C.c_double(eta), C.byref(v),
                                  C.byref(d)))
    return v.value, d.value


def von_mises_field(s, u, E=1.0, nu=0.3):
    """fea_von_mises_field: centroid von Mises per solid voxel (0 elsewhere)."""
    uu = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(s.nvox)
    _check(lib().oracle_von_mises(C.byref(s.p), C.c_double(E), C.c_double(nu), _ptr(uu),
                                  _ptr(out)))
    return out


def hex8_stress(E, nu, h, ue, xi=0.0, eta=0.0, zeta=0.0):
    ue = np.ascontiguousarray(ue, dtype=np.float64)
    sig = np.zeros(6)
    vm = C.c_double(0)
    _check(lib().oracle_hex8_stress(C.c_double(E), C.c_double(nu), C.c_double(h), _ptr(ue),
                                    C.c_double(xi), C.c_double(eta), C.c_double(zeta), _ptr(sig),
                                    C.byref(vm)))
    return sig, vm.value


def float_to_half(x):
    """IEEE half bits (uint16) of float32(x), round to nearest even."""
    f = np.ascontiguousarray(x, dtype=np.float32).ravel()
    out = np.zeros(f.size, dtype=np.uint16)
    _check(lib().oracle_float_to_half(f.ctypes.data_as(C.c_void_p), C.c_size_t(f.size),
                                      out.ctypes.data_as(C.c_void_p)))
    return out


def fnv1a64(data, basis=None):
    """design_fingerprint's FNV-1a 64 (the reference's offset basis unless
    another is given)."""
    b = bytes(data)
    buf = C.create_string_buffer(b, len(b))
    if basis is None:
        return int(lib().oracle_fnv1a64(buf, C.c_size_t(len(b))))
    return int(lib().oracle_fnv1a64_basis(buf, C.c_size_t(len(b)), C.c_uint64(basis)))


def optimize_mma(s, volfrac, radius, iterations, E0=1.0, nu=0.3, penalty=3.0, rho_min=1e-3,
                 move=0.2, change_tol=0.0, cg_tol=1e-8, plateau_window=10, plateau_tol=1e-3,
                 plateau_min_drop=0.05, plateau_flat_windows=3, mask=None, projection=None,
                 eta=0.5, adaptive=None):
    """simp_optimize with SimpUpdater::MMA (plateau stop); with a design mask
    the masked loop with mma_update_masked. projection = (beta0, beta_max)
    turns on SimpOptions::mma_projection (beta continuation); adaptive =
    (grow, shrink, osc_lo, osc_hi, max, min) turns on SimpOptions::adaptive_move."""
    ad = None if adaptive is None else np.ascontiguousarray(adaptive, dtype=np.float64)
    b0, bmax = projection if projection else (1.0, 32.0)
    design = np.zeros(s.nvox)
    xphys = np.zeros(s.nvox)
    hist = np.zeros(4 * iterations)
    it = C.c_int(0)
    comp = C.c_double(0)
    vf = C.c_double(0)
    m, mp = (None, None) if mask is None else _mask(s, mask)  # m keeps the buffer alive
    _check(lib().oracle_optimize_mma(
        C.byref(s.p), C.c_double(E0), C.c_double(nu), C.c_double(penalty), C.c_double(rho_min),
        C.c_double(volfrac), C.c_double(radius), C.c_double(move), int(iterations),
        C.c_double(change_tol), C.c_double(cg_tol), int(plateau_window), C.c_double(plateau_tol),
        C.c_double(plateau_min_drop), int(plateau_flat_windows),
        int(projection is not None), C.c_double(b0), C.c_double(bmax), C.c_double(eta),
        _ptr(ad), mp, _ptr(design), _ptr(xphys),
        _ptr(hist), C.byref(it), C.byref(comp), C.byref(vf)))
    n = it.value
    return {"design": design, "xphys": xphys, "history": hist[: 4 * n].reshape(n, 4),
            "iterations": n, "compliance": comp.value, "volume_fraction": vf.value}


def mma_continuation_move(move, beta):
    """mma_continuation_move (core/src/simp/simp.cpp:251-254)."""
    f = lib().oracle_mma_continuation_move
    f.restype = C.c_double
    return f(C.c_double(move), C.c_double(beta))


def mma_steps(s, radius, xs, dcs, volfrac, move=0.2, rho_min=1e-3):
    """Successive mma_update calls (one state, mma_iter 1..n) on the given
    designs and compliance sensitivities; returns (new designs, lambdas)."""
    xs = np.ascontiguousarray(xs, dtype=np.float64).reshape(-1, s.nvox)
    dcs = np.ascontiguousarray(dcs, dtype=np.float64).reshape(-1, s.nvox)
    n = xs.shape[0]
    out = np.zeros_like(xs)
    lam = np.zeros(n)
    _check(lib().oracle_mma_steps(C.byref(s.p), C.c_double(radius), n, _ptr(xs), _ptr(dcs),
                                  C.c_double(volfrac), C.c_double(move), C.c_double(rho_min),
                                  _ptr(out), _ptr(lam)))
    return out, lam


def mma_plateau(c, window=10, rel_tol=1e-3, min_drop=0.05, flat_windows=3):
    cc = np.ascontiguousarray(c, dtype=np.float64)
    out = C.c_int(0)
    _check(lib().oracle_mma_plateau(_ptr(cc), cc.size, int(window), C.c_double(rel_tol),
                                    C.c_double(min_drop), int(flat_windows), C.byref(out)))
    return bool(out.value)
