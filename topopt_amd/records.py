"""Run records byte-compatible with the reference, over the C ABI.

  iterations.csv       IterationCsvWriter, core/include/topopt/observability.hpp:60-214
  density snapshots    write_density_snapshot / read_density_f16,
                       core/src/simp/observability.cpp:362-425
  design.bin v1        write_design_file / read_design_file / design_fingerprint,
                       core/include/topopt/design_store.hpp:40-100
  fields.bin v1        write_fields_file, core/include/topopt/fields.hpp:28-70
  PROGRESS / VARIANT   the CLI stdout lines, core/src/cli/run_job.cpp:10223-10259
  run_info.json        run_info_json / write_run_info (the keys of this solve),
                       core/src/simp/observability.cpp:476-1284, 1761-1767

The writers live in libtopopt_hip.so (formats.cpp); the arrays they carry are
formed on the device (Problem.fields_f32, Problem.density_f16).
"""
import ctypes as C

import numpy as np

from . import _check, _p, lib


class IterationObs(C.Structure):
    """topohip_iteration_obs (SimpIterationObservation + IterationPhaseTimes)."""
    _fields_ = [
        ("iteration", C.c_int), ("compliance", C.c_double), ("change", C.c_double),
        ("volume_fraction", C.c_double),
        ("cg_iterations", C.c_int), ("cg_used_multigrid", C.c_int), ("cg_mg_levels", C.c_int),
        ("cg_hier_built", C.c_int), ("cg_mg_cycles_attempted", C.c_int),
        ("beta", C.c_double), ("active_fraction", C.c_double),
        ("plateau", C.c_int), ("infeasible", C.c_int), ("recycle_dim", C.c_int),
        ("geneo_dim", C.c_int), ("geneo_action", C.c_int), ("geneo_burn", C.c_int),
        ("geneo_threshold", C.c_int),
        ("total_ms", C.c_double), ("tail_prev_ms", C.c_double), ("filter_ms", C.c_double),
        ("project_ms", C.c_double), ("solve_ms", C.c_double), ("fea_ms", C.c_double),
        ("sens_ms", C.c_double), ("update_ms", C.c_double), ("analysis_ms", C.c_double),
        ("observe_ms", C.c_double), ("residual_ms", C.c_double),
        ("solver_build_ms", C.c_double), ("solver_mg_build_ms", C.c_double),
        ("solver_mg_ms", C.c_double), ("solver_cg_ms", C.c_double),
        ("solver_geneo_setup_ms", C.c_double), ("solver_geneo_apply_ms", C.c_double),
        ("solver_recycle_ms", C.c_double),
        ("fea_solves", C.c_longlong), ("matvecs", C.c_longlong),
        ("rss_mb", C.c_double), ("peak_rss_mb", C.c_double), ("compressed_mb", C.c_double),
        ("available_mb", C.c_double), ("major_faults", C.c_longlong), ("swapins", C.c_longlong),
    ]

    @classmethod
    def make(cls, **kw):
        """Defaults as the reference's: active_fraction 1, memory columns -1."""
        o = cls()
        o.active_fraction = 1.0
        o.rss_mb = o.peak_rss_mb = o.compressed_mb = o.available_mb = -1.0
        o.major_faults = o.swapins = -1
        for k, v in kw.items():
            setattr(o, k, v)
        return o

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


ITERATION_CB = C.CFUNCTYPE(None, C.POINTER(IterationObs), C.c_void_p)


class GridFrame(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("origin", C.c_double * 3), ("spacing", C.c_double)]

    @classmethod
    def make(cls, nx, ny, nz, spacing=1.0, origin=(0.0, 0.0, 0.0)):
        g = cls()
        g.nx, g.ny, g.nz, g.spacing = nx, ny, nz, spacing
        g.origin[:] = [float(v) for v in origin]
        return g


class DesignVariant(C.Structure):
    _fields_ = [
        ("requested_volume_fraction", C.c_double), ("achieved_volume_fraction", C.c_double),
        ("margin_worst_case", C.c_double), ("margin_effective", C.c_double),
        ("max_von_mises_mpa", C.c_double), ("accepted", C.c_int), ("iterations", C.c_int),
        ("applied_build_dir", C.c_double * 3), ("build_direction_auto_applied", C.c_int),
        ("export_baked", C.c_int), ("fingerprint", C.c_uint64), ("density", C.c_void_p),
    ]


class FieldsVariant(C.Structure):
    _fields_ = [
        ("requested_volume_fraction", C.c_double), ("mass_grams", C.c_double),
        ("support_volume_voxels", C.c_int), ("von_mises", C.c_void_p),
        ("displacement", C.c_void_p),
    ]


def _setup():
    L = lib()
    if getattr(L, "_records_ready", False):
        return L
    L.topohip_iterations_csv_header.restype = C.c_char_p
    L.topohip_csv_rows.restype = C.c_size_t
    L.topohip_csv_rows.argtypes = [C.c_void_p]
    L.topohip_design_fingerprint.restype = C.c_uint64
    L.topohip_design_fingerprint.argtypes = [C.c_void_p, C.c_size_t]
    L.topohip_iterations_csv_row.argtypes = [C.c_size_t, C.POINTER(IterationObs), C.c_longlong,
                                             C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]
    L.topohip_csv_append.argtypes = [C.c_void_p, C.c_size_t, C.POINTER(IterationObs)]
    L.topohip_csv_append_at.argtypes = [C.c_void_p, C.c_size_t, C.POINTER(IterationObs),
                                        C.c_longlong]
    L.topohip_csv_close.argtypes = [C.c_void_p]
    L.topohip_read_density_f16.argtypes = [C.c_char_p, C.c_size_t, C.c_void_p]
    L.topohip_snapshots_capture.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_size_t, C.c_int, C.c_int, C.POINTER(C.c_int)]
    L.topohip_snapshots_destroy.argtypes = [C.c_void_p]
    L.topohip_snapshots_counts.argtypes = [C.c_void_p, C.POINTER(C.c_size_t),
                                           C.POINTER(C.c_size_t)]
    L._records_ready = True
    return L


def iterations_csv_header():
    return _setup().topohip_iterations_csv_header().decode()


def iterations_csv_row(rung, obs, wall_ms):
    buf = C.create_string_buffer(2048)
    n = C.c_size_t(0)
    _check(_setup().topohip_iterations_csv_row(rung, C.byref(obs), wall_ms, buf, 2048,
                                               C.byref(n)))
    return buf.raw[:n.value].decode()


class CsvWriter:
    """IterationCsvWriter: header on open, one flushed row per append."""

    def __init__(self, path):
        self._h = C.c_void_p()
        _check(_setup().topohip_csv_open(str(path).encode(), C.byref(self._h)))

    def append(self, rung, obs):
        _check(_setup().topohip_csv_append(self._h, rung, C.byref(obs)))

    def append_at(self, rung, obs, wall_ms):
        _check(_setup().topohip_csv_append_at(self._h, rung, C.byref(obs), wall_ms))

    def rows(self):
        return int(_setup().topohip_csv_rows(self._h))

    def close(self):
        if self._h:
            _check(_setup().topohip_csv_close(self._h))
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read_density_f16(path, count):
    out = np.zeros(count, dtype=np.float32)
    _check(_setup().topohip_read_density_f16(str(path).encode(), count, _p(out)))
    return out


def design_fingerprint(density):
    d = np.ascontiguousarray(density, dtype=np.float64).ravel()
    return int(_setup().topohip_design_fingerprint(_p(d), d.size))


_DESIGN_KEYS = ("requested_volume_fraction", "achieved_volume_fraction", "margin_worst_case",
                "margin_effective", "max_von_mises_mpa", "accepted", "iterations",
                "build_direction_auto_applied", "export_baked")


def write_design_file(path, grid, variants):
    """variants: dicts with a `density` array and the StoredDesign scalars."""
    keep, arr = [], (DesignVariant * max(1, len(variants)))()
    for i, v in enumerate(variants):
        d = np.ascontiguousarray(v["density"], dtype=np.float64).ravel()
        keep.append(d)
        for k in _DESIGN_KEYS:
            if k in v:
                setattr(arr[i], k, v[k])
        arr[i].applied_build_dir[:] = list(v.get("applied_build_dir", (0.0, 0.0, 1.0)))
        arr[i].density = d.ctypes.data
    _check(_setup().topohip_write_design_file(str(path).encode(), C.byref(grid), len(variants),
                                              arr))


def read_design_file(path):
    L = _setup()
    g = GridFrame()
    n = C.c_int(0)
    _check(L.topohip_read_design_header(str(path).encode(), C.byref(g), C.byref(n)))
    out = {"grid": g, "variants": []}
    for i in range(n.value):
        m = DesignVariant()
        d = np.zeros(g.nx * g.ny * g.nz)
        _check(L.topohip_read_design_variant(str(path).encode(), i, C.byref(m), _p(d)))
        v = {k: getattr(m, k) for k in _DESIGN_KEYS}
        v["applied_build_dir"] = tuple(m.applied_build_dir)
        v["fingerprint"] = int(m.fingerprint)
        v["density"] = d
        out["variants"].append(v)
    return out


def write_fields_file(path, grid, variants):
    """variants: dicts with float32 `von_mises` and `displacement` arrays and
    requested_volume_fraction, mass_grams, support_volume_voxels."""
    keep, arr = [], (FieldsVariant * max(1, len(variants)))()
    for i, v in enumerate(variants):
        arr[i].requested_volume_fraction = v.get("requested_volume_fraction", 0.0)
        arr[i].mass_grams = v.get("mass_grams", 0.0)
        arr[i].support_volume_voxels = v.get("support_volume_voxels", 0)
        for k in ("von_mises", "displacement"):
            a = v.get(k)
            if a is not None:
                a = np.ascontiguousarray(a, dtype=np.float32).ravel()
                keep.append(a)
                setattr(arr[i], k, a.ctypes.data)
    _check(_setup().topohip_write_fields_file(str(path).encode(), C.byref(grid), len(variants),
                                              arr))


class RunInfo(C.Structure):
    """topohip_run_info: the run_info.json keys this library writes."""
    _fields_ = [
        ("cli_version", C.c_char_p), ("fingerprint", C.c_char_p), ("mode", C.c_char_p),
        ("material", C.c_char_p), ("source_format", C.c_char_p), ("resolution", C.c_int),
        ("load_source", C.c_char_p), ("solver", C.c_char_p),
        ("cg_multigrid_observed", C.c_int), ("cg_multigrid", C.c_int), ("mg_levels", C.c_int),
        ("mg_mode", C.c_char_p), ("mixed_precision", C.c_int), ("matfree_threads", C.c_int),
        ("warm_start_inherit", C.c_int), ("solved_grid_dofs", C.c_longlong),
        ("projection", C.c_int), ("n_rungs", C.c_int),
        ("rung_non_convergent", C.c_void_p), ("rung_non_convergent_iteration", C.c_void_p),
        ("rung_non_convergent_residual", C.c_void_p), ("n_ladder", C.c_int),
        ("ladder", C.c_void_p), ("ladder_mode", C.c_char_p), ("created_wall_ms", C.c_longlong),
        ("iteration_csv", C.c_int), ("density_snapshots", C.c_int), ("snapshot_every", C.c_int),
        ("snapshot_cap", C.c_int),
    ]


def _run_info(kw):
    """RunInfo from keyword values; the arrays are kept alive on the struct."""
    ri = RunInfo()
    keep = []
    for k, v in kw.items():
        if k in ("ladder", "rung_non_convergent", "rung_non_convergent_iteration",
                 "rung_non_convergent_residual"):
            dt = np.float64 if k in ("ladder", "rung_non_convergent_residual") else np.int32
            a = np.ascontiguousarray(v, dtype=dt)
            keep.append(a)
            setattr(ri, k, a.ctypes.data)
            if k == "ladder":
                ri.n_ladder = a.size
            else:
                ri.n_rungs = a.size
        elif isinstance(v, str):
            setattr(ri, k, v.encode())
        else:
            setattr(ri, k, v)
    ri._keep = keep
    return ri


def run_info_json(**kw):
    """run_info.json text of the given RunInfo fields (names as topohip_run_info)."""
    ri = _run_info(kw)
    n = C.c_size_t(0)
    buf = C.create_string_buffer(1 << 16)
    _check(_setup().topohip_run_info_json(C.byref(ri), buf, C.c_size_t(1 << 16), C.byref(n)))
    return buf.value.decode()


def write_run_info(path, **kw):
    ri = _run_info(kw)
    _check(_setup().topohip_write_run_info(str(path).encode(), C.byref(ri)))


def format_progress(rung, rungs, iteration):
    buf = C.create_string_buffer(256)
    _check(_setup().topohip_format_progress(C.c_size_t(rung), C.c_size_t(rungs), int(iteration),
                                            buf, 256))
    return buf.value.decode()


def format_variant(vf, achieved, printed, margin, mesh):
    buf = C.create_string_buffer(4096)
    _check(_setup().topohip_format_variant(C.c_double(vf), C.c_double(achieved),
                                           C.c_double(printed), C.c_double(margin),
                                           str(mesh).encode(), buf, 4096))
    return buf.value.decode()


class SnapshotCapture:
    """SnapshotCapture: the density-snapshot cadence (every_n) and per-job cap
    with oldest eviction; rung boundaries are always kept."""

    def __init__(self, directory, every_n, cap):
        self._h = C.c_void_p()
        _check(_setup().topohip_snapshots_create(str(directory).encode(), int(every_n), int(cap),
                                                 C.byref(self._h)))

    def capture(self, problem, rung, iteration, boundary=False, density=None,
                origin=(0.0, 0.0, 0.0)):
        d = None if density is None else np.ascontiguousarray(density, dtype=np.float64).ravel()
        o = np.ascontiguousarray(origin, dtype=np.float64)
        w = C.c_int(0)
        _check(_setup().topohip_snapshots_capture(self._h, problem._h, _p(o), _p(d),
                                                  C.c_size_t(rung), int(iteration),
                                                  int(bool(boundary)), C.byref(w)))
        return bool(w.value)

    def counts(self):
        w = C.c_size_t(0)
        e = C.c_size_t(0)
        _check(_setup().topohip_snapshots_counts(self._h, C.byref(w), C.byref(e)))
        return w.value, e.value

    def close(self):
        if self._h:
            _check(_setup().topohip_snapshots_destroy(self._h))
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
