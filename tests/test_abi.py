"""C-ABI checks that need no GPU.

* libtopopt_hip.so loads and exports every function include/topopt_hip.h
  declares;
* the ctypes mirror's structs have the C header's size and field offsets
  (a tiny C program compiled with gcc against the header prints them);
* host-side validation reaches the caller as the reference's exception
  taxonomy (std::invalid_argument -> TOPOHIP_EINVAL / ValueError) before any
  device call is made.
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import topopt_amd as ta

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "topopt_hip.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    return sorted(set(re.findall(r"\b(topohip_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = _declared_functions()
    for must in ["topohip_problem_create", "topohip_problem_destroy", "topohip_set_loads",
                 "topohip_ku_apply_f32", "topohip_ku_apply_f64", "topohip_solve",
                 "topohip_simp_compliance", "topohip_filter", "topohip_oc_update",
                 "topohip_optimize", "topohip_server_create", "topohip_server_submit",
                 "topohip_server_poll", "topohip_server_destroy", "topohip_last_error"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(ta.LIB_PATH)
    missing = [n for n in _declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", ta.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (topohip_[a-z0-9_]+)\b", out))
    assert set(_declared_functions()) <= exported


_LAYOUT_C = r"""
#include <stdio.h>
#include <stddef.h>
#include "topopt_hip.h"
#define F(T, f) printf(#T " " #f " %zu\n", offsetof(T, f));
#define S(T) printf(#T " sizeof %zu\n", sizeof(T));
int main(void) {
  S(topohip_cginfo) F(topohip_cginfo, converged) F(topohip_cginfo, residual)
  F(topohip_cginfo, matvecs) F(topohip_cginfo, t_build_ms) F(topohip_cginfo, t_total_ms)
  F(topohip_cginfo, residual_replacements) F(topohip_cginfo, coarse_dofs)
  S(topohip_options) F(topohip_options, volfrac) F(topohip_options, cg_tol)
  F(topohip_options, max_iterations) F(topohip_options, cg_max_iterations)
  F(topohip_options, pc_p0) F(topohip_options, pc_per_stage)
  S(topohip_result) F(topohip_result, compliance) F(topohip_result, t_oc_ms)
  F(topohip_result, cg_iterations_total) F(topohip_result, cancelled)
  F(topohip_result, deadline_exceeded)
  S(topohip_job) F(topohip_job, volfrac) F(topohip_job, iterations) F(topohip_job, precision)
  S(topohip_job_result) F(topohip_job_result, compliance) F(topohip_job_result, t_end_ms)
  F(topohip_job_result, status)
  return 0;
}
"""


def test_struct_layout_matches_header(tmp_path):
    src = tmp_path / "layout.c"
    exe = tmp_path / "layout"
    src.write_text(_LAYOUT_C)
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    py = {"topohip_cginfo": ta.CgInfo, "topohip_options": ta.Options, "topohip_result": ta.Result,
          "topohip_job": ta.Job, "topohip_job_result": ta.JobResult}
    n = 0
    for line in filter(None, lines):
        tname, field, value = line.split()
        cls = py[tname]
        got = C.sizeof(cls) if field == "sizeof" else getattr(cls, field).offset
        assert got == int(value), (tname, field, got, value)
        n += 1
    assert n == 29


def test_version_string():
    assert "gfx950" in ta.version()


@pytest.mark.parametrize("kw,fragment", [
    (dict(nx=0, ny=4, nz=4), "grid dimensions"),
    (dict(nx=4, ny=4, nz=4, poisson=0.5), "poisson"),
    (dict(nx=4, ny=4, nz=4, h=0.0), "element_size"),
    (dict(nx=4, ny=4, nz=4, bcs=[(125, 0, 0.0)]), "BC index out of range"),
    (dict(nx=4, ny=4, nz=4, bcs=[(0, 3, 0.0)]), "BC index out of range"),
    (dict(nx=4, ny=4, nz=4, precision=7), "precision"),
])
def test_invalid_arguments_raise_before_device_work(kw, fragment):
    nx, ny, nz = kw.pop("nx"), kw.pop("ny"), kw.pop("nz")
    with pytest.raises(ValueError, match=fragment):
        ta.Problem(nx, ny, nz, **kw)


def test_solid_tag_size_checked():
    with pytest.raises(ValueError):
        ta.Problem(4, 4, 4, solid=np.ones(10, dtype=np.uint8))


def test_create_with_null_out_is_einval():
    code = ta.lib().topohip_problem_create(2, 2, 2, C.c_double(1.0), C.c_double(0.3), None, 0,
                                           None, None, None, 0, None)
    assert code == ta.EINVAL
    assert "out" in ta._err()


def test_job_footprint_grows_with_grid():
    f = ta.lib().topohip_job_footprint
    small = f(C.byref(ta.Job(48, 16, 16, 0, 0.3, 1.5, 3.0, 0.2, 1e-6, 60, 0)))
    large = f(C.byref(ta.Job(192, 64, 64, 1, 0.3, 1.5, 3.0, 0.2, 1e-6, 60, 0)))
    f64 = f(C.byref(ta.Job(192, 64, 64, 1, 0.3, 1.5, 3.0, 0.2, 1e-6, 60, 1)))
    assert 0 < small < large
    # 192x64x64 holds ~2.4 M DOF: at least ~10 FP64 vectors of that length
    assert large > 10 * 8 * 2_446_275
    assert f64 > 0
