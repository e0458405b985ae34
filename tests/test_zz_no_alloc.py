"""The allocation-free OC entry point (topohip_oc_update_noalloc) against the
allocating one (topohip_oc_update) through the C ABI: two calls in a row on
one GPU, the second on a smaller grid, each bit-identical to topohip_oc_update
on the same inputs (new design, active set, lambda, bisection count); and its
limits rejected with TOPOHIP_EINVAL. Device buffers for the inputs come from
the HIP runtime through ctypes."""
import ctypes as C

import numpy as np
import pytest

import topopt_amd as ta

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
    return _hip


class DevBuf:
    def __init__(self, nbytes):
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), C.c_size_t(max(1, nbytes))) == 0
        self.n = nbytes

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.n
        assert hip().hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0

    def download(self, dtype, count):
        out = np.zeros(count, dtype=dtype)
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        hip().hipFree(self.p)


def noalloc_oc(p, radius, x, dc, volfrac, move=0.2, rho_min=1e-3):
    n = x.size
    dx, ddc, dxn, dact = DevBuf(8 * n), DevBuf(8 * n), DevBuf(8 * n), DevBuf(n)
    try:
        dx.upload(x)
        ddc.upload(dc)
        nb = C.c_int(0)
        lam = C.c_double(0)
        code = ta.lib().topohip_oc_update_noalloc(
            p._h, C.c_double(radius), dx.p, ddc.p, C.c_double(volfrac), C.c_double(move),
            C.c_double(rho_min), dxn.p, dact.p, C.byref(nb), C.byref(lam))
        ta._check(code)
        return dxn.download(np.float64, n), dact.download(np.uint8, n), nb.value, lam.value
    finally:
        for b in (dx, ddc, dxn, dact):
            b.free()


@pytest.mark.parametrize("order", ["large_then_small"])
def test_noalloc_oc_matches_allocating_entry(order):
    cases = [((20, 10, 8), 1.5, 0.3, 11), ((12, 6, 5), 2.0, 0.4, 12)]
    for (nx, ny, nz), r, vf, seed in cases:
        rng = np.random.default_rng(seed)
        n = nx * ny * nz
        x = rng.uniform(1e-3, 1, n)
        dc = -rng.uniform(0, 5, n)
        with ta.Problem(nx, ny, nz) as p:
            ref = p.oc_update(r, x, dc, vf)  # set-up: also builds the filter of radius r
            out = noalloc_oc(p, r, x, dc, vf)
            again = noalloc_oc(p, r, x, dc, vf)
        print(f"noalloc {nx}x{ny}x{nz}: lambda {out[3]!r} (alloc {ref[3]!r}), "
              f"bisections {out[2]} ({ref[2]})", flush=True)
        assert np.array_equal(out[0], ref[0]) and np.array_equal(out[1], ref[1])
        assert out[2] == ref[2] and out[3] == ref[3]
        assert np.array_equal(again[0], ref[0])


def test_noalloc_oc_rejects_what_it_cannot_do():
    nx, ny, nz = 8, 4, 4
    x = np.full(nx * ny * nz, 0.3)
    dc = -np.ones(nx * ny * nz)
    with ta.Problem(nx, ny, nz) as p:
        with pytest.raises(ValueError, match="filter"):
            noalloc_oc(p, 1.5, x, dc, 0.3)  # no filter of this radius yet
    nx, ny, nz = 256, 96, 86  # 2113536 voxels > 2^21
    with ta.Problem(nx, ny, nz) as p:
        n = nx * ny * nz
        p.filter(1.5, np.zeros(n))
        with pytest.raises(ValueError, match="allocation-free limit"):
            noalloc_oc(p, 1.5, np.full(n, 0.3), -np.ones(n), 0.3)
