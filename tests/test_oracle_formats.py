"""Oracle pins for the result fields and the on-disk formats (CPU only).

hex8_stress: the reference's closed forms (core/tests/unit/test_fea.cpp:303-390).
von_mises_field: uniform-strain and rigid-motion fields on a small grid with
void voxels. float_to_half: numpy's IEEE float16 conversion for every finite
float class, plus the reference's anchors and round-trip bound
(core/tests/unit/test_observability.cpp:55-74). fnv1a64: the fingerprints a
design.bin written by the reference stores (tests/golden/design_bin_reference_
4x4x4.bin, from tools/topopt-worker/e2e/.../out/design.bin), and the published
FNV-1a 64-bit test vectors under the published offset basis (the reference's
basis, 1469598103934665603, is one digit shorter; core/src/io/design_store.cpp:78).
"""
import os
import struct

import numpy as np
import pytest

import oracle

E, NU, H = 200.0, 0.3, 1.0
X = np.array([0, 1, 1, 0, 0, 1, 1, 0], float)
Y = np.array([0, 0, 1, 1, 0, 0, 1, 1], float)
Z = np.array([0, 0, 0, 0, 1, 1, 1, 1], float)
C_ = E / ((1 + NU) * (1 - 2 * NU))
G_ = E / (2 * (1 + NU))


def test_hex8_stress_uniaxial_strain():
    a = 1e-3
    ue = np.zeros(24)
    ue[0::3] = a * X
    sig, vm = oracle.hex8_stress(E, NU, H, ue)
    sxx, syy = C_ * (1 - NU) * a, C_ * NU * a
    assert np.allclose(sig, [sxx, syy, syy, 0, 0, 0], rtol=0, atol=1e-9 * sxx)
    assert abs(vm - abs(sxx - syy)) <= 1e-9 * vm


def test_hex8_stress_pure_shear_each_plane():
    g = 2e-3
    for (comp, coord, voigt) in [(0, Y, 3), (1, Z, 4), (2, X, 5)]:
        ue = np.zeros(24)
        ue[comp::3] = g * coord
        sig, vm = oracle.hex8_stress(E, NU, H, ue)
        want = np.zeros(6)
        want[voigt] = G_ * g
        assert np.allclose(sig, want, rtol=0, atol=1e-9 * G_ * g), (comp, sig)
        assert abs(vm - np.sqrt(3) * G_ * g) <= 1e-9 * vm


def test_hex8_stress_rigid_motions_are_stress_free():
    ue = np.tile([0.7, -0.3, 1.1], 8)
    sig, vm = oracle.hex8_stress(E, NU, H, ue)
    assert np.abs(sig).max() <= 1e-9 and vm <= 1e-9
    w = np.array([0.2, -0.5, 0.3])  # infinitesimal rotation u = w x r
    r = np.stack([X, Y, Z], 1)
    ue = np.cross(w, r).ravel()
    sig, vm = oracle.hex8_stress(E, NU, H, ue, 0.3, -0.6, 0.1)
    assert np.abs(sig).max() <= 1e-12 and vm <= 1e-12


def test_hex8_stress_rejects_bad_material():
    for e, nu, h in [(0.0, 0.3, 1.0), (1.0, 0.5, 1.0), (1.0, 0.3, 0.0)]:
        with pytest.raises(RuntimeError):
            oracle.hex8_stress(e, nu, h, np.zeros(24))


def test_von_mises_field_uniform_strain_and_voids():
    nx, ny, nz, h = 4, 3, 2, 2.5
    solid = np.ones(nx * ny * nz, np.uint8)
    solid[[1, 7, 20]] = 0
    s = oracle.Setup(nx, ny, nz, h=h, solid=solid, bcs=[], loads=[])
    k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    a, b = 1e-3, -4e-4
    u = np.zeros((nz + 1, ny + 1, nx + 1, 3))
    u[..., 0] = a * i * h + 0.37       # eps_xx = a, plus a translation
    u[..., 2] = b * j * h              # gamma_yz = b
    vm = oracle.von_mises_field(s, u.ravel(), E=E, nu=NU)
    sxx, syy, tyz = C_ * (1 - NU) * a, C_ * NU * a, G_ * b
    want = np.sqrt((sxx - syy) ** 2 + 3 * tyz ** 2)
    assert np.all(vm[solid == 0] == 0)
    assert np.allclose(vm[solid == 1], want, rtol=1e-12, atol=0)


def _f16_cases():
    rng = np.random.default_rng(3)
    bits = rng.integers(0, 2**32, 200000, dtype=np.uint64).astype(np.uint32)
    f = bits.view(np.float32)
    f = f[np.isfinite(f)]
    edge = np.array([0.0, -0.0, 1.0, 0.5, 65504.0, 65519.996, 65520.0, -65520.0, 1e-8,
                     2.0 ** -24, 2.0 ** -25, -(2.0 ** -25), 2.0 ** -25 * 1.0000001,
                     1.5 * 2.0 ** -24, 2.5 * 2.0 ** -24, 6.1e-5, 6.103515625e-05,
                     1 + 2.0 ** -11, 1 + 3 * 2.0 ** -11, np.inf, -np.inf], np.float32)
    return np.concatenate([f, edge, np.linspace(0, 1, 100001, dtype=np.float32)])


def test_float_to_half_matches_ieee_round_to_nearest_even():
    f = _f16_cases()
    with np.errstate(over="ignore"):
        want = f.astype(np.float16).view(np.uint16)
    got = oracle.float_to_half(f)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, (f[bad[:5]], got[bad[:5]], want[bad[:5]])


def test_float_to_half_reference_anchors():
    v = np.arange(1001, dtype=np.float32) / np.float32(1000)
    rt = oracle.float_to_half(v).view(np.float16).astype(np.float32)
    assert np.abs(rt - v).max() <= 5e-4
    h = oracle.float_to_half(np.array([0.0, 1.0, 0.5, np.nan, -np.nan], np.float32))
    assert list(h[:3]) == [0x0000, 0x3C00, 0x3800]
    assert h[3] & 0x7FFF == 0x7E00 and h[4] & 0x7FFF == 0x7E00


def test_fnv1a64_published_vectors():
    basis = 0xCBF29CE484222325
    assert oracle.fnv1a64(b"", basis) == 0xCBF29CE484222325
    assert oracle.fnv1a64(b"a", basis) == 0xAF63DC4C8601EC8C
    assert oracle.fnv1a64(b"foobar", basis) == 0x85944171F73967E8
    assert oracle.fnv1a64(b"") == 1469598103934665603


def test_fnv1a64_matches_reference_written_design_bin():
    b = open(os.path.join(os.path.dirname(__file__), "golden",
                          "design_bin_reference_4x4x4.bin"), "rb").read()
    nx, ny, nz = struct.unpack_from("<3i", b, 4)
    count = struct.unpack_from("<i", b, 48)[0]
    at = 56
    for _ in range(count):
        fp, n = struct.unpack_from("<Qq", b, at + 80)
        assert n == nx * ny * nz
        dens = b[at + 96:at + 96 + 8 * n]
        assert oracle.fnv1a64(dens) == fp
        assert oracle.fnv1a64(dens, 0xCBF29CE484222325) != fp
        at += 96 + 8 * n
    assert at == len(b) and count == 2
