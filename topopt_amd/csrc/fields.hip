// fields.hip — per-voxel result fields and the narrowed arrays of the run
// records, formed from device-resident vectors so only the narrowed bytes
// cross PCIe.
//
// von_mises_k restates fea_von_mises_field (core/src/fea/assembly.cpp:976-996)
// with hex8_stress at the centroid (core/src/fea/hex_element.cpp:242-316). At
// the centroid every shape-function derivative is +-q along its own axis, so
// B u reduces to signed sums of the eight nodal displacements; the sums are
// accumulated in the reference's column order (the terms B leaves zero add
// nothing) and without contraction, so the field matches a plain FP64
// evaluation of the reference formula bit for bit.
#include "kernels.hpp"

namespace topohip {

namespace {

constexpr int FB = 256;

// The kernels' products and sums are rounded one at a time.
#pragma clang fp contract(off)

__global__ void __launch_bounds__(FB)
von_mises_k(Geom g, VmCoef c, const uint8_t* __restrict__ solid, const double* __restrict__ u,
            double* __restrict__ vm64, float* __restrict__ vm32) {
  const long long e = xcd_tile(blockIdx.x, gridDim.x) * FB + threadIdx.x;
  if (e >= g.elems()) return;
  double vm = 0.0;
  if (!solid || solid[e]) {
    const int i = static_cast<int>(e % g.ex);
    const int j = static_cast<int>((e / g.ex) % g.ey);
    const int k = static_cast<int>(e / (static_cast<long long>(g.ex) * g.ey));
    const long long sy = g.nnx, sz = static_cast<long long>(g.nnx) * g.nny;
    const long long n0 = k * sz + j * sy + i;
    // reference local node order: bottom face counter-clockwise, then top
    const long long nd[8] = {n0,      n0 + 1,      n0 + sy + 1,      n0 + sy,
                             n0 + sz, n0 + sz + 1, n0 + sz + sy + 1, n0 + sz + sy};
    constexpr int sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    constexpr int syg[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    constexpr int szg[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    double ux[8], uy[8], uz[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      ux[a] = u[3 * nd[a]];
      uy[a] = u[3 * nd[a] + 1];
      uz[a] = u[3 * nd[a] + 2];
    }
    double e0 = 0.0, e1 = 0.0, e2 = 0.0, e3 = 0.0, e4 = 0.0, e5 = 0.0;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const double dx = sx[a] > 0 ? c.q : -c.q;
      const double dy = syg[a] > 0 ? c.q : -c.q;
      const double dz = szg[a] > 0 ? c.q : -c.q;
      e0 = e0 + dx * ux[a];
      e1 = e1 + dy * uy[a];
      e2 = e2 + dz * uz[a];
      e3 = e3 + dy * ux[a];
      e3 = e3 + dx * uy[a];
      e4 = e4 + dz * uy[a];
      e4 = e4 + dy * uz[a];
      e5 = e5 + dz * ux[a];
      e5 = e5 + dx * uz[a];
    }
    const double s0 = (c.d0 * e0 + c.d1 * e1) + c.d1 * e2;
    const double s1 = (c.d1 * e0 + c.d0 * e1) + c.d1 * e2;
    const double s2 = (c.d1 * e0 + c.d1 * e1) + c.d0 * e2;
    const double s3 = c.gs * e3, s4 = c.gs * e4, s5 = c.gs * e5;
    const double a01 = s0 - s1, a12 = s1 - s2, a20 = s2 - s0;
    vm = __builtin_sqrt(0.5 * ((a01 * a01 + a12 * a12) + a20 * a20) +
                        3.0 * ((s3 * s3 + s4 * s4) + s5 * s5));
  }
  if (vm64) vm64[e] = vm;
  if (vm32) vm32[e] = static_cast<float>(vm);
}

__global__ void __launch_bounds__(FB)
narrow_f32_k(const double* __restrict__ in, float* __restrict__ out, long long n) {
  const long long i0 = (static_cast<long long>(blockIdx.x) * FB + threadIdx.x) * 4;
  if (i0 + 4 <= n) {
    const double2 a = *reinterpret_cast<const double2*>(in + i0);
    const double2 b = *reinterpret_cast<const double2*>(in + i0 + 2);
    *reinterpret_cast<float4*>(out + i0) =
        make_float4(static_cast<float>(a.x), static_cast<float>(a.y), static_cast<float>(b.x),
                    static_cast<float>(b.y));
  } else {
    for (long long i = i0; i < n; ++i) out[i] = static_cast<float>(in[i]);
  }
}

// float -> binary16 with round-to-nearest-even on the dropped bits, subnormals
// by shifting in the implicit bit, overflow to Inf and NaN to the quiet NaN
// (core/src/simp/observability.cpp:35-75 defines these results).
__device__ inline uint16_t f32_to_f16_bits(float v) {
  const uint32_t x = __float_as_uint(v);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const int biased = static_cast<int>((x >> 23) & 0xFFu);
  const uint32_t mant = x & 0x7FFFFFu;
  if (biased == 255) return static_cast<uint16_t>(sign | 0x7C00u | (mant ? 0x200u : 0u));
  const int e = biased - 112;
  if (e >= 31) return static_cast<uint16_t>(sign | 0x7C00u);
  if (e <= 0) {
    if (e < -10) return static_cast<uint16_t>(sign);
    const uint32_t m = mant | 0x800000u;
    const int sh = 14 - e;
    uint32_t hm = m >> sh;
    const uint32_t rem = m & ((1u << sh) - 1u), mid = 1u << (sh - 1);
    hm += (rem > mid || (rem == mid && (hm & 1u))) ? 1u : 0u;
    return static_cast<uint16_t>(sign | hm);
  }
  uint32_t hv = sign | (static_cast<uint32_t>(e) << 10) | (mant >> 13);
  const uint32_t rem = mant & 0x1FFFu;
  hv += (rem > 0x1000u || (rem == 0x1000u && (hv & 1u))) ? 1u : 0u;
  return static_cast<uint16_t>(hv);
}

__global__ void __launch_bounds__(FB)
density_f16_k(const double* __restrict__ in, uint16_t* __restrict__ out, long long n) {
  const long long i0 = (static_cast<long long>(blockIdx.x) * FB + threadIdx.x) * 4;
  if (i0 + 4 <= n) {
    const double2 a = *reinterpret_cast<const double2*>(in + i0);
    const double2 b = *reinterpret_cast<const double2*>(in + i0 + 2);
    const uint32_t lo = f32_to_f16_bits(static_cast<float>(a.x)) |
                        (static_cast<uint32_t>(f32_to_f16_bits(static_cast<float>(a.y))) << 16);
    const uint32_t hi = f32_to_f16_bits(static_cast<float>(b.x)) |
                        (static_cast<uint32_t>(f32_to_f16_bits(static_cast<float>(b.y))) << 16);
    *reinterpret_cast<uint2*>(out + i0) = make_uint2(lo, hi);
  } else {
    for (long long i = i0; i < n; ++i) out[i] = f32_to_f16_bits(static_cast<float>(in[i]));
  }
}

unsigned blocks_for(long long n, long long per_block) {
  return static_cast<unsigned>((n + per_block - 1) / per_block);
}

}  // namespace

void launch_von_mises(const Geom& g, const VmCoef& c, const uint8_t* solid, const double* u,
                      double* vm64, float* vm32, hipStream_t s) {
  const long long n = g.elems();
  if (n == 0) return;
  hipLaunchKernelGGL(von_mises_k, dim3(xcd_grid((n + FB - 1) / FB)), dim3(FB), 0, s, g, c, solid,
                     u, vm64, vm32);
  TOPOHIP_CHECK(hipGetLastError());
}

void launch_narrow_f32(const double* in, float* out, long long n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(narrow_f32_k, dim3(blocks_for(n, 4LL * FB)), dim3(FB), 0, s, in, out, n);
  TOPOHIP_CHECK(hipGetLastError());
}

void launch_density_f16(const double* in, uint16_t* out, long long n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(density_f16_k, dim3(blocks_for(n, 4LL * FB)), dim3(FB), 0, s, in, out, n);
  TOPOHIP_CHECK(hipGetLastError());
}

}  // namespace topohip
