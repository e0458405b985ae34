// galerkin.hip — multigrid hierarchy build on the device.
//
// The reference forms A1 = P0^T A0 P0 element-locally without assembling A0
// (galerkin_a1_from_prolong, core/src/fea/multigrid.cpp:1543-1868) and deeper
// levels by sparse Galerkin products (build_hierarchy :557-716). On a
// structured grid with 2x vertex coarsening every fine element lies in exactly
// one coarse element and its nodes interpolate only from that coarse
// element's 8 nodes, so every level is the sum of per-coarse-element 24x24
// blocks:
//   K1_E = Mc ( sum_c E_c W_c^T (Mf Ke Mf) W_c ) Mc
//   K_E  = Mc ( sum_c W_c^T K_c W_c ) Mc            (levels >= 2)
// with W_c the trilinear weights of child c and Mf/Mc the active-DOF masks
// (rows/columns the reference's P drops). Children whose nodes are all active
// use the 8 precomputed S_c = W_c^T Ke W_c blocks (the reference's colour
// cache, :1760-1800). Element matrices are kept element-major in FP64; the
// levels the V-cycle applies are then assembled into node stencils
// (stencil.hip) and the coarsest into plane blocks (coarse.hip).
#include "kernels.hpp"

namespace topohip {

namespace {

// 1-D trilinear weight of coarse node bit b for fine position p in {0,1,2}
// (axis_weights, core/src/fea/multigrid.cpp:494-506, within one coarse cell).
__device__ inline double w1(int p, int b) {
  if (p == 1) return 0.5;
  return (p == 0) == (b == 0) ? 1.0 : 0.0;
}
// W_c(fine local dof a, coarse local dof s) for child c (bit-ordered).
__device__ inline double wc(int c, int a, int s) {
  if ((a % 3) != (s % 3)) return 0.0;
  const int ma = a / 3, ms = s / 3;
  return w1(((c >> 0) & 1) + ((ma >> 0) & 1), (ms >> 0) & 1) *
         w1(((c >> 1) & 1) + ((ma >> 1) & 1), (ms >> 1) & 1) *
         w1(((c >> 2) & 1) + ((ma >> 2) & 1), (ms >> 2) & 1);
}

__device__ inline long long node_of(const Geom& g, int a, int b, int c) {
  return (static_cast<long long>(c) * g.nny + b) * g.nnx + a;
}

template <typename T>
__global__ void diag_fine_k(Geom g, const double* __restrict__ E, const double* __restrict__ kd,
                            const uint8_t* __restrict__ mask, T* __restrict__ dinv) {
  const long long n = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (n >= g.nodes()) return;
  const int a = static_cast<int>(n % g.nnx);
  const int b = static_cast<int>((n / g.nnx) % g.nny);
  const int c = static_cast<int>(n / (static_cast<long long>(g.nnx) * g.nny));
  double d[3] = {0.0, 0.0, 0.0};
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) {
        const int i = a - dx, j = b - dy, k = c - dz;
        if (i < 0 || j < 0 || k < 0 || i >= g.ex || j >= g.ey || k >= g.ez) continue;
        const double e = E[(static_cast<long long>(k) * g.ey + j) * g.ex + i];
        if (e == 0.0) continue;
        const int m = dx + 2 * dy + 4 * dz;
        for (int q = 0; q < 3; ++q) d[q] += e * kd[3 * m + q];
      }
  const int mk = mask ? mask[n] : 7;
  for (int q = 0; q < 3; ++q)
    dinv[n * 3 + q] = ((mk >> q) & 1) && d[q] > 0.0 ? static_cast<T>(1.0 / d[q]) : T(0);
}

// Coarse elements none of whose 27 fine nodes has an inactive DOF: every
// child uses its precomputed S_c, so K1_E[q] = sum_c E_c S_c[q] (c ascending,
// void children skipped) needs no LDS and no barrier. One thread per entry q,
// GF_EPB coarse elements per block; the S_c column stays in registers.
constexpr int GF_EPB = 8;
__global__ void __launch_bounds__(576)
galerkin_fine_fast_k(Geom gc, Geom gf, const double* __restrict__ E,
                     const uint8_t* __restrict__ mc, const double* __restrict__ S8,
                     double* __restrict__ Kout) {
  const int q = threadIdx.x;
  double sc[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) sc[c] = S8[c * 576 + q];
  const int r = q / 24, s = q % 24;
  const int mr = r / 3, ms = s / 3;
  const long long ne = gc.elems();
  for (int i = 0; i < GF_EPB; ++i) {
    const long long Eid = static_cast<long long>(blockIdx.x) * GF_EPB + i;
    if (Eid >= ne) return;
    const int I = static_cast<int>(Eid % gc.ex);
    const int J = static_cast<int>((Eid / gc.ex) % gc.ey);
    const int K = static_cast<int>(Eid / (static_cast<long long>(gc.ex) * gc.ey));
    double acc = 0.0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int fi = 2 * I + (c & 1), fj = 2 * J + ((c >> 1) & 1), fk = 2 * K + (c >> 2);
      const double e = E[(static_cast<long long>(fk) * gf.ey + fj) * gf.ex + fi];
      if (e != 0.0) acc += e * sc[c];
    }
    const long long nr = node_of(gc, I + (mr & 1), J + ((mr >> 1) & 1), K + (mr >> 2));
    const long long ns = node_of(gc, I + (ms & 1), J + ((ms >> 1) & 1), K + (ms >> 2));
    const int kr = mc ? mc[nr] : 7, ks = mc ? mc[ns] : 7;
    if (!((kr >> (r % 3)) & 1) || !((ks >> (s % 3)) & 1)) acc = 0.0;
    Kout[Eid * 576 + q] = acc;
  }
}

// One block (576 threads) per coarse element of `list` (or per coarse element
// when list is null): children touching an inactive fine DOF form
// W_c^T (Mf Ke Mf) W_c explicitly.
__global__ void __launch_bounds__(576)
galerkin_fine_k(Geom gc, Geom gf, const double* __restrict__ E, const uint8_t* __restrict__ mf,
                const uint8_t* __restrict__ mc, const double* __restrict__ S8,
                const double* __restrict__ Ke, const int* __restrict__ list,
                double* __restrict__ Kout) {
  __shared__ double T[576];
  __shared__ int s_mask[8];
  const long long Eid = list ? list[blockIdx.x] : blockIdx.x;
  const int I = static_cast<int>(Eid % gc.ex);
  const int J = static_cast<int>((Eid / gc.ex) % gc.ey);
  const int K = static_cast<int>(Eid / (static_cast<long long>(gc.ex) * gc.ey));
  const int q = threadIdx.x;
  const int r = q / 24, s = q % 24;
  double acc = 0.0;
  for (int c = 0; c < 8; ++c) {
    const int fi = 2 * I + (c & 1), fj = 2 * J + ((c >> 1) & 1), fk = 2 * K + (c >> 2);
    const double e = E[(static_cast<long long>(fk) * gf.ey + fj) * gf.ex + fi];
    if (e == 0.0) continue;
    if (q < 8) {
      const long long fn = node_of(gf, fi + (q & 1), fj + ((q >> 1) & 1), fk + (q >> 2));
      s_mask[q] = mf ? mf[fn] : 7;
    }
    __syncthreads();
    bool masked = false;
    for (int m = 0; m < 8; ++m) masked |= s_mask[m] != 7;
    if (!masked) {
      acc += e * S8[c * 576 + q];
    } else {
      {
        const int aa = r, ss = s;  // T[a][s] = sum_b Ke[a][b] M_b W(b, s)
        double t = 0.0;
        for (int mb = 0; mb < 8; ++mb) {
          const int bb = 3 * mb + (ss % 3);
          if (!((s_mask[mb] >> (ss % 3)) & 1)) continue;
          t += Ke[aa * 24 + bb] * wc(c, bb, ss);
        }
        T[aa * 24 + ss] = t;
      }
      __syncthreads();
      double sv = 0.0;
      for (int ma = 0; ma < 8; ++ma) {
        const int aa = 3 * ma + (r % 3);
        if (!((s_mask[ma] >> (r % 3)) & 1)) continue;
        sv += wc(c, aa, r) * T[aa * 24 + s];
      }
      acc += e * sv;
    }
    __syncthreads();
  }
  const int mr = r / 3, ms = s / 3;
  const long long nr = node_of(gc, I + (mr & 1), J + ((mr >> 1) & 1), K + (mr >> 2));
  const long long ns = node_of(gc, I + (ms & 1), J + ((ms >> 1) & 1), K + (ms >> 2));
  const int kr = mc ? mc[nr] : 7, ks = mc ? mc[ns] : 7;
  if (!((kr >> (r % 3)) & 1) || !((ks >> (s % 3)) & 1)) acc = 0.0;
  Kout[Eid * 576 + q] = acc;
}

// One block (576 threads) per coarse element; the 8 child blocks are staged
// in LDS together so a coarse element costs three barriers.
__global__ void __launch_bounds__(576)
galerkin_coarse_k(Geom gc, Geom gf, const double* __restrict__ Kf,
                  const uint8_t* __restrict__ mc, double* __restrict__ Kout) {
  __shared__ double Ks[8][576];
  const long long Eid = blockIdx.x;
  const int I = static_cast<int>(Eid % gc.ex);
  const int J = static_cast<int>((Eid / gc.ex) % gc.ey);
  const int K = static_cast<int>(Eid / (static_cast<long long>(gc.ex) * gc.ey));
  const int q = threadIdx.x;
  const int r = q / 24, s = q % 24;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int fi = 2 * I + (c & 1), fj = 2 * J + ((c >> 1) & 1), fk = 2 * K + (c >> 2);
    const long long fe = (static_cast<long long>(fk) * gf.ey + fj) * gf.ex + fi;
    Ks[c][q] = Kf[fe * 576 + q];
  }
  __syncthreads();
  double t[8];  // T_c[r][s] = sum_b K_c[r][b] W_c(b, s)
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    double v = 0.0;
    for (int mb = 0; mb < 8; ++mb) {
      const int bb = 3 * mb + (s % 3);
      v += Ks[c][r * 24 + bb] * wc(c, bb, s);
    }
    t[c] = v;
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 8; ++c) Ks[c][q] = t[c];
  __syncthreads();
  double acc = 0.0;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    double sv = 0.0;
    for (int ma = 0; ma < 8; ++ma) {
      const int aa = 3 * ma + (r % 3);
      sv += wc(c, aa, r) * Ks[c][aa * 24 + s];
    }
    acc += sv;
  }
  const int mr = r / 3, ms = s / 3;
  const long long nr = node_of(gc, I + (mr & 1), J + ((mr >> 1) & 1), K + (mr >> 2));
  const long long ns = node_of(gc, I + (ms & 1), J + ((ms >> 1) & 1), K + (ms >> 2));
  const int kr = mc ? mc[nr] : 7, ks = mc ? mc[ns] : 7;
  if (!((kr >> (r % 3)) & 1) || !((ks >> (s % 3)) & 1)) acc = 0.0;
  Kout[Eid * 576 + q] = acc;
}

template <typename T>
__global__ void diag_dense_k(Geom g, const double* __restrict__ K, const uint8_t* __restrict__ mask,
                             T* __restrict__ dinv) {
  const long long n = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (n >= g.nodes()) return;
  const int a = static_cast<int>(n % g.nnx);
  const int b = static_cast<int>((n / g.nnx) % g.nny);
  const int c = static_cast<int>(n / (static_cast<long long>(g.nnx) * g.nny));
  double d[3] = {0.0, 0.0, 0.0};
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) {
        const int i = a - dx, j = b - dy, k = c - dz;
        if (i < 0 || j < 0 || k < 0 || i >= g.ex || j >= g.ey || k >= g.ez) continue;
        const long long e = (static_cast<long long>(k) * g.ey + j) * g.ex + i;
        const int m = dx + 2 * dy + 4 * dz;
        for (int q = 0; q < 3; ++q) d[q] += K[e * 576 + (3 * m + q) * 24 + 3 * m + q];
      }
  const int mk = mask ? mask[n] : 7;
  for (int q = 0; q < 3; ++q)
    dinv[n * 3 + q] = ((mk >> q) & 1) && d[q] > 0.0 ? static_cast<T>(1.0 / d[q]) : T(0);
}

}  // namespace

template <typename T>
void launch_diag_fine(const Geom& g, const double* E, const double* kd, const uint8_t* mask, T* dinv,
                      hipStream_t s) {
  const long long n = g.nodes();
  hipLaunchKernelGGL((diag_fine_k<T>), dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                     s, g, E, kd, mask, dinv);
  TOPOHIP_CHECK(hipGetLastError());
}
template void launch_diag_fine<float>(const Geom&, const double*, const double*, const uint8_t*,
                                      float*, hipStream_t);
template void launch_diag_fine<double>(const Geom&, const double*, const double*, const uint8_t*,
                                       double*, hipStream_t);

void launch_galerkin_fine(const Geom& gc, const Geom& gf, const double* E, const uint8_t* mask_f,
                          const uint8_t* mask_c, const double* S8, const double* Ke,
                          const int* special, int n_special, double* Kout, hipStream_t s) {
  const long long ne = gc.elems();
  hipLaunchKernelGGL(galerkin_fine_fast_k, dim3(static_cast<unsigned>((ne + GF_EPB - 1) / GF_EPB)),
                     dim3(576), 0, s, gc, gf, E, mask_c, S8, Kout);
  TOPOHIP_CHECK(hipGetLastError());
  // Elements with an inactive fine DOF are recomputed (after the fast pass on
  // the same stream) by the masked path.
  if (n_special > 0) {
    hipLaunchKernelGGL(galerkin_fine_k, dim3(static_cast<unsigned>(n_special)), dim3(576), 0, s,
                       gc, gf, E, mask_f, mask_c, S8, Ke, special, Kout);
    TOPOHIP_CHECK(hipGetLastError());
  }
}

void launch_galerkin_coarse(const Geom& gc, const Geom& gf, const double* Kf, const uint8_t* mask_c,
                            double* Kout, hipStream_t s) {
  hipLaunchKernelGGL(galerkin_coarse_k, dim3(static_cast<unsigned>(gc.elems())), dim3(576), 0, s,
                     gc, gf, Kf, mask_c, Kout);
  TOPOHIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_diag_dense(const Geom& g, const double* K, const uint8_t* mask, T* dinv, hipStream_t s) {
  const long long n = g.nodes();
  hipLaunchKernelGGL((diag_dense_k<T>), dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                     s, g, K, mask, dinv);
  TOPOHIP_CHECK(hipGetLastError());
}
template void launch_diag_dense<float>(const Geom&, const double*, const uint8_t*, float*,
                                       hipStream_t);
template void launch_diag_dense<double>(const Geom&, const double*, const uint8_t*, double*,
                                        hipStream_t);

}  // namespace topohip
