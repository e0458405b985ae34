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
                double* __restrict__ Kout, int compact) {
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
  Kout[(compact ? static_cast<long long>(blockIdx.x) : Eid) * 576 + q] = acc;
}

// Level-1 element matrix entries without a stored level-1 operator: K1_E[q]
// for a level-1 element with no inactive DOF among its 27 fine nodes is
// sum_c E_c S_c[q] (c ascending); the others come from the compact buffer the
// masked path wrote (slot = spmap[E]).
struct L1Src {
  Geom g1, g0;
  const double* E;      // fine moduli
  const double* S8;     // S_c blocks (global; kernels may stage them in LDS)
  const double* K1sp;   // compact level-1 matrices of the listed elements
  const int* spmap;     // level-1 element -> slot in K1sp or -1
};

// Level-1 stencil straight from the fine moduli: thread (h, n) forms the 3x3
// block of level-1 node n towards its stored neighbour o = 13 + h by summing
// the K1_E blocks of the elements that hold both nodes, in the order (and
// with the values) the element-matrix path would assemble them; the centre
// thread also writes the Jacobi inverse diagonal.
template <typename T>
__global__ void __launch_bounds__(256)
stencil_l1_k(L1Src src, const uint8_t* __restrict__ mask, T* __restrict__ S, T* __restrict__ dinv) {
  __shared__ double s8[8 * 576];
  for (int i = threadIdx.x; i < 8 * 576; i += blockDim.x) s8[i] = src.S8[i];
  __syncthreads();
  const Geom g = src.g1, g0 = src.g0;
  const long long nodes = g.nodes();
  const long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (t >= nodes * 14) return;
  const long long ni = t % nodes;
  const int h = static_cast<int>(t / nodes);
  const int o = 13 + h;
  const int ox = o % 3 - 1, oy = (o / 3) % 3 - 1, oz = o / 9 - 1;
  const int a = static_cast<int>(ni % g.nnx);
  const int b = static_cast<int>((ni / g.nnx) % g.nny);
  const int c = static_cast<int>(ni / (static_cast<long long>(g.nnx) * g.nny));
  const int aj = a + ox, bj = b + oy, cj = c + oz;
  double blk[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (aj >= 0 && bj >= 0 && cj >= 0 && aj < g.nnx && bj < g.nny && cj < g.nnz) {
    for (int dz = 0; dz < 2; ++dz)
      for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx) {
          const int i = a - dx, j = b - dy, k = c - dz;
          if (i < 0 || j < 0 || k < 0 || i >= g.ex || j >= g.ey || k >= g.ez) continue;
          const int bx = aj - i, by = bj - j, bz = cj - k;
          if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;
          const long long e = (static_cast<long long>(k) * g.ey + j) * g.ex + i;
          const int mi = dx + 2 * dy + 4 * dz, mj = bx + 2 * by + 4 * bz;
          const int slot = src.spmap[e];
          if (slot >= 0) {
            const double* Kq = src.K1sp + static_cast<long long>(slot) * 576;
            for (int p = 0; p < 3; ++p)
              for (int q = 0; q < 3; ++q) blk[p * 3 + q] += Kq[(3 * mi + p) * 24 + 3 * mj + q];
          } else {
            double ec[8];
#pragma unroll
            for (int cc = 0; cc < 8; ++cc)
              ec[cc] = src.E[(static_cast<long long>(2 * k + (cc >> 2)) * g0.ey + 2 * j + ((cc >> 1) & 1)) *
                                 g0.ex + 2 * i + (cc & 1)];
#pragma unroll
            for (int p = 0; p < 3; ++p)
#pragma unroll
              for (int q = 0; q < 3; ++q) {
                const int qq = (3 * mi + p) * 24 + 3 * mj + q;
                double acc = 0.0;  // (void children have E = 0 and add nothing)
#pragma unroll
                for (int cc = 0; cc < 8; ++cc) acc += ec[cc] * s8[cc * 576 + qq];
                blk[p * 3 + q] += acc;
              }
          }
        }
  }
#pragma unroll
  for (int pq = 0; pq < 9; ++pq) S[(static_cast<long long>(h) * 9 + pq) * nodes + ni] = static_cast<T>(blk[pq]);
  if (h == 0) {
    const int mk = mask ? mask[ni] : 7;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double d = blk[q * 3 + q];
      dinv[ni * 3 + q] = ((mk >> q) & 1) && d > 0.0 ? static_cast<T>(1.0 / d) : T(0);
    }
  }
}

// Level-2 element matrices straight from the fine moduli for elements with no
// inactive DOF among their 125 fine nodes: K2_E = sum_{c2, c1} E_{c2 c1} T_{c2 c1}
// with T_{c2 c1} = W_{c2}^T S_{c1} W_{c2} precomputed (64 blocks of 24 x 24),
// c2 the level-1 child and c1 its fine child. A workgroup covers GL2_EPB
// elements, thread q entry q of each: the elements' 64 moduli are staged in
// LDS once, and the thread streams its coefficients 8 at a time (one c2,
// the next 8 in flight) into GL2_EPB partial sums (c ascending per entry).
// Elements of `special` are recomputed afterwards.
constexpr int GL2_EPB = 16;
__global__ void __launch_bounds__(576)
galerkin_l2_direct_k(Geom g2, Geom g0, const double* __restrict__ E, const double* __restrict__ T64,
                     const uint8_t* __restrict__ mc, double* __restrict__ Kout) {
  __shared__ double ev[GL2_EPB][64];
  const int q = threadIdx.x;
  const long long ne = g2.elems();
  const long long E0 = static_cast<long long>(blockIdx.x) * GL2_EPB;
  if (q < 64) {  // lane c of the first wave loads modulus c of every element
    const int c2 = q / 8, c1 = q % 8;
    const int ox = 2 * (c2 & 1) + (c1 & 1), oy = 2 * ((c2 >> 1) & 1) + ((c1 >> 1) & 1);
    const int oz = 2 * (c2 >> 2) + (c1 >> 2);
#pragma unroll
    for (int i = 0; i < GL2_EPB; ++i) {
      const long long Eid = E0 + i < ne ? E0 + i : ne - 1;  // uniform
      const int I = static_cast<int>(Eid % g2.ex);
      const int J = static_cast<int>((Eid / g2.ex) % g2.ey);
      const int K = static_cast<int>(Eid / (static_cast<long long>(g2.ex) * g2.ey));
      ev[i][q] = E[(static_cast<long long>(4 * K + oz) * g0.ey + 4 * J + oy) * g0.ex + 4 * I + ox];
    }
  }
  __syncthreads();
  double acc[GL2_EPB];
#pragma unroll
  for (int i = 0; i < GL2_EPB; ++i) acc[i] = 0.0;
  double t8[8], n8[8];
#pragma unroll
  for (int c1 = 0; c1 < 8; ++c1) t8[c1] = T64[c1 * 576 + q];
#pragma unroll 1
  for (int c2 = 0; c2 < 8; ++c2) {
    if (c2 + 1 < 8) {  // next chunk in flight while this one is used
#pragma unroll
      for (int c1 = 0; c1 < 8; ++c1) n8[c1] = T64[((c2 + 1) * 8 + c1) * 576 + q];
    }
#pragma unroll
    for (int i = 0; i < GL2_EPB; ++i)
#pragma unroll
      for (int c1 = 0; c1 < 8; ++c1) acc[i] += ev[i][c2 * 8 + c1] * t8[c1];  // (void: E = 0)
#pragma unroll
    for (int c1 = 0; c1 < 8; ++c1) t8[c1] = n8[c1];
  }
  const int r = q / 24, s = q % 24;
  const int mr = r / 3, ms = s / 3;
#pragma unroll
  for (int i = 0; i < GL2_EPB; ++i) {
    const long long Eid = E0 + i;
    if (Eid >= ne) return;
    const int I = static_cast<int>(Eid % g2.ex);
    const int J = static_cast<int>((Eid / g2.ex) % g2.ey);
    const int K = static_cast<int>(Eid / (static_cast<long long>(g2.ex) * g2.ey));
    const long long nr = node_of(g2, I + (mr & 1), J + ((mr >> 1) & 1), K + (mr >> 2));
    const long long ns = node_of(g2, I + (ms & 1), J + ((ms >> 1) & 1), K + (ms >> 2));
    const int kr = mc ? mc[nr] : 7, ks = mc ? mc[ns] : 7;
    const bool keep = ((kr >> (r % 3)) & 1) && ((ks >> (s % 3)) & 1);
    Kout[Eid * 576 + q] = keep ? acc[i] : 0.0;
  }
}

// Level-2 elements of `list` by the Galerkin product over their 8 level-1
// children, whose matrices come from L1Src (no stored level-1 operator).
__global__ void __launch_bounds__(576)
galerkin_l2_special_k(Geom g2, L1Src src, const int* __restrict__ list,
                      const uint8_t* __restrict__ mc, double* __restrict__ Kout) {
  __shared__ double Ks[8][576];
  const Geom g1 = src.g1, g0 = src.g0;
  const long long Eid = list[blockIdx.x];
  const int I = static_cast<int>(Eid % g2.ex);
  const int J = static_cast<int>((Eid / g2.ex) % g2.ey);
  const int K = static_cast<int>(Eid / (static_cast<long long>(g2.ex) * g2.ey));
  const int q = threadIdx.x;
  const int r = q / 24, s = q % 24;
  double sc[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) sc[c] = src.S8[c * 576 + q];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int fi = 2 * I + (c & 1), fj = 2 * J + ((c >> 1) & 1), fk = 2 * K + (c >> 2);
    const long long fe = (static_cast<long long>(fk) * g1.ey + fj) * g1.ex + fi;
    const int slot = src.spmap[fe];
    double v;
    if (slot >= 0) {
      v = src.K1sp[static_cast<long long>(slot) * 576 + q];
    } else {
      v = 0.0;  // (void children have E = 0 and add nothing)
#pragma unroll
      for (int cc = 0; cc < 8; ++cc)
        v += src.E[(static_cast<long long>(2 * fk + (cc >> 2)) * g0.ey + 2 * fj + ((cc >> 1) & 1)) *
                   g0.ex + 2 * fi + (cc & 1)] * sc[cc];
    }
    Ks[c][q] = v;
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
  const long long nr = node_of(g2, I + (mr & 1), J + ((mr >> 1) & 1), K + (mr >> 2));
  const long long ns = node_of(g2, I + (ms & 1), J + ((ms >> 1) & 1), K + (ms >> 2));
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
                       gc, gf, E, mask_f, mask_c, S8, Ke, special, Kout, 0);
    TOPOHIP_CHECK(hipGetLastError());
  }
}

void launch_galerkin_fine_compact(const Geom& gc, const Geom& gf, const double* E,
                                  const uint8_t* mask_f, const uint8_t* mask_c, const double* S8,
                                  const double* Ke, const int* special, int n_special,
                                  double* K1sp, hipStream_t s) {
  if (n_special <= 0) return;
  hipLaunchKernelGGL(galerkin_fine_k, dim3(static_cast<unsigned>(n_special)), dim3(576), 0, s, gc,
                     gf, E, mask_f, mask_c, S8, Ke, special, K1sp, 1);
  TOPOHIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_stencil_l1(const Geom& g1, const Geom& g0, const double* E, const double* S8,
                       const double* K1sp, const int* spmap, const uint8_t* mask1, T* S, T* dinv,
                       hipStream_t s) {
  const L1Src src{g1, g0, E, S8, K1sp, spmap};
  const long long t = g1.nodes() * 14;
  hipLaunchKernelGGL((stencil_l1_k<T>), dim3(static_cast<unsigned>((t + 255) / 256)), dim3(256), 0,
                     s, src, mask1, S, dinv);
  TOPOHIP_CHECK(hipGetLastError());
}
template void launch_stencil_l1<float>(const Geom&, const Geom&, const double*, const double*,
                                       const double*, const int*, const uint8_t*, float*, float*,
                                       hipStream_t);
template void launch_stencil_l1<double>(const Geom&, const Geom&, const double*, const double*,
                                        const double*, const int*, const uint8_t*, double*, double*,
                                        hipStream_t);

void launch_galerkin_l2(const Geom& g2, const Geom& g1, const Geom& g0, const double* E,
                        const double* T64, const double* S8, const double* K1sp, const int* spmap,
                        const uint8_t* mask2, const int* special, int n_special, double* Kout,
                        hipStream_t s) {
  const long long ne = g2.elems();
  hipLaunchKernelGGL(galerkin_l2_direct_k, dim3(static_cast<unsigned>((ne + GL2_EPB - 1) / GL2_EPB)),
                     dim3(576), 0, s, g2, g0, E, T64, mask2, Kout);
  TOPOHIP_CHECK(hipGetLastError());
  if (n_special > 0) {
    const L1Src src{g1, g0, E, S8, K1sp, spmap};
    hipLaunchKernelGGL(galerkin_l2_special_k, dim3(static_cast<unsigned>(n_special)), dim3(576), 0,
                       s, g2, src, special, mask2, Kout);
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
