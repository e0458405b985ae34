// ku.hip — matrix-free hex8 operator apply on a structured level.
//
// Replaces the colour-ordered element loop of core/src/fea/matfree.cpp:75-123,
// 535-553 (gather 24 / 24x24 column AXPY / scatter 24). Design for gfx950:
//
//  * A workgroup owns a (TX-1) x (TY-1) tile of nodes in x-y and marches a
//    z-chunk of node planes. Each lane computes one element column, so the
//    block computes a TX x TY element tile (one halo element on the low x and
//    y side) and TZ+1 element layers; node planes are staged in LDS and
//    double-buffered along z.
//  * Node-owned deterministic accumulation: every element writes its 8 nodal
//    results to LDS and the owning lane sums its (at most) 8 element
//    contributions in a fixed order. No atomics, bit-identical run to run.
//  * Fine level: the 24x24 element stiffness is applied in the Hadamard
//    (parity-mode) basis, where it has 45 non-zeros instead of 576: forward
//    butterfly (72 add/sub), 45 FMA, scale by E_e, inverse butterfly.
//  * Coarse levels: a dense per-element 24x24 (Galerkin product), read
//    entry-major so each load instruction is coalesced across the wave.
#include "kernels.hpp"

namespace topohip {

namespace {

constexpr int NT = 256;

template <typename T>
__device__ inline void hadamard3(T (&a)[24]) {
#pragma unroll
  for (int bit = 1; bit < 8; bit <<= 1) {
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      if (m & bit) continue;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T p = a[3 * m + c];
        const T q = a[3 * (m | bit) + c];
        a[3 * m + c] = p + q;
        a[3 * (m | bit) + c] = p - q;
      }
    }
  }
}

// y = E * Ke * u in bit-ordered local DOFs via the modal basis.
template <typename T>
__device__ inline void modal_apply(T (&u)[24], T (&v)[24], T e, const Modal<T>& mc) {
  hadamard3(u);
#pragma unroll
  for (int s = 3; s < 24; ++s) u[s] *= e;
#pragma unroll
  for (int r = 0; r < 24; ++r) v[r] = T(0);
#define TOPOHIP_X(k, r, s) v[r] += mc.c[k] * u[s];
  TOPOHIP_MODAL_PATTERN(TOPOHIP_X)
#undef TOPOHIP_X
  hadamard3(v);
}

template <typename T>
__device__ inline void dense_apply(const T (&u)[24], T (&v)[24], const T* __restrict__ K,
                                   long long ne, long long e) {
#pragma unroll
  for (int r = 0; r < 24; ++r) {
    T s = T(0);
#pragma unroll
    for (int c = 0; c < 24; ++c) s += K[(r * 24 + c) * ne + e] * u[c];
    v[r] = s;
  }
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

template <typename T, int TX, bool DENSE, int EPI, bool DOT>
__global__ void __launch_bounds__(NT) ku_tiled(KuOp<T> a) {
  constexpr int TY = NT / TX;
  constexpr int PLANE = (TX + 1) * (TY + 1) * 3;
  constexpr int NPF = (PLANE + NT - 1) / NT;  // prefetch registers per lane
  if (a.done && *a.done) return;
  __shared__ T s_u[2][PLANE];
  __shared__ T s_c[8][NT * 3];
  __shared__ double s_red[NT / 64];

  const Geom g = a.g;
  const int t = threadIdx.x;
  const int tx = t % TX, ty = t / TX;
  const int i0 = blockIdx.x * (TX - 1), j0 = blockIdx.y * (TY - 1), k0 = blockIdx.z * a.tz;
  const int ei = i0 - 1 + tx, ej = j0 - 1 + ty;
  const int na = i0 + tx, nb = j0 + ty;
  const bool owner = tx < TX - 1 && ty < TY - 1 && na < g.nnx && nb < g.nny;
  const bool exy = ei >= 0 && ei < g.ex && ej >= 0 && ej < g.ey;
  const uint8_t* __restrict__ mask = a.mask;
  const T* __restrict__ x = a.x;

  // Node plane c is fetched into registers one layer ahead (software
  // pipelining) and written to LDS at the top of the next layer.
  T pf[NPF];
  auto fetch_plane = [&](int c) {
#pragma unroll
    for (int i = 0; i < NPF; ++i) {
      const int q = t + i * NT;
      T v = T(0);
      if (q < PLANE) {
        const int comp = q % 3;
        const int n = q / 3;
        const int lx = n % (TX + 1), ly = n / (TX + 1);
        const int ga = i0 - 1 + lx, gb = j0 - 1 + ly;
        if (c >= 0 && c < g.nnz && ga >= 0 && ga < g.nnx && gb >= 0 && gb < g.nny) {
          const long long node = (static_cast<long long>(c) * g.nny + gb) * g.nnx + ga;
          const bool kept = !mask || ((mask[node] >> comp) & 1);
          if (kept) v = x[node * 3 + comp];
        }
      }
      pf[i] = v;
    }
  };
  auto store_plane = [&](int buf) {
#pragma unroll
    for (int i = 0; i < NPF; ++i) {
      const int q = t + i * NT;
      if (q < PLANE) s_u[buf][q] = pf[i];
    }
  };

  int cur = 0;
  fetch_plane(k0 - 1);
  store_plane(cur);
  fetch_plane(k0);
  T acc[3] = {T(0), T(0), T(0)};
  T xown[3] = {T(0), T(0), T(0)};
  double dotp = 0.0;
  const int kend = k0 + a.tz;
  for (int k = k0 - 1; k < kend; ++k) {
    const int nxt = cur ^ 1;
    store_plane(nxt);  // plane k + 1
    if (k + 1 < kend) fetch_plane(k + 2);
    __syncthreads();
    if (owner) {
#pragma unroll
      for (int c = 0; c < 3; ++c) xown[c] = s_u[cur][((ty + 1) * (TX + 1) + (tx + 1)) * 3 + c];
    }
    T yl[24];
    bool has = exy && k >= 0 && k < g.ez;
    const long long eidx = (static_cast<long long>(k) * g.ey + ej) * g.ex + ei;
    T e = T(0);
    if (!DENSE && has) {
      e = a.E[eidx];
      has = e != T(0);
    }
    if (has) {
      T ul[24];
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const int bx = m & 1, by = (m >> 1) & 1, bz = m >> 2;
        const T* pl = bz ? s_u[nxt] : s_u[cur];
        const int base = ((ty + by) * (TX + 1) + (tx + bx)) * 3;
#pragma unroll
        for (int c = 0; c < 3; ++c) ul[3 * m + c] = pl[base + c];
      }
      if constexpr (DENSE) {
        dense_apply(ul, yl, a.Kel, g.elems(), eidx);
      } else {
        modal_apply(ul, yl, e, a.mc);
      }
    } else {
#pragma unroll
      for (int r = 0; r < 24; ++r) yl[r] = T(0);
    }
#pragma unroll
    for (int m = 0; m < 8; ++m)
#pragma unroll
      for (int c = 0; c < 3; ++c) s_c[m][t * 3 + c] = yl[3 * m + c];
    __syncthreads();
    if (owner) {
      // Node (na, nb) is corner m of element (na-1+bx', nb-1+by') with
      // bit x = 1 - (element offset along x), etc. Fixed summation order.
      const int t00 = t, t10 = t + 1, t01 = t + TX, t11 = t + TX + 1;
      T bot[3], top[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        bot[c] = ((s_c[3][t00 * 3 + c] + s_c[2][t10 * 3 + c]) + s_c[1][t01 * 3 + c]) +
                 s_c[0][t11 * 3 + c];
        top[c] = ((s_c[7][t00 * 3 + c] + s_c[6][t10 * 3 + c]) + s_c[5][t01 * 3 + c]) +
                 s_c[4][t11 * 3 + c];
      }
      if (k >= k0 && k < g.nnz) {
        const long long node = (static_cast<long long>(k) * g.nny + nb) * g.nnx + na;
        const int mk = mask ? mask[node] : 7;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const T ax = acc[c] + bot[c];
          const bool kept = (mk >> c) & 1;
          T out = T(0);
          if (kept) {
            if constexpr (EPI == EPI_AX) {
              out = ax;
            } else if constexpr (EPI == EPI_RES) {
              out = a.b[node * 3 + c] - ax;
            } else {
              out = xown[c] + a.omega * a.dinv[node * 3 + c] * (a.b[node * 3 + c] - ax);
            }
            if constexpr (DOT) dotp += static_cast<double>(xown[c]) * static_cast<double>(ax);
          }
          a.y[node * 3 + c] = out;
        }
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[c] = top[c];
    }
    cur = nxt;
  }
  if constexpr (DOT) {
    dotp = wave_sum(dotp);
    if ((t & 63) == 0) s_red[t >> 6] = dotp;
    __syncthreads();
    if (t == 0) {
      double s = 0.0;
      for (int w = 0; w < NT / 64; ++w) s += s_red[w];
      const int blin = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
      a.partials[blin] = s;
    }
  }
}

template <typename T, int TX, bool DENSE>
void dispatch_tx(const KuOp<T>& op, hipStream_t s) {
  constexpr int TY = NT / TX;
  const dim3 grid((op.g.nnx + TX - 2) / (TX - 1), (op.g.nny + TY - 2) / (TY - 1),
                  (op.g.nnz + op.tz - 1) / op.tz);
  const bool dot = op.partials != nullptr;
#define TOPOHIP_L(E, D) hipLaunchKernelGGL((ku_tiled<T, TX, DENSE, E, D>), grid, dim3(NT), 0, s, op)
  if (op.epi == EPI_AX) {
    if (dot) TOPOHIP_L(EPI_AX, true); else TOPOHIP_L(EPI_AX, false);
  } else if (op.epi == EPI_RES) {
    if (dot) TOPOHIP_L(EPI_RES, true); else TOPOHIP_L(EPI_RES, false);
  } else {
    if (dot) TOPOHIP_L(EPI_JAC, true); else TOPOHIP_L(EPI_JAC, false);
  }
#undef TOPOHIP_L
  TOPOHIP_CHECK(hipGetLastError());
}

template <typename T, bool DENSE>
void dispatch(const KuOp<T>& op, hipStream_t s) {
  if (op.tx == 32) dispatch_tx<T, 32, DENSE>(op, s);
  else if (op.tx == 64) dispatch_tx<T, 64, DENSE>(op, s);
  else dispatch_tx<T, 16, DENSE>(op, s);
}

}  // namespace

int ku_blocks(const Geom& g, int tz, int tx) {
  if (tx != 32 && tx != 64) tx = 16;
  const int ty = NT / tx;
  return ((g.nnx + tx - 2) / (tx - 1)) * ((g.nny + ty - 2) / (ty - 1)) * ((g.nnz + tz - 1) / tz);
}

template <typename T>
void launch_ku_fine(const KuOp<T>& op, hipStream_t s) {
  dispatch<T, false>(op, s);
}
template <typename T>
void launch_ku_dense(const KuOp<T>& op, hipStream_t s) {
  dispatch<T, true>(op, s);
}

template void launch_ku_fine<float>(const KuOp<float>&, hipStream_t);
template void launch_ku_fine<double>(const KuOp<double>&, hipStream_t);
template void launch_ku_dense<float>(const KuOp<float>&, hipStream_t);
template void launch_ku_dense<double>(const KuOp<double>&, hipStream_t);

}  // namespace topohip
