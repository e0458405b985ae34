// stencil.hip — assembled coarse-level operators as 27-point node stencils.
//
// Below the fine level the Galerkin operators (A_l = P^T A_{l-1} P,
// core/src/fea/multigrid.cpp:557-716, 1543-1868) are stored the way the
// structured grid sees them: each node couples to its 27 neighbours through a
// 3x3 block. The blocks are assembled once per hierarchy build from the
// per-coarse-element Galerkin matrices (fixed element order, no atomics) and
// applied by a gather whose coefficient loads are coalesced (entry-major
// layout). The operator is symmetric, so only the centre block
// and the 13 blocks towards lexicographically larger neighbours are stored
// (offset index o = 13 + h, h = 0..13); the block towards a smaller neighbour
// m is the transpose of the block m stores towards this node. That halves the
// stored coefficients, and the transposed reads of a neighbour's blocks mostly
// hit the caches.
#include "kernels.hpp"

namespace topohip {

namespace {

constexpr int SH = 14;  // stored blocks per node

template <typename T>
__global__ void assemble_stencil_k(Geom g, const double* __restrict__ K, T* __restrict__ S) {
  const long long nodes = g.nodes();
  const long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (t >= nodes * SH) return;
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
          for (int p = 0; p < 3; ++p)
            for (int q = 0; q < 3; ++q) blk[p * 3 + q] += K[e * 576 + (3 * mi + p) * 24 + 3 * mj + q];
        }
  }
  for (int pq = 0; pq < 9; ++pq) S[(static_cast<long long>(h) * 9 + pq) * nodes + ni] = static_cast<T>(blk[pq]);
}

// Apply: a workgroup of 9 waves covers 64 consecutive nodes; wave j gathers
// the 3 neighbours (dx = -1, 0, 1) of row j = (dz + 1) * 3 + (dy + 1) for its
// lane's node, the 9 row sums meet in LDS and 3 waves finish the 64 x 3 outputs
// in a fixed order. Neighbour indices are clamped into the grid and invalid
// terms weighted by zero, so every load is used unconditionally and a wave
// issues them all up front instead of behind a branch each.
constexpr int SA_NODES = 64;
constexpr int SA_ROWS = 9;

template <typename T, int EPI>
__global__ void __launch_bounds__(SA_NODES * SA_ROWS)
stencil_apply_k(Geom g, const T* __restrict__ S, const uint8_t* __restrict__ mask,
                const T* __restrict__ x, T* __restrict__ y, const T* __restrict__ b,
                const T* __restrict__ dinv, T omega, const int* done) {
  __shared__ T s_acc[SA_ROWS][3][SA_NODES];
  const int rc = blockIdx.y;
  if (case_done(done, rc)) return;
  const long long nodes = g.nodes();
  x += rc * 3 * nodes;
  y += rc * 3 * nodes;
  if (b) b += rc * 3 * nodes;
  const int lane = threadIdx.x % SA_NODES, row = threadIdx.x / SA_NODES;
  const long long n0 = xcd_tile(blockIdx.x, gridDim.x) * SA_NODES;
  const long long n = n0 + lane;
  const bool live = n < nodes;
  const long long nn = live ? n : nodes - 1;
  const int a = static_cast<int>(nn % g.nnx);
  const int bb = static_cast<int>((nn / g.nnx) % g.nny);
  const int c = static_cast<int>(nn / (static_cast<long long>(g.nnx) * g.nny));
  const int oy = row % 3 - 1, oz = row / 3 - 1;
  const int bj = bb + oy, cj = c + oz;
  const bool rowok = bj >= 0 && cj >= 0 && bj < g.nny && cj < g.nnz;
  long long m[3];
  int mm[3];
  bool ok[3];
#pragma unroll
  for (int dx = 0; dx < 3; ++dx) {
    const int aj = a + dx - 1;
    ok[dx] = rowok && aj >= 0 && aj < g.nnx;
    m[dx] = ok[dx] ? (static_cast<long long>(cj) * g.nny + bj) * g.nnx + aj : nn;
    mm[dx] = 7;
  }
  if (mask) {
#pragma unroll
    for (int dx = 0; dx < 3; ++dx) mm[dx] = mask[m[dx]];
  }
  // The three waves that finish the outputs load their epilogue operands
  // now, so those loads overlap the row gathers instead of following the
  // barrier as a second memory round trip (except the FP64 sweep, where the
  // held operands would cost a wave per SIMD).
  constexpr bool EARLY = sizeof(T) == 4 || EPI != EPI_JAC;
  const int pe = threadIdx.x / SA_NODES;  // output component of a finishing wave
  const bool fin = threadIdx.x < 3 * SA_NODES && live;
  int mk_e = 7;
  T b_e = T(0), x_e = T(0), d_e = T(0);
  auto epilogue_operands = [&]() {
    if (mask) mk_e = mask[n];
    if constexpr (EPI != EPI_AX) b_e = b[n * 3 + pe];
    if constexpr (EPI == EPI_JAC) {
      x_e = x[n * 3 + pe];
      d_e = dinv[n * 3 + pe];
    }
  };
  if (EARLY && fin) epilogue_operands();
  T acc[3] = {T(0), T(0), T(0)};
#pragma unroll
  for (int dx = 0; dx < 3; ++dx) {
    const int o = row * 3 + dx;  // (dz+1)*9 + (dy+1)*3 + (dx+1)
    const int mk = ok[dx] ? mm[dx] : 0;
    T xv[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) xv[q] = x[m[dx] * 3 + q] * static_cast<T>((mk >> q) & 1);
    // own block towards m (o >= 13), else the transpose of m's block towards n
    const T* Sb = o >= 13 ? S + static_cast<long long>(o - 13) * 9 * nodes + nn
                          : S + static_cast<long long>(13 - o) * 9 * nodes + m[dx];
    T blk[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) blk[e] = Sb[e * nodes];
    if (o >= 13) {
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int q = 0; q < 3; ++q) acc[p] += blk[p * 3 + q] * xv[q];
    } else {
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int q = 0; q < 3; ++q) acc[p] += blk[q * 3 + p] * xv[q];
    }
  }
#pragma unroll
  for (int p = 0; p < 3; ++p) s_acc[row][p][lane] = acc[p];
  __syncthreads();
  if (!fin) return;
  if (!EARLY) epilogue_operands();
  const int p = pe;
  T sum = s_acc[0][p][lane];
#pragma unroll
  for (int r = 1; r < SA_ROWS; ++r) sum += s_acc[r][p][lane];
  T out = T(0);
  if ((mk_e >> p) & 1) {
    if constexpr (EPI == EPI_AX) out = sum;
    else if constexpr (EPI == EPI_RES) out = b_e - sum;
    else out = x_e + omega * d_e * (b_e - sum);
  }
  y[n * 3 + p] = out;
}

}  // namespace

template <typename T>
void launch_assemble_stencil(const Geom& g, const double* Kem, T* S, hipStream_t s) {
  const long long t = g.nodes() * SH;
  hipLaunchKernelGGL((assemble_stencil_k<T>), dim3(static_cast<unsigned>((t + 255) / 256)), dim3(256),
                     0, s, g, Kem, S);
  TOPOHIP_CHECK(hipGetLastError());
}
template void launch_assemble_stencil<float>(const Geom&, const double*, float*, hipStream_t);
template void launch_assemble_stencil<double>(const Geom&, const double*, double*, hipStream_t);

template <typename T>
void launch_stencil_apply(const Geom& g, const T* S, const uint8_t* mask, const T* x, T* y,
                          const T* b, const T* dinv, double omega, int epi, const int* done,
                          hipStream_t s, int nr) {
  const long long n = g.nodes();
  const dim3 grid(xcd_grid((n + SA_NODES - 1) / SA_NODES), nr), blk(SA_NODES * SA_ROWS);
  const T om = static_cast<T>(omega);
  if (epi == EPI_AX)
    hipLaunchKernelGGL((stencil_apply_k<T, EPI_AX>), grid, blk, 0, s, g, S, mask, x, y, b, dinv, om, done);
  else if (epi == EPI_RES)
    hipLaunchKernelGGL((stencil_apply_k<T, EPI_RES>), grid, blk, 0, s, g, S, mask, x, y, b, dinv, om, done);
  else
    hipLaunchKernelGGL((stencil_apply_k<T, EPI_JAC>), grid, blk, 0, s, g, S, mask, x, y, b, dinv, om, done);
  TOPOHIP_CHECK(hipGetLastError());
}
template void launch_stencil_apply<float>(const Geom&, const float*, const uint8_t*, const float*,
                                          float*, const float*, const float*, double, int,
                                          const int*, hipStream_t, int);
template void launch_stencil_apply<double>(const Geom&, const double*, const uint8_t*, const double*,
                                           double*, const double*, const double*, double, int,
                                           const int*, hipStream_t, int);

}  // namespace topohip
