// coarse.hip — exact FP64 solve of the coarsest multigrid level by block
// cyclic reduction over grid planes.
//
// The reference factors its coarsest level (<= 6000 DOF) with a sparse LDL^T
// (core/src/fea/multigrid.cpp:690-694) and solves it exactly in every V-cycle
// (:798-801). On the device the level is ordered plane by plane along its
// longest node axis, which makes it block tridiagonal with m x m blocks
// (m = 3 * nodes per plane):  A = tridiag(L_i, D_i, L_{i+1}^T).
// Block cyclic reduction eliminates the odd planes, leaving a block
// tridiagonal Schur complement on the even ones, and recurses:
//   Dinv_i = D_i^-1,  P_i = Dinv_i L_i,  Q_i = Dinv_i L_{i+1}^T      (i odd)
//   D'_j = D_j - L_j Q_{j-1} - L_{j+1}^T P_{j+1}
//   L'_j = -L_j P_{j-1}                                               (j even)
// and the solve runs
//   b'_j = b_j - Q_{j-1}^T b_{j-1} - P_{j+1}^T b_{j+1}        (forward)
//   x_i  = Dinv_i b_i - P_i x_{i-1} - Q_i x_{i+1}             (backward)
// It is an exact direct solve (block Cholesky in a nested-dissection plane
// order), with log2(planes) sequential steps, each a batch of independent
// m x m FP64 inversions / GEMMs / GEMVs.
#include <algorithm>
#include <vector>

#include "coarse.hpp"
#include "../../include/topopt_hip.h"

namespace topohip {

namespace {

constexpr int IV_MP = 256;  // blocks are padded to 256 x 256 (m <= 256)
constexpr int IV_NB = 32;   // pivots per step = output tile edge

// Batched inversion of SPD m x m blocks with the symmetric sweep operator,
// one launch per 32-pivot block B and one workgroup per 32 x 32 output tile:
// sweeping B maps
//   A_BB -> Sw = -A_BB^-1,   A_iB -> U_i = A_iB A_BB^-1 = -A_iB Sw,
//   A_ij -> A_ij - U_i A_jB^T                         (i, j outside B),
// so after every block is swept the matrix holds -A^-1. Each step reads the
// previous state and writes the next one (ping-pong between two padded
// 256 x 256 buffers per block), so tiles of one step never race. Pad entries
// (index >= m) enter as the identity, which leaves the inverse of the m x m
// block in its leading corner. Every workgroup sweeps the 32 x 32 pivot block
// itself (cheap) instead of waiting for another launch: each lane keeps its 4
// entries of the block in registers and only the next pivot's row and column
// go through LDS, one barrier per pivot. All global reads of a step (the
// pivot block, the two panels and the output tile) are issued up front.
// offs holds (source, destination) offset pairs; step 0 reads the source,
// the last step writes the negated result to the destination.
constexpr int IV_LD = IV_NB + 2;  // LDS row stride (even: 16-byte aligned rows)

// 1 / d from the hardware reciprocal estimate and two Newton steps (the
// estimate's relative error is below 2^-22, so the result is within an ulp).
__device__ inline double recip(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  r = fma(r, fma(-d, r, 1.0), r);
  return r;
}

__global__ void __launch_bounds__(256)
inv_step_k(double* W, const long long* offs, int m, int step, int nsteps, double* scratch) {
  __shared__ double Sw[IV_NB][IV_LD];
  __shared__ double Ai[IV_NB][IV_LD];
  __shared__ double AjT[IV_NB][IV_LD];  // transposed: AjT[e][r] = A_JB[r][e]
  __shared__ double Ui[IV_NB][IV_LD];
  __shared__ double Uj[IV_NB][IV_LD];
  __shared__ double rowb[2][IV_NB], colb[2][IV_NB];
  const int tiles = (m + IV_NB - 1) / IV_NB;
  const int I = blockIdx.x / tiles, J = blockIdx.x % tiles;
  const int b = blockIdx.y;
  const double* src = W + offs[2 * b];
  double* dst = W + offs[2 * b + 1];
  const long long pb = static_cast<long long>(IV_MP) * IV_MP;
  const double* in = scratch + (2LL * b + (step & 1)) * pb;
  double* out = scratch + (2LL * b + ((step + 1) & 1)) * pb;
  const bool first = step == 0, last = step + 1 == nsteps;
  auto rd = [&](int i, int j) -> double {
    if (first) return (i < m && j < m) ? src[static_cast<long long>(i) * m + j] : (i == j ? 1.0 : 0.0);
    return in[i * IV_MP + j];
  };
  const int t = threadIdx.x;
  const int tr = t / 8, tc = (t % 8) * 4;  // lane entries (tr, tc .. tc + 3) of every tile
  const int k0 = step * IV_NB;
  const int i = I * IV_NB + tr;
  const bool gen = I != step && J != step;  // output tile outside the pivot row/column
  double e[4], ai[4], aj[4], aij[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    e[q] = rd(k0 + tr, k0 + tc + q);
    ai[q] = rd(i, k0 + tc + q);
    aj[q] = rd(J * IV_NB + tr, k0 + tc + q);
    if (gen) aij[q] = rd(i, J * IV_NB + tc + q);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    Ai[tr][tc + q] = ai[q];
    AjT[tc + q][tr] = aj[q];
  }
  if (tr == 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) rowb[0][tc + q] = e[q];
  }
  if (tc == 0) colb[0][tr] = e[0];
  __syncthreads();
  // 32 single-pivot sweeps of the pivot block, in registers.
  for (int p = 0; p < IV_NB; ++p) {
    const int cur = p & 1, nxt = cur ^ 1;
    const double inv = recip(rowb[cur][p]);
    const double cp = colb[cur][tr];  // S[tr][p]
    double rp[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) rp[q] = rowb[cur][tc + q];  // S[p][tc + q]
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int sb = tc + q;
      if (tr == p && sb == p) e[q] = -inv;
      else if (tr == p || sb == p) e[q] = e[q] * inv;
      else e[q] = e[q] - cp * (rp[q] * inv);
    }
    if (p + 1 < IV_NB) {  // publish the next pivot's row and column
      if (tr == p + 1) {
#pragma unroll
        for (int q = 0; q < 4; ++q) rowb[nxt][tc + q] = e[q];
      }
      const int d = p + 1 - tc;
      if (d >= 0 && d < 4) colb[nxt][tr] = d == 0 ? e[0] : d == 1 ? e[1] : d == 2 ? e[2] : e[3];
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) Sw[tr][tc + q] = e[q];
  __syncthreads();
  // U_I = -A_IB Sw (every tile but the pivot row's), U_J = -A_JB Sw (the
  // pivot row's tiles only). The branches are uniform over the workgroup.
  const bool need_ui = I != step, need_uj = I == step && J != step;
  if (need_ui) {
    double si[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < IV_NB; ++k) {
      const double a = Ai[tr][k];
#pragma unroll
      for (int q = 0; q < 4; ++q) si[q] -= a * Sw[k][tc + q];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) Ui[tr][tc + q] = si[q];
  }
  if (need_uj) {
    double sj[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < IV_NB; ++k) {
      const double c = AjT[k][tr];
#pragma unroll
      for (int q = 0; q < 4; ++q) sj[q] -= c * Sw[k][tc + q];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) Uj[tr][tc + q] = sj[q];
  }
  __syncthreads();
  double v[4];
  if (I == step && J == step) {
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = Sw[tr][tc + q];
  } else if (J == step) {
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = Ui[tr][tc + q];
  } else if (I == step) {
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = Uj[tc + q][tr];
  } else {
    double sacc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < IV_NB; ++k) {
      const double u = Ui[tr][k];
#pragma unroll
      for (int q = 0; q < 4; ++q) sacc[q] += u * AjT[k][tc + q];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = aij[q] - sacc[q];
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int j = J * IV_NB + tc + q;
    if (last) {
      if (i < m && j < m) dst[static_cast<long long>(i) * m + j] = -v[q];
    } else {
      out[i * IV_MP + j] = v[q];
    }
  }
}

// Batched C = beta*D + sum_t alpha_t op(A_t) op(B_t): op(A) m x m (m <= 256),
// op(B) m x n, C and D m x n, B, C and D with row stride ld (square jobs:
// n = ld = m), and Ct = C^T as well when a square job names one. One
// workgroup per 32 x 32 output tile; wave w multiplies the k range
// [64w, 64w + 64) for the whole tile on the FP64 matrix cores (2 x 2 blocks of
// 16 x 16, k = 4 per instruction, two batches of 8 instructions per term with
// every operand of a batch loaded first), and the four partial tiles are added
// in LDS in a fixed order. Operands past row/column m (or n) are read at a
// clamped in-bounds address and zeroed by a bit mask, not a branch. alpha_t
// scales the A operand (every job uses alpha = +-1, exact).
typedef double dbl4 __attribute__((ext_vector_type(4)));
constexpr int GT = 32;

__device__ inline double wave_sum64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

__device__ inline double ld_keep(const double* p, long long i, bool keep) {
  const unsigned long long u = static_cast<unsigned long long>(__double_as_longlong(p[i]));
  return __longlong_as_double(static_cast<long long>(u & (keep ? ~0ULL : 0ULL)));
}

__global__ void __launch_bounds__(256) gemm_jobs_k(double* W, const CrGemm* jobs, int m, int n,
                                                   int ld) {
  __shared__ double red[4][GT][GT + 1];
  const CrGemm& jb = jobs[blockIdx.y];  // read in place: indexing a local copy by term would spill
  const int tcols = (n + GT - 1) / GT;
  const int tr = blockIdx.x / tcols, tc = blockIdx.x % tcols;
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  const int li = l % 16, lk = l / 16;
  // v_mfma_f64_16x16x4f64 lane layout: A[i][k] from lane (i + 16k), B[k][j]
  // from lane (j + 16k); result register r of lane l is D[l/16 + 4r][l%16].
  // Lane (li, lk) supplies k = kb + s at instruction s of a batch, kb its
  // batch base, so each lane walks 8 consecutive k.
  dbl4 acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) acc[x][y] = dbl4{0.0, 0.0, 0.0, 0.0};
  for (int t = 0; t < jb.nterms; ++t) {
    const double* A = W + jb.a[t];
    const double* B = W + jb.b[t];
    const long long sa = jb.ta[t] ? m : 1, ra = jb.ta[t] ? 1 : m;    // A(r, k) = A[r*ra + k*sa]
    const long long sb = jb.tb[t] ? 1 : ld, cb = jb.tb[t] ? ld : 1;  // B(k, c) = B[k*sb + c*cb]
    const double al = jb.alpha[t];
#pragma unroll 1
    for (int h = 0; h < 2; ++h) {
      const int kb = 64 * w + 32 * h + 8 * lk;
      double a[2][8], b[2][8];
#pragma unroll
      for (int x = 0; x < 2; ++x) {
        const int r = tr * GT + 16 * x + li, c = tc * GT + 16 * x + li;
        const int rc = r < m ? r : m - 1, cc = c < n ? c : n - 1;
#pragma unroll
        for (int s8 = 0; s8 < 8; ++s8) {
          const int k = kb + s8;
          const int kc = k < m ? k : m - 1;
          a[x][s8] = al * ld_keep(A, rc * ra + kc * sa, r < m && k < m);
          b[x][s8] = ld_keep(B, kc * sb + cc * cb, k < m && c < n);
        }
      }
#pragma unroll
      for (int s8 = 0; s8 < 8; ++s8)
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
          for (int y = 0; y < 2; ++y)
            acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x][s8], b[y][s8], acc[x][y], 0, 0, 0);
    }
  }
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) red[w][16 * x + lk + 4 * r][16 * y + li] = acc[x][y][r];
  __syncthreads();
  auto tile_sum = [&](int i, int j) {
    return ((red[0][i][j] + red[1][i][j]) + red[2][i][j]) + red[3][i][j];
  };
  const int i = threadIdx.x / 8, j0 = (threadIdx.x % 8) * 4;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int gr = tr * GT + i, gc = tc * GT + j0 + q;
    if (gr < m && gc < n) {
      const long long o = static_cast<long long>(gr) * ld + gc;
      double v = tile_sum(i, j0 + q);
      if (jb.beta != 0.0) v += jb.beta * W[jb.d + o];
      W[jb.c + o] = v;
    }
  }
  if (jb.ct >= 0 && jb.beta == 0.0 && n == m && ld == m) {  // C^T, written along its rows
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int gr = tr * GT + j0 + q, gc = tc * GT + i;  // C[gr][gc] -> Ct[gc][gr]
      if (gr < m && gc < m) W[jb.ct + static_cast<long long>(gc) * m + gr] = tile_sum(j0 + q, i);
    }
  }
}

// y = M v for one dense n x n FP64 matrix (the inverse of the coarse solver's
// tail level): one wave per row, 8 strides of 64 lanes per batch with every
// load of a batch issued first, then the fixed shuffle tree.
__global__ void __launch_bounds__(256) gemv_dense_k(const double* W, double* V, long long vsz,
                                                    long long moff, long long voff,
                                                    long long ooff, int n, const int* done) {
  const int rc = blockIdx.z;
  if (case_done(done, rc)) return;
  V += rc * vsz;
  const int lane = threadIdx.x & 63;
  const int o = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (o >= n) return;  // whole wave
  const double* M = W + moff + static_cast<long long>(o) * n;
  const double* v = V + voff;
  double s = 0.0;
  for (int k0 = 0; k0 < n; k0 += 512) {
    double pm[8], pv[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int k = k0 + lane + 64 * i;
      const int kc = k < n ? k : n - 1;
      pm[i] = ld_keep(M, kc, k < n);
      pv[i] = v[kc];
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) s += pm[i] * pv[i];
  }
  s = wave_sum64(s);
  if (lane == 0) V[ooff + o] = s;
}

// Batched y = beta*y0 + sum_t alpha_t M_t v_t (length-m vectors, FP64; M_t
// row-major, one row per output). One wave per output row: the lanes stride
// the row (coalesced), then a fixed shuffle tree adds the 64 partial sums, so
// each reduction level is one launch with ~m/4 * jobs workgroups and the result
// is deterministic. (The plan stores every plane block in both orientations,
// so each product reads its matrix along the rows.)
__global__ void __launch_bounds__(256) gemv_rows_k(const double* W, double* V, long long vsz,
                                                   const CrGemv* jobs, int m, const int* done) {
  const int rc = blockIdx.z;
  if (case_done(done, rc)) return;
  V += rc * vsz;
  const CrGemv& jb = jobs[blockIdx.x];
  const int lane = threadIdx.x & 63;
  const int o = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (o >= m) return;  // whole wave
  // Fixed trip counts (3 terms, m <= IV_MP): absent terms and the tail of the
  // row are clamped to valid addresses and weighted by zero, so all loads of
  // a lane are used unconditionally and issued together.
  static_assert(IV_MP == 4 * 64, "gemv_rows_k covers a row in 4 strides of 64 lanes");
  double s = 0.0;
  if (jb.nterms > 0) {  // (jobs without terms are plain copies y = y0)
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    const bool tok = t < jb.nterms;
    const double* M = W + (tok ? jb.m[t] : jb.m[0]) + static_cast<long long>(o) * m;
    const double* v = V + (tok ? jb.v[t] : jb.v[0]);
    double st = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int k = lane + 64 * i;
      const int kc = k < m ? k : m - 1;
      const double f = k < m ? 1.0 : 0.0;
      st += (M[kc] * f) * v[kc];
    }
    s += (tok ? jb.alpha[t] : 0.0) * st;
  }
  }
  s = wave_sum64(s);
  if (lane == 0) {
    if (jb.beta != 0.0) s += jb.beta * V[jb.y0 + o];
    V[jb.out + o] = s;
  }
}

// A = I (n x n).
__global__ void eye_k(double* A, long long n) {
  const long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (t < n * n) A[t] = (t / n == t % n) ? 1.0 : 0.0;
}

// Blocks of the coarsest operator in plane order, from the element matrices.
__global__ void cr_assemble_k(Geom g, const double* __restrict__ K, const uint8_t* __restrict__ mask,
                              const int* __restrict__ plane_of, const int* __restrict__ local_of,
                              int m, double* W, long long offD, long long offL) {
  const long long nodes = g.nodes();
  const long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (t >= nodes * 27) return;
  const long long ni = t % nodes;
  const int o = static_cast<int>(t / nodes);
  const int ox = o % 3 - 1, oy = (o / 3) % 3 - 1, oz = o / 9 - 1;
  const int a = static_cast<int>(ni % g.nnx);
  const int b = static_cast<int>((ni / g.nnx) % g.nny);
  const int c = static_cast<int>(ni / (static_cast<long long>(g.nnx) * g.nny));
  const int aj = a + ox, bj = b + oy, cj = c + oz;
  if (aj < 0 || bj < 0 || cj < 0 || aj >= g.nnx || bj >= g.nny || cj >= g.nnz) return;
  const long long nj = (static_cast<long long>(cj) * g.nny + bj) * g.nnx + aj;
  const int pi = plane_of[ni], pj = plane_of[nj];
  long long base;
  if (pj == pi) base = offD + static_cast<long long>(pi) * m * m;
  else if (pj == pi - 1) base = offL + static_cast<long long>(pi) * m * m;
  else return;
  double blk[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
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
  const int li = local_of[ni], lj = local_of[nj];
  const int mki = mask ? mask[ni] : 7;
  for (int p = 0; p < 3; ++p)
    for (int q = 0; q < 3; ++q) {
      double v = blk[p * 3 + q];
      if (ni == nj && p == q && !((mki >> p) & 1)) v = 1.0;  // identity on inactive DOFs
      W[base + static_cast<long long>(li + p) * m + lj + q] = v;
    }
}

// Node-ordered level vector (T) <-> plane-ordered FP64 vector.
template <typename T>
__global__ void cr_gather_k(const T* __restrict__ v, const int* __restrict__ dst, long long n,
                            double* V, long long vsz, long long off, const int* done) {
  const int rc = blockIdx.y;
  if (case_done(done, rc)) return;
  const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (i < n) V[rc * vsz + off + dst[i]] = static_cast<double>(v[rc * n + i]);
}
template <typename T>
__global__ void cr_scatter_k(const double* V, long long vsz, long long off,
                             const int* __restrict__ dst, long long n, T* __restrict__ v,
                             const int* done) {
  const int rc = blockIdx.y;
  if (case_done(done, rc)) return;
  const long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (i < n) v[rc * n + i] = static_cast<T>(V[rc * vsz + off + dst[i]]);
}

}  // namespace

bool CoarseCR::plan(const Geom& g, hipStream_t s) {
  geom = g;
  const int dims[3] = {g.nnx, g.nny, g.nnz};
  axis = 0;
  for (int d = 1; d < 3; ++d)
    if (dims[d] > dims[axis]) axis = d;
  planes = dims[axis];
  const int u = axis == 0 ? 1 : 0, w = axis == 2 ? 1 : 2;
  m = 3 * dims[u] * dims[w];
  if (m > IV_MP) return false;
  const long long nodes = g.nodes();
  std::vector<int> plane_of(static_cast<size_t>(nodes)), local_of(static_cast<size_t>(nodes));
  std::vector<int> dst(static_cast<size_t>(3 * nodes));
  for (int c = 0; c < g.nnz; ++c)
    for (int b = 0; b < g.nny; ++b)
      for (int a = 0; a < g.nnx; ++a) {
        const int idx[3] = {a, b, c};
        const long long n = (static_cast<long long>(c) * g.nny + b) * g.nnx + a;
        plane_of[n] = idx[axis];
        local_of[n] = 3 * (idx[w] * dims[u] + idx[u]);
        for (int q = 0; q < 3; ++q)
          dst[3 * n + q] = plane_of[n] * m + local_of[n] + q;
      }
  // Level sizes and workspace layout.
  lv.clear();
  long long off = 0, voff = 0;  // matrices in W, vectors (per load case) in V
  const long long mm = static_cast<long long>(m) * m;
  int n = planes;
  while (true) {
    CrLevel L;
    L.n = n;
    L.D = off; off += n * mm;
    L.L = off; off += n * mm;
    L.Dinv = off; off += n * mm;
    L.P = off; off += n * mm;
    L.Q = off; off += n * mm;
    L.Pt = off; off += n * mm;
    L.Qt = off; off += n * mm;
    L.b = voff; voff += static_cast<long long>(n) * m;
    L.x = voff; voff += static_cast<long long>(n) * m;
    lv.push_back(L);
    if (n == 1) break;
    n = (n + 1) / 2;
  }
  wsize = off;
  vsize = voff;
  // Job tables.
  std::vector<long long> inv_offs;
  std::vector<CrGemm> gemm_p, gemm_d;
  std::vector<CrGemv> fwd, bwd;
  lv_inv.clear(); lv_gp.clear(); lv_gd.clear(); lv_fw.clear(); lv_bw.clear();
  for (size_t r = 0; r < lv.size(); ++r) {
    const CrLevel& L = lv[r];
    const int nr = L.n;
    lv_inv.push_back(static_cast<int>(inv_offs.size() / 2));
    lv_gp.push_back(static_cast<int>(gemm_p.size()));
    lv_gd.push_back(static_cast<int>(gemm_d.size()));
    lv_fw.push_back(static_cast<int>(fwd.size()));
    lv_bw.push_back(static_cast<int>(bwd.size()));
    if (nr == 1) {
      inv_offs.push_back(L.D);  // final block: Dinv = D^-1
      inv_offs.push_back(L.Dinv);
      CrGemv j{};
      j.out = L.x; j.nterms = 1; j.m[0] = L.Dinv; j.v[0] = L.b; j.alpha[0] = 1.0;
      bwd.push_back(j);
      continue;
    }
    const CrLevel& N = lv[r + 1];
    for (int i = 1; i < nr; i += 2) {
      inv_offs.push_back(L.D + i * mm);
      inv_offs.push_back(L.Dinv + i * mm);
    }
    for (int i = 1; i < nr; i += 2) {
      CrGemm g1{};  // P_i = Dinv_i L_i, and P_i^T
      g1.c = L.P + i * mm; g1.ct = L.Pt + i * mm; g1.nterms = 1; g1.a[0] = L.Dinv + i * mm;
      g1.b[0] = L.L + i * mm; g1.alpha[0] = 1.0;
      gemm_p.push_back(g1);
      if (i + 1 < nr) {
        CrGemm g2{};  // Q_i = Dinv_i L_{i+1}^T, and Q_i^T
        g2.c = L.Q + i * mm; g2.ct = L.Qt + i * mm; g2.nterms = 1; g2.a[0] = L.Dinv + i * mm;
        g2.b[0] = L.L + (i + 1) * mm; g2.tb[0] = 1; g2.alpha[0] = 1.0;
        gemm_p.push_back(g2);
      }
    }
    for (int j = 0; j < nr; j += 2) {
      CrGemm gd{};  // D'_j = D_j - L_j Q_{j-1} - L_{j+1}^T P_{j+1}
      gd.c = N.D + (j / 2) * mm; gd.d = L.D + j * mm; gd.beta = 1.0;
      if (j >= 1) {
        gd.a[gd.nterms] = L.L + j * mm; gd.b[gd.nterms] = L.Q + (j - 1) * mm;
        gd.alpha[gd.nterms++] = -1.0;
      }
      if (j + 1 < nr) {
        gd.a[gd.nterms] = L.L + (j + 1) * mm; gd.ta[gd.nterms] = 1;
        gd.b[gd.nterms] = L.P + (j + 1) * mm; gd.alpha[gd.nterms++] = -1.0;
      }
      gemm_d.push_back(gd);
      if (j >= 2) {
        CrGemm gl{};  // L'_j = -L_j P_{j-1}
        gl.c = N.L + (j / 2) * mm; gl.nterms = 1; gl.a[0] = L.L + j * mm;
        gl.b[0] = L.P + (j - 1) * mm; gl.alpha[0] = -1.0;
        gemm_d.push_back(gl);
      }
      CrGemv f{};  // b'_j = b_j - Q_{j-1}^T b_{j-1} - P_{j+1}^T b_{j+1}
      f.out = N.b + static_cast<long long>(j / 2) * m; f.y0 = L.b + static_cast<long long>(j) * m;
      f.beta = 1.0;
      if (j >= 1) {
        f.m[f.nterms] = L.Qt + (j - 1) * mm; f.v[f.nterms] = L.b + static_cast<long long>(j - 1) * m;
        f.alpha[f.nterms++] = -1.0;
      }
      if (j + 1 < nr) {
        f.m[f.nterms] = L.Pt + (j + 1) * mm; f.v[f.nterms] = L.b + static_cast<long long>(j + 1) * m;
        f.alpha[f.nterms++] = -1.0;
      }
      fwd.push_back(f);
      CrGemv c{};  // x_j = x'_{j/2}
      c.out = L.x + static_cast<long long>(j) * m; c.y0 = N.x + static_cast<long long>(j / 2) * m;
      c.beta = 1.0;
      bwd.push_back(c);
    }
    for (int i = 1; i < nr; i += 2) {
      CrGemv bi{};  // x_i = Dinv_i b_i - P_i x_{i-1} - Q_i x_{i+1}; x_{i+-1} read from the coarser level
      bi.out = L.x + static_cast<long long>(i) * m;
      bi.m[0] = L.Dinv + i * mm; bi.v[0] = L.b + static_cast<long long>(i) * m;
      bi.alpha[0] = 1.0; bi.nterms = 1;
      bi.m[bi.nterms] = L.P + i * mm; bi.v[bi.nterms] = N.x + static_cast<long long>((i - 1) / 2) * m;
      bi.alpha[bi.nterms++] = -1.0;
      if (i + 1 < nr) {
        bi.m[bi.nterms] = L.Q + i * mm; bi.v[bi.nterms] = N.x + static_cast<long long>((i + 1) / 2) * m;
        bi.alpha[bi.nterms++] = -1.0;
      }
      bwd.push_back(bi);
    }
  }
  lv_inv.push_back(static_cast<int>(inv_offs.size() / 2));
  lv_gp.push_back(static_cast<int>(gemm_p.size()));
  lv_gd.push_back(static_cast<int>(gemm_d.size()));
  lv_fw.push_back(static_cast<int>(fwd.size()));
  lv_bw.push_back(static_cast<int>(bwd.size()));
  // Dense tail: the first reduction level dl whose whole system has at most
  // kDenseMax unknowns is solved by one GEMV with its explicit inverse, which
  // the factorisation forms by running the remaining reduction on the
  // identity: each length-m vector of a level r >= dl becomes an m x nden
  // block (in W), and every GEMV job of those levels the same GEMM job.
  dl = 0;
  while (static_cast<long long>(lv[dl].n) * m > kDenseMax) ++dl;
  nden = static_cast<long long>(lv[dl].n) * m;
  std::vector<long long> bm(lv.size(), -1), xm(lv.size(), -1);
  for (size_t r = dl; r < lv.size(); ++r) {
    bm[r] = off; off += static_cast<long long>(lv[r].n) * m * nden;
    xm[r] = off; off += static_cast<long long>(lv[r].n) * m * nden;
  }
  xm_dense = xm[dl];
  eye_off = bm[dl];
  wsize = off;
  auto to_mat = [&](long long o) {
    for (size_t r = dl; r < lv.size(); ++r) {
      const long long span = static_cast<long long>(lv[r].n) * m;
      if (o >= lv[r].b && o < lv[r].b + span) return bm[r] + (o - lv[r].b) * nden;
      if (o >= lv[r].x && o < lv[r].x + span) return xm[r] + (o - lv[r].x) * nden;
    }
    throw Error(TOPOHIP_EINTERNAL, "coarse solver: vector outside the dense tail");
  };
  auto as_gemm = [&](const CrGemv& v) {
    CrGemm g{};
    g.c = to_mat(v.out);
    g.beta = v.beta;
    if (v.beta != 0.0) g.d = to_mat(v.y0);
    g.nterms = v.nterms;
    for (int t = 0; t < v.nterms; ++t) {
      g.a[t] = v.m[t];
      g.b[t] = to_mat(v.v[t]);
      g.alpha[t] = v.alpha[t];
    }
    return g;
  };
  std::vector<CrGemm> mj;
  lv_mj.clear();
  for (size_t r = dl; r + 1 < lv.size(); ++r) {
    lv_mj.push_back(static_cast<int>(mj.size()));
    for (int j = lv_fw[r]; j < lv_fw[r + 1]; ++j) mj.push_back(as_gemm(fwd[j]));
  }
  for (int r = static_cast<int>(lv.size()) - 1; r >= dl; --r) {
    lv_mj.push_back(static_cast<int>(mj.size()));
    for (int j = lv_bw[r]; j < lv_bw[r + 1]; ++j) mj.push_back(as_gemm(bwd[j]));
  }
  lv_mj.push_back(static_cast<int>(mj.size()));
  // Uploads go through the problem's own stream: a legacy-stream copy would
  // have to order itself against other problems' streams, which may be
  // capturing a graph at that moment.
  auto up = [&](auto& buf, const auto& host) {
    buf.ensure(host.size());
    if (!host.empty())
      TOPOHIP_CHECK(hipMemcpyAsync(buf.p, host.data(), host.size() * sizeof(host[0]),
                                   hipMemcpyHostToDevice, s));
  };
  W.ensure(static_cast<size_t>(wsize));
  hipLaunchKernelGGL(eye_k, dim3(static_cast<unsigned>((nden * nden + 255) / 256)), dim3(256), 0, s,
                     W.p + eye_off, nden);
  TOPOHIP_CHECK(hipGetLastError());
  ncases = 0;
  reserve_cases(1);
  size_t maxinv = 1;
  for (size_t r = 0; r + 1 < lv_inv.size(); ++r)
    maxinv = std::max(maxinv, static_cast<size_t>(lv_inv[r + 1] - lv_inv[r]));
  inv_scratch.ensure(maxinv * 2 * IV_MP * IV_MP);
  up(d_inv, inv_offs);
  up(d_gp, gemm_p);
  up(d_gd, gemm_d);
  up(d_mj, mj);
  up(d_fw, fwd);
  up(d_bw, bwd);
  up(d_plane, plane_of);
  up(d_local, local_of);
  up(d_dst, dst);
  TOPOHIP_CHECK(hipStreamSynchronize(s));  // host tables go out of scope
  return true;
}

void CoarseCR::build(const double* Kem, const uint8_t* mask, hipStream_t s) {
  const Geom& g = geom;
  const long long mm = static_cast<long long>(m) * m;
  launch_fill<double>(W.p + lv[0].D, 0.0, 2LL * planes * mm, s);  // D and L of level 0
  const long long t = g.nodes() * 27;
  hipLaunchKernelGGL(cr_assemble_k, dim3(static_cast<unsigned>((t + 255) / 256)), dim3(256), 0, s, g,
                     Kem, mask, d_plane.p, d_local.p, m, W.p, lv[0].D, lv[0].L);
  TOPOHIP_CHECK(hipGetLastError());
  const int tiles = (m + 31) / 32;
  for (size_t r = 0; r < lv.size(); ++r) {
    const int ni = lv_inv[r + 1] - lv_inv[r];
    if (ni > 0) {
      const int nt = (m + IV_NB - 1) / IV_NB;
      for (int st = 0; st < nt; ++st) {
        hipLaunchKernelGGL(inv_step_k, dim3(nt * nt, ni), dim3(256), 0, s, W.p,
                           d_inv.p + 2 * lv_inv[r], m, st, nt, inv_scratch.p);
        TOPOHIP_CHECK(hipGetLastError());
      }
    }
    const int np = lv_gp[r + 1] - lv_gp[r];
    if (np > 0) {
      hipLaunchKernelGGL(gemm_jobs_k, dim3(tiles * tiles, np), dim3(256), 0, s, W.p,
                         d_gp.p + lv_gp[r], m, m, m);
      TOPOHIP_CHECK(hipGetLastError());
    }
    const int nd = lv_gd[r + 1] - lv_gd[r];
    if (nd > 0) {
      hipLaunchKernelGGL(gemm_jobs_k, dim3(tiles * tiles, nd), dim3(256), 0, s, W.p,
                         d_gd.p + lv_gd[r], m, m, m);
      TOPOHIP_CHECK(hipGetLastError());
    }
  }
  // Dense tail: A_dl^-1 = the reduction solve of levels >= dl applied to the
  // identity (every vector of those levels is an m x nden block row).
  const int nt = static_cast<int>(nden);
  const unsigned mtiles = static_cast<unsigned>(tiles * ((nt + 31) / 32));
  for (size_t i = 0; i + 1 < lv_mj.size(); ++i) {
    const int nj = lv_mj[i + 1] - lv_mj[i];
    if (nj == 0) continue;
    hipLaunchKernelGGL(gemm_jobs_k, dim3(mtiles, nj), dim3(256), 0, s, W.p, d_mj.p + lv_mj[i], m,
                       nt, nt);
    TOPOHIP_CHECK(hipGetLastError());
  }
}

void CoarseCR::reserve_cases(int nr) {
  if (nr <= ncases) return;
  V.ensure(static_cast<size_t>(vsize) * nr);
  ncases = nr;
}

void CoarseCR::gemv(const CrGemv* jobs, int n, const int* done, hipStream_t s, int nr) {
  if (n <= 0) return;
  hipLaunchKernelGGL(gemv_rows_k, dim3(n, (m + 3) / 4, nr), dim3(256), 0, s, W.p, V.p, vsize, jobs,
                     m, done);
}

CrMap CoarseCR::rhs_map() {
  CrMap cm;
  cm.dst = d_dst.p;
  cm.V = V.p;
  cm.vstride = vsize;
  cm.off = lv[0].b;
  return cm;
}

CrMap CoarseCR::sol_map() {
  CrMap cm = rhs_map();
  cm.off = lv[0].x;
  return cm;
}

// Forward reduction down to the dense level, its explicit inverse, backward
// substitution back up.
void CoarseCR::reduce_solve(const int* done, hipStream_t s, int nr) {
  for (int r = 0; r < dl; ++r) gemv(d_fw.p + lv_fw[r], lv_fw[r + 1] - lv_fw[r], done, s, nr);
  const int nt = static_cast<int>(nden);
  hipLaunchKernelGGL(gemv_dense_k, dim3((nt + 3) / 4, 1, nr), dim3(256), 0, s, W.p, V.p, vsize,
                     xm_dense, lv[dl].b, lv[dl].x, nt, done);
  TOPOHIP_CHECK(hipGetLastError());
  for (int r = dl - 1; r >= 0; --r) gemv(d_bw.p + lv_bw[r], lv_bw[r + 1] - lv_bw[r], done, s, nr);
}

void CoarseCR::solve_in_place(const int* done, hipStream_t s, int nr) {
  if (nr > ncases) throw Error(TOPOHIP_EINTERNAL, "coarse solve: load-case space not reserved");
  reduce_solve(done, s, nr);
}

template <typename T>
void CoarseCR::solve(const T* b, T* x, const int* done, hipStream_t s, int nr) {
  if (nr > ncases) throw Error(TOPOHIP_EINTERNAL, "coarse solve: load-case space not reserved");
  const long long n = 3 * geom.nodes();
  const dim3 grid(static_cast<unsigned>((n + 255) / 256), nr);
  hipLaunchKernelGGL((cr_gather_k<T>), grid, dim3(256), 0, s, b, d_dst.p, n, V.p, vsize, lv[0].b,
                     done);
  reduce_solve(done, s, nr);
  hipLaunchKernelGGL((cr_scatter_k<T>), grid, dim3(256), 0, s, V.p, vsize, lv[0].x, d_dst.p, n, x,
                     done);
  TOPOHIP_CHECK(hipGetLastError());
}
template void CoarseCR::solve<float>(const float*, float*, const int*, hipStream_t, int);
template void CoarseCR::solve<double>(const double*, double*, const int*, hipStream_t, int);

}  // namespace topohip
