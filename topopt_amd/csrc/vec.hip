// vec.hip — grid transfers, reductions and the device-resident PCG scalars.
//
// Transfers restate the trilinear vertex prolongation of
// core/src/fea/multigrid.cpp:494-506 (axis_weights) and :586-654 as stencils on
// the structured node grid: restriction is its transpose (R = P^T), both
// masked by the level's active-DOF bits so inactive (fixed / void) DOFs carry
// zero, exactly as P drops inactive rows and columns.
//
// Reductions are two-pass and fixed-order (per-block partials, then a single
// block sums them in index order), so every scalar is bit-identical run to run.
#include "kernels.hpp"

namespace topohip {

namespace {

constexpr int RB = 256;  // reduction block size
// Fine levels with at least this many nodes are prolonged a node pair per thread.
constexpr long long kPairProlongMin = 262144;

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  return v;
}

// Block sum of one value per thread; result valid in thread 0.
__device__ inline double block_sum(double v) {
  __shared__ double sh[RB / 64];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < static_cast<int>(blockDim.x) / 64; ++w) s += sh[w];
  return s;
}

template <typename Tf, typename Tc>
__global__ void restrict_k(Geom gf, Geom gc, const Tf* __restrict__ rf,
                           const uint8_t* __restrict__ mask_c, Tc* __restrict__ bc,
                           Tc* __restrict__ xc, const Tc* __restrict__ dinv_c, double omega,
                           long long sf, long long sc, CrMap cm, const int* done) {
  const int rc = blockIdx.y;
  if (case_done(done, rc)) return;
  rf += rc * sf;
  if (bc) bc += rc * sc;
  if (xc) xc += rc * sc;
  const long long nodes = gc.nodes();
  const long long n = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (n >= nodes) return;
  const int A = static_cast<int>(n % gc.nnx);
  const int B = static_cast<int>((n / gc.nnx) % gc.nny);
  const int C = static_cast<int>(n / (static_cast<long long>(gc.nnx) * gc.nny));
  const int mk = mask_c ? mask_c[n] : 7;
  // 27 fine nodes in a fixed order; out-of-grid ones are clamped to the
  // coincident node 2*(A, B, C) and weighted by zero. (Every load is used
  // unconditionally, so the compiler issues them up front rather than
  // behind a branch each.)
  double s[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int dz = -1; dz <= 1; ++dz) {
    const int c = 2 * C + dz;
    const bool okz = c >= 0 && c < gf.nnz;
    const double wz = dz == 0 ? 1.0 : 0.5;
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy) {
      const int b = 2 * B + dy;
      const bool oky = okz && b >= 0 && b < gf.nny;
      const double wy = wz * (dy == 0 ? 1.0 : 0.5);
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx) {
        const int a = 2 * A + dx;
        const bool ok = oky && a >= 0 && a < gf.nnx;
        const double w = ok ? wy * (dx == 0 ? 1.0 : 0.5) : 0.0;
        const long long fn = ok ? (static_cast<long long>(c) * gf.nny + b) * gf.nnx + a
                                : (static_cast<long long>(2 * C) * gf.nny + 2 * B) * gf.nnx + 2 * A;
#pragma unroll
        for (int q = 0; q < 3; ++q) s[q] += w * static_cast<double>(rf[fn * 3 + q]);
      }
    }
  }
  if (cm.dst) {  // straight into the coarse solver's plane-ordered FP64 right-hand side
#pragma unroll
    for (int q = 0; q < 3; ++q)
      cm.V[rc * cm.vstride + cm.off + cm.dst[n * 3 + q]] = ((mk >> q) & 1) ? s[q] : 0.0;
    return;
  }
  Tc dv[3] = {Tc(0), Tc(0), Tc(0)};
  if (xc) {
#pragma unroll
    for (int q = 0; q < 3; ++q) dv[q] = dinv_c[n * 3 + q];
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const bool kept = (mk >> q) & 1;
    const Tc v = kept ? static_cast<Tc>(s[q]) : Tc(0);
    bc[n * 3 + q] = v;
    if (xc) xc[n * 3 + q] = kept ? static_cast<Tc>(omega * static_cast<double>(dv[q]) *
                                                   static_cast<double>(v))
                                 : Tc(0);
  }
}

template <typename Tc, typename Tf>
__global__ void prolong_add_k(Geom gc, Geom gf, const Tc* __restrict__ xc,
                              const uint8_t* __restrict__ mask_f, Tf* __restrict__ xf,
                              long long sc, long long sf, CrMap cm, const int* done) {
  const int rc = blockIdx.y;
  if (case_done(done, rc)) return;
  if (xc) xc += rc * sc;
  xf += rc * sf;
  const long long nodes = gf.nodes();
  const long long n = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (n >= nodes) return;
  const int a = static_cast<int>(n % gf.nnx);
  const int b = static_cast<int>((n / gf.nnx) % gf.nny);
  const int c = static_cast<int>(n / (static_cast<long long>(gf.nnx) * gf.nny));
  // (the mask only gates the stores: a branch on it would hold every load
  // behind the mask's)
  const int mk = mask_f ? mask_f[n] : 7;
  // The 1, 2, 4 or 8 coarse nodes of fine node (a, b, c), in a fixed order;
  // the second neighbour along an even axis is absent: its (in-grid) load is
  // weighted by zero, so every load is used and issued up front.
  const int ia[2] = {a >> 1, (a + 1) >> 1}, ib[2] = {b >> 1, (b + 1) >> 1};
  const int ic[2] = {c >> 1, (c + 1) >> 1};
  const double wa = (a & 1) ? 0.5 : 1.0, wb = (b & 1) ? 0.5 : 1.0, wc = (c & 1) ? 0.5 : 1.0;
  double w[8], v[8][3];
  long long cn[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const int x = t >> 2, y = (t >> 1) & 1, z = t & 1;
    const bool ok = (x == 0 || (a & 1)) && (y == 0 || (b & 1)) && (z == 0 || (c & 1));
    w[t] = ok ? wa * wb * wc : 0.0;
    cn[t] = (static_cast<long long>(ic[z]) * gc.nny + ib[y]) * gc.nnx + ia[x];
  }
  if (cm.dst) {  // coarse values straight from the coarse solver's FP64 solution
    int d[8][3];
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
      for (int q = 0; q < 3; ++q) d[t][q] = cm.dst[cn[t] * 3 + q];
    const double* V = cm.V + rc * cm.vstride + cm.off;
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
      for (int q = 0; q < 3; ++q) v[t][q] = V[d[t][q]];
  } else {
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
      for (int q = 0; q < 3; ++q) v[t][q] = static_cast<double>(xc[cn[t] * 3 + q]);
  }
  double s[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int t = 0; t < 8; ++t)
#pragma unroll
    for (int q = 0; q < 3; ++q) s[q] += w[t] * v[t][q];
  Tf cur[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) cur[q] = xf[n * 3 + q];
#pragma unroll
  for (int q = 0; q < 3; ++q)
    if ((mk >> q) & 1) xf[n * 3 + q] = static_cast<Tf>(static_cast<double>(cur[q]) + s[q]);
}

// prolong_add_k for a pair of fine nodes (2i, 2i + 1) along x per thread: both
// gather from coarse x = i and i + 1, so the pair issues 24 coarse loads
// instead of 48. Each node's weights, terms and their order are prolong_add_k's
// (the even node's x = 1 terms read coarse i with weight 0, as there), so the
// result is bit-identical. For large fine levels, where the loads dominate.
template <typename Tc, typename Tf>
__global__ void prolong_add_pair_k(Geom gc, Geom gf, const Tc* __restrict__ xc,
                                   const uint8_t* __restrict__ mask_f, Tf* __restrict__ xf,
                                   long long sc, long long sf, const int* done) {
  const int rc = blockIdx.y;
  if (case_done(done, rc)) return;
  xc += rc * sc;
  xf += rc * sf;
  const int px = (gf.nnx + 1) / 2;  // pairs per fine row; the last may hold only its even node
  const long long pairs = static_cast<long long>(px) * gf.nny * gf.nnz;
  const long long q = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (q >= pairs) return;
  const int i = static_cast<int>(q % px);
  const int b = static_cast<int>((q / px) % gf.nny);
  const int c = static_cast<int>(q / (static_cast<long long>(px) * gf.nny));
  const int a0 = 2 * i;
  const bool has_odd = a0 + 1 < gf.nnx;
  const int i1 = has_odd ? i + 1 : i;
  const int ib[2] = {b >> 1, (b + 1) >> 1}, ic[2] = {c >> 1, (c + 1) >> 1};
  const double wb = (b & 1) ? 0.5 : 1.0, wc = (c & 1) ? 0.5 : 1.0;
  double v0[4][3], v1[4][3];  // coarse x = i and i1 at (y, z) = (yz >> 1, yz & 1)
#pragma unroll
  for (int yz = 0; yz < 4; ++yz) {
    const long long row = (static_cast<long long>(ic[yz & 1]) * gc.nny + ib[yz >> 1]) * gc.nnx;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      v0[yz][k] = static_cast<double>(xc[(row + i) * 3 + k]);
      v1[yz][k] = static_cast<double>(xc[(row + i1) * 3 + k]);
    }
  }
  auto node = [&](int a, bool odd) {
    const long long n = (static_cast<long long>(c) * gf.nny + b) * gf.nnx + a;
    const int mk = mask_f ? mask_f[n] : 7;  // gates the stores only (see prolong_add_k)
    const double wa = odd ? 0.5 : 1.0;
    double s[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int x = t >> 2, y = (t >> 1) & 1, z = t & 1;
      const bool ok = (x == 0 || odd) && (y == 0 || (b & 1)) && (z == 0 || (c & 1));
      const double w = ok ? wa * wb * wc : 0.0;
      const double* v = (x == 1 && odd) ? v1[2 * y + z] : v0[2 * y + z];
#pragma unroll
      for (int k = 0; k < 3; ++k) s[k] += w * v[k];
    }
    Tf cur[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) cur[k] = xf[n * 3 + k];
#pragma unroll
    for (int k = 0; k < 3; ++k)
      if ((mk >> k) & 1) xf[n * 3 + k] = static_cast<Tf>(static_cast<double>(cur[k]) + s[k]);
  };
  node(a0, false);
  if (has_odd) node(a0 + 1, true);
}

__global__ void sum_partials_k(const double* __restrict__ p, int n, double* out, int pstride) {
  p += static_cast<long long>(blockIdx.x) * pstride;
  out += blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += p[i];
  s = block_sum(s);
  if (threadIdx.x == 0) *out = s;
}

__global__ void max_partials_k(const double* __restrict__ p, int n, double* out) {
  __shared__ double sh[RB / 64];
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, p[i]);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int w = 0; w < RB / 64; ++w) r = fmax(r, sh[w]);
    *out = r;
  }
}

template <typename TA, typename TB>
__global__ void dot_k(const TA* __restrict__ a, const TB* __restrict__ b, long long n,
                      double* partials, int pstride, const int* done) {
  const int rc = blockIdx.y;
  if (case_done(done, rc)) return;
  a += rc * n;
  b += rc * n;
  partials += static_cast<long long>(rc) * pstride;
  double s = 0.0;
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    s += static_cast<double>(a[i]) * static_cast<double>(b[i]);
  s = block_sum(s);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

template <typename T>
__global__ void sum_vec_k(const T* __restrict__ a, long long n, double* partials) {
  double s = 0.0;
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    s += static_cast<double>(a[i]);
  s = block_sum(s);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

template <typename T>
__global__ void fill_k(T* p, T v, long long n) {
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    p[i] = v;
}

template <typename TI, typename TO>
__global__ void convert_k(const TI* __restrict__ in, TO* __restrict__ out, long long n) {
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    out[i] = static_cast<TO>(in[i]);
}

template <typename T>
__global__ void jacobi0_k(T* __restrict__ x, const T* __restrict__ dinv, const T* __restrict__ b,
                          double omega, long long n, const int* done) {
  const int rc = blockIdx.y;
  if (case_done(done, rc)) return;
  x += rc * n;
  b += rc * n;
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    x[i] = static_cast<T>(omega * static_cast<double>(dinv[i]) * static_cast<double>(b[i]));
}

__global__ void add_k(const double* __restrict__ x, const double* __restrict__ up,
                      double* __restrict__ y, long long n) {
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    y[i] = x[i] + up[i];
}

// ---- PCG scalar kernels (single thread of a single block reads partials) ---
__device__ inline double fixed_sum(const double* p, int n) {
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += p[i];
  return block_sum(s);
}

// fixed_sum with the result in every thread of the block.
__device__ inline double fixed_sum_all(const double* p, int n) {
  __shared__ double r;
  const double s = fixed_sum(p, n);
  if (threadIdx.x == 0) r = s;
  __syncthreads();
  return r;
}

__global__ void cg_init_k(CgState* st, const double* p, int n, int pstride, double tol, int cap) {
  st += blockIdx.x;
  p += static_cast<long long>(blockIdx.x) * pstride;
  const double bb = fixed_sum(p, n);
  if (threadIdx.x == 0) {
    st->bnorm2 = bb;
    const double th = tol * sqrt(bb);
    st->thresh2 = th * th;
    st->done = bb > 0.0 ? 0 : 1;
    st->fail = 0;
    st->iters = 0;
    st->cap = cap;
    st->rr = bb;
    st->pfresh = 1;
    st->it0 = 0;
  }
}

// Restart: rz = rz_new = <r, z>, and the next A.p takes p = z. (In the
// iteration, the A.p forms beta = <r,z>_new / rz from the preconditioner's
// partials and records <r,z>_new, cg_xr uses it for alpha and clears the
// restart mark, cg_check moves it to rz.)
__global__ void cg_set_rz_k(CgState* st, const double* p, int n, int pstride) {
  st += blockIdx.x;
  p += static_cast<long long>(blockIdx.x) * pstride;
  if (st->done) return;
  const double rz = fixed_sum(p, n);
  if (threadIdx.x == 0) {
    st->rz = rz;
    st->rz_new = rz;
    st->pfresh = 1;
    st->it0 = st->iters;
    if (!isfinite(rz)) { st->fail = 1; st->done = 1; }
  }
}

__global__ void cg_check_k(CgState* st, const double* p, int n, int pstride) {
  st += blockIdx.x;
  p += static_cast<long long>(blockIdx.x) * pstride;
  if (st->done) return;
  const double rr = fixed_sum(p, n);
  if (threadIdx.x == 0) {
    st->rr = rr;
    st->iters += 1;
    if (!isfinite(rr)) { st->fail = 1; st->done = 1; }
    else if (rr <= st->thresh2) st->done = 1;
    else if (st->cap > 0 && st->iters >= st->cap) { st->fail = 2; st->done = 1; }
    st->rz = st->rz_new;
  }
}

__global__ void cg_resume_k(CgState* st, const double* p, int n, int pstride) {
  st += blockIdx.x;
  p += static_cast<long long>(blockIdx.x) * pstride;
  const double rr = fixed_sum(p, n);
  if (threadIdx.x == 0) {
    st->rr = rr;
    st->done = rr <= st->thresh2 ? 1 : 0;
    st->fail = 0;
  }
}

// alpha = rz_new / pAp with pAp the fixed-order sum of the case's A.p partials
// (every block forms the same value; block 0 records it and the breakdown
// test, pAp <= 0 or non-finite), then x += alpha p, r -= alpha Ap, rr partials.
template <typename TP, typename TA>
__global__ void cg_xr_k(CgState* st, double* __restrict__ x, double* __restrict__ r,
                        const TP* __restrict__ p, const TA* __restrict__ Ap, float* r32,
                        long long n, const double* pap, int npap, double* partials,
                        int pstride, const double* __restrict__ jdinv, double* __restrict__ jz,
                        double* __restrict__ rzp, const TP* __restrict__ p_prev) {
  const int rc = blockIdx.y;
  st += rc;
  if (st->done) return;
  // Paired solution updates (p_prev set): an even iteration since the last
  // restart defers its x += alpha p, the next one applies both in sequence.
  const int j = st->iters - st->it0;
  const bool defer = p_prev && (j & 1) == 0;
  const bool pair = p_prev && (j & 1) == 1;
  const double alpha_prev = pair ? st->alpha_prev : 0.0;
  x += rc * n;
  r += rc * n;
  p += rc * n;
  if (p_prev) p_prev += rc * n;
  Ap += rc * n;
  if (r32) r32 += rc * n;
  if (jz) jz += rc * n;
  const double pAp = fixed_sum_all(pap + static_cast<long long>(rc) * pstride, npap);
  partials += static_cast<long long>(rc) * pstride;
  const bool first = blockIdx.x == 0 && threadIdx.x == 0;
  if (!(pAp > 0.0) || !isfinite(pAp)) {
    if (first) {
      st->pAp = pAp;
      st->alpha = 0.0;
      st->fail = 1;
      st->done = 1;
    }
    return;
  }
  const double alpha = st->rz_new / pAp;
  if (first) {
    st->pAp = pAp;
    st->alpha = alpha;
    st->pfresh = 0;
    if (defer) st->alpha_prev = alpha;
  }
  double s = 0.0, s2 = 0.0;
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    if (pair) {
      const double xp = x[i] + alpha_prev * static_cast<double>(p_prev[i]);
      x[i] = xp + alpha * static_cast<double>(p[i]);
    } else if (!defer) {
      x[i] += alpha * static_cast<double>(p[i]);
    }
    const double ri = r[i] - alpha * static_cast<double>(Ap[i]);
    r[i] = ri;
    if (r32) r32[i] = static_cast<float>(ri);
    s += ri * ri;
    if (jz) {
      // the Jacobi preconditioner z = D^-1 r and <r, z>, as jacobi0_k (omega 1)
      // and dot_k form them over the same block partition
      const double zi = 1.0 * jdinv[i] * ri;
      jz[i] = zi;
      s2 += ri * zi;
    }
  }
  s = block_sum(s);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
  if (jz) {
    __syncthreads();  // block_sum's LDS is reused
    s2 = block_sum(s2);
    if (threadIdx.x == 0) rzp[static_cast<long long>(rc) * pstride + blockIdx.x] = s2;
  }
}

// The deferred solution update a stopped recurrence owes (an odd count of
// iterations since its restart): x += alpha_prev p for the cases in the
// pending mask, p from buffer p1 for the cases in the parity mask (else p0);
// each case's first block then moves it0 to iters, so nothing is owed twice.
__global__ void cg_xfinish_k(CgState* st, double* __restrict__ x, const double* __restrict__ p0,
                             const double* __restrict__ p1, long long n, unsigned pend,
                             unsigned par) {
  const int rc = blockIdx.y;
  if (!((pend >> rc) & 1u)) return;
  st += rc;
  const double a = st->alpha_prev;
  const double* p = ((par >> rc) & 1u) ? p1 : p0;
  x += rc * n;
  p += rc * n;
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    x[i] += a * p[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) st->it0 = st->iters;
}

// The separate direction update (grids too large for the A.p to carry it):
// p_out = z + beta p_in with beta = <r,z>_new / rz, <r,z>_new the fixed-order
// sum of the case's partials (every block forms the same value; block 0
// records it and the breakdown test). restart: p_out = z.
template <typename TZ>
__global__ void cg_p_k(CgState* st, const TZ* __restrict__ z, const double* __restrict__ p_in,
                       double* __restrict__ p_out, long long n, int restart, const double* rzp,
                       int nrz, int pstride) {
  const int rc = blockIdx.y;
  st += rc;
  if (st->done) return;
  z += rc * n;
  p_in += rc * n;
  p_out += rc * n;
  double beta = 0.0;
  if (!restart) {
    const double rzn = fixed_sum_all(rzp + static_cast<long long>(rc) * pstride, nrz);
    const bool first = blockIdx.x == 0 && threadIdx.x == 0;
    if (!isfinite(rzn)) {
      if (first) {
        st->rz_new = rzn;
        st->beta = 0.0;
        st->fail = 1;
        st->done = 1;
      }
      return;
    }
    beta = rzn / st->rz;
    if (first) {
      st->rz_new = rzn;
      st->beta = beta;
    }
  }
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    p_out[i] = restart ? static_cast<double>(z[i]) : fma(beta, p_in[i], static_cast<double>(z[i]));
}

// ---- Chebyshev smoother ---------------------------------------------------------
// coef (device, per level): [0] lambda_max estimate, [1] 1/theta, then for
// step k = 1..degree-1 the pair (c1_k, c2_k) at [2k], [2k+1].
// (xo receives the updated x; it may be x itself.)
template <typename T>
__global__ void cheb_first_k(const T* x, T* xo, T* __restrict__ d, const T* r,
                             const T* __restrict__ dinv, const double* __restrict__ coef,
                             long long n, int zero_start, const int* done) {
  const int rc = blockIdx.y;
  if (case_done(done, rc)) return;
  x += rc * n;
  xo += rc * n;
  d += rc * n;
  r += rc * n;
  const double it = coef[1];
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    const T di = static_cast<T>(static_cast<double>(dinv[i]) * static_cast<double>(r[i]) * it);
    d[i] = di;
    xo[i] = zero_start ? di : x[i] + di;
  }
}

template <typename T>
__global__ void cheb_step_k(const T* x, T* xo, T* __restrict__ d, const T* r,
                            const T* __restrict__ dinv, const double* __restrict__ coef, int k,
                            long long n, const int* done) {
  const int rc = blockIdx.y;
  if (case_done(done, rc)) return;
  x += rc * n;
  xo += rc * n;
  d += rc * n;
  r += rc * n;
  const T c1 = static_cast<T>(coef[2 * k]), c2 = static_cast<T>(coef[2 * k + 1]);
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    const T di = c1 * d[i] + c2 * (dinv[i] * r[i]);
    d[i] = di;
    xo[i] = x[i] + di;
  }
}

// Power iteration of lambda_max(D^-1 A). start: v = cheb_start(3 node +
// comp) on kept DOFs (dinv != 0), 0 elsewhere.
template <typename T>
__global__ void cheb_start_k(T* __restrict__ v, const T* __restrict__ dinv, long long n) {
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    const unsigned h = static_cast<unsigned>(i + 1) * 2654435761u;
    const double v0 = static_cast<double>(h >> 8) * (1.0 / 16777216.0) - 0.5;
    v[i] = dinv[i] != T(0) ? static_cast<T>(v0) : T(0);
  }
}
// w = D^-1 y into v, partials of |w|^2.
template <typename T>
__global__ void cheb_scale_k(T* __restrict__ v, const T* __restrict__ y, const T* __restrict__ dinv,
                             long long n, double* __restrict__ partials) {
  double s = 0.0;
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    const double w = static_cast<double>(dinv[i]) * static_cast<double>(y[i]);
    v[i] = static_cast<T>(w);
    s += w * w;
  }
  s = block_sum(s);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}
// v /= ||v|| (every block re-sums the partials in a fixed order).
template <typename T>
__global__ void cheb_norm_k(T* __restrict__ v, long long n, const double* __restrict__ partials,
                            int np) {
  const double inv = 1.0 / sqrt(fixed_sum_all(partials, np));
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    v[i] = static_cast<T>(static_cast<double>(v[i]) * inv);
}
// partials of v.y (first np) and v.D v (next np), D = 1 / dinv on kept DOFs.
template <typename T>
__global__ void cheb_rq_k(const T* __restrict__ v, const T* __restrict__ y,
                          const T* __restrict__ dinv, long long n, double* __restrict__ partials) {
  double a = 0.0, b = 0.0;
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    const double vi = static_cast<double>(v[i]);
    a += vi * static_cast<double>(y[i]);
    if (dinv[i] != T(0)) b += vi * vi / static_cast<double>(dinv[i]);
  }
  a = block_sum(a);
  __syncthreads();
  b = block_sum(b);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = a;
    partials[gridDim.x + blockIdx.x] = b;
  }
}
// lambda = (v.Av) / (v.Dv); the interval [lo, hi] * lambda and the step
// coefficients of the degree-`degree` Chebyshev smoother.
__global__ void cheb_coef_k(const double* __restrict__ partials, int np, double lo, double hi,
                            int degree, double* __restrict__ coef) {
  const double vav = fixed_sum_all(partials, np);
  const double vdv = fixed_sum_all(partials + np, np);
  if (threadIdx.x != 0) return;
  const double lam = vav / vdv;
  const double theta = 0.5 * (hi + lo) * lam, delta = 0.5 * (hi - lo) * lam;
  coef[0] = lam;
  coef[1] = 1.0 / theta;
  const double sigma = theta / delta;
  double rho = 1.0 / sigma;
  for (int k = 1; k < degree; ++k) {
    const double rho_new = 1.0 / (2.0 * sigma - rho);
    coef[2 * k] = rho_new * rho;
    coef[2 * k + 1] = 2.0 * rho_new / delta;
    rho = rho_new;
  }
}

int grid_for(long long n, int bs, int cap) {
  long long b = (n + bs - 1) / bs;
  if (b > cap) b = cap;
  if (b < 1) b = 1;
  return static_cast<int>(b);
}

}  // namespace

template <typename Tf, typename Tc>
void launch_restrict(const Geom& gf, const Geom& gc, const Tf* rf, const uint8_t* mask_c, Tc* bc,
                     Tc* xc, const Tc* dinv_c, double omega, const int* done, hipStream_t s,
                     int nr, CrMap cm) {
  const long long n = gc.nodes();
  hipLaunchKernelGGL((restrict_k<Tf, Tc>), dim3(xcd_grid((n + 255) / 256), nr),
                     dim3(256), 0, s, gf, gc, rf, mask_c, bc, xc, dinv_c, omega, gf.ndof(),
                     gc.ndof(), cm, done);
  TOPOHIP_CHECK(hipGetLastError());
}

template <typename Tc, typename Tf>
void launch_prolong_add(const Geom& gc, const Geom& gf, const Tc* xc, const uint8_t* mask_f,
                        Tf* xf, const int* done, hipStream_t s, int nr, CrMap cm) {
  const long long n = gf.nodes();
  if (!cm.dst && n >= kPairProlongMin) {
    const long long pairs = static_cast<long long>((gf.nnx + 1) / 2) * gf.nny * gf.nnz;
    hipLaunchKernelGGL((prolong_add_pair_k<Tc, Tf>), dim3(xcd_grid((pairs + 255) / 256), nr),
                       dim3(256), 0, s, gc, gf, xc, mask_f, xf, gc.ndof(), gf.ndof(), done);
  } else {
    hipLaunchKernelGGL((prolong_add_k<Tc, Tf>), dim3(xcd_grid((n + 255) / 256), nr),
                       dim3(256), 0, s, gc, gf, xc, mask_f, xf, gc.ndof(), gf.ndof(), cm, done);
  }
  TOPOHIP_CHECK(hipGetLastError());
}

template void launch_restrict<float, float>(const Geom&, const Geom&, const float*, const uint8_t*,
                                            float*, float*, const float*, double, const int*,
                                            hipStream_t, int, CrMap);
template void launch_restrict<double, double>(const Geom&, const Geom&, const double*,
                                              const uint8_t*, double*, double*, const double*,
                                              double, const int*, hipStream_t, int, CrMap);
template void launch_restrict<float, double>(const Geom&, const Geom&, const float*,
                                             const uint8_t*, double*, double*, const double*,
                                             double, const int*, hipStream_t, int, CrMap);
template void launch_prolong_add<float, float>(const Geom&, const Geom&, const float*,
                                               const uint8_t*, float*, const int*, hipStream_t,
                                               int, CrMap);
template void launch_prolong_add<double, double>(const Geom&, const Geom&, const double*,
                                                 const uint8_t*, double*, const int*, hipStream_t,
                                                 int, CrMap);
template void launch_prolong_add<double, float>(const Geom&, const Geom&, const double*,
                                                const uint8_t*, float*, const int*, hipStream_t,
                                                int, CrMap);

void launch_sum_partials(const double* partials, int n, double* out, hipStream_t s, int nr,
                         int pstride) {
  hipLaunchKernelGGL(sum_partials_k, dim3(nr), dim3(RB), 0, s, partials, n, out, pstride);
  TOPOHIP_CHECK(hipGetLastError());
}
void launch_max_partials(const double* partials, int n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(max_partials_k, dim3(1), dim3(RB), 0, s, partials, n, out);
  TOPOHIP_CHECK(hipGetLastError());
}

int dot_blocks(long long n) { return grid_for(n, RB * 4, 1024); }

template <typename TA, typename TB>
void launch_dot(const TA* a, const TB* b, long long n, double* partials, int nblocks,
                const int* done, hipStream_t s, int nr, int pstride) {
  hipLaunchKernelGGL((dot_k<TA, TB>), dim3(nblocks, nr), dim3(RB), 0, s, a, b, n, partials,
                     pstride, done);
  TOPOHIP_CHECK(hipGetLastError());
}
template void launch_dot<double, double>(const double*, const double*, long long, double*, int,
                                         const int*, hipStream_t, int, int);
template void launch_dot<double, float>(const double*, const float*, long long, double*, int,
                                        const int*, hipStream_t, int, int);
template void launch_dot<float, float>(const float*, const float*, long long, double*, int,
                                       const int*, hipStream_t, int, int);

template <typename T>
void launch_sum_vec(const T* v, long long n, double* partials, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL((sum_vec_k<T>), dim3(nblocks), dim3(RB), 0, s, v, n, partials);
  TOPOHIP_CHECK(hipGetLastError());
}
template void launch_sum_vec<double>(const double*, long long, double*, int, hipStream_t);

template <typename T>
void launch_fill(T* p, T v, long long n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL((fill_k<T>), dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, p, v, n);
  TOPOHIP_CHECK(hipGetLastError());
}
template void launch_fill<float>(float*, float, long long, hipStream_t);
template void launch_fill<double>(double*, double, long long, hipStream_t);
template void launch_fill<uint8_t>(uint8_t*, uint8_t, long long, hipStream_t);

template <typename TI, typename TO>
void launch_convert(const TI* in, TO* out, long long n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL((convert_k<TI, TO>), dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, in, out,
                     n);
  TOPOHIP_CHECK(hipGetLastError());
}
template void launch_convert<double, float>(const double*, float*, long long, hipStream_t);
template void launch_convert<float, double>(const float*, double*, long long, hipStream_t);

template <typename T>
void launch_jacobi0(T* x, const T* dinv, const T* b, double omega, long long n, const int* done,
                    hipStream_t s, int nr) {
  hipLaunchKernelGGL((jacobi0_k<T>), dim3(grid_for(n, 256, 4096), nr), dim3(256), 0, s, x, dinv,
                     b, omega, n, done);
  TOPOHIP_CHECK(hipGetLastError());
}
template void launch_jacobi0<float>(float*, const float*, const float*, double, long long,
                                    const int*, hipStream_t, int);
template void launch_jacobi0<double>(double*, const double*, const double*, double, long long,
                                     const int*, hipStream_t, int);

void launch_add(const double* x, const double* up, double* y, long long n, hipStream_t s) {
  hipLaunchKernelGGL(add_k, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, x, up, y, n);
  TOPOHIP_CHECK(hipGetLastError());
}

void launch_cg_init(CgState* st, const double* p, int n, double tol, int cap, hipStream_t s,
                    int nr, int pstride) {
  hipLaunchKernelGGL(cg_init_k, dim3(nr), dim3(RB), 0, s, st, p, n, pstride, tol, cap);
  TOPOHIP_CHECK(hipGetLastError());
}
#define TOPOHIP_CG_SCALAR(NAME, KERNEL)                                                    \
  void NAME(CgState* st, const double* p, int n, hipStream_t s, int nr, int pstride) {     \
    hipLaunchKernelGGL(KERNEL, dim3(nr), dim3(RB), 0, s, st, p, n, pstride);               \
    TOPOHIP_CHECK(hipGetLastError());                                                      \
  }
TOPOHIP_CG_SCALAR(launch_cg_set_rz, cg_set_rz_k)
TOPOHIP_CG_SCALAR(launch_cg_resume, cg_resume_k)
TOPOHIP_CG_SCALAR(launch_cg_check, cg_check_k)
#undef TOPOHIP_CG_SCALAR

template <typename TP, typename TA>
void launch_cg_xr(CgState* st, double* x, double* r, const TP* p, const TA* Ap, float* r32,
                  long long n, const double* pap, int npap, double* partials, int nblocks,
                  hipStream_t s, int nr, int pstride, const double* jdinv, double* jz,
                  double* rzp, const TP* p_prev) {
  if (jz && (nblocks != dot_blocks(n) || !jdinv || !rzp))
    throw HipError("cg_xr: the Jacobi step needs D^-1, the <r,z> partials and dot_blocks(n) blocks");
  hipLaunchKernelGGL((cg_xr_k<TP, TA>), dim3(nblocks, nr), dim3(RB), 0, s, st, x, r, p, Ap, r32,
                     n, pap, npap, partials, pstride, jdinv, jz, rzp, p_prev);
  TOPOHIP_CHECK(hipGetLastError());
}
template void launch_cg_xr<double, double>(CgState*, double*, double*, const double*,
                                           const double*, float*, long long, const double*, int,
                                           double*, int, hipStream_t, int, int, const double*,
                                           double*, double*, const double*);
template void launch_cg_xr<float, float>(CgState*, double*, double*, const float*, const float*,
                                         float*, long long, const double*, int, double*, int,
                                         hipStream_t, int, int, const double*, double*, double*,
                                         const float*);

void launch_cg_xfinish(CgState* st, double* x, const double* p0, const double* p1, long long n,
                       unsigned pend, unsigned par, hipStream_t s, int nr) {
  if (!pend) return;
  hipLaunchKernelGGL(cg_xfinish_k, dim3(grid_for(n, 256, 2048), nr), dim3(256), 0, s, st, x, p0,
                     p1, n, pend, par);
  TOPOHIP_CHECK(hipGetLastError());
}

template <typename TZ>
void launch_cg_p(CgState* st, const TZ* z, const double* p_in, double* p_out, long long n,
                 int restart, const double* rzp, int nrz, hipStream_t s, int nr, int pstride) {
  // every block re-sums the <r,z> partials, so the grid stays at 1024 blocks
  hipLaunchKernelGGL((cg_p_k<TZ>), dim3(grid_for(n, 256, 1024), nr), dim3(256), 0, s, st, z, p_in,
                     p_out, n, restart, rzp, nrz, pstride);
  TOPOHIP_CHECK(hipGetLastError());
}
template void launch_cg_p<float>(CgState*, const float*, const double*, double*, long long, int,
                                 const double*, int, hipStream_t, int, int);
template void launch_cg_p<double>(CgState*, const double*, const double*, double*, long long, int,
                                  const double*, int, hipStream_t, int, int);

int cheb_blocks(long long n) { return grid_for(n, RB * 4, 1024); }

template <typename T>
void launch_cheb_first(const T* x, T* xo, T* d, const T* r, const T* dinv, const double* coef,
                       long long n, bool zero_start, const int* done, hipStream_t s, int nr) {
  hipLaunchKernelGGL((cheb_first_k<T>), dim3(grid_for(n, 256, 4096), nr), dim3(256), 0, s, x, xo,
                     d, r, dinv, coef, n, zero_start ? 1 : 0, done);
  TOPOHIP_CHECK(hipGetLastError());
}
template <typename T>
void launch_cheb_step(const T* x, T* xo, T* d, const T* r, const T* dinv, const double* coef,
                      int k, long long n, const int* done, hipStream_t s, int nr) {
  hipLaunchKernelGGL((cheb_step_k<T>), dim3(grid_for(n, 256, 4096), nr), dim3(256), 0, s, x, xo,
                     d, r, dinv, coef, k, n, done);
  TOPOHIP_CHECK(hipGetLastError());
}
template <typename T>
void launch_cheb_power(int step, T* v, const T* y, const T* dinv, long long n, double* partials,
                       double* coef, int degree, double lo, double hi, hipStream_t s) {
  const int nb = cheb_blocks(n);
  const dim3 g(nb), b(RB);
  if (step == 0) {
    hipLaunchKernelGGL((cheb_start_k<T>), dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, v, dinv,
                       n);
  } else if (step == 1) {
    hipLaunchKernelGGL((cheb_scale_k<T>), g, b, 0, s, v, y, dinv, n, partials);
    hipLaunchKernelGGL((cheb_norm_k<T>), g, b, 0, s, v, n, partials, nb);
  } else {
    hipLaunchKernelGGL((cheb_rq_k<T>), g, b, 0, s, v, y, dinv, n, partials);
    hipLaunchKernelGGL(cheb_coef_k, dim3(1), b, 0, s, partials, nb, lo, hi, degree, coef);
  }
  TOPOHIP_CHECK(hipGetLastError());
}
#define TOPOHIP_CHEB_INST(T)                                                                     \
  template void launch_cheb_first<T>(const T*, T*, T*, const T*, const T*, const double*,       \
                                     long long, bool, const int*, hipStream_t, int);             \
  template void launch_cheb_step<T>(const T*, T*, T*, const T*, const T*, const double*, int,    \
                                    long long, const int*, hipStream_t, int);                    \
  template void launch_cheb_power<T>(int, T*, const T*, const T*, long long, double*, double*,   \
                                     int, double, double, hipStream_t);
TOPOHIP_CHEB_INST(float)
TOPOHIP_CHEB_INST(double)
#undef TOPOHIP_CHEB_INST

}  // namespace topohip
