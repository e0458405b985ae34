// simp.hip — SIMP material law, compliance sensitivity, density filter and
// the optimality-criteria lambda search, all FP64.
//
// Restates core/src/simp/simp.cpp: the law :299-303 / :627-665, the
// sensitivity sweep :732-826, the cone filter and its transpose :834-944 and
// oc_update :948-1032. The filter and OC kernels evaluate the same operations
// in the same order as the reference (gather in the reference's neighbour
// order, no fused multiply-add), so the filtered fields and the OC candidate
// are bit-identical to an FP64 CPU evaluation on identical inputs.
#pragma clang fp contract(off)
#include <mutex>

#include "kernels.hpp"

namespace topohip {

namespace {

constexpr int RB = 256;

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
__device__ inline double block_sum(double v) {
  __shared__ double sh[RB / 64];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < RB / 64; ++w) s += sh[w];
  return s;
}
__device__ inline double block_max(double v) {
  __shared__ double sh[RB / 64];
  v = wave_max(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < RB / 64; ++w) s = fmax(s, sh[w]);
  return s;
}

// N block sums at once (one barrier); the results are valid in thread 0.
template <int N>
__device__ inline void block_sum_n(double (&v)[N]) {
  __shared__ double sh[N][RB / 64];
#pragma unroll
  for (int j = 0; j < N; ++j) {
    const double w = wave_sum(v[j]);
    if ((threadIdx.x & 63) == 0) sh[j][threadIdx.x >> 6] = w;
  }
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      double s = 0.0;
      for (int w = 0; w < RB / 64; ++w) s += sh[j][w];
      v[j] = s;
    }
}

__global__ void youngs_k(const double* __restrict__ xphys, const uint8_t* __restrict__ solid,
                         long long n, double E0, double p, double rho_min, double* __restrict__ E64,
                         float* __restrict__ E32) {
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < n;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    double v = 0.0;
    if (!solid || solid[e]) {
      const double rho = fmin(1.0, fmax(rho_min, xphys[e]));
      v = pow(rho, p) * E0;
    }
    if (E64) E64[e] = v;
    if (E32) E32[e] = static_cast<float>(v);
  }
}

#pragma clang fp contract(fast)
__global__ void sensitivity_k(Geom g, const double* __restrict__ u, const double* __restrict__ xphys,
                              const uint8_t* __restrict__ solid, Modal<double> mc, double E0,
                              double p, double rho_min, double* __restrict__ dc, int accumulate,
                              double* __restrict__ partials) {
  double cacc = 0.0;
  const long long ne = g.elems();
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < ne;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    if (solid && !solid[e]) {
      if (!accumulate) dc[e] = 0.0;
      continue;
    }
    const int i = static_cast<int>(e % g.ex);
    const int j = static_cast<int>((e / g.ex) % g.ey);
    const int k = static_cast<int>(e / (static_cast<long long>(g.ex) * g.ey));
    double a[24];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const long long node =
          (static_cast<long long>(k + (m >> 2)) * g.nny + j + ((m >> 1) & 1)) * g.nnx + i + (m & 1);
#pragma unroll
      for (int c = 0; c < 3; ++c) a[3 * m + c] = u[node * 3 + c];
    }
#pragma unroll
    for (int bit = 1; bit < 8; bit <<= 1)
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        if (m & bit) continue;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double x = a[3 * m + c], y = a[3 * (m | bit) + c];
          a[3 * m + c] = x + y;
          a[3 * (m | bit) + c] = x - y;
        }
      }
    double q = 0.0;
#define TOPOHIP_X(kk, r, s) q += mc.c[kk] * a[r] * a[s];
    TOPOHIP_MODAL_PATTERN(TOPOHIP_X)
#undef TOPOHIP_X
    const double rho = fmin(1.0, fmax(rho_min, xphys[e]));
    cacc += pow(rho, p) * E0 * q;
    const double d = -p * pow(rho, p - 1.0) * E0 * q;
    dc[e] = accumulate ? dc[e] + d : d;
  }
  cacc = block_sum(cacc);
  if (threadIdx.x == 0) partials[blockIdx.x] = cacc;
}
#pragma clang fp contract(off)

__global__ void filter_hs_k(Geom g, const int* __restrict__ off, const double* __restrict__ w,
                            int np, const uint8_t* __restrict__ solid, double* __restrict__ hs) {
  const long long n = g.elems();
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < n;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    if (solid && !solid[e]) { hs[e] = 0.0; continue; }
    const int i = static_cast<int>(e % g.ex);
    const int j = static_cast<int>((e / g.ex) % g.ey);
    const int k = static_cast<int>(e / (static_cast<long long>(g.ex) * g.ey));
    double s = 0.0;
    for (int p = 0; p < np; ++p) {
      const int ii = i + off[3 * p], jj = j + off[3 * p + 1], kk = k + off[3 * p + 2];
      if (ii < 0 || ii >= g.ex || jj < 0 || jj >= g.ey || kk < 0 || kk >= g.ez) continue;
      const long long q = (static_cast<long long>(kk) * g.ey + jj) * g.ex + ii;
      if (solid && !solid[q]) continue;
      s += w[p];
    }
    hs[e] = s;
  }
}

// Forward: out_e = (sum_p w_p x_{e+off_p}) / Hs_e in the reference's neighbour
// order (dk, dj, di ascending). Adjoint: out_j = sum over e = j - off_p in
// ascending e of w_p * (x_e / Hs_e), the order in which the reference's
// scatter accumulates into out[j].
__global__ void filter_k(Geom g, const int* __restrict__ off, const double* __restrict__ w, int np,
                         const uint8_t* __restrict__ solid, const double* __restrict__ hs,
                         const double* __restrict__ x, double* __restrict__ out, int adjoint) {
  const long long n = g.elems();
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < n;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    if (solid && !solid[e]) { out[e] = 0.0; continue; }
    const int i = static_cast<int>(e % g.ex);
    const int j = static_cast<int>((e / g.ex) % g.ey);
    const int k = static_cast<int>(e / (static_cast<long long>(g.ex) * g.ey));
    double acc = 0.0;
    if (!adjoint) {
      if (!(hs[e] > 0.0)) { out[e] = 0.0; continue; }
      for (int p = 0; p < np; ++p) {
        const int ii = i + off[3 * p], jj = j + off[3 * p + 1], kk = k + off[3 * p + 2];
        if (ii < 0 || ii >= g.ex || jj < 0 || jj >= g.ey || kk < 0 || kk >= g.ez) continue;
        const long long q = (static_cast<long long>(kk) * g.ey + jj) * g.ex + ii;
        if (solid && !solid[q]) continue;
        acc += w[p] * x[q];
      }
      out[e] = acc / hs[e];
    } else {
      for (int p = np - 1; p >= 0; --p) {
        const int ii = i - off[3 * p], jj = j - off[3 * p + 1], kk = k - off[3 * p + 2];
        if (ii < 0 || ii >= g.ex || jj < 0 || jj >= g.ey || kk < 0 || kk >= g.ez) continue;
        const long long q = (static_cast<long long>(kk) * g.ey + jj) * g.ex + ii;
        if (solid && !solid[q]) continue;
        if (!(hs[q] > 0.0)) continue;
        const double se = x[q] / hs[q];
        acc += w[p] * se;
      }
      out[e] = acc;
    }
  }
}

__device__ inline double oc_lambda(const OcState* st) {
  return st->phase == 0 ? st->l2 : 0.5 * (st->l1 + st->l2);
}

// OC candidate for one voxel: oc_update's `candidate` lambda
// (core/src/simp/simp.cpp:972-991). Returns the active-set code.
__device__ inline int oc_candidate(double xe, double dce, double dve, double lambda, double move,
                                   double rho_min, double& xt) {
  double be = 0.0;
  if (dve > 0.0) be = -dce / (lambda * dve);
  if (be < 0.0) be = 0.0;
  xt = xe * sqrt(be);
  const double lo = fmax(rho_min, xe - move);
  const double hi = fmin(1.0, xe + move);
  int act = 0;
  if (xt < lo) { xt = lo; act = 1; }
  if (xt > hi) { xt = hi; act = 2; }
  return act;
}

__global__ void oc_init_k(OcState* st, double target) {
  st->l1 = 0.0;
  st->l2 = 1.0;
  st->lmid = 0.0;
  st->target = target;
  st->vol = 0.0;
  st->phase = 0;
  st->steps = 0;
  st->gap = INFINITY;
}

__global__ void oc_trial_k(const double* __restrict__ x, const double* __restrict__ dcf,
                           const double* __restrict__ dv, const uint8_t* __restrict__ solid,
                           long long n, double move, double rho_min, const OcState* st,
                           double* __restrict__ partials) {
  if (st->phase == 2) return;
  const double lambda = oc_lambda(st);
  double s = 0.0;
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < n;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    if (solid && !solid[e]) continue;
    double xt;
    oc_candidate(x[e], dcf[e], dv[e], lambda, move, rho_min, xt);
    s += xt * dv[e];
  }
  s = block_sum(s);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// Advance the lambda search by one trial: the bracket loop `l2 *= 2 while
// vol(l2) > target && l2 < 1e30`, then at most 100 bisections while
// (l2 - l1) > 1e-6 (l1 + l2).
__global__ void oc_decide_k(OcState* st, const double* __restrict__ partials, int n) {
  if (st->phase == 2) return;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += partials[i];
  s = block_sum(s);
  if (threadIdx.x != 0) return;
  st->vol = s;
  st->gap = fmin(st->gap, fabs(s - st->target));
  if (st->phase == 0) {
    if (s > st->target) {
      st->l2 *= 2.0;
      if (st->l2 < 1e30) return;  // bracket again
    }
    st->phase = 1;
  } else {
    const double lmid = 0.5 * (st->l1 + st->l2);
    if (s > st->target) st->l1 = lmid; else st->l2 = lmid;
    st->steps += 1;
  }
  if (!(st->steps < 100 && (st->l2 - st->l1) > 1e-6 * (st->l1 + st->l2))) st->phase = 2;
}

// Lambdas of the next pass: the bracket doublings l2, 2 l2, ... below 1e30,
// or the midpoints of the next three bisection levels in heap order (node j's
// children 2j+1: lower half, 2j+2: upper half), each formed as the sequential
// search forms it, 0.5 (l1 + l2) of the interval it would hold.
__device__ inline void oc_plan(OcState* st) {
  if (st->phase == 0) {
    double l = st->l2;
    int k = 0;
    while (k < kLamPass && l < 1e30) {
      st->lam[k++] = l;
      l *= 2.0;
    }
    st->nlam = k;
  } else if (st->phase == 1) {
    const double l1 = st->l1, l2 = st->l2;
    const double m0 = 0.5 * (l1 + l2);
    const double mL = 0.5 * (l1 + m0), mR = 0.5 * (m0 + l2);
    st->lam[0] = m0;
    st->lam[1] = mL;
    st->lam[2] = mR;
    st->lam[3] = 0.5 * (l1 + mL);
    st->lam[4] = 0.5 * (mL + m0);
    st->lam[5] = 0.5 * (m0 + mR);
    st->lam[6] = 0.5 * (mR + l2);
    st->nlam = kLamPass;
  } else {
    st->nlam = 0;
  }
}

__global__ void oc_init_multi_k(OcState* st, double target) {
  st->l1 = 0.0;
  st->l2 = 1.0;
  st->lmid = 0.0;
  st->target = target;
  st->vol = 0.0;
  st->phase = 0;
  st->steps = 0;
  st->passes = 0;
  st->gap = INFINITY;
  oc_plan(st);
}

__global__ void oc_trial_multi_k(const double* __restrict__ x, const double* __restrict__ dcf,
                                 const double* __restrict__ dv, const uint8_t* __restrict__ solid,
                                 long long n, double move, double rho_min, const OcState* st,
                                 double* __restrict__ partials) {
  if (st->phase == 2) return;
  const int nl = st->nlam;
  double lam[kLamPass], s[kLamPass];
#pragma unroll
  for (int j = 0; j < kLamPass; ++j) {
    lam[j] = j < nl ? st->lam[j] : 1.0;
    s[j] = 0.0;
  }
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < n;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    if (solid && !solid[e]) continue;
    const double xe = x[e], dce = dcf[e], dve = dv[e];
#pragma unroll
    for (int j = 0; j < kLamPass; ++j) {
      if (j >= nl) break;
      double xt;
      oc_candidate(xe, dce, dve, lam[j], move, rho_min, xt);
      s[j] += xt * dve;
    }
  }
  block_sum_n<kLamPass>(s);
  if (threadIdx.x == 0)
    for (int j = 0; j < nl; ++j) partials[j * gridDim.x + blockIdx.x] = s[j];
}

// The sequential search's decisions over the pass's volumes: bracketing
// doubles l2 while vol(l2) > target and l2 < 1e30; bisection moves l1 or l2 to
// the midpoint; both stop as oc_decide_k does.
__global__ void oc_decide_multi_k(OcState* st, const double* __restrict__ partials, int n) {
  if (st->phase == 2) return;
  const int nl = st->nlam;
  double S[kLamPass];
#pragma unroll
  for (int j = 0; j < kLamPass; ++j) {
    double a = 0.0;
    if (j < nl)
      for (int i = threadIdx.x; i < n; i += blockDim.x) a += partials[j * n + i];
    S[j] = a;
  }
  block_sum_n<kLamPass>(S);
  if (threadIdx.x != 0) return;
  st->passes += 1;
  auto stop = [&] { return !(st->steps < 100 && (st->l2 - st->l1) > 1e-6 * (st->l1 + st->l2)); };
  if (st->phase == 0) {
    for (int j = 0; j < nl; ++j) {
      st->vol = S[j];
      st->gap = fmin(st->gap, fabs(S[j] - st->target));
      if (S[j] > st->target) {
        st->l2 *= 2.0;
        if (st->l2 < 1e30) continue;
      }
      st->phase = 1;
      if (stop()) st->phase = 2;
      break;
    }
  } else {
    int node = 0;
    for (int lev = 0; lev < 3; ++lev) {
      const double lmid = st->lam[node];
      st->vol = S[node];
      st->gap = fmin(st->gap, fabs(S[node] - st->target));
      if (S[node] > st->target) {
        st->l1 = lmid;
        node = 2 * node + 2;
      } else {
        st->l2 = lmid;
        node = 2 * node + 1;
      }
      st->steps += 1;
      if (stop()) {
        st->phase = 2;
        break;
      }
    }
  }
  oc_plan(st);
}

__global__ void oc_final_k(const double* __restrict__ x, const double* __restrict__ dcf,
                           const double* __restrict__ dv, const uint8_t* __restrict__ solid,
                           long long n, double move, double rho_min, const OcState* st,
                           double* __restrict__ xnew, uint8_t* __restrict__ active,
                           double* __restrict__ partials, int keep) {
  const double lambda = 0.5 * (st->l1 + st->l2);
  double m = 0.0;
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < n;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    if (solid && !solid[e]) {  // plain: 0 off the design set; masked: x carried through
      const double xo = keep ? x[e] : 0.0;
      xnew[e] = xo;
      if (active) active[e] = 0;
      m = fmax(m, fabs(xo - x[e]));
      continue;
    }
    double xt;
    const int act = oc_candidate(x[e], dcf[e], dv[e], lambda, move, rho_min, xt);
    xnew[e] = xt;
    if (active) active[e] = static_cast<uint8_t>(act);
    m = fmax(m, fabs(xt - x[e]));
  }
  m = block_max(m);
  if (threadIdx.x == 0) partials[blockIdx.x] = m;
}

int grid_for(long long n, int bs, int cap) {
  long long b = (n + bs - 1) / bs;
  if (b > cap) b = cap;
  if (b < 1) b = 1;
  return static_cast<int>(b);
}

}  // namespace

void launch_youngs(const double* xphys, const uint8_t* solid, long long n, double E0,
                   double penalty, double rho_min, double* E64, float* E32, hipStream_t s) {
  hipLaunchKernelGGL(youngs_k, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, xphys, solid, n, E0,
                     penalty, rho_min, E64, E32);
  TOPOHIP_CHECK(hipGetLastError());
}

int sens_blocks(const Geom& g) { return grid_for(g.elems(), RB, 2048); }

void launch_sensitivity(const Geom& g, const double* u, const double* xphys, const uint8_t* solid,
                        const Modal<double>& mc, double E0, double penalty, double rho_min,
                        double* dc, int accumulate, double* partials, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(sensitivity_k, dim3(nblocks), dim3(RB), 0, s, g, u, xphys, solid, mc, E0,
                     penalty, rho_min, dc, accumulate, partials);
  TOPOHIP_CHECK(hipGetLastError());
}

void launch_filter_hs(const Geom& g, const Stencil& st, const uint8_t* solid, double* hs,
                      hipStream_t s) {
  hipLaunchKernelGGL(filter_hs_k, dim3(grid_for(g.elems(), 256, 8192)), dim3(256), 0, s, g, st.off,
                     st.w, st.n, solid, hs);
  TOPOHIP_CHECK(hipGetLastError());
}

void launch_filter(const Geom& g, const Stencil& st, const uint8_t* solid, const double* hs,
                   const double* x, double* out, int adjoint, hipStream_t s) {
  hipLaunchKernelGGL(filter_k, dim3(grid_for(g.elems(), 256, 8192)), dim3(256), 0, s, g, st.off,
                     st.w, st.n, solid, hs, x, out, adjoint);
  TOPOHIP_CHECK(hipGetLastError());
}

void launch_oc_init(OcState* st, double target, hipStream_t s) {
  hipLaunchKernelGGL(oc_init_k, dim3(1), dim3(1), 0, s, st, target);
  TOPOHIP_CHECK(hipGetLastError());
}

int oc_blocks(long long n) { return grid_for(n, RB * 4, kOcMaxBlocks); }

// ---- work memory of the allocation-free OC update ------------------------------
// One array per GPU for the life of the process (kNoAllocScratchBytes, about
// 16 MiB). Host code never passes `scratch` itself to a kernel: its name is a
// host-side stand-in, so every use goes through the address from
// hipGetSymbolAddress below. Every call on one GPU shares the array; calls
// that overlap on one GPU race.
namespace {
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

void* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) throw HipError("noalloc_scratch: device id out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = 0;
    TOPOHIP_CHECK(hipGetDevice(&cur));
    if (cur != device) TOPOHIP_CHECK(hipSetDevice(device));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != device) TOPOHIP_CHECK(hipSetDevice(cur));
    TOPOHIP_CHECK(e);
    if (!ptr) throw HipError("noalloc_scratch: hipGetSymbolAddress returned no address");
    g_scratch_ptr[device] = ptr;
  }
  return g_scratch_ptr[device];
}

void launch_oc_trial(const double* x, const double* dcf, const double* dv, const uint8_t* solid,
                     long long n, double move, double rho_min, const OcState* st,
                     double* partials, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(oc_trial_k, dim3(nblocks), dim3(RB), 0, s, x, dcf, dv, solid, n, move,
                     rho_min, st, partials);
  TOPOHIP_CHECK(hipGetLastError());
}

void launch_oc_init_multi(OcState* st, double target, hipStream_t s) {
  hipLaunchKernelGGL(oc_init_multi_k, dim3(1), dim3(1), 0, s, st, target);
  TOPOHIP_CHECK(hipGetLastError());
}
void launch_oc_trial_multi(const double* x, const double* dcf, const double* dv,
                           const uint8_t* solid, long long n, double move, double rho_min,
                           const OcState* st, double* partials, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(oc_trial_multi_k, dim3(nblocks), dim3(RB), 0, s, x, dcf, dv, solid, n, move,
                     rho_min, st, partials);
  TOPOHIP_CHECK(hipGetLastError());
}
void launch_oc_decide_multi(OcState* st, const double* partials, int n, hipStream_t s) {
  hipLaunchKernelGGL(oc_decide_multi_k, dim3(1), dim3(RB), 0, s, st, partials, n);
  TOPOHIP_CHECK(hipGetLastError());
}

void launch_oc_decide(OcState* st, const double* partials, int n, hipStream_t s) {
  hipLaunchKernelGGL(oc_decide_k, dim3(1), dim3(RB), 0, s, st, partials, n);
  TOPOHIP_CHECK(hipGetLastError());
}

void launch_oc_final(const double* x, const double* dcf, const double* dv, const uint8_t* solid,
                     long long n, double move, double rho_min, const OcState* st, double* xnew,
                     uint8_t* active, double* partials, int nblocks, hipStream_t s, int keep) {
  hipLaunchKernelGGL(oc_final_k, dim3(nblocks), dim3(RB), 0, s, x, dcf, dv, solid, n, move, rho_min,
                     st, xnew, active, partials, keep);
  TOPOHIP_CHECK(hipGetLastError());
}

// ---- Heaviside projection (core/src/simp/simp.cpp:352-379) and the projected
// OC update (oc_update_projected, :1647-1716) -----------------------------------
namespace {
__device__ inline double hproj(double rt, double beta, double eta) {
  const double den = tanh(beta * eta) + tanh(beta * (1.0 - eta));
  return (tanh(beta * eta) + tanh(beta * (rt - eta))) / den;
}
__device__ inline double hproj_d(double rt, double beta, double eta) {
  const double den = tanh(beta * eta) + tanh(beta * (1.0 - eta));
  const double ch = cosh(beta * (rt - eta));
  return beta / (ch * ch * den);
}
// Projected OC candidate: UNCLAMPED dv in the ratio; a non-finite ratio (0/0)
// holds the voxel (Be = 1).
__device__ inline double oc_candidate_proj(double xe, double dce, double dve, double lambda,
                                           double move, double rho_min) {
  double be = -dce / (lambda * dve);
  if (!isfinite(be)) be = 1.0;
  if (be < 0.0) be = 0.0;
  double xt = xe * sqrt(be);
  const double lo = fmax(rho_min, xe - move);
  const double hi = fmin(1.0, xe + move);
  if (xt < lo) xt = lo;
  if (xt > hi) xt = hi;
  return xt;
}

// project_solid: design voxels projected, the rest copied.
__global__ void project_k(const double* __restrict__ xt, double* __restrict__ xp,
                          const uint8_t* __restrict__ solid, long long n, double beta, double eta) {
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < n;
       e += static_cast<long long>(gridDim.x) * blockDim.x)
    xp[e] = (!solid || solid[e]) ? hproj(xt[e], beta, eta) : xt[e];
}
// dproj = rho_bar'(xtilde), dc_phys = dc * dproj on design voxels (0 elsewhere).
__global__ void proj_chain_k(const double* __restrict__ xt, const double* __restrict__ dc,
                             const uint8_t* __restrict__ solid, long long n, double beta,
                             double eta, double* __restrict__ dcp, double* __restrict__ dpr) {
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < n;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    const bool d = !solid || solid[e];
    const double g = d ? hproj_d(xt[e], beta, eta) : 0.0;
    dpr[e] = g;
    dcp[e] = d ? dc[e] * g : 0.0;
  }
}
// Candidate of the current lambda trial (0 off the design set).
__global__ void oc_cand_proj_k(const double* __restrict__ x, const double* __restrict__ dcf,
                               const double* __restrict__ dvf, const uint8_t* __restrict__ solid,
                               long long n, double move, double rho_min, const OcState* st,
                               double* __restrict__ xc) {
  if (st->phase == 2) return;
  const double lambda = oc_lambda(st);
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < n;
       e += static_cast<long long>(gridDim.x) * blockDim.x)
    xc[e] = (!solid || solid[e]) ? oc_candidate_proj(x[e], dcf[e], dvf[e], lambda, move, rho_min)
                                 : 0.0;
}
// Projected volume partials of a filtered candidate, over design voxels.
__global__ void proj_sum_k(const double* __restrict__ xtc, const uint8_t* __restrict__ solid,
                           long long n, double beta, double eta, const OcState* st,
                           double* __restrict__ partials) {
  if (st->phase == 2) return;
  double s = 0.0;
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < n;
       e += static_cast<long long>(gridDim.x) * blockDim.x)
    if (!solid || solid[e]) s += hproj(xtc[e], beta, eta);
  s = block_sum(s);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}
// (keep: voxels off the design set keep their x, as the masked loop's frozen
// voxels do; otherwise they are 0, as off the solid set)
__global__ void oc_final_proj_k(const double* __restrict__ x, const double* __restrict__ dcf,
                                const double* __restrict__ dvf, const uint8_t* __restrict__ solid,
                                long long n, double move, double rho_min, const OcState* st,
                                double* __restrict__ xnew, double* __restrict__ partials,
                                int keep) {
  const double lambda = 0.5 * (st->l1 + st->l2);
  double m = 0.0;
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < n;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    const double xt = (!solid || solid[e])
                          ? oc_candidate_proj(x[e], dcf[e], dvf[e], lambda, move, rho_min)
                          : (keep ? x[e] : 0.0);
    xnew[e] = xt;
    m = fmax(m, fabs(xt - x[e]));
  }
  m = block_max(m);
  if (threadIdx.x == 0) partials[blockIdx.x] = m;
}
}  // namespace

void launch_project(const double* xt, double* xp, const uint8_t* solid, long long n, double beta,
                    double eta, hipStream_t s) {
  hipLaunchKernelGGL(project_k, dim3(grid_for(n, RB, 4096)), dim3(RB), 0, s, xt, xp, solid, n,
                     beta, eta);
  TOPOHIP_CHECK(hipGetLastError());
}
void launch_proj_chain(const double* xt, const double* dc, const uint8_t* solid, long long n,
                       double beta, double eta, double* dcp, double* dpr, hipStream_t s) {
  hipLaunchKernelGGL(proj_chain_k, dim3(grid_for(n, RB, 4096)), dim3(RB), 0, s, xt, dc, solid, n,
                     beta, eta, dcp, dpr);
  TOPOHIP_CHECK(hipGetLastError());
}
void launch_oc_cand_proj(const double* x, const double* dcf, const double* dvf,
                         const uint8_t* solid, long long n, double move, double rho_min,
                         const OcState* st, double* xc, hipStream_t s) {
  hipLaunchKernelGGL(oc_cand_proj_k, dim3(grid_for(n, RB, 4096)), dim3(RB), 0, s, x, dcf, dvf,
                     solid, n, move, rho_min, st, xc);
  TOPOHIP_CHECK(hipGetLastError());
}
void launch_proj_sum(const double* xtc, const uint8_t* solid, long long n, double beta, double eta,
                     const OcState* st, double* partials, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(proj_sum_k, dim3(nblocks), dim3(RB), 0, s, xtc, solid, n, beta, eta, st,
                     partials);
  TOPOHIP_CHECK(hipGetLastError());
}
void launch_oc_final_proj(const double* x, const double* dcf, const double* dvf,
                          const uint8_t* solid, long long n, double move, double rho_min,
                          const OcState* st, double* xnew, double* partials, int nblocks,
                          hipStream_t s, bool keep) {
  hipLaunchKernelGGL(oc_final_proj_k, dim3(nblocks), dim3(RB), 0, s, x, dcf, dvf, solid, n, move,
                     rho_min, st, xnew, partials, keep ? 1 : 0);
  TOPOHIP_CHECK(hipGetLastError());
}

// apply_mask_pins (core/src/simp/simp.cpp:2649-2654): FrozenSolid -> 1, FrozenVoid -> 0.
__global__ void mask_pins_k(double* __restrict__ xphys, const uint8_t* __restrict__ eff,
                            long long n) {
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < n;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    if (eff[e] == 1) xphys[e] = 1.0;
    else if (eff[e] == 2) xphys[e] = 0.0;
  }
}
void launch_mask_pins(double* xphys, const uint8_t* eff, long long n, hipStream_t s) {
  hipLaunchKernelGGL(mask_pins_k, dim3(grid_for(n, RB, 4096)), dim3(RB), 0, s, xphys, eff, n);
  TOPOHIP_CHECK(hipGetLastError());
}

// ---- MMA (Svanberg 1987), one volume constraint ---------------------------------
// mma_update (core/src/simp/simp.cpp:1433-1488) over mma_update_joint
// (core/src/simp/mma_joint.cpp:18-118): asymptotes, the separable convex
// coefficients and the move box per design voxel, then the 1-D dual search
// for the volume multiplier with its decisions taken on the device.
namespace {

// per-block max |dcf| over the design set (the scale-invariance normalisation)
__global__ void mma_scale_k(const double* __restrict__ dcf, const uint8_t* __restrict__ dset,
                            long long n, double* __restrict__ partials) {
  double m = 0.0;
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < n;
       e += static_cast<long long>(gridDim.x) * blockDim.x)
    if (!dset || dset[e]) m = fmax(m, fabs(dcf[e]));
  m = block_max(m);
  if (threadIdx.x == 0) partials[blockIdx.x] = m;
}

__global__ void mma_coef_k(MmaArgs a, const double* __restrict__ x, const double* __restrict__ x1,
                           const double* __restrict__ x2, const double* __restrict__ dcf,
                           const double* __restrict__ dv, const uint8_t* __restrict__ dset,
                           long long n, int iter, double move, double xmin,
                           const double* __restrict__ scale, double* __restrict__ partials) {
  const double asyinit = 0.5, asyincr = 1.2, asydecr = 0.7, albefa = 0.1, raa0 = 1e-5;
  const double xmax = 1.0, xrange = xmax - xmin;
  const double sc = *scale;
  double bsum = 0.0;
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < n;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    if (dset && !dset[e]) continue;
    const double xe = x[e];
    double L, U;
    if (iter <= 2) {
      L = xe - asyinit * xrange;
      U = xe + asyinit * xrange;
    } else {
      const double o1 = x1[e], o2 = x2[e];
      const double sgn = (xe - o1) * (o1 - o2);
      const double gam = sgn < 0.0 ? asydecr : (sgn > 0.0 ? asyincr : 1.0);
      L = xe - gam * (o1 - a.low[e]);
      U = xe + gam * (a.upp[e] - o1);
      L = fmin(L, xe - 0.01 * xrange);
      L = fmax(L, xe - 10.0 * xrange);
      U = fmax(U, xe + 0.01 * xrange);
      U = fmin(U, xe + 10.0 * xrange);
    }
    a.low[e] = L;
    a.upp[e] = U;
    const double d = sc > 0.0 ? dcf[e] / sc : dcf[e];
    const double ux1 = U - xe, xl1 = xe - L;
    const double dcp = fmax(d, 0.0), dcm = fmax(-d, 0.0);
    const double dvp = fmax(dv[e], 0.0), dvm = fmax(-dv[e], 0.0);
    const double r = raa0 / xrange;
    const double p0 = ux1 * ux1 * (1.001 * dcp + 0.001 * dcm + r);
    const double q0 = xl1 * xl1 * (0.001 * dcp + 1.001 * dcm + r);
    const double p1 = ux1 * ux1 * (1.001 * dvp + 0.001 * dvm + r);
    const double q1 = xl1 * xl1 * (0.001 * dvp + 1.001 * dvm + r);
    a.p0[e] = p0;
    a.q0[e] = q0;
    a.p1[e] = p1;
    a.q1[e] = q1;
    bsum += p1 / ux1 + q1 / xl1;
    a.al[e] = fmax(fmax(xmin, L + albefa * (xe - L)), xe - move * xrange);
    a.be[e] = fmin(fmin(xmax, U - albefa * (U - xe)), xe + move * xrange);
  }
  bsum = block_sum(bsum);
  if (threadIdx.x == 0) partials[blockIdx.x] = bsum;
}

// Lambdas of the next pass (as oc_plan): first 0 and the doublings 1, 2, 4,
// ... that follow a positive g(0); then further doublings of l2 while below
// 1e30; then three bisection levels of midpoints in heap order.
__device__ inline void mma_plan(MmaSearch* st) {
  if (st->phase == 0) {
    st->lam[0] = 0.0;
    double l = 1.0;
    for (int k = 1; k < kLamPass; ++k) {
      st->lam[k] = l;
      l *= 2.0;
    }
    st->nlam = kLamPass;
  } else if (st->phase == 1) {
    double l = st->l2;
    int k = 0;
    while (k < kLamPass && l < 1e30) {
      st->lam[k++] = l;
      l *= 2.0;
    }
    st->nlam = k;
  } else if (st->phase == 2) {
    const double l1 = st->l1, l2 = st->l2;
    const double m0 = 0.5 * (l1 + l2);
    const double mL = 0.5 * (l1 + m0), mR = 0.5 * (m0 + l2);
    st->lam[0] = m0;
    st->lam[1] = mL;
    st->lam[2] = mR;
    st->lam[3] = 0.5 * (l1 + mL);
    st->lam[4] = 0.5 * (mL + m0);
    st->lam[5] = 0.5 * (m0 + mR);
    st->lam[6] = 0.5 * (mR + l2);
    st->nlam = kLamPass;
  } else {
    st->nlam = 0;
  }
}

// b = -(vol - target) + sum of the coefficient partials; the search starts at
// lambda = 0.
__global__ void mma_init_k(MmaSearch* st, const double* __restrict__ partials, int n,
                           const double* __restrict__ vol, double target) {
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += partials[i];
  s = block_sum(s);
  if (threadIdx.x != 0) return;
  st->b = -(*vol - target) + s;
  st->l1 = 0.0;
  st->l2 = 1.0;
  st->lambda = 0.0;
  st->phase = 0;
  st->it = 0;
  st->steps = 0;
  mma_plan(st);
}

__device__ inline double mma_candidate(const MmaArgs& a, long long e, double lam) {
  const double pl = sqrt(a.p0[e] + lam * a.p1[e]);
  const double ql = sqrt(a.q0[e] + lam * a.q1[e]);
  double xt = (pl * a.low[e] + ql * a.upp[e]) / (pl + ql);
  if (xt < a.al[e]) xt = a.al[e];
  if (xt > a.be[e]) xt = a.be[e];
  return xt;
}

__global__ void mma_trial_k(MmaArgs a, const uint8_t* __restrict__ dset, long long n,
                            const MmaSearch* st, double* __restrict__ partials) {
  if (st->phase == 3) return;
  const int nl = st->nlam;
  double lam[kLamPass], s[kLamPass];
#pragma unroll
  for (int j = 0; j < kLamPass; ++j) {
    lam[j] = j < nl ? st->lam[j] : 0.0;
    s[j] = 0.0;
  }
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < n;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    if (dset && !dset[e]) continue;
    const double L = a.low[e], U = a.upp[e], p1 = a.p1[e], q1 = a.q1[e];
#pragma unroll
    for (int j = 0; j < kLamPass; ++j) {
      if (j >= nl) break;
      const double xt = mma_candidate(a, e, lam[j]);
      s[j] += p1 / (U - xt) + q1 / (xt - L);
    }
  }
  block_sum_n<kLamPass>(s);
  if (threadIdx.x == 0)
    for (int j = 0; j < nl; ++j) partials[j * gridDim.x + blockIdx.x] = s[j];
}

// The sequential search's decisions over the pass's values of g: g(0) > 0
// opens the bracket [0, 1], doubled while l2 < 1e30 and g(l2) > 0; then at
// most 100 bisections while (l2 - l1) > 1e-9 (1 + l1 + l2); lambda =
// (l1 + l2) / 2 (0 when g(0) <= 0).
__global__ void mma_decide_k(MmaSearch* st, const double* __restrict__ partials, int n) {
  if (st->phase == 3) return;
  const int nl = st->nlam;
  double S[kLamPass];
#pragma unroll
  for (int j = 0; j < kLamPass; ++j) {
    double a = 0.0;
    if (j < nl)
      for (int i = threadIdx.x; i < n; i += blockDim.x) a += partials[j * n + i];
    S[j] = a;
  }
  block_sum_n<kLamPass>(S);
  if (threadIdx.x != 0) return;
  auto bisect_or_finish = [&]() {
    if (st->it < 100 && (st->l2 - st->l1) > 1e-9 * (1.0 + st->l1 + st->l2)) {
      st->phase = 2;
    } else {
      st->lambda = 0.5 * (st->l1 + st->l2);
      st->phase = 3;
    }
  };
  int j = 0;
  if (st->phase == 0) {
    st->steps += 1;
    if (!(S[0] - st->b > 0.0)) {
      st->lambda = 0.0;
      st->phase = 3;
    } else {
      st->l1 = 0.0;
      st->l2 = 1.0;
      st->phase = 1;
    }
    j = 1;
  }
  if (st->phase == 1) {
    for (; j < nl; ++j) {
      st->steps += 1;
      if (S[j] - st->b > 0.0) {
        st->l2 *= 2.0;
        if (st->l2 < 1e30) continue;
      }
      bisect_or_finish();
      break;
    }
  } else if (st->phase == 2) {
    int node = 0;
    for (int lev = 0; lev < 3 && st->phase == 2; ++lev) {
      st->steps += 1;
      const double lm = st->lam[node];
      if (S[node] - st->b > 0.0) {
        st->l1 = lm;
        node = 2 * node + 2;
      } else {
        st->l2 = lm;
        node = 2 * node + 1;
      }
      st->it += 1;
      bisect_or_finish();
    }
  }
  mma_plan(st);
}

// xnew = candidate(lambda) on the design set, 0 elsewhere; change partials.
__global__ void mma_final_k(MmaArgs a, const double* __restrict__ x,
                            const uint8_t* __restrict__ dset, long long n, const MmaSearch* st,
                            double* __restrict__ xnew, double* __restrict__ partials) {
  const double lam = st->lambda;
  double m = 0.0;
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < n;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    const double xt = (!dset || dset[e]) ? mma_candidate(a, e, lam) : 0.0;
    xnew[e] = xt;
    m = fmax(m, fabs(xt - x[e]));
  }
  m = block_max(m);
  if (threadIdx.x == 0) partials[blockIdx.x] = m;
}

}  // namespace

void launch_mma_scale(const double* dcf, const uint8_t* dset, long long n, double* partials,
                      int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(mma_scale_k, dim3(nblocks), dim3(RB), 0, s, dcf, dset, n, partials);
  TOPOHIP_CHECK(hipGetLastError());
}
void launch_mma_coef(const MmaArgs& a, const double* x, const double* x1, const double* x2,
                     const double* dcf, const double* dv, const uint8_t* dset, long long n,
                     int iter, double move, double xmin, const double* scale, double* partials,
                     int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(mma_coef_k, dim3(nblocks), dim3(RB), 0, s, a, x, x1, x2, dcf, dv, dset, n,
                     iter, move, xmin, scale, partials);
  TOPOHIP_CHECK(hipGetLastError());
}
void launch_mma_init(MmaSearch* st, const double* partials, int n, const double* vol,
                     double target, hipStream_t s) {
  hipLaunchKernelGGL(mma_init_k, dim3(1), dim3(RB), 0, s, st, partials, n, vol, target);
  TOPOHIP_CHECK(hipGetLastError());
}
void launch_mma_trial(const MmaArgs& a, const uint8_t* dset, long long n, const MmaSearch* st,
                      double* partials, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(mma_trial_k, dim3(nblocks), dim3(RB), 0, s, a, dset, n, st, partials);
  TOPOHIP_CHECK(hipGetLastError());
}
void launch_mma_decide(MmaSearch* st, const double* partials, int n, hipStream_t s) {
  hipLaunchKernelGGL(mma_decide_k, dim3(1), dim3(RB), 0, s, st, partials, n);
  TOPOHIP_CHECK(hipGetLastError());
}
namespace {
// oscillation_fraction's counts (core/src/simp/simp.cpp:262-276) per block:
// design voxels that moved on both of the last two steps, and those whose
// direction reversed (integers, summed exactly as doubles).
__global__ void osc_count_k(const double* __restrict__ x, const double* __restrict__ x1,
                            const double* __restrict__ x2, const uint8_t* __restrict__ dset,
                            long long n, double* __restrict__ moved,
                            double* __restrict__ reversed) {
  double m = 0.0, r = 0.0;
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < n;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    if (dset && !dset[e]) continue;
    const double d1 = x[e] - x1[e], d2 = x1[e] - x2[e];
    if (d1 == 0.0 || d2 == 0.0) continue;
    m += 1.0;
    if (d1 * d2 < 0.0) r += 1.0;
  }
  m = block_sum(m);
  __syncthreads();  // block_sum's LDS is reused
  r = block_sum(r);
  if (threadIdx.x == 0) {
    moved[blockIdx.x] = m;
    reversed[blockIdx.x] = r;
  }
}
}  // namespace

void launch_osc_count(const double* x, const double* x1, const double* x2, const uint8_t* dset,
                      long long n, double* moved, double* reversed, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(osc_count_k, dim3(nblocks), dim3(RB), 0, s, x, x1, x2, dset, n, moved,
                     reversed);
  TOPOHIP_CHECK(hipGetLastError());
}

void launch_mma_final(const MmaArgs& a, const double* x, const uint8_t* dset, long long n,
                      const MmaSearch* st, double* xnew, double* partials, int nblocks,
                      hipStream_t s) {
  hipLaunchKernelGGL(mma_final_k, dim3(nblocks), dim3(RB), 0, s, a, x, dset, n, st, xnew,
                     partials);
  TOPOHIP_CHECK(hipGetLastError());
}

}  // namespace topohip
