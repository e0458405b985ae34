// ku.hip — matrix-free hex8 operator apply on the fine level.
//
// Replaces the colour-ordered element loop of core/src/fea/matfree.cpp:75-123,
// 535-553 (gather 24 / 24x24 column AXPY / scatter 24). Design for gfx950:
//
//  * A workgroup of kTX x kTY = 32 x 8 lanes owns a 31 x 7 tile of nodes in
//    x-y and marches a z-chunk of node planes (a multiple of 4, autotuned per
//    grid and precision). Each lane handles one
//    element column (one halo element on the low x and y side); node planes
//    are staged in LDS, double buffered, with the next plane (and the next
//    layer's moduli, masks and epilogue operands) prefetched into registers
//    one layer ahead.
//  * The element matrix is applied in the Hadamard (parity-mode) basis of the
//    8 corners, where E*Ke has 45 non-zeros taking 8 distinct values (the
//    grouped core below: 45 multiply-adds including the scaling by E). The
//    butterfly factorises by axis, and its x-y stages act on one node plane at
//    a time, so each lane transforms the 4-node face of its column once per
//    plane and shares it between the element below and the element above; on
//    the way back the z stage splits each element's result into a bottom and a
//    top face, the two faces that meet on a plane are added first, and the
//    inverse x-y stages run once per plane.
//  * Node-owned deterministic accumulation: every lane writes its face's 4
//    corner results to LDS and the owning lane sums its 4 face contributions
//    in a fixed order. No atomics, bit-identical run to run.
#include <string>
#include <type_traits>

#include "kernels.hpp"

namespace topohip {

namespace {

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// A value every lane holds equally, moved to scalar registers.
__device__ inline double uniform(double v) {
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
  const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(b));
  const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(b >> 32));
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// Forward x-y butterfly of a 4-node face (corner q = bx + 2 by, 3 components).
template <typename T>
__device__ inline void face_fwd(const T (&u)[12], T (&F)[12]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const T a0 = u[c] + u[3 + c], a1 = u[c] - u[3 + c];
    const T a2 = u[6 + c] + u[9 + c], a3 = u[6 + c] - u[9 + c];
    F[c] = a0 + a2;
    F[6 + c] = a0 - a2;
    F[3 + c] = a1 + a3;
    F[9 + c] = a1 - a3;
  }
}

// The K.u kernel is compiled without automatic FP contraction and spells out
// its fused multiply-adds, so every tile shape (whose code the compiler
// schedules differently) produces the same bits.
#pragma clang fp contract(off)

// v = E * C A for the 24 modal amplitudes A, with C in its grouped form
// (8 distinct values): 45 multiply-adds, no per-entry coefficient registers.
template <typename T>
__device__ inline void modal_core(const T (&A)[24], T (&v)[24], T e, const ModalGroups<T>& mg) {
  const T ea = e * mg.amb, eb = e * mg.b, ed = e * mg.d, ee = e * mg.e, ef = e * mg.f;
  const T eg = e * mg.gmh, eh = e * mg.h, ek = e * mg.k;
  v[0] = v[1] = v[2] = T(0);  // rigid translations
  const T sn = (A[3] + A[7]) + A[14];
  const T bn = eb * sn;
  v[3] = fma(ea, A[3], bn);
  v[7] = fma(ea, A[7], bn);
  v[14] = fma(ea, A[14], bn);
  v[4] = v[6] = ed * (A[4] + A[6]);
  v[5] = v[12] = ed * (A[5] + A[12]);
  v[8] = v[13] = ed * (A[8] + A[13]);
  v[9] = fma(ee, A[9], ef * A[20]);
  v[20] = fma(ee, A[20], ef * A[9]);
  v[10] = fma(ee, A[10], ef * A[17]);
  v[17] = fma(ee, A[17], ef * A[10]);
  v[15] = fma(ee, A[15], ef * A[19]);
  v[19] = fma(ee, A[19], ef * A[15]);
  const T su = (A[11] + A[16]) + A[18];
  const T hu = eh * su;
  v[11] = fma(eg, A[11], hu);
  v[16] = fma(eg, A[16], hu);
  v[18] = fma(eg, A[18], hu);
  v[21] = ek * A[21];
  v[22] = ek * A[22];
  v[23] = ek * A[23];
}

// Inputs x must be zero on inactive DOFs (every solver vector is); the mask
// only decides which outputs are kept.
// DOTK: 0 none, 1 <x, Ax> over kept DOFs, 2 <rdot, y> over kept DOFs (FP64 partials).
//
// Launch shape: kTX x kTY lanes (one element column each) over a z-chunk of
// a.tz node planes, a.tz a multiple of the slab ku_slab<T>(). The FP64 dot
// partials are one per (slab, tile): each lane sums its node's terms over the
// slab's planes in plane order and the workgroup adds its four waves in a
// fixed order, so the partials (and every PCG scalar formed from them) do not
// depend on the z-chunk the autotuner picks.
// PU: the fused PCG direction update of the FP64 A.p (KuOp::pw), with z in
// FP32 (1) or FP64 (2); 0 applies x as it is.
template <typename T, int EPI, int DOTK, int PU = 0>
__global__ void __launch_bounds__(kTX * kTY, PU != 0 ? 3 : 1) ku_face(KuOp<T> a) {
  constexpr int TX = kTX, TY = kTY;
  constexpr int NT = TX * TY;
  constexpr int PW = TX + 1;                  // staged plane: PW x (TY + 1) nodes
  constexpr int PLANE = PW * (TY + 1) * 3;
  constexpr int NPF = (PLANE + NT - 1) / NT;  // prefetch registers per lane
  constexpr int SLAB = ku_slab<T>();          // planes per dot partial
  // re-read plane k from LDS (see step): the FP64 residual and smoothing
  // variants and the FP32 sweeps, which gain a wave per SIMD by it (the FP32
  // sweep from zero: 24.0 -> 20.9 us at 192x64x64, 574 -> 593 us at
  // 512x256x256; the FP64 plain and A.p applies measured no gain or slower)
  constexpr bool RECOMP = (sizeof(T) == 8 && EPI != EPI_AX) || EPI == EPI_RES0 ||
                          (sizeof(T) == 4 && EPI == EPI_JAC);
  // LATE: the FP32 Jacobi sweep loads plane k's owner operands (mask, b, D^-1,
  // the FP64 r) in step k after the modal core, in flight across the corner
  // exchange, instead of a step ahead: with RECOMP 120 -> 81 registers, five
  // waves per SIMD (24.1 -> 20.1 us at 192x64x64, 597 -> 576 us at 512x256x256)
  constexpr bool LATE = sizeof(T) == 4 && EPI == EPI_JAC;
  const Geom g = a.g;
  // Logical tile (x fastest, then y, z-chunk, load case) of this workgroup,
  // XCD-contiguous so that tiles sharing halo planes share an L2.
  const int gx = (g.nnx + TX - 2) / (TX - 1), gy = (g.nny + TY - 2) / (TY - 1);
  const int nchunk = (g.nnz + a.tz - 1) / a.tz;
  const long long tile = xcd_tile(blockIdx.x, gridDim.x);
  if (tile >= static_cast<long long>(gx) * gy * nchunk * a.nrhs) return;
  const int bx = static_cast<int>(tile % gx);
  const int by = static_cast<int>((tile / gx) % gy);
  const int bz = static_cast<int>(tile / (static_cast<long long>(gx) * gy));
  const int rc = bz / nchunk;  // load case
  const int chunk = bz - rc * nchunk;
  if (case_done(a.done, rc)) return;
  __shared__ T s_u[2][PLANE];
  __shared__ T s_c[4][NT * 3];
  __shared__ double s_red[NT / 64];

  const int t = threadIdx.x;
  const int tx = t % TX, ty = t / TX;
  // Direction update p = z + beta * x: beta from the <r,z> partials, summed in
  // the same fixed order by every workgroup (lanes stride the partials, waves
  // reduce by shuffles, the four wave sums are added in wave order). Formed
  // after the first plane's loads are issued, so its latency overlaps theirs;
  // false: a breakdown (non-finite <r,z>), recorded by the case's first tile.
  double beta = 0.0;
  bool fresh = false;
  if constexpr (PU != 0) fresh = a.st[rc].pfresh != 0;
  auto form_beta = [&]() -> bool {
    if constexpr (PU != 0) {
      if (fresh) return true;
      CgState* cs = a.st + rc;
      const double* rp = a.rzp + static_cast<long long>(rc) * a.pstride;
      double s = 0.0;
      for (int i = t; i < a.nrz; i += NT) s += rp[i];
      s = wave_sum(s);
      if ((t & 63) == 0) s_red[t >> 6] = s;
      __syncthreads();
      double rzn = 0.0;
#pragma unroll
      for (int w = 0; w < NT / 64; ++w) rzn += s_red[w];
      rzn = uniform(rzn);
      __syncthreads();  // s_red is reused by the dot partials
      const bool lead = bx == 0 && by == 0 && chunk == 0 && t == 0;
      if (!isfinite(rzn)) {
        if (lead) {
          cs->rz_new = rzn;
          cs->beta = 0.0;
          cs->fail = 1;
          cs->done = 1;
        }
        return false;
      }
      beta = uniform(rzn / cs->rz);
      if (lead) {
        cs->rz_new = rzn;
        cs->beta = beta;
      }
    }
    return true;
  };
  const int i0 = bx * (TX - 1), j0 = by * (TY - 1);
  const int k0 = chunk * a.tz;
  const int kend = min(k0 + a.tz, g.nnz);
  const int ei = i0 - 1 + tx, ej = j0 - 1 + ty;
  const int na = i0 + tx, nb = j0 + ty;
  const bool owner = tx < TX - 1 && ty < TY - 1 && na < g.nnx && nb < g.nny;
  const bool exy = ei >= 0 && ei < g.ex && ej >= 0 && ej < g.ey;
  const long long vo = rc * g.ndof();       // case offset of x, y, b
  const long long pn = static_cast<long long>(g.nnx) * g.nny;  // nodes per plane
  const long long pe = static_cast<long long>(g.ex) * g.ey;    // elements per layer
  const T* __restrict__ xs = a.x ? a.x + vo : nullptr;
  T* __restrict__ ys = a.y + vo;
  const T* __restrict__ bs = a.b ? a.b + vo : nullptr;
  const T* __restrict__ ds = a.dinv;
  T* __restrict__ xw = a.xw ? a.xw + vo : nullptr;
  const double* __restrict__ rd = a.rdot ? a.rdot + vo : nullptr;
  const uint8_t* __restrict__ mask = a.mask;
  const T* __restrict__ E = a.E;
  using TZ = typename std::conditional<PU == 1, float, double>::type;
  const TZ* __restrict__ zs = nullptr;
  double* __restrict__ pws = nullptr;
  if constexpr (PU == 1) zs = a.pz32 + vo;
  if constexpr (PU == 2) zs = a.pz64 + vo;
  if constexpr (PU != 0) pws = a.pw + vo;
  // Loop-invariant in-plane offsets (node 0 for a halo position outside the
  // grid, whose value is zeroed through `valid`).
  int foff[NPF];
  unsigned valid = 0;
#pragma unroll
  for (int i = 0; i < NPF; ++i) {
    const int q = t + i * NT;
    foff[i] = 0;
    if (q < PLANE) {
      const int n = q / 3, comp = q - 3 * n;
      const int ga = i0 - 1 + n % PW, gb = j0 - 1 + n / PW;
      if (ga >= 0 && ga < g.nnx && gb >= 0 && gb < g.nny) {
        foff[i] = (gb * g.nnx + ga) * 3 + comp;
        valid |= 1u << i;
      }
    }
  }
  const int eoff = exy ? ej * g.ex + ei : 0;
  const int noff = owner ? nb * g.nnx + na : 0;

  T pf[NPF], pd[NPF];  // plane values in flight (pd: D^-1 of EPI_RES0)
  TZ pz[NPF];          // PU: z of the plane in flight
  bool pin = false;    // EPI_RES0: the plane in flight lies inside the grid
  auto fetch_plane = [&](int c, T (&pf)[NPF]) {
    if constexpr (PU != 0) {
      // the previous direction (not read after a restart) and z
      if (c >= 0 && c < g.nnz) {
        const T* xp = xs + c * pn * 3;
        const TZ* zp = zs + c * pn * 3;
#pragma unroll
        for (int i = 0; i < NPF; ++i) {
          pf[i] = (valid >> i) & 1u && !fresh ? xp[foff[i]] : T(0);
          pz[i] = (valid >> i) & 1u ? zp[foff[i]] : TZ(0);
        }
      } else {
#pragma unroll
        for (int i = 0; i < NPF; ++i) {
          pf[i] = T(0);
          pz[i] = TZ(0);
        }
      }
    } else if constexpr (EPI == EPI_RES0) {
      // one damped-Jacobi sweep from zero, formed on the fly: x0 = omega D^-1 b.
      // The plane index is clamped into the grid and a plane outside it is
      // zeroed where the values are used, so pd is written in one place only.
      pin = c >= 0 && c < g.nnz;
      const int cc = min(max(c, 0), g.nnz - 1);
      const T* bp = bs + cc * pn * 3;
      const T* dp = ds + cc * pn * 3;
#pragma unroll
      for (int i = 0; i < NPF; ++i) {
        pf[i] = bp[foff[i]];
        pd[i] = dp[foff[i]];
      }
    } else if (c >= 0 && c < g.nnz) {
      const T* xp = xs + c * pn * 3;
#pragma unroll
      for (int i = 0; i < NPF; ++i) pf[i] = (valid >> i) & 1u ? xp[foff[i]] : T(0);
    } else {
#pragma unroll
      for (int i = 0; i < NPF; ++i) pf[i] = T(0);
    }
  };
  // The loaded values are first used here, a layer after their loads were
  // issued. EPI_RES0 loads b and D^-1 unconditionally at in-grid addresses
  // and zeroes halo positions and planes outside the grid here (measured
  // faster than predicated loads for that variant; the others load predicated).
  auto store_plane = [&](int buf, const T (&pf)[NPF]) {
#pragma unroll
    for (int i = 0; i < NPF; ++i) {
      const int q = t + i * NT;
      T v = pf[i];
      if constexpr (PU != 0) {
        // p = z + beta * x (zero off the grid, where both loads were skipped)
        const T z = static_cast<T>(pz[i]);
        v = fresh ? z : fma(beta, v, z);
      }
      if constexpr (EPI == EPI_RES0) {
        v = static_cast<T>(a.omega_d * static_cast<double>(pd[i]) * static_cast<double>(v));
        v = ((valid >> i) & 1u) && pin ? v : T(0);
      }
      if (i + 1 < NPF || q < PLANE) s_u[buf][q] = v;
    }
  };
  auto load_face = [&](int buf, T (&F)[12]) {
    T u[12];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int base = ((ty + (q >> 1)) * PW + tx + (q & 1)) * 3;
#pragma unroll
      for (int c = 0; c < 3; ++c) u[3 * q + c] = s_u[buf][base + c];
    }
    face_fwd(u, F);
  };
  auto load_own = [&](int buf, T (&xo)[3]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) xo[c] = s_u[buf][((ty + 1) * PW + (tx + 1)) * 3 + c];
  };
  // Modulus of layer k as loaded; lanes off the grid's element columns and
  // layers outside it get zero where it is used (use_modulus).
  auto modulus = [&](int k) -> T {
    if (k < 0 || k >= g.ez) return T(0);
    return E[k * pe + eoff];
  };
  // Owner-node operands of plane c (fetched one layer ahead).
  struct Own {
    int mk;
    T b[3], d[3];
    double r[3];
  };
  auto fetch_owner = [&](int c, Own& o) {
    if (c < 0 || c >= g.nnz) return;
    const long long node = c * pn + noff;
    o.mk = mask ? mask[node] : 7;
    if constexpr (EPI != EPI_AX) {
#pragma unroll
      for (int q = 0; q < 3; ++q) o.b[q] = bs[node * 3 + q];
    }
    if constexpr (EPI == EPI_JAC) {
#pragma unroll
      for (int q = 0; q < 3; ++q) o.d[q] = ds[node * 3 + q];
    }
    if constexpr (DOTK == 2) {
#pragma unroll
      for (int q = 0; q < 3; ++q) o.r[q] = rd[node * 3 + q];
    }
  };

  T F0[12], F1[12], X0[3] = {T(0), T(0), T(0)}, X1[3], Vt[12];
  Own O0{}, O1{};
  T e0, e1;
#pragma unroll
  for (int r = 0; r < 12; ++r) Vt[r] = T(0);
  double dotp = 0.0;
  int pending = -1;  // slab whose wave sums wait in s_red
  double* const part = DOTK != 0 ? a.partials + static_cast<long long>(rc) * a.pstride : nullptr;
  auto flush_pending = [&]() {
    if (t == 0) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < NT / 64; ++w) s += s_red[w];
      part[(static_cast<long long>(pending) * gy + by) * gx + bx] = s;
    }
    pending = -1;
  };

  fetch_plane(k0 - 1, pf);
  if (!form_beta()) return;
  store_plane(0, pf);
  fetch_plane(k0, pf);
  e0 = modulus(k0 - 1);

  // One element layer k (between node planes k and k + 1): plane k + 1 goes to
  // LDS buffer nxt, its face is transformed into Fn; layer k is applied and
  // plane k's outputs (whose other half came from layer k - 1 through Vt) are
  // finished. Fp / xo / e / oc hold plane k's face, own values, layer k's
  // modulus and plane k's owner operands; the *n arguments receive plane k + 1's.
  // With RECOMP plane k's face and own values are not carried from the
  // previous step but formed again from LDS buffer cur, which still holds
  // plane k: 15 LDS reads and 24 adds more per step for 30 fewer registers
  // carried through the modal core.
  auto step = [&](int k, int cur, T (&Fp)[12], T (&Fn)[12], const T (&xo)[3], T (&xn)[3],
                  const T& e, T& en, const Own& oc, Own& on) {
    const int nxt = cur ^ 1;
    store_plane(nxt, pf);  // plane k + 1
    if (k + 1 < kend) fetch_plane(k + 2, pf);
    en = modulus(k + 1);
    if constexpr (!LATE) fetch_owner(k + 1, on);
    __syncthreads();
    if constexpr (DOTK != 0) {
      if (pending >= 0) flush_pending();
    }
    if (RECOMP || k == k0 - 1) load_face(cur, Fp);
    load_face(nxt, Fn);
    if constexpr (!RECOMP) load_own(nxt, xn);
    T A[24], v[24];
#pragma unroll
    for (int q = 0; q < 12; ++q) {
      A[q] = Fp[q] + Fn[q];
      A[12 + q] = Fp[q] - Fn[q];
    }
    modal_core(A, v, exy ? e : T(0), a.mg);
    T G[12];
#pragma unroll
    for (int q = 0; q < 12; ++q) {
      G[q] = Vt[q] + (v[q] + v[12 + q]);
      Vt[q] = v[q] - v[12 + q];
    }
    T xk[3];  // plane k's own values (read before the barrier below: the next
              // step's plane store may overwrite buffer cur after it)
    if constexpr (RECOMP) {
      load_own(cur, xk);
    } else {
#pragma unroll
      for (int c = 0; c < 3; ++c) xk[c] = xo[c];
    }
    Own ol{};
    if constexpr (LATE) {
      if (k >= k0) fetch_owner(k, ol);
    }
    const Own& ow = LATE ? ol : oc;
    if (k >= k0) {
      // Inverse x-y butterfly of plane k's face, corner results to LDS.
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T b0 = G[c] + G[3 + c], b1 = G[c] - G[3 + c];
        const T b2 = G[6 + c] + G[9 + c], b3 = G[6 + c] - G[9 + c];
        s_c[0][t * 3 + c] = b0 + b2;
        s_c[2][t * 3 + c] = b0 - b2;
        s_c[1][t * 3 + c] = b1 + b3;
        s_c[3][t * 3 + c] = b1 - b3;
      }
    }
    __syncthreads();
    if (k >= k0 && owner) {
      // Node (na, nb) is corner 3 of column (tx, ty), 2 of (tx+1, ty),
      // 1 of (tx, ty+1) and 0 of (tx+1, ty+1). Fixed summation order.
      const int t00 = t, t10 = t + 1, t01 = t + TX, t11 = t + TX + 1;
      T* yp = ys + (k * pn + noff) * 3;
      T* xwp = nullptr;
      if constexpr (EPI == EPI_RES0) xwp = xw + (k * pn + noff) * 3;
      T ax[3];
#pragma unroll
      for (int c = 0; c < 3; ++c)
        ax[c] = ((s_c[3][t00 * 3 + c] + s_c[2][t10 * 3 + c]) + s_c[1][t01 * 3 + c]) +
                s_c[0][t11 * 3 + c];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const bool kept = (ow.mk >> c) & 1;
        T out;
        if constexpr (EPI == EPI_AX) {
          out = ax[c];
        } else if constexpr (EPI == EPI_RES || EPI == EPI_RES0) {
          out = ow.b[c] - ax[c];
        } else {
          out = fma(a.omega * ow.d[c], ow.b[c] - ax[c], xk[c]);
        }
        out = kept ? out : T(0);
        if constexpr (DOTK == 1)
          dotp = fma(static_cast<double>(xk[c]), kept ? static_cast<double>(ax[c]) : 0.0, dotp);
        if constexpr (DOTK == 2) dotp = fma(ow.r[c], static_cast<double>(out), dotp);
        yp[c] = out;
        if constexpr (EPI == EPI_RES0) xwp[c] = xk[c];
        if constexpr (PU != 0) pws[(k * pn + noff) * 3 + c] = xk[c];
      }
    }
    if constexpr (DOTK != 0) {
      // end of a slab (or of the grid): wave sums to LDS, written out by
      // thread 0 after the next barrier
      if (k >= k0 && ((k + 1) % SLAB == 0 || k + 1 == g.nnz)) {
        const double w = wave_sum(dotp);
        if ((t & 63) == 0) s_red[t >> 6] = w;
        dotp = 0.0;
        pending = k / SLAB;
      }
    }
  };

  // (Unrolling this loop by two, so the hand-over below became a renaming,
  // measured slower: the doubled body holds more registers, costing a wave
  // per SIMD in FP64.)
  for (int k = k0 - 1, cur = 0; k < kend; ++k, cur ^= 1) {
    step(k, cur, F0, F1, X0, X1, e0, e1, O0, O1);
    if constexpr (!RECOMP) {
#pragma unroll
      for (int q = 0; q < 12; ++q) F0[q] = F1[q];
#pragma unroll
      for (int q = 0; q < 3; ++q) X0[q] = X1[q];
    }
    e0 = e1;
    if constexpr (!LATE) O0 = O1;
  }
  if constexpr (DOTK != 0) {
    __syncthreads();
    if (pending >= 0) flush_pending();
  }
}

template <typename T>
void dispatch(const KuOp<T>& op, hipStream_t s) {
  constexpr int slab = ku_slab<T>();
  if (op.tz < slab || op.tz % slab != 0)
    throw HipError("K.u z-chunk must be a positive multiple of " + std::to_string(slab));
  const long long tiles = static_cast<long long>((op.g.nnx + kTX - 2) / (kTX - 1)) *
                          ((op.g.nny + kTY - 2) / (kTY - 1)) * ((op.g.nnz + op.tz - 1) / op.tz) *
                          op.nrhs;
  const dim3 grid(xcd_grid(tiles));
  const bool dot = op.partials != nullptr;
  if (op.pw) {
    // fused PCG direction update: FP64 A.p with <p, Ap> partials only
    if (sizeof(T) != 8 || op.epi != EPI_AX || !dot || !op.st || !op.rzp || !op.x ||
        (op.pz32 == nullptr) == (op.pz64 == nullptr) ||
        static_cast<const void*>(op.pw) == static_cast<const void*>(op.x))
      throw HipError("K.u direction update needs the FP64 A.p with partials, the state, "
                     "the <r,z> partials, the previous direction, one z and a separate output");
    if constexpr (sizeof(T) == 8) {
      if (op.pz32)
        hipLaunchKernelGGL((ku_face<T, EPI_AX, 1, 1>), grid, dim3(kTX * kTY), 0, s, op);
      else
        hipLaunchKernelGGL((ku_face<T, EPI_AX, 1, 2>), grid, dim3(kTX * kTY), 0, s, op);
    }
    TOPOHIP_CHECK(hipGetLastError());
    return;
  }
#define TOPOHIP_L(E, D) hipLaunchKernelGGL((ku_face<T, E, D>), grid, dim3(kTX * kTY), 0, s, op)
  if (op.epi == EPI_AX) {
    if (dot) TOPOHIP_L(EPI_AX, 1); else TOPOHIP_L(EPI_AX, 0);
  } else if (op.epi == EPI_RES) {
    if (dot) TOPOHIP_L(EPI_RES, 1); else TOPOHIP_L(EPI_RES, 0);
  } else if (op.epi == EPI_RES0) {
    if (dot || !op.xw || !op.b || !op.dinv)
      throw HipError("K.u EPI_RES0 needs b, dinv and xw and takes no partials");
    TOPOHIP_L(EPI_RES0, 0);
  } else {
    if (dot && !op.rdot) throw HipError("K.u EPI_JAC partials need rdot");
    if (dot) TOPOHIP_L(EPI_JAC, 2); else TOPOHIP_L(EPI_JAC, 0);
  }
#undef TOPOHIP_L
  TOPOHIP_CHECK(hipGetLastError());
}

}  // namespace

template <typename T>
int ku_partials(const Geom& g) {
  constexpr int slab = ku_slab<T>();
  return ((g.nnx + kTX - 2) / (kTX - 1)) * ((g.nny + kTY - 2) / (kTY - 1)) *
         ((g.nnz + slab - 1) / slab);
}
template int ku_partials<float>(const Geom&);
template int ku_partials<double>(const Geom&);

template <typename T>
void launch_ku_fine(const KuOp<T>& op, hipStream_t s) {
  dispatch<T>(op, s);
}

template void launch_ku_fine<float>(const KuOp<float>&, hipStream_t);
template void launch_ku_fine<double>(const KuOp<double>&, hipStream_t);

}  // namespace topohip
