// kernels.hpp — host-side launchers of the engine's HIP kernels.
#pragma once

#include "common.hpp"

namespace topohip {

// EPI_RES0 (fine operator only): x = omega*dinv .* b is formed on the fly from
// b and dinv (one damped-Jacobi sweep from zero), written to xw, and y = b - A x.
enum Epilogue { EPI_AX = 0, EPI_RES = 1, EPI_JAC = 2, EPI_RES0 = 3 };

// One tiled operator apply over the fine structured level: the modal element
// operator scaled by per-element moduli E. Epilogues: y = A x, y = b - A x,
// y = x + omega*dinv*(b - A x), and EPI_RES0 above.
// With `partials` set, each block also writes an FP64 sum over its owned kept
// DOFs to partials[block]: sum(x . A x) for EPI_AX / EPI_RES, sum(rdot . y)
// for EPI_JAC (rdot required).
template <typename T>
struct KuOp {
  Geom g;
  const T* x = nullptr;
  T* y = nullptr;
  const T* b = nullptr;
  const T* dinv = nullptr;
  T omega = T(0);
  double omega_d = 0.0;           // EPI_RES0: omega in FP64 (x = T(omega_d * dinv * b))
  T* xw = nullptr;                // EPI_RES0: x written here
  const double* rdot = nullptr;   // EPI_JAC with partials: FP64 vector dotted with y
  const uint8_t* mask = nullptr;  // per node: bit c = DOF c kept; null = all kept
  const T* E = nullptr;           // fine level
  ModalGroups<T> mg;              // grouped modal coefficients of Ke(E=1)
  int epi = EPI_AX;
  double* partials = nullptr;
  int tz = 8;                     // node planes per workgroup (z-march length)
  const int* done = nullptr;      // case r exits early when done[r * kDoneStride] != 0
  int nrhs = 1;                   // stacked load cases (x, y, b stride g.ndof())
  int pstride = 0;                // partials of case r at partials + r * pstride
  // Fused PCG direction update (FP64 EPI_AX with partials only): the applied
  // vector is p = z + beta * x, formed where each node value is staged, with
  // x the previous direction, z the preconditioned residual (pz32 or pz64) and
  // beta = rz_new / st->rz, rz_new the fixed-order sum of the case's nrz <r,z>
  // partials at rzp (every workgroup forms the same value; the case's first
  // tile records rz_new and beta, and a non-finite rz_new is a breakdown).
  // After a restart (st->pfresh) p = z. The owners write p to pw, which must
  // not alias x (neighbouring tiles still read x as their halo).
  const float* pz32 = nullptr;
  const double* pz64 = nullptr;
  double* pw = nullptr;
  CgState* st = nullptr;
  const double* rzp = nullptr;
  int nrz = 0;
};

// Fine K.u launch shape: kTX x kTY lanes per workgroup (one element column
// each, owning (kTX-1) x (kTY-1) nodes) over a z-chunk of tz node planes, tz a
// multiple of the precision's slab ku_slab<T>(); the FP64 dot partials are one
// per (slab, x-y tile), ku_partials<T>(g) per load case, whatever tz is.
// The FP64 applies hold three waves per SIMD against the FP32 applies' five,
// and chunks in multiples of 7 planes fill their resident workgroups better
// (192x64x64: a 7-plane chunk puts the grid's 700 workgroups in one round,
// where 4 planes need 1190 in two and 8 planes 630 of 9 steps each).
constexpr int kTX = 32, kTY = 8;
template <typename T>
constexpr int ku_slab() { return sizeof(T) == 8 ? 7 : 4; }
template <typename T> int ku_partials(const Geom& g);
template <typename T> void launch_ku_fine(const KuOp<T>& op, hipStream_t s);

// ---- transfers between a fine level f and the coincident coarse level c ----
// Coarse solver's plane-ordered FP64 vector space: when dst is set, the
// restriction writes (and the prolongation reads) V[case * vstride + off +
// dst[3 * node + comp]] instead of a node-ordered coarse vector.
struct CrMap {
  const int* dst = nullptr;
  double* V = nullptr;
  long long vstride = 0;
  long long off = 0;
};
// bc = mask_c .* P^T rf; optionally xc = omega * dinv_c .* bc (pre-smooth from 0).
// With nr load cases the vectors are stacked with the level's ndof as stride.
template <typename Tf, typename Tc>
void launch_restrict(const Geom& gf, const Geom& gc, const Tf* rf, const uint8_t* mask_c,
                     Tc* bc, Tc* xc, const Tc* dinv_c, double omega, const int* done,
                     hipStream_t s, int nr = 1, CrMap cm = CrMap());
// xf += mask_f .* P xc
template <typename Tc, typename Tf>
void launch_prolong_add(const Geom& gc, const Geom& gf, const Tc* xc, const uint8_t* mask_f,
                        Tf* xf, const int* done, hipStream_t s, int nr = 1, CrMap cm = CrMap());

// ---- reductions / vector kernels -------------------------------------------
// out[r] = sum(partials[r*pstride + 0..n)) in a fixed order (one block per case).
void launch_sum_partials(const double* partials, int n, double* out, hipStream_t s, int nr = 1,
                         int pstride = 0);
// partial dot products over n entries: partials[r*pstride + block] = sum a[i]*b[i]
// for each of nr stacked case vectors (stride n).
template <typename TA, typename TB>
void launch_dot(const TA* a, const TB* b, long long n, double* partials, int nblocks,
                const int* done, hipStream_t s, int nr = 1, int pstride = 0);
int dot_blocks(long long n);
// x = omega * dinv .* b (one damped-Jacobi sweep from a zero start).
template <typename T>
void launch_jacobi0(T* x, const T* dinv, const T* b, double omega, long long n, const int* done,
                    hipStream_t s, int nr = 1);
// Chebyshev smoother (coef: device [lambda, 1/theta, (c1_k, c2_k) for k >= 1]):
// first step d = D^-1 r / theta, xo = d (zero_start) or xo = x + d; step k:
// d = c1_k d + c2_k D^-1 r, xo = x + d (xo may be x). Stacked cases of stride n.
template <typename T>
void launch_cheb_first(const T* x, T* xo, T* d, const T* r, const T* dinv, const double* coef,
                       long long n, bool zero_start, const int* done, hipStream_t s, int nr = 1);
template <typename T>
void launch_cheb_step(const T* x, T* xo, T* d, const T* r, const T* dinv, const double* coef,
                      int k, long long n, const int* done, hipStream_t s, int nr = 1);
// Power iteration for lambda_max(D^-1 A) over one vector of length n:
// step 0 writes the start vector into v; step 1 turns y = A v into
// v = D^-1 y / ||D^-1 y||; step 2 forms the Rayleigh quotient (v.y)/(v.Dv) and
// the Chebyshev coefficients of [lo, hi] * lambda into coef (partials: at
// least 2 * cheb_blocks(n)).
template <typename T>
void launch_cheb_power(int step, T* v, const T* y, const T* dinv, long long n, double* partials,
                       double* coef, int degree, double lo, double hi, hipStream_t s);
int cheb_blocks(long long n);
// y = x + up (scatter prescribed values into the solution).
void launch_add(const double* x, const double* up, double* y, long long n, hipStream_t s);
template <typename T> void launch_fill(T* p, T v, long long n, hipStream_t s);
template <typename TI, typename TO> void launch_convert(const TI* in, TO* out, long long n,
                                                        hipStream_t s);

// PCG pieces (device-resident scalars in CgState). Every launcher covers nr
// load cases: states st[0..nr), partials of case r at partials + r*pstride,
// vectors stacked with stride n.
// alpha = rz_new / pAp (pAp = fixed-order sum of the npap A.p partials at
// pap, breakdown sets fail/done); x += alpha*p; r -= alpha*Ap; rr partials
// to `partials` (a different buffer); r32 = float(r) (nullable).
// With jz set (the Jacobi-preconditioned CG): also z = jdinv .* r into jz and the
// <r, z> partials into rzp (case stride pstride), over nblocks = dot_blocks(n),
// the partition launch_dot uses, so the sums are dot_k's.
// With p_prev set the solution updates are paired: iteration j = iters - it0
// since the last restart defers its x += alpha p when j is even (recording
// alpha in alpha_prev) and applies x += alpha_prev p_prev, then x += alpha p,
// when j is odd (p_prev: the previous iteration's direction). A recurrence
// that stops after an even iteration owes its last update: launch_cg_xfinish.
template <typename TP, typename TA>
void launch_cg_xr(CgState* st, double* x, double* r, const TP* p, const TA* Ap, float* r32,
                  long long n, const double* pap, int npap, double* partials, int nblocks,
                  hipStream_t s, int nr = 1, int pstride = 0, const double* jdinv = nullptr,
                  double* jz = nullptr, double* rzp = nullptr, const TP* p_prev = nullptr);
// x += alpha_prev * p for the cases r with bit r of pend set, p from p1 where bit
// r of par is set (else p0); then it0 = iters for them (nr <= 32 cases).
void launch_cg_xfinish(CgState* st, double* x, const double* p0, const double* p1, long long n,
                       unsigned pend, unsigned par, hipStream_t s, int nr = 1);
// rr = sum(partials); iters++; done if rr <= thresh2; rz = rz_new
void launch_cg_check(CgState* st, const double* partials, int n, hipStream_t s, int nr = 1,
                     int pstride = 0);
// start: bnorm2, thresh2, rz from partials etc.
void launch_cg_init(CgState* st, const double* bb_partials, int nb, double tol, int cap,
                    hipStream_t s, int nr = 1, int pstride = 0);
// rr = sum(partials) of a recomputed residual; done = rr <= thresh2 (resume).
void launch_cg_resume(CgState* st, const double* partials, int n, hipStream_t s, int nr = 1,
                      int pstride = 0);
// p_out = z + beta*p_in with rz_new = fixed-order sum of the nrz <r,z> partials
// at rzp and beta = rz_new / rz (restart: p_out = z, partials unused); the
// direction update of grids whose A.p does not carry it (p_in != p_out).
template <typename TZ>
void launch_cg_p(CgState* st, const TZ* z, const double* p_in, double* p_out, long long n,
                 int restart, const double* rzp, int nrz, hipStream_t s, int nr = 1,
                 int pstride = 0);
// restart: rz = rz_new = sum(partials); the next A.p takes p = z (pfresh)
void launch_cg_set_rz(CgState* st, const double* partials, int n, hipStream_t s, int nr = 1,
                      int pstride = 0);

// ---- hierarchy build ---------------------------------------------------------
// Fine Jacobi inverse diagonal: dinv = 1 / sum_e E_e Ke(l,l) on kept DOFs, else 0.
template <typename T>
void launch_diag_fine(const Geom& g, const double* E, const double* kediag, const uint8_t* mask,
                      T* dinv, hipStream_t s);
// Level-1 element matrices: K1_E = Mc (sum_c E_c W_c^T (Mf Ke Mf) W_c) Mc.
// Level-1 Galerkin blocks; `special` lists the n_special coarse elements with
// an inactive DOF among their 27 fine nodes (they take the masked path).
void launch_galerkin_fine(const Geom& gc, const Geom& gf, const double* E, const uint8_t* mask_f,
                          const uint8_t* mask_c, const double* S8, const double* Ke,
                          const int* special, int n_special, double* Kout, hipStream_t s);
// The masked path only, for the n_special listed level-1 elements, written
// compactly: K1sp[i] is the matrix of special[i].
void launch_galerkin_fine_compact(const Geom& gc, const Geom& gf, const double* E,
                                  const uint8_t* mask_f, const uint8_t* mask_c, const double* S8,
                                  const double* Ke, const int* special, int n_special,
                                  double* K1sp, hipStream_t s);
// Level-1 stencil (symmetric half storage, see launch_assemble_stencil) and
// Jacobi inverse diagonal straight from the fine moduli, with no stored
// level-1 element matrices: non-special elements from the S_c blocks, the
// others from K1sp via spmap (level-1 element -> slot or -1).
template <typename T>
void launch_stencil_l1(const Geom& g1, const Geom& g0, const double* E, const double* S8,
                       const double* K1sp, const int* spmap, const uint8_t* mask1, T* S, T* dinv,
                       hipStream_t s);
// Level-2 element matrices straight from the fine moduli (T64: the 64 blocks
// W_{c2}^T S_{c1} W_{c2}); the n_special listed elements (an inactive DOF among
// their 125 fine nodes) by the Galerkin product over level-1 children formed
// as in launch_stencil_l1.
void launch_galerkin_l2(const Geom& g2, const Geom& g1, const Geom& g0, const double* E,
                        const double* T64, const double* S8, const double* K1sp, const int* spmap,
                        const uint8_t* mask2, const int* special, int n_special, double* Kout,
                        hipStream_t s);
// Deeper element matrices: K_E = Mc (sum_c W_c^T K_c W_c) Mc.
void launch_galerkin_coarse(const Geom& gc, const Geom& gf, const double* Kf,
                            const uint8_t* mask_c, double* Kout, hipStream_t s);
// dinv of an element-form level.
template <typename T>
void launch_diag_dense(const Geom& g, const double* K, const uint8_t* mask, T* dinv,
                       hipStream_t s);
// ---- SIMP kernels --------------------------------------------------------------
// E = clamp(xphys, rho_min, 1)^p * E0 on solid voxels (0 elsewhere).
void launch_youngs(const double* xphys, const uint8_t* solid, long long n, double E0,
                   double penalty, double rho_min, double* E64, float* E32, hipStream_t s);
// q_e = u_e^T K1 u_e; c partials = sum E_e q_e; dc_e = -p rho^(p-1) E0 q_e (sum over rhs).
void launch_sensitivity(const Geom& g, const double* u, const double* xphys, const uint8_t* solid,
                        const Modal<double>& mc, double E0, double penalty, double rho_min,
                        double* dc, int accumulate, double* partials, int nblocks, hipStream_t s);
int sens_blocks(const Geom& g);

struct Stencil {
  int n = 0;
  int* off = nullptr;    // device: 3 ints per point (di, dj, dk), lexicographic (dk, dj, di)
  double* w = nullptr;   // device: weight per point
};
// out = H x / Hs (adjoint = 0) or H^T (x / Hs) (adjoint = 1); hs precomputed.
void launch_filter(const Geom& g, const Stencil& st, const uint8_t* solid, const double* hs,
                   const double* x, double* out, int adjoint, hipStream_t s);
void launch_filter_hs(const Geom& g, const Stencil& st, const uint8_t* solid, double* hs,
                      hipStream_t s);

// OC state: the lambda search runs on the device.
// Up to kLamPass lambdas are tried in one pass over the voxels: the next
// bracket doublings, or the three bisection levels ahead (7 midpoints in heap
// order). The decisions then replay the sequential search exactly.
constexpr int kLamPass = 7;
struct OcState {
  double l1, l2, lmid, target, vol;
  int phase;      // 0 = bracketing, 1 = bisecting, 2 = done
  int steps;      // bisection steps taken
  int nlam;
  int passes;     // multi-lambda passes that did work (oc_decide_multi)
  double lam[kLamPass];
  double gap;     // smallest |vol - target| over the search's volume decisions
};
void launch_oc_init(OcState* st, double target, hipStream_t s);

// Allocation-free OC update: its work memory is one file-scope device array per
// GPU (`scratch`, simp.hip), laid out as below. Voxel counts above
// kNoAllocMaxVoxels are rejected. oc_blocks() never exceeds kOcMaxBlocks.
constexpr long long kNoAllocMaxVoxels = 1LL << 21;
constexpr int kOcMaxBlocks = 1024;
constexpr size_t kScrState = 0;                                            // OcState
constexpr size_t kScrLam = 256;                                            // kLamPass x blocks
constexpr size_t kScrChange = kScrLam + sizeof(double) * kLamPass * kOcMaxBlocks;  // blocks
constexpr size_t kScrScal = kScrChange + sizeof(double) * kOcMaxBlocks;    // 32 doubles
constexpr size_t kScrDcf = kScrScal + 256;                                 // voxels
constexpr size_t kNoAllocScratchBytes = kScrDcf + sizeof(double) * kNoAllocMaxVoxels;
static_assert(sizeof(OcState) <= kScrLam, "OcState must fit its scratch slot");
static_assert(kScrDcf % 256 == 0, "scratch slots stay 256-byte aligned");
// Device address of that GPU's scratch array (hipGetSymbolAddress once per
// device, with the device made current; cached for the process).
void* noalloc_scratch(int device);
// One lambda trial: volume partials of candidate(lambda_trial(st)).
void launch_oc_trial(const double* x, const double* dcf, const double* dv, const uint8_t* solid,
                     long long n, double move, double rho_min, const OcState* st,
                     double* partials, int nblocks, hipStream_t s);
void launch_oc_decide(OcState* st, const double* partials, int n, hipStream_t s);
// The multi-lambda search of the plain update: init plans the first pass; a
// trial writes the volume partials of every planned lambda (kLamPass x nblocks);
// decide replays the sequential decisions over them and plans the next pass.
void launch_oc_init_multi(OcState* st, double target, hipStream_t s);
void launch_oc_trial_multi(const double* x, const double* dcf, const double* dv,
                           const uint8_t* solid, long long n, double move, double rho_min,
                           const OcState* st, double* partials, int nblocks, hipStream_t s);
void launch_oc_decide_multi(OcState* st, const double* partials, int n, hipStream_t s);
int oc_blocks(long long n);
// Final candidate at lambda = (l1+l2)/2 + active set + change partials (max).
// Voxels off the design set `solid` get 0 (keep = 0, plain oc_update) or keep
// their x (keep = 1, oc_update_masked).
void launch_oc_final(const double* x, const double* dcf, const double* dv, const uint8_t* solid,
                     long long n, double move, double rho_min, const OcState* st, double* xnew,
                     uint8_t* active, double* change_partials, int nblocks, hipStream_t s,
                     int keep = 0);
// Heaviside projection (threshold eta, sharpness beta) of design voxels.
void launch_project(const double* xt, double* xp, const uint8_t* solid, long long n, double beta,
                    double eta, hipStream_t s);
// dpr = projection derivative at xt, dcp = dc * dpr (design voxels; 0 elsewhere).
void launch_proj_chain(const double* xt, const double* dc, const uint8_t* solid, long long n,
                       double beta, double eta, double* dcp, double* dpr, hipStream_t s);
// Projected OC: the candidate of the current lambda trial, the projected
// volume partials of its filtered field, and the final candidate + change.
void launch_oc_cand_proj(const double* x, const double* dcf, const double* dvf,
                         const uint8_t* solid, long long n, double move, double rho_min,
                         const OcState* st, double* xc, hipStream_t s);
void launch_proj_sum(const double* xtc, const uint8_t* solid, long long n, double beta, double eta,
                     const OcState* st, double* partials, int nblocks, hipStream_t s);
// keep: voxels off the design set keep x (the masked loop), else they are 0.
void launch_oc_final_proj(const double* x, const double* dcf, const double* dvf,
                          const uint8_t* solid, long long n, double move, double rho_min,
                          const OcState* st, double* xnew, double* partials, int nblocks,
                          hipStream_t s, bool keep = false);
// MMA: per-voxel state and coefficients (each n long) and the dual search.
struct MmaArgs {
  double *low, *upp, *p0, *q0, *p1, *q1, *al, *be;
};
struct MmaSearch {
  double l1, l2, b, lambda;
  int phase;  // 0 test lambda = 0, 1 bracket doubling, 2 bisection, 3 done
  int it, steps, nlam;
  double lam[kLamPass];  // lambdas of the current pass (as OcState)
};
void launch_mma_scale(const double* dcf, const uint8_t* dset, long long n, double* partials,
                      int nblocks, hipStream_t s);
// Asymptotes (iter <= 2: initial half-span), coefficients and move box; the
// partials of sum p1/(U-x) + q1/(x-L). scale: device max |dcf|.
void launch_mma_coef(const MmaArgs& a, const double* x, const double* x1, const double* x2,
                     const double* dcf, const double* dv, const uint8_t* dset, long long n,
                     int iter, double move, double xmin, const double* scale, double* partials,
                     int nblocks, hipStream_t s);
void launch_mma_init(MmaSearch* st, const double* partials, int n, const double* vol,
                     double target, hipStream_t s);
// partials: kLamPass x nblocks (one set per planned lambda)
void launch_mma_trial(const MmaArgs& a, const uint8_t* dset, long long n, const MmaSearch* st,
                      double* partials, int nblocks, hipStream_t s);
void launch_mma_decide(MmaSearch* st, const double* partials, int n, hipStream_t s);
void launch_mma_final(const MmaArgs& a, const double* x, const uint8_t* dset, long long n,
                      const MmaSearch* st, double* xnew, double* partials, int nblocks,
                      hipStream_t s);
// oscillation_fraction's per-block counts over the design set dset (null = all):
// voxels that moved on both steps x2 -> x1 -> x, and those that reversed.
void launch_osc_count(const double* x, const double* x1, const double* x2, const uint8_t* dset,
                      long long n, double* moved, double* reversed, int nblocks, hipStream_t s);
// Design-mask pins on a physical density: eff 1 (FrozenSolid) -> 1, 2 (FrozenVoid) -> 0.
void launch_mask_pins(double* xphys, const uint8_t* eff, long long n, hipStream_t s);
void launch_max_partials(const double* partials, int n, double* out, hipStream_t s);
template <typename T>
void launch_sum_vec(const T* v, long long n, double* partials, int nblocks, hipStream_t s);

}  // namespace topohip

namespace topohip {
// ---- assembled coarse levels: 27-point node stencil of 3x3 blocks ----------
// Symmetric half storage: block (o, p, q) of node n for neighbour
// o = (dz+1)*9 + (dy+1)*3 + (dx+1) >= 13 (centre and lexicographically larger
// neighbours; row component p, column component q) at
// S[((o-13)*9 + p*3 + q) * nodes + n], 126 coefficients per node. Blocks
// towards smaller neighbours are the transposes stored at those neighbours.
template <typename T>
void launch_assemble_stencil(const Geom& g, const double* Kem, T* S, hipStream_t s);
// y = A x / b - A x / x + omega*dinv*(b - A x) on a stencil level (masked).
template <typename T>
void launch_stencil_apply(const Geom& g, const T* S, const uint8_t* mask, const T* x, T* y,
                          const T* b, const T* dinv, double omega, int epi, const int* done,
                          hipStream_t s, int nr = 1);
}  // namespace topohip

namespace topohip {
// ---- result fields (fields.hip) -----------------------------------------------
// Centroid stress constants of the isotropic hex8: q = (2/h)/8 is |dN/dx| at
// the centroid; D has d0 on the normal diagonal, d1 off it, gs on the shear
// diagonal (core/src/fea/hex_element.cpp:242-316).
struct VmCoef {
  double q, d0, d1, gs;
};
// Von Mises at every solid voxel's centroid from the full displacement u
// (3 * nodes, reference DOF order); 0 on empty voxels. Either output nullable.
void launch_von_mises(const Geom& g, const VmCoef& c, const uint8_t* solid, const double* u,
                      double* vm64, float* vm32, hipStream_t s);
// out[i] = float(in[i]) (round to nearest even).
void launch_narrow_f32(const double* in, float* out, long long n, hipStream_t s);
// out[i] = IEEE half of float(in[i]), round to nearest even.
void launch_density_f16(const double* in, uint16_t* out, long long n, hipStream_t s);
}  // namespace topohip
