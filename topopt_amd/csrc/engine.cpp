// engine.cpp — problem handle, multigrid hierarchy and the device-resident
// PCG / OC drivers of the voxel-SIMP engine.
//
// The solve restates solve_mgcg_matfree (core/src/fea/multigrid.cpp:2098-2319)
// with mf_mgpcg (:1211-1318) and the V-cycle (:1143-1160, :798-819); the
// reduced system and void gate restate mf_build_reduced
// (core/src/fea/matfree.cpp:670-834). Vectors stay full-grid (3 per node) with
// inactive DOFs held at zero, which is the reduced system embedded in the grid.
#include "engine.hpp"
#include "formats.hpp"


#include <algorithm>
#include <array>
#include <chrono>
#include <functional>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <map>
#include <mutex>
#include <random>

namespace topohip {

namespace {

constexpr double kOmega = 0.6;       // damped Jacobi weight (multigrid.cpp:72)
constexpr int kMinCoarseElems = 2;   // coarsen.hpp:76
constexpr int kCoarseDofCap = 6000;  // coarsen.hpp:80
constexpr int kMgIterBudget = 300;   // multigrid.cpp:108
// Chebyshev smoother option: lambda_max(D^-1 A) from kChebPowerSteps power
// steps per smoothed level at every hierarchy build, smoothing the interval
// [kChebLow, kChebHigh] * lambda_max (measured on the oracle: 0.3 keeps the MBB
// and cantilever counts lowest; 0.1 and 1/30 smooth too little of the top).
constexpr int kChebPowerSteps = 10;
constexpr double kChebLow = 0.3, kChebHigh = 1.1;

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

void sync(Problem& P) { TOPOHIP_CHECK(hipStreamSynchronize(P.s)); }

// The solve deadline of the calling thread (absolute steady-clock ms; 0 =
// disarmed), as fea_set_solve_deadline_ms (core/include/topopt/fea.hpp:
// 440-455, core/src/fea/matfree.cpp:1267-1273): thread-local and sticky.
thread_local double t_solve_deadline_ms = 0.0;

bool deadline_passed() {
  return t_solve_deadline_ms > 0.0 && now_ms() >= t_solve_deadline_ms;
}

// The per-run multigrid stagnation latch of the calling thread
// (core/src/fea/multigrid.cpp:123-150, 2138-2155, 2261-2275; reset and query
// core/include/topopt/fea.hpp:1081-1091): a solve that builds a hierarchy and
// does not converge with it (FP32 V-cycle, then the FP64 retry) is a
// stagnation; kMgLatchThreshold consecutive ones latch multigrid off, so
// later solves of the thread skip the hierarchy and go straight to Jacobi-CG,
// and a solve that converges clears the count. Sticky until the caller
// resets it at the start of a run.
constexpr int kMgLatchThreshold = 3;
thread_local int t_mg_stagnations = 0;
thread_local bool t_mg_latched = false;
// The measurement-only re-arm period (multigrid.cpp:152-180, 2144-2156,
// 2248-2262): 0 (the default) keeps the latch for the run; n > 0 clears it
// after n latched solves so one solve attempts multigrid again, with the
// stagnation count one short of the threshold so a stagnating retry re-latches
// at once. Attempts and carries (re-armed solves that converged) are counted.
thread_local int t_mg_rearm_period = 0;
thread_local int t_mg_latched_solves = 0;
thread_local long long t_mg_rearm_attempts = 0, t_mg_rearm_carries = 0;
thread_local bool t_mg_rearmed = false;  // the current solve re-armed the latch

template <typename T>
void h2d(T* d, const T* h, size_t n, hipStream_t s) {
  TOPOHIP_CHECK(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, s));
}
template <typename T>
void d2h(T* h, const T* d, size_t n, hipStream_t s) {
  TOPOHIP_CHECK(hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, s));
  TOPOHIP_CHECK(hipStreamSynchronize(s));
}

// 1-D trilinear weight of coarse node bit b for fine position p in {0,1,2}.
double w1h(int p, int b) {
  if (p == 1) return 0.5;
  return (p == 0) == (b == 0) ? 1.0 : 0.0;
}

int popcount3(uint8_t v) { return (v & 1) + ((v >> 1) & 1) + ((v >> 2) & 1); }

}  // namespace

// Isotropic hex8 stiffness by 2x2x2 Gauss quadrature of B^T D B on a cube of
// edge h (the element of core/src/fea/hex_element.cpp:23-114), with D written
// as lambda * m m^T + mu * diag(2,2,2,1,1,1) in Voigt order
// [xx, yy, zz, gxy, gyz, gzx]. Reference DOF order (bottom CCW, top CCW).
void hex8_stiffness_host(double E, double nu, double h, double* K) {
  if (!(E > 0.0)) throw Error(TOPOHIP_EINVAL, "hex8_stiffness: youngs_modulus must be > 0");
  if (!(h > 0.0)) throw Error(TOPOHIP_EINVAL, "hex8_stiffness: element_size must be > 0");
  if (!(nu > -1.0 && nu < 0.5))
    throw Error(TOPOHIP_EINVAL, "hex8_stiffness: poisson must be in (-1, 0.5)");
  const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double mu = E / (2.0 * (1.0 + nu));
  double D[6][6] = {};
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) D[i][j] = lam;
    D[i][i] += 2.0 * mu;
    D[i + 3][i + 3] = mu;
  }
  static const int cx[8] = {0, 1, 1, 0, 0, 1, 1, 0};
  static const int cy[8] = {0, 0, 1, 1, 0, 0, 1, 1};
  static const int cz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
  std::memset(K, 0, 576 * sizeof(double));
  const double gq = 1.0 / std::sqrt(3.0);
  const double jac = 0.125 * h * h * h;  // det of the (h/2) I map
  for (int q = 0; q < 8; ++q) {
    const double xi = (q & 1) ? gq : -gq, et = (q & 2) ? gq : -gq, ze = (q & 4) ? gq : -gq;
    double G[8][3];
    for (int a = 0; a < 8; ++a) {
      const double sx = cx[a] ? 1.0 : -1.0, sy = cy[a] ? 1.0 : -1.0, sz = cz[a] ? 1.0 : -1.0;
      G[a][0] = (2.0 / h) * 0.125 * sx * (1.0 + sy * et) * (1.0 + sz * ze);
      G[a][1] = (2.0 / h) * 0.125 * sy * (1.0 + sx * xi) * (1.0 + sz * ze);
      G[a][2] = (2.0 / h) * 0.125 * sz * (1.0 + sx * xi) * (1.0 + sy * et);
    }
    double B[6][24] = {};
    for (int a = 0; a < 8; ++a) {
      B[0][3 * a] = G[a][0];
      B[1][3 * a + 1] = G[a][1];
      B[2][3 * a + 2] = G[a][2];
      B[3][3 * a] = G[a][1]; B[3][3 * a + 1] = G[a][0];
      B[4][3 * a + 1] = G[a][2]; B[4][3 * a + 2] = G[a][1];
      B[5][3 * a] = G[a][2]; B[5][3 * a + 2] = G[a][0];
    }
    double DB[6][24];
    for (int i = 0; i < 6; ++i)
      for (int c = 0; c < 24; ++c) {
        double s = 0.0;
        for (int k = 0; k < 6; ++k) s += D[i][k] * B[k][c];
        DB[i][c] = s;
      }
    for (int r = 0; r < 24; ++r)
      for (int c = 0; c < 24; ++c) {
        double s = 0.0;
        for (int k = 0; k < 6; ++k) s += B[k][r] * DB[k][c];
        K[r * 24 + c] += jac * s;
      }
  }
}

Problem::~Problem() {
  std::lock_guard<std::recursive_mutex> lk(device_api_mutex());
  for (auto& g : cg_graph)
    if (g) (void)hipGraphExecDestroy(g);
  if (st_h) (void)hipHostFree(st_h);
  if (oc_h) (void)hipHostFree(oc_h);
  if (mma_h) (void)hipHostFree(mma_h);
  if (scal_h) (void)hipHostFree(scal_h);
  filters.clear();
  L.clear();
  if (s) (void)hipStreamDestroy(s);
}

namespace {

// Element constants: bit-ordered Ke, its modal coefficients (checked against
// the 45-entry pattern), the diagonal and the 8 child Galerkin blocks
// S_c = W_c^T Ke W_c.
void element_constants(Problem& P) {
  hex8_stiffness_host(1.0, P.nu, P.h, P.Ke);
  for (int mr = 0; mr < 8; ++mr)
    for (int cr = 0; cr < 3; ++cr)
      for (int ms = 0; ms < 8; ++ms)
        for (int cs = 0; cs < 3; ++cs)
          P.Kb[(3 * mr + cr) * 24 + 3 * ms + cs] =
              P.Ke[(3 * kRefOfBit[mr] + cr) * 24 + 3 * kRefOfBit[ms] + cs];
  for (int i = 0; i < 24; ++i) P.kdiag[i] = P.Kb[i * 24 + i];
  // M = H Kb H (H = parity Hadamard per component, symmetric).
  auto had = [](int q, int m) { return (__builtin_popcount(q & m) & 1) ? -1.0 : 1.0; };
  double M[576];
  double mx = 0.0;
  for (int r = 0; r < 24; ++r)
    for (int c = 0; c < 24; ++c) {
      double s = 0.0;
      for (int a = 0; a < 8; ++a)
        for (int b = 0; b < 8; ++b)
          s += had(r / 3, a) * P.Kb[(3 * a + r % 3) * 24 + 3 * b + c % 3] * had(c / 3, b);
      M[r * 24 + c] = s;
      mx = std::max(mx, std::fabs(s));
    }
  bool in_pat[576] = {};
  int k = 0;
#define TOPOHIP_X(kk, rr, cc)               \
  in_pat[(rr) * 24 + (cc)] = true;          \
  P.mc64.c[kk] = M[(rr) * 24 + (cc)] / 64.0; \
  P.mc32.c[kk] = static_cast<float>(P.mc64.c[kk]); \
  ++k;
  TOPOHIP_MODAL_PATTERN(TOPOHIP_X)
#undef TOPOHIP_X
  (void)k;
  for (int i = 0; i < 576; ++i)
    if (!in_pat[i] && std::fabs(M[i]) > 1e-12 * mx)
      throw Error(TOPOHIP_EINTERNAL, "element stiffness outside the modal pattern");
  // Grouped form used by the K.u kernel; every entry of M is checked against it.
  {
    auto m = [&](int r, int c) { return M[r * 24 + c] / 64.0; };
    const double a = m(3, 3), b = m(3, 7), d = m(4, 4), e = m(9, 9), f = m(9, 20);
    const double g = m(11, 11), h = m(11, 16), kk = m(21, 21);
    double R[576] = {};
    auto set_block = [&](std::initializer_list<int> idx, double diag, double off) {
      for (int r : idx)
        for (int c : idx) R[r * 24 + c] = r == c ? diag : off;
    };
    set_block({3, 7, 14}, a, b);
    set_block({4, 6}, d, d);
    set_block({5, 12}, d, d);
    set_block({8, 13}, d, d);
    set_block({9, 20}, e, f);
    set_block({10, 17}, e, f);
    set_block({15, 19}, e, f);
    set_block({11, 16, 18}, g, h);
    set_block({21}, kk, 0.0);
    set_block({22}, kk, 0.0);
    set_block({23}, kk, 0.0);
    for (int i = 0; i < 576; ++i)
      if (std::fabs(M[i] / 64.0 - R[i]) > 1e-12 * mx / 64.0)
        throw Error(TOPOHIP_EINTERNAL, "element stiffness does not have the grouped modal form");
    P.mg64 = ModalGroups<double>{a - b, b, d, e, f, g - h, h, kk};
    P.mg32 = ModalGroups<float>{static_cast<float>(a - b), static_cast<float>(b),
                                static_cast<float>(d), static_cast<float>(e),
                                static_cast<float>(f), static_cast<float>(g - h),
                                static_cast<float>(h), static_cast<float>(kk)};
  }
  for (int c = 0; c < 8; ++c) {
    double W[576];
    for (int a = 0; a < 24; ++a)
      for (int s = 0; s < 24; ++s) {
        double w = 0.0;
        if (a % 3 == s % 3) {
          const int ma = a / 3, ms = s / 3;
          w = w1h((c & 1) + (ma & 1), ms & 1) * w1h(((c >> 1) & 1) + ((ma >> 1) & 1), (ms >> 1) & 1) *
              w1h(((c >> 2) & 1) + ((ma >> 2) & 1), (ms >> 2) & 1);
        }
        W[a * 24 + s] = w;
      }
    double KW[576];
    for (int a = 0; a < 24; ++a)
      for (int s = 0; s < 24; ++s) {
        double v = 0.0;
        for (int b = 0; b < 24; ++b) v += P.Kb[a * 24 + b] * W[b * 24 + s];
        KW[a * 24 + s] = v;
      }
    for (int r = 0; r < 24; ++r)
      for (int s = 0; s < 24; ++s) {
        double v = 0.0;
        for (int a = 0; a < 24; ++a) v += W[a * 24 + r] * KW[a * 24 + s];
        P.S8[c * 576 + r * 24 + s] = v;
      }
  }
  // Two coarsenings at once: T[c2 * 8 + c1] = W_{c2}^T S_{c1} W_{c2}, the
  // level-2 contribution of fine element c1 of level-1 child c2 per unit modulus.
  P.T64.assign(64 * 576, 0.0);
  for (int c2 = 0; c2 < 8; ++c2) {
    double W[576];
    for (int a = 0; a < 24; ++a)
      for (int s = 0; s < 24; ++s) {
        double w = 0.0;
        if (a % 3 == s % 3) {
          const int ma = a / 3, ms = s / 3;
          w = w1h((c2 & 1) + (ma & 1), ms & 1) *
              w1h(((c2 >> 1) & 1) + ((ma >> 1) & 1), (ms >> 1) & 1) *
              w1h(((c2 >> 2) & 1) + ((ma >> 2) & 1), (ms >> 2) & 1);
        }
        W[a * 24 + s] = w;
      }
    for (int c1 = 0; c1 < 8; ++c1) {
      double SW[576];
      for (int a = 0; a < 24; ++a)
        for (int s = 0; s < 24; ++s) {
          double v = 0.0;
          for (int b = 0; b < 24; ++b) v += P.S8[c1 * 576 + a * 24 + b] * W[b * 24 + s];
          SW[a * 24 + s] = v;
        }
      for (int r = 0; r < 24; ++r)
        for (int s = 0; s < 24; ++s) {
          double v = 0.0;
          for (int a = 0; a < 24; ++a) v += W[a * 24 + r] * SW[a * 24 + s];
          P.T64[(c2 * 8 + c1) * 576 + r * 24 + s] = v;
        }
    }
  }
}

// Coarsening plan (build_mf_hierarchy + build_hierarchy acceptance rules,
// core/src/fea/multigrid.cpp:557-716, 1344-1515; unpadded geometric branch).
void plan_levels(Problem& P) {
  P.L.clear();
  auto lv0 = std::make_unique<Level>();
  lv0->g = P.g;
  const long long nn = P.g.nodes();
  lv0->mask_h.assign(static_cast<size_t>(nn), 0);
  long long nact = 0;
  for (long long n = 0; n < nn; ++n) {
    const uint8_t m = P.touched_h[n] ? static_cast<uint8_t>(~P.fixed_h[n] & 7) : 0;
    lv0->mask_h[n] = m;
    nact += popcount3(m);
  }
  lv0->nact = nact;
  P.L.push_back(std::move(lv0));
  P.mg_plan_ok = false;
  P.l1_direct = false;

  auto coarsen = [](const Level& f, std::unique_ptr<Level>& out) -> bool {
    if ((f.g.ex & 1) || (f.g.ey & 1) || (f.g.ez & 1)) return false;
    if (f.g.ex / 2 < kMinCoarseElems || f.g.ey / 2 < kMinCoarseElems || f.g.ez / 2 < kMinCoarseElems)
      return false;
    out = std::make_unique<Level>();
    out->g = make_geom(f.g.ex / 2, f.g.ey / 2, f.g.ez / 2);
    const Geom& gc = out->g;
    out->mask_h.assign(static_cast<size_t>(gc.nodes()), 0);
    long long na = 0;
    for (int c = 0; c < gc.nnz; ++c)
      for (int b = 0; b < gc.nny; ++b)
        for (int a = 0; a < gc.nnx; ++a) {
          const long long fn = (static_cast<long long>(2 * c) * f.g.nny + 2 * b) * f.g.nnx + 2 * a;
          const long long cn = (static_cast<long long>(c) * gc.nny + b) * gc.nnx + a;
          out->mask_h[cn] = f.mask_h[fn];
          na += popcount3(f.mask_h[fn]);
        }
    out->nact = na;
    return true;
  };

  std::unique_ptr<Level> l1;
  if (!coarsen(*P.L[0], l1) || l1->nact == 0) return;
  std::vector<std::unique_ptr<Level>> sub;
  sub.push_back(std::move(l1));
  while (true) {
    std::unique_ptr<Level> nx;
    if (!coarsen(*sub.back(), nx)) break;
    if (nx->nact == 0) break;
    const long long nc = nx->nact;
    sub.push_back(std::move(nx));
    if (nc <= kCoarseDofCap) break;
  }
  const bool ok_sub = sub.size() >= 2 && sub.back()->nact <= kCoarseDofCap;
  if (!ok_sub) {
    if (sub[0]->nact > kCoarseDofCap) return;
    sub.resize(1);
  }
  for (auto& l : sub) P.L.push_back(std::move(l));
  P.mg_plan_ok = true;
  P.l1_direct = P.L.size() >= 3;
}

template <typename T> T* Kapp(Level& l);
template <> float* Kapp<float>(Level& l) { return l.K32.p; }
template <> double* Kapp<double>(Level& l) { return l.K64.p; }
template <typename T> T* Dinv(Level& l);
template <> float* Dinv<float>(Level& l) { return l.dinv32.p; }
template <> double* Dinv<double>(Level& l) { return l.dinv64.p; }
template <typename T> T* Bv(Level& l);
template <> float* Bv<float>(Level& l) { return l.b32.p; }
template <> double* Bv<double>(Level& l) { return l.b64.p; }
template <typename T> T* Xv(Level& l);
template <> float* Xv<float>(Level& l) { return l.x32.p; }
template <> double* Xv<double>(Level& l) { return l.x64.p; }
template <typename T> T* Tv(Level& l);
template <> float* Tv<float>(Level& l) { return l.t32.p; }
template <> double* Tv<double>(Level& l) { return l.t64.p; }
template <typename T> T* Dv(Level& l);
template <> float* Dv<float>(Level& l) { return l.d32.p; }
template <> double* Dv<double>(Level& l) { return l.d64.p; }
template <typename T> const T* Ev(Problem& P);
template <> const float* Ev<float>(Problem& P) { return P.E32.p; }
template <> const double* Ev<double>(Problem& P) { return P.E64.p; }
template <typename T> const ModalGroups<T>& Mg(Problem& P);
template <> const ModalGroups<float>& Mg<float>(Problem& P) { return P.mg32; }
template <> const ModalGroups<double>& Mg<double>(Problem& P) { return P.mg64; }

void alloc_levels(Problem& P) {
  const bool mixed = P.precision == TOPOHIP_PREC_MIXED;
  for (size_t l = 0; l < P.L.size(); ++l) {
    Level& lv = *P.L[l];
    const long long nd = lv.g.ndof();
    lv.mask.ensure(static_cast<size_t>(lv.g.nodes()));
    h2d(lv.mask.p, lv.mask_h.data(), lv.mask_h.size(), P.s);
    // FP64 diagonal always (Jacobi-CG fallback on the fine level).
    lv.dinv64.ensure(static_cast<size_t>(nd));
    if (mixed) lv.dinv32.ensure(static_cast<size_t>(nd));
    if (l > 0) {
      if (!(l == 1 && P.l1_direct)) lv.Kem.ensure(static_cast<size_t>(lv.g.elems()) * 576);
      if (l == 1) {
        // Masks are fixed per problem (BCs and solid tags), so the coarse
        // elements that need the masked Galerkin path are listed once.
        const Level& f = *P.L[0];
        std::vector<int> sp;
        for (int K = 0; K < lv.g.ez; ++K)
          for (int J = 0; J < lv.g.ey; ++J)
            for (int I = 0; I < lv.g.ex; ++I) {
              bool any = false;
              for (int c = 0; c < 3 && !any; ++c)
                for (int b = 0; b < 3 && !any; ++b)
                  for (int a = 0; a < 3 && !any; ++a)
                    any = f.mask_h[(static_cast<long long>(2 * K + c) * f.g.nny + 2 * J + b) *
                                       f.g.nnx + 2 * I + a] != 7;
              if (any) sp.push_back((K * lv.g.ey + J) * lv.g.ex + I);
            }
        lv.n_special = static_cast<int>(sp.size());
        if (!sp.empty()) {
          lv.special.ensure(sp.size());
          h2d(lv.special.p, sp.data(), sp.size(), P.s);
          sync(P);
        }
        if (P.l1_direct) {
          lv.K1sp.ensure(std::max<size_t>(1, sp.size()) * 576);
          std::vector<int> map(static_cast<size_t>(lv.g.elems()), -1);
          for (size_t i = 0; i < sp.size(); ++i) map[static_cast<size_t>(sp[i])] = static_cast<int>(i);
          lv.spmap.ensure(map.size());
          h2d(lv.spmap.p, map.data(), map.size(), P.s);
          sync(P);
        }
      }
      if (l == 2 && P.l1_direct) {
        // Level-2 elements whose 125 fine nodes are not all fully active take
        // the Galerkin product over their level-1 children.
        const Level& f = *P.L[0];
        std::vector<int> sp;
        for (int K = 0; K < lv.g.ez; ++K)
          for (int J = 0; J < lv.g.ey; ++J)
            for (int I = 0; I < lv.g.ex; ++I) {
              bool any = false;
              for (int c = 0; c < 5 && !any; ++c)
                for (int b = 0; b < 5 && !any; ++b)
                  for (int a = 0; a < 5 && !any; ++a)
                    any = f.mask_h[(static_cast<long long>(4 * K + c) * f.g.nny + 4 * J + b) *
                                       f.g.nnx + 4 * I + a] != 7;
              if (any) sp.push_back((K * lv.g.ey + J) * lv.g.ex + I);
            }
        lv.n_special = static_cast<int>(sp.size());
        if (!sp.empty()) {
          lv.special.ensure(sp.size());
          h2d(lv.special.p, sp.data(), sp.size(), P.s);
          sync(P);
        }
      }
      const bool coarsest = l + 1 == P.L.size();
      if (!coarsest) {
        // symmetric half stencil: centre + 13 neighbour blocks of 3x3
        if (mixed) lv.K32.ensure(static_cast<size_t>(lv.g.nodes()) * 126);
        else lv.K64.ensure(static_cast<size_t>(lv.g.nodes()) * 126);
      }
    }
  }
  if (P.mg_plan_ok && !P.cr.plan(P.L.back()->g, P.s)) P.mg_plan_ok = false;
}

int max_partials(const Problem& P) {
  const int ku = std::max(ku_partials<float>(P.g), ku_partials<double>(P.g));
  return std::max(std::max(ku, 2 * cheb_blocks(P.g.ndof())),
                  std::max(dot_blocks(P.g.ndof()), sens_blocks(P.g)));
}

template <typename T>
int ku_tz(const Problem& P) { return sizeof(T) == 8 ? P.tz64 : P.tz32; }

// ---- V-cycle ----------------------------------------------------------------
// All level vectors hold nc stacked load cases (stride: the level's ndof);
// every launch covers the nc cases, and a case whose PCG is done is skipped.
// Coarse levels l >= 1: restrict the level-(l-1) residual rprev, smooth,
// recurse, prolong, smooth. Result lives in t_l (x_l at the coarsest).
// Chebyshev smoothing of one level (degree operator applies from x, one
// fewer from zero): res(x, t) forms t = b - A x; the last step writes xo
// (which may be t: each entry is read before it is overwritten).
template <typename T, class Res>
void cheb_smooth(Problem& P, Level& lv, T* x, T* xo, T* t, const T* b, bool zero_start,
                 const int* done, int nc, Res res) {
  const long long n = lv.g.ndof();
  T* d = Dv<T>(lv);
  const T* dinv = Dinv<T>(lv);
  const double* coef = lv.cheb.p;
  const int m = P.cheb_degree;
  if (!zero_start) res(x, t);
  launch_cheb_first<T>(x, m == 1 ? xo : x, d, zero_start ? b : t, dinv, coef, n, zero_start, done,
                       P.s, nc);
  for (int k = 1; k < m; ++k) {
    res(x, t);
    launch_cheb_step<T>(x, k + 1 == m ? xo : x, d, t, dinv, coef, k, n, done, P.s, nc);
  }
}

template <typename T>
void coarse_cycle(Problem& P, size_t l, const T* rprev, const int* done, int nc) {
  Level& f = *P.L[l - 1];
  Level& c = *P.L[l];
  const bool coarsest = l + 1 == P.L.size();
  if (coarsest) {
    // restrict into the coarse solver's plane-ordered FP64 space, solve there;
    // the prolongation reads the solution from the same space
    launch_restrict<T, T>(f.g, c.g, rprev, c.mask.p, static_cast<T*>(nullptr), nullptr, nullptr,
                          0.0, done, P.s, nc, P.cr.rhs_map());
    P.cr.solve_in_place(done, P.s, nc);
    return;
  }
  if (P.smoother == 1) {
    auto res = [&](const T* x, T* t) {
      launch_stencil_apply<T>(c.g, Kapp<T>(c), c.mask.p, x, t, Bv<T>(c), nullptr, 0.0, EPI_RES,
                              done, P.s, nc);
    };
    launch_restrict<T, T>(f.g, c.g, rprev, c.mask.p, Bv<T>(c), static_cast<T*>(nullptr), nullptr,
                          0.0, done, P.s, nc);
    cheb_smooth<T>(P, c, Xv<T>(c), Xv<T>(c), Tv<T>(c), Bv<T>(c), true, done, nc, res);
    res(Xv<T>(c), Tv<T>(c));
    coarse_cycle<T>(P, l + 1, Tv<T>(c), done, nc);
    Level& n = *P.L[l + 1];
    const bool ncoarsest = l + 2 == P.L.size();
    launch_prolong_add<T, T>(n.g, c.g, ncoarsest ? nullptr : Tv<T>(n), c.mask.p, Xv<T>(c), done,
                             P.s, nc, ncoarsest ? P.cr.sol_map() : CrMap());
    // the level's result goes to t_l, where the parent's prolongation reads it
    cheb_smooth<T>(P, c, Xv<T>(c), Tv<T>(c), Tv<T>(c), Bv<T>(c), false, done, nc, res);
    return;
  }
  launch_restrict<T, T>(f.g, c.g, rprev, c.mask.p, Bv<T>(c), Xv<T>(c), Dinv<T>(c), kOmega, done,
                        P.s, nc);
  launch_stencil_apply<T>(c.g, Kapp<T>(c), c.mask.p, Xv<T>(c), Tv<T>(c), Bv<T>(c), nullptr, 0.0,
                          EPI_RES, done, P.s, nc);
  coarse_cycle<T>(P, l + 1, Tv<T>(c), done, nc);
  Level& n = *P.L[l + 1];
  const bool ncoarsest = l + 2 == P.L.size();
  launch_prolong_add<T, T>(n.g, c.g, ncoarsest ? nullptr : Tv<T>(n), c.mask.p, Xv<T>(c), done,
                           P.s, nc, ncoarsest ? P.cr.sol_map() : CrMap());
  launch_stencil_apply<T>(c.g, Kapp<T>(c), c.mask.p, Xv<T>(c), Tv<T>(c), Bv<T>(c), Dinv<T>(c),
                          kOmega, EPI_JAC, done, P.s, nc);
}

template <typename T>
KuOp<T> fine_op(Problem& P, const int* done, int nc) {
  KuOp<T> op;
  op.g = P.g;
  op.mask = P.L[0]->mask.p;
  op.E = Ev<T>(P);
  op.mg = Mg<T>(P);
  op.tz = ku_tz<T>(P);
  op.done = done;
  op.nrhs = nc;
  op.pstride = P.pstride;
  return op;
}

// z = M^{-1} b0 on the fine level (mf_v_cycle with 1+1 damped Jacobi sweeps;
// the first sweep starts from zero so it is the scaling omega D^-1 b0, which
// the first operator apply forms on the fly). With rz_partials set, the last
// apply also writes the per-block partials of <r64, z> there.
template <typename T>
void vcycle(Problem& P, const T* b0, T* z, const int* done, int nc, const double* r64,
            double* rz_partials) {
  Level& f = *P.L[0];
  if (P.smoother == 1) {
    auto res = [&](const T* x, T* t) {
      KuOp<T> op = fine_op<T>(P, done, nc);
      op.x = x;
      op.y = t;
      op.b = b0;
      op.epi = EPI_RES;
      launch_ku_fine<T>(op, P.s);
    };
    cheb_smooth<T>(P, f, Xv<T>(f), Xv<T>(f), Tv<T>(f), b0, true, done, nc, res);
    res(Xv<T>(f), Tv<T>(f));
    coarse_cycle<T>(P, 1, Tv<T>(f), done, nc);
    Level& c = *P.L[1];
    const bool ccoarsest = P.L.size() == 2;
    launch_prolong_add<T, T>(c.g, f.g, ccoarsest ? nullptr : Tv<T>(c), f.mask.p, Xv<T>(f), done,
                             P.s, nc, ccoarsest ? P.cr.sol_map() : CrMap());
    cheb_smooth<T>(P, f, Xv<T>(f), z, Tv<T>(f), b0, false, done, nc, res);
    launch_dot<double, T>(r64, z, f.g.ndof(), rz_partials, dot_blocks(f.g.ndof()), done, P.s, nc,
                          P.pstride);
    return;
  }
  KuOp<T> op = fine_op<T>(P, done, nc);
  op.b = b0;
  op.dinv = Dinv<T>(f);
  op.omega_d = kOmega;
  op.xw = Xv<T>(f);
  op.y = Tv<T>(f);
  op.epi = EPI_RES0;
  launch_ku_fine<T>(op, P.s);
  coarse_cycle<T>(P, 1, Tv<T>(f), done, nc);
  Level& c = *P.L[1];
  const bool ccoarsest = P.L.size() == 2;
  launch_prolong_add<T, T>(c.g, f.g, ccoarsest ? nullptr : Tv<T>(c), f.mask.p, Xv<T>(f), done,
                           P.s, nc, ccoarsest ? P.cr.sol_map() : CrMap());
  op.epi = EPI_JAC;
  op.x = Xv<T>(f);
  op.xw = nullptr;
  op.omega = static_cast<T>(kOmega);
  op.y = z;
  op.partials = rz_partials;
  op.rdot = r64;
  launch_ku_fine<T>(op, P.s);
}

template <typename T>
void fine_apply(Problem& P, const T* x, T* y, const T* b, int epi, double* partials,
                const int* done, int nc) {
  KuOp<T> op = fine_op<T>(P, done, nc);
  op.x = x;
  op.y = y;
  op.b = b;
  op.epi = epi;
  op.partials = partials;
  launch_ku_fine<T>(op, P.s);
}

enum class Prec { MG32, MG64, JACOBI };

const int* done_flags(Problem& P) { return &P.st.p->done; }

// Preconditioner application z = M^{-1} r (mixed: r32 must hold float(r)),
// with the partials of <r, z> written to P.partials; returns their count per
// case. The V-cycle's last operator apply forms them; Jacobi uses a dot pass.
// Count of <r, z> partials per case the preconditioner writes (the V-cycle's
// last fine apply forms them in the V-cycle's precision; Jacobi and the
// Chebyshev smoother use a dot pass).
int rz_count(const Problem& P, Prec pc) {
  const long long nd = P.g.ndof();
  if (pc == Prec::JACOBI || P.smoother == 1) return dot_blocks(nd);
  return pc == Prec::MG32 ? ku_partials<float>(P.g) : ku_partials<double>(P.g);
}

// Preconditioner application z = M^{-1} r (mixed: r32 must hold float(r)),
// with the partials of <r, z> written to P.rzpart (rz_count per case).
void precondition_rz(Problem& P, Prec pc, const int* done, int nc) {
  const long long nd = P.g.ndof();
  if (pc == Prec::MG32) {
    vcycle<float>(P, P.r32.p, P.z32.p, done, nc, P.r.p, P.rzpart.p);
    return;
  }
  if (pc == Prec::MG64) {
    vcycle<double>(P, P.r.p, P.z64.p, done, nc, P.r.p, P.rzpart.p);
    return;
  }
  launch_jacobi0<double>(P.z64.p, P.L[0]->dinv64.p, P.r.p, 1.0, nd, done, P.s, nc);
  launch_dot<double, double>(P.r.p, P.z64.p, nd, P.rzpart.p, dot_blocks(nd), done, P.s, nc,
                             P.pstride);
}

double* direction(Problem& P, int i) { return i == 0 ? P.p.p : P.p2.p; }

// Paired solution updates in cg_xr (up to 32 cases: the width of the masks
// that settle a stopped case's deferred update).
bool paired_updates(int nc) { return nc <= 32; }

// The PCG's A.p carries the direction update p = z + beta p while its grid
// fits in about one round of resident workgroups (the FP64 apply holds three
// waves per SIMD: three workgroups per CU). Past that the separate update is
// cheaper: the launches then overlap one another's latency, and the apply's
// extra z and p streams cost it more than the update kernel's own pass
// (192x64x64: 700 tiles on 768 slots, fused 29.5 us against 22.4 + 11.2 us;
// 512x256x256: 8177 tiles, fused 1184 us against 601 us + the update's pass).
// TOPOHIP_PCG_DIRECTION=fused or =separate overrides the choice (both give
// the same bits: the same fixed-order beta and the same fma).
bool fused_direction(const Problem& P, int nc) {
  if (const char* env = std::getenv("TOPOHIP_PCG_DIRECTION")) {
    if (std::strcmp(env, "fused") == 0) return true;
    if (std::strcmp(env, "separate") == 0) return false;
  }
  const Geom& g = P.g;
  const long long tiles = static_cast<long long>((g.nnx + kTX - 2) / (kTX - 1)) *
                          ((g.nny + kTY - 2) / (kTY - 1)) * ((g.nnz + P.tz64 - 1) / P.tz64) * nc;
  return 2 * tiles <= 3LL * 3 * P.cus;
}

// Restart the recurrence of every case that is not done from its current r:
// z = M^-1 r, rz = r.z, and p = z: the next A.p takes it (cg_set_rz marks it),
// or, without the fused update, it is written to the current direction buffer.
void restart(Problem& P, Prec pc, int nc) {
  const long long nd = P.g.ndof();
  const int* done = done_flags(P);
  if (pc == Prec::MG32) launch_convert<double, float>(P.r.p, P.r32.p, nd * nc, P.s);
  precondition_rz(P, pc, done, nc);
  launch_cg_set_rz(P.st.p, P.rzpart.p, rz_count(P, pc), P.s, nc, P.pstride);
  if (fused_direction(P, nc)) return;
  double* p = direction(P, P.pcur);
  if (pc == Prec::MG32)
    launch_cg_p<float>(P.st.p, P.z32.p, p, p, nd, 1, nullptr, 0, P.s, nc, P.pstride);
  else
    launch_cg_p<double>(P.st.p, P.z64.p, p, p, nd, 1, nullptr, 0, P.s, nc, P.pstride);
}

// One PCG iteration at direction-buffer parity par (the next iteration runs
// at par ^ 1). Fused (fused_direction): buffer par holds the previous
// direction and the FP64 A.p forms p = z + beta p into buffer par ^ 1 (beta
// from the previous preconditioner's <r, z> partials; p = z after a restart).
// Separate: buffer par holds p, and cg_p writes the next direction into buffer
// par ^ 1 after the preconditioner. Between them x, r and the stopping test,
// then z = M^-1 r with its <r, z> partials.
void iteration(Problem& P, Prec pc, int nc, int par) {
  const long long nd = P.g.ndof();
  const int* done = done_flags(P);
  const int kb = ku_partials<double>(P.g);  // A.p is FP64
  const int db = dot_blocks(nd);
  // A.p partials go to `partials`, the rr partials of cg_xr to `partials2`
  // (cg_xr re-sums the former in every block while it writes the latter).
  // A.p is FP64 in every form (it defines the residual, as mf_mgpcg's
  // mf_fine_matvec, core/src/fea/multigrid.cpp:1286); the mixed form's FP32
  // V-cycle reads float(r), which cg_xr writes alongside r.
  const bool fused = fused_direction(P, nc);
  // the direction this iteration applies: formed by the A.p into buffer
  // par ^ 1, or (separate update) already in buffer par
  double* p = direction(P, fused ? par ^ 1 : par);
  KuOp<double> op = fine_op<double>(P, done, nc);
  op.y = P.Ap64.p;
  op.epi = EPI_AX;
  op.partials = P.partials.p;
  if (fused) {
    op.x = direction(P, par);
    op.pw = p;
    if (pc == Prec::MG32) op.pz32 = P.z32.p; else op.pz64 = P.z64.p;
    op.st = P.st.p;
    op.rzp = P.rzpart.p;
    op.nrz = rz_count(P, pc);
  } else {
    op.x = p;
  }
  launch_ku_fine<double>(op, P.s);
  // (Jacobi-CG: cg_xr also forms z = D^-1 r and the <r, z> partials; the
  // solution updates are paired over two iterations: the previous direction
  // is the other buffer, fused or not)
  const bool jac = pc == Prec::JACOBI;
  launch_cg_xr<double, double>(P.st.p, P.x.p, P.r.p, p, P.Ap64.p,
                               pc == Prec::MG32 ? P.r32.p : nullptr, nd, P.partials.p, kb,
                               P.partials2.p, db, P.s, nc, P.pstride,
                               jac ? P.L[0]->dinv64.p : nullptr, jac ? P.z64.p : nullptr,
                               jac ? P.rzpart.p : nullptr,
                               paired_updates(nc) ? direction(P, fused ? par : par ^ 1) : nullptr);
  launch_cg_check(P.st.p, P.partials2.p, db, P.s, nc, P.pstride);
  if (!jac) precondition_rz(P, pc, done, nc);
  if (fused) return;
  // the next direction z + beta p into buffer par ^ 1, as the fused A.p would
  if (pc == Prec::MG32)
    launch_cg_p<float>(P.st.p, P.z32.p, p, direction(P, par ^ 1), nd, 0, P.rzpart.p,
                       rz_count(P, pc), P.s, nc, P.pstride);
  else
    launch_cg_p<double>(P.st.p, P.z64.p, p, direction(P, par ^ 1), nd, 0, P.rzpart.p,
                        rz_count(P, pc), P.s, nc, P.pstride);
}

void read_state(Problem& P, int nc) { d2h(P.st_h, P.st.p, static_cast<size_t>(nc), P.s); }

// rt = b - A x in FP64 for every case; scal_h[r] = ||rt_r||^2.
void true_residual(Problem& P, int nc) {
  const long long nd = P.g.ndof();
  P.matvecs += nc;
  fine_apply<double>(P, P.x.p, P.rt.p, P.b.p, EPI_RES, nullptr, nullptr, nc);
  launch_dot<double, double>(P.rt.p, P.rt.p, nd, P.partials.p, dot_blocks(nd), nullptr, P.s, nc,
                             P.pstride);
  launch_sum_partials(P.partials.p, dot_blocks(nd), P.scal.p, P.s, nc, P.pstride);
  d2h(P.scal_h, P.scal.p, static_cast<size_t>(nc), P.s);
}

void set_done(Problem& P, int r, int v) {
  TOPOHIP_CHECK(hipMemcpyAsync(&P.st.p[r].done, &v, sizeof(int), hipMemcpyHostToDevice, P.s));
  sync(P);
}

// Block PCG over nc stacked load cases. Each case keeps its own recurrence
// (alpha, beta, stopping test, breakdown checks, as mf_mgpcg,
// core/src/fea/multigrid.cpp:1211-1318), but every operator apply, V-cycle
// and vector kernel is one launch over all cases, and a converged case drops
// out of the launches. x must hold the start vectors. Cases with skip[r] are
// left untouched and count as solved. A case is solved when its FP64 true
// residual meets ||b - A x|| <= tol ||b||; the mixed form restarts a case
// whose true residual misses it from r = b - A x (residual replacement).
void pcg(Problem& P, Prec pc, int nc, double tol, int cap, bool warm, const bool* skip,
         int* iters, double* resid, bool* ok) {
  const long long nd = P.g.ndof();
  const int db = dot_blocks(nd);
  if (warm) {
    fine_apply<double>(P, P.x.p, P.r.p, P.b.p, EPI_RES, nullptr, nullptr, nc);
  } else {
    TOPOHIP_CHECK(hipMemcpyAsync(P.r.p, P.b.p, nd * nc * sizeof(double),
                                 hipMemcpyDeviceToDevice, P.s));
  }
  launch_dot<double, double>(P.b.p, P.b.p, nd, P.partials.p, db, nullptr, P.s, nc, P.pstride);
  // A warm-started MG-PCG is not tested before its first iteration (the
  // stopping test follows each update, core/src/fea/multigrid.cpp:1250-1300);
  // a warm-started Jacobi-CG returns at once when its initial residual already
  // meets the tolerance (core/src/fea/matfree.cpp:956-967).
  launch_cg_init(P.st.p, P.partials.p, db, tol, cap, P.s, nc, P.pstride);
  if (warm && pc == Prec::JACOBI) {
    launch_dot<double, double>(P.r.p, P.r.p, nd, P.partials.p, db, nullptr, P.s, nc, P.pstride);
    launch_cg_resume(P.st.p, P.partials.p, db, P.s, nc, P.pstride);
  }
  for (int r = 0; r < nc; ++r)
    if (skip && skip[r]) set_done(P, r, 1);
  restart(P, pc, nc);
  read_state(P, nc);
  for (int r = 0; r < nc; ++r) {
    iters[r] = 0;
    resid[r] = 0.0;
    ok[r] = skip && skip[r];
    if (!ok[r] && P.st_h[r].bnorm2 == 0.0) {
      launch_fill<double>(P.x.p + r * nd, 0.0, nd, P.s);
      ok[r] = true;  // done was set by the init
    }
  }
  auto all_done = [&] {
    for (int r = 0; r < nc; ++r)
      if (!P.st_h[r].done) return false;
    return true;
  };
  // The iteration body is captured into HIP graphs (all pointers are fixed
  // while the capacity stands) and replayed; every kernel in them skips cases
  // whose done flag is set. Per preconditioner kind there are four: a batch of
  // kBatch iterations and a single one, each starting from either direction
  // buffer (a batch keeps the buffer parity, a single iteration flips it).
  // Relaxed capture: the body makes no synchronising calls, and other threads
  // (other problems of the serving loop) keep allocating and copying meanwhile.
  constexpr int kBatch = 4;
  const int gi = static_cast<int>(pc);
  for (int par = 0; par < 2; ++par)
    for (int len = 0; len < 2; ++len) {
      const int k = (gi * 2 + par) * 2 + len;
      hipGraphExec_t& gx = P.cg_graph[k];
      if (gx && P.cg_graph_nc[k] == nc) continue;
      std::lock_guard<std::recursive_mutex> lk(device_api_mutex());
      if (gx) (void)hipGraphExecDestroy(gx);
      gx = nullptr;
      hipGraph_t gr;
      TOPOHIP_CHECK(hipStreamBeginCapture(P.s, hipStreamCaptureModeRelaxed));
      const int n = len == 0 ? kBatch : 1;
      for (int i = 0; i < n; ++i) iteration(P, pc, nc, par ^ (i & 1));
      TOPOHIP_CHECK(hipStreamEndCapture(P.s, &gr));
      TOPOHIP_CHECK(hipGraphInstantiate(&gx, gr, nullptr, nullptr, 0));
      TOPOHIP_CHECK(hipGraphDestroy(gr));
      P.cg_graph_nc[k] = nc;
    }
  auto launch_iterations = [&](int n) {
    for (; n >= kBatch; n -= kBatch)
      TOPOHIP_CHECK(hipGraphLaunch(P.cg_graph[(gi * 2 + P.pcur) * 2], P.s));
    for (; n > 0; --n) {
      TOPOHIP_CHECK(hipGraphLaunch(P.cg_graph[(gi * 2 + P.pcur) * 2 + 1], P.s));
      P.pcur ^= 1;
    }
  };
  const long long per_it = pc == Prec::JACOBI ? 1 : 3;
  // Iterations launched before each read-back of the states. The first group
  // is as many as the previous solve of this kind needed, less a margin
  // (consecutive design iterations take similar counts); after each read-back
  // the next group is the count at which the slowest unfinished case reaches
  // its threshold if its residual keeps falling at the rate seen since the
  // previous read-back. Iterations past convergence are no-ops (every kernel
  // checks the done flag) but still cost their launches, and a read-back
  // costs a round trip, so the prediction saves both.
  std::vector<double> rr_seen(nc);
  std::vector<int> it_seen(nc);
  auto remember = [&]() {
    for (int r = 0; r < nc; ++r) {
      rr_seen[r] = P.st_h[r].rr;
      it_seen[r] = P.st_h[r].iters;
    }
  };
  auto predict = [&]() {
    int need = 1;
    for (int r = 0; r < nc; ++r) {
      const CgState& st = P.st_h[r];
      if (st.done) continue;
      const int di = st.iters - it_seen[r];
      if (di <= 0 || !(st.rr > 0.0) || !(st.rr < rr_seen[r]) || !(st.thresh2 > 0.0)) {
        need = std::max(need, kBatch);
        continue;
      }
      const double rate = std::log(st.rr / rr_seen[r]) / di;  // < 0
      const double k = std::ceil(std::log(st.thresh2 / st.rr) / rate);
      need = std::max(need, static_cast<int>(std::min(64.0, std::max(1.0, k))));
    }
    return need;
  };
  remember();
  // Paired solution updates: case r's recurrence restarted at iteration it0
  // from direction parity par0; its iteration i ran at parity par0 ^ (i & 1)
  // and applied the direction in buffer (that parity) ^ 1 when the A.p forms it
  // (fused), in buffer (that parity) otherwise. A recurrence that stopped after
  // an odd count of iterations owes its last, deferred update.
  const bool fusedd = fused_direction(P, nc);
  int par0 = P.pcur;
  auto finish_x = [&]() {
    if (!paired_updates(nc)) return;
    unsigned pend = 0, par = 0;
    for (int r = 0; r < nc; ++r) {
      const int c = P.st_h[r].iters - P.st_h[r].it0;
      if ((c & 1) == 0) continue;
      pend |= 1u << r;
      if ((par0 ^ ((c - 1) & 1) ^ (fusedd ? 1 : 0)) != 0) par |= 1u << r;
      P.st_h[r].it0 = P.st_h[r].iters;  // the kernel does the same on the device
    }
    launch_cg_xfinish(P.st.p, P.x.p, P.p.p, P.p2.p, nd, pend, par, P.s, nc);
  };
  int next = std::max(kBatch, std::min(64, P.pcg_hint[gi] - 3));
  for (long long guard = 0;; ++guard) {
    if (guard > 4LL * cap + 64) throw Error(TOPOHIP_EINTERNAL, "pcg: iteration state stalled");
    // An armed solve deadline is polled before every group of iterations (the
    // reference polls every 256 iterations, matfree.cpp:988-1001); when it has
    // passed the unfinished cases stop with the state they reached.
    if (deadline_passed()) {
      finish_x();
      for (int r = 0; r < nc; ++r) {
        if (ok[r] || (skip && skip[r])) continue;
        iters[r] = P.st_h[r].iters;
        resid[r] = P.st_h[r].bnorm2 > 0.0 ? std::sqrt(P.st_h[r].rr / P.st_h[r].bnorm2) : 0.0;
      }
      P.deadline_hit = true;
      return;
    }
    if (!all_done()) {
      long long before = 0;
      for (int r = 0; r < nc; ++r) before += P.st_h[r].iters;
      launch_iterations(next);
      read_state(P, nc);
      long long after = 0;
      for (int r = 0; r < nc; ++r) after += P.st_h[r].iters;
      P.matvecs += (after - before) * per_it;
      if (!all_done()) {
        next = predict();
        remember();
        continue;
      }
    }
    // Every recurrence has stopped: confirm with the FP64 true residuals.
    finish_x();
    bool any_pending = false;
    for (int r = 0; r < nc; ++r)
      if (!ok[r] && !P.st_h[r].fail) any_pending = true;
    if (!any_pending) break;
    true_residual(P, nc);
    bool restarted = false;
    for (int r = 0; r < nc; ++r) {
      if (ok[r]) continue;
      const CgState& st = P.st_h[r];
      iters[r] = st.iters;
      if (st.fail) {
        resid[r] = std::sqrt(st.rr / st.bnorm2);
        continue;
      }
      const double rr = P.scal_h[r];
      resid[r] = std::sqrt(rr / st.bnorm2);
      if (!std::isfinite(rr)) {
        set_done(P, r, 1);
        P.st_h[r].fail = 1;
        continue;
      }
      if (rr <= st.thresh2) {
        ok[r] = true;
        continue;
      }
      if (st.iters >= cap) {
        P.st_h[r].fail = 2;
        continue;
      }
      // Residual replacement for this case: continue from r = b - A x.
      ++P.replacements;
      TOPOHIP_CHECK(hipMemcpyAsync(P.r.p + r * nd, P.rt.p + r * nd, nd * sizeof(double),
                                   hipMemcpyDeviceToDevice, P.s));
      launch_dot<double, double>(P.r.p + r * nd, P.r.p + r * nd, nd,
                                 P.partials.p + static_cast<long long>(r) * P.pstride, db,
                                 nullptr, P.s);
      launch_cg_resume(P.st.p + r, P.partials.p + static_cast<long long>(r) * P.pstride, db, P.s);
      restarted = true;
    }
    if (!restarted) break;
    restart(P, pc, nc);
    read_state(P, nc);
    remember();
    par0 = P.pcur;
    next = 2;  // a residual replacement near the threshold needs a few more
  }
  int most = 0;
  for (int r = 0; r < nc; ++r) {
    if (!(skip && skip[r])) {
      iters[r] = P.st_h[r].iters;
      // a recurrence that stopped on breakdown or at the cap reports the
      // residual it reached (SolverNonConvergence's residual)
      if (!ok[r] && P.st_h[r].fail && P.st_h[r].bnorm2 > 0.0)
        resid[r] = std::sqrt(P.st_h[r].rr / P.st_h[r].bnorm2);
    }
    most = std::max(most, P.st_h[r].iters);
  }
  P.pcg_hint[gi] = most;
}

}  // namespace

void reset_mg_stagnation_latch() {
  t_mg_stagnations = 0;
  t_mg_latched = false;
  t_mg_latched_solves = 0;
  t_mg_rearm_attempts = 0;
  t_mg_rearm_carries = 0;
}

// (the period survives the reset, as the reference's harness posture does)
void set_mg_rearm_period(int period) {
  t_mg_rearm_period = period > 0 ? period : 0;
  t_mg_latched_solves = 0;
}
int mg_rearm_period() { return t_mg_rearm_period; }
long long mg_rearm_attempts() { return t_mg_rearm_attempts; }
long long mg_rearm_carries() { return t_mg_rearm_carries; }

bool mg_stagnation_latched() { return t_mg_latched; }

double set_solve_deadline_ms(double absolute_steady_ms) {
  const double prev = t_solve_deadline_ms;
  t_solve_deadline_ms = absolute_steady_ms > 0.0 ? absolute_steady_ms : 0.0;
  return prev;
}

double steady_ms() { return now_ms(); }

// Launch-shape autotuner: time the fine-level apply for each z-chunk length
// on the problem's own buffers and keep the fastest, separately for the FP32
// and the FP64 applies (override: TOPOHIP_KU_TILE=tz or 32,8,tz, taken by each
// precision whose slab divides it, or tz32:tz64). The choice changes the
// launch shape only: every z-chunk computes bit-identical applies (each node
// sums its face contributions in a fixed order) and identical dot partials
// (one per fixed slab of planes), so results do not depend on what the tuner
// picks.
void autotune_ku(Problem& P) {
  bool have32 = false, have64 = false;
  if (const char* env = std::getenv("TOPOHIP_KU_TILE")) {
    int a = 0, b = 0, c = 0;
    if (std::sscanf(env, "%d:%d", &a, &b) == 2) {
      if (a >= ku_slab<float>() && a % ku_slab<float>() == 0) { P.tz32 = a; have32 = true; }
      if (b >= ku_slab<double>() && b % ku_slab<double>() == 0) { P.tz64 = b; have64 = true; }
    } else {
      const int n = std::sscanf(env, "%d,%d,%d", &a, &b, &c);
      const int tz = n == 3 ? (a == kTX && b == kTY ? c : 0) : (n == 1 ? a : 0);
      if (tz >= ku_slab<float>() && tz % ku_slab<float>() == 0) { P.tz32 = tz; have32 = true; }
      if (tz >= ku_slab<double>() && tz % ku_slab<double>() == 0) { P.tz64 = tz; have64 = true; }
    }
  }
  const bool overridden = have32 || have64;
  const bool mixed = P.precision == TOPOHIP_PREC_MIXED;
  if (!mixed) have32 = true;  // the FP64 form launches FP64 applies only
  if ((have32 && have64) || P.g.elems() < 4096) return;
  // The choice is cached per (device, grid, precision) for the life of the
  // process, so problems of one shape launch alike (an override bypasses it).
  static std::mutex mu;
  static std::map<std::array<int, 5>, std::array<int, 2>> cache;
  const std::array<int, 5> key{P.device, P.g.ex, P.g.ey, P.g.ez, P.precision};
  if (!overridden) {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
      if (!have32) P.tz32 = it->second[0];
      if (!have64) P.tz64 = it->second[1];
      return;
    }
  }
  const long long nd = P.g.ndof();
  launch_fill<double>(P.E64.p, 0.5, P.g.elems(), P.s);
  launch_convert<double, float>(P.E64.p, P.E32.p, P.g.elems(), P.s);
  launch_fill<double>(P.x.p, 1.0, nd, P.s);
  if (mixed) launch_fill<float>(P.r32.p, 1.0f, nd, P.s);
  hipEvent_t e0, e1;
  TOPOHIP_CHECK(hipEventCreate(&e0));
  TOPOHIP_CHECK(hipEventCreate(&e1));
  auto timed = [&](const std::function<void()>& run) {
    run();
    TOPOHIP_CHECK(hipEventRecord(e0, P.s));
    for (int i = 0; i < 5; ++i) run();
    TOPOHIP_CHECK(hipEventRecord(e1, P.s));
    TOPOHIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    TOPOHIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    return ms;
  };
  // Three rounds over the candidates, each keeping its fastest round, after a
  // few untimed launches: the clocks ramp during the first launches of a
  // process, which would otherwise penalise whichever candidate runs first.
  auto pick = [&](const std::array<int, 4>& tzs, const std::function<void(int)>& launch) {
    for (int i = 0; i < 8; ++i) launch(tzs[0]);
    std::array<float, 4> best;
    best.fill(1e30f);
    for (int round = 0; round < 3; ++round)
      for (int c = 0; c < 4; ++c)
        best[c] = std::min(best[c], timed([&]() { launch(tzs[c]); }));
    int c = 0;
    for (int k = 1; k < 4; ++k)
      if (best[k] < best[c]) c = k;
    return tzs[c];
  };
  if (!have32)
    P.tz32 = pick({4, 8, 16, 32}, [&](int tz) {
      KuOp<float> op;
      op.g = P.g; op.E = P.E32.p; op.mg = P.mg32; op.tz = tz;
      op.x = P.r32.p; op.y = P.z32.p; op.epi = EPI_AX;  // (scratch at set-up)
      launch_ku_fine<float>(op, P.s);
    });
  if (!have64)
    P.tz64 = pick({7, 14, 21, 28}, [&](int tz) {
      KuOp<double> op;
      op.g = P.g; op.E = P.E64.p; op.mg = P.mg64; op.tz = tz;
      op.x = P.x.p; op.y = P.Ap64.p; op.epi = EPI_AX;
      launch_ku_fine<double>(op, P.s);
    });
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (overridden) return;
  std::lock_guard<std::mutex> lk(mu);
  // a concurrent tuner may have won: every problem of the key launches alike
  const auto& won = cache.emplace(key, std::array<int, 2>{P.tz32, P.tz64}).first->second;
  if (!have32) P.tz32 = won[0];
  if (!have64) P.tz64 = won[1];
}

void drop_graphs(Problem& P) {
  std::lock_guard<std::recursive_mutex> lk(device_api_mutex());
  for (auto& g : P.cg_graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

// Capacity for nb stacked load cases: solver vectors, level work vectors,
// PCG states, partial sums and the coarse solver's vector space. Growing it
// moves buffers, so the captured PCG graphs are dropped.
void ensure_cases(Problem& P, int nb) {
  nb = std::max(nb, 1);
  const int pstride = max_partials(P) + 8;
  if (nb <= P.nb && pstride <= P.pstride) return;
  sync(P);
  drop_graphs(P);
  nb = std::max(nb, P.nb);
  P.pstride = std::max(pstride, P.pstride);
  const bool mixed = P.precision == TOPOHIP_PREC_MIXED;
  const size_t nd = static_cast<size_t>(P.g.ndof()) * nb;
  P.b.ensure(nd); P.x.ensure(nd); P.r.ensure(nd); P.p.ensure(nd); P.p2.ensure(nd); P.rt.ensure(nd);
  P.u.ensure(nd); P.Ap64.ensure(nd); P.z64.ensure(nd);
  if (mixed) { P.r32.ensure(nd); P.z32.ensure(nd); }
  const bool cheb = P.smoother == 1;
  for (size_t l = 0; l < P.L.size(); ++l) {
    Level& lv = *P.L[l];
    const size_t ndl = static_cast<size_t>(lv.g.ndof()) * nb;
    const bool smoothed = P.mg_plan_ok && l + 1 < P.L.size();
    if (mixed) {
      lv.x32.ensure(ndl);
      lv.t32.ensure(ndl);
      if (l > 0) lv.b32.ensure(ndl);
      if (cheb && smoothed) lv.d32.ensure(ndl);
    }
    if (!mixed || P.mg64_alloc) {
      lv.x64.ensure(ndl);
      lv.t64.ensure(ndl);
      if (l > 0) lv.b64.ensure(ndl);
      if (cheb && smoothed) lv.d64.ensure(ndl);
    }
    if (cheb && smoothed) lv.cheb.ensure(static_cast<size_t>(2 * P.cheb_degree + 2));
  }
  P.st.ensure(static_cast<size_t>(nb));
  P.scal.ensure(static_cast<size_t>(std::max(8, nb)));
  if (nb > P.nb || !P.st_h) {
    std::lock_guard<std::recursive_mutex> lk(device_api_mutex());
    if (P.st_h) (void)hipHostFree(P.st_h);
    if (P.scal_h) (void)hipHostFree(P.scal_h);
    P.st_h = nullptr;
    P.scal_h = nullptr;
    TOPOHIP_CHECK(hipHostMalloc(&P.st_h, sizeof(CgState) * nb));
    TOPOHIP_CHECK(hipHostMalloc(&P.scal_h, sizeof(double) * std::max(8, nb)));
  }
  P.partials.ensure(static_cast<size_t>(nb) * P.pstride);
  P.partials2.ensure(static_cast<size_t>(nb) * P.pstride);
  P.rzpart.ensure(static_cast<size_t>(nb) * P.pstride);
  if (P.mg_plan_ok) P.cr.reserve_cases(nb);
  P.nb = nb;
}

// FP64 V-cycle data of the mixed form, built on demand for the FP64 retry of a
// mixed MG-PCG that did not converge (the reference retries its FP32 V-cycle
// in FP64 before the Jacobi-CG fallback, core/src/fea/multigrid.cpp:2098-2319).
void ensure_mg64(Problem& P) {
  if (P.precision != TOPOHIP_PREC_MIXED || !P.mg_plan_ok) return;
  if (!P.mg64_alloc) {
    for (size_t l = 1; l + 1 < P.L.size(); ++l)
      P.L[l]->K64.ensure(static_cast<size_t>(P.L[l]->g.nodes()) * 126);
    P.mg64_alloc = true;
    const int nb = P.nb;
    P.nb = 0;  // re-run the capacity pass so the FP64 level vectors exist
    ensure_cases(P, nb);
  }
  if (!P.mg64_ready) {
    for (size_t l = 1; l + 1 < P.L.size(); ++l) {
      Level& c = *P.L[l];
      if (l == 1 && P.l1_direct) {  // (its special matrices are current: same hierarchy)
        launch_stencil_l1<double>(c.g, P.L[0]->g, P.E64.p, P.dS8.p, c.K1sp.p, c.spmap.p,
                                  c.mask.p, c.K64.p, c.dinv64.p, P.s);
        continue;
      }
      launch_diag_dense<double>(c.g, c.Kem.p, c.mask.p, c.dinv64.p, P.s);
      launch_assemble_stencil<double>(c.g, c.Kem.p, c.K64.p, P.s);
    }
    P.mg64_ready = true;
  }
}

void set_smoother(Problem& P, int kind, int degree) {
  if (kind != 0 && kind != 1) throw Error(TOPOHIP_EINVAL, "smoother: kind must be 0 or 1");
  if (degree < 1 || degree > 8) throw Error(TOPOHIP_EINVAL, "smoother: degree must be 1..8");
  sync(P);
  drop_graphs(P);  // the captured iteration bodies smooth the old way
  P.smoother = kind;
  P.cheb_degree = kind == 1 ? degree : 1;
  const int nb = P.nb;
  P.nb = 0;  // re-run the capacity pass: direction vectors and coefficient slots
  ensure_cases(P, nb);
}

std::unique_ptr<Problem> create_problem(int nx, int ny, int nz, double h, double nu,
                                        const uint8_t* solid, int n_bc, const int32_t* bc_node,
                                        const int32_t* bc_comp, const double* bc_val,
                                        int precision, int device) {
  if (nx < 1 || ny < 1 || nz < 1) throw Error(TOPOHIP_EINVAL, "grid dimensions must be >= 1");
  if (precision != TOPOHIP_PREC_MIXED && precision != TOPOHIP_PREC_F64)
    throw Error(TOPOHIP_EINVAL, "unknown precision");
  auto P = std::make_unique<Problem>();
  P->g = make_geom(nx, ny, nz);
  P->h = h;
  P->nu = nu;
  P->precision = precision;
  element_constants(*P);  // validates h and nu
  // Host-side validation and masks first: no HIP call before the inputs are
  // known to be good.
  const long long nv = P->g.elems(), nn = P->g.nodes();
  P->all_solid = solid == nullptr;
  if (solid) {
    P->solid_h.assign(solid, solid + nv);
    for (auto& v : P->solid_h) v = v ? 1 : 0;
  }
  P->n_solid = 0;
  P->touched_h.assign(static_cast<size_t>(nn), 0);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const long long e = (static_cast<long long>(k) * ny + j) * nx + i;
        if (solid && !P->solid_h[e]) continue;
        ++P->n_solid;
        for (int m = 0; m < 8; ++m)
          P->touched_h[(static_cast<long long>(k + (m >> 2)) * P->g.nny + j + ((m >> 1) & 1)) *
                           P->g.nnx + i + (m & 1)] = 1;
      }
  P->fixed_h.assign(static_cast<size_t>(nn), 0);
  P->up_h.assign(static_cast<size_t>(3 * nn), 0.0);
  for (int i = 0; i < n_bc; ++i) {
    if (bc_node[i] < 0 || bc_node[i] >= nn || bc_comp[i] < 0 || bc_comp[i] > 2)
      throw Error(TOPOHIP_EINVAL, "fea_solve_mgcg_matfree: BC index out of range");
    P->fixed_h[bc_node[i]] |= static_cast<uint8_t>(1 << bc_comp[i]);
    P->up_h[3 * static_cast<long long>(bc_node[i]) + bc_comp[i]] = bc_val[i];
    if (bc_val[i] != 0.0) P->nonzero_bc = true;
  }

  TOPOHIP_CHECK(hipSetDevice(device));
  P->device = device;
  TOPOHIP_CHECK(hipDeviceGetAttribute(&P->cus, hipDeviceAttributeMultiprocessorCount, device));
  P->scratch = noalloc_scratch(device);  // set-up of the allocation-free entry point
  {
    std::lock_guard<std::recursive_mutex> lk(device_api_mutex());
    TOPOHIP_CHECK(hipStreamCreateWithFlags(&P->s, hipStreamNonBlocking));
  }
  if (solid) {
    P->solid.ensure(static_cast<size_t>(nv));
    h2d(P->solid.p, P->solid_h.data(), static_cast<size_t>(nv), P->s);
  }
  P->up.ensure(static_cast<size_t>(3 * nn));
  h2d(P->up.p, P->up_h.data(), P->up_h.size(), P->s);
  P->dKb.ensure(576);
  h2d(P->dKb.p, P->Kb, 576, P->s);
  P->dS8.ensure(8 * 576);
  h2d(P->dS8.p, P->S8, 8 * 576, P->s);
  P->dT64.ensure(64 * 576);
  h2d(P->dT64.p, P->T64.data(), 64 * 576, P->s);
  P->dkdiag.ensure(24);
  h2d(P->dkdiag.p, P->kdiag, 24, P->s);
  P->E64.ensure(static_cast<size_t>(nv));
  P->E32.ensure(static_cast<size_t>(nv));
  {
    std::lock_guard<std::recursive_mutex> lk(device_api_mutex());
    TOPOHIP_CHECK(hipHostMalloc(&P->oc_h, sizeof(OcState)));
    TOPOHIP_CHECK(hipHostMalloc(&P->mma_h, sizeof(MmaSearch)));
  }
  P->oc.ensure(1);
  P->mma_st.ensure(1);
  plan_levels(*P);
  alloc_levels(*P);
  ensure_cases(*P, 1);
  autotune_ku(*P);
  ensure_cases(*P, 1);  // partial sums sized for the tuned tile
  sync(*P);
  return P;
}

// Void-DOF gate of mf_build_reduced (core/src/fea/matfree.cpp:782-814): every
// free DOF void -> singular; a load on an untouched free DOF -> no equilibrium.
void void_gate(const Problem& P, const std::vector<double>& rhs_full) {
  const long long nn = P.g.nodes();
  long long nfree = 0, ntouched_free = 0;
  double rmax = 0.0;
  for (long long n = 0; n < nn; ++n)
    for (int c = 0; c < 3; ++c) {
      if ((P.fixed_h[n] >> c) & 1) continue;
      ++nfree;
      if (P.touched_h[n]) ++ntouched_free;
      rmax = std::max(rmax, std::fabs(rhs_full[3 * n + c]));
    }
  if (nfree > 0 && ntouched_free == 0)
    throw Error(TOPOHIP_EUNDERCONSTRAINED,
                "fea_solve_mgcg_matfree: singular system (no stiffness — every free DOF is void)");
  for (long long n = 0; n < nn; ++n) {
    if (P.touched_h[n]) continue;
    for (int c = 0; c < 3; ++c)
      if (!((P.fixed_h[n] >> c) & 1) && std::fabs(rhs_full[3 * n + c]) > 1e-9 * rmax)
        throw Error(TOPOHIP_EUNDERCONSTRAINED,
                    "fea_solve_mgcg_matfree: under-constrained system (load applied to a void DOF "
                    "with no stiffness — no equilibrium possible)");
  }
}

void set_loads(Problem& P, int n_rhs, const int32_t* counts, const int32_t* node,
               const int32_t* comp, const double* val) {
  if (n_rhs < 1) throw Error(TOPOHIP_EINVAL, "n_rhs must be >= 1");
  const long long nn = P.g.nodes();
  const long long nd = 3 * nn;
  P.f_h.assign(static_cast<size_t>(n_rhs), std::vector<double>(static_cast<size_t>(nd), 0.0));
  long long q = 0;
  for (int r = 0; r < n_rhs; ++r) {
    if (counts[r] < 0) throw Error(TOPOHIP_EINVAL, "negative load count");
    for (int i = 0; i < counts[r]; ++i, ++q) {
      if (node[q] < 0 || node[q] >= nn || comp[q] < 0 || comp[q] > 2)
        throw Error(TOPOHIP_EINVAL, "fea_solve_mgcg_matfree: load index out of range");
      P.f_h[r][3 * static_cast<long long>(node[q]) + comp[q]] += val[q];
    }
  }
  // With zero prescribed values the gate and the masked RHS depend only on the
  // loads: check and upload them once here.
  if (!P.nonzero_bc)
    for (int r = 0; r < n_rhs; ++r) void_gate(P, P.f_h[r]);
  P.nrhs = n_rhs;
  ensure_cases(P, n_rhs);
  P.f.ensure(static_cast<size_t>(n_rhs * nd));
  const Level& f0 = *P.L[0];
  for (int r = 0; r < n_rhs; ++r) {
    std::vector<double> bm = P.f_h[r];
    if (!P.nonzero_bc)
      for (long long n = 0; n < nn; ++n)
        for (int c = 0; c < 3; ++c)
          if (!((f0.mask_h[n] >> c) & 1)) bm[3 * n + c] = 0.0;
    h2d(P.f.p + r * nd, bm.data(), static_cast<size_t>(nd), P.s);
  }
  sync(P);
}

// lambda_max(D^-1 A) of every smoothed level by kChebPowerSteps power steps
// (the level's x / t work vectors hold v and A v) and the Chebyshev
// coefficients derived from it, all on the device: captured PCG graphs read
// them from there, so a rebuilt hierarchy needs no new capture.
template <typename T>
void cheb_estimate(Problem& P) {
  for (size_t l = 0; l + 1 < P.L.size(); ++l) {
    Level& lv = *P.L[l];
    const long long n = lv.g.ndof();
    T* v = Xv<T>(lv);
    T* y = Tv<T>(lv);
    const T* dinv = Dinv<T>(lv);
    auto apply = [&]() {
      if (l == 0) {
        KuOp<T> op = fine_op<T>(P, nullptr, 1);
        op.x = v;
        op.y = y;
        op.epi = EPI_AX;
        launch_ku_fine<T>(op, P.s);
      } else {
        launch_stencil_apply<T>(lv.g, Kapp<T>(lv), lv.mask.p, v, y, nullptr, nullptr, 0.0, EPI_AX,
                                nullptr, P.s, 1);
      }
    };
    launch_cheb_power<T>(0, v, nullptr, dinv, n, P.partials.p, nullptr, 0, 0.0, 0.0, P.s);
    for (int k = 0; k < kChebPowerSteps; ++k) {
      apply();
      if (k + 1 < kChebPowerSteps)
        launch_cheb_power<T>(1, v, y, dinv, n, P.partials.p, nullptr, 0, 0.0, 0.0, P.s);
    }
    launch_cheb_power<T>(2, v, y, dinv, n, P.partials.p, lv.cheb.p, P.cheb_degree, kChebLow,
                         kChebHigh, P.s);
  }
}

void build_hierarchy(Problem& P, Timers& t) {
  const double t0 = now_ms();
  // re-arm check, once per solve call before its build (a block solve counts
  // once and re-arms all its load cases together)
  t_mg_rearmed = false;
  if (t_mg_latched && t_mg_rearm_period > 0 && ++t_mg_latched_solves >= t_mg_rearm_period) {
    t_mg_latched = false;
    t_mg_latched_solves = 0;
    t_mg_stagnations = kMgLatchThreshold - 1;
    t_mg_rearmed = true;
    ++t_mg_rearm_attempts;
  }
  P.mg64_ready = false;
  const long long nv = P.g.elems();
  const bool mixed = P.precision == TOPOHIP_PREC_MIXED;
  launch_convert<double, float>(P.E64.p, P.E32.p, nv, P.s);
  Level& f = *P.L[0];
  launch_diag_fine<double>(f.g, P.E64.p, P.dkdiag.p, f.mask.p, f.dinv64.p, P.s);
  if (mixed) launch_diag_fine<float>(f.g, P.E64.p, P.dkdiag.p, f.mask.p, f.dinv32.p, P.s);
  // (a latched thread skips the levels: the next solve goes straight to Jacobi-CG)
  if (P.mg_plan_ok && !t_mg_latched) {
    for (size_t l = 1; l < P.L.size(); ++l) {
      Level& c = *P.L[l];
      Level& pf = *P.L[l - 1];
      if (l == 1 && P.l1_direct) {
        // stencil and diagonal straight from the fine moduli
        launch_galerkin_fine_compact(c.g, pf.g, P.E64.p, pf.mask.p, c.mask.p, P.dS8.p, P.dKb.p,
                                     c.special.p, c.n_special, c.K1sp.p, P.s);
        if (mixed)
          launch_stencil_l1<float>(c.g, pf.g, P.E64.p, P.dS8.p, c.K1sp.p, c.spmap.p, c.mask.p,
                                   c.K32.p, c.dinv32.p, P.s);
        else
          launch_stencil_l1<double>(c.g, pf.g, P.E64.p, P.dS8.p, c.K1sp.p, c.spmap.p, c.mask.p,
                                    c.K64.p, c.dinv64.p, P.s);
        continue;
      }
      if (l == 1) {
        launch_galerkin_fine(c.g, pf.g, P.E64.p, pf.mask.p, c.mask.p, P.dS8.p, P.dKb.p,
                             c.special.p, c.n_special, c.Kem.p, P.s);
      } else if (l == 2 && P.l1_direct) {
        const Level& f0 = *P.L[0];
        launch_galerkin_l2(c.g, pf.g, f0.g, P.E64.p, P.dT64.p, P.dS8.p, pf.K1sp.p, pf.spmap.p,
                           c.mask.p, c.special.p, c.n_special, c.Kem.p, P.s);
      } else {
        launch_galerkin_coarse(c.g, pf.g, pf.Kem.p, c.mask.p, c.Kem.p, P.s);
      }
      const bool coarsest = l + 1 == P.L.size();
      if (!coarsest) {
        if (mixed) {
          launch_diag_dense<float>(c.g, c.Kem.p, c.mask.p, c.dinv32.p, P.s);
          launch_assemble_stencil<float>(c.g, c.Kem.p, c.K32.p, P.s);
        } else {
          launch_diag_dense<double>(c.g, c.Kem.p, c.mask.p, c.dinv64.p, P.s);
          launch_assemble_stencil<double>(c.g, c.Kem.p, c.K64.p, P.s);
        }
      }
    }
    Level& c = *P.L.back();
    P.cr.build(c.Kem.p, c.mask.p, P.s);
    if (P.smoother == 1) {
      if (mixed) cheb_estimate<float>(P);
      else cheb_estimate<double>(P);
    }
  }
  sync(P);
  t.mg_build += now_ms() - t0;
}

void solve_cases(Problem& P, int nc, double tol, int max_it, const double* d_guess,
                 topohip_cginfo* info) {
  const double t0 = now_ms();
  const long long nn = P.g.nodes(), nd = 3 * nn;
  if (nc < 1 || nc > P.nrhs) throw Error(TOPOHIP_EINVAL, "no load case set");
  if (d_guess && nc != 1) throw Error(TOPOHIP_EINTERNAL, "warm start needs a single load case");
  ensure_cases(P, nc);
  P.u_valid = false;
  Level& f = *P.L[0];
  if (P.nonzero_bc) {
    // rhs = f - K(E) up depends on the moduli: form it on the device, check
    // the void gate, and mask it (core/src/fea/matfree.cpp:745-834).
    KuOp<double> op;
    op.g = P.g;
    op.E = P.E64.p;
    op.mg = P.mg64;
    op.tz = P.tz64;
    op.x = P.up.p;
    op.y = P.rt.p;
    op.epi = EPI_AX;
    launch_ku_fine<double>(op, P.s);
    std::vector<double> kup(static_cast<size_t>(nd));
    d2h(kup.data(), P.rt.p, static_cast<size_t>(nd), P.s);
    for (int r = 0; r < nc; ++r) {
      std::vector<double> rhs_full = P.f_h[r];
      for (long long d = 0; d < nd; ++d) rhs_full[d] -= kup[d];
      void_gate(P, rhs_full);
      for (long long n = 0; n < nn; ++n)
        for (int c = 0; c < 3; ++c)
          if (!((f.mask_h[n] >> c) & 1)) rhs_full[3 * n + c] = 0.0;
      h2d(P.b.p + r * nd, rhs_full.data(), static_cast<size_t>(nd), P.s);
    }
  } else {
    // Zero prescribed values: the masked RHS were formed and gated by set_loads.
    TOPOHIP_CHECK(hipMemcpyAsync(P.b.p, P.f.p, nc * nd * sizeof(double), hipMemcpyDeviceToDevice,
                                 P.s));
  }
  const long long ng = f.nact;
  const int cap = max_it > 0 ? max_it : static_cast<int>(std::max<long long>(1000, 2 * ng));
  std::vector<topohip_cginfo> d(static_cast<size_t>(nc));
  for (auto& di : d) di.converged = 1;
  P.replacements = 0;
  P.deadline_hit = false;
  const long long mv0 = P.matvecs;
  // SolverDeadlineExceeded: the solve was abandoned at the deadline with the
  // iterations and residual it had reached (core/include/topopt/fea.hpp:455-469).
  auto check_deadline = [&](const std::vector<int>& it, const std::vector<double>& res,
                            const bool* ok) {
    if (!P.deadline_hit) return;
    for (int r = 0; r < nc; ++r) {
      if (ok[r]) continue;
      d[r].converged = 0;
      d[r].iterations = it[r];
      d[r].residual = res[r];
      d[r].matvecs = P.matvecs - mv0;
      d[r].t_total_ms = now_ms() - t0;
    }
    if (info) std::copy(d.begin(), d.end(), info);
    throw Error(TOPOHIP_EDEADLINE, "fea_solve_mgcg_matfree: the solve deadline passed before CG "
                                   "reached the requested tolerance");
  };
  if (ng > 0) {
    bool warm = false;
    std::vector<double> gh;
    if (d_guess) {
      // masked copy of the guess: x = guess on active DOFs; a non-finite
      // entry makes the solve cold, as the reference's initial_guess handling.
      gh.resize(static_cast<size_t>(nd));
      d2h(gh.data(), d_guess, static_cast<size_t>(nd), P.s);
      warm = true;
      for (long long n = 0; n < nn && warm; ++n)
        for (int c = 0; c < 3; ++c) {
          if (!((f.mask_h[n] >> c) & 1)) { gh[3 * n + c] = 0.0; continue; }
          if (!std::isfinite(gh[3 * n + c])) { warm = false; break; }
        }
      if (warm) h2d(P.x.p, gh.data(), static_cast<size_t>(nd), P.s);
    }
    if (!warm) launch_fill<double>(P.x.p, 0.0, nd * nc, P.s);
    std::vector<int> it(static_cast<size_t>(nc), 0);
    std::vector<double> res(static_cast<size_t>(nc), 0.0);
    std::unique_ptr<bool[]> ok(new bool[nc]());
    const bool try_mg = P.mg_plan_ok && !t_mg_latched;
    if (try_mg) {
      const double tm = now_ms();
      const Prec pc = P.precision == TOPOHIP_PREC_MIXED ? Prec::MG32 : Prec::MG64;
      pcg(P, pc, nc, tol, std::min(cap, kMgIterBudget), warm, nullptr, it.data(), res.data(),
          ok.get());
      check_deadline(it, res, ok.get());
      const double tmg = now_ms() - tm;
      for (int r = 0; r < nc; ++r) {
        d[r].hier_built = 1;
        d[r].mg_cycles_attempted = it[r];
        d[r].t_mg_ms = tmg;
        if (ok[r]) {
          d[r].iterations = it[r];
          d[r].residual = res[r];
          d[r].used_multigrid = 1;
          d[r].mg_levels = static_cast<int>(P.L.size());
          d[r].coarse_dofs = static_cast<int>(P.L.back()->nact);
        }
      }
    }
    bool all_ok = true;
    for (int r = 0; r < nc; ++r) all_ok = all_ok && ok[r];
    if (!all_ok && try_mg && P.precision == TOPOHIP_PREC_MIXED) {
      // FP64 V-cycle retry of the cases the mixed solve did not finish.
      ensure_mg64(P);
      const double tm = now_ms();
      std::unique_ptr<bool[]> skip(new bool[nc]);
      for (int r = 0; r < nc; ++r) {
        skip[r] = ok[r];
        if (ok[r]) continue;
        if (warm) h2d(P.x.p, gh.data(), static_cast<size_t>(nd), P.s);
        else launch_fill<double>(P.x.p + r * nd, 0.0, nd, P.s);
      }
      std::vector<int> it64(static_cast<size_t>(nc), 0);
      pcg(P, Prec::MG64, nc, tol, std::min(cap, kMgIterBudget), warm, skip.get(), it64.data(),
          res.data(), ok.get());
      check_deadline(it64, res, ok.get());
      const double tmg = now_ms() - tm;
      all_ok = true;
      for (int r = 0; r < nc; ++r) {
        all_ok = all_ok && ok[r];
        if (skip[r]) continue;
        d[r].mg_cycles_attempted += it64[r];
        d[r].t_mg_ms += tmg;
        it[r] = it64[r];
        if (ok[r]) {
          d[r].iterations = it64[r];
          d[r].residual = res[r];
          d[r].used_multigrid = 1;
          d[r].mg_levels = static_cast<int>(P.L.size());
          d[r].coarse_dofs = static_cast<int>(P.L.back()->nact);
        }
      }
    }
    // Latch bookkeeping when a hierarchy was built and attempted: each load
    // case counts as the reference's one-case solve would, in case order (a
    // converging case clears the count, a stagnating one advances it); the
    // latch then holds from the next solve on.
    if (try_mg) {
      for (int r = 0; r < nc; ++r) {
        if (ok[r]) {
          t_mg_stagnations = 0;
        } else if (++t_mg_stagnations >= kMgLatchThreshold) {
          t_mg_latched = true;
        }
      }
      // a re-armed attempt carries when every case converged with multigrid
      if (t_mg_rearmed && all_ok) ++t_mg_rearm_carries;
      t_mg_rearmed = false;
    }
    if (!all_ok) {
      // Exact Jacobi-PCG fallback (core/src/fea/matfree.cpp:841-1076), FP64, for
      // the cases the multigrid solve did not finish; the others keep their x.
      const double tc = now_ms();
      std::unique_ptr<bool[]> skip(new bool[nc]);
      for (int r = 0; r < nc; ++r) {
        skip[r] = ok[r];
        if (ok[r]) continue;
        if (warm) h2d(P.x.p, gh.data(), static_cast<size_t>(nd), P.s);
        else launch_fill<double>(P.x.p + r * nd, 0.0, nd, P.s);
      }
      pcg(P, Prec::JACOBI, nc, tol, cap, warm, skip.get(), it.data(), res.data(), ok.get());
      check_deadline(it, res, ok.get());
      const double tcg = now_ms() - tc;
      for (int r = 0; r < nc; ++r) {
        if (skip[r]) continue;
        d[r].iterations = it[r];
        d[r].residual = res[r];
        d[r].converged = ok[r];
        d[r].used_multigrid = 0;
        d[r].mg_levels = 0;
        d[r].t_cg_ms = tcg;
      }
      for (int r = 0; r < nc; ++r)
        if (!ok[r]) {
          d[r].matvecs = P.matvecs - mv0;
          d[r].residual_replacements = P.replacements;
          d[r].t_total_ms = now_ms() - t0;
          if (info) std::copy(d.begin(), d.end(), info);
          throw Error(TOPOHIP_ENONCONVERGENCE,
                      "fea_solve_mgcg_matfree: CG did not reach the requested tolerance within "
                      "max_iterations");
        }
    }
  } else {
    launch_fill<double>(P.x.p, 0.0, nd * nc, P.s);
  }
  for (int r = 0; r < nc; ++r) launch_add(P.x.p + r * nd, P.up.p, P.u.p + r * nd, nd, P.s);
  sync(P);
  P.u_valid = true;
  for (auto& di : d) {
    di.matvecs = P.matvecs - mv0;
    di.residual_replacements = P.replacements;
    di.t_total_ms = now_ms() - t0;
  }
  if (info) std::copy(d.begin(), d.end(), info);
}

// Density filter of the given radius over the design set `dset` (device, one
// byte per voxel, null = every voxel): stencil, Hs and dv = H^T(1/Hs).
std::unique_ptr<FilterCache> make_filter(Problem& P, double radius, const uint8_t* dset) {
  if (!(radius > 0.0)) throw Error(TOPOHIP_EINVAL, "make_density_filter: radius must be > 0");
  auto fc = std::make_unique<FilterCache>();
  fc->radius = radius;
  const int R = static_cast<int>(std::ceil(radius));
  std::vector<int> off;
  std::vector<double> w;
  // Neighbour order of make_density_filter (core/src/simp/simp.cpp:862-880).
  for (int dk = -R; dk <= R; ++dk)
    for (int dj = -R; dj <= R; ++dj)
      for (int di = -R; di <= R; ++di) {
        const double dist = std::sqrt(static_cast<double>(di * di + dj * dj + dk * dk));
        const double hw = radius - dist;
        if (hw <= 0.0) continue;
        off.push_back(di);
        off.push_back(dj);
        off.push_back(dk);
        w.push_back(hw);
      }
  fc->off.ensure(off.size());
  fc->w.ensure(w.size());
  h2d(fc->off.p, off.data(), off.size(), P.s);
  h2d(fc->w.p, w.data(), w.size(), P.s);
  fc->st.n = static_cast<int>(w.size());
  fc->st.off = fc->off.p;
  fc->st.w = fc->w.p;
  const long long nv = P.g.elems();
  fc->hs.ensure(static_cast<size_t>(nv));
  fc->dv.ensure(static_cast<size_t>(nv));
  const uint8_t* sol = dset;
  launch_filter_hs(P.g, fc->st, sol, fc->hs.p, P.s);
  // dv = H^T (ones / Hs) over design voxels (ones on non-design voxels are
  // ignored by the adjoint, which skips them).
  DBuf<double> ones;
  ones.ensure(static_cast<size_t>(nv));
  launch_fill<double>(ones.p, 1.0, nv, P.s);
  launch_filter(P.g, fc->st, sol, fc->hs.p, ones.p, fc->dv.p, 1, P.s);
  sync(P);
  return fc;
}

FilterCache& filter_for(Problem& P, double radius) {
  if (!(radius > 0.0)) throw Error(TOPOHIP_EINVAL, "make_density_filter: radius must be > 0");
  for (auto& fc : P.filters)
    if (fc->radius == radius) return *fc;
  P.filters.push_back(make_filter(P, radius, P.all_solid ? nullptr : P.solid.p));
  return *P.filters.back();
}

double simp_compliance_dev(Problem& P, const double* d_xphys, double E0, double penalty,
                           double rho_min, double tol, int max_it, topohip_cginfo* info,
                           const double* d_guess) {
  if (!(E0 > 0.0)) throw Error(TOPOHIP_EINVAL, "simp: youngs_modulus (E0) must be > 0");
  if (!(penalty > 0.0)) throw Error(TOPOHIP_EINVAL, "simp: penalty (p) must be > 0");
  if (!(rho_min > 0.0 && rho_min <= 1.0))
    throw Error(TOPOHIP_EINVAL, "simp: density_min must be in (0, 1]");
  if (P.nrhs < 1) throw Error(TOPOHIP_EINVAL, "no load case set");
  const long long nv = P.g.elems();
  const uint8_t* sol = P.all_solid ? nullptr : P.solid.p;
  launch_youngs(d_xphys, sol, nv, E0, penalty, rho_min, P.E64.p, nullptr, P.s);
  Timers t;
  build_hierarchy(P, t);
  P.dc.ensure(static_cast<size_t>(nv));
  const int sb = sens_blocks(P.g);
  const long long nd = P.g.ndof();
  std::vector<topohip_cginfo> ci(static_cast<size_t>(P.nrhs));
  if (d_guess && P.nrhs != 1)
    throw Error(TOPOHIP_EINVAL, "simp_compliance: an initial guess needs a single load case");
  solve_cases(P, P.nrhs, tol, max_it, d_guess, ci.data());
  // Multi-load objective: c = sum_r c_r, dc = sum_r dc_r.
  double c = 0.0;
  topohip_cginfo agg = ci[0];
  for (int r = 0; r < P.nrhs; ++r) {
    if (r > 0) {
      agg.iterations = std::max(agg.iterations, ci[r].iterations);
      agg.mg_cycles_attempted = std::max(agg.mg_cycles_attempted, ci[r].mg_cycles_attempted);
      agg.residual = std::max(agg.residual, ci[r].residual);
      agg.converged = agg.converged && ci[r].converged;
      agg.used_multigrid = agg.used_multigrid && ci[r].used_multigrid;
    }
    launch_sensitivity(P.g, P.u.p + r * nd, d_xphys, sol, P.mc64, E0, penalty, rho_min, P.dc.p,
                       r > 0, P.partials.p, sb, P.s);
    launch_sum_partials(P.partials.p, sb, P.scal.p, P.s);
    d2h(P.scal_h, P.scal.p, 1, P.s);
    c += P.scal_h[0];
  }
  agg.t_mg_build_ms = t.mg_build;
  if (info) *info = agg;
  return c;
}

// Work memory of one OC search: device state and partial sums, their pinned
// host read-back slots, and the filtered sensitivities.
struct OcWork {
  OcState* st;
  OcState* st_h;
  double* lam;   // kLamPass x oc_blocks
  double* chg;   // oc_blocks
  double* scal;  // >= 1
  double* scal_h;
  double* dcf;   // voxels
};

void oc_search(Problem& P, FilterCache& fc, const uint8_t* dset, double n_design, bool keep,
               const double* d_x, const double* d_dc, double volfrac, double move,
               double rho_min, double* d_xnew, uint8_t* d_active, int* bisections,
               double* lambda, double* change, const OcWork& w) {
  if (!(volfrac > 0.0 && volfrac <= 1.0))
    throw Error(TOPOHIP_EINVAL, "oc_update: volume_fraction must be in (0, 1]");
  if (!(move > 0.0)) throw Error(TOPOHIP_EINVAL, "oc_update: move must be > 0");
  if (!(rho_min > 0.0 && rho_min <= 1.0))
    throw Error(TOPOHIP_EINVAL, "oc_update: density_min must be in (0, 1]");
  const long long nv = P.g.elems();
  const uint8_t* sol = dset;
  launch_filter(P.g, fc.st, sol, fc.hs.p, d_dc, w.dcf, 1, P.s);
  const double target = volfrac * n_design;
  launch_oc_init_multi(w.st, target, P.s);
  const int ob = oc_blocks(nv);
  // Each pass tries up to 7 lambdas (~100 bracket + 100 bisection trials at
  // most). The first round launches as many passes as the previous search
  // needed (consecutive design iterations bisect to similar depths), later
  // rounds 3; passes after the search ends are no-ops, each round ends with
  // a read-back of the state.
  int passes = P.oc_pass_hint > 0 ? P.oc_pass_hint : 6;
  for (int round = 0;; ++round) {
    if (round > 80) throw Error(TOPOHIP_EINTERNAL, "oc_update: lambda search did not finish");
    for (int i = 0; i < passes; ++i) {
      launch_oc_trial_multi(d_x, w.dcf, fc.dv.p, sol, nv, move, rho_min, w.st, w.lam, ob, P.s);
      launch_oc_decide_multi(w.st, w.lam, ob, P.s);
    }
    d2h(w.st_h, w.st, 1, P.s);
    if (w.st_h->phase == 2) break;
    passes = 3;
  }
  P.oc_pass_hint = w.st_h->passes;
  launch_oc_final(d_x, w.dcf, fc.dv.p, sol, nv, move, rho_min, w.st, d_xnew, d_active, w.chg, ob,
                  P.s, keep ? 1 : 0);
  launch_max_partials(w.chg, ob, w.scal, P.s);
  d2h(w.scal_h, w.scal, 1, P.s);
  if (bisections) *bisections = w.st_h->steps;
  if (lambda) *lambda = 0.5 * (w.st_h->l1 + w.st_h->l2);
  if (change) *change = w.scal_h[0];
  // Decision margin: the volume tested against the target closest to it, as a
  // fraction of the target, beside the rounding bound (n - 1) u of a sum of n
  // non-negative terms (the reference's sequential order; the device's block
  // tree is tighter). A gap above the bound decides the same in either order.
  P.oc_gap_rel = target > 0.0 ? w.st_h->gap / target : 0.0;
  P.oc_gap_min = std::min(P.oc_gap_min, P.oc_gap_rel);
  P.oc_gap_bound = std::max(0.0, n_design - 1.0) * std::ldexp(1.0, -53);
}

// One OC step (oc_update, core/src/simp/simp.cpp:948-1032) on the design set
// dset with fc built on the same set: oc_update (keep = false: voxels off the
// set get 0) or oc_update_masked (keep = true: they keep their x; budget
// volfrac * n_design). Work memory: the handle's own buffers.
void oc_update_core(Problem& P, FilterCache& fc, const uint8_t* dset, double n_design, bool keep,
                    const double* d_x, const double* d_dc, double volfrac, double move,
                    double rho_min, double* d_xnew, uint8_t* d_active, int* bisections,
                    double* lambda, double* change) {
  const long long nv = P.g.elems();
  P.dcf.ensure(static_cast<size_t>(nv));
  P.lam_partials.ensure(static_cast<size_t>(kLamPass) * oc_blocks(nv));
  const OcWork w{P.oc.p, P.oc_h, P.lam_partials.p, P.partials.p, P.scal.p, P.scal_h, P.dcf.p};
  oc_search(P, fc, dset, n_design, keep, d_x, d_dc, volfrac, move, rho_min, d_xnew, d_active,
            bisections, lambda, change, w);
}

void oc_update_dev(Problem& P, double radius, const double* d_x, const double* d_dc,
                   double volfrac, double move, double rho_min, double* d_xnew, uint8_t* d_active,
                   int* bisections, double* lambda, double* change) {
  FilterCache& fc = filter_for(P, radius);
  oc_update_core(P, fc, P.all_solid ? nullptr : P.solid.p, static_cast<double>(P.n_solid), false,
                 d_x, d_dc, volfrac, move, rho_min, d_xnew, d_active, bisections, lambda, change);
}

// oc_update on device buffers with no allocation: the work memory is the
// GPU's file-scope scratch array and the handle's pinned read-back slots, and
// the density filter of `radius` must already be on the handle (built at
// set-up by any filter or OC call with that radius).
void oc_update_noalloc(Problem& P, double radius, const double* d_x, const double* d_dc,
                       double volfrac, double move, double rho_min, double* d_xnew,
                       uint8_t* d_active, int* bisections, double* lambda) {
  const long long nv = P.g.elems();
  if (nv > kNoAllocMaxVoxels)
    throw Error(TOPOHIP_EINVAL, "oc_update_noalloc: " + std::to_string(nv) +
                                    " voxels exceed the allocation-free limit of " +
                                    std::to_string(kNoAllocMaxVoxels));
  if (!(radius > 0.0)) throw Error(TOPOHIP_EINVAL, "make_density_filter: radius must be > 0");
  FilterCache* fc = nullptr;
  for (auto& f : P.filters)
    if (f->radius == radius) fc = f.get();
  if (!fc)
    throw Error(TOPOHIP_EINVAL, "oc_update_noalloc: the density filter of this radius is not on "
                                "the handle; build it at set-up with topohip_filter or "
                                "topohip_oc_update at the same radius");
  if (!P.scratch) throw Error(TOPOHIP_EINTERNAL, "oc_update_noalloc: no scratch address");
  char* base = static_cast<char*>(P.scratch);
  const OcWork w{reinterpret_cast<OcState*>(base + kScrState), P.oc_h,
                 reinterpret_cast<double*>(base + kScrLam),
                 reinterpret_cast<double*>(base + kScrChange),
                 reinterpret_cast<double*>(base + kScrScal), P.scal_h,
                 reinterpret_cast<double*>(base + kScrDcf)};
  oc_search(P, *fc, P.all_solid ? nullptr : P.solid.p, static_cast<double>(P.n_solid), false, d_x,
            d_dc, volfrac, move, rho_min, d_xnew, d_active, bisections, lambda, nullptr, w);
  sync(P);
}

// mma_update over mma_update_joint (core/src/simp/simp.cpp:1433-1488,
// core/src/simp/mma_joint.cpp:18-118). The volume is the sum of the filtered
// design; the search decisions run on the device, read back once per 8 trials.
void mma_update_core(Problem& P, FilterCache& fc, const uint8_t* dset, double n_design,
                     int mma_iter, const double* d_x, const double* d_dc, double volfrac,
                     double move, double rho_min, double* d_xnew, double* lambda,
                     double* change, const double* d_xt, double beta, double eta) {
  if (!(volfrac > 0.0 && volfrac <= 1.0))
    throw Error(TOPOHIP_EINVAL, "mma_update: volume_fraction must be in (0, 1]");
  if (!(move > 0.0)) throw Error(TOPOHIP_EINVAL, "mma_update: move must be > 0");
  if (!(rho_min > 0.0 && rho_min < 1.0))
    throw Error(TOPOHIP_EINVAL, "mma_update: density_min must be in (0, 1)");
  if (mma_iter < 1) throw Error(TOPOHIP_EINVAL, "mma_update: mma_iter must be >= 1");
  const long long nv = P.g.elems();
  const size_t n = static_cast<size_t>(nv);
  for (auto* b : {&P.dcf, &P.mma_low, &P.mma_upp, &P.mma_x1, &P.mma_x2, &P.xcand}) b->ensure(n);
  if (mma_iter > 1 && mma_iter != P.mma_last + 1)
    throw Error(TOPOHIP_EINVAL, "mma_update: mma_iter must continue the run's state (1, 2, 3, ...)");
  P.mma_c.ensure(6 * n);
  MmaArgs a{P.mma_low.p, P.mma_upp.p, P.mma_c.p, P.mma_c.p + n, P.mma_c.p + 2 * n,
            P.mma_c.p + 3 * n, P.mma_c.p + 4 * n, P.mma_c.p + 5 * n};
  const int ob = oc_blocks(nv), vb = dot_blocks(nv);
  const double* dv = fc.dv.p;
  if (d_xt) {
    // projected: dproj and dc * dproj on the design set through the filter
    // transpose; the volume is the projected filtered design's (0 elsewhere)
    for (auto* b : {&P.dpr, &P.dcp, &P.dvf}) b->ensure(n);
    launch_proj_chain(d_xt, d_dc, dset, nv, beta, eta, P.dcp.p, P.dpr.p, P.s);
    launch_filter(P.g, fc.st, dset, fc.hs.p, P.dcp.p, P.dcf.p, 1, P.s);
    launch_filter(P.g, fc.st, dset, fc.hs.p, P.dpr.p, P.dvf.p, 1, P.s);
    launch_project(d_xt, P.xcand.p, dset, nv, beta, eta, P.s);
    dv = P.dvf.p;
  } else {
    launch_filter(P.g, fc.st, dset, fc.hs.p, d_dc, P.dcf.p, 1, P.s);
    launch_filter(P.g, fc.st, dset, fc.hs.p, d_x, P.xcand.p, 0, P.s);
  }
  launch_mma_scale(P.dcf.p, dset, nv, P.partials.p, ob, P.s);
  launch_max_partials(P.partials.p, ob, P.scal.p, P.s);
  launch_sum_vec<double>(P.xcand.p, nv, P.partials.p, vb, P.s);
  launch_sum_partials(P.partials.p, vb, P.scal.p + 1, P.s);
  launch_mma_coef(a, d_x, P.mma_x1.p, P.mma_x2.p, P.dcf.p, dv, dset, nv, mma_iter, move,
                  rho_min, P.scal.p, P.partials2.p, ob, P.s);
  launch_mma_init(P.mma_st.p, P.partials2.p, ob, P.scal.p + 1, volfrac * n_design, P.s);
  P.lam_partials.ensure(static_cast<size_t>(kLamPass) * ob);
  // each pass tries up to 7 lambdas: at most 1 + ~100 doublings + 100 bisections
  for (int round = 0;; ++round) {
    if (round > 40) throw Error(TOPOHIP_EINTERNAL, "mma_update: dual search did not finish");
    for (int i = 0; i < 6; ++i) {
      launch_mma_trial(a, dset, nv, P.mma_st.p, P.lam_partials.p, ob, P.s);
      launch_mma_decide(P.mma_st.p, P.lam_partials.p, ob, P.s);
    }
    d2h(P.mma_h, P.mma_st.p, 1, P.s);
    if (P.mma_h->phase == 3) break;
  }
  launch_mma_final(a, d_x, dset, nv, P.mma_st.p, d_xnew, P.partials.p, ob, P.s);
  launch_max_partials(P.partials.p, ob, P.scal.p + 2, P.s);
  // roll the history: xold2 <- xold1, xold1 <- x
  std::swap(P.mma_x1.p, P.mma_x2.p);
  std::swap(P.mma_x1.n, P.mma_x2.n);
  TOPOHIP_CHECK(hipMemcpyAsync(P.mma_x1.p, d_x, n * sizeof(double), hipMemcpyDeviceToDevice, P.s));
  d2h(P.scal_h, P.scal.p, 3, P.s);
  P.mma_last = mma_iter;
  if (lambda) *lambda = P.mma_h->lambda;
  if (change) *change = P.scal_h[2];
}

bool mma_objective_plateau(const double* c, int n, int window, double rel_tol, double min_drop,
                           int min_flat_windows) {
  if (window <= 0 || n < window + 1) return false;
  double best_now = c[0];
  for (int i = 1; i < n; ++i) best_now = std::min(best_now, c[i]);
  if (min_drop > 0.0 && !(best_now <= c[0] * (1.0 - min_drop))) {
    // the progress gate is closed: only the flatness escape may fire
    bool escape = false;
    const int span = min_flat_windows * window;
    if (min_flat_windows > 0 && n >= span + 1) {
      double best_before = c[0];
      for (int i = 1; i < n - span; ++i) best_before = std::min(best_before, c[i]);
      if (best_before > 0.0) escape = (best_before - best_now) / best_before < rel_tol;
    }
    if (!escape) return false;
  }
  double best_prev = c[0];
  for (int i = 1; i < n - window; ++i) best_prev = std::min(best_prev, c[i]);
  if (!(best_prev > 0.0)) return false;
  return (best_prev - best_now) / best_prev < rel_tol;
}

// Host-side design mask of the masked loop: eff (effective_design_mask,
// core/src/simp/simp.cpp:2610-2633: empty voxels -> FrozenVoid) and the design
// set (solid Active voxels). The handle's grid is the analysis grid: a
// FrozenVoid voxel must be empty in it (analysis_grid_for_mask, :2640-2645).
void design_mask_of(const Problem& P, const uint8_t* mask, std::vector<uint8_t>& eff,
                    std::vector<uint8_t>& dset, long long& n_design) {
  const long long nv = P.g.elems();
  eff.assign(static_cast<size_t>(nv), 0);
  dset.assign(static_cast<size_t>(nv), 0);
  n_design = 0;
  for (long long e = 0; e < nv; ++e) {
    const bool solid = P.all_solid || P.solid_h[e];
    if (mask[e] > 2) throw Error(TOPOHIP_EINVAL, "simp_optimize: mask value must be 0, 1 or 2");
    if (mask[e] == 2 && solid)
      throw Error(TOPOHIP_EINVAL,
                  "simp_optimize: a FrozenVoid voxel is solid in the problem grid (create the "
                  "problem on the analysis grid: FrozenVoid voxels empty)");
    eff[e] = solid ? mask[e] : 2;
    if (eff[e] == 0) { dset[e] = 1; ++n_design; }
  }
}

// oc_update_projected (core/src/simp/simp.cpp:1647-1716) over the design set
// dset with its filter fc; d_xt = filter_density(x). Masked (keep), it is
// oc_update_masked_projected (:2801-2855): dset the Active voxels, fc their
// Active-only filter, and the frozen voxels keep their x.
void oc_update_projected_core(Problem& P, FilterCache& fc, const uint8_t* dset, double n_design,
                              bool keep, const double* d_x, const double* d_xt, const double* d_dc,
                              double beta, double eta, double volfrac, double move,
                              double rho_min, double* d_xnew, double* change) {
  const long long nv = P.g.elems();
  for (auto* b : {&P.dpr, &P.dcp, &P.dcf, &P.dvf, &P.xcand, &P.xcf})
    b->ensure(static_cast<size_t>(nv));
  launch_proj_chain(d_xt, d_dc, dset, nv, beta, eta, P.dcp.p, P.dpr.p, P.s);
  launch_filter(P.g, fc.st, dset, fc.hs.p, P.dcp.p, P.dcf.p, 1, P.s);
  launch_filter(P.g, fc.st, dset, fc.hs.p, P.dpr.p, P.dvf.p, 1, P.s);
  launch_oc_init(P.oc.p, volfrac * n_design, P.s);
  const int ob = oc_blocks(nv);
  for (int round = 0;; ++round) {
    if (round > 40) throw Error(TOPOHIP_EINTERNAL, "oc_update: lambda search did not finish");
    for (int i = 0; i < 8; ++i) {
      launch_oc_cand_proj(d_x, P.dcf.p, P.dvf.p, dset, nv, move, rho_min, P.oc.p, P.xcand.p, P.s);
      launch_filter(P.g, fc.st, dset, fc.hs.p, P.xcand.p, P.xcf.p, 0, P.s);
      launch_proj_sum(P.xcf.p, dset, nv, beta, eta, P.oc.p, P.partials.p, ob, P.s);
      launch_oc_decide(P.oc.p, P.partials.p, ob, P.s);
    }
    d2h(P.oc_h, P.oc.p, 1, P.s);
    if (P.oc_h->phase == 2) break;
  }
  launch_oc_final_proj(d_x, P.dcf.p, P.dvf.p, dset, nv, move, rho_min, P.oc.p, d_xnew,
                       P.partials.p, ob, P.s, keep);
  launch_max_partials(P.partials.p, ob, P.scal.p, P.s);
  d2h(P.scal_h, P.scal.p, 1, P.s);
  if (change) *change = P.scal_h[0];
}

void optimize(Problem& P, const topohip_options& o, double* design_out, double* xphys_out,
              double* history, topohip_result* res, const uint8_t* mask,
              const std::vector<ProjStage>* proj, double eta, const MmaOpts* mma) {
  if (!(o.volfrac > 0.0 && o.volfrac <= 1.0))
    throw Error(TOPOHIP_EINVAL, "simp_optimize: volume_fraction must be in (0, 1]");
  if (!(o.move > 0.0)) throw Error(TOPOHIP_EINVAL, "simp_optimize: move must be > 0");
  if (o.max_iterations < 1) throw Error(TOPOHIP_EINVAL, "simp_optimize: max_iterations must be >= 1");
  if (proj && !proj->empty()) {
    if (!(eta > 0.0 && eta < 1.0)) throw Error(TOPOHIP_EINVAL, "heaviside_project: eta must be in (0, 1)");
    for (const ProjStage& ps : *proj)
      if (!(std::isfinite(ps.beta) && ps.beta > 0.0) || !(ps.move > 0.0) || ps.iterations < 1)
        throw Error(TOPOHIP_EINVAL, "simp_optimize: projection stage needs beta > 0, move > 0, "
                                    "iterations >= 1");
  }
  if (o.change_tol < 0.0) throw Error(TOPOHIP_EINVAL, "simp_optimize: change_tol must be >= 0");
  // validate_updater_options: projection stages run with the OC updater only.
  // Masked, the MMA update moves the Active voxels and returns 0 elsewhere
  // (mma_update_masked, core/src/simp/simp.cpp:2736-2783).
  if (mma && proj && !proj->empty())
    throw Error(TOPOHIP_EINVAL, "simp_optimize: projection stages need the OC updater");
  // SimpOptions::mma_projection (validate_mma_projection_options,
  // core/src/simp/simp.cpp:112-135)
  const bool mma_cont = mma && mma->projection;
  // validate_adaptive_move_options (core/src/simp/simp.cpp:140-163)
  if (mma && mma->adaptive) {
    if (!(mma->ad_grow > 1.0) || !std::isfinite(mma->ad_grow))
      throw Error(TOPOHIP_EINVAL, "simp_optimize: adaptive_move_grow must be finite and > 1");
    if (!(mma->ad_shrink > 0.0 && mma->ad_shrink < 1.0))
      throw Error(TOPOHIP_EINVAL, "simp_optimize: adaptive_move_shrink must be in (0, 1)");
    if (!(mma->ad_osc_lo >= 0.0 && mma->ad_osc_lo <= mma->ad_osc_hi && mma->ad_osc_hi <= 1.0))
      throw Error(TOPOHIP_EINVAL, "simp_optimize: adaptive_move_osc thresholds must satisfy "
                                  "0 <= lo <= hi <= 1");
    if (!(mma->ad_min > 0.0) || !(mma->ad_max >= mma->ad_min) || !(mma->ad_max < 1.0))
      throw Error(TOPOHIP_EINVAL, "simp_optimize: adaptive_move bounds must satisfy "
                                  "0 < min <= max < 1");
  }
  double adaptive_move_val = 0.0;
  if (mma_cont) {
    if (!(eta > 0.0 && eta < 1.0))
      throw Error(TOPOHIP_EINVAL, "simp_optimize: projection_eta must be in (0, 1)");
    if (!(mma->beta0 > 0.0) || !std::isfinite(mma->beta0))
      throw Error(TOPOHIP_EINVAL, "simp_optimize: mma_projection_beta0 must be finite and > 0");
    if (!(mma->beta_max >= mma->beta0) || !std::isfinite(mma->beta_max))
      throw Error(TOPOHIP_EINVAL, "simp_optimize: mma_projection_beta_max must be finite and "
                                  ">= mma_projection_beta0");
  }
  int mma_iter = 0;
  P.oc_gap_min = 1e300;
  std::vector<double> curve;  // compliance history for the MMA plateau stop
  std::vector<std::pair<double, int>> stages;
  if (o.pc_per_stage > 0) {
    if (!(o.pc_p0 > 0.0) || !(o.pc_p1 >= o.pc_p0) || !(o.pc_step > 0.0))
      throw Error(TOPOHIP_EINVAL, "penalty_continuation_ramp: bad arguments");
    const long long n = static_cast<long long>((o.pc_p1 - o.pc_p0) / o.pc_step + 1e-9);
    for (long long k = 0; k <= n; ++k)
      stages.emplace_back(o.pc_p0 + static_cast<double>(k) * o.pc_step, o.pc_per_stage);
  }
  auto penalty_at = [&](int it) {
    if (stages.empty()) return o.penalty;
    long long rem = it;
    for (auto& s : stages) {
      rem -= s.second;
      if (rem <= 0) return s.first;
    }
    return stages.back().first;
  };
  const double t0 = now_ms();
  const long long nv = P.g.elems();
  // Design set: the solid voxels (plain loop) or, with a design mask, the solid
  // Active voxels (masked loop, core/src/simp/simp.cpp:2989-3689), whose
  // filter admits only Active neighbours and whose OC update moves only them.
  const uint8_t* sol = P.all_solid ? nullptr : P.solid.p;
  FilterCache* fcp = nullptr;
  std::unique_ptr<FilterCache> masked_fc;
  DBuf<uint8_t> d_eff, d_dset;
  double nd = static_cast<double>(P.n_solid);
  std::vector<double> x0(static_cast<size_t>(nv), 0.0);
  if (mask) {
    std::vector<uint8_t> eff, dset;
    long long n_design = 0;
    design_mask_of(P, mask, eff, dset, n_design);
    d_eff.ensure(eff.size());
    d_dset.ensure(dset.size());
    h2d(d_eff.p, eff.data(), eff.size(), P.s);
    h2d(d_dset.p, dset.data(), dset.size(), P.s);
    sol = d_dset.p;
    nd = static_cast<double>(n_design);
    masked_fc = make_filter(P, o.rmin, sol);
    fcp = masked_fc.get();
    // Active at the target fraction, FrozenSolid at 1, the rest 0 (:3073-3080)
    for (long long e = 0; e < nv; ++e)
      x0[e] = eff[e] == 0 ? o.volfrac : eff[e] == 1 ? 1.0 : 0.0;
  } else {
    fcp = &filter_for(P, o.rmin);
    // x = volfrac on solid voxels (simp_uniform_density, simp.cpp:305-312)
    for (long long e = 0; e < nv; ++e)
      if (P.all_solid || P.solid_h[e]) x0[e] = o.volfrac;
  }
  FilterCache& fc = *fcp;
  P.xd.ensure(static_cast<size_t>(nv));
  P.xphys.ensure(static_cast<size_t>(nv));
  P.xnew.ensure(static_cast<size_t>(nv));
  h2d(P.xd.p, x0.data(), static_cast<size_t>(nv), P.s);
  topohip_result r{};
  double t_solve = 0, t_filter = 0, t_oc = 0, t_build = 0;
  const int vb = dot_blocks(nv);
  // Stage plan (build_stage_plan, core/src/simp/simp.cpp:1053-1068): one
  // unprojected stage, or the projected stages in order.
  // (cont: the single dynamic stage of MMA with projection, whose beta starts
  // at beta0 and doubles on plateaus, :1036-1068, :2461-2491)
  struct Stage {
    bool project;
    double beta, move;
    int iterations;
    bool cont;
  };
  std::vector<Stage> plan;
  double mma_beta = mma_cont ? mma->beta0 : 0.0;
  size_t stage_start = 0;  // curve index where the current beta stage began
  if (mma_cont) {
    plan.push_back({false, 0.0, o.move, o.max_iterations, true});
    P.xtil.ensure(static_cast<size_t>(nv));
  } else if (proj && !proj->empty()) {
    for (const ProjStage& ps : *proj) plan.push_back({true, ps.beta, ps.move, ps.iterations});
    P.xtil.ensure(static_cast<size_t>(nv));
  } else {
    plan.push_back({false, 0.0, o.move, o.max_iterations});
  }
  // filtered (and, in a projected stage, projected) density of the design
  // into P.xphys; the filtered one stays in P.xtil for the OC chain rule
  auto physical = [&](bool project, double beta) {
    if (project) {
      launch_filter(P.g, fc.st, sol, fc.hs.p, P.xd.p, P.xtil.p, 0, P.s);
      launch_project(P.xtil.p, P.xphys.p, sol, nv, beta, eta, P.s);
    } else {
      launch_filter(P.g, fc.st, sol, fc.hs.p, P.xd.p, P.xphys.p, 0, P.s);
    }
  };
  int it = 0;
  double t_prev_obs = -1.0;  // end of the previous iteration's observation
  for (const Stage& st : plan) {
    bool stage_converged = false;
    for (int si = 0; si < st.iterations; ++si, ++it) {
      // SimpOptions::cancel, polled once per design iteration at its start
      // (core/include/topopt/simp.hpp:1505-1521, core/src/simp/simp.cpp:2159-2163)
      if (P.cancel && __atomic_load_n(P.cancel, __ATOMIC_ACQUIRE) != 0) {
        r.cancelled = 1;
        break;
      }
      const double t_top = now_ms();
      double ta = t_top;
      const double cur_beta = st.cont ? mma_beta : st.beta;
      // mma_continuation_move (:251-254): move * min(1, 8 / beta)
      double cur_move = st.cont && cur_beta > 8.0 ? st.move * (8.0 / cur_beta) : st.move;
      if (mma && mma->adaptive) {
        // adaptive move (:2179-2194, oscillation_fraction / adapt_move_value
        // :262-300): seeded at the reference move for the first two MMA steps
        // (again after a beta advance restarts them), then shrunk above osc_hi,
        // grown below osc_lo and clamped; the reading is over the design set
        if (mma_iter < 2) {
          adaptive_move_val = cur_move;
        } else {
          const int ob = oc_blocks(nv);
          launch_osc_count(P.xd.p, P.mma_x1.p, P.mma_x2.p, sol, nv, P.partials.p,
                           P.partials2.p, ob, P.s);
          launch_sum_partials(P.partials.p, ob, P.scal.p, P.s);
          launch_sum_partials(P.partials2.p, ob, P.scal.p + 1, P.s);
          d2h(P.scal_h, P.scal.p, 2, P.s);
          const double osc = P.scal_h[0] > 0.0 ? P.scal_h[1] / P.scal_h[0] : 0.0;
          double v = adaptive_move_val;
          if (osc > mma->ad_osc_hi) v *= mma->ad_shrink;
          else if (osc < mma->ad_osc_lo) v *= mma->ad_grow;
          adaptive_move_val = std::min(std::max(v, mma->ad_min), mma->ad_max);
        }
        cur_move = adaptive_move_val;
      }
      physical(st.project || st.cont, cur_beta);
      if (mask) launch_mask_pins(P.xphys.p, d_eff.p, nv, P.s);
      sync(P);
      const double ms_filter = now_ms() - ta;
      t_filter += ms_filter;
      ta = now_ms();
      topohip_cginfo ci{};
      double c = 0.0;
      try {
        c = simp_compliance_dev(P, P.xphys.p, o.E0, penalty_at(r.iterations + 1), o.rho_min,
                                o.cg_tol, o.cg_max_iterations, &ci);
      } catch (const Error& e) {
        // SolverDeadlineExceeded is-a SolverNonConvergence: either ends the run
        if (e.code != TOPOHIP_ENONCONVERGENCE && e.code != TOPOHIP_EDEADLINE) throw;
        r.non_convergent = 1;
        if (e.code == TOPOHIP_EDEADLINE) r.deadline_exceeded = 1;
        break;
      }
      const double ms_solve = now_ms() - ta;
      t_solve += ms_solve;
      t_build += ci.t_mg_build_ms;
      ta = now_ms();
      double change = 0.0;
      if (st.cont)
        mma_update_core(P, fc, sol, nd, ++mma_iter, P.xd.p, P.dc.p, o.volfrac, cur_move,
                        o.rho_min, P.xnew.p, nullptr, &change, P.xtil.p, cur_beta, eta);
      else if (st.project)
        oc_update_projected_core(P, fc, sol, nd, mask != nullptr, P.xd.p, P.xtil.p, P.dc.p,
                                 st.beta, eta, o.volfrac, st.move, o.rho_min, P.xnew.p, &change);
      else if (mma)
        mma_update_core(P, fc, sol, nd, ++mma_iter, P.xd.p, P.dc.p, o.volfrac, cur_move, o.rho_min,
                        P.xnew.p, nullptr, &change);
      else
        oc_update_core(P, fc, sol, nd, mask != nullptr, P.xd.p, P.dc.p, o.volfrac, st.move,
                       o.rho_min, P.xnew.p, nullptr, nullptr, nullptr, &change);
      std::swap(P.xd.p, P.xnew.p);
      const double ms_update = now_ms() - ta;
      t_oc += ms_update;
      if (r.iterations == 0) r.initial_compliance = c;
      ++r.iterations;
      r.cg_iterations_total += ci.iterations;
      ta = now_ms();
      physical(st.project, st.beta);  // (the continuation stage reports the filtered volume)
      launch_sum_vec<double>(P.xphys.p, nv, P.partials.p, vb, P.s);
      launch_sum_partials(P.partials.p, vb, P.scal.p, P.s);
      d2h(P.scal_h, P.scal.p, 1, P.s);
      const double ms_analysis = now_ms() - ta;
      t_filter += ms_analysis;
      const double vf = nd > 0 ? P.scal_h[0] / nd : 0.0;
      if (history) {
        history[4 * it + 0] = c;
        history[4 * it + 1] = change;
        history[4 * it + 2] = vf;
        history[4 * it + 3] = ci.iterations;
      }
      // stage_should_stop (core/src/simp/simp.cpp:1219-1232): the MMA plateau
      // detector on the compliance history, else the change test
      curve.push_back(c);
      const bool use_plateau = mma && mma->plateau_window > 0;
      bool plateau = false, stop = false;
      if (st.cont) {
        // beta continuation: a plateau of the current beta stage (the progress
        // gate on the first stage only) doubles beta, capped, and restarts the
        // MMA asymptotes; a plateau at the cap ends the run
        const int n0 = static_cast<int>(stage_start);
        plateau = mma_objective_plateau(curve.data() + n0, static_cast<int>(curve.size()) - n0,
                                        mma->plateau_window, mma->plateau_tol,
                                        n0 == 0 ? mma->plateau_min_drop : 0.0,
                                        mma->plateau_flat_windows);
        if (plateau) {
          if (mma_beta < mma->beta_max) {
            mma_beta = std::min(mma_beta * 2.0, mma->beta_max);
            stage_start = curve.size();
            mma_iter = 0;
          } else {
            stop = true;
          }
        }
      } else {
        plateau = use_plateau &&
                  mma_objective_plateau(curve.data(), static_cast<int>(curve.size()),
                                        mma->plateau_window, mma->plateau_tol,
                                        mma->plateau_min_drop, mma->plateau_flat_windows);
        stop = use_plateau ? plateau : change < o.change_tol;
      }
      if (P.observer) {
        // SimpIterationObservation + IterationPhaseTimes
        // (core/include/topopt/simp.hpp:870-1000). The projection and mask
        // pins run inside the filter's launch sequence: their time is in
        // filter_ms. The sensitivity sweep is solve_ms minus the linear solve.
        ta = now_ms();
        topohip_iteration_obs ob{};
        ob.iteration = it + 1;
        ob.compliance = c;
        ob.change = change;
        ob.volume_fraction = vf;
        ob.cg_iterations = ci.iterations;
        ob.cg_used_multigrid = ci.used_multigrid;
        ob.cg_mg_levels = ci.mg_levels;
        ob.cg_hier_built = ci.hier_built;
        ob.cg_mg_cycles_attempted = ci.mg_cycles_attempted;
        ob.beta = st.project || st.cont ? cur_beta : 0.0;
        ob.plateau = plateau ? 1 : 0;
        ob.active_fraction = 1.0;
        ob.tail_prev_ms = t_prev_obs < 0.0 ? 0.0 : t_top - t_prev_obs;
        ob.filter_ms = ms_filter;
        ob.solve_ms = ms_solve;
        ob.fea_ms = ci.t_total_ms;
        ob.sens_ms = ms_solve > ci.t_total_ms ? ms_solve - ci.t_total_ms : 0.0;
        ob.update_ms = ms_update;
        ob.analysis_ms = ms_analysis;
        ob.solver_build_ms = ci.t_build_ms;
        ob.solver_mg_build_ms = ci.t_mg_build_ms;
        ob.solver_mg_ms = ci.t_mg_ms;
        ob.solver_cg_ms = ci.t_cg_ms;
        ob.fea_solves = 1;
        ob.matvecs = ci.matvecs;
        sample_process_memory(ob);
        const double t_obs = now_ms();
        ob.observe_ms = t_obs - ta;
        ob.total_ms = t_obs - t_top;
        ob.residual_ms = ob.total_ms - (ob.filter_ms + ob.project_ms + ob.solve_ms +
                                        ob.update_ms + ob.analysis_ms + ob.observe_ms);
        P.observer(&ob, P.observer_user);
        t_prev_obs = now_ms();
      }
      if (stop) {
        stage_converged = true;
        break;
      }
    }
    r.converged = stage_converged && !r.cancelled ? 1 : 0;
    if (r.non_convergent || r.cancelled) break;
  }
  physical(plan.back().project || plan.back().cont,
           plan.back().cont ? mma_beta : plan.back().beta);
  // volume fraction of the unpinned physical density (active_volfrac)
  launch_sum_vec<double>(P.xphys.p, nv, P.partials.p, vb, P.s);
  launch_sum_partials(P.partials.p, vb, P.scal.p, P.s);
  d2h(P.scal_h, P.scal.p, 1, P.s);
  r.volume_fraction = nd > 0 ? P.scal_h[0] / nd : 0.0;
  if (mask) launch_mask_pins(P.xphys.p, d_eff.p, nv, P.s);
  if (r.non_convergent) {
    r.compliance = history && r.iterations > 0 ? history[4 * (r.iterations - 1)] : 0.0;
  } else {
    try {
      const double ta = now_ms();
      topohip_cginfo ci{};
      r.compliance = simp_compliance_dev(P, P.xphys.p, o.E0, o.penalty, o.rho_min, o.cg_tol,
                                         o.cg_max_iterations, &ci);
      t_solve += now_ms() - ta;
    } catch (const Error& e) {
      if (e.code != TOPOHIP_ENONCONVERGENCE && e.code != TOPOHIP_EDEADLINE) throw;
      r.non_convergent = 1;
      if (e.code == TOPOHIP_EDEADLINE) r.deadline_exceeded = 1;
      r.converged = 0;
    }
  }
  if (design_out) d2h(design_out, P.xd.p, static_cast<size_t>(nv), P.s);
  if (xphys_out) d2h(xphys_out, P.xphys.p, static_cast<size_t>(nv), P.s);
  r.t_total_ms = now_ms() - t0;
  r.t_solve_ms = t_solve;
  r.t_build_ms = t_build;
  r.t_filter_ms = t_filter;
  r.t_oc_ms = t_oc;
  if (res) *res = r;
}

// fea_von_mises_field (core/src/fea/assembly.cpp:976-996) and the float
// arrays of fields.bin (core/src/io/fields.cpp:107-118), formed on the device.
void result_fields(Problem& P, double E, const double* u_host, double* vm64, float* vm32,
                   float* disp32) {
  const double nu = P.nu, h = P.h;
  if (!(E > 0.0)) throw Error(TOPOHIP_EINVAL, "hex8_stress: youngs_modulus must be > 0");
  if (!(h > 0.0)) throw Error(TOPOHIP_EINVAL, "hex8_stress: element_size must be > 0");
  if (!(nu > -1.0 && nu < 0.5))
    throw Error(TOPOHIP_EINVAL, "hex8_stress: poisson must be in (-1, 0.5)");
  const long long nv = P.g.elems(), nd = P.g.ndof();
  const double* du = P.u.p;
  if (u_host) {
    P.fu.ensure(static_cast<size_t>(nd));
    h2d(P.fu.p, u_host, static_cast<size_t>(nd), P.s);
    du = P.fu.p;
  } else if (!P.u_valid) {
    throw Error(TOPOHIP_EINVAL, "no displacement: pass u or solve first");
  }
  // D and the centroid derivative as hex8_stress forms them (:258-266, :270-281)
  const double c = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
  VmCoef k{(2.0 / h) * 0.125, c * (1.0 - nu), c * nu, c * (1.0 - 2.0 * nu) / 2.0};
  if (vm64) P.fvm.ensure(static_cast<size_t>(nv));
  if (vm32) P.fvm32.ensure(static_cast<size_t>(nv));
  if (vm64 || vm32)
    launch_von_mises(P.g, k, P.all_solid ? nullptr : P.solid.p, du, vm64 ? P.fvm.p : nullptr,
                     vm32 ? P.fvm32.p : nullptr, P.s);
  if (disp32) {
    P.fu32.ensure(static_cast<size_t>(nd));
    launch_narrow_f32(du, P.fu32.p, nd, P.s);
  }
  if (vm64) d2h(vm64, P.fvm.p, static_cast<size_t>(nv), P.s);
  if (vm32) d2h(vm32, P.fvm32.p, static_cast<size_t>(nv), P.s);
  if (disp32) d2h(disp32, P.fu32.p, static_cast<size_t>(nd), P.s);
  sync(P);
}

// The halves of write_density_snapshot (core/src/simp/observability.cpp:374-377).
void density_halves(Problem& P, const double* d_host, uint16_t* out) {
  const long long nv = P.g.elems();
  const double* d = P.xphys.p;
  if (d_host) {
    P.fu.ensure(static_cast<size_t>(nv));
    h2d(P.fu.p, d_host, static_cast<size_t>(nv), P.s);
    d = P.fu.p;
  } else if (!d || P.xphys.n < static_cast<size_t>(nv)) {
    throw Error(TOPOHIP_EINVAL, "no physical density: pass one or optimise first");
  }
  P.fh16.ensure(static_cast<size_t>(nv));
  launch_density_f16(d, P.fh16.p, nv, P.s);
  d2h(out, P.fh16.p, static_cast<size_t>(nv), P.s);
}

}  // namespace topohip
