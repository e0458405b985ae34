// coarse.hpp — block cyclic reduction solver of the coarsest level.
#pragma once

#include <vector>

#include "engine_buf.hpp"
#include "kernels.hpp"

namespace topohip {

// C = beta*D + sum_t alpha_t op(A_t) op(B_t) on blocks of a workspace
// (offsets in doubles); with ct >= 0 (square jobs, beta = 0) also Ct = C^T.
struct CrGemm {
  long long c, d;
  long long ct = -1;
  long long a[3], b[3];
  int ta[3], tb[3];
  double alpha[3];
  double beta;
  int nterms;
};
// out = beta*y0 + sum_t alpha_t M_t v_t on length-m vectors, M_t row-major
// (matrix offsets index W, vector offsets a load case's vector space V).
struct CrGemv {
  long long out, y0;
  long long m[3], v[3];
  double alpha[3];
  double beta;
  int nterms;
};
struct CrLevel {
  int n = 0;  // planes
  long long D = 0, L = 0, Dinv = 0, P = 0, Q = 0, Pt = 0, Qt = 0;  // in W
  long long b = 0, x = 0;                                           // in V
};

class CoarseCR {
 public:
  // Plans the reduction for a level of geometry g (false if a plane holds
  // more than 256 DOFs).
  bool plan(const Geom& g, hipStream_t s);
  // Factor from the level's element-major FP64 element matrices.
  void build(const double* Kem, const uint8_t* mask, hipStream_t s);
  // Vector space for nr load cases solved together.
  void reserve_cases(int nr);
  // x = A^-1 b for nr stacked load cases (node-ordered level vectors, stride
  // 3*nodes; FP64 arithmetic inside).
  template <typename T>
  void solve(const T* b, T* x, const int* done, hipStream_t s, int nr = 1);
  // The plane-ordered right-hand side / solution space of load case r
  // (restriction writes rhs_map() directly, prolongation reads sol_map()),
  // and the solve between them.
  CrMap rhs_map();
  CrMap sol_map();
  void solve_in_place(const int* done, hipStream_t s, int nr = 1);
  int block() const { return m; }
  int nplanes() const { return planes; }
  size_t bytes() const {
    return static_cast<size_t>(wsize + vsize * ncases) * sizeof(double);
  }

 private:
  Geom geom;
  int axis = 0, planes = 0, m = 0;
  long long wsize = 0, vsize = 0;
  int ncases = 0;
  std::vector<CrLevel> lv;
  std::vector<int> lv_inv, lv_gp, lv_gd, lv_fw, lv_bw;
  // Dense tail: levels >= dl are solved with the explicit inverse of level
  // dl's system (nden x nden at W + xm_dense), formed by the GEMM jobs
  // d_mj (one launch per lv_mj range) from the identity at W + eye_off.
  // 1024 balances the inverse's formation (once per hierarchy build, its
  // GEMMs scale with the tail size) against the GEMV launches it saves in
  // every V-cycle: at 192x64x64 (planes of 243 unknowns) the tail starts at
  // 4 planes (972 unknowns); 2048 (7 planes) and 512 (2 planes) measured
  // 2 % slower per design iteration.
  static constexpr long long kDenseMax = 1024;
  int dl = 0;
  long long nden = 0, xm_dense = 0, eye_off = 0;
  std::vector<int> lv_mj;
  void gemv(const CrGemv* jobs, int n, const int* done, hipStream_t s, int nr);
  void reduce_solve(const int* done, hipStream_t s, int nr);
  DBuf<double> W, V, inv_scratch;
  DBuf<long long> d_inv;
  DBuf<CrGemm> d_gp, d_gd, d_mj;
  DBuf<CrGemv> d_fw, d_bw;
  DBuf<int> d_plane, d_local, d_dst;
};

}  // namespace topohip
