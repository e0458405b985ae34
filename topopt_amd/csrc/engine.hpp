// engine.hpp — device-resident problem handle of the voxel-SIMP engine.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../../include/topopt_hip.h"
#include "coarse.hpp"
#include "engine_buf.hpp"
#include "kernels.hpp"

namespace topohip {

// One level of the multigrid hierarchy (level 0 = the fine, matrix-free one).
struct Level {
  Geom g;
  std::vector<uint8_t> mask_h;  // per node: bit c = DOF c active
  DBuf<uint8_t> mask;
  long long nact = 0;           // active DOFs
  DBuf<double> Kem;             // element-major FP64 matrices (levels >= 1)
  DBuf<int> special;            // levels 1, 2: elements with an inactive fine DOF
  int n_special = 0;
  DBuf<double> K1sp;            // level 1 (no stored operator): matrices of `special`
  DBuf<int> spmap;              // level 1 (no stored operator): element -> K1sp slot or -1
  DBuf<float> K32;              // entry-major apply copies
  DBuf<double> K64;
  DBuf<float> dinv32;
  DBuf<double> dinv64;
  DBuf<float> b32, x32, t32;    // V-cycle work (mixed form), stacked load cases
  DBuf<double> b64, x64, t64;   // V-cycle work (FP64 form), stacked load cases
  DBuf<float> d32;              // Chebyshev smoother: direction (stacked cases)
  DBuf<double> d64;
  DBuf<double> cheb;            // Chebyshev smoother: [lambda, 1/theta, (c1_k, c2_k)...]
};

struct FilterCache {
  double radius = 0.0;
  Stencil st;
  DBuf<int> off;
  DBuf<double> w;
  DBuf<double> hs, dv;
};

struct Timers {
  double build = 0, mg_build = 0, mg = 0, cg = 0, total = 0;
};

struct Problem {
  int device = 0;
  int cus = 256;  // compute units of the device
  hipStream_t s = nullptr;
  void* scratch = nullptr;  // this device's allocation-free work array (noalloc_scratch)
  Geom g;
  double h = 1.0, nu = 0.3;
  int precision = TOPOHIP_PREC_MIXED;
  // fine K.u z-chunks of the FP32 and FP64 applies (autotuned; multiples of
  // ku_slab<float>() and ku_slab<double>())
  int tz32 = 8, tz64 = 7;
  // V-cycle smoother: 0 damped Jacobi (the reference's), 1 Chebyshev of degree
  // cheb_degree on every smoothed level
  int smoother = 0;
  int cheb_degree = 1;

  bool all_solid = true;
  std::vector<uint8_t> solid_h;
  DBuf<uint8_t> solid;
  long long n_solid = 0;
  std::vector<uint8_t> fixed_h;    // per node: bit c = Dirichlet-fixed
  std::vector<uint8_t> touched_h;  // per node: 1 if a solid element touches it
  std::vector<double> up_h;        // prescribed displacement (full)
  bool nonzero_bc = false;
  DBuf<double> up;

  double Ke[576];   // reference DOF order
  double Kb[576];   // parity-bit DOF order
  double S8[8 * 576];
  std::vector<double> T64;   // W_{c2}^T S_{c1} W_{c2}, 64 blocks (level-2 elements from E)
  double kdiag[24];
  Modal<float> mc32;
  Modal<double> mc64;
  ModalGroups<float> mg32;   // K.u form of the modal coefficients
  ModalGroups<double> mg64;
  DBuf<double> dKb, dS8, dkdiag, dT64;

  int nrhs = 0;
  std::vector<std::vector<double>> f_h;
  DBuf<double> f;  // nrhs * ndof

  DBuf<double> E64;
  DBuf<float> E32;

  std::vector<std::unique_ptr<Level>> L;
  bool mg_plan_ok = false;
  // Level 1 is a stencil level (three or more levels): its operator is built
  // straight from the fine moduli and no level-1 element matrices are stored.
  bool l1_direct = false;
  bool mg64_alloc = false;   // mixed form: FP64 V-cycle buffers exist (retry path)
  bool mg64_ready = false;   // ... and hold the current hierarchy
  CoarseCR cr;  // exact coarsest-level solver

  // PCG vectors
  // PCG vectors, nb stacked load cases each (stride ndof)
  DBuf<double> b, x, r, p, Ap64, z64, rt, u;
  // The A.p forms the next search direction from the current one (p, p2
  // alternate: neighbouring tiles still read the current one as their halo);
  // pcur is the buffer that holds the current direction.
  DBuf<double> p2;
  int pcur = 0;
  bool u_valid = false;  // P.u holds the displacements of a completed solve
  DBuf<float> r32, z32;  // mixed form: the V-cycle's FP32 input and output
  DBuf<double> partials, partials2;  // nb cases x pstride block partial sums
  DBuf<double> rzpart;               // the <r,z> partials the next A.p re-sums
  DBuf<CgState> st;
  CgState* st_h = nullptr;
  DBuf<double> scal;
  double* scal_h = nullptr;

  // SIMP work
  std::vector<std::unique_ptr<FilterCache>> filters;
  DBuf<OcState> oc;
  OcState* oc_h = nullptr;
  // OC decision margins: last search, smallest since the last reset (each
  // optimisation run resets it) and the rounding bound, relative to the target
  double oc_gap_rel = 0.0, oc_gap_min = 1e300, oc_gap_bound = 0.0;
  DBuf<double> xd, xphys, dc, xnew, dcf;
  DBuf<double> xtil, dpr, dcp, dvf, xcand, xcf;  // projected OC work
  DBuf<uint8_t> active;

  // MMA updater state of the current run: asymptotes, the last two designs,
  // the coefficient arrays and the dual search
  DBuf<double> mma_low, mma_upp, mma_x1, mma_x2, mma_c;
  DBuf<MmaSearch> mma_st;
  DBuf<double> lam_partials;  // kLamPass x blocks: the multi-lambda search passes
  MmaSearch* mma_h = nullptr;
  int mma_last = 0;  // mma_iter of the last completed step

  // result fields and run records
  DBuf<double> fu, fvm;
  DBuf<double> guess;  // warm-start displacement uploaded by the host entry points
  DBuf<float> fvm32, fu32;
  DBuf<uint16_t> fh16;
  topohip_iteration_cb observer = nullptr;  // called after every design iteration
  void* observer_user = nullptr;

  int nb = 0;        // stacked load-case capacity of the solver buffers
  int pstride = 0;   // partial sums per case
  long long matvecs = 0;
  int replacements = 0;
  bool deadline_hit = false;      // the last solve stopped at the thread's solve deadline
  const int* cancel = nullptr;    // caller-owned cancel flag of the design loops (nullable)
  // Captured PCG iterations per (preconditioner kind, direction buffer at the
  // start, length: kCgBatch iterations or one), index (kind * 2 + parity) * 2 + len.
  static constexpr int kCgGraphs = 12;
  hipGraphExec_t cg_graph[kCgGraphs] = {};
  int cg_graph_nc[kCgGraphs] = {};  // case count captured
  int pcg_hint[3] = {0, 0, 0};  // iterations of the last solve per preconditioner kind
  int oc_pass_hint = 0;         // multi-lambda passes the last OC search needed

  Problem() = default;
  Problem(const Problem&) = delete;
  Problem& operator=(const Problem&) = delete;
  ~Problem();
};

std::unique_ptr<Problem> create_problem(int nx, int ny, int nz, double h, double nu,
                                        const uint8_t* solid, int n_bc, const int32_t* bc_node,
                                        const int32_t* bc_comp, const double* bc_val,
                                        int precision, int device);
void set_loads(Problem& P, int n_rhs, const int32_t* counts, const int32_t* node,
               const int32_t* comp, const double* val);
// E64 (device) -> E32, diagonals, Galerkin levels and the coarsest inverse.
void build_hierarchy(Problem& P, Timers& t);
// Solve load cases 0..nc-1 together (block PCG) with the current moduli;
// leaves the full displacements in P.u (stride ndof). d_guess (nc == 1 only)
// is a warm start. info[r] per case.
void solve_cases(Problem& P, int nc, double tol, int max_it, const double* d_guess,
                 topohip_cginfo* info);
FilterCache& filter_for(Problem& P, double radius);
// d_guess (nullable, single load case): warm-start displacement (3 * nodes).
double simp_compliance_dev(Problem& P, const double* d_xphys, double E0, double penalty,
                           double rho_min, double tol, int max_it, topohip_cginfo* info,
                           const double* d_guess = nullptr);
void oc_update_noalloc(Problem& P, double radius, const double* d_x, const double* d_dc,
                       double volfrac, double move, double rho_min, double* d_xnew,
                       uint8_t* d_active, int* bisections, double* lambda);
void oc_update_dev(Problem& P, double radius, const double* d_x, const double* d_dc,
                   double volfrac, double move, double rho_min, double* d_xnew,
                   uint8_t* d_active, int* bisections, double* lambda, double* change);
// One projected (Heaviside beta-continuation) stage: ProjectionStage,
// core/include/topopt/simp.hpp:518-524.
struct ProjStage {
  double beta, move;
  int iterations;
};
// SimpUpdater::MMA and its plateau stop (core/include/topopt/simp.hpp:1119,
// 1256-1276); window <= 0 keeps the change_tol test.
struct MmaOpts {
  int plateau_window = 10;
  double plateau_tol = 1e-3;
  double plateau_min_drop = 0.05;
  int plateau_flat_windows = 3;
  // SimpOptions::mma_projection: Heaviside projection with beta continuation
  // (one dynamic stage, beta0 doubling to beta_max on plateaus)
  bool projection = false;
  double beta0 = 1.0, beta_max = 32.0;
  // SimpOptions::adaptive_move: the move evolves from the oscillation fraction
  bool adaptive = false;
  double ad_grow = 1.2, ad_shrink = 0.7, ad_osc_lo = 0.15, ad_osc_hi = 0.30;
  double ad_max = 0.5, ad_min = 0.02;
};
// One MMA step (mma_update, core/src/simp/simp.cpp:1433-1488) on the design
// set dset with its filter: mma_iter is 1-based (the asymptotes restart at 1
// and 2); the state in P carries over between calls of one run.
// d_xt (nullable) = filter_density(x) makes it mma_update_projected
// (:1509-1636, masked :2861-2990): dc and dv chained through the projection
// derivative at (beta, eta) before the filter transpose, the constraint on the
// projected volume sum H(xtilde) over the design set.
void mma_update_core(Problem& P, FilterCache& fc, const uint8_t* dset, double n_design,
                     int mma_iter, const double* d_x, const double* d_dc, double volfrac,
                     double move, double rho_min, double* d_xnew, double* lambda,
                     double* change, const double* d_xt = nullptr, double beta = 1.0,
                     double eta = 0.5);
// mma_objective_plateau, core/src/simp/simp.cpp:1081-1134.
bool mma_objective_plateau(const double* c, int n, int window, double rel_tol, double min_drop,
                           int min_flat_windows);
// mask (nullable): one MaskValue byte per voxel (0 Active, 1 FrozenSolid,
// 2 FrozenVoid) for the masked loop; FrozenVoid voxels must be empty in P's grid.
// proj (nullable, with eta): projected stages run in order instead of one
// unprojected stage of o.max_iterations at o.move (plain loop only).
void optimize(Problem& P, const topohip_options& o, double* design_out, double* xphys_out,
              double* history, topohip_result* res, const uint8_t* mask = nullptr,
              const std::vector<ProjStage>* proj = nullptr, double eta = 0.5,
              const MmaOpts* mma = nullptr);
std::unique_ptr<FilterCache> make_filter(Problem& P, double radius, const uint8_t* dset);
void design_mask_of(const Problem& P, const uint8_t* mask, std::vector<uint8_t>& eff,
                    std::vector<uint8_t>& dset, long long& n_design);
void hex8_stiffness_host(double E, double nu, double h, double* K576);
// Select the V-cycle smoother (0 Jacobi, 1 Chebyshev of `degree`) for later solves.
void set_smoother(Problem& P, int kind, int degree);
// fea_set_solve_deadline_ms for the calling thread; returns the previous value.
double set_solve_deadline_ms(double absolute_steady_ms);
// fea_matfree_reset_mg_stagnation_latch / fea_matfree_mg_stagnation_latched
// for the calling thread.
void reset_mg_stagnation_latch();
bool mg_stagnation_latched();
// fea_matfree_set_mg_rearm_period and its counters (core/include/topopt/
// fea.hpp:1093-1116): measurement-only, per thread, 0 = off (the default).
void set_mg_rearm_period(int period);
int mg_rearm_period();
long long mg_rearm_attempts();
long long mg_rearm_carries();
double steady_ms();
// fea_von_mises_field of u_host (3 * nodes; null = P.u of the last solve) for
// Young's modulus E: FP64 field into vm64 and/or its float narrowing into
// vm32, and the displacement narrowed to float into disp32 (outputs nullable).
void result_fields(Problem& P, double E, const double* u_host, double* vm64, float* vm32,
                   float* disp32);
// Binary16 halves of float(d) per voxel (d_host null = P.xphys).
void density_halves(Problem& P, const double* d_host, uint16_t* out);

}  // namespace topohip
