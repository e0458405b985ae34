// oracle.hpp — FP64 CPU restatement of the reference voxel-SIMP inner loop.
//
// TEST INFRASTRUCTURE ONLY. This is the checker the HIP path is compared
// against (and the CPU baseline that is timed beside it). Nothing in the
// product library links, loads or calls it; only tests/, the smoke check and
// the bench's cpu_baseline leg do.
//
// Every function names the reference file:line it restates. Paths are relative
// to the reference checkout (core/...). What pins each part (tests in
// tests/test_oracle.py unless named otherwise):
//   hex8_stiffness      reference fixture hex_stiffness_reference.json (nu 0.30
//                       and 0.40, all 576 entries); symmetry, 6 rigid modes
//   apply_full          explicitly assembled K on small grids; K.t = 0,
//                       symmetry, positivity, linearity in E; void voxels
//   build_reduced       void-gate cases of core/tests/test_matfree.cpp:351-441
//   mf_mgpcg / solve_*  dense direct solve of the reduced system (MG and
//                       Jacobi-CG routes); uniform-strain patch test;
//                       mesh-independent MG iteration counts
//   build_hierarchy     level count and DOF sizes of the reference rule
//   simp_compliance     c = f.u; central differences of c in x
//   DensityFilter       O(N^2) brute force of the cone weights; stencil sizes
//   oc_update           brute-force bisection re-run, bounds, move limit,
//                       volume target
//   penalty_at          continuation-ramp values; heaviside: the reference's
//                       projection_unit_reference spot values (1e-12)
//   simp_optimize       structural gates (monotone compliance, volume,
//                       xPhys = filter(design))
//   hex8_stress         the reference's closed forms (uniaxial strain, pure
//                       shear, rigid translation; core/tests/unit/test_fea.cpp:303-390)
//   float_to_half       numpy's IEEE float16 conversion (every finite input),
//                       the reference's anchors and round-trip bound
//                       (core/tests/unit/test_observability.cpp:55-74)
//   fnv1a64             the fingerprints stored in design.bin files the reference
//                       wrote (tests/golden/design_bin_reference_4x4x4.bin), and
//                       the published FNV-1a 64-bit vectors once the reference's
//                       offset basis is swapped for the published one
//   mma_update          (tests/test_oracle_mma.py) convergence to the closed-form
//                       optimum of a separable problem (sum c/x at fixed volume),
//                       bounds and move box on every step; the loop's descent
//   mma_objective_plateau  the reference's hand-built curves and verdicts
//                       (core/tests/unit/test_simp.cpp:1245-1435)
//   simp_compliance     warm start: a guess at the solution stops after one
//   (initial_guess)     iteration with the same compliance; a neighbouring field
//                       cuts the iteration count and keeps the cold answer
//   masked loop         (tests/test_oracle_masked.py) all-Active mask equals
//                       the plain loop bit for bit; masked filter against a
//                       brute-force evaluation of its definition; no bleed of
//                       frozen x; OC keeps frozen x, box/move limits, Active
//                       budget; pins (FrozenSolid 1 / FrozenVoid 0) and the
//                       final solve on the analysis grid
// Parity unpinned (no reference-held vector exists and the reference cannot
// be built here without Eigen): the reference's exact MG-PCG iteration counts
// and CgInfo traces, the SimplicialLDLT rounding of its coarsest solve, and its
// per-iteration compliance trajectory at the benchmark sizes.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace oracle {

constexpr int kDof = 24;
constexpr int kNumColors = 8;

// Voxel grid: x-fastest, then y, then z (core/include/topopt/voxel.hpp:64-69).
// `solid` is one byte per voxel (non-zero = solid); empty vector = all solid.
struct Grid {
  int nx = 0, ny = 0, nz = 0;
  double h = 1.0;
  std::vector<std::uint8_t> solid;

  std::size_t voxels() const {
    return static_cast<std::size_t>(nx) * static_cast<std::size_t>(ny) *
           static_cast<std::size_t>(nz);
  }
  std::size_t index(int i, int j, int k) const {
    return (static_cast<std::size_t>(k) * static_cast<std::size_t>(ny) +
            static_cast<std::size_t>(j)) *
               static_cast<std::size_t>(nx) +
           static_cast<std::size_t>(i);
  }
  bool is_solid(int i, int j, int k) const {
    return solid.empty() || solid[index(i, j, k)] != 0;
  }
  std::size_t solid_count() const;
  int nnx() const { return nx + 1; }
  int nny() const { return ny + 1; }
  int nnz() const { return nz + 1; }
  int nodes() const { return nnx() * nny() * nnz(); }
  // core/src/fea/assembly.cpp:41-43
  int node_index(int a, int b, int c) const { return (c * nny() + b) * nnx() + a; }
};

struct DirichletBC { int node = 0; int comp = 0; double value = 0.0; };
struct NodalLoad { int node = 0; int comp = 0; double value = 0.0; };

// core/include/topopt/fea.hpp:418-426
struct SolverNonConvergence : public std::runtime_error {
  int iterations;
  double residual;
  SolverNonConvergence(const std::string& w, int it, double r)
      : std::runtime_error(w), iterations(it), residual(r) {}
};

// Subset of CgInfo (core/include/topopt/fea.hpp:272-381).
struct CgInfo {
  bool converged = false;
  int iterations = 0;
  double residual = 0.0;
  bool used_multigrid = false;
  int mg_levels = 0;
  bool hier_built = false;
  int mg_cycles_attempted = 0;
  long long matvecs = 0;
  double t_build_ms = 0.0, t_mg_build_ms = 0.0, t_mg_ms = 0.0, t_cg_ms = 0.0,
         t_total_ms = 0.0;
};

// ---- L1 element ------------------------------------------------------------
// hex8_stiffness(E, nu, h): core/src/fea/hex_element.cpp:23-114. Row-major 24x24.
void hex8_stiffness(double E, double nu, double h, double* K576);
// Hex8 node order (bottom CCW then top CCW): core/src/fea/assembly.cpp:45-54.
std::array<int, 8> element_nodes(const Grid& g, int i, int j, int k);
// core/src/fea/assembly.cpp:70-116 (self-weight, 1/8 lumped per corner node).
std::vector<NodalLoad> self_weight_loads(const Grid& g, double density,
                                         double gravity, const double dir[3]);

// ---- L2 matrix-free operator ----------------------------------------------
// MfElem (core/src/fea/fea_matfree.hpp:56-59).
struct MfElem {
  double factor = 1.0;
  int edof[kDof];
};

void set_threads(int n);  // 0 = hardware concurrency
int threads();

// Colour-sorted element table: core/src/fea/matfree.cpp:422-472.
std::vector<MfElem> build_elems(const Grid& g, const std::vector<double>* youngs,
                                std::vector<int>* color_offsets);
// y = K x over the full stiffness, colour 0..7 in fixed order, column-AXPY
// mul+add per element: core/src/fea/matfree.cpp:75-123, 535-553.
void apply_full(const std::vector<MfElem>& elems, const std::vector<int>& coff,
                const double* Ke, const double* x, double* y, int ndof);
// fea_matfree_apply graded overload: core/src/fea/matfree.cpp:1085-1104,1190.
std::vector<double> matfree_apply(const Grid& g, const std::vector<double>& youngs,
                                  double nu, const std::vector<double>& u);

// Reduced, void-gated system: core/src/fea/matfree.cpp:670-834.
struct Reduced {
  int ndof = 0;
  int ng = 0;
  double Ke[kDof * kDof];
  std::vector<MfElem> elems;
  std::vector<int> color_offsets;
  std::vector<double> up;           // prescribed field (full)
  std::vector<int> kept_global;     // kept id -> global dof
  std::vector<double> rg;           // reduced RHS
  std::vector<double> invdiag;      // 1 / diag
  mutable std::vector<double> xfull, yfull;
  mutable long long matvecs = 0;
  // core/src/fea/matfree.cpp:595-608
  void apply_kgg(const double* xg, double* yg) const;
};
Reduced build_reduced(const Grid& g, const std::vector<double>& youngs, double nu,
                      const std::vector<DirichletBC>& bcs,
                      const std::vector<NodalLoad>& loads, CgInfo* info);

// Jacobi-PCG (Eigen ConjugateGradient algorithm): core/src/fea/matfree.cpp:841-1076.
void cg_solve(const Reduced& m, double tol, int max_it, std::vector<double>& x,
              int& iters, double& err, bool& converged);

// ---- L2 multigrid ---------------------------------------------------------
// Compressed sparse row matrix (the restatement's stand-in for Eigen::SparseMatrix).
struct Csr {
  int rows = 0, cols = 0;
  std::vector<int> ptr, col;
  std::vector<double> val;
  void spmv(const double* x, double* y) const;   // y = A x
  void spmv_t(const double* x, double* y) const; // y = A^T x (y sized cols)
};

// One coarse level: core/src/fea/multigrid.cpp:474-488.
struct Level {
  int nx = 0, ny = 0, nz = 0;  // node dims
  int n = 0;
  Csr A;
  std::vector<double> Dinv;
  Csr P;  // prolongation coarse(level+1) -> this
  std::vector<int> active;
  bool coarsest = false;
  // Envelope Cholesky factor of the coarsest operator (oracle.cpp).
  struct Factor {
    int n = 0;
    std::vector<int> perm, first;
    std::vector<std::size_t> off;
    std::vector<double> L;
  } chol;
};

// Chebyshev interval of one level's smoother: theta = (lU + lL) / 2,
// delta = (lU - lL) / 2 for the bounds [lL, lU] on the spectrum of D^-1 A.
struct ChebLevel {
  double lambda = 0.0;  // power-iteration estimate of lambda_max(D^-1 A)
  double theta = 0.0, delta = 0.0;
};

struct Hierarchy {
  const Reduced* m = nullptr;
  Csr P0;                     // fine kept rows x level-1 cols
  std::vector<Level> coarse;  // levels 1..L
  std::vector<ChebLevel> cheb;  // Chebyshev smoother only: levels 0..L-2
  int levels() const { return 1 + static_cast<int>(coarse.size()); }
};

// V-cycle smoother: 0 = damped Jacobi, omega 0.6, 1 + 1 sweeps (the
// reference's, core/src/fea/multigrid.cpp:72,1128-1160); 1 = Chebyshev of the
// given degree (degree operator applies per smoothing) on the Jacobi-scaled
// operator, an option of this engine with no reference counterpart. Process-wide.
void set_smoother(int kind, int degree);
// Chebyshev constants: power-iteration steps for lambda_max and the interval
// [kChebLow, kChebHigh] * lambda_max it smooths.
constexpr int kChebPowerSteps = 10;
constexpr double kChebLow = 0.3, kChebHigh = 1.1;
// The deterministic start vector of the power iteration at a level DOF whose
// node-major index is g = 3 * node + component.
inline double cheb_start(long long g) {
  const unsigned h = static_cast<unsigned>(g + 1) * 2654435761u;
  return static_cast<double>(h >> 8) * (1.0 / 16777216.0) - 0.5;
}

// build_mf_hierarchy: core/src/fea/multigrid.cpp:1344-1515 (geometric branch,
// unpadded), galerkin_a1_from_prolong :1543-1868, build_hierarchy :557-716.
bool build_mf_hierarchy(const Reduced& m, int nnx, int nny, int nnz, Hierarchy& H);
// mf_v_cycle: core/src/fea/multigrid.cpp:1143-1160 (+ v_cycle :798-819).
void mf_v_cycle(const Hierarchy& H, const std::vector<double>& b,
                std::vector<double>& x);
// mf_mgpcg: core/src/fea/multigrid.cpp:1211-1318.
bool mf_mgpcg(const Hierarchy& H, const std::vector<double>& b, double tol,
              int max_it, std::vector<double>& x, int& iters, double& resid);

// solve_mgcg_matfree / fea_solve_mgcg_matfree (graded):
// core/src/fea/multigrid.cpp:2098-2319, 2540-2570. Returns full u (3*nodes).
std::vector<double> solve_mgcg_matfree(const Grid& g, const std::vector<double>& youngs,
                                       double nu, const std::vector<DirichletBC>& bcs,
                                       const std::vector<NodalLoad>& loads,
                                       double tol, int max_it, CgInfo* info,
                                       const std::vector<double>* initial_guess = nullptr);
// fea_solve_cg_matfree (graded, Jacobi-CG only): core/src/fea/matfree.cpp:1190-1290.
std::vector<double> solve_cg_matfree(const Grid& g, const std::vector<double>& youngs,
                                     double nu, const std::vector<DirichletBC>& bcs,
                                     const std::vector<NodalLoad>& loads, double tol,
                                     int max_it, CgInfo* info);

// ---- L3 SIMP ----------------------------------------------------------------
struct SimpParams {
  double E0 = 1.0, nu = 0.3, penalty = 3.0, density_min = 1e-3;
};
// clamp_density + simp_youngs: core/src/simp/simp.cpp:44-46, 299-303.
double simp_youngs(const SimpParams& p, double rho);

struct SimpCompliance {
  double compliance = 0.0;
  std::vector<double> dc;  // per voxel
  std::vector<double> u;   // full displacement
  CgInfo cg;
};
// simp_compliance (MultigridCG_Matfree route + sensitivity sweep):
// core/src/simp/simp.cpp:567-828.
SimpCompliance simp_compliance(const Grid& g, const SimpParams& p,
                               const std::vector<double>& density,
                               const std::vector<DirichletBC>& bcs,
                               const std::vector<NodalLoad>& loads, double tol,
                               int max_it, bool use_mg = true,
                               const std::vector<double>* initial_guess = nullptr);

// DensityFilter (core/include/topopt/simp.hpp:387-406; core/src/simp/simp.cpp:834-944).
struct DensityFilter {
  std::size_t voxel_count = 0;
  std::vector<std::size_t> row_start, neighbor;
  std::vector<double> weight, weight_sum;
  std::vector<double> filter_density(const std::vector<double>& x) const;
  std::vector<double> filter_sensitivity(const std::vector<double>& g) const;
};
DensityFilter make_density_filter(const Grid& g, double radius);

// oc_update: core/src/simp/simp.cpp:948-1032. `bisections` (optional) reports
// the number of bisection steps taken and `lambda` the multiplier used.
std::vector<double> oc_update(const Grid& g, const DensityFilter& f,
                              const std::vector<double>& x,
                              const std::vector<double>& dcompliance,
                              double volfrac, double move, double density_min,
                              int* bisections = nullptr, double* lambda = nullptr);

// penalty_continuation_ramp / penalty_at_iteration: core/src/simp/simp.cpp:381-427.
struct PenaltyStage { double penalty; int iterations; };
double penalty_at_iteration(const std::vector<PenaltyStage>& s, int iteration);
std::vector<PenaltyStage> penalty_continuation_ramp(double p0, double p1, double step,
                                                    int per_stage);

// One beta-continuation stage (ProjectionStage, core/include/topopt/simp.hpp:518-524).
struct ProjStage { double beta; double move; int iterations; };
// heaviside_continuation_schedule (core/src/simp/simp.cpp:367-373).
std::vector<ProjStage> heaviside_continuation_schedule();

struct SimpOptions {
  double volume_fraction = 0.5;
  double filter_radius = 1.5;
  double move = 0.2;
  int max_iterations = 200;
  double change_tol = 0.01;
  double cg_tolerance = 1e-8;
  int cg_max_iterations = 0;
  bool use_mg = true;
  std::vector<PenaltyStage> penalty_continuation;
  // Projected stages (SimpOptions::projection / projection_eta): when set, the
  // loop runs them in order (build_stage_plan, simp.cpp:1053-1068) instead of
  // one unprojected stage of max_iterations at `move`.
  std::vector<ProjStage> projection;
  double projection_eta = 0.5;
  // SimpOptions::updater == MMA and its plateau stop (core/include/topopt/simp.hpp:
  // 1119, 1256-1276): plain (unprojected) stage only.
  bool mma = false;
  int mma_plateau_window = 10;
  double mma_plateau_tol = 1e-3;
  double mma_plateau_min_drop = 0.05;
  int mma_plateau_flat_windows = 3;
  // SimpOptions::mma_projection (core/include/topopt/simp.hpp:1488-1500): MMA
  // with Heaviside projection and beta continuation, one dynamic stage
  // (build_stage_plan, simp.cpp:1036-1068): beta starts at beta0 and doubles
  // (capped at beta_max) on each plateau of the current beta stage.
  bool mma_projection = false;
  double mma_projection_beta0 = 1.0;
  double mma_projection_beta_max = 32.0;
  // SimpOptions::adaptive_move (simp.hpp:1108-1118, MMA only): the move limit
  // evolves from the oscillation fraction of the design voxels' last two steps.
  bool adaptive_move = false;
  double adaptive_move_grow = 1.2, adaptive_move_shrink = 0.7;
  double adaptive_move_osc_lo = 0.15, adaptive_move_osc_hi = 0.30;
  double adaptive_move_max = 0.5, adaptive_move_min = 0.02;
};
struct SimpIteration { double compliance, change, volume_fraction; int cg_iterations; };
struct SimpResult {
  std::vector<double> design, physical_density;
  std::vector<SimpIteration> history;
  double compliance = 0.0, initial_compliance = 0.0, volume_fraction = 0.0;
  int iterations = 0;
  bool converged = false, non_convergent = false;
};
// simp_optimize plain OC loop: core/src/simp/simp.cpp:2018-2608 (single
// unprojected stage, OC updater, no active domain, cold solves).
SimpResult simp_optimize(const Grid& g, const SimpParams& p,
                         const std::vector<DirichletBC>& bcs,
                         const std::vector<NodalLoad>& loads, const SimpOptions& o);

// ---- Passive regions: the masked OC loop -----------------------------------
// MaskValue (core/include/topopt/voxel.hpp:42-46): one byte per voxel.
enum : std::uint8_t { kActive = 0, kFrozenSolid = 1, kFrozenVoid = 2 };
// effective_design_mask (core/src/simp/simp.cpp:2610-2633): Empty voxels become
// FrozenVoid. (Load/Fixture tags -> FrozenSolid: the oracle grid carries only
// solid/empty, so that part is the caller's.)
std::vector<std::uint8_t> effective_design_mask(const Grid& g, const std::vector<std::uint8_t>& mask);
// analysis_grid_for_mask (simp.cpp:2640-2645): FrozenVoid voxels become Empty.
Grid analysis_grid_for_mask(const Grid& g, const std::vector<std::uint8_t>& eff);
// apply_mask_pins (simp.cpp:2649-2654): FrozenSolid -> 1, FrozenVoid -> 0.
void apply_mask_pins(const std::vector<std::uint8_t>& eff, std::vector<double>& xphys);
// make_density_filter(grid, radius, mask) (simp.cpp:834-890): neighbours and
// rows restricted to solid Active voxels (no bleed across the mask boundary).
DensityFilter make_density_filter(const Grid& g, double radius,
                                  const std::vector<std::uint8_t>& mask);
// oc_update_masked (simp.cpp:2662-2734): only Active voxels move, frozen ones
// keep their x; budget volfrac * n_active on the masked filter's volume.
std::vector<double> oc_update_masked(const Grid& g, const DensityFilter& f,
                                     const std::vector<std::uint8_t>& eff,
                                     const std::vector<double>& x,
                                     const std::vector<double>& dcompliance, double volfrac,
                                     double move, double density_min, int* bisections = nullptr,
                                     double* lambda = nullptr);
// simp_optimize(grid, params, bcs, loads, options, mask) (simp.cpp:2989-3689)
// restricted like simp_optimize above: OC (or MMA, unprojected) updater, one
// unprojected stage or the projected stages of options.projection (OC only,
// :3160-3440), no SEMDOT / lattice coupling / adaptive move, cold solves.
SimpResult simp_optimize_masked(const Grid& g, const SimpParams& p,
                                const std::vector<DirichletBC>& bcs,
                                const std::vector<NodalLoad>& loads, const SimpOptions& o,
                                const std::vector<std::uint8_t>& mask);

// ---- Heaviside projection (core/src/simp/simp.cpp:352-379) ---------------
// Element stress at natural coordinates (xi, eta, zeta): core/src/fea/hex_element.cpp:242-316
// (sigma = D B u, Voigt [xx, yy, zz, xy, yz, zx], engineering shear).
struct Hex8Stress { double sigma[6]; double von_mises; };
Hex8Stress hex8_stress(double E, double nu, double h, const double* u_elem, double xi = 0.0,
                       double eta = 0.0, double zeta = 0.0);
// core/src/fea/assembly.cpp:976-996: centroid von Mises per solid voxel, 0 elsewhere.
std::vector<double> von_mises_field(const Grid& g, double E, double nu,
                                    const std::vector<double>& u);
// core/src/simp/observability.cpp:35-75: float -> IEEE half, round to nearest even.
std::uint16_t float_to_half(float v);
// core/src/simp/design_store.cpp (design_fingerprint): 64-bit FNV-1a over bytes.
std::uint64_t fnv1a64(const void* data, std::size_t n,
                      std::uint64_t basis = 1469598103934665603ULL);

// MMA (Svanberg 1987) with one volume constraint: mma_update,
// core/src/simp/simp.cpp:1433-1488 over mma_update_joint,
// core/src/simp/mma_joint.cpp:18-118. The state carries the asymptotes and the
// last two iterates across the calls of one run; mma_iter is 1-based.
struct MmaState { std::vector<double> low, upp, xold1, xold2; };
// dset (nullable): the design set of mma_update_masked (:2736-2783, the Active
// voxels; f then the Active-only filter); null = the solid voxels.
// xtilde (nullable) = filter_density(x) makes it mma_update_projected
// (simp.cpp:1509-1636; masked :2861-2990): both sensitivities chained through
// the projection derivative at (beta, eta) before the filter transpose, and
// the constraint on the projected volume sum H(xtilde) over the design set.
std::vector<double> mma_update(MmaState& st, int mma_iter, const Grid& g, const DensityFilter& f,
                               const std::vector<double>& x, const std::vector<double>& dcomp,
                               double volfrac, double move, double density_min,
                               double* lambda_out = nullptr,
                               const std::vector<std::uint8_t>* dset = nullptr,
                               const std::vector<double>* xtilde = nullptr, double beta = 1.0,
                               double eta = 0.5);
// mma_continuation_move (simp.cpp:251-254): move * min(1, 8 / beta).
double mma_continuation_move(double move, double beta);
// oscillation_fraction / adapt_move_value (simp.cpp:262-300): among the design
// voxels that moved on both of the last two steps, the fraction whose direction
// reversed; the move shrinks above osc_hi, grows below osc_lo, clamped.
double oscillation_fraction(const std::vector<double>& x, const std::vector<double>& xold1,
                            const std::vector<double>& xold2,
                            const std::vector<std::size_t>& dof);
double adapt_move_value(double cur, double f_osc, const SimpOptions& o);
// core/src/simp/simp.cpp:1081-1134.
bool mma_objective_plateau(const std::vector<double>& c, int window, double rel_tol,
                           double min_drop, int min_flat_windows);

double heaviside_project(double rt, double beta, double eta);
double heaviside_project_derivative(double rt, double beta, double eta);
// project_active (simp.cpp:2789-2794): the projection on Active voxels only.
void project_active(const std::vector<std::uint8_t>& eff, std::vector<double>& x, double beta,
                    double eta);
// oc_update_masked_projected (simp.cpp:2801-2855): oc_update_projected over the
// Active set with the Active-only filter; frozen voxels keep their x and the
// volume constraint is on the projected Active density (target volfrac *
// n_active). xtilde = filter_density(x) of that filter.
std::vector<double> oc_update_masked_projected(const Grid& g, const DensityFilter& f,
                                               const std::vector<std::uint8_t>& eff,
                                               const std::vector<double>& x,
                                               const std::vector<double>& xtilde,
                                               const std::vector<double>& dcompliance,
                                               double beta, double eta, double volfrac,
                                               double move, double density_min);
// oc_update_projected (simp.cpp:1647-1716): sensitivities chained through the
// projection derivative before the filter transpose, UNCLAMPED dv in the OC
// ratio (a non-finite ratio holds the voxel), volume on the projected
// candidate. xtilde = filter_density(x).
std::vector<double> oc_update_projected(const Grid& g, const DensityFilter& f,
                                        const std::vector<double>& x,
                                        const std::vector<double>& xtilde,
                                        const std::vector<double>& dcompliance, double beta,
                                        double eta, double volfrac, double move,
                                        double density_min, int* bisections = nullptr,
                                        double* lambda = nullptr);

}  // namespace oracle
