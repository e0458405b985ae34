// oracle_capi.cpp — extern "C" surface of the FP64 oracle, for the Python test
// suite and the bench's CPU-baseline leg (ctypes). TEST INFRASTRUCTURE ONLY.
// Status codes match the product library's (include/topopt_hip.h).
#include <cstring>
#include <string>

#include "oracle.hpp"

namespace {
thread_local std::string g_err;

struct OracleProblem {
  int nx, ny, nz;
  double h;
  const unsigned char* solid;  // may be null (all solid)
  int nbc;
  const int* bc_node;
  const int* bc_comp;
  const double* bc_val;
  int nload;
  const int* ld_node;
  const int* ld_comp;
  const double* ld_val;
};

struct OracleInfo {
  int converged, iterations, used_multigrid, mg_levels, hier_built, mg_cycles_attempted;
  double residual;
  long long matvecs;
  double t_build_ms, t_mg_build_ms, t_mg_ms, t_cg_ms, t_total_ms;
};

oracle::Grid grid_of(const OracleProblem* p) {
  oracle::Grid g;
  g.nx = p->nx; g.ny = p->ny; g.nz = p->nz; g.h = p->h;
  if (g.nx < 1 || g.ny < 1 || g.nz < 1) throw std::invalid_argument("grid dims must be >= 1");
  if (p->solid) g.solid.assign(p->solid, p->solid + g.voxels());
  return g;
}
std::vector<oracle::DirichletBC> bcs_of(const OracleProblem* p) {
  std::vector<oracle::DirichletBC> v(static_cast<std::size_t>(p->nbc));
  for (int i = 0; i < p->nbc; ++i) v[i] = {p->bc_node[i], p->bc_comp[i], p->bc_val[i]};
  return v;
}
std::vector<oracle::NodalLoad> loads_of(const OracleProblem* p) {
  std::vector<oracle::NodalLoad> v(static_cast<std::size_t>(p->nload));
  for (int i = 0; i < p->nload; ++i) v[i] = {p->ld_node[i], p->ld_comp[i], p->ld_val[i]};
  return v;
}
void fill_info(const oracle::CgInfo& c, OracleInfo* o) {
  if (!o) return;
  o->converged = c.converged; o->iterations = c.iterations;
  o->used_multigrid = c.used_multigrid; o->mg_levels = c.mg_levels;
  o->hier_built = c.hier_built; o->mg_cycles_attempted = c.mg_cycles_attempted;
  o->residual = c.residual; o->matvecs = c.matvecs;
  o->t_build_ms = c.t_build_ms; o->t_mg_build_ms = c.t_mg_build_ms;
  o->t_mg_ms = c.t_mg_ms; o->t_cg_ms = c.t_cg_ms; o->t_total_ms = c.t_total_ms;
}

template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const oracle::SolverNonConvergence& e) {
    g_err = e.what();
    return 3;
  } catch (const std::invalid_argument& e) {
    g_err = e.what();
    return 1;
  } catch (const std::runtime_error& e) {
    g_err = e.what();
    return 2;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 6;
  }
}
}  // namespace

extern "C" {

const char* oracle_last_error() { return g_err.c_str(); }
void oracle_set_threads(int n) { oracle::set_threads(n); }
int oracle_threads() { return oracle::threads(); }

int oracle_hex8_stiffness(double E, double nu, double h, double* K) {
  return guarded([&] { oracle::hex8_stiffness(E, nu, h, K); });
}

int oracle_apply(const OracleProblem* p, const double* E, double nu, const double* u,
                 double* y) {
  return guarded([&] {
    const oracle::Grid g = grid_of(p);
    const std::size_t nd = 3 * static_cast<std::size_t>(g.nodes());
    std::vector<double> Ev(E, E + g.voxels()), uv(u, u + nd);
    const std::vector<double> yv = oracle::matfree_apply(g, Ev, nu, uv);
    std::memcpy(y, yv.data(), nd * sizeof(double));
  });
}

int oracle_solve(const OracleProblem* p, const double* E, double nu, double tol, int max_it,
                 int use_mg, const double* guess, double* u, OracleInfo* info) {
  return guarded([&] {
    const oracle::Grid g = grid_of(p);
    const std::size_t nd = 3 * static_cast<std::size_t>(g.nodes());
    std::vector<double> Ev(E, E + g.voxels());
    oracle::CgInfo ci;
    std::vector<double> gv;
    if (guess) gv.assign(guess, guess + nd);
    std::vector<double> uv;
    try {
      uv = use_mg ? oracle::solve_mgcg_matfree(g, Ev, nu, bcs_of(p), loads_of(p), tol, max_it,
                                               &ci, guess ? &gv : nullptr)
                  : oracle::solve_cg_matfree(g, Ev, nu, bcs_of(p), loads_of(p), tol, max_it, &ci);
    } catch (...) {
      fill_info(ci, info);
      throw;
    }
    fill_info(ci, info);
    std::memcpy(u, uv.data(), nd * sizeof(double));
  });
}

// Hierarchy shape for the given system: levels (incl. the fine one) and the
// DOF count of each level (up to 16), without solving.
int oracle_hierarchy(const OracleProblem* p, const double* E, double nu, int* levels,
                     int* dofs) {
  return guarded([&] {
    const oracle::Grid g = grid_of(p);
    std::vector<double> Ev(E, E + g.voxels());
    oracle::Reduced m = oracle::build_reduced(g, Ev, nu, bcs_of(p), loads_of(p), nullptr);
    oracle::Hierarchy H;
    if (!oracle::build_mf_hierarchy(m, g.nnx(), g.nny(), g.nnz(), H)) {
      *levels = 0;
      return;
    }
    *levels = H.levels();
    dofs[0] = m.ng;
    for (std::size_t i = 0; i < H.coarse.size() && i < 15; ++i) dofs[i + 1] = H.coarse[i].n;
  });
}

// Operator of hierarchy level `level` (1 = A1 from galerkin_a1, 2.. = the
// sparse P^T A P levels) as CSR over that level's active DOFs, numbered
// node-major then component. Two calls: with ptr == NULL only *n and *nnz are
// set; then ptr (n + 1), col and val (nnz) are filled. Level 0 is the fine
// reduced operator's row count only (*nnz = 0).
int oracle_level_csr(const OracleProblem* p, const double* E, double nu, int level, int* n,
                     int* nnz, int* ptr, int* col, double* val) {
  return guarded([&] {
    const oracle::Grid g = grid_of(p);
    std::vector<double> Ev(E, E + g.voxels());
    oracle::Reduced m = oracle::build_reduced(g, Ev, nu, bcs_of(p), loads_of(p), nullptr);
    if (level == 0) {
      *n = m.ng;
      *nnz = 0;
      return;
    }
    oracle::Hierarchy H;
    if (!oracle::build_mf_hierarchy(m, g.nnx(), g.nny(), g.nnz(), H))
      throw std::invalid_argument("no multigrid hierarchy for this grid");
    if (level < 1 || level >= H.levels()) throw std::invalid_argument("level out of range");
    const oracle::Csr& A = H.coarse[static_cast<std::size_t>(level - 1)].A;
    *n = A.rows;
    *nnz = static_cast<int>(A.col.size());
    if (!ptr) return;
    std::memcpy(ptr, A.ptr.data(), A.ptr.size() * sizeof(int));
    std::memcpy(col, A.col.data(), A.col.size() * sizeof(int));
    std::memcpy(val, A.val.data(), A.val.size() * sizeof(double));
  });
}

int oracle_set_smoother(int kind, int degree) {
  return guarded([&] { oracle::set_smoother(kind, degree); });
}

// lambda_max(D^-1 A) estimates of the Chebyshev smoother per level (0 = fine)
// for this system, from a hierarchy built with the current smoother setting.
int oracle_cheb_lambdas(const OracleProblem* p, const double* E, double nu, int* n_out,
                        double* lambdas) {
  return guarded([&] {
    const oracle::Grid g = grid_of(p);
    std::vector<double> Ev(E, E + g.voxels());
    oracle::Reduced m = oracle::build_reduced(g, Ev, nu, bcs_of(p), loads_of(p), nullptr);
    oracle::Hierarchy H;
    if (!oracle::build_mf_hierarchy(m, g.nnx(), g.nny(), g.nnz(), H))
      throw std::invalid_argument("no multigrid hierarchy for this grid");
    *n_out = static_cast<int>(H.cheb.size());
    for (std::size_t i = 0; i < H.cheb.size() && i < 16; ++i) lambdas[i] = H.cheb[i].lambda;
  });
}

int oracle_simp_compliance(const OracleProblem* p, const double* density, double E0, double nu,
                           double penalty, double rho_min, double tol, int max_it, int use_mg,
                           double* c, double* dc, double* u, OracleInfo* info) {
  return guarded([&] {
    const oracle::Grid g = grid_of(p);
    oracle::SimpParams sp{E0, nu, penalty, rho_min};
    std::vector<double> d(density, density + g.voxels());
    const oracle::SimpCompliance r =
        oracle::simp_compliance(g, sp, d, bcs_of(p), loads_of(p), tol, max_it, use_mg != 0);
    fill_info(r.cg, info);
    *c = r.compliance;
    if (dc) std::memcpy(dc, r.dc.data(), r.dc.size() * sizeof(double));
    if (u) std::memcpy(u, r.u.data(), r.u.size() * sizeof(double));
  });
}

// As oracle_simp_compliance with a warm start (guess: 3 * nodes).
int oracle_simp_compliance_warm(const OracleProblem* p, const double* density, double E0,
                                double nu, double penalty, double rho_min, double tol, int max_it,
                                const double* guess, double* c, double* dc, double* u,
                                OracleInfo* info) {
  return guarded([&] {
    const oracle::Grid g = grid_of(p);
    oracle::SimpParams sp{E0, nu, penalty, rho_min};
    std::vector<double> d(density, density + g.voxels());
    const std::vector<double> gv(guess, guess + 3 * static_cast<std::size_t>(g.nodes()));
    const oracle::SimpCompliance r =
        oracle::simp_compliance(g, sp, d, bcs_of(p), loads_of(p), tol, max_it, true, &gv);
    fill_info(r.cg, info);
    *c = r.compliance;
    if (dc) std::memcpy(dc, r.dc.data(), r.dc.size() * sizeof(double));
    if (u) std::memcpy(u, r.u.data(), r.u.size() * sizeof(double));
  });
}

// adjoint == 0: xPhys = filter_density(x); adjoint != 0: filter_sensitivity(x).
int oracle_filter(const OracleProblem* p, double radius, const double* x, double* out,
                  int adjoint) {
  return guarded([&] {
    const oracle::Grid g = grid_of(p);
    const oracle::DensityFilter f = oracle::make_density_filter(g, radius);
    std::vector<double> xv(x, x + g.voxels());
    const std::vector<double> r = adjoint ? f.filter_sensitivity(xv) : f.filter_density(xv);
    std::memcpy(out, r.data(), r.size() * sizeof(double));
  });
}

// Filter weights: row nnz and Hs per voxel (for stencil checks).
int oracle_filter_rows(const OracleProblem* p, double radius, int* nnz, double* hs) {
  return guarded([&] {
    const oracle::Grid g = grid_of(p);
    const oracle::DensityFilter f = oracle::make_density_filter(g, radius);
    for (std::size_t e = 0; e < f.voxel_count; ++e) {
      nnz[e] = static_cast<int>(f.row_start[e + 1] - f.row_start[e]);
      hs[e] = f.weight_sum[e];
    }
  });
}

int oracle_oc_update(const OracleProblem* p, double radius, const double* x, const double* dc,
                     double volfrac, double move, double rho_min, double* x_new, int* bisections,
                     double* lambda) {
  return guarded([&] {
    const oracle::Grid g = grid_of(p);
    const oracle::DensityFilter f = oracle::make_density_filter(g, radius);
    std::vector<double> xv(x, x + g.voxels()), dv(dc, dc + g.voxels());
    const std::vector<double> r =
        oracle::oc_update(g, f, xv, dv, volfrac, move, rho_min, bisections, lambda);
    std::memcpy(x_new, r.data(), r.size() * sizeof(double));
  });
}

// Plain OC loop. hist receives 4 doubles per iteration (compliance, change,
// volume_fraction, cg_iterations). If p_cont_stages > 0 a continuation ramp
// (p0 -> p1 by step, per_stage iterations each) is armed.
int oracle_optimize(const OracleProblem* p, double E0, double nu, double penalty, double rho_min,
                    double volfrac, double radius, double move, int max_iterations,
                    double change_tol, double cg_tol, int use_mg, double pc_p0, double pc_p1,
                    double pc_step, int pc_per_stage, double* design, double* xphys,
                    double* hist, int* iterations, double* compliance, double* final_volfrac) {
  return guarded([&] {
    const oracle::Grid g = grid_of(p);
    oracle::SimpParams sp{E0, nu, penalty, rho_min};
    oracle::SimpOptions o;
    o.volume_fraction = volfrac;
    o.filter_radius = radius;
    o.move = move;
    o.max_iterations = max_iterations;
    o.change_tol = change_tol;
    o.cg_tolerance = cg_tol;
    o.use_mg = use_mg != 0;
    if (pc_per_stage > 0)
      o.penalty_continuation = oracle::penalty_continuation_ramp(pc_p0, pc_p1, pc_step, pc_per_stage);
    const oracle::SimpResult r = oracle::simp_optimize(g, sp, bcs_of(p), loads_of(p), o);
    if (design) std::memcpy(design, r.design.data(), r.design.size() * sizeof(double));
    if (xphys)
      std::memcpy(xphys, r.physical_density.data(), r.physical_density.size() * sizeof(double));
    for (std::size_t i = 0; i < r.history.size(); ++i) {
      hist[4 * i + 0] = r.history[i].compliance;
      hist[4 * i + 1] = r.history[i].change;
      hist[4 * i + 2] = r.history[i].volume_fraction;
      hist[4 * i + 3] = r.history[i].cg_iterations;
    }
    *iterations = r.iterations;
    *compliance = r.compliance;
    *final_volfrac = r.volume_fraction;
    if (r.non_convergent) throw oracle::SolverNonConvergence("oracle_optimize: non-convergent", 0, 0);
  });
}

// The plain loop with the MMA updater and its plateau stop (window <= 0: the
// change_tol test); other arguments as oracle_optimize.
double oracle_mma_continuation_move(double move, double beta) {
  return oracle::mma_continuation_move(move, beta);
}

int oracle_optimize_mma(const OracleProblem* p, double E0, double nu, double penalty,
                        double rho_min, double volfrac, double radius, double move,
                        int max_iterations, double change_tol, double cg_tol, int window,
                        double plateau_tol, double min_drop, int flat_windows,
                        int mma_projection, double beta0, double beta_max, double eta,
                        const double* adaptive, const std::uint8_t* mask, double* design,
                        double* xphys, double* hist, int* iterations, double* compliance,
                        double* final_volfrac) {
  return guarded([&] {
    const oracle::Grid g = grid_of(p);
    oracle::SimpParams sp{E0, nu, penalty, rho_min};
    oracle::SimpOptions o;
    o.volume_fraction = volfrac;
    o.filter_radius = radius;
    o.move = move;
    o.max_iterations = max_iterations;
    o.change_tol = change_tol;
    o.cg_tolerance = cg_tol;
    o.mma = true;
    o.mma_plateau_window = window;
    o.mma_plateau_tol = plateau_tol;
    o.mma_plateau_min_drop = min_drop;
    o.mma_plateau_flat_windows = flat_windows;
    o.mma_projection = mma_projection != 0;
    o.mma_projection_beta0 = beta0;
    o.mma_projection_beta_max = beta_max;
    o.projection_eta = eta;
    if (adaptive) {  // grow, shrink, osc_lo, osc_hi, max, min
      o.adaptive_move = true;
      o.adaptive_move_grow = adaptive[0];
      o.adaptive_move_shrink = adaptive[1];
      o.adaptive_move_osc_lo = adaptive[2];
      o.adaptive_move_osc_hi = adaptive[3];
      o.adaptive_move_max = adaptive[4];
      o.adaptive_move_min = adaptive[5];
    }
    const oracle::SimpResult r =
        mask ? oracle::simp_optimize_masked(g, sp, bcs_of(p), loads_of(p), o,
                                            std::vector<std::uint8_t>(mask, mask + g.voxels()))
             : oracle::simp_optimize(g, sp, bcs_of(p), loads_of(p), o);
    if (design) std::memcpy(design, r.design.data(), r.design.size() * sizeof(double));
    if (xphys)
      std::memcpy(xphys, r.physical_density.data(), r.physical_density.size() * sizeof(double));
    for (std::size_t i = 0; i < r.history.size(); ++i) {
      hist[4 * i + 0] = r.history[i].compliance;
      hist[4 * i + 1] = r.history[i].change;
      hist[4 * i + 2] = r.history[i].volume_fraction;
      hist[4 * i + 3] = r.history[i].cg_iterations;
    }
    *iterations = r.iterations;
    *compliance = r.compliance;
    *final_volfrac = r.volume_fraction;
    if (r.non_convergent) throw oracle::SolverNonConvergence("oracle_optimize_mma: non-convergent", 0, 0);
  });
}

// n_steps successive mma_update calls (mma_iter 1..n_steps, one state) on the
// given designs x[s] and compliance sensitivities dc[s] (each N long).
int oracle_mma_steps(const OracleProblem* p, double radius, int n_steps, const double* x,
                     const double* dc, double volfrac, double move, double rho_min, double* out,
                     double* lambdas) {
  return guarded([&] {
    const oracle::Grid g = grid_of(p);
    const oracle::DensityFilter f = oracle::make_density_filter(g, radius);
    const std::size_t N = g.voxels();
    oracle::MmaState st;
    for (int s = 0; s < n_steps; ++s) {
      const std::vector<double> xs(x + s * N, x + (s + 1) * N), ds(dc + s * N, dc + (s + 1) * N);
      double lam = 0.0;
      const std::vector<double> xn =
          oracle::mma_update(st, s + 1, g, f, xs, ds, volfrac, move, rho_min, &lam);
      std::memcpy(out + s * N, xn.data(), N * sizeof(double));
      if (lambdas) lambdas[s] = lam;
    }
  });
}

int oracle_mma_plateau(const double* c, int n, int window, double rel_tol, double min_drop,
                       int flat_windows, int* out) {
  return guarded([&] {
    *out = oracle::mma_objective_plateau(std::vector<double>(c, c + n), window, rel_tol, min_drop,
                                         flat_windows) ? 1 : 0;
  });
}

// Projected (Heaviside beta-continuation) OC loop: n_stages stages of
// (beta, move, iterations), threshold eta; other arguments as oracle_optimize.
int oracle_optimize_projected(const OracleProblem* p, const uint8_t* mask, double E0, double nu,
                              double penalty,
                              double rho_min, double volfrac, double radius, int n_stages,
                              const double* betas, const double* moves, const int* iters,
                              double eta, double change_tol, double cg_tol, int use_mg,
                              double* design, double* xphys, double* hist, int* iterations,
                              double* compliance, double* final_volfrac) {
  return guarded([&] {
    const oracle::Grid g = grid_of(p);
    oracle::SimpParams sp{E0, nu, penalty, rho_min};
    oracle::SimpOptions o;
    o.volume_fraction = volfrac;
    o.filter_radius = radius;
    o.change_tol = change_tol;
    o.cg_tolerance = cg_tol;
    o.use_mg = use_mg != 0;
    o.projection_eta = eta;
    for (int s = 0; s < n_stages; ++s) o.projection.push_back({betas[s], moves[s], iters[s]});
    const oracle::SimpResult r =
        mask ? oracle::simp_optimize_masked(g, sp, bcs_of(p), loads_of(p), o,
                                            std::vector<uint8_t>(mask, mask + g.voxels()))
             : oracle::simp_optimize(g, sp, bcs_of(p), loads_of(p), o);
    if (design) std::memcpy(design, r.design.data(), r.design.size() * sizeof(double));
    if (xphys)
      std::memcpy(xphys, r.physical_density.data(), r.physical_density.size() * sizeof(double));
    for (std::size_t i = 0; i < r.history.size(); ++i) {
      hist[4 * i + 0] = r.history[i].compliance;
      hist[4 * i + 1] = r.history[i].change;
      hist[4 * i + 2] = r.history[i].volume_fraction;
      hist[4 * i + 3] = r.history[i].cg_iterations;
    }
    *iterations = r.iterations;
    *compliance = r.compliance;
    *final_volfrac = r.volume_fraction;
    if (r.non_convergent) throw oracle::SolverNonConvergence("oracle_optimize: non-convergent", 0, 0);
  });
}

// ---- masked (passive-region) variants: mask = one MaskValue byte per voxel ----
int oracle_filter_masked(const OracleProblem* p, double radius, const uint8_t* mask,
                         const double* x, double* out, int adjoint) {
  return guarded([&] {
    const oracle::Grid g = grid_of(p);
    const std::vector<uint8_t> eff =
        oracle::effective_design_mask(g, std::vector<uint8_t>(mask, mask + g.voxels()));
    const oracle::DensityFilter f = oracle::make_density_filter(g, radius, eff);
    std::vector<double> xv(x, x + g.voxels());
    const std::vector<double> r = adjoint ? f.filter_sensitivity(xv) : f.filter_density(xv);
    std::memcpy(out, r.data(), r.size() * sizeof(double));
  });
}

int oracle_filter_rows_masked(const OracleProblem* p, double radius, const uint8_t* mask,
                              int* nnz, double* hs) {
  return guarded([&] {
    const oracle::Grid g = grid_of(p);
    const std::vector<uint8_t> eff =
        oracle::effective_design_mask(g, std::vector<uint8_t>(mask, mask + g.voxels()));
    const oracle::DensityFilter f = oracle::make_density_filter(g, radius, eff);
    for (std::size_t e = 0; e < f.voxel_count; ++e) {
      nnz[e] = static_cast<int>(f.row_start[e + 1] - f.row_start[e]);
      hs[e] = f.weight_sum[e];
    }
  });
}

int oracle_oc_update_masked(const OracleProblem* p, double radius, const uint8_t* mask,
                            const double* x, const double* dc, double volfrac, double move,
                            double rho_min, double* x_new, int* bisections, double* lambda) {
  return guarded([&] {
    const oracle::Grid g = grid_of(p);
    const std::vector<uint8_t> eff =
        oracle::effective_design_mask(g, std::vector<uint8_t>(mask, mask + g.voxels()));
    const oracle::DensityFilter f = oracle::make_density_filter(g, radius, eff);
    std::vector<double> xv(x, x + g.voxels()), dv(dc, dc + g.voxels());
    const std::vector<double> r = oracle::oc_update_masked(g, f, eff, xv, dv, volfrac, move,
                                                           rho_min, bisections, lambda);
    std::memcpy(x_new, r.data(), r.size() * sizeof(double));
  });
}

// Masked OC loop; arguments as oracle_optimize plus the mask.
int oracle_optimize_masked(const OracleProblem* p, const uint8_t* mask, double E0, double nu,
                           double penalty, double rho_min, double volfrac, double radius,
                           double move, int max_iterations, double change_tol, double cg_tol,
                           int use_mg, double pc_p0, double pc_p1, double pc_step,
                           int pc_per_stage, double* design, double* xphys, double* hist,
                           int* iterations, double* compliance, double* final_volfrac) {
  return guarded([&] {
    const oracle::Grid g = grid_of(p);
    oracle::SimpParams sp{E0, nu, penalty, rho_min};
    oracle::SimpOptions o;
    o.volume_fraction = volfrac;
    o.filter_radius = radius;
    o.move = move;
    o.max_iterations = max_iterations;
    o.change_tol = change_tol;
    o.cg_tolerance = cg_tol;
    o.use_mg = use_mg != 0;
    if (pc_per_stage > 0)
      o.penalty_continuation = oracle::penalty_continuation_ramp(pc_p0, pc_p1, pc_step, pc_per_stage);
    const oracle::SimpResult r = oracle::simp_optimize_masked(
        g, sp, bcs_of(p), loads_of(p), o, std::vector<uint8_t>(mask, mask + g.voxels()));
    if (design) std::memcpy(design, r.design.data(), r.design.size() * sizeof(double));
    if (xphys)
      std::memcpy(xphys, r.physical_density.data(), r.physical_density.size() * sizeof(double));
    for (std::size_t i = 0; i < r.history.size(); ++i) {
      hist[4 * i + 0] = r.history[i].compliance;
      hist[4 * i + 1] = r.history[i].change;
      hist[4 * i + 2] = r.history[i].volume_fraction;
      hist[4 * i + 3] = r.history[i].cg_iterations;
    }
    *iterations = r.iterations;
    *compliance = r.compliance;
    *final_volfrac = r.volume_fraction;
    if (r.non_convergent) throw oracle::SolverNonConvergence("oracle_optimize: non-convergent", 0, 0);
  });
}

// Self-weight load vector (full, 3*nodes) for the given grid.
int oracle_self_weight(const OracleProblem* p, double density, double gravity, double dx,
                       double dy, double dz, double* f) {
  return guarded([&] {
    const oracle::Grid g = grid_of(p);
    const double d[3] = {dx, dy, dz};
    const auto loads = oracle::self_weight_loads(g, density, gravity, d);
    std::memset(f, 0, 3 * static_cast<std::size_t>(g.nodes()) * sizeof(double));
    for (const auto& l : loads) f[3 * static_cast<std::size_t>(l.node) + l.comp] += l.value;
  });
}

int oracle_penalty_at(double p0, double p1, double step, int per_stage, int iteration,
                      double* out) {
  return guarded([&] {
    *out = oracle::penalty_at_iteration(oracle::penalty_continuation_ramp(p0, p1, step, per_stage),
                                        iteration);
  });
}

int oracle_von_mises(const OracleProblem* p, double E, double nu, const double* u, double* vm) {
  return guarded([&] {
    const oracle::Grid g = grid_of(p);
    const std::vector<double> uv(u, u + 3 * static_cast<std::size_t>(g.nodes()));
    const auto f = oracle::von_mises_field(g, E, nu, uv);
    std::memcpy(vm, f.data(), f.size() * sizeof(double));
  });
}

int oracle_hex8_stress(double E, double nu, double h, const double* ue, double xi, double eta,
                       double zeta, double* sigma6, double* vm) {
  return guarded([&] {
    const oracle::Hex8Stress st = oracle::hex8_stress(E, nu, h, ue, xi, eta, zeta);
    std::memcpy(sigma6, st.sigma, sizeof(st.sigma));
    *vm = st.von_mises;
  });
}

int oracle_float_to_half(const float* in, std::size_t n, std::uint16_t* out) {
  return guarded([&] {
    for (std::size_t i = 0; i < n; ++i) out[i] = oracle::float_to_half(in[i]);
  });
}

std::uint64_t oracle_fnv1a64(const void* data, std::size_t n) { return oracle::fnv1a64(data, n); }
std::uint64_t oracle_fnv1a64_basis(const void* data, std::size_t n, std::uint64_t basis) {
  return oracle::fnv1a64(data, n, basis);
}

int oracle_heaviside(double rt, double beta, double eta, double* value, double* deriv) {
  return guarded([&] {
    *value = oracle::heaviside_project(rt, beta, eta);
    *deriv = oracle::heaviside_project_derivative(rt, beta, eta);
  });
}

}  // extern "C"
