// capi.cpp — extern "C" boundary (include/topopt_hip.h). Converts exceptions
// to status codes and keeps the message per calling thread.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <type_traits>
#include <vector>
#include <random>
#include <string>

#include "engine.hpp"
#include "formats.hpp"

using namespace topohip;

struct topohip_problem {
  std::unique_ptr<Problem> P;
  // bench buffers (device-resident inputs)
  DBuf<double> bu64, by64;
  DBuf<float> bu32, by32;
  bool bench_ready = false;
};

namespace {
thread_local std::string g_err;
int g_device = 0;

template <class F>
int guarded(F&& f) {
  try {
    f();
    return TOPOHIP_OK;
  } catch (const Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const HipError& e) {
    g_err = e.what();
    return TOPOHIP_EDEVICE;
  } catch (const std::bad_alloc& e) {
    g_err = "out of memory";
    return TOPOHIP_EDEVICE;
  } catch (const std::exception& e) {
    g_err = e.what();
    return TOPOHIP_EINTERNAL;
  }
}

void need(const void* p, const char* what) {
  if (!p) throw Error(TOPOHIP_EINVAL, std::string(what) + " must not be NULL");
}

// Upload host youngs (validated > 0 on solid voxels, as mf_build_elems,
// core/src/fea/matfree.cpp:442-448) into P.E64.
void upload_youngs(Problem& P, const double* youngs) {
  const long long nv = P.g.elems();
  for (long long e = 0; e < nv; ++e) {
    if (!P.all_solid && !P.solid_h[e]) continue;
    if (!(youngs[e] > 0.0))
      throw Error(TOPOHIP_EINVAL,
                  "fea_solve_mgcg_matfree: per-voxel Young's modulus must be > 0 on a solid voxel");
  }
  std::vector<double> Eh(youngs, youngs + nv);
  if (!P.all_solid)
    for (long long e = 0; e < nv; ++e)
      if (!P.solid_h[e]) Eh[e] = 0.0;
  TOPOHIP_CHECK(hipMemcpyAsync(P.E64.p, Eh.data(), nv * sizeof(double), hipMemcpyHostToDevice, P.s));
  TOPOHIP_CHECK(hipStreamSynchronize(P.s));
}

template <typename T>
void apply_host(Problem& P, const double* youngs, const double* u, double* y) {
  upload_youngs(P, youngs);
  const long long nd = P.g.ndof(), nv = P.g.elems();
  DBuf<double> u64, y64;
  u64.ensure(nd);
  y64.ensure(nd);
  TOPOHIP_CHECK(hipMemcpyAsync(u64.p, u, nd * sizeof(double), hipMemcpyHostToDevice, P.s));
  KuOp<T> op;
  op.g = P.g;
  op.tz = sizeof(T) == 8 ? P.tz64 : P.tz32;
  op.epi = EPI_AX;
  if constexpr (sizeof(T) == 4) {
    DBuf<float> u32, y32;
    u32.ensure(nd);
    y32.ensure(nd);
    launch_convert<double, float>(P.E64.p, P.E32.p, nv, P.s);
    launch_convert<double, float>(u64.p, u32.p, nd, P.s);
    op.E = P.E32.p;
    op.mg = P.mg32;
    op.x = u32.p;
    op.y = y32.p;
    launch_ku_fine<float>(op, P.s);
    launch_convert<float, double>(y32.p, y64.p, nd, P.s);
    TOPOHIP_CHECK(hipStreamSynchronize(P.s));
  } else {
    op.E = P.E64.p;
    op.mg = P.mg64;
    op.x = u64.p;
    op.y = y64.p;
    launch_ku_fine<double>(op, P.s);
  }
  TOPOHIP_CHECK(hipMemcpyAsync(y, y64.p, nd * sizeof(double), hipMemcpyDeviceToHost, P.s));
  TOPOHIP_CHECK(hipStreamSynchronize(P.s));
}
}  // namespace

namespace {
// Buffer sets cycled through by the microbenchmarks so that their working set
// (sets x bytes per launch) exceeds twice the 256 MiB last-level cache: every
// launch then reads its inputs from HBM, as the design loop's operands at
// these sizes largely are.
int rotation_sets(double bytes_per_launch) {
  const double want = 2.0 * 256.0 * 1024.0 * 1024.0;
  return static_cast<int>(std::min(64.0, std::max(1.0, std::ceil(want / bytes_per_launch))));
}

double time_launches(hipStream_t s, int reps, const std::function<void(int)>& run) {
  hipEvent_t e0, e1;
  TOPOHIP_CHECK(hipEventCreate(&e0));
  TOPOHIP_CHECK(hipEventCreate(&e1));
  run(0);
  TOPOHIP_CHECK(hipEventRecord(e0, s));
  for (int i = 0; i < reps; ++i) run(i + 1);
  TOPOHIP_CHECK(hipEventRecord(e1, s));
  TOPOHIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  TOPOHIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return static_cast<double>(ms) / reps;
}

// One operand set of a fine K.u launch in precision T.
template <typename T>
struct KuSet {
  T *x, *y, *b, *d, *xw, *E;
  double* r;
};
}  // namespace

extern "C" {

const char* topohip_version(void) { return "topopt_hip 0.1 (gfx950)"; }

int topohip_last_error(char* buf, size_t n) {
  if (buf && n > 0) {
    std::strncpy(buf, g_err.c_str(), n - 1);
    buf[n - 1] = '\0';
  }
  return static_cast<int>(g_err.size());
}

int topohip_set_device(int device) {
  return guarded([&] {
    TOPOHIP_CHECK(hipSetDevice(device));
    g_device = device;
  });
}

int topohip_problem_create(int nx, int ny, int nz, double h, double poisson, const uint8_t* solid,
                           int n_bc, const int32_t* bc_node, const int32_t* bc_comp,
                           const double* bc_val, int precision, topohip_problem** out) {
  return guarded([&] {
    need(out, "out");
    if (n_bc > 0) { need(bc_node, "bc_node"); need(bc_comp, "bc_comp"); need(bc_val, "bc_val"); }
    auto h_ = std::make_unique<topohip_problem>();
    h_->P = create_problem(nx, ny, nz, h, poisson, solid, n_bc, bc_node, bc_comp, bc_val,
                           precision, g_device);
    *out = h_.release();
  });
}

int topohip_problem_destroy(topohip_problem* p) {
  return guarded([&] { delete p; });
}

int topohip_set_loads(topohip_problem* p, int n_rhs, const int32_t* counts, const int32_t* node,
                      const int32_t* comp, const double* val) {
  return guarded([&] {
    need(p, "problem");
    need(counts, "counts");
    set_loads(*p->P, n_rhs, counts, node, comp, val);
  });
}

int topohip_ku_apply_f32(topohip_problem* p, const double* youngs, const double* u, double* y) {
  return guarded([&] {
    need(p, "problem"); need(youngs, "youngs"); need(u, "u"); need(y, "y");
    apply_host<float>(*p->P, youngs, u, y);
  });
}

int topohip_ku_apply_f64(topohip_problem* p, const double* youngs, const double* u, double* y) {
  return guarded([&] {
    need(p, "problem"); need(youngs, "youngs"); need(u, "u"); need(y, "y");
    apply_host<double>(*p->P, youngs, u, y);
  });
}

int topohip_solve(topohip_problem* p, const double* youngs, double tol, int max_it,
                  const double* guess, double* u_out, topohip_cginfo* info) {
  return guarded([&] {
    need(p, "problem"); need(youngs, "youngs"); need(u_out, "u_out");
    Problem& P = *p->P;
    upload_youngs(P, youngs);
    Timers t;
    build_hierarchy(P, t);
    const long long nd = P.g.ndof();
    if (guess) {
      P.guess.ensure(nd);
      TOPOHIP_CHECK(hipMemcpyAsync(P.guess.p, guess, nd * sizeof(double), hipMemcpyHostToDevice, P.s));
    }
    topohip_cginfo ci{};
    try {
      solve_cases(P, 1, tol, max_it, guess ? P.guess.p : nullptr, &ci);
    } catch (...) {
      ci.t_mg_build_ms = t.mg_build;
      if (info) *info = ci;
      throw;
    }
    ci.t_mg_build_ms = t.mg_build;
    if (info) *info = ci;
    TOPOHIP_CHECK(hipMemcpyAsync(u_out, P.u.p, nd * sizeof(double), hipMemcpyDeviceToHost, P.s));
    TOPOHIP_CHECK(hipStreamSynchronize(P.s));
  });
}

int topohip_simp_compliance(topohip_problem* p, const double* xphys, double E0, double penalty,
                            double rho_min, double tol, int max_it, double* c, double* dc,
                            double* u_out, topohip_cginfo* info) {
  return guarded([&] {
    need(p, "problem"); need(xphys, "xphys"); need(c, "c");
    Problem& P = *p->P;
    const long long nv = P.g.elems(), nd = P.g.ndof();
    P.xphys.ensure(nv);
    TOPOHIP_CHECK(hipMemcpyAsync(P.xphys.p, xphys, nv * sizeof(double), hipMemcpyHostToDevice, P.s));
    *c = simp_compliance_dev(P, P.xphys.p, E0, penalty, rho_min, tol, max_it, info);
    if (dc)
      TOPOHIP_CHECK(hipMemcpyAsync(dc, P.dc.p, nv * sizeof(double), hipMemcpyDeviceToHost, P.s));
    if (u_out)  // load case 0
      TOPOHIP_CHECK(hipMemcpyAsync(u_out, P.u.p, nd * sizeof(double), hipMemcpyDeviceToHost, P.s));
    TOPOHIP_CHECK(hipStreamSynchronize(P.s));
  });
}

int topohip_simp_compliance_warm(topohip_problem* p, const double* xphys, double E0,
                                 double penalty, double rho_min, double tol, int max_it,
                                 const double* guess, double* c, double* dc, double* u_out,
                                 topohip_cginfo* info) {
  return guarded([&] {
    need(p, "problem"); need(xphys, "xphys"); need(c, "c");
    Problem& P = *p->P;
    const long long nv = P.g.elems(), nd = P.g.ndof();
    P.xphys.ensure(nv);
    TOPOHIP_CHECK(hipMemcpyAsync(P.xphys.p, xphys, nv * sizeof(double), hipMemcpyHostToDevice, P.s));
    if (guess) {
      P.guess.ensure(nd);
      TOPOHIP_CHECK(hipMemcpyAsync(P.guess.p, guess, nd * sizeof(double), hipMemcpyHostToDevice, P.s));
    }
    *c = simp_compliance_dev(P, P.xphys.p, E0, penalty, rho_min, tol, max_it, info,
                             guess ? P.guess.p : nullptr);
    if (dc)
      TOPOHIP_CHECK(hipMemcpyAsync(dc, P.dc.p, nv * sizeof(double), hipMemcpyDeviceToHost, P.s));
    if (u_out)
      TOPOHIP_CHECK(hipMemcpyAsync(u_out, P.u.p, nd * sizeof(double), hipMemcpyDeviceToHost, P.s));
    TOPOHIP_CHECK(hipStreamSynchronize(P.s));
  });
}

int topohip_filter(topohip_problem* p, double radius, const double* x, double* out, int adjoint) {
  return guarded([&] {
    need(p, "problem"); need(x, "x"); need(out, "out");
    Problem& P = *p->P;
    FilterCache& fc = filter_for(P, radius);
    const long long nv = P.g.elems();
    DBuf<double> xi, xo;
    xi.ensure(nv);
    xo.ensure(nv);
    TOPOHIP_CHECK(hipMemcpyAsync(xi.p, x, nv * sizeof(double), hipMemcpyHostToDevice, P.s));
    launch_filter(P.g, fc.st, P.all_solid ? nullptr : P.solid.p, fc.hs.p, xi.p, xo.p, adjoint ? 1 : 0,
                  P.s);
    TOPOHIP_CHECK(hipMemcpyAsync(out, xo.p, nv * sizeof(double), hipMemcpyDeviceToHost, P.s));
    TOPOHIP_CHECK(hipStreamSynchronize(P.s));
  });
}

int topohip_oc_update(topohip_problem* p, double radius, const double* x, const double* dc,
                      double volfrac, double move, double rho_min, double* x_new,
                      uint8_t* active_set, int* bisections, double* lambda) {
  return guarded([&] {
    need(p, "problem"); need(x, "x"); need(dc, "dc"); need(x_new, "x_new");
    Problem& P = *p->P;
    const long long nv = P.g.elems();
    DBuf<double> xi, di, xo;
    DBuf<uint8_t> act;
    xi.ensure(nv); di.ensure(nv); xo.ensure(nv); act.ensure(nv);
    TOPOHIP_CHECK(hipMemcpyAsync(xi.p, x, nv * sizeof(double), hipMemcpyHostToDevice, P.s));
    TOPOHIP_CHECK(hipMemcpyAsync(di.p, dc, nv * sizeof(double), hipMemcpyHostToDevice, P.s));
    oc_update_dev(P, radius, xi.p, di.p, volfrac, move, rho_min, xo.p, act.p, bisections, lambda,
                  nullptr);
    TOPOHIP_CHECK(hipMemcpyAsync(x_new, xo.p, nv * sizeof(double), hipMemcpyDeviceToHost, P.s));
    if (active_set)
      TOPOHIP_CHECK(hipMemcpyAsync(active_set, act.p, nv, hipMemcpyDeviceToHost, P.s));
    TOPOHIP_CHECK(hipStreamSynchronize(P.s));
  });
}

int topohip_oc_update_noalloc(topohip_problem* p, double radius, const double* x,
                              const double* dc, double volfrac, double move, double rho_min,
                              double* x_new, uint8_t* active_set, int* bisections,
                              double* lambda) {
  return guarded([&] {
    need(p, "problem"); need(x, "x"); need(dc, "dc"); need(x_new, "x_new");
    oc_update_noalloc(*p->P, radius, x, dc, volfrac, move, rho_min, x_new, active_set,
                      bisections, lambda);
  });
}

int topohip_oc_margin(topohip_problem* p, double* last_gap, double* run_min_gap,
                      double* bound) {
  return guarded([&] {
    need(p, "problem");
    const Problem& P = *p->P;
    if (last_gap) *last_gap = P.oc_gap_rel;
    if (run_min_gap) *run_min_gap = P.oc_gap_min;
    if (bound) *bound = P.oc_gap_bound;
  });
}

double topohip_set_solve_deadline_ms(double absolute_steady_ms) {
  return set_solve_deadline_ms(absolute_steady_ms);
}

double topohip_steady_ms(void) { return steady_ms(); }

void topohip_reset_mg_stagnation_latch(void) { reset_mg_stagnation_latch(); }

int topohip_mg_stagnation_latched(void) { return mg_stagnation_latched() ? 1 : 0; }

void topohip_set_mg_rearm_period(int period) { set_mg_rearm_period(period); }

int topohip_mg_rearm_period(void) { return mg_rearm_period(); }

long long topohip_mg_rearm_attempts(void) { return mg_rearm_attempts(); }

long long topohip_mg_rearm_carries(void) { return mg_rearm_carries(); }

int topohip_set_smoother(topohip_problem* p, int kind, int degree) {
  return guarded([&] {
    need(p, "problem");
    set_smoother(*p->P, kind, degree);
  });
}

int topohip_smoother_lambdas(topohip_problem* p, int n, double* lambdas, int* levels) {
  return guarded([&] {
    need(p, "problem");
    Problem& P = *p->P;
    int k = 0;
    if (P.smoother == 1 && P.mg_plan_ok) {
      for (size_t l = 0; l + 1 < P.L.size(); ++l, ++k) {
        if (k >= n || !lambdas) continue;
        double v = 0.0;
        TOPOHIP_CHECK(hipMemcpyAsync(&v, P.L[l]->cheb.p, sizeof(double), hipMemcpyDeviceToHost, P.s));
        TOPOHIP_CHECK(hipStreamSynchronize(P.s));
        lambdas[k] = v;
      }
    }
    if (levels) *levels = k;
  });
}

int topohip_set_cancel_flag(topohip_problem* p, const int* flag) {
  return guarded([&] {
    need(p, "problem");
    p->P->cancel = flag;
  });
}

int topohip_optimize(topohip_problem* p, const topohip_options* opt, double* design_out,
                     double* xphys_out, double* history, topohip_result* result) {
  return guarded([&] {
    need(p, "problem"); need(opt, "options");
    optimize(*p->P, *opt, design_out, xphys_out, history, result);
  });
}

int topohip_mma_update(topohip_problem* p, double radius, int mma_iter, const double* x,
                       const double* dc, double volfrac, double move, double rho_min,
                       double* x_new, double* lambda, double* change) {
  return guarded([&] {
    need(p, "problem"); need(x, "x"); need(dc, "dc"); need(x_new, "x_new");
    Problem& P = *p->P;
    const long long nv = P.g.elems();
    DBuf<double> xi, di, xo;
    xi.ensure(nv); di.ensure(nv); xo.ensure(nv);
    TOPOHIP_CHECK(hipMemcpyAsync(xi.p, x, nv * sizeof(double), hipMemcpyHostToDevice, P.s));
    TOPOHIP_CHECK(hipMemcpyAsync(di.p, dc, nv * sizeof(double), hipMemcpyHostToDevice, P.s));
    FilterCache& fc = filter_for(P, radius);
    mma_update_core(P, fc, P.all_solid ? nullptr : P.solid.p, static_cast<double>(P.n_solid),
                    mma_iter, xi.p, di.p, volfrac, move, rho_min, xo.p, lambda, change);
    TOPOHIP_CHECK(hipMemcpyAsync(x_new, xo.p, nv * sizeof(double), hipMemcpyDeviceToHost, P.s));
    TOPOHIP_CHECK(hipStreamSynchronize(P.s));
  });
}

// topohip_mma_options -> MmaOpts (NULL = the reference's defaults, adaptive move off).
MmaOpts mma_opts_of(const topohip_mma_options* mma) {
  MmaOpts m;
  if (!mma) return m;
  m.plateau_window = mma->plateau_window;
  m.plateau_tol = mma->plateau_tol;
  m.plateau_min_drop = mma->plateau_min_drop;
  m.plateau_flat_windows = mma->plateau_flat_windows;
  m.adaptive = mma->adaptive_move != 0;
  if (m.adaptive) {
    m.ad_grow = mma->adaptive_move_grow;
    m.ad_shrink = mma->adaptive_move_shrink;
    m.ad_osc_lo = mma->adaptive_move_osc_lo;
    m.ad_osc_hi = mma->adaptive_move_osc_hi;
    m.ad_max = mma->adaptive_move_max;
    m.ad_min = mma->adaptive_move_min;
  }
  return m;
}

int topohip_optimize_mma(topohip_problem* p, const topohip_options* opt,
                         const topohip_mma_options* mma, const uint8_t* mask, double* design_out,
                         double* xphys_out, double* history, topohip_result* result) {
  return guarded([&] {
    need(p, "problem"); need(opt, "options");
    MmaOpts m = mma_opts_of(mma);
    optimize(*p->P, *opt, design_out, xphys_out, history, result, mask, nullptr, 0.5, &m);
  });
}

int topohip_mma_plateau(const double* c, int n, int window, double rel_tol, double min_drop,
                        int flat_windows, int* out) {
  return guarded([&] {
    need(out, "out");
    if (n > 0) need(c, "c");
    *out = mma_objective_plateau(c, n, window, rel_tol, min_drop, flat_windows) ? 1 : 0;
  });
}

int topohip_optimize_mma_projected(topohip_problem* p, const topohip_options* opt,
                                   const topohip_mma_options* mma, double beta0,
                                   double beta_max, double eta, const uint8_t* mask,
                                   double* design_out, double* xphys_out, double* history,
                                   topohip_result* result) {
  return guarded([&] {
    need(p, "problem"); need(opt, "options");
    MmaOpts m = mma_opts_of(mma);
    m.projection = true;
    m.beta0 = beta0;
    m.beta_max = beta_max;
    optimize(*p->P, *opt, design_out, xphys_out, history, result, mask, nullptr, eta, &m);
  });
}

int topohip_optimize_masked(topohip_problem* p, const uint8_t* mask, const topohip_options* opt,
                            double* design_out, double* xphys_out, double* history,
                            topohip_result* result) {
  return guarded([&] {
    need(p, "problem"); need(mask, "mask"); need(opt, "options");
    optimize(*p->P, *opt, design_out, xphys_out, history, result, mask);
  });
}

int topohip_optimize_projected(topohip_problem* p, const topohip_options* opt, int n_stages,
                               const double* beta, const double* move, const int* iterations,
                               double eta, double* design_out, double* xphys_out, double* history,
                               topohip_result* result) {
  return guarded([&] {
    need(p, "problem"); need(opt, "options");
    if (n_stages < 1) throw Error(TOPOHIP_EINVAL, "simp_optimize: projection needs >= 1 stage");
    need(beta, "beta"); need(move, "move"); need(iterations, "iterations");
    std::vector<ProjStage> st;
    for (int i = 0; i < n_stages; ++i) st.push_back({beta[i], move[i], iterations[i]});
    optimize(*p->P, *opt, design_out, xphys_out, history, result, nullptr, &st, eta);
  });
}

int topohip_optimize_masked_projected(topohip_problem* p, const uint8_t* mask,
                                      const topohip_options* opt, int n_stages, const double* beta,
                                      const double* move, const int* iterations, double eta,
                                      double* design_out, double* xphys_out, double* history,
                                      topohip_result* result) {
  return guarded([&] {
    need(p, "problem"); need(mask, "mask"); need(opt, "options");
    if (n_stages < 1) throw Error(TOPOHIP_EINVAL, "simp_optimize: projection needs >= 1 stage");
    need(beta, "beta"); need(move, "move"); need(iterations, "iterations");
    std::vector<ProjStage> st;
    for (int i = 0; i < n_stages; ++i) st.push_back({beta[i], move[i], iterations[i]});
    optimize(*p->P, *opt, design_out, xphys_out, history, result, mask, &st, eta);
  });
}

int topohip_filter_masked(topohip_problem* p, double radius, const uint8_t* mask,
                          const double* x, double* out, int adjoint) {
  return guarded([&] {
    need(p, "problem"); need(mask, "mask"); need(x, "x"); need(out, "out");
    Problem& P = *p->P;
    std::vector<uint8_t> eff, dset;
    long long n_design = 0;
    design_mask_of(P, mask, eff, dset, n_design);
    const long long nv = P.g.elems();
    DBuf<uint8_t> d_dset;
    d_dset.ensure(static_cast<size_t>(nv));
    TOPOHIP_CHECK(hipMemcpyAsync(d_dset.p, dset.data(), nv, hipMemcpyHostToDevice, P.s));
    std::unique_ptr<FilterCache> fc = make_filter(P, radius, d_dset.p);
    DBuf<double> xi, xo;
    xi.ensure(nv);
    xo.ensure(nv);
    TOPOHIP_CHECK(hipMemcpyAsync(xi.p, x, nv * sizeof(double), hipMemcpyHostToDevice, P.s));
    launch_filter(P.g, fc->st, d_dset.p, fc->hs.p, xi.p, xo.p, adjoint ? 1 : 0, P.s);
    TOPOHIP_CHECK(hipMemcpyAsync(out, xo.p, nv * sizeof(double), hipMemcpyDeviceToHost, P.s));
    TOPOHIP_CHECK(hipStreamSynchronize(P.s));
  });
}

int topohip_bench_prepare(topohip_problem* p, uint64_t seed) {
  return guarded([&] {
    need(p, "problem");
    Problem& P = *p->P;
    const long long nd = P.g.ndof(), nv = P.g.elems();
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> uu(-1.0, 1.0), ur(1e-3, 1.0);
    std::vector<double> u(nd), E(nv);
    for (auto& v : u) v = uu(rng);
    for (auto& v : E) { const double r = ur(rng); v = r * r * r; }
    p->bu64.ensure(nd); p->by64.ensure(nd); p->bu32.ensure(nd); p->by32.ensure(nd);
    TOPOHIP_CHECK(hipMemcpyAsync(p->bu64.p, u.data(), nd * sizeof(double), hipMemcpyHostToDevice, P.s));
    TOPOHIP_CHECK(hipMemcpyAsync(P.E64.p, E.data(), nv * sizeof(double), hipMemcpyHostToDevice, P.s));
    launch_convert<double, float>(p->bu64.p, p->bu32.p, nd, P.s);
    launch_convert<double, float>(P.E64.p, P.E32.p, nv, P.s);
    TOPOHIP_CHECK(hipStreamSynchronize(P.s));
    p->bench_ready = true;
  });
}

int topohip_ku_tile(topohip_problem* p, int* tx, int* ty, int* tz) {
  return guarded([&] {
    need(p, "problem");
    const Problem& P = *p->P;
    if (tx) *tx = kTX;
    if (ty) *ty = kTY;
    if (tz) *tz = P.tz64;
  });
}

int topohip_ku_tiles(topohip_problem* p, int* tz32, int* tz64) {
  return guarded([&] {
    need(p, "problem");
    const Problem& P = *p->P;
    if (tz32) *tz32 = P.tz32;
    if (tz64) *tz64 = P.tz64;
  });
}


int topohip_bench_ku_variant(topohip_problem* p, int variant, int reps, double* ms_per_apply,
                             int* sets_used) {
  return guarded([&] {
    need(p, "problem"); need(ms_per_apply, "ms_per_apply");
    if (!p->bench_ready) throw Error(TOPOHIP_EINVAL, "call topohip_bench_prepare first");
    if (reps < 1) throw Error(TOPOHIP_EINVAL, "reps must be >= 1");
    if (variant < 0 || variant > 5) throw Error(TOPOHIP_EINVAL, "unknown K.u variant");
    Problem& P = *p->P;
    const long long nd = P.g.ndof(), nv = P.g.elems();
    const bool f64 = variant == 1 || variant == 2 || variant == 5;
    const int w = f64 ? 8 : 4;
    // algorithmic bytes of the variant (the per-launch figures of the bench)
    double bytes = 2.0 * w * nd + static_cast<double>(w) * nv;
    if (variant == 3) bytes += 3.0 * w * nd;              // b, dinv read; x0 written
    if (variant == 4) bytes += 3.0 * w * nd + 8.0 * nd;   // x, b, dinv + FP64 r read
    if (variant == 5) bytes += 8.0 * nd + 4.0 * nd;       // p written, FP32 z read
    if (variant == 5) {
      // the case-0 state the direction update reads: beta = 0.5 from nrz
      // partials of 1 and rz = 2 nrz, not after a restart
      const int nrz = ku_partials<float>(P.g);
      launch_fill<double>(P.rzpart.p, 1.0, nrz, P.s);
      CgState st{};
      st.rz = 2.0 * nrz;
      st.rz_new = st.rz;
      TOPOHIP_CHECK(hipMemcpyAsync(P.st.p, &st, sizeof(st), hipMemcpyHostToDevice, P.s));
    }
    const int ns = rotation_sets(bytes);
    if (sets_used) *sets_used = ns;
    // one arena: per set x, y, b, dinv, xw, E in T and r in FP64
    const size_t per = static_cast<size_t>(5 * nd + nv) * w + static_cast<size_t>(nd) * 8;
    DBuf<unsigned char> arena;
    arena.ensure(per * ns);
    auto fill_sets = [&](auto tag) {
      using T = decltype(tag);
      std::vector<KuSet<T>> S(static_cast<size_t>(ns));
      const T* u = f64 ? reinterpret_cast<const T*>(p->bu64.p) : reinterpret_cast<const T*>(p->bu32.p);
      const T* e = f64 ? reinterpret_cast<const T*>(P.E64.p) : reinterpret_cast<const T*>(P.E32.p);
      for (int k = 0; k < ns; ++k) {
        unsigned char* b0 = arena.p + per * k;
        KuSet<T>& q = S[k];
        q.x = reinterpret_cast<T*>(b0);
        q.y = q.x + nd; q.b = q.y + nd; q.d = q.b + nd; q.xw = q.d + nd; q.E = q.xw + nd;
        q.r = reinterpret_cast<double*>(q.E + nv);
        TOPOHIP_CHECK(hipMemcpyAsync(q.x, u, nd * sizeof(T), hipMemcpyDeviceToDevice, P.s));
        TOPOHIP_CHECK(hipMemcpyAsync(q.b, u, nd * sizeof(T), hipMemcpyDeviceToDevice, P.s));
        TOPOHIP_CHECK(hipMemcpyAsync(q.E, e, nv * sizeof(T), hipMemcpyDeviceToDevice, P.s));
        launch_fill<T>(q.d, T(0.5), nd, P.s);
        TOPOHIP_CHECK(hipMemcpyAsync(q.r, p->bu64.p, nd * sizeof(double), hipMemcpyDeviceToDevice, P.s));
        if (variant == 5) launch_fill<float>(reinterpret_cast<float*>(q.xw), 0.25f, nd, P.s);
      }
      return S;
    };
    auto op_of = [&](auto& q, auto mg) {
      using T = std::remove_pointer_t<decltype(q.x)>;
      KuOp<T> op;
      op.g = P.g; op.tz = sizeof(T) == 8 ? P.tz64 : P.tz32; op.mg = mg; op.E = q.E;
      op.mask = P.L[0]->mask.p;
      op.pstride = P.pstride;
      op.y = q.y;
      if (variant == 3) {
        op.epi = EPI_RES0; op.b = q.b; op.dinv = q.d; op.omega_d = 0.6; op.xw = q.xw;
      } else if (variant == 4) {
        op.epi = EPI_JAC; op.x = q.x; op.b = q.b; op.dinv = q.d; op.omega = T(0.6);
        op.partials = P.partials.p; op.rdot = q.r;
      } else {
        op.epi = EPI_AX; op.x = q.x;
        if (variant == 2 || variant == 5) op.partials = P.partials.p;
        if constexpr (sizeof(T) == 8) {
          if (variant == 5) {
            // z (FP32) in the set's x0 space, the new direction to its b
            op.pz32 = reinterpret_cast<const float*>(q.xw);
            op.pw = q.b;
            op.st = P.st.p;
            op.rzp = P.rzpart.p;
            op.nrz = ku_partials<float>(P.g);
          }
        }
      }
      return op;
    };
    if (f64) {
      auto S = fill_sets(double());
      *ms_per_apply = time_launches(P.s, reps, [&](int i) {
        launch_ku_fine<double>(op_of(S[i % ns], P.mg64), P.s);
      });
    } else {
      auto S = fill_sets(float());
      *ms_per_apply = time_launches(P.s, reps, [&](int i) {
        launch_ku_fine<float>(op_of(S[i % ns], P.mg32), P.s);
      });
    }
    TOPOHIP_CHECK(hipStreamSynchronize(P.s));
  });
}

int topohip_bench_ku(topohip_problem* p, int precision, int reps, double* ms_per_apply) {
  return topohip_bench_ku_variant(p, precision == TOPOHIP_PREC_F64 ? 1 : 0, reps, ms_per_apply,
                                  nullptr);
}

// ---- result fields ------------------------------------------------------------

int topohip_bench_fields(topohip_problem* p, int reps, double* ms3) {
  return guarded([&] {
    need(p, "problem"); need(ms3, "ms3");
    if (!p->bench_ready) throw Error(TOPOHIP_EINVAL, "call topohip_bench_prepare first");
    if (reps < 1) throw Error(TOPOHIP_EINVAL, "reps must be >= 1");
    Problem& P = *p->P;
    const long long nv = P.g.elems(), nd = P.g.ndof();
    const double c = 1.0 / ((1.0 + P.nu) * (1.0 - 2.0 * P.nu));
    const VmCoef k{(2.0 / P.h) * 0.125, c * (1.0 - P.nu), c * P.nu, c * (1.0 - 2.0 * P.nu) / 2.0};
    const uint8_t* sol = P.all_solid ? nullptr : P.solid.p;
    // per rotated set: the FP64 input (displacement, or density in its first
    // nv entries) and the three outputs
    const size_t per = static_cast<size_t>(nd) * 8 + static_cast<size_t>(nv) * 4 +
                       static_cast<size_t>(nd) * 4 + static_cast<size_t>(nv) * 2 + 256;
    const int ns = rotation_sets(static_cast<double>(12 * nd));
    DBuf<unsigned char> arena;
    arena.ensure(per * ns);
    std::vector<double*> in(ns);
    std::vector<float*> vm(ns), du(ns);
    std::vector<uint16_t*> h16(ns);
    for (int s = 0; s < ns; ++s) {
      unsigned char* b0 = arena.p + per * s;
      in[s] = reinterpret_cast<double*>(b0);
      vm[s] = reinterpret_cast<float*>(in[s] + nd);
      du[s] = vm[s] + nv;
      h16[s] = reinterpret_cast<uint16_t*>(du[s] + nd);
      TOPOHIP_CHECK(hipMemcpyAsync(in[s], p->bu64.p, nd * sizeof(double), hipMemcpyDeviceToDevice, P.s));
    }
    for (int which = 0; which < 3; ++which) {
      ms3[which] = time_launches(P.s, reps, [&](int i) {
        const int s = i % ns;
        if (which == 0) launch_von_mises(P.g, k, sol, in[s], nullptr, vm[s], P.s);
        else if (which == 1) launch_narrow_f32(in[s], du[s], nd, P.s);
        else launch_density_f16(in[s], h16[s], nv, P.s);
      });
    }
    TOPOHIP_CHECK(hipStreamSynchronize(P.s));
  });
}

int topohip_von_mises_field(topohip_problem* p, double youngs, const double* u, double* vm_out) {
  return guarded([&] {
    need(p, "problem"); need(vm_out, "vm_out");
    result_fields(*p->P, youngs, u, vm_out, nullptr, nullptr);
  });
}

int topohip_fields_f32(topohip_problem* p, double youngs, const double* u, float* vm_out,
                       float* disp_out) {
  return guarded([&] {
    need(p, "problem");
    result_fields(*p->P, youngs, u, nullptr, vm_out, disp_out);
  });
}

int topohip_density_f16(topohip_problem* p, const double* density, uint16_t* out) {
  return guarded([&] {
    need(p, "problem"); need(out, "out");
    density_halves(*p->P, density, out);
  });
}

int topohip_set_iteration_observer(topohip_problem* p, topohip_iteration_cb cb, void* user) {
  return guarded([&] {
    need(p, "problem");
    p->P->observer = cb;
    p->P->observer_user = cb ? user : nullptr;
  });
}

// ---- run records ------------------------------------------------------------------

const char* topohip_iterations_csv_header(void) { return kIterationCsvHeader; }

int topohip_iterations_csv_row(size_t rung, const topohip_iteration_obs* obs, long long wall_ms,
                               char* buf, size_t n, size_t* len) {
  return guarded([&] {
    need(obs, "obs"); need(buf, "buf");
    const std::string row = iterations_csv_row(rung, *obs, wall_ms);
    if (row.size() + 1 > n) throw Error(TOPOHIP_EINVAL, "buffer too small for the row");
    std::memcpy(buf, row.c_str(), row.size() + 1);
    if (len) *len = row.size();
  });
}

}  // extern "C"

struct topohip_csv_writer {
  std::unique_ptr<CsvWriter> w;
};

extern "C" {

int topohip_csv_open(const char* path, topohip_csv_writer** out) {
  return guarded([&] {
    need(path, "path"); need(out, "out");
    auto h = std::make_unique<topohip_csv_writer>();
    h->w = std::make_unique<CsvWriter>(path);
    *out = h.release();
  });
}

int topohip_csv_append(topohip_csv_writer* w, size_t rung, const topohip_iteration_obs* obs) {
  return guarded([&] {
    need(w, "writer"); need(obs, "obs");
    w->w->append_at(rung, *obs, wall_clock_ms());
  });
}

int topohip_csv_append_at(topohip_csv_writer* w, size_t rung, const topohip_iteration_obs* obs,
                          long long wall_ms) {
  return guarded([&] {
    need(w, "writer"); need(obs, "obs");
    w->w->append_at(rung, *obs, wall_ms);
  });
}

size_t topohip_csv_rows(const topohip_csv_writer* w) { return w ? w->w->rows() : 0; }

int topohip_csv_close(topohip_csv_writer* w) {
  return guarded([&] { delete w; });
}

int topohip_write_density_snapshot(topohip_problem* p, const char* dir, const char* basename,
                                   const double* origin, const double* density, size_t rung,
                                   int iteration, int boundary) {
  return guarded([&] {
    need(p, "problem"); need(dir, "dir"); need(basename, "basename");
    Problem& P = *p->P;
    std::vector<uint16_t> halves(static_cast<size_t>(P.g.elems()));
    density_halves(P, density, halves.data());
    const double zero[3] = {0.0, 0.0, 0.0};
    write_density_snapshot(dir, basename, P.g.ex, P.g.ey, P.g.ez, P.h, origin ? origin : zero,
                           halves, rung, iteration, boundary != 0);
  });
}

}  // extern "C"

struct topohip_snapshots {
  std::unique_ptr<SnapshotCapture> cap;
};

extern "C" {

int topohip_snapshots_create(const char* dir, int every_n, int cap, topohip_snapshots** out) {
  return guarded([&] {
    need(dir, "dir"); need(out, "out");
    auto h = std::make_unique<topohip_snapshots>();
    h->cap = std::make_unique<SnapshotCapture>(dir, every_n, cap);
    *out = h.release();
  });
}

int topohip_snapshots_capture(topohip_snapshots* sc, topohip_problem* p, const double* origin,
                              const double* density, size_t rung, int iteration, int boundary,
                              int* written) {
  return guarded([&] {
    need(sc, "snapshots"); need(p, "problem");
    const bool b = boundary != 0;
    const bool w = sc->cap->wants(iteration, b);
    if (written) *written = w ? 1 : 0;
    if (!w) return;
    Problem& P = *p->P;
    std::vector<uint16_t> halves(static_cast<size_t>(P.g.elems()));
    density_halves(P, density, halves.data());
    const double zero[3] = {0.0, 0.0, 0.0};
    const std::string base = SnapshotCapture::basename(rung, iteration, b);
    write_density_snapshot(sc->cap->dir(), base, P.g.ex, P.g.ey, P.g.ez, P.h,
                           origin ? origin : zero, halves, rung, iteration, b);
    sc->cap->written(base, b);
  });
}

int topohip_snapshots_counts(const topohip_snapshots* sc, size_t* written, size_t* evicted) {
  return guarded([&] {
    need(sc, "snapshots");
    if (written) *written = sc->cap->n_written();
    if (evicted) *evicted = sc->cap->n_evicted();
  });
}

int topohip_snapshots_destroy(topohip_snapshots* sc) {
  return guarded([&] { delete sc; });
}

int topohip_read_density_f16(const char* path, size_t count, float* out) {
  return guarded([&] {
    need(path, "path"); need(out, "out");
    const std::vector<float> v = read_density_f16(path, count);
    std::copy(v.begin(), v.end(), out);
  });
}

uint64_t topohip_design_fingerprint(const double* density, size_t n) {
  return density ? design_fingerprint(density, n) : design_fingerprint(nullptr, 0);
}

int topohip_write_design_file(const char* path, const topohip_grid_frame* grid, int n_variants,
                              const topohip_design_variant* variants) {
  return guarded([&] {
    need(path, "path"); need(grid, "grid");
    write_design_file(path, *grid, n_variants, variants);
  });
}

int topohip_read_design_header(const char* path, topohip_grid_frame* grid, int* n_variants) {
  return guarded([&] {
    need(path, "path");
    const DesignFile f = read_design_file(path);
    if (grid) *grid = f.grid;
    if (n_variants) *n_variants = static_cast<int>(f.meta.size());
  });
}

int topohip_read_design_variant(const char* path, int index, topohip_design_variant* meta,
                                double* density_out) {
  return guarded([&] {
    need(path, "path");
    const DesignFile f = read_design_file(path);
    if (index < 0 || index >= static_cast<int>(f.meta.size()))
      throw Error(TOPOHIP_EINVAL, "design.bin: no variant " + std::to_string(index));
    if (meta) {
      *meta = f.meta[static_cast<size_t>(index)];
      meta->density = nullptr;
    }
    if (density_out) {
      const auto& d = f.density[static_cast<size_t>(index)];
      std::copy(d.begin(), d.end(), density_out);
    }
  });
}

int topohip_write_fields_file(const char* path, const topohip_grid_frame* grid, int n_variants,
                              const topohip_fields_variant* variants) {
  return guarded([&] {
    need(path, "path"); need(grid, "grid");
    write_fields_file(path, *grid, n_variants, variants);
  });
}

namespace {
int copy_line(const std::string& line, char* buf, size_t n) {
  if (!buf || line.size() + 1 > n) {
    g_err = "buffer too small for the line";
    return TOPOHIP_EINVAL;
  }
  std::memcpy(buf, line.c_str(), line.size() + 1);
  return TOPOHIP_OK;
}
}  // namespace

int topohip_format_progress(size_t rung, size_t rungs, int iteration, char* buf, size_t n) {
  return copy_line(progress_line(rung, rungs, iteration), buf, n);
}

int topohip_run_info_json(const topohip_run_info* info, char* buf, size_t n, size_t* len) {
  return guarded([&] {
    need(info, "info");
    if (info->n_rungs < 0 || info->n_ladder < 0 || (info->n_ladder > 0 && !info->ladder))
      throw Error(TOPOHIP_EINVAL, "run_info: bad ladder or rung counts");
    const std::string js = run_info_json(*info);
    if (len) *len = js.size();
    if (!buf || js.size() + 1 > n) throw Error(TOPOHIP_EINVAL, "buffer too small for run_info");
    std::memcpy(buf, js.c_str(), js.size() + 1);
  });
}

int topohip_write_run_info(const char* path, const topohip_run_info* info) {
  return guarded([&] {
    need(path, "path"); need(info, "info");
    if (info->n_rungs < 0 || info->n_ladder < 0 || (info->n_ladder > 0 && !info->ladder))
      throw Error(TOPOHIP_EINVAL, "run_info: bad ladder or rung counts");
    const std::string js = run_info_json(*info);
    std::FILE* f = std::fopen(path, "wb");
    if (!f) throw Error(TOPOHIP_EINVAL, std::string("write_run_info: cannot open ") + path);
    const size_t w = std::fwrite(js.data(), 1, js.size(), f);
    const int c = std::fclose(f);
    if (w != js.size() || c != 0)
      throw Error(TOPOHIP_EINVAL, std::string("write_run_info: failed writing ") + path);
  });
}

int topohip_format_variant(double vf, double achieved, double printed, double margin,
                           const char* mesh, char* buf, size_t n) {
  return copy_line(variant_line(vf, achieved, printed, margin, mesh ? mesh : ""), buf, n);
}

}  // extern "C"
