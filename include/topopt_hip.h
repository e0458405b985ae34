/* topopt_hip.h — C ABI of the MI355X voxel-SIMP engine (libtopopt_hip.so).
 *
 * Drop-in boundary for the reference core's matrix-free multigrid solve and
 * the SIMP inner loop around it. Every entry point below names the reference
 * interface it replaces (paths relative to the reference checkout). Plain C
 * types only: host buffers are caller-owned, device buffers are handle-owned.
 *
 * Numbering is the reference's: voxel (i,j,k) -> (k*ny + j)*nx + i; node
 * (a,b,c) -> (c*(ny+1) + b)*(nx+1) + a; DOF = 3*node + component (x,y,z);
 * displacement vectors are interleaved xyz FP64 of length 3*(nx+1)(ny+1)(nz+1)
 * (core/include/topopt/voxel.hpp:64-69, core/src/fea/assembly.cpp:41-54).
 *
 * Status codes map 1:1 onto the reference exception taxonomy
 * (core/include/topopt/fea.hpp:418-469):
 *   TOPOHIP_EINVAL           std::invalid_argument (sizes, indices, material)
 *   TOPOHIP_EUNDERCONSTRAINED std::runtime_error from the void-DOF gate
 *   TOPOHIP_ENONCONVERGENCE  SolverNonConvergence (iterations/residual in info)
 *   TOPOHIP_EDEADLINE        SolverDeadlineExceeded
 *   TOPOHIP_EDEVICE          a HIP runtime failure
 * topohip_last_error() returns the message of the calling thread's last error.
 *
 * Threading: one handle per host thread. A problem handle owns one HIP stream;
 * calls on different handles may run concurrently.
 */
#ifndef TOPOPT_HIP_H
#define TOPOPT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  TOPOHIP_OK = 0,
  TOPOHIP_EINVAL = 1,
  TOPOHIP_EUNDERCONSTRAINED = 2,
  TOPOHIP_ENONCONVERGENCE = 3,
  TOPOHIP_EDEADLINE = 4,
  TOPOHIP_EDEVICE = 5,
  TOPOHIP_EINTERNAL = 6
};

/* Solver precision of a problem handle.
 *   TOPOHIP_PREC_MIXED: the multigrid V-cycle in FP32 (its K.u applies,
 *     stencils, transfers; FP64 coarsest solve) preconditioning an FP64 PCG
 *     (A.p, x, r, p and the dot products FP64), with an FP64 true-residual
 *     check before convergence is declared — the reference's mixed-precision
 *     MG-PCG (core/src/fea/multigrid.cpp:1211-1318).
 *   TOPOHIP_PREC_F64: every operator apply and the whole V-cycle in FP64 (the
 *     plain FP64 form). */
enum { TOPOHIP_PREC_MIXED = 0, TOPOHIP_PREC_F64 = 1 };

typedef struct topohip_problem topohip_problem;
typedef struct topohip_server topohip_server;

/* Subset of CgInfo (core/include/topopt/fea.hpp:272-381) plus the engine's own
 * counters. */
typedef struct {
  int converged;
  int iterations;            /* outer PCG iterations */
  int used_multigrid;
  int mg_levels;
  int hier_built;
  int mg_cycles_attempted;
  double residual;           /* final ||b - A x|| / ||b|| (FP64 true residual) */
  long long matvecs;         /* fine-level operator applies */
  double t_build_ms, t_mg_build_ms, t_mg_ms, t_cg_ms, t_total_ms;
  int residual_replacements; /* mixed form: FP64 true-residual restarts */
  int coarse_dofs;           /* size of the coarsest (direct) level */
} topohip_cginfo;

/* Create a problem on device 0 (or the device set by topohip_set_device).
 * Replaces the setup half of mf_build_reduced (core/src/fea/matfree.cpp:670-834):
 * grid, Poisson ratio, element size, solid tags (NULL = every voxel solid) and
 * Dirichlet BCs (node, component, value) as in DirichletBC
 * (core/include/topopt/fea.hpp:180-184). Validates indices like the reference
 * (TOPOHIP_EINVAL on a BC index out of range or a bad material). */
int topohip_problem_create(int nx, int ny, int nz, double h, double poisson,
                           const uint8_t* solid, int n_bc, const int32_t* bc_node,
                           const int32_t* bc_comp, const double* bc_val, int precision,
                           topohip_problem** out);
int topohip_problem_destroy(topohip_problem* p);

/* Load cases: n_rhs right-hand sides; case r holds counts[r] consecutive
 * (node, comp, value) entries of the concatenated arrays, summed per DOF as
 * NodalLoad vectors are (core/include/topopt/fea.hpp:188-192;
 * core/src/fea/matfree.cpp:726-732). */
int topohip_set_loads(topohip_problem* p, int n_rhs, const int32_t* counts,
                      const int32_t* node, const int32_t* comp, const double* val);

/* y = K(E) u over the full stiffness of the solid voxels (no BCs), matrix-free.
 * Mirrors fea_matfree_apply's graded overload (core/include/topopt/fea.hpp:
 * 763-766; core/src/fea/matfree.cpp:1085-1104): youngs per voxel (> 0 on solid
 * voxels), u and y of length 3*nodes. _f32 computes in FP32, _f64 in FP64. */
int topohip_ku_apply_f32(topohip_problem* p, const double* youngs, const double* u,
                         double* y);
int topohip_ku_apply_f64(topohip_problem* p, const double* youngs, const double* u,
                         double* y);

/* Multigrid-preconditioned CG solve of the BC-reduced, void-gated system for
 * load case 0. Mirrors fea_solve_mgcg_matfree's graded overload
 * (core/include/topopt/fea.hpp:707-715; core/src/fea/multigrid.cpp:2098-2319):
 * stop on ||r|| <= tol*||b||; max_it <= 0 selects max(1000, 2*ng) capped at the
 * 300-cycle MG budget, then the Jacobi-CG fallback; `guess` (NULL = cold start)
 * of the wrong length is ignored exactly like initial_guess. u_out receives the
 * full displacement with prescribed values scattered in. */
int topohip_solve(topohip_problem* p, const double* youngs, double tol, int max_it,
                  const double* guess, double* u_out, topohip_cginfo* info);

/* Compliance and sensitivity: simp_compliance's MultigridCG_Matfree route and
 * sensitivity sweep (core/src/simp/simp.cpp:567-828). E_e = clamp(xphys_e,
 * rho_min, 1)^penalty * E0; c = sum E_e u_e^T K1 u_e; dc_e = -p rho^(p-1) E0 q_e.
 * With n_rhs > 1 load cases the solves run as one block PCG and c, dc are the
 * sums over cases. u_out (nullable) receives load case 0's displacement. */
int topohip_simp_compliance(topohip_problem* p, const double* xphys, double E0,
                            double penalty, double rho_min, double tol, int max_it,
                            double* c, double* dc, double* u_out, topohip_cginfo* info);

/* simp_compliance with a warm start (its initial_guess argument,
 * core/src/simp/simp.cpp:567-828, as the PLSM state solve calls it at a loose
 * tolerance, core/src/simp/plsm.cpp:671-689): guess (nullable) is a full
 * displacement, 3 * nodes, used as the MG-PCG start on the free DOFs (a
 * non-finite entry makes the solve cold, as the reference's handling). One load
 * case only (TOPOHIP_EINVAL otherwise). Other arguments as
 * topohip_simp_compliance. */
int topohip_simp_compliance_warm(topohip_problem* p, const double* xphys, double E0,
                                 double penalty, double rho_min, double tol, int max_it,
                                 const double* guess, double* c, double* dc, double* u_out,
                                 topohip_cginfo* info);

/* Density filter with cone weights max(0, r - dist) over solid voxels
 * (make_density_filter / filter_density / filter_sensitivity,
 * core/src/simp/simp.cpp:834-944). adjoint == 0: out = H x / Hs;
 * adjoint != 0: out = H^T (x / Hs). FP64. */
int topohip_filter(topohip_problem* p, double radius, const double* x, double* out,
                   int adjoint);

/* Optimality-criteria update with lambda bisection on the filtered volume
 * (oc_update, core/src/simp/simp.cpp:948-1032). active_set (nullable) receives
 * per voxel 0 = interior, 1 = at the lower clamp, 2 = at the upper clamp.
 * bisections / lambda (nullable) report the search. FP64 throughout. */
int topohip_oc_update(topohip_problem* p, double radius, const double* x, const double* dc,
                      double volfrac, double move, double rho_min, double* x_new,
                      uint8_t* active_set, int* bisections, double* lambda);

/* Allocation-free variant of topohip_oc_update. No reference counterpart: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). x, dc, x_new
 * and active_set (nullable) are DEVICE pointers on the handle's device, nx*ny*nz
 * entries each, ready when the call starts; the call returns when x_new is
 * written. The work memory is one static device array per GPU (about 16 MiB,
 * held until the process ends) plus the handle's own buffers, so the call
 * makes no hipMalloc/hipFree. Limits: at most 2^21 (2097152) voxels, and the
 * density filter of `radius` must already be on the handle (built at set-up
 * by topohip_filter or topohip_oc_update with that radius); otherwise
 * TOPOHIP_EINVAL. Every call on one GPU shares the array: calls that overlap
 * on one GPU race. Results are bit-identical to topohip_oc_update. */
int topohip_oc_update_noalloc(topohip_problem* p, double radius, const double* x,
                              const double* dc, double volfrac, double move, double rho_min,
                              double* x_new, uint8_t* active_set, int* bisections,
                              double* lambda);

/* Decision margin of the OC volume bisection (the test `vol > target` of
 * oc_update, core/src/simp/simp.cpp:1002-1031): the smallest |vol - target|
 * over the volumes the last search tested (last_gap) and over every OC step
 * since the last topohip_optimize started (run_min_gap), both relative to the
 * target, and the rounding bound (n - 1) * 2^-53 of a sum of the n
 * non-negative per-voxel terms (bound). A gap above the bound decides the same
 * way whatever the summation order, so the device's block-tree volume and the
 * reference's sequential sum take the same bisection path. Outputs nullable.
 * No reference counterpart (a diagnostic of the device's reduction order). */
int topohip_oc_margin(topohip_problem* p, double* last_gap, double* run_min_gap, double* bound);

/* The plain OC design loop (simp_optimize, core/src/simp/simp.cpp:2018-2608):
 * filter -> solve -> sensitivity -> OC, change = max|x_new - x|, stop when
 * change < change_tol, then a final solve on filter(design). Optional
 * penalty continuation ramp pc_p0 -> pc_p1 by pc_step, pc_per_stage iterations
 * each (penalty_continuation_ramp, core/src/simp/simp.cpp:402-427); pc_per_stage
 * 0 disables it. */
typedef struct {
  double volfrac, rmin, move, penalty, E0, rho_min, change_tol, cg_tol;
  int max_iterations, cg_max_iterations;
  double pc_p0, pc_p1, pc_step;
  int pc_per_stage;
} topohip_options;

typedef struct {
  int iterations, converged, non_convergent;
  double compliance, initial_compliance, volume_fraction;
  double t_total_ms, t_solve_ms, t_build_ms, t_filter_ms, t_sens_ms, t_oc_ms;
  long long cg_iterations_total;
  int cancelled;          /* stopped by the cancel flag (topohip_set_cancel_flag) */
  int deadline_exceeded;  /* a solve passed the thread's solve deadline */
} topohip_result;

/* history (nullable) receives 4 doubles per iteration: compliance, change,
 * volume fraction, outer PCG iterations. */
int topohip_optimize(topohip_problem* p, const topohip_options* opt, double* design_out,
                     double* xphys_out, double* history, topohip_result* result);

/* ---- MMA updater (SimpUpdater::MMA, core/include/topopt/simp.hpp:622-639) ---- */

/* The plateau stop of an MMA run (SimpOptions::mma_plateau_*,
 * core/include/topopt/simp.hpp:1256-1276); plateau_window <= 0 keeps the
 * change_tol test. Defaults: 10, 1e-3, 0.05, 3.
 * adaptive_move != 0 turns on SimpOptions::adaptive_move (simp.hpp:1108-1118;
 * core/src/simp/simp.cpp:140-163, 262-300, 2179-2194): from the third MMA step
 * of a stage the move limit is multiplied by adaptive_move_shrink when the
 * fraction of design voxels whose last two steps reversed direction (among
 * those that moved on both) exceeds adaptive_move_osc_hi, by
 * adaptive_move_grow when it is below adaptive_move_osc_lo, and clamped to
 * [adaptive_move_min, adaptive_move_max]; the first two steps use the stage's
 * move. Reference defaults 1.2, 0.7, 0.15, 0.30, 0.5, 0.02; invalid values are
 * TOPOHIP_EINVAL. Zero-filled = off. */
typedef struct {
  int plateau_window;
  double plateau_tol, plateau_min_drop;
  int plateau_flat_windows;
  int adaptive_move;
  double adaptive_move_grow, adaptive_move_shrink, adaptive_move_osc_lo, adaptive_move_osc_hi;
  double adaptive_move_max, adaptive_move_min;
} topohip_mma_options;

/* simp_optimize with options.updater = MMA (core/src/simp/simp.cpp:2305-2311):
 * the unprojected loop with mma_update as the design update and
 * stage_should_stop (:1219-1232) as its stop rule. mask (nullable) selects the
 * masked loop as topohip_optimize_masked, with mma_update_masked
 * (:2736-2783, :3351-3357): only Active voxels move, the others come back 0.
 * mma NULL = the reference's defaults. Outputs as topohip_optimize. */
int topohip_optimize_mma(topohip_problem* p, const topohip_options* opt,
                         const topohip_mma_options* mma, const uint8_t* mask, double* design_out,
                         double* xphys_out, double* history, topohip_result* result);
/* simp_optimize with options.updater = MMA and options.mma_projection
 * (core/include/topopt/simp.hpp:1488-1500; one dynamic stage, build_stage_plan
 * core/src/simp/simp.cpp:1036-1068): each iteration projects the filtered
 * design at the current beta (project_solid, or project_active with a mask),
 * solves, and takes mma_update_projected (:1509-1636; masked
 * mma_update_masked_projected :2861-2990: both sensitivities chained through
 * the projection derivative before the filter transpose, dc normalised, the
 * constraint on the projected volume) with the move damped to
 * move * min(1, 8 / beta) (mma_continuation_move, :251-254). A plateau of the
 * current beta stage's compliance (mma_objective_plateau, the progress gate on
 * the first stage only) doubles beta, capped at beta_max, and restarts the MMA
 * asymptotes; a plateau at the cap ends the run (:2461-2491); opt->max_iterations
 * is the safety cap. The reported per-iteration volume fraction is the filtered
 * design's; the final physical density is projected at the last beta.
 * TOPOHIP_EINVAL: beta0 not finite and > 0, beta_max < beta0 or not finite,
 * eta outside (0, 1). mma, mask and outputs as topohip_optimize_mma. */
int topohip_optimize_mma_projected(topohip_problem* p, const topohip_options* opt,
                                   const topohip_mma_options* mma, double beta0,
                                   double beta_max, double eta, const uint8_t* mask,
                                   double* design_out, double* xphys_out, double* history,
                                   topohip_result* result);

/* One mma_update (core/src/simp/simp.cpp:1433-1488, core/src/simp/mma_joint.cpp:
 * 18-118) on the handle's solid voxels with the density filter of `radius`:
 * mma_iter is 1-based and the handle keeps the state (asymptotes, last two
 * designs) between calls, as MmaState does; restart a run at mma_iter 1.
 * lambda (nullable): the volume multiplier; change (nullable): max |x_new - x|. */
int topohip_mma_update(topohip_problem* p, double radius, int mma_iter, const double* x,
                       const double* dc, double volfrac, double move, double rho_min,
                       double* x_new, double* lambda, double* change);

/* mma_objective_plateau (core/src/simp/simp.cpp:1081-1134) on a compliance
 * history c[0..n); *out = 1 when it fires. */
int topohip_mma_plateau(const double* c, int n, int window, double rel_tol, double min_drop,
                        int flat_windows, int* out);

/* The design loop with Heaviside projection stages (SimpOptions::projection /
 * projection_eta, core/include/topopt/simp.hpp:518-524; loop
 * core/src/simp/simp.cpp:2018-2608 with build_stage_plan :1053-1068): stage s
 * runs iterations[s] OC updates at sharpness beta[s] with move limit move[s]
 * (opt->move and opt->max_iterations are not used). Each iteration projects
 * the filtered design (project_solid, :1637-1645), solves, and takes the
 * projected OC step (oc_update_projected, :1647-1716: sensitivities chained
 * through the projection derivative before the filter transpose, unclamped dv,
 * a non-finite ratio holds the voxel, volume on the projected candidate);
 * change < change_tol ends a stage. The final physical density is projected
 * at the last stage's beta. The reference schedule of its Gate V2 test is
 * heaviside_continuation_schedule (:367-373). history needs
 * 4 * sum(iterations) doubles. */
int topohip_optimize_projected(topohip_problem* p, const topohip_options* opt, int n_stages,
                               const double* beta, const double* move, const int* iterations,
                               double eta, double* design_out, double* xphys_out, double* history,
                               topohip_result* result);

/* ---- passive regions: the masked design loop ------------------------------
 * mask: one byte per voxel, the reference's MaskValue (core/include/topopt/
 * voxel.hpp:42-46): 0 Active, 1 FrozenSolid, 2 FrozenVoid.
 *
 * topohip_optimize_masked mirrors simp_optimize(grid, params, bcs, loads,
 * options, mask) (core/src/simp/simp.cpp:2989-3689) with the OC updater and
 * one unprojected stage: the effective mask (effective_design_mask, :2610-2633:
 * empty voxels -> FrozenVoid), the Active-only filter (make_density_filter
 * with mask, :834-890, no bleed across the mask boundary), pins after every
 * filter (apply_mask_pins, :2649-2654: FrozenSolid -> 1, FrozenVoid -> 0),
 * oc_update_masked (:2662-2734: frozen voxels keep x, budget
 * volfrac * n_active), volume fraction over the Active voxels. The handle must
 * be created on the analysis grid (analysis_grid_for_mask, :2640-2645):
 * FrozenVoid voxels empty; a FrozenVoid voxel that is solid in the handle's
 * grid is TOPOHIP_EINVAL. Load/Fixture voxels are the caller's FrozenSolid.
 * design_out / xphys_out / history / result as topohip_optimize (xphys_out
 * pinned). */
int topohip_optimize_masked(topohip_problem* p, const uint8_t* mask, const topohip_options* opt,
                            double* design_out, double* xphys_out, double* history,
                            topohip_result* result);
/* The masked loop with Heaviside projection stages: simp_optimize(grid,
 * params, bcs, loads, options, mask) with options.projection
 * (core/src/simp/simp.cpp:2989-3689; the projected branch :3196-3204,
 * :3306-3313, :3411-3413, :3609-3611). Each iteration filters with the
 * Active-only filter, projects the Active voxels (project_active,
 * :2789-2794), pins, solves, and takes oc_update_masked_projected
 * (:2801-2855: the chain rule through the projection on the Active set, the
 * Active-only filter transposes, unclamped dv with 0/0 holding the voxel,
 * frozen voxels keep x, the volume constraint on the projected Active density
 * with target volfrac * n_active); the reported volume fraction is the
 * projected Active density's. Stages, history and result as
 * topohip_optimize_projected; the handle and mask as topohip_optimize_masked
 * (xphys_out pinned). */
int topohip_optimize_masked_projected(topohip_problem* p, const uint8_t* mask,
                                      const topohip_options* opt, int n_stages, const double* beta,
                                      const double* move, const int* iterations, double eta,
                                      double* design_out, double* xphys_out, double* history,
                                      topohip_result* result);
/* make_density_filter(grid, radius, mask) applied to x (adjoint = 0:
 * filter_density; 1: filter_sensitivity), core/src/simp/simp.cpp:834-944. */
int topohip_filter_masked(topohip_problem* p, double radius, const uint8_t* mask,
                          const double* x, double* out, int adjoint);

/* ---- device-resident benchmarking entry (inputs already in HBM) -------------
 * Runs `reps` back-to-back applies of the fine operator on copies of the
 * handle's bench inputs (random u, E = rho^3 set by topohip_bench_prepare) and
 * returns the mean kernel time measured with HIP events on the handle's
 * stream. precision: TOPOHIP_PREC_MIXED = FP32 apply, TOPOHIP_PREC_F64. */
int topohip_bench_prepare(topohip_problem* p, uint64_t seed);
int topohip_bench_ku(topohip_problem* p, int precision, int reps, double* ms_per_apply);
/* The fine-operator kernel variants the design loop launches, timed the same
 * way: variant 0 = FP32 y = K u, 1 = FP64 y = K u, 2 = FP64 A.p with the
 * fused <p, Ap> partials (the PCG's operator apply), 3 = FP32 zero-start
 * damped-Jacobi residual (the V-cycle's pre-smoothing apply), 4 = FP32
 * damped-Jacobi sweep with the fused <r, z> partials (post-smoothing), 5 =
 * the PCG's FP64 A.p as the design loop launches it in the mixed form: the
 * direction update p = z + beta p (FP32 z) fused in, p written, <p, Ap>
 * partials. Every
 * launch cycles to the next of *sets_used (nullable) operand sets, enough that
 * they span more than twice the 256 MiB last-level cache, so inputs come from
 * HBM. The launch shape is the handle's tile of that precision
 * (topohip_ku_tiles). No reference counterpart (topohip_bench_ku is variants
 * 0 / 1). */
int topohip_bench_ku_variant(topohip_problem* p, int variant, int reps, double* ms_per_apply,
                             int* sets_used);
/* The fine K.u launch shape of the handle's FP64 applies (the PCG's A.p):
 * tx x ty lanes per workgroup over a z-chunk of tz node planes (autotuned at
 * creation, or TOPOHIP_KU_TILE). No reference counterpart. */
int topohip_ku_tile(topohip_problem* p, int* tx, int* ty, int* tz);
/* The z-chunks of the FP32 and the FP64 applies (multiples of 4 and of 7
 * planes: the dot-partial slabs of the two precisions). No reference
 * counterpart. */
int topohip_ku_tiles(topohip_problem* p, int* tz32, int* tz64);
/* Average device time (ms) of the result-field kernels over reps launches on
 * the bench inputs (topohip_bench_prepare), rotated over operand sets that
 * span more than twice the last-level cache: ms3[0] von Mises to float32,
 * ms3[1] displacement narrowing, ms3[2] density halves. No reference
 * counterpart. */
int topohip_bench_fields(topohip_problem* p, int reps, double* ms3);

/* ---- serving: persistent continuous-batching loop ---------------------------
 * Jobs of uneven grid sizes are admitted as device memory (budget_bytes) and
 * slots (max_slots) allow; each admitted job keeps its density, displacement
 * and hierarchy resident and runs on its own stream, interleaved with the
 * others. There is no reference counterpart beyond the worker's FIFO queue
 * (tools/topopt-worker/topopt_worker.py:356-459). */
typedef struct {
  int nx, ny, nz;
  int bc_kind;        /* 0 = cantilever, 1 = MBB (see DESIGN) */
  double volfrac, rmin, penalty, move, cg_tol;
  int iterations;
  int precision;
} topohip_job;

typedef struct {
  int state;          /* 0 queued, 1 running, 2 done, 3 failed */
  int iterations;
  double compliance, volume_fraction;
  double t_submit_ms, t_start_ms, t_end_ms; /* server clock */
  int status;
} topohip_job_result;

int topohip_server_create(size_t budget_bytes, int max_slots, topohip_server** out);
int topohip_server_submit(topohip_server* s, const topohip_job* job, int64_t* job_id);
int topohip_server_poll(topohip_server* s, int64_t job_id, topohip_job_result* out);
/* Blocks until every submitted job has finished. */
int topohip_server_drain(topohip_server* s);
/* Copies a finished job's physical density (nx*ny*nz doubles). */
int topohip_server_fetch_density(topohip_server* s, int64_t job_id, double* xphys);
/* Message of a failed job (empty when it succeeded). */
int topohip_server_job_error(topohip_server* s, int64_t job_id, char* buf, size_t n);
int topohip_server_destroy(topohip_server* s);
/* Run records per job (dir NULL or "" turns them off; affects jobs started
 * afterwards): <dir>/job_<id>/iterations.csv (one row per design iteration),
 * stdout.txt (the CLI's PROGRESS lines as iterations complete, then the
 * VARIANT line), design.bin and fields.bin (one variant; von Mises of the
 * final solve formed on the device) — the files the LAN worker and the app
 * read from a run directory (tools/topopt-worker/topopt_worker.py:231-266). */
int topohip_server_set_output(topohip_server* s, const char* dir);
/* Device bytes a job of this shape keeps resident (admission estimate). */
size_t topohip_job_footprint(const topohip_job* job);

/* ---- result fields ----------------------------------------------------------
 * Formed on the device from the handle's grid. u (3 * nodes doubles) is
 * nullable everywhere below: NULL takes the handle's resident displacement of
 * its last solve (load case 0). */

/* fea_von_mises_field (core/include/topopt/fea.hpp:1324-1331,
 * core/src/fea/assembly.cpp:976-996): von Mises stress at each solid voxel's
 * centroid for Young's modulus `youngs` and the handle's Poisson ratio and
 * spacing; 0 on empty voxels. vm_out: nx*ny*nz doubles. */
int topohip_von_mises_field(topohip_problem* p, double youngs, const double* u, double* vm_out);

/* The two float32 arrays of a fields.bin variant block (write_fields_file,
 * core/src/io/fields.cpp:54-119): the von Mises field narrowed to float
 * (vm_out, nx*ny*nz) and the displacement narrowed to float (disp_out,
 * 3 * nodes). Either output may be NULL. */
int topohip_fields_f32(topohip_problem* p, double youngs, const double* u, float* vm_out,
                       float* disp_out);

/* The halves of a density snapshot (write_density_snapshot,
 * core/src/simp/observability.cpp:362-384): float_to_half(float(d[e])) per
 * voxel, IEEE binary16 round-to-nearest-even. density NULL = the handle's
 * resident physical density (after topohip_optimize*). */
int topohip_density_f16(topohip_problem* p, const double* density, uint16_t* out);

/* ---- per-iteration observation ----------------------------------------------
 * SimpIterationObservation + IterationPhaseTimes
 * (core/include/topopt/simp.hpp:870-1000). Fields for subsystems this library
 * does not have (plateau detector, infeasibility verdict, Krylov recycling,
 * GenEO) are always 0, which is what the reference writes when they are off.
 * The memory columns follow process_memory (core/src/simp/observability.cpp:
 * 116-205): negative = not answered. */
typedef struct {
  int iteration;              /* 1-based, monotone across stages */
  double compliance;          /* of xPhys at the start of the step */
  double change;
  double volume_fraction;     /* achieved after the step */
  int cg_iterations, cg_used_multigrid, cg_mg_levels, cg_hier_built, cg_mg_cycles_attempted;
  double beta;                /* stage beta while projecting, else 0 */
  double active_fraction;     /* 1 (the full analysis grid is solved) */
  int plateau, infeasible, recycle_dim, geneo_dim, geneo_action, geneo_burn, geneo_threshold;
  double total_ms, tail_prev_ms, filter_ms, project_ms, solve_ms, fea_ms, sens_ms, update_ms,
      analysis_ms, observe_ms, residual_ms;
  double solver_build_ms, solver_mg_build_ms, solver_mg_ms, solver_cg_ms,
      solver_geneo_setup_ms, solver_geneo_apply_ms, solver_recycle_ms;
  long long fea_solves, matvecs;
  double rss_mb, peak_rss_mb, compressed_mb, available_mb;
  long long major_faults, swapins;
} topohip_iteration_obs;

/* SimpOptions::observe (core/include/topopt/simp.hpp:1534-1542): called on
 * the optimising thread after every design iteration of topohip_optimize,
 * topohip_optimize_masked and topohip_optimize_projected on this handle. NULL
 * clears it. The callback must not call back into the same handle. */
typedef void (*topohip_iteration_cb)(const topohip_iteration_obs* obs, void* user);
int topohip_set_iteration_observer(topohip_problem* p, topohip_iteration_cb cb, void* user);

/* ---- run records (host formats, byte-compatible with the reference) -------- */

/* iterations.csv: IterationCsvWriter (core/include/topopt/observability.hpp:
 * 60-214, core/src/simp/observability.cpp:252-330). The header is
 * kIterationCsvHeader; a row is formatted exactly as append_at formats it. */
const char* topohip_iterations_csv_header(void);
/* Formats one row (with its newline) into buf; *len (nullable) receives its
 * length. EINVAL if n is too small. */
int topohip_iterations_csv_row(size_t rung, const topohip_iteration_obs* obs, long long wall_ms,
                               char* buf, size_t n, size_t* len);
typedef struct topohip_csv_writer topohip_csv_writer;
/* Opens (truncating) path and writes the header; every append flushes. */
int topohip_csv_open(const char* path, topohip_csv_writer** out);
/* append stamps the epoch wall clock in ms (wall_clock_ms); append_at takes it. */
int topohip_csv_append(topohip_csv_writer* w, size_t rung, const topohip_iteration_obs* obs);
int topohip_csv_append_at(topohip_csv_writer* w, size_t rung, const topohip_iteration_obs* obs,
                          long long wall_ms);
size_t topohip_csv_rows(const topohip_csv_writer* w);
int topohip_csv_close(topohip_csv_writer* w);

/* Density snapshot: <dir>/<basename>.f16 (raw little-endian halves, x-fastest)
 * and the <dir>/<basename>.json sidecar, as write_density_snapshot
 * (core/src/simp/observability.cpp:362-408); dir is created if absent. The
 * halves are formed on the device (topohip_density_f16); dims and spacing are
 * the handle's, origin (3 doubles, nullable = 0) is the grid's minimum corner. */
int topohip_write_density_snapshot(topohip_problem* p, const char* dir, const char* basename,
                                   const double* origin, const double* density, size_t rung,
                                   int iteration, int boundary);
/* SnapshotCapture (core/include/topopt/observability.hpp:236-259,
 * core/src/simp/observability.cpp:426-455): the density-snapshot feed's
 * cadence and cap. Rung boundaries are always written and never evicted;
 * other iterations every every_n (<= 0: 1); once more than cap (> 0)
 * per-iteration snapshots exist the oldest (.f16 and .json) is deleted.
 * Basenames are snap_rung<r>_iter<4-digit i>[_boundary]. */
typedef struct topohip_snapshots topohip_snapshots;
int topohip_snapshots_create(const char* dir, int every_n, int cap, topohip_snapshots** out);
/* Consider one event; *written (nullable) = 1 when a snapshot was written.
 * density NULL = the handle's resident physical density. */
int topohip_snapshots_capture(topohip_snapshots* sc, topohip_problem* p, const double* origin,
                              const double* density, size_t rung, int iteration, int boundary,
                              int* written);
int topohip_snapshots_counts(const topohip_snapshots* sc, size_t* written, size_t* evicted);
int topohip_snapshots_destroy(topohip_snapshots* sc);
/* read_density_f16 (core/src/simp/observability.cpp:410-425): EINVAL unless
 * the file holds exactly count halves. */
int topohip_read_density_f16(const char* path, size_t count, float* out);

/* The solved grid a design.bin / fields.bin container indexes to. */
typedef struct {
  int nx, ny, nz;
  double origin[3];
  double spacing;
} topohip_grid_frame;

/* design.bin v1 (core/include/topopt/design_store.hpp:40-100,
 * core/src/io/design_store.cpp:75-243). */
typedef struct {
  double requested_volume_fraction, achieved_volume_fraction;
  double margin_worst_case, margin_effective, max_von_mises_mpa;
  int accepted, iterations;
  double applied_build_dir[3];
  int build_direction_auto_applied, export_baked;
  uint64_t fingerprint;       /* read: as stored (verified); write: ignored */
  const double* density;      /* write: nx*ny*nz doubles */
} topohip_design_variant;
/* design_fingerprint: FNV-1a 64 over the little-endian bytes of the doubles. */
uint64_t topohip_design_fingerprint(const double* density, size_t n);
/* write_design_file: atomic (written beside, then renamed). */
int topohip_write_design_file(const char* path, const topohip_grid_frame* grid, int n_variants,
                              const topohip_design_variant* variants);
/* read_design_file in two steps: the header, then variant `index` (its
 * density into density_out, nx*ny*nz doubles, nullable). Version, header
 * sanity, counts and each fingerprint are checked (EINVAL on a mismatch). */
int topohip_read_design_header(const char* path, topohip_grid_frame* grid, int* n_variants);
int topohip_read_design_variant(const char* path, int index, topohip_design_variant* meta,
                                double* density_out);

/* fields.bin v1 (core/include/topopt/fields.hpp:28-70, core/src/io/fields.cpp:54-140):
 * one block per variant passed (the caller passes the accepted ones). Arrays
 * are float32 (topohip_fields_f32) and both are required: a variant without
 * them is EINVAL, as the reference refuses fields that do not match the grid.
 * Atomic (written beside, then renamed). */
typedef struct {
  double requested_volume_fraction, mass_grams;
  int support_volume_voxels;
  const float* von_mises;     /* nx*ny*nz */
  const float* displacement;  /* 3*(nx+1)(ny+1)(nz+1) */
} topohip_fields_variant;
int topohip_write_fields_file(const char* path, const topohip_grid_frame* grid, int n_variants,
                              const topohip_fields_variant* variants);

/* run_info.json, the run version record (RunInfo / run_info_json /
 * write_run_info, core/include/topopt/observability.hpp:260-268,
 * core/src/simp/observability.cpp:476-1284, 1761-1767): this library writes
 * the keys of the solve it runs (identity, solver, multigrid outcome,
 * precision, warm start, per-rung non-convergence, ladder, timestamp, the
 * record switches) in the reference's order and number formats (%.10g reals,
 * true/false, JSON null for a multigrid outcome not yet observed). Keys of
 * subsystems it does not have (GeneO, Krylov recycling beyond its off echo,
 * lattices, PLSM, preflight, active domain, print settings) are not written.
 * NULL strings are written as "". */
typedef struct {
  const char *cli_version, *fingerprint, *mode, *material, *source_format;
  int resolution;
  const char *load_source, *solver;
  int cg_multigrid_observed, cg_multigrid, mg_levels;
  const char* mg_mode;       /* NULL or "" = null */
  int mixed_precision, matfree_threads, warm_start_inherit;
  long long solved_grid_dofs;
  int projection;
  int n_rungs;               /* rung_* arrays (nullable when 0) */
  const int* rung_non_convergent;
  const int* rung_non_convergent_iteration;
  const double* rung_non_convergent_residual;
  int n_ladder;
  const double* ladder;
  const char* ladder_mode;
  long long created_wall_ms;
  int iteration_csv, density_snapshots, snapshot_every, snapshot_cap;
} topohip_run_info;
/* The record's text into buf (NUL-terminated); *len (nullable) receives its
 * length. EINVAL if n is too small. */
int topohip_run_info_json(const topohip_run_info* info, char* buf, size_t n, size_t* len);
/* write_run_info: the same text into path (truncating). */
int topohip_write_run_info(const char* path, const topohip_run_info* info);

/* The CLI's stdout protocol lines the LAN worker parses
 * (core/src/cli/run_job.cpp:10223-10259, tools/topopt-worker/topopt_worker.py:231-266),
 * newline included. */
int topohip_format_progress(size_t rung, size_t rungs, int iteration, char* buf, size_t n);
int topohip_format_variant(double vf, double achieved, double printed, double margin,
                           const char* mesh, char* buf, size_t n);

/* ---- V-cycle smoother ------------------------------------------------------------
 * kind 0 (the default): damped Jacobi, omega 0.6, one pre- and one
 * post-sweep per level, the reference's smoother (core/src/fea/multigrid.cpp:72,
 * 1128-1160). kind 1: Chebyshev of `degree` (1..8) on the Jacobi-scaled
 * operator of each smoothed level, an option of this engine with no reference
 * counterpart: degree operator applies per smoothing, the interval
 * [0.3, 1.1] * lambda_max(D^-1 A) with lambda_max from 10 power steps at every
 * hierarchy build. Applies to later solves of this handle. */
int topohip_set_smoother(topohip_problem* p, int kind, int degree);
/* The Chebyshev smoother's lambda_max estimates of the last hierarchy build,
 * one per smoothed level (level 0 = fine): up to n into lambdas, the count
 * into *levels (0 with the Jacobi smoother). */
int topohip_smoother_lambdas(topohip_problem* p, int n, double* lambdas, int* levels);

/* ---- solve deadline and cancellation -----------------------------------------
 * fea_set_solve_deadline_ms (core/include/topopt/fea.hpp:440-469,
 * core/src/fea/matfree.cpp:1267-1273): an absolute steady-clock time in ms
 * (topohip_steady_ms' clock) for every solve of the CALLING THREAD; <= 0
 * disarms it (the default). Thread-local and sticky; returns the previous
 * value. Armed, it is polled before each batch of PCG iterations: a solve
 * still running when it passes stops with TOPOHIP_EDEADLINE (the reference's
 * SolverDeadlineExceeded, a SolverNonConvergence) and the iterations and
 * residual it had reached in the info struct; a design loop ends as
 * non-convergent with deadline_exceeded set. A run that does not reach it is
 * bit-identical to a disarmed one. */
double topohip_set_solve_deadline_ms(double absolute_steady_ms);
double topohip_steady_ms(void);
/* The per-run multigrid stagnation latch of the CALLING THREAD:
 * fea_matfree_reset_mg_stagnation_latch / fea_matfree_mg_stagnation_latched
 * (core/include/topopt/fea.hpp:1081-1091; core/src/fea/multigrid.cpp:123-150,
 * 2138-2155, 2261-2275). A solve that builds a hierarchy and does not converge
 * with it (FP32 V-cycle, then the FP64 retry, each within the 300-cycle
 * budget) is a stagnation; 3 consecutive stagnations latch multigrid off, so
 * later solves of the thread skip the hierarchy build and go straight to
 * Jacobi-CG (hier_built = 0, used_multigrid = 0), and a converging solve
 * clears the count. A block solve counts its load cases in order, as one-case
 * solves would, and the latch holds from the next solve on. Sticky: the
 * driver resets it at the start of each run; the query returns 1 when
 * latched. */
void topohip_reset_mg_stagnation_latch(void);
int topohip_mg_stagnation_latched(void);
/* The latch's measurement-only re-arm period of the CALLING THREAD:
 * fea_matfree_set_mg_rearm_period / _mg_rearm_period / _mg_rearm_attempts /
 * _mg_rearm_carries (core/include/topopt/fea.hpp:1093-1116;
 * core/src/fea/multigrid.cpp:152-180, 2144-2156, 2248-2262, 2436-2454).
 * 0 (the default) keeps the latch for the run. n > 0: once latched, the n-th
 * latched solve call clears the latch and attempts multigrid again, with the
 * stagnation count one short of the threshold, so a stagnating retry
 * re-latches at once; attempts counts those solves, carries the ones that
 * converged (a block solve counts once, and carries when all its load cases
 * converged). The period survives topohip_reset_mg_stagnation_latch, which
 * zeroes the counters; setting it restarts the latched-solve count. */
void topohip_set_mg_rearm_period(int period);
int topohip_mg_rearm_period(void);
long long topohip_mg_rearm_attempts(void);
long long topohip_mg_rearm_carries(void);
/* SimpOptions::cancel (core/include/topopt/simp.hpp:1505-1521): a caller-owned
 * flag (NULL clears) polled once per design iteration, at its start, by
 * topohip_optimize, topohip_optimize_masked, topohip_optimize_mma and
 * topohip_optimize_projected on this handle. When nonzero the loop stops:
 * result.cancelled = 1, converged = 0, and the outputs are still consistent
 * (one final solve on the partially optimised density). The flag must outlive
 * the call; another thread may set it at any time. */
int topohip_set_cancel_flag(topohip_problem* p, const int* flag);

/* ---- misc ------------------------------------------------------------------ */
int topohip_set_device(int device);
int topohip_last_error(char* buf, size_t n);
const char* topohip_version(void);

#ifdef __cplusplus
}
#endif

#endif /* TOPOPT_HIP_H */
