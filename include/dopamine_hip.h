/* C ABI of libdopamine_hip.so: the MI355X (gfx950) implementation of DOPAMINE's RK3W
 * fractional-step hot path.
 *
 * Two levels, both plain C (no torch types):
 *  1. The pressure-backend ABI of src/solver/pressure_backend_petsc.c, symbol for symbol
 *     (declared in fvm_pressure_backend.h), so src/solver/mod_pressure_solver.f90 links unchanged.
 *  2. A solver-handle ABI mirroring the FVSolver type-bound procedures of
 *     src/solver/mod_fv_solver.f90:270-285 (init / advance / project_once / compute_dt /
 *     to_device / sync_host / divergence_metrics) for a Fortran host that keeps its own Mesh.
 *
 * All mesh indices are 0-based here; a Fortran binding subtracts 1 from its 1-based arrays.
 */
#ifndef DOPAMINE_HIP_H
#define DOPAMINE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mesh arrays as the reference's Mesh SoA mirror holds them (src/mesh/mod_mesh.f90:164-186 and
 * the LS stencil of :105-110), 0-based, host memory, read once by dopamine_solver_create. */
typedef struct dopamine_mesh_desc {
  int32_t n_cells, n_faces, n_int_faces, n_bnd_faces, n_patches;
  const double *cell_centroid, *cell_volume;           /* 3*nc, nc */
  const int32_t *cell_face_offset, *cell_faces;        /* nc+1, CSR ascending face order */
  const int32_t *face_owner, *face_neighbour, *face_patch;
  const double *face_area, *face_centre, *face_normal, *face_sf, *face_dcf, *face_gc, *face_skew, *face_eps;
  const int32_t *patch_bc, *patch_start, *patch_nfaces;
  const int32_t *periodic_pair;                        /* per boundary face, -1 = none */
  const double *periodic_pair_dcf, *periodic_gc;
  const int32_t *ls_offset, *ls_cell, *ls_face;        /* LS stencil CSR; ls_cell -1 = boundary face */
  const double *ls_weight;                             /* 3 per entry */
} dopamine_mesh_desc;

/* Numerics and physics (solver_params_t, src/solver/mod_fv_solver.f90:41-110). */
typedef struct dopamine_params {
  double nu;
  double f_const[3];
  int32_t scheme;            /* 1 VANLEER_TVD, 2 KOREN_TVD, 3 CDS, 4 KEP, 5 QUICK (src/mod_parameters.f90:54-60) */
  int32_t gradient;          /* 0 LEAST_SQUARES, 1 GREEN_GAUSS */
  int32_t gg_iterations;
  int32_t sgs;               /* 0 none, 1 Vreman, 2 Smagorinsky, 3 AMD, 4 dynamic Smagorinsky */
  double c_vreman;           /* reference C_VREMAN = 0.025 */
  double cs_smag;            /* 0.1 */
  int32_t wall_model;        /* 0 log-law (Reichardt), 1 none, 2 MOST (Monin-Obukhov, rough wall) */
  double wall_model_height;  /* log-law sampling height; <= 0: wall-adjacent cell */
  int32_t n_nonortho_corr;   /* 0 = auto */
  double min_cos_theta;
  int32_t rhie_chow_dual;
  int32_t nonortho_check;
  int32_t rotational;
  double p_tol;
  int32_t p_max_iter;
  double cfl_max;
  double dt;
  const double *patch_p_val;  /* n_patches */
  const double *patch_U_val;  /* 3*n_patches */
  int32_t precision;          /* 0: FP64 pressure solve; 1: FP32-SpMV CG inside FP64 iterative refinement */
  int32_t cg_variant;         /* 0: two passes per iteration (KSPCG's Hestenes-Stiefel recurrence);
                                 1: one pass per iteration (single-reduction recurrence) */
  double c_amd;               /* AMD constant (sgs 3), reference AMD_C_DEFAULT = 1/3 */
  int32_t amd_wall_gradient;  /* AMD: log-law wall-normal gradient in wall-model cells */
  /* temperature: solved when boussinesq != 0 (solve_temperature, mod_fv_solver.f90:503) */
  int32_t boussinesq;
  double alpha_T, beta, T_ref, g[3], Pr_t;
  int32_t t_scheme;           /* temperature_scheme: 1 VANLEER_TVD, 2 KOREN_TVD, 3 CDS, 5 QUICK */
  const double *patch_T_val;  /* n_patches: fixedTemperature / inlet temperature */
  /* Pressure preconditioner: 0 Jacobi (the CG passes above); 1 learned: the first learn_collect
   * pressure solves run Jacobi CG and keep residual snapshots (learn_capacity of them), the
   * message-passing network is then trained (learn_iters Adam steps of learn_batch residuals,
   * rate learn_lr, seed learn_seed) and later solves run flexible CG with it, falling back to
   * Jacobi CG when a solve fails or needs more iterations than the Jacobi solves did; 2 the
   * aggregation-multigrid V-cycle in flexible CG (in FP32 inside FP64 refinement with precision 1). */
  int32_t precond;
  int32_t learn_collect, learn_capacity, learn_iters, learn_batch, learn_seed;
  double learn_lr;
  /* MOST wall model (wall_model 2, src/abl/mod_abl_wall.f90): roughness lengths (most_z0_h <= 0:
   * most_z0_m e^-2), surface temperature, initial Obukhov length (0 = neutral), and whether the
   * first stage of a step updates the domain-mean L. The cell temperature is T when boussinesq,
   * else T_ref; the heat flux enters the temperature equation. */
  double most_z0_m, most_z0_h, T_surface, L_obukhov;
  int32_t compute_L;
  /* cell_projection (src/solver/mod_fv_solver.f90:91-92): 0 'LS' (correct_velocity_psolver_dual),
   * 1 'EXACT' (V-orthogonal projection by flexible PCG with the pressure solve as preconditioner,
   * src/numerics/mod_cell_projection.f90); cell_projection_tol its relative energy tolerance. */
  int32_t cell_projection;
  double cell_projection_tol;
  /* momentum forcing (compute_explicit_rhs and body_force_now, src/solver/mod_fv_solver.f90:
   * 1661-1700, 2904-2923): Coriolis -2 omega_rot x U when use_rotation; harmonic forcing
   * forcing_amp cos(forcing_omega t + forcing_phase) along forcing_dir when use_harmonic; with
   * mass_flux the uniform force (U_bulk_target - U_bulk) / dt, set at each step start
   * (update_mass_flux_forcing, :960-983), replaces f_const (which must then be 0). */
  int32_t use_rotation;
  double omega_rot[3];
  int32_t use_harmonic;
  double forcing_amp, forcing_omega, forcing_phase, forcing_dir[3];
  int32_t mass_flux;
  double U_bulk_target[3];
  /* passive scalars (n_scalars <= 4; compute_explicit_rhs :1719-1743, apply_passive_scalar_bc,
   * src/bc/mod_bc.f90:582-679): advection scheme (as t_scheme), molecular diffusivity and
   * turbulent Schmidt number per scalar; per patch and scalar ([4 patch + s], 4 n_patches entries)
   * the BC mode (1 Dirichlet, 2 Neumann, 3 zero gradient), value, flux, flux frequency, phase.
   * Fields "C<s>", "C<s>_bnd", "grad_C<s>". */
  int32_t n_scalars, scalar_scheme;
  double scalar_diffusivity[4], scalar_pr_t[4];
  const int32_t *scalar_bc_mode;
  const double *scalar_bc_value, *scalar_bc_flux, *scalar_bc_omega, *scalar_bc_phase;
  /* boundary extensions, per patch (n_patches entries; null: all off): patch_outlet_conv 1 = the
   * outlet's boundary velocity is a convected RK state (patch_outlet_mode 'convective',
   * src/bc/mod_bc.f90:681-786); patch_backflow_beta > 0 the backflow-traction coefficient of the
   * outlet pressure (:789-818); patch_z0_m / patch_z0_h the MOST roughness of a wall patch (0
   * inherits most_z0_m / most_z0_h; a negative z0_h derives it from z0_m,
   * src/abl/mod_abl_wall.f90:249-262). The log-law wall model (wall_model 0) samples the velocity
   * wall_model_height above the wall (> 0: the cell-to-cell walk of find_sample_cell) and filters
   * it in time: wall_model_filter_time > 0 an exponential filter of that time constant,
   * wall_model_filter_tf_fac > 0 the adaptive T_f = f y / (kappa u_tau), advanced once per step
   * (src/sgs/mod_wall_model.f90:126-313). */
  const int32_t *patch_outlet_conv;
  const double *patch_backflow_beta, *patch_z0_m, *patch_z0_h;
  double wall_model_filter_time, wall_model_filter_tf_fac;
} dopamine_params;

/* Create a solver on the current HIP device (mesh + params copied to HBM). NULL on error. */
void *dopamine_solver_create(const dopamine_mesh_desc *mesh, const dopamine_params *params);
void dopamine_solver_destroy(void *solver);
/* Device (HIP) index the library runs on; -1 if none is usable. */
int dopamine_device_count(void);
const char *dopamine_last_error(void);

/* Copy a host field in / out. Names: "U" (3*nc, component-fastest), "p" (nc), "nu_sgs", "U_bnd",
 * "p_bnd", "grad_U" (9*nc, grad(i,j) = du_j/dx_i at [9c+3j+i]), "U_face" (nf), "tau_w_bnd",
 * "rhs_p", "H", with boussinesq "T" (nc), "T_bnd", "grad_T" (3*nc), with wall_model 2 "q_w_bnd"
 * and "L_bnd" (one per boundary face), with passive scalars "C<s>" (nc), "C<s>_bnd", "grad_C<s>".
 * set_field("U"), set_field("T") and set_field("C<s>") re-apply the boundary conditions as
 * solver_init does. */
int dopamine_solver_set_field(void *solver, const char *name, const double *host);
int dopamine_solver_get_field(void *solver, const char *name, double *host, int64_t len);

/* solver_project_once (src/solver/mod_fv_solver.f90:837-917). A pressure solve that stops at its
 * iteration limit is repeated from p = 0 with four times the limit, then soft-accepted or failed
 * with -997 as the reference backend does (src/solver/pressure_backend_petsc.c:878-937); every
 * non-zero return leaves its reason in dopamine_last_error(). */
int dopamine_solver_project_once(void *solver, double dt);
/* solver_advance (src/solver/mod_fv_solver.f90:920-957): CFL clamp, one RK3W step. Returns -2
 * when a pressure solve failed (after the retry above), -3 when the CFL/diffusion limit is below
 * 1e-6 dt; the reason is in dopamine_last_error(). */
int dopamine_solver_advance(void *solver, double dt_req, double *dt_used);
/* solver_state_is_finite (src/solver/mod_fv_solver.f90:3925-3945): 1 when U and p hold no NaN or
 * Inf (the sum of U and p over the cells, reduced on the device, is finite), 0 otherwise, negative
 * on error. The reference's time loop checks it every 10 steps (src/main.f90:434-442). */
int dopamine_solver_state_is_finite(void *solver);
/* max over the cells of |sum of the cell's projected face fluxes| (m^3/s, the discrete divergence
 * integrated over the cell; solver_divergence_metrics, mod_fv_solver.f90:3641-3713, reports the
 * same sum divided by the cell volume). No counterpart of its own in the reference. */
int dopamine_solver_cell_net_flux(void *solver, double *out);
/* Pressure solves that needed the retry from p = 0 (see dopamine_solver_project_once). */
int64_t dopamine_solver_retries(void *solver);
/* solver_compute_dt (src/solver/mod_fv_solver.f90:3864-3899). */
double dopamine_solver_compute_dt(void *solver);
/* compute_stats_sums + solver_divergence_metrics: {E, eps, Umax, Ub, tauw_mean, nu_sgs_mean,
 * max_abs_divU, volavg_abs_divU} (src/stats/mod_stats.f90:271-336). */
int dopamine_solver_stats(void *solver, double out[8]);
double dopamine_solver_time(void *solver);
/* Cell-centred velocity divergence of solver_divergence_metrics (mod_fv_solver.f90:3641-3713):
 * {max |div u|, volume-averaged |div u|}, after a gradient-free velocity boundary refresh. */
int dopamine_solver_cell_divergence(void *solver, double out[2]);
/* Append one row to the stats CSV at path in the format of src/stats/mod_stats.f90:68-70, 241-249
 * (header when the file is new or append == 0). dt is the requested step (main.f90:452 passes it
 * before solver_advance's CFL clamp). *E0 <= 0 takes this row's E as the reference energy and
 * returns it. */
int dopamine_solver_stats_csv(void *solver, const char *path, int step, double dt, double *E0, int append);
/* Checkpoint in the reference's per-rank format (write_checkpoint, src/io/mod_checkpoint_io.f90:15-150):
 * dir/processor<rank>/checkpoint_<step, 8 digits>.bin, Fortran unformatted sequential records
 * (header, time, global ids, U, p, T, nu_sgs, U_face, wall traction [, q_w, L for MOST]); rank 0
 * also writes dir/checkpoint_latest. */
int dopamine_solver_write_checkpoint(void *solver, const char *dir, int step, int rank);
/* read_checkpoint (:153-283) followed by solver_refresh_after_restart (mod_fv_solver.f90:601-609).
 * Returns the checkpoint's step (>= 0) or a negative error (dimension or partition mismatch). */
int dopamine_solver_read_checkpoint(void *solver, const char *dir, int step, int rank);
/* {pressure solves, total CG iterations, last CG iteration count, refinement sweeps}. */
int dopamine_solver_counters(void *solver, int64_t out[4]);
/* {skip_nonortho, skip_grad_p, n_nonortho_corr}. */
int dopamine_solver_nonortho_info(void *solver, int32_t out[3]);
/* Learned preconditioner state: {phase (0 collecting, 1 in use, 2 switched off), solves
 * collected, residual samples kept, learned solves, their CG iterations, fallbacks to Jacobi CG,
 * mean Jacobi CG iterations per solve, training time in ms}. -1 when precond != 1. */
int dopamine_solver_learned_info(void *solver, double out[8]);
/* Preconditioner solves of the exact cell projection that ran the learned flexible CG (the policy
 * in use): {solves, their CG iterations}. -1 when precond != 1. No reference counterpart (the
 * reference's projection preconditions with its PETSc pressure solve). */
int dopamine_solver_learned_precond_info(void *solver, int64_t out[2]);
/* Exact cell projection counters: out = {projections, flexible-PCG iterations} so far. */
int dopamine_solver_projection_info(void *solver, int64_t out[2]);
/* Multigrid hierarchy of the pressure solve (precond 2): out[0] levels L, out[1] 1 when the first
 * coarsening was aggressive, out[2..2+L) cells per level, out[2+L..2+2L) off-diagonal entries per
 * level, out[2+2L..2+3L) their sliced-ELL slots with padding (both 0 for level 0, which uses the
 * solver's ELL lists); max >= 2 + 3L. No reference counterpart (BoomerAMG's hierarchy lives inside
 * PETSc). */
int dopamine_solver_multigrid_info(void *solver, int64_t *out, int max);
/* Mean training loss of each Adam step (up to max values); returns the count. */
int dopamine_solver_learned_loss(void *solver, double *out, int max);
/* Wait for all work queued by the solver. */
int dopamine_solver_sync(void *solver);

/* ---- Single routines of the reference's numerics and SGS modules on the solver's mesh, for the
 * Fortran drop-in modules (dopamine_amd/fortran/dopamine_modules.f90). Host arrays in the
 * reference's layout (component fastest: U(3, n_cells), grad_U(3, 3, n_cells) with
 * grad(i, j) = du_j/dx_i, U_bnd(3, n_bnd_faces)); each call copies its inputs to the GPU, runs the
 * routine's kernel on the solver's device mesh and copies the result back. 0 or a negative error. */
/* {n_cells, n_faces, n_bnd_faces} of the solver's mesh. */
int dopamine_op_sizes(void *solver, int64_t out[3]);
/* ls_gradient_vector / ls_gradient_scalar (src/numerics/mod_gradient.f90:425, 307). */
int dopamine_op_ls_gradient_vector(void *solver, const double *U, const double *U_bnd, double *grad_U);
int dopamine_op_ls_gradient_scalar(void *solver, const double *phi, const double *phi_bnd, double *grad_phi);
/* green_gauss_gradient_vector / _scalar (mod_gradient.f90:172, 38): one sweep; grad_prev (or NULL)
 * is the previous sweep's gradient for the skewness correction. */
int dopamine_op_gg_gradient_vector(void *solver, const double *U, const double *U_bnd, const double *grad_prev,
                                   double *grad_U);
int dopamine_op_gg_gradient_scalar(void *solver, const double *phi, const double *phi_bnd, const double *grad_prev,
                                   double *grad_phi);
/* nu_t from grad_U: model 1 vreman_sgs_viscosity (coef = C_v; src/sgs/mod_vreman.f90:50), 2
 * smagorinsky_sgs (coef = C_s; :295), 3 amd_sgs_viscosity (coef = c_amd, with the log-law
 * wall-cell gradient when wall_gradient != 0, nu the molecular viscosity; src/sgs/mod_amd.f90:69);
 * dt_stab > 0 caps nu_t at 0.4 V^(2/3) / dt_stab. */
int dopamine_op_sgs_viscosity(void *solver, int model, const double *U, const double *grad_U, double coef, double nu,
                              int wall_gradient, double dt_stab, double *nu_sgs);
/* dynamic_smagorinsky_sgs (src/sgs/mod_dynamic_smagorinsky.f90:39): Germano-Lilly nu_sgs(n_cells) from
 * U(3, n_cells), grad_U(3, 3, n_cells) and U_bnd(3, n_bnd); dt_stab > 0 caps it at 0.4 Delta^2 / dt_stab
 * (the solver handle needs an SGS model other than none, so the filter-width tensor exists). */
int dopamine_op_dynamic_smagorinsky(void *solver, const double *U, const double *grad_U, const double *U_bnd,
                                    double dt_stab, double *nu_sgs);
/* apply_wall_model_bc (src/sgs/mod_wall_model.f90:176): tau_w_bnd(3, n_bnd) of the log-law wall
 * model. h_wm: the sampling height, which must equal the solver's wall_model_height (the sampling
 * cells are found at create); t_filt > 0 / tf_fac > 0: the fixed / adaptive time filter, advanced
 * by dt, its state (seeded by the first filtered call) kept in the solver between calls. */
int dopamine_op_wall_model(void *solver, const double *U, double nu, double h_wm, double t_filt, double dt,
                           double tf_fac, double *tau_w_bnd);
/* advection_rhs_vector (src/numerics/mod_advection.f90:679): rhs = the advective term (overwritten);
 * scheme VANLEER_TVD 1, KOREN_TVD 2, CDS 3, KEP 4, QUICK 5; face_flux(n_faces) the projected flux. */
int dopamine_op_advection_rhs_vector(void *solver, int scheme, const double *U, const double *U_bnd,
                                     const double *face_flux, const double *grad_U, double *rhs);
/* viscous_rhs_vector (src/numerics/mod_viscous.f90:37): adds the viscous term to rhs. */
int dopamine_op_viscous_rhs_vector(void *solver, double nu, const double *U, const double *nu_sgs,
                                   const double *grad_U, const double *U_bnd, double *rhs);
/* advection_rhs_scalar (mod_advection.f90:56): rhs(n_cells) overwritten; scheme 1, 2, 3 or 5. */
int dopamine_op_advection_rhs_scalar(void *solver, int scheme, const double *phi, const double *U,
                                     const double *phi_bnd, const double *U_bnd, const double *face_flux,
                                     const double *grad_phi, double *rhs);
/* viscous_rhs_scalar (mod_viscous.f90:286): adds (alpha + nu_t / Pr_t) diffusion to rhs. */
int dopamine_op_viscous_rhs_scalar(void *solver, const double *phi, double alpha_mol, const double *nu_sgs,
                                   double pr_t, const double *grad_phi, const double *phi_bnd, double *rhs);

/* Kernel classes with device timing. */
enum {
  DOPAMINE_K_CG_DIR = 0,   /* pressure CG, first pass of an iteration: p update + operator apply
                              + p.q (cg_variant 0 and the flexible CG), or the whole iteration
                              (cg_variant 1) */
  DOPAMINE_K_CG_UPD = 1,   /* pressure CG, second pass (cg_variant 0): x/r update, Jacobi, dots */
  DOPAMINE_K_EXPLICIT = 2, /* advection + viscous + forcing + RK3 predictor */
  DOPAMINE_K_LS_GRAD = 3,  /* least-squares velocity gradient with the fused SGS viscosity */
  DOPAMINE_K_MG_DOWN = 4,  /* multigrid cycle, finest level: pre-smoothing and residual */
  DOPAMINE_K_MG_UP = 5     /* multigrid cycle, finest level: post-smoothing and the flexible-CG partials */
};
/* Device timing of one kernel class: an event pair on the solver's stream that the dispatch of
 * each launch stamps at the kernel's start and end (hipExtLaunchKernelGGL).
 * enable != 0 clears the totals and starts recording, 0 stops. Measurement only: no counterpart
 * in the reference (its NVTX ranges, src/io/nvtx_shim.c, are host-side markers). */
int dopamine_solver_kernel_timing(void *solver, int kernel, int enable);
/* Launches timed since enabling and their summed device duration in milliseconds. CG passes that
 * run after convergence (gated no-ops) are not counted. */
int dopamine_solver_kernel_time(void *solver, int kernel, int64_t *launches, double *total_ms);
/* fvm_petsc_solve_pressure without any allocation. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). Same arguments and return codes as fvm_petsc_solve_pressure
 * (fvm_pressure_backend.h), with rhs and solution always device pointers; the operator arrays are
 * host arrays as there. It runs the Jacobi CG (KSPCG's two-pass recurrence, with the reference's
 * retry and soft accept) whatever the selector, on the null stream. Its device memory is a
 * file-scope array of the library, one copy per GPU, about 20 MB: operators of at most 131072
 * rows and 8 off-diagonal entries per row; larger ones return -61 and -62 with the reason in
 * dopamine_backend_last_error(). Calls share that memory and the host assembly arrays: calls that
 * overlap (on one GPU or from several host threads) race. */
int fvm_petsc_solve_pressure_noalloc(int n_cells_local, int n_cells_global, int n_faces, const int *local_to_global,
                                     const int *face_owner, const int *face_neighbor_local,
                                     const int *face_neighbor_global, const int *skip_face, const double *diagonal,
                                     const double *rhs, const double *face_coeff, int reference_cell,
                                     int max_iterations, double rtol, double atol, const int *dirichlet_rows,
                                     const double *dirichlet_values, int n_dirichlet_rows, double *solution,
                                     int *iteration_count, double *residual_norm);
/* Reason of the last non-zero return of fvm_petsc_solve_pressure_noalloc ("" after success). */
const char *dopamine_backend_last_error(void);
/* Outcome of the last fvm_petsc_solve_pressure call of the pressure backend (no counterpart in the
 * reference): {return code of the first solve, its iterations, 1 if it was repeated from x = 0,
 * iterations of both solves}. */
void dopamine_backend_last_solve(int out[4]);
/* Measurement utility (no counterpart in the reference): copies n doubles between two fresh HBM
 * buffers reps times with a plain grid-stride kernel and returns the mean kernel duration; the
 * known 16 n bytes per launch calibrate the HBM traffic counters and the achievable bandwidth. */
int dopamine_hbm_copy_probe(int64_t n, int reps, double *avg_ms);

#ifdef __cplusplus
}
#endif
#endif
