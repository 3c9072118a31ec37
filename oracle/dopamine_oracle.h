/* ORACLE — test infrastructure only.
 *
 * FP64 CPU restatement of DOPAMINE's RK3W fractional-step hot path, used as the parity checker
 * for the HIP product path. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it. Nothing in dopamine_amd/ links or calls it.
 *
 * It is written as the reference's own sequential face-scatter loops (one face at a time,
 * owner/neighbour accumulation), each routine citing the reference file:line it follows.
 * Indices are 0-based; a face neighbour of -1 marks a boundary face.
 */
#ifndef DOPAMINE_ORACLE_H
#define DOPAMINE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mesh arrays (caller-owned, read-only), same meaning as the mesh builder's view. */
typedef struct om_mesh {
  int32_t n_cells, n_faces, n_int_faces, n_bnd_faces, n_patches;
  const double *cell_centroid, *cell_volume;
  const int32_t *cell_face_offset, *cell_faces;
  const int32_t *face_owner, *face_neighbour, *face_patch;
  const double *face_area, *face_centre, *face_normal, *face_sf, *face_dcf, *face_gc, *face_skew, *face_eps;
  const int32_t *patch_bc, *patch_start, *patch_nfaces;
  const int32_t *periodic_pair;
  const double *periodic_pair_dcf, *periodic_gc;
  const int32_t *ls_offset, *ls_cell, *ls_face;
  const double *ls_weight;
} om_mesh;

enum { OM_SCHEME_VANLEER = 1, OM_SCHEME_KOREN = 2, OM_SCHEME_CDS = 3, OM_SCHEME_KEP = 4, OM_SCHEME_QUICK = 5 };
enum { OM_GRAD_LS = 0, OM_GRAD_GG = 1 };
enum { OM_SGS_NONE = 0, OM_SGS_VREMAN = 1, OM_SGS_SMAGORINSKY = 2, OM_SGS_AMD = 3, OM_SGS_DYNAMIC = 4 };
enum { OM_WM_LOGLAW = 0, OM_WM_NONE = 1, OM_WM_MOST = 2 };

typedef struct om_params {
  double nu;
  double f_const[3];
  int scheme;            /* OM_SCHEME_* */
  int gradient;          /* OM_GRAD_* */
  int gg_iterations;
  int sgs;               /* OM_SGS_* */
  double c_vreman;       /* reference value 0.025 */
  double cs_smag;        /* 0.1 */
  int wall_model;        /* OM_WM_* */
  double wall_model_height;
  int n_nonortho_corr;   /* 0 = auto */
  double min_cos_theta;
  int rhie_chow_dual;
  int nonortho_check;
  int rotational;
  double p_tol;
  int p_max_iter;
  double cfl_max;
  double dt;             /* requested (maximum) time step */
  const double *patch_p_val;  /* n_patches, outlet pressure */
  const double *patch_U_val;  /* 3*n_patches, inlet / moving-wall velocity */
  double c_amd;               /* AMD constant, reference 1/3 */
  int amd_wall_gradient;      /* AMD: log-law wall-normal gradient in wall-model cells */
  /* temperature (solved when boussinesq != 0) */
  int boussinesq;
  double alpha_T, beta, T_ref, g[3], Pr_t;
  int t_scheme;               /* OM_SCHEME_* (CDS, QUICK, VANLEER, KOREN) */
  const double *patch_T_val;  /* n_patches, fixed-temperature / inlet value */
  /* MOST wall model (wall_model == OM_WM_MOST): roughness lengths (z0_h <= 0: z0_m e^-2),
   * surface temperature, initial Obukhov length (0 = neutral), domain-mean L update */
  double most_z0_m, most_z0_h, T_surface, L_obukhov;
  int compute_L;
  /* cell_projection: OM_PROJ_LS (correct_velocity_psolver_dual) or OM_PROJ_EXACT (V-orthogonal
   * projection by flexible PCG, src/numerics/mod_cell_projection.f90), its energy tolerance */
  int cell_projection;
  double cell_projection_tol;
  /* momentum forcing (compute_explicit_rhs, src/solver/mod_fv_solver.f90:1661-1700): Coriolis
   * -2 Omega x U; harmonic forcing amp cos(omega t + phase) along dir; constant mass flux: the
   * uniform force (U_bulk_target - U_bulk) / dt set at each step start in place of f_const */
  int use_rotation;
  double omega_rot[3];
  int use_harmonic;
  double forcing_amp, forcing_omega, forcing_phase, forcing_dir[3];
  int mass_flux;
  double U_bulk_target[3];
  /* passive scalars (n_scalars <= OM_MAX_SCALARS): advection scheme, molecular diffusivity and
   * turbulent Schmidt number per scalar; per patch and scalar ([OM_MAX_SCALARS p + s]) the BC mode
   * (OM_SCALAR_DIRICHLET, _NEUMANN, _ZERO_GRADIENT), value, flux, flux frequency and phase */
  int n_scalars;
  int scalar_scheme;
  double scalar_diffusivity[4], scalar_pr_t[4];
  const int *scalar_bc_mode;
  const double *scalar_bc_value, *scalar_bc_flux, *scalar_bc_omega, *scalar_bc_phase;
  /* boundary extensions, per patch (NULL: all off): convective outlet (patch_outlet_mode
   * 'convective', src/bc/mod_bc.f90:681-786), backflow-traction coefficient (:789-818), MOST
   * roughness lengths (0 inherits most_z0_m / most_z0_h; a negative z0_h derives it from z0_m,
   * src/abl/mod_abl_wall.f90:249-262); the log-law wall model's exponential time filter
   * (wall_model_filter_time > 0) and adaptive filter factor (wall_model_filter_tf_fac > 0:
   * T_f = f y / (kappa u_tau)), src/sgs/mod_wall_model.f90:176-313 */
  const int *patch_outlet_conv;
  const double *patch_backflow_beta, *patch_z0_m, *patch_z0_h;
  double wall_model_filter_time, wall_model_filter_tf_fac;
} om_params;

enum { OM_MAX_SCALARS = 4, OM_SCALAR_DIRICHLET = 1, OM_SCALAR_NEUMANN = 2, OM_SCALAR_ZERO_GRADIENT = 3 };

enum { OM_PROJ_LS = 0, OM_PROJ_EXACT = 1 };

typedef struct om_solver om_solver;

om_solver *om_create(const om_mesh *m, const om_params *p);
void om_free(om_solver *s);
/* Pointer to an internal array by name ("U","p","p_bnd","U_bnd","nu_sgs","grad_U","U_face",
 * "face_flux","tau_w_bnd","w_gp","rhs_p","T","T_bnd","grad_T","q_w_bnd","L_bnd", and per
 * passive scalar s "C<s>", "C<s>_bnd", "grad_C<s>"), with its length in doubles. */
double *om_array(om_solver *s, const char *name, int64_t *len);
/* After U (and p) were written from outside: refresh boundary values as solver_init does. */
void om_apply_all_bc(om_solver *s);
int om_project_once(om_solver *s, double dt);
/* One solver_advance: CFL clamp of dt_req, RK3W step; *dt_used receives the step taken. */
int om_advance(om_solver *s, double dt_req, double *dt_used);
/* compute_stats_sums + divergence metrics: out = {E, eps, Umax, Ub, tauw_mean, nu_sgs_mean,
 * max_abs_divU, volavg_abs_divU}. */
void om_stats(om_solver *s, double out[8]);
double om_time(const om_solver *s);
/* Cell-centred divergence metrics {max, volume mean} (src/solver/mod_fv_solver.f90:3641-3713). */
void om_divergence_cell(om_solver *s, double out[2]);
/* Boundary refresh, then the closures a RK3W stage starts with (velocity gradients into "grad_U",
 * SGS viscosity into "nu_sgs" with stability cap 0.4 V^(2/3) / dt_stab, wall traction into
 * "tau_w_bnd"), on the current U, without advancing. */
int om_eval_closures(om_solver *s, double dt_stab);
/* single operator routines (gradient, advection, viscous, SGS, wall model) on caller-set fields */
void om_op_gradient_vector(om_solver *s, int gg, int iterations, double *out);
void om_op_gradient_scalar(om_solver *s, int gg, int iterations, const double *phi, const double *pb, double *out);
void om_op_advection_vector(om_solver *s, int scheme, double *out);
void om_op_viscous_vector(om_solver *s, double *out);
void om_op_advection_scalar(om_solver *s, int scheme, const double *phi, const double *pb, const double *g,
                            double *out);
void om_op_viscous_scalar(om_solver *s, const double *phi, const double *pb, const double *g, double alpha,
                          double prt, double *out);
void om_op_sgs(om_solver *s, int model, double coef, int wall_gradient, double dt_stab, double *out);
void om_op_wall_model(om_solver *s, double dt, double *out);
/* outlet_dn_u (src/bc/mod_bc.f90:775-786): n.grad u at an outlet face from the face and owner
 * values, the owner gradient along the offset (gdd) and along n - (n.d)d (gt), |d| and n.d */
double om_outlet_dn_u(double u_b, double u_o, double gdd, double gt, double len_d, double nd);
/* the wall-model sampling cell and distance of boundary face b, the filtered wall-parallel
 * velocity (3 per boundary face) and whether the filter is seeded */
int om_wall_sample(const om_solver *s, int b, double *y);
int om_wall_filter_state(const om_solver *s, double *out);
/* per-patch convective-outlet speed and reference speed of the last stage (2 per patch) */
void om_outlet_speeds(const om_solver *s, double *out);
/* Wall shear stress tau_w = u_tau^2 of the Reichardt-law Newton solve for a wall-parallel speed
 * u_par sampled at wall distance y_w (src/sgs/mod_wall_model.f90:50-115). */
double om_wall_tau(double u_par, double y_w, double nu);
/* AMD closure at one point (src/sgs/mod_amd.f90:31-66), column-major alpha(i,j) = du_j/dx_i and
 * G(i,j), no stability cap. */
double om_amd_point(const double alpha[9], const double G[9], double c_amd);
/* MOST surface-layer solve at one wall face (src/abl/mod_abl_wall.f90:155-223): speed u_par and
 * cell temperature T_o at wall distance y_w over a surface at T_surface, roughness z0_m, z0_h,
 * starting Obukhov length L_init. out = {u_star, T_star, L}. */
void om_most_iterate(double u_par, double T_o, double T_surface, double y_w, double z0_m, double z0_h,
                     double L_init, double out[3]);
/* Exact cell projection operators on the solver's pressure faces (cp_divergence,
 * cp_adjoint_gradient, cp_project of src/numerics/mod_cell_projection.f90:89-303): out = B u
 * (nc), v = V^-1 B^T chi (3*nc), and u <- u - V^-1 B^T chi with (B V^-1 B^T) chi = B u + bconst by
 * flexible PCG, preconditioned by the identity (precond 0) or by the pressure solve at rtol 0.05,
 * 50 iterations (precond 1, project_cells_exact). Returns the iteration count. */
void om_proj_divergence(om_solver *s, const double *u, double *out);
void om_proj_adjoint(om_solver *s, const double *chi, double *v);
int om_proj_project(om_solver *s, double *u, const double *bconst, int precond, double tol, int max_it);
/* {projections, total projection iterations} of the time steps so far. */
void om_proj_counters(const om_solver *s, int64_t out[2]);
/* Over-relaxed non-orthogonal k-vector with the cos(theta) floor min_cos (nonortho_kfactor_limited,
 * src/numerics/mod_nonortho.f90:65-98): k = Sf - (|Sf|^2 / max(Sf.d, min_cos |Sf| |d|)) d. */
void om_kfactor_limited(const double sf[3], const double d_on[3], double min_cos, double k[3]);
/* Businger-Dyer stability corrections psi_m, psi_h (src/abl/mod_abl_profiles.f90:51-85). */
double om_psi_m(double zeta);
double om_psi_h(double zeta);
/* Pressure-operator face lists ("ps_owner", "ps_nb", "ps_skip"), 0-based. */
int32_t *om_int_array(om_solver *s, const char *name, int64_t *len);
/* Counters: total pressure solves, total CG iterations, last CG iteration count. */
void om_counters(const om_solver *s, int64_t out[3]);
int om_nonortho_info(const om_solver *s, int out[3]); /* skip_nonortho, skip_grad_p, n_nonortho_corr */

/* Standalone pressure solve with the reference's operator assembly convention
 * (src/solver/pressure_backend_petsc.c:296-365): A_ii = diag, A_on = A_no = -coeff;
 * Jacobi-preconditioned CG, unpreconditioned 2-norm stop ||r|| <= max(rtol*||b||, atol);
 * constant null space projected from b and from each preconditioned residual when
 * nullspace != 0. Returns 0 when converged, -1 otherwise. 0-based owner/neighbour. */
int om_cg_jacobi(int n, int nfaces, const int32_t *owner, const int32_t *neighbour, const double *coeff,
                 const double *diag, const double *b, double *x, double rtol, double atol, int max_it,
                 int nullspace, int *its, double *resid);

#ifdef __cplusplus
}
#endif
#endif
