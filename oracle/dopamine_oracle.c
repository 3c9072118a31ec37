/* ORACLE — test infrastructure only (see dopamine_oracle.h).
 *
 * FP64 CPU restatement of one DOPAMINE RK3W step and its pressure projection, restated from the
 * reference Fortran as sequential face-scatter loops. Every routine cites what it follows.
 * parity pins: tests/test_oracle_regression.py (reference regression CSVs), tests/test_oracle_units.py
 * (closed forms and reference unit-test known answers).
 */
#include "dopamine_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define SMALL 1.0e-30
#define NONORTHO_SIN2_TOL 1.0e-10
#define SKEW_BLEND_LO 0.15
#define SKEW_BLEND_HI 0.25
#define REFINE_SKEW_TOL 0.15
#define KAPPA 0.41
#define BLOG 5.2
#define NU_SGS_STAB_FRAC 0.4

enum {
  BC_INLET = 1, BC_OUTLET = 2, BC_WALL_NOSLIP = 3, BC_WALL_MODEL = 4, BC_SLIP = 5, BC_PERIODIC = 6,
  BC_SYMMETRY = 7, BC_FIXEDTEMP = 8, BC_ZEROGRADIENT = 9, BC_WALL_VELOCITY = 10, BC_ABL_INLET = 11
};

/* Wray RK3 coefficients (src/mod_parameters.f90:64-71). */
static const double RK3_A[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
static const double RK3_B[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0};
static const double RK3_G[3] = {8.0 / 15.0, 2.0 / 15.0, 1.0 / 3.0};
static const double RK3_C[3] = {0.0, 8.0 / 15.0, 2.0 / 3.0};

/* grad(i,j,c) = d u_j / d x_i, column-major like the reference grad_U(3,3,nc). */
#define GI(c, i, j) (9 * (size_t)(c) + 3 * (j) + (i))

struct om_solver {
  om_mesh m;
  om_params p;
  int nc, nf, nint, nb, npatch;
  double *U, *P, *p_old, *nu_sgs, *U_bnd, *p_bnd, *phi_prev, *phi_prev_bnd, *tau_w;
  double *grad_U, *gU_prev, *face_flux, *U_face, *inv_vol;
  double *H, *H0, *conv, *visc, *rhs_p, *div_w, *phi_it;
  double *gp, *gp_rc, *gp_cell, *p_tot, *p_tot_bnd, *gp_prev;
  double *Tinv, *outlet_pb, *Gf, *ws_y, *div_diag;
  int *per_nb, *per_bf, *nper_bf, n_per_bf, n_nper_bf, n_wm_patches, n_outlet;
  int *ws_cell;
  char *built;
  /* pressure operator (src/solver/mod_pressure_solver.f90:76-96) */
  int npf, n_dir;
  int *ps_owner, *ps_nb, *ps_fmesh, *ps_skip, *dir_rows;
  double *ps_coeff, *ps_diag, *dir_vals;
  double *k_int, *k_bnd;
  char *act_int, *act_bnd;
  int skip_nonortho, skip_grad_p, n_corr, diag_interface;
  double V_min, V_total, p_atol_factor, time;
  int u_face_ready;
  int64_t n_solves, n_iters, last_iters;
  double *cg_r, *cg_z, *cg_p, *cg_q, *cg_b;
  double *dt_flux, *dt_diff;
  /* temperature */
  double *T, *T_bnd, *grad_T, *gT_prev, *H_T, *H_T0, *conv_T, *visc_T;
  /* MOST: per-face heat flux and Obukhov length, the configuration's L */
  double *q_w, *L_bnd, L_cfg;
  /* exact cell projection: pressure faces that are not skipped, outlet boundary faces */
  int cp_n, cp_nout;
  int *cp_own, *cp_nbr, *cp_bcell;
  double *cp_sf, *cp_gc, *cp_bsf;
  int64_t n_proj, n_proj_its;
  double mff[3]; /* constant mass flux forcing of the current step */
  /* passive scalars */
  double *C[OM_MAX_SCALARS], *C_bnd[OM_MAX_SCALARS], *grad_C[OM_MAX_SCALARS], *H_C[OM_MAX_SCALARS],
      *H_C0[OM_MAX_SCALARS], *gC_prev;
  /* convective outlet and backflow traction (src/bc/mod_bc.f90:61-63): per patch the mode, the
   * traction coefficient, the convection speed and the reference speed; per boundary face the
   * previous stage's boundary tendency; whether a stage has advanced the boundary state */
  int out_active, out_ready;
  int *out_conv;
  double *out_beta, *out_uc, *out_uref, *out_gprev;
  /* log-law wall model: filtered wall-parallel velocity per boundary face, whether it is seeded,
   * the time step the first stage of the current step hands the filter */
  double *wm_filt, wm_dt;
  int wm_filt_ready;
  /* MOST roughness lengths per patch */
  double *z0m_p, *z0h_p;
};

/* body_force_now (src/solver/mod_fv_solver.f90:2904-2923): constant or mass-flux force plus the
 * harmonic forcing at the solver's current (stage) time. */
static void body_force_now(const om_solver *s, double f[3]) {
  for (int k = 0; k < 3; ++k) f[k] = s->p.mass_flux ? s->mff[k] : s->p.f_const[k];
  if (s->p.use_harmonic) {
    const double *d = s->p.forcing_dir;
    const double dn = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double hd[3] = {1.0, 0.0, 0.0};
    if (dn > 1.0e-30)
      for (int k = 0; k < 3; ++k) hd[k] = d[k] / dn;
    const double hs = s->p.forcing_amp * cos(s->p.forcing_omega * s->time + s->p.forcing_phase);
    for (int k = 0; k < 3; ++k) f[k] = f[k] + hs * hd[k];
  }
}

static double *dz(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }
static int *iz(size_t n) { return (int *)calloc(n ? n : 1, sizeof(int)); }

static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static int bc_of_face(const om_solver *s, int f) {
  const int pid = s->m.face_patch[f];
  if (pid >= 0 && pid < s->npatch) return s->m.patch_bc[pid];
  return BC_ZEROGRADIENT;
}

/* refine_blend_weight (src/mesh/mod_mesh.f90:491-500) */
static double refine_blend_weight(double skew) {
  double t = (skew - SKEW_BLEND_LO) / (SKEW_BLEND_HI - SKEW_BLEND_LO);
  t = fmax(0.0, fmin(1.0, t));
  return 1.0 - t * t * (3.0 - 2.0 * t);
}

/* nonortho_kfactor_limited (src/numerics/mod_nonortho.f90:65-98) */
static void kfactor_limited(const double sf[3], const double d_on[3], double min_cos, double k[3]) {
  const double d_on_len = sqrt(d_on[0] * d_on[0] + d_on[1] * d_on[1] + d_on[2] * d_on[2]);
  if (d_on_len < 1.0e-30) {
    k[0] = k[1] = k[2] = 0.0;
    return;
  }
  const double inv_len = 1.0 / d_on_len;
  const double s_dot_dhat = dot3(sf, d_on) * inv_len;
  double d_len = s_dot_dhat * d_on_len;
  const double sf_sq = sf[0] * sf[0] + sf[1] * sf[1] + sf[2] * sf[2];
  if (min_cos > 0.0) d_len = fmax(d_len, min_cos * sqrt(sf_sq) * d_on_len);
  if (d_len > 1.0e-30) {
    for (int i = 0; i < 3; ++i) k[i] = sf[i] - (sf_sq / d_len) * d_on[i];
  } else {
    for (int i = 0; i < 3; ++i) k[i] = sf[i] - (s_dot_dhat * inv_len) * d_on[i];
  }
}

/* nonortho_cross_flux (src/numerics/mod_nonortho.f90:120-128) */
static double cross_flux(const double *go, const double *gn, double gc, const double *k) {
  double v[3];
  for (int i = 0; i < 3; ++i) v[i] = gc * go[i] + (1.0 - gc) * gn[i];
  return dot3(v, k);
}

/* resolve_periodic_nb (src/bc/mod_bc.f90:822-859), single rank: partner through the pair. */
static int resolve_periodic_nb(const om_solver *s, int b) {
  const int o = s->m.face_owner[s->nint + b];
  const int pair = s->m.periodic_pair[b];
  if (pair >= 0 && pair < s->nb) {
    const int nb = s->m.face_owner[s->nint + pair];
    if (nb >= 0 && nb != o) return nb;
  }
  return -1;
}

/* ---------------------------------------------------------------- boundary conditions */

/* apply_velocity_bc (src/bc/mod_bc.f90:196-379). filter: 0 all, 1 non-periodic, 2 periodic only. */
static void apply_velocity_bc(om_solver *s, int filter, const double *gU) {
  for (int b = 0; b < s->nb; ++b) {
    const int f = s->nint + b;
    const int bc = bc_of_face(s, f);
    if (filter == 2 && bc != BC_PERIODIC) continue;
    if (filter == 1 && bc == BC_PERIODIC) continue;
    const int o = s->m.face_owner[f], pid = s->m.face_patch[f];
    const double *n = &s->m.face_normal[3 * f];
    double *ub = &s->U_bnd[3 * b];
    const double *uo = &s->U[3 * o];
    switch (bc) {
      case BC_INLET:
        for (int k = 0; k < 3; ++k) ub[k] = s->p.patch_U_val[3 * pid + k];
        break;
      case BC_OUTLET:
      case BC_ZEROGRADIENT: {
        double u[3] = {uo[0], uo[1], uo[2]};
        if (gU) {
          const double d[3] = {s->m.face_centre[3 * f] - s->m.cell_centroid[3 * o],
                               s->m.face_centre[3 * f + 1] - s->m.cell_centroid[3 * o + 1],
                               s->m.face_centre[3 * f + 2] - s->m.cell_centroid[3 * o + 2]};
          for (int k = 0; k < 3; ++k)
            u[k] = u[k] + gU[GI(o, 0, k)] * d[0] + gU[GI(o, 1, k)] * d[1] + gU[GI(o, 2, k)] * d[2];
        }
        /* a convective outlet face is an RK state advanced by outlet_advance_boundary; only its
         * start-up value is extrapolated (src/bc/mod_bc.f90:304-311) */
        if (bc == BC_OUTLET && s->out_ready && pid >= 0 && pid < s->npatch && s->out_conv[pid]) break;
        for (int k = 0; k < 3; ++k) ub[k] = u[k];
        break;
      }
      case BC_WALL_NOSLIP:
      case BC_WALL_MODEL:
        ub[0] = ub[1] = ub[2] = 0.0;
        break;
      case BC_WALL_VELOCITY: {
        double u[3];
        for (int k = 0; k < 3; ++k) u[k] = s->p.patch_U_val[3 * pid + k];
        const double un = u[0] * n[0] + u[1] * n[1] + u[2] * n[2];
        for (int k = 0; k < 3; ++k) ub[k] = u[k] - un * n[k];
        break;
      }
      case BC_PERIODIC: {
        double u[3] = {uo[0], uo[1], uo[2]};
        const int nb = s->per_nb[b];
        if (nb >= 0) {
          const double w = s->m.periodic_gc[b];
          for (int k = 0; k < 3; ++k) u[k] = w * u[k] + (1.0 - w) * s->U[3 * nb + k];
        }
        for (int k = 0; k < 3; ++k) ub[k] = u[k];
        break;
      }
      case BC_SLIP:
      case BC_SYMMETRY: {
        const double un = uo[0] * n[0] + uo[1] * n[1] + uo[2] * n[2];
        for (int k = 0; k < 3; ++k) ub[k] = uo[k] - un * n[k];
        break;
      }
      default:
        break;
    }
  }
}

/* apply_pressure_bc (src/bc/mod_bc.f90:383-486). */
static void apply_pressure_bc(om_solver *s, const double *p, double *pb, int filter, const double *gp,
                              double outlet_scale, const double *pob) {
  for (int b = 0; b < s->nb; ++b) {
    const int f = s->nint + b;
    const int bc = bc_of_face(s, f);
    if (filter == 2 && bc != BC_PERIODIC) continue;
    if (filter == 1 && bc == BC_PERIODIC) continue;
    const int o = s->m.face_owner[f], pid = s->m.face_patch[f];
    switch (bc) {
      case BC_OUTLET:
        pb[b] = outlet_scale * (pob ? pob[b] : s->p.patch_p_val[pid]);
        break;
      case BC_PERIODIC: {
        const int nb = s->per_nb[b];
        pb[b] = (nb >= 0) ? s->m.periodic_gc[b] * p[o] + (1.0 - s->m.periodic_gc[b]) * p[nb] : p[o];
        break;
      }
      case BC_WALL_NOSLIP:
      case BC_WALL_VELOCITY:
      case BC_WALL_MODEL:
      case BC_SLIP:
      case BC_SYMMETRY: {
        /* Kim-Moin Neumann with a_wall = 0 (BCConfig default) */
        const double *n = &s->m.face_normal[3 * f];
        const double an = 0.0 * n[0] + 0.0 * n[1] + 0.0 * n[2];
        double d[3];
        for (int i = 0; i < 3; ++i) d[i] = s->m.face_centre[3 * f + i] - s->m.cell_centroid[3 * o + i];
        const double d_ortho = fabs(d[0] * n[0] + d[1] * n[1] + d[2] * n[2]);
        pb[b] = p[o] - an * d_ortho;
        break;
      }
      default:
        if (gp) {
          double d[3];
          for (int i = 0; i < 3; ++i) d[i] = s->m.face_centre[3 * f + i] - s->m.cell_centroid[3 * o + i];
          pb[b] = p[o] + gp[3 * o] * d[0] + gp[3 * o + 1] * d[1] + gp[3 * o + 2] * d[2];
        } else {
          pb[b] = p[o];
        }
    }
  }
}

/* apply_temperature_bc (src/bc/mod_bc.f90:489-580). filter: 0 all, 1 non-periodic, 2 periodic. */
/* apply_passive_scalar_bc (src/bc/mod_bc.f90:582-679) for scalar k: periodic faces interpolate
 * with the partner cell, other faces by the patch's mode: Dirichlet value, Neumann flux q (times
 * cos(omega t + phase) when omega != 0) as C_o + q d_cf / (D + nu_t / Pr_t) with the temperature's
 * Pr_t, else zero gradient. */
static void apply_scalar_bc(om_solver *s, int k, int filter) {
  double *C = s->C[k], *Cb = s->C_bnd[k];
  for (int b = 0; b < s->nb; ++b) {
    const int f = s->nint + b;
    const int bc = bc_of_face(s, f);
    if (filter == 2 && bc != BC_PERIODIC) continue;
    if (filter == 1 && bc == BC_PERIODIC) continue;
    const int o = s->m.face_owner[f], pid = s->m.face_patch[f];
    if (bc == BC_PERIODIC) {
      const int nb = s->per_nb[b];
      Cb[b] = (nb >= 0) ? s->m.periodic_gc[b] * C[o] + (1.0 - s->m.periodic_gc[b]) * C[nb] : C[o];
      continue;
    }
    const int q = OM_MAX_SCALARS * pid + k;
    const int mode = s->p.scalar_bc_mode ? s->p.scalar_bc_mode[q] : OM_SCALAR_ZERO_GRADIENT;
    if (mode == OM_SCALAR_DIRICHLET) {
      Cb[b] = s->p.scalar_bc_value[q];
    } else if (mode == OM_SCALAR_NEUMANN) {
      double fl = s->p.scalar_bc_flux[q];
      if (fabs(s->p.scalar_bc_omega[q]) > SMALL) fl = fl * cos(s->p.scalar_bc_omega[q] * s->time + s->p.scalar_bc_phase[q]);
      const double af = s->p.scalar_diffusivity[k] + s->nu_sgs[o] / s->p.Pr_t;
      const double dcf = fmax(s->m.face_dcf[f], SMALL);
      Cb[b] = (af > SMALL) ? C[o] + fl * dcf / af : C[o];
    } else {
      Cb[b] = C[o];
    }
  }
}

static void apply_temperature_bc(om_solver *s, int filter, const double *gT) {
  if (!s->p.boussinesq) return;
  for (int b = 0; b < s->nb; ++b) {
    const int f = s->nint + b;
    const int bc = bc_of_face(s, f);
    if (filter == 2 && bc != BC_PERIODIC) continue;
    if (filter == 1 && bc == BC_PERIODIC) continue;
    const int o = s->m.face_owner[f], pid = s->m.face_patch[f];
    switch (bc) {
      case BC_INLET:
      case BC_FIXEDTEMP:
        s->T_bnd[b] = s->p.patch_T_val[pid];
        break;
      case BC_PERIODIC: {
        const int nb = s->per_nb[b];
        s->T_bnd[b] = (nb >= 0) ? s->m.periodic_gc[b] * s->T[o] + (1.0 - s->m.periodic_gc[b]) * s->T[nb] : s->T[o];
        break;
      }
      default:
        if (gT) {
          double d[3];
          for (int i = 0; i < 3; ++i) d[i] = s->m.face_centre[3 * f + i] - s->m.cell_centroid[3 * o + i];
          s->T_bnd[b] = s->T[o] + gT[3 * o] * d[0] + gT[3 * o + 1] * d[1] + gT[3 * o + 2] * d[2];
        } else {
          s->T_bnd[b] = s->T[o];
        }
    }
  }
}

/* ---------------------------------------------------------------- gradients */

/* ls_gradient_vector (src/numerics/mod_gradient.f90:425-556, CPU branch) */
static void ls_gradient_vector(const om_solver *s, const double *U, const double *Ub, double *g) {
  memset(g, 0, sizeof(double) * 9 * (size_t)s->nc);
  for (int c = 0; c < s->nc; ++c) {
    for (int k = s->m.ls_offset[c]; k < s->m.ls_offset[c + 1]; ++k) {
      const int nb = s->m.ls_cell[k];
      double du[3];
      if (nb < 0) {
        const int f = s->m.ls_face[k];
        const int bi = (f >= 0) ? f - s->nint : -1;
        const double *un = (bi >= 0 && bi < s->nb) ? &Ub[3 * bi] : &U[3 * c];
        for (int j = 0; j < 3; ++j) du[j] = un[j] - U[3 * c + j];
      } else {
        for (int j = 0; j < 3; ++j) du[j] = U[3 * nb + j] - U[3 * c + j];
      }
      const double *w = &s->m.ls_weight[3 * (size_t)k];
      for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) g[GI(c, i, j)] = g[GI(c, i, j)] + w[i] * du[j];
    }
  }
}

/* ls_gradient_scalar (src/numerics/mod_gradient.f90:307-422, CPU branch) */
static void ls_gradient_scalar(const om_solver *s, const double *phi, const double *pb, double *g) {
  memset(g, 0, sizeof(double) * 3 * (size_t)s->nc);
  for (int c = 0; c < s->nc; ++c) {
    for (int k = s->m.ls_offset[c]; k < s->m.ls_offset[c + 1]; ++k) {
      const int nb = s->m.ls_cell[k];
      double pn;
      if (nb < 0) {
        const int f = s->m.ls_face[k];
        const int bi = (f >= 0) ? f - s->nint : -1;
        pn = (bi >= 0 && bi < s->nb) ? pb[bi] : phi[c];
      } else {
        pn = phi[nb];
      }
      const double d = pn - phi[c];
      const double *w = &s->m.ls_weight[3 * (size_t)k];
      for (int i = 0; i < 3; ++i) g[3 * c + i] = g[3 * c + i] + w[i] * d;
    }
  }
}

/* green_gauss_gradient_vector (src/numerics/mod_gradient.f90:172-304) */
static void gg_gradient_vector(const om_solver *s, const double *U, const double *Ub, const double *gprev, double *g) {
  memset(g, 0, sizeof(double) * 9 * (size_t)s->nc);
  for (int f = 0; f < s->nf; ++f) {
    const int o = s->m.face_owner[f], nb = s->m.face_neighbour[f];
    const double *sf = &s->m.face_sf[3 * f];
    double Uf[3];
    if (nb >= 0) {
      const double gc = s->m.face_gc[f];
      if (gprev) {
        const double *eps = &s->m.face_eps[3 * f];
        for (int k = 0; k < 3; ++k) {
          double gm[3];
          for (int i = 0; i < 3; ++i) gm[i] = gc * gprev[GI(o, i, k)] + (1.0 - gc) * gprev[GI(nb, i, k)];
          Uf[k] = gc * U[3 * o + k] + (1.0 - gc) * U[3 * nb + k] + dot3(gm, eps);
        }
        const double w = refine_blend_weight(s->m.face_skew[f]);
        if (w < 1.0)
          for (int k = 0; k < 3; ++k) {
            const double pl = gc * U[3 * o + k] + (1.0 - gc) * U[3 * nb + k];
            Uf[k] = pl + w * (Uf[k] - pl);
          }
      } else {
        for (int k = 0; k < 3; ++k) Uf[k] = gc * U[3 * o + k] + (1.0 - gc) * U[3 * nb + k];
      }
    } else {
      for (int k = 0; k < 3; ++k) Uf[k] = Ub[3 * (f - s->nint) + k];
    }
    for (int k = 0; k < 3; ++k)
      for (int i = 0; i < 3; ++i) {
        g[GI(o, i, k)] = g[GI(o, i, k)] + Uf[k] * sf[i];
        if (nb >= 0) g[GI(nb, i, k)] = g[GI(nb, i, k)] - Uf[k] * sf[i];
      }
  }
  for (int c = 0; c < s->nc; ++c)
    if (s->m.cell_volume[c] > SMALL) {
      const double iv = 1.0 / s->m.cell_volume[c];
      for (int q = 0; q < 9; ++q) g[9 * (size_t)c + q] = g[9 * (size_t)c + q] * iv;
    }
}

/* green_gauss_gradient_scalar (src/numerics/mod_gradient.f90:38-168) */
static void gg_gradient_scalar(const om_solver *s, const double *phi, const double *pb, const double *gprev, double *g) {
  memset(g, 0, sizeof(double) * 3 * (size_t)s->nc);
  for (int f = 0; f < s->nf; ++f) {
    const int o = s->m.face_owner[f], nb = s->m.face_neighbour[f];
    const double *sf = &s->m.face_sf[3 * f];
    double pf;
    if (nb >= 0) {
      const double gc = s->m.face_gc[f];
      if (gprev) {
        double gm[3];
        for (int i = 0; i < 3; ++i) gm[i] = gc * gprev[3 * o + i] + (1.0 - gc) * gprev[3 * nb + i];
        const double pl = gc * phi[o] + (1.0 - gc) * phi[nb];
        pf = pl + dot3(gm, &s->m.face_eps[3 * f]);
        const double w = refine_blend_weight(s->m.face_skew[f]);
        if (w < 1.0) pf = pl + w * (pf - pl);
      } else {
        pf = gc * phi[o] + (1.0 - gc) * phi[nb];
      }
      for (int i = 0; i < 3; ++i) {
        g[3 * o + i] = g[3 * o + i] + pf * sf[i];
        g[3 * nb + i] = g[3 * nb + i] - pf * sf[i];
      }
    } else {
      pf = pb[f - s->nint];
      for (int i = 0; i < 3; ++i) g[3 * o + i] = g[3 * o + i] + pf * sf[i];
    }
  }
  for (int c = 0; c < s->nc; ++c)
    if (s->m.cell_volume[c] > SMALL) {
      const double iv = 1.0 / s->m.cell_volume[c];
      for (int i = 0; i < 3; ++i) g[3 * c + i] = g[3 * c + i] * iv;
    }
}

/* compute_all_gradients (src/solver/mod_fv_solver.f90:1003-1077), velocity only */
static void compute_all_gradients(om_solver *s) {
  const int temp = s->p.boussinesq, ns = s->p.n_scalars;
  if (s->p.gradient == OM_GRAD_GG) {
    gg_gradient_vector(s, s->U, s->U_bnd, NULL, s->grad_U);
    if (temp) gg_gradient_scalar(s, s->T, s->T_bnd, NULL, s->grad_T);
    for (int k = 0; k < ns; ++k) gg_gradient_scalar(s, s->C[k], s->C_bnd[k], NULL, s->grad_C[k]);
    for (int it = 2; it <= s->p.gg_iterations; ++it) {
      memcpy(s->gU_prev, s->grad_U, sizeof(double) * 9 * (size_t)s->nc);
      gg_gradient_vector(s, s->U, s->U_bnd, s->gU_prev, s->grad_U);
      if (temp) {
        memcpy(s->gT_prev, s->grad_T, sizeof(double) * 3 * (size_t)s->nc);
        gg_gradient_scalar(s, s->T, s->T_bnd, s->gT_prev, s->grad_T);
      }
      for (int k = 0; k < ns; ++k) {
        memcpy(s->gC_prev, s->grad_C[k], sizeof(double) * 3 * (size_t)s->nc);
        gg_gradient_scalar(s, s->C[k], s->C_bnd[k], s->gC_prev, s->grad_C[k]);
      }
    }
  } else {
    ls_gradient_vector(s, s->U, s->U_bnd, s->grad_U);
    if (temp) ls_gradient_scalar(s, s->T, s->T_bnd, s->grad_T);
    for (int k = 0; k < ns; ++k) ls_gradient_scalar(s, s->C[k], s->C_bnd[k], s->grad_C[k]);
  }
}

/* gradient_of_scalar (src/solver/mod_fv_solver.f90:2341-2369) */
static void gradient_of_scalar(om_solver *s, const double *fld, const double *fb, double *g) {
  if (s->p.gradient == OM_GRAD_GG) {
    gg_gradient_scalar(s, fld, fb, NULL, g);
    for (int it = 2; it <= s->p.gg_iterations; ++it) {
      memcpy(s->gp_prev, g, sizeof(double) * 3 * (size_t)s->nc);
      gg_gradient_scalar(s, fld, fb, s->gp_prev, g);
    }
  } else {
    ls_gradient_scalar(s, fld, fb, g);
  }
}

static void compute_pressure_gradient(om_solver *s) {
  if (s->skip_grad_p) return;
  gradient_of_scalar(s, s->P, s->p_bnd, s->gp);
}

/* refresh_halos_and_bc + apply_all_bc_subset (src/solver/mod_fv_solver.f90:985-1000, 2025-2060):
 * single rank, so the non-periodic and periodic passes run back to back. */
static void refresh_bc(om_solver *s) {
  for (int pass = 1; pass <= 2; ++pass) {
    apply_velocity_bc(s, pass, s->grad_U);
    apply_pressure_bc(s, s->P, s->p_bnd, pass, s->skip_grad_p ? NULL : s->gp, 1.0, s->outlet_pb);
    apply_temperature_bc(s, pass, s->grad_T);
    for (int k = 0; k < s->p.n_scalars; ++k) apply_scalar_bc(s, k, pass);
  }
}

static void velocity_bc_both(om_solver *s) {
  apply_velocity_bc(s, 1, s->grad_U);
  apply_velocity_bc(s, 2, NULL);
}

/* ---------------------------------------------------------------- SGS and wall model */

static double det3(const double *A) {
  /* A column-major 3x3: A(i,j) = A[3j+i] (det3x3, src/sgs/mod_vreman.f90:38-44) */
#define A_(i, j) A[3 * ((j)-1) + ((i)-1)]
  return A_(1, 1) * (A_(2, 2) * A_(3, 3) - A_(2, 3) * A_(3, 2)) - A_(1, 2) * (A_(2, 1) * A_(3, 3) - A_(2, 3) * A_(3, 1)) +
         A_(1, 3) * (A_(2, 1) * A_(3, 2) - A_(2, 2) * A_(3, 1));
#undef A_
}

/* build_filter_tensor + smooth_filter_tensor (src/sgs/mod_vreman.f90:168-264) */
static void build_filter_tensor(om_solver *s) {
  const int nc = s->nc;
  for (int c = 0; c < nc; ++c) {
    const double V = s->m.cell_volume[c];
    const double d2_iso = pow(V, 2.0 / 3.0);
    double g[9] = {0};
    for (int q = s->m.cell_face_offset[c]; q < s->m.cell_face_offset[c + 1]; ++q) {
      const int f = s->m.cell_faces[q];
      double dx[3];
      for (int i = 0; i < 3; ++i) dx[i] = s->m.face_centre[3 * f + i] - s->m.cell_centroid[3 * c + i];
      const double a = s->m.face_area[f];
      for (int n = 0; n < 3; ++n)
        for (int mm = 0; mm < 3; ++mm) g[3 * n + mm] = g[3 * n + mm] + a * dx[mm] * dx[n];
    }
    const double det = det3(g);
    const double t = (g[0] + g[4] + g[8]) / 3.0;
    const double tr3 = t * t * t;
    double *G = &s->Gf[9 * (size_t)c];
    if (tr3 > 0.0 && det > 1.0e-12 * tr3) {
      const double sc = pow(V * V / det, 1.0 / 3.0);
      for (int q = 0; q < 9; ++q) G[q] = sc * g[q];
    } else {
      for (int q = 0; q < 9; ++q) G[q] = 0.0;
      G[0] = G[4] = G[8] = d2_iso;
    }
  }
  double *gnb = dz(9 * (size_t)nc), *ws = dz(nc);
  for (int pass = 0; pass < 2; ++pass) {
    memset(gnb, 0, sizeof(double) * 9 * (size_t)nc);
    memset(ws, 0, sizeof(double) * nc);
    for (int f = 0; f < s->nf; ++f) {
      const int o = s->m.face_owner[f], nb = s->m.face_neighbour[f];
      if (nb < 0) continue;
      const double a = s->m.face_area[f];
      for (int q = 0; q < 9; ++q) gnb[9 * (size_t)o + q] = gnb[9 * (size_t)o + q] + a * s->Gf[9 * (size_t)nb + q];
      ws[o] = ws[o] + a;
      for (int q = 0; q < 9; ++q) gnb[9 * (size_t)nb + q] = gnb[9 * (size_t)nb + q] + a * s->Gf[9 * (size_t)o + q];
      ws[nb] = ws[nb] + a;
    }
    for (int c = 0; c < nc; ++c)
      if (ws[c] > SMALL)
        for (int q = 0; q < 9; ++q)
          s->Gf[9 * (size_t)c + q] = (1.0 - 0.5) * s->Gf[9 * (size_t)c + q] + 0.5 * gnb[9 * (size_t)c + q] / ws[c];
  }
  free(gnb);
  free(ws);
}

/* vreman_sgs_viscosity (src/sgs/mod_vreman.f90:50-153) */
static void vreman(om_solver *s, double dt_stab) {
  for (int c = 0; c < s->nc; ++c) {
    const double *al = &s->grad_U[9 * (size_t)c]; /* al[3j+i] = alpha(i,j) */
    const double *G = &s->Gf[9 * (size_t)c];      /* G[3n+m] = G(m,n) */
    double beta[9] = {0};
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i <= j; ++i) {
        double b = 0.0;
        for (int n = 0; n < 3; ++n)
          for (int mm = 0; mm < 3; ++mm) b = b + G[3 * n + mm] * al[3 * i + mm] * al[3 * j + n];
        beta[3 * j + i] = b;
      }
    double asq = 0.0;
    /* alpha(1,1)**2 + alpha(1,2)**2 + alpha(1,3)**2 + alpha(2,1)**2 + ... (row-wise) */
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) asq = asq + al[3 * j + i] * al[3 * j + i];
#define B_(i, j) beta[3 * ((j)-1) + ((i)-1)]
    const double Bb = B_(1, 1) * B_(2, 2) - B_(1, 2) * B_(1, 2) + B_(1, 1) * B_(3, 3) - B_(1, 3) * B_(1, 3) +
                      B_(2, 2) * B_(3, 3) - B_(2, 3) * B_(2, 3);
#undef B_
    double nu = (asq > SMALL && Bb > 0.0) ? s->p.c_vreman * sqrt(Bb / asq) : 0.0;
    const double cap = NU_SGS_STAB_FRAC * pow(s->m.cell_volume[c], 2.0 / 3.0) / fmax(dt_stab, SMALL);
    s->nu_sgs[c] = fmin(nu, cap);
  }
}

/* smagorinsky_sgs (src/sgs/mod_vreman.f90:295-364) */
static void smagorinsky(om_solver *s, double dt_stab) {
  for (int c = 0; c < s->nc; ++c) {
    const double *g = &s->grad_U[9 * (size_t)c];
    double S[9], ss = 0.0;
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) S[3 * j + i] = 0.5 * (g[3 * j + i] + g[3 * i + j]);
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) ss = ss + S[3 * j + i] * S[3 * j + i];
    const double Sn = sqrt(2.0 * ss);
    const double delta = fmax(pow(fmax(det3(&s->Gf[9 * (size_t)c]), 0.0), 1.0 / 6.0), SMALL);
    const double csd = s->p.cs_smag * delta;
    double nu = csd * csd * Sn;
    const double cap = NU_SGS_STAB_FRAC * delta * delta / fmax(dt_stab, SMALL);
    s->nu_sgs[c] = fmin(nu, cap);
  }
}

/* amd_point_viscosity (src/sgs/mod_amd.f90:31-66): al[3j+i] = alpha(i,j), G[3j+i] = G(i,j) */
static double amd_point(const double *al, const double *G, double c_amd) {
  double asq = 0.0;
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) asq = asq + al[3 * j + i] * al[3 * j + i];
  double num = 0.0;
  for (int n = 0; n < 3; ++n)
    for (int m = 0; m < 3; ++m) {
      double beta = 0.0;
      for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) beta = beta + G[3 * j + i] * al[3 * m + i] * al[3 * n + j];
      num = num - beta * al[3 * n + m];
    }
  const double trg = G[0] + G[4] + G[8];
  if (asq > 1.0e-20 && num > 1.0e-12 * trg * asq * sqrt(asq)) return c_amd * num / asq;
  return 0.0;
}

double om_amd_point(const double alpha[9], const double G[9], double c_amd) { return amd_point(alpha, G, c_amd); }

static double wall_utau(double u_par, double y_w, double nu);

/* amd_sgs_viscosity (src/sgs/mod_amd.f90:69-207): point closure and cap in every cell, then, for
 * wall-model faces in boundary order, the owner's alpha with its wall-parallel wall-normal
 * derivative swapped for the log-law u_tau / (kappa y) (last face wins). */
static void amd(om_solver *s, double dt_stab) {
  const double c = s->p.c_amd;
  for (int cc = 0; cc < s->nc; ++cc) {
    double nu = amd_point(&s->grad_U[9 * (size_t)cc], &s->Gf[9 * (size_t)cc], c);
    const double cap = NU_SGS_STAB_FRAC * pow(s->m.cell_volume[cc], 2.0 / 3.0) / fmax(dt_stab, SMALL);
    s->nu_sgs[cc] = fmin(nu, cap);
  }
  if (!s->p.amd_wall_gradient) return;
  for (int b = 0; b < s->nb; ++b) {
    const int f = s->nint + b;
    if (bc_of_face(s, f) != BC_WALL_MODEL) continue;
    const int o = s->m.face_owner[f];
    double nrm[3], dx[3];
    for (int i = 0; i < 3; ++i) {
      nrm[i] = -s->m.face_normal[3 * f + i];
      dx[i] = s->m.face_centre[3 * f + i] - s->m.cell_centroid[3 * o + i];
    }
    const double y_w = fabs(dot3(dx, nrm));
    const double *u = &s->U[3 * (size_t)o];
    const double un = dot3(u, nrm);
    double ut[3];
    for (int i = 0; i < 3; ++i) ut[i] = u[i] - un * nrm[i];
    const double ut_mag = sqrt(ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2]);
    if (ut_mag <= 1.0e-14 || y_w <= SMALL) continue;
    const double u_tau = wall_utau(ut_mag, y_w, s->p.nu);
    const double g_wall = u_tau / (KAPPA * y_w);
    double al[9], d_old[3], d_t[3];
    memcpy(al, &s->grad_U[9 * (size_t)o], sizeof(al));
    for (int j = 0; j < 3; ++j) d_old[j] = nrm[0] * al[3 * j + 0] + nrm[1] * al[3 * j + 1] + nrm[2] * al[3 * j + 2];
    const double dn = dot3(d_old, nrm);
    for (int j = 0; j < 3; ++j) d_t[j] = g_wall * ut[j] / ut_mag - (d_old[j] - dn * nrm[j]);
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) al[3 * j + i] = al[3 * j + i] + nrm[i] * d_t[j];
    const double nu = amd_point(al, &s->Gf[9 * (size_t)o], c);
    const double cap = NU_SGS_STAB_FRAC * pow(s->m.cell_volume[o], 2.0 / 3.0) / fmax(dt_stab, SMALL);
    s->nu_sgs[o] = fmin(nu, cap);
  }
}

/* dynamic_smagorinsky_sgs (src/sgs/mod_dynamic_smagorinsky.f90:39-186): Germano-Lilly with the
 * compact test filter (the volume-weighted mean of a cell and its face neighbours, periodic
 * partners through the pair's owner, :238-304), Cs^2 = <L_ij M_ij> / <M_ij M_ij> clipped at 0,
 * nu_t = Cs^2 Delta^2 |S|, Delta = det(G)^(1/6) of the Vreman filter tensor (mod_vreman.f90:367-381),
 * capped by 0.4 Delta^2 / dt_stab. ALPHA_SQ = 4 (test-to-grid filter width ratio squared). */
static void dyn_test_filter(const om_solver *s, int ncomp, const double *in, double *out) {
  const int nc = s->nc;
  for (int c = 0; c < nc; ++c) {
    double wsum = s->m.cell_volume[c];
    for (int k = 0; k < ncomp; ++k) out[(size_t)ncomp * c + k] = wsum * in[(size_t)ncomp * c + k];
    for (int q = s->m.cell_face_offset[c]; q < s->m.cell_face_offset[c + 1]; ++q) {
      const int fj = s->m.cell_faces[q];
      if (fj < 0) continue;
      int nb = s->m.face_neighbour[fj];
      if (nb == c) nb = s->m.face_owner[fj];
      if (nb < 0 && fj >= s->nint) {
        const int pair = s->m.periodic_pair[fj - s->nint];
        if (pair >= 0) nb = s->m.face_owner[s->nint + pair];
      }
      if (nb < 0 || nb >= nc || nb == c) continue;
      const double w = s->m.cell_volume[nb];
      wsum = wsum + w;
      for (int k = 0; k < ncomp; ++k)
        out[(size_t)ncomp * c + k] = out[(size_t)ncomp * c + k] + w * in[(size_t)ncomp * nb + k];
    }
    for (int k = 0; k < ncomp; ++k) out[(size_t)ncomp * c + k] = out[(size_t)ncomp * c + k] / fmax(wsum, SMALL);
  }
}

/* |S| of a gradient g[3j+i] = d u_j / d x_i, summed over S(i,j) in column order (s_norm_of, :188-201) */
static double dyn_s_norm(const double *g) {
  double ss = 0.0;
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) {
      const double sij = 0.5 * (g[3 * j + i] + g[3 * i + j]);
      ss = ss + sij * sij;
    }
  return sqrt(2.0 * ss);
}

static void dynamic_smag(om_solver *s, double dt_stab) {
  const int nc = s->nc;
  double *fb = dz(15 * (size_t)nc), *fh = dz(15 * (size_t)nc), *uh = dz(3 * (size_t)nc);
  double *gh = dz(9 * (size_t)nc), *lm = dz(2 * (size_t)nc), *lmh = dz(2 * (size_t)nc);
  for (int c = 0; c < nc; ++c) {
    const double *u = &s->U[3 * (size_t)c], *g = &s->grad_U[9 * (size_t)c];
    double *f = &fb[15 * (size_t)c];
    /* S(i,j) from grad_U(i,j) = g[3j+i] */
    const double S11 = g[0], S22 = g[4], S33 = g[8];
    const double S12 = 0.5 * (g[3] + g[1]), S13 = 0.5 * (g[6] + g[2]), S23 = 0.5 * (g[7] + g[5]);
    const double Sn = sqrt(2.0 * (S11 * S11 + S22 * S22 + S33 * S33 + 2.0 * (S12 * S12 + S13 * S13 + S23 * S23)));
    f[0] = u[0]; f[1] = u[1]; f[2] = u[2];
    f[3] = u[0] * u[0]; f[4] = u[0] * u[1]; f[5] = u[0] * u[2];
    f[6] = u[1] * u[1]; f[7] = u[1] * u[2]; f[8] = u[2] * u[2];
    f[9] = Sn * S11; f[10] = Sn * S12; f[11] = Sn * S13; f[12] = Sn * S22; f[13] = Sn * S23; f[14] = Sn * S33;
  }
  dyn_test_filter(s, 15, fb, fh);
  for (int c = 0; c < nc; ++c)
    for (int k = 0; k < 3; ++k) uh[3 * (size_t)c + k] = fh[15 * (size_t)c + k];
  gg_gradient_vector(s, uh, s->U_bnd, NULL, gh);
  for (int c = 0; c < nc; ++c) {
    /* germano_lilly_terms (:203-236) */
    const double *h = &fh[15 * (size_t)c], *g = &gh[9 * (size_t)c];
    double Sh[9], ss = 0.0;
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) Sh[3 * j + i] = 0.5 * (g[3 * j + i] + g[3 * i + j]);
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) ss = ss + Sh[3 * j + i] * Sh[3 * j + i];
    const double Shn = sqrt(2.0 * ss);
    const double L11 = h[3] - h[0] * h[0], L12 = h[4] - h[0] * h[1], L13 = h[5] - h[0] * h[2];
    const double L22 = h[6] - h[1] * h[1], L23 = h[7] - h[1] * h[2], L33 = h[8] - h[2] * h[2];
    const double d = fmax(pow(fmax(det3(&s->Gf[9 * (size_t)c]), 0.0), 1.0 / 6.0), SMALL);
    const double d2 = d * d;
    /* Sh[3j+i] = Shat(i,j); symmetric */
    const double M11 = d2 * (2.0 * (h[9] - 4.0 * Shn * Sh[0])), M12 = d2 * (2.0 * (h[10] - 4.0 * Shn * Sh[3]));
    const double M13 = d2 * (2.0 * (h[11] - 4.0 * Shn * Sh[6])), M22 = d2 * (2.0 * (h[12] - 4.0 * Shn * Sh[4]));
    const double M23 = d2 * (2.0 * (h[13] - 4.0 * Shn * Sh[7])), M33 = d2 * (2.0 * (h[14] - 4.0 * Shn * Sh[8]));
    lm[2 * (size_t)c] = L11 * M11 + L22 * M22 + L33 * M33 + 2.0 * (L12 * M12 + L13 * M13 + L23 * M23);
    lm[2 * (size_t)c + 1] = M11 * M11 + M22 * M22 + M33 * M33 + 2.0 * (M12 * M12 + M13 * M13 + M23 * M23);
  }
  dyn_test_filter(s, 2, lm, lmh);
  for (int c = 0; c < nc; ++c) {
    const double d = fmax(pow(fmax(det3(&s->Gf[9 * (size_t)c]), 0.0), 1.0 / 6.0), SMALL);
    double nu = fmax(lmh[2 * (size_t)c] / fmax(lmh[2 * (size_t)c + 1], SMALL), 0.0) * (d * d) *
                dyn_s_norm(&s->grad_U[9 * (size_t)c]);
    const double cap = NU_SGS_STAB_FRAC * (d * d) / fmax(dt_stab, SMALL);
    s->nu_sgs[c] = fmin(nu, cap);
  }
  free(fb); free(fh); free(uh); free(gh); free(lm); free(lmh);
}

/* reichardt_uplus / wall_stress_newton (src/sgs/mod_wall_model.f90:50-115) */
static void reichardt(double yp, double cbuf, double *up, double *dup) {
  const double A1 = 11.0, A2 = 0.33, A1I = 1.0 / 11.0;
  const double e1 = exp(-yp * A1I), e2 = exp(-A2 * yp);
  *up = log(1.0 + KAPPA * yp) / KAPPA + cbuf * (1.0 - e1 - yp * A1I * e2);
  *dup = 1.0 / (1.0 + KAPPA * yp) + cbuf * (A1I * e1 - A1I * e2 + yp * A1I * A2 * e2);
  (void)A1;
}

static double wall_utau(double u_par, double y_w, double nu) {
  const double cbuf = BLOG - log(KAPPA) / KAPPA;
  if (u_par <= SMALL || y_w <= SMALL) return 0.0;
  double ut = fmax(0.01 * u_par, sqrt(nu * u_par / (y_w + SMALL)));
  const double inv_nu = 1.0 / (nu + SMALL), ywn = y_w * inv_nu;
  for (int it = 0; it < 50; ++it) {
    double up, dup;
    reichardt(ut * ywn, cbuf, &up, &dup);
    const double F = u_par / (ut + SMALL) - up;
    const double dF = -u_par / (ut * ut + SMALL) - dup * ywn;
    double un = ut - F / (dF + copysign(SMALL, dF));
    un = fmin(fmax(un, 0.1 * ut), 10.0 * ut);
    if (fabs(un - ut) < 1.0e-8 * un) {
      ut = un;
      break;
    }
    ut = un;
  }
  return ut;
}

static double wall_tau(double u_par, double y_w, double nu) {
  const double ut = wall_utau(u_par, y_w, nu);
  return ut * ut;
}

/* apply_wall_model_bc (src/sgs/mod_wall_model.f90:176-313): the wall-parallel velocity of the
 * sampling cell, optionally time-filtered (exponential with w = dt / (T + dt), or adaptive with
 * eps = min(1, dt kappa u_tau / (f y)) from the previous filtered velocity; stages after the
 * first pass dt = 0, so the filter advances once per step; the first filtered call seeds it),
 * then the Reichardt law. dt_use is the step's dt at the first stage, else 0. */
static void wall_model_dt(om_solver *s, double dt_use) {
  const double t_use = s->p.wall_model_filter_time, tf_use = s->p.wall_model_filter_tf_fac;
  const int adaptive = tf_use > SMALL;
  const int do_filter = t_use > SMALL || adaptive;
  const int seed = !s->wm_filt_ready;
  const double w_filt = (do_filter && !adaptive) ? dt_use / (t_use + dt_use) : 0.0;
  for (int b = 0; b < s->nb; ++b) {
    const int f = s->nint + b;
    if (bc_of_face(s, f) != BC_WALL_MODEL) continue;
    const double *n = &s->m.face_normal[3 * f];
    const int c = s->ws_cell[b];
    const double y_w = s->ws_y[b];
    const double *u = &s->U[3 * c];
    const double un = u[0] * n[0] + u[1] * n[1] + u[2] * n[2];
    double up[3] = {u[0] - un * n[0], u[1] - un * n[1], u[2] - un * n[2]};
    if (do_filter) {
      double *uf = &s->wm_filt[3 * b];
      if (seed) {
        for (int k = 0; k < 3; ++k) uf[k] = up[k];
      } else {
        double eps = w_filt;
        if (adaptive) {
          eps = 0.0;
          if (dt_use > 0.0) {
            const double prev = sqrt(uf[0] * uf[0] + uf[1] * uf[1] + uf[2] * uf[2]);
            const double ut_prev = wall_utau(prev, y_w, s->p.nu);
            eps = 1.0;
            if (ut_prev > SMALL) eps = fmin(1.0, dt_use * KAPPA * ut_prev / (tf_use * y_w));
          }
        }
        for (int k = 0; k < 3; ++k) uf[k] = uf[k] + eps * (up[k] - uf[k]);
      }
      for (int k = 0; k < 3; ++k) up[k] = uf[k];
    }
    const double mag = sqrt(up[0] * up[0] + up[1] * up[1] + up[2] * up[2]);
    const double tau = wall_tau(mag, y_w, s->p.nu);
    const double sc = -tau / (mag + SMALL);
    for (int k = 0; k < 3; ++k) s->tau_w[3 * b + k] = sc * up[k];
  }
  if (do_filter) s->wm_filt_ready = 1;
}

/* find_sample_cell (src/sgs/mod_wall_model.f90:126-173): from the wall face's owner, step to the
 * face neighbour farthest from the wall along the inward normal until the distance reaches h_wm
 * (at most 20 steps; interior faces only) */
static void find_sample_cell(const om_solver *s, int b, double h_wm, int *cell_out, double *y_out) {
  const om_mesh *m = &s->m;
  const int fg = s->nint + b;
  const double *n = &m->face_normal[3 * fg], *xw = &m->face_centre[3 * fg];
  const double n_in[3] = {-n[0], -n[1], -n[2]};
  int c = m->face_owner[fg];
  *cell_out = c;
  {
    const double d[3] = {m->cell_centroid[3 * c] - xw[0], m->cell_centroid[3 * c + 1] - xw[1],
                         m->cell_centroid[3 * c + 2] - xw[2]};
    *y_out = fabs(dot3(d, n_in)) + SMALL;
  }
  if (h_wm <= 0.0) return;
  for (int step = 1; step <= 20; ++step) {
    if (*y_out >= h_wm) break;
    int best = -1;
    double best_dist = *y_out;
    for (int q = m->cell_face_offset[c]; q < m->cell_face_offset[c + 1]; ++q) {
      const int fj = m->cell_faces[q];
      if (fj < 0 || fj >= s->nf) continue;
      if (m->face_neighbour[fj] < 0) continue;
      int other = m->face_owner[fj];
      if (other == c) other = m->face_neighbour[fj];
      const double d[3] = {m->cell_centroid[3 * other] - xw[0], m->cell_centroid[3 * other + 1] - xw[1],
                           m->cell_centroid[3 * other + 2] - xw[2]};
      const double dist = dot3(d, n_in);
      if (dist > best_dist) {
        best_dist = dist;
        best = other;
      }
    }
    if (best < 0) break;
    c = best;
    *cell_out = best;
    *y_out = best_dist;
  }
}

int om_wall_sample(const om_solver *s, int b, double *y) {
  if (y) *y = s->ws_y[b];
  return s->ws_cell[b];
}

int om_wall_filter_state(const om_solver *s, double *out) {
  if (out) memcpy(out, s->wm_filt, sizeof(double) * 3 * (size_t)s->nb);
  return s->wm_filt_ready;
}

/* ---------------------------------------------------------------- convective outlet */

/* outlet_dn_u (src/bc/mod_bc.f90:775-786): the Taylor expansion along the owner-to-face offset
 * removes the curvature term of the one-sided difference */
static double outlet_dn_u(double u_b, double u_o, double gdd, double gt, double len_d, double nd) {
  return nd * (2.0 * (u_b - u_o) / len_d - gdd) + gt;
}

double om_outlet_dn_u(double u_b, double u_o, double gdd, double gt, double len_d, double nd) {
  return outlet_dn_u(u_b, u_o, gdd, gt, len_d, nd);
}

/* outlet_advance_boundary (src/bc/mod_bc.f90:681-772): per outlet patch with a convective mode or
 * a backflow traction, the area means of u.n (clipped at 0: the convection speed) and of |u| (the
 * reference speed) from the stage-start cell velocity; then the convective faces' boundary
 * velocity advances one RK3W stage of du_b/dt = -U_c n.grad u. */
static void outlet_advance_boundary(om_solver *s, double dt, double a_k, double b_k) {
  if (!s->out_active) return;
  const om_mesh *m = &s->m;
  double *stat = dz(3 * (size_t)s->npatch);
  for (int b = 0; b < s->nb; ++b) {
    const int f = s->nint + b, q = m->face_patch[f];
    if (q < 0 || q >= s->npatch || m->patch_bc[q] != BC_OUTLET) continue;
    if (!s->out_conv[q] && s->out_beta[q] <= 0.0) continue;
    const int o = m->face_owner[f];
    const double *u = &s->U[3 * o], *n = &m->face_normal[3 * f];
    const double area = m->face_area[f];
    stat[3 * q] = stat[3 * q] + area * dot3(u, n);
    stat[3 * q + 1] = stat[3 * q + 1] + area * sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    stat[3 * q + 2] = stat[3 * q + 2] + area;
  }
  for (int q = 0; q < s->npatch; ++q) {
    if (stat[3 * q + 2] <= 0.0) continue;
    s->out_uref[q] = stat[3 * q + 1] / stat[3 * q + 2];
    s->out_uc[q] = fmax(stat[3 * q] / stat[3 * q + 2], 0.0);
  }
  free(stat);
  for (int b = 0; b < s->nb; ++b) {
    const int f = s->nint + b, q = m->face_patch[f];
    if (q < 0 || q >= s->npatch || m->patch_bc[q] != BC_OUTLET || !s->out_conv[q]) continue;
    const int o = m->face_owner[f];
    double dl[3];
    for (int i = 0; i < 3; ++i) dl[i] = m->face_centre[3 * f + i] - m->cell_centroid[3 * o + i];
    const double len_d = fmax(sqrt(dot3(dl, dl)), SMALL);
    for (int i = 0; i < 3; ++i) dl[i] = dl[i] / len_d;
    const double *n = &m->face_normal[3 * f];
    const double nd = dot3(n, dl);
    double tn[3];
    for (int i = 0; i < 3; ++i) tn[i] = n[i] - nd * dl[i];
    const double uc = s->out_uc[q];
    for (int k = 0; k < 3; ++k) {
      const double gdd = s->grad_U[GI(o, 0, k)] * dl[0] + s->grad_U[GI(o, 1, k)] * dl[1] + s->grad_U[GI(o, 2, k)] * dl[2];
      const double gt = s->grad_U[GI(o, 0, k)] * tn[0] + s->grad_U[GI(o, 1, k)] * tn[1] + s->grad_U[GI(o, 2, k)] * tn[2];
      const double g = -uc * outlet_dn_u(s->U_bnd[3 * b + k], s->U[3 * o + k], gdd, gt, len_d, nd);
      s->U_bnd[3 * b + k] = s->U_bnd[3 * b + k] + dt * (a_k * g + b_k * s->out_gprev[3 * b + k]);
      s->out_gprev[3 * b + k] = g;
    }
  }
  s->out_ready = 1;
}

/* update_outlet_pressure (src/bc/mod_bc.f90:789-818): p_b = p_val - beta |u|^2 S0 / 2 with the
 * backflow switch S0 = (1 - tanh(u.n / (U_ref delta))) / 2, delta = 0.05 (mod_parameters.f90:76) */
static void update_outlet_pressure(om_solver *s) {
  if (!s->out_active) return;
  const om_mesh *m = &s->m;
  for (int b = 0; b < s->nb; ++b) {
    const int f = s->nint + b, q = m->face_patch[f];
    if (q < 0 || q >= s->npatch || m->patch_bc[q] != BC_OUTLET || s->out_beta[q] <= 0.0) continue;
    const int o = m->face_owner[f];
    const double *u = &s->U[3 * o];
    const double un = dot3(u, &m->face_normal[3 * f]);
    const double u2 = dot3(u, u);
    const double ur = fmax(s->out_uref[q], SMALL);
    const double s0 = 0.5 * (1.0 - tanh(un / (ur * 0.05)));
    s->outlet_pb[b] = (s->p.patch_p_val ? s->p.patch_p_val[q] : 0.0) - s->out_beta[q] * 0.5 * u2 * s0;
  }
}

void om_outlet_speeds(const om_solver *s, double *out) {
  for (int q = 0; q < s->npatch; ++q) {
    out[2 * q] = s->out_uc[q];
    out[2 * q + 1] = s->out_uref[q];
  }
}

/* ---------------------------------------------------------------- MOST wall model */
#define G_STD 9.81
#define L_CLIP 5000.0
#define KB_INV 2.0
#define PI_D 3.14159265358979323846

/* abl_psi_m (src/abl/mod_abl_profiles.f90:51-68): stable -5 zeta, unstable Paulson (1970) */
double om_psi_m(double zeta) {
  if (fabs(zeta) < 1.0e-6) return 0.0;
  if (zeta > 0.0) return -5.0 * zeta;
  const double x = pow(1.0 - 16.0 * zeta, 0.25);
  return 2.0 * log((1.0 + x) / 2.0) + log((1.0 + x * x) / 2.0) - 2.0 * atan(x) + PI_D / 2.0;
}

/* abl_psi_h (:71-85) */
double om_psi_h(double zeta) {
  if (fabs(zeta) < 1.0e-6) return 0.0;
  if (zeta > 0.0) return -5.0 * zeta;
  const double y = pow(1.0 - 16.0 * zeta, 0.5);
  return 2.0 * log((1.0 + y) / 2.0);
}

/* most_residuals (src/abl/mod_abl_wall.f90:260-284) */
static void most_residuals(double u_star, double T_star, double L, double u_par, double T_o, double T_surf,
                           double y_w, double z0_m, double z0_h, double *Fu, double *FT) {
  double dm = log(y_w / z0_m) - om_psi_m(y_w / L) + om_psi_m(z0_m / L);
  double dh = log(y_w / z0_h) - om_psi_h(y_w / L) + om_psi_h(z0_h / L);
  if (fabs(dm) < SMALL) dm = copysign(SMALL, dm);
  if (fabs(dh) < SMALL) dh = copysign(SMALL, dh);
  *Fu = u_star - KAPPA * u_par / dm;
  *FT = T_star - KAPPA * (T_o - T_surf) / dh;
}

/* L_from (:226-240): L = u*^2 T_ref / (kappa g T*), clipped to [0.001, L_CLIP] in magnitude */
static double L_from(double u_star, double T_star, double T_ref) {
  if (fabs(T_star) > SMALL) {
    double L = u_star * u_star * T_ref / (KAPPA * G_STD * T_star);
    L = copysign(fmax(fabs(L), 0.001), L);
    if (fabs(L) > L_CLIP) L = copysign(L_CLIP, L);
    return L;
  }
  return L_CLIP;
}

/* most_iterate (:155-223): relaxed Newton on (Fu, FT) with a finite-difference Jacobian */
void om_most_iterate(double u_par, double T_o, double T_surf, double y_w, double z0_m, double z0_h, double L_init,
                     double out[3]) {
  const double T_ref = T_o;
  double u_star = fmax(SMALL, KAPPA * u_par / fmax(SMALL, log(y_w / z0_m)));
  double T_star = 0.1 * (T_o - T_surf);
  double L = fabs(L_init) < SMALL ? L_CLIP : L_init;
  for (int it = 0; it < 20; ++it) {
    L = copysign(fmax(fabs(L), 0.001), L);
    if (fabs(L) > L_CLIP) L = copysign(L_CLIP, L);
    double Fu, FT, Fup, FTp;
    most_residuals(u_star, T_star, L, u_par, T_o, T_surf, y_w, z0_m, z0_h, &Fu, &FT);
    if (fabs(Fu) < 1.0e-8 && fabs(FT) < 1.0e-8) break;
    const double du = fmax(1.0e-6, 1.0e-6 * fabs(u_star));
    const double dT = fmax(1.0e-6, 1.0e-6 * fabs(T_star));
    most_residuals(u_star + du, T_star, L_from(u_star + du, T_star, T_ref), u_par, T_o, T_surf, y_w, z0_m, z0_h,
                   &Fup, &FTp);
    const double a11 = (Fup - Fu) / du, a21 = (FTp - FT) / du;
    most_residuals(u_star, T_star + dT, L_from(u_star, T_star + dT, T_ref), u_par, T_o, T_surf, y_w, z0_m, z0_h,
                   &Fup, &FTp);
    const double a12 = (Fup - Fu) / dT, a22 = (FTp - FT) / dT;
    const double den = a11 * a22 - a12 * a21;
    if (fabs(den) < SMALL) break;
    const double dus = -(a22 * Fu - a12 * FT) / den;
    const double dts = -(a11 * FT - a21 * Fu) / den;
    u_star = u_star + 0.5 * dus;
    T_star = T_star + 0.5 * dts;
    u_star = fmax(SMALL, fmin(u_star, 0.5 * u_par));
    L = L_from(u_star, T_star, T_ref);
  }
  out[0] = u_star;
  out[1] = T_star;
  out[2] = L;
}

/* abl_wall_model (src/abl/mod_abl_wall.f90:64-153): per wall-model face, the MOST fluxes from the
 * owner cell; tau_w = -u*^2 u_par/|u_par|, q_w = -u* T*; at the first stage of a step the
 * configuration's L becomes the mean over the wall faces (compute_L). */
static void most_wall(om_solver *s, int update_L) {
  double L_sum = 0.0;
  int nwf = 0;
  for (int b = 0; b < s->nb; ++b) {
    const int f = s->nint + b;
    if (bc_of_face(s, f) != BC_WALL_MODEL) continue;
    const int o = s->m.face_owner[f];
    const double *n = &s->m.face_normal[3 * f];
    const double y_w = s->ws_y[b];
    const double T_o = s->p.boussinesq ? s->T[o] : s->p.T_ref;
    const double *u = &s->U[3 * o];
    const double un = u[0] * n[0] + u[1] * n[1] + u[2] * n[2];
    double up[3] = {u[0] - un * n[0], u[1] - un * n[1], u[2] - un * n[2]};
    const double mag = sqrt(up[0] * up[0] + up[1] * up[1] + up[2] * up[2]);
    for (int k = 0; k < 3; ++k) up[k] = mag > SMALL ? up[k] / mag : 0.0;
    double L_guess = s->L_bnd[b];
    if (fabs(L_guess) < SMALL) L_guess = s->L_cfg;
    double r[3];
    const int q = s->m.face_patch[f];
    om_most_iterate(mag, T_o, s->p.T_surface, y_w, s->z0m_p[q], s->z0h_p[q], L_guess, r);
    s->L_bnd[b] = r[2];
    L_sum = L_sum + r[2];
    nwf++;
    for (int k = 0; k < 3; ++k) s->tau_w[3 * b + k] = -r[0] * r[0] * up[k];
    s->q_w[b] = -r[0] * r[1];
  }
  if (s->p.compute_L && update_L && nwf > 0) s->L_cfg = L_sum / (double)nwf;
}

/* ---------------------------------------------------------------- explicit terms */

static int periodic_partner_soa(const om_solver *s, int b, int o) {
  /* nb_per resolution of the SoA kernels (mod_advection.f90:1108-1124): periodic_neighbor is
   * unset on one rank, so the partner comes from the pair's owner. */
  int nb_per = -1;
  const int pair = s->m.periodic_pair[b];
  if (pair >= 0) nb_per = s->m.face_owner[s->nint + pair];
  if (nb_per == o) nb_per = -1;
  return nb_per;
}

/* advection_rhs_vector_kep_soa (src/numerics/mod_advection.f90:1013-1197) */
static void advection_kep(om_solver *s, double *rhs) {
  const int nc = s->nc;
  memset(rhs, 0, sizeof(double) * 3 * (size_t)nc);
  memset(s->div_w, 0, sizeof(double) * nc);
  for (int f = 0; f < s->nf; ++f) {
    const int o = s->m.face_owner[f], nb = s->m.face_neighbour[f];
    const double *sf = &s->m.face_sf[3 * f];
    int nb_per = -1, bc = -1;
    double Uf[3], Fn;
    if (nb >= 0) {
      const double gc = s->m.face_gc[f];
      for (int k = 0; k < 3; ++k) Uf[k] = gc * s->U[3 * o + k] + (1.0 - gc) * s->U[3 * nb + k];
      Fn = s->U_face[f];
    } else {
      const int b = f - s->nint;
      bc = bc_of_face(s, f);
      if (bc == BC_PERIODIC) {
        nb_per = periodic_partner_soa(s, b, o);
        if (nb_per >= 0) {
          const double w = s->m.periodic_gc[b];
          for (int k = 0; k < 3; ++k) Uf[k] = w * s->U[3 * o + k] + (1.0 - w) * s->U[3 * nb_per + k];
        } else {
          for (int k = 0; k < 3; ++k) Uf[k] = s->U_bnd[3 * b + k];
        }
      } else {
        for (int k = 0; k < 3; ++k) Uf[k] = s->U_bnd[3 * b + k];
      }
      if (bc == BC_PERIODIC || bc == BC_OUTLET) {
        Fn = s->U_face[f];
      } else if (bc == BC_WALL_NOSLIP || bc == BC_WALL_MODEL || bc == BC_WALL_VELOCITY || bc == BC_SLIP ||
                 bc == BC_SYMMETRY) {
        Fn = 0.0;
      } else {
        Fn = dot3(Uf, sf);
      }
    }
    s->div_w[o] = s->div_w[o] + Fn;
    if (nb >= 0) s->div_w[nb] = s->div_w[nb] - Fn;
    for (int k = 0; k < 3; ++k) {
      double phi;
      if (nb >= 0)
        phi = 0.5 * (s->U[3 * o + k] + s->U[3 * nb + k]);
      else if (nb_per >= 0)
        phi = 0.5 * (s->U[3 * o + k] + s->U[3 * nb_per + k]);
      else
        phi = Uf[k];
      rhs[3 * o + k] = rhs[3 * o + k] - Fn * phi * s->inv_vol[o];
      if (nb >= 0) rhs[3 * nb + k] = rhs[3 * nb + k] + Fn * phi * s->inv_vol[nb];
    }
  }
  for (int c = 0; c < nc; ++c)
    for (int k = 0; k < 3; ++k) rhs[3 * c + k] = rhs[3 * c + k] + 0.5 * s->U[3 * c + k] * s->div_w[c] * s->inv_vol[c];
}

/* advection_rhs_vector_cds_soa (src/numerics/mod_advection.f90:1200-1392) */
static void advection_cds(om_solver *s, double *rhs) {
  memset(rhs, 0, sizeof(double) * 3 * (size_t)s->nc);
  for (int f = 0; f < s->nf; ++f) {
    const int o = s->m.face_owner[f], nb = s->m.face_neighbour[f];
    const double *sf = &s->m.face_sf[3 * f];
    int b = 0, bc = -1, nb_per = -1;
    double Fn;
    if (nb >= 0) {
      Fn = s->U_face[f];
    } else {
      b = f - s->nint;
      bc = bc_of_face(s, f);
      if (bc == BC_PERIODIC || bc == BC_OUTLET)
        Fn = s->U_face[f];
      else if (bc == BC_WALL_NOSLIP || bc == BC_WALL_MODEL || bc == BC_WALL_VELOCITY || bc == BC_SLIP ||
               bc == BC_SYMMETRY)
        Fn = 0.0;
      else
        Fn = s->U_bnd[3 * b] * sf[0] + s->U_bnd[3 * b + 1] * sf[1] + s->U_bnd[3 * b + 2] * sf[2];
    }
    int has_rec = (nb >= 0);
    if (!has_rec && bc == BC_PERIODIC) {
      nb_per = periodic_partner_soa(s, b, o);
      has_rec = (nb_per >= 0);
    }
    const int nbs = has_rec ? (nb >= 0 ? nb : nb_per) : -1;
    for (int k = 0; k < 3; ++k) {
      double phi;
      if (has_rec) {
        const double gc = (f >= s->nint) ? s->m.periodic_gc[b] : s->m.face_gc[f];
        const double pl = gc * s->U[3 * o + k] + (1.0 - gc) * s->U[3 * nbs + k];
        if (f < s->nint) {
          double gm[3];
          for (int i = 0; i < 3; ++i) gm[i] = gc * s->grad_U[GI(o, i, k)] + (1.0 - gc) * s->grad_U[GI(nbs, i, k)];
          phi = pl + dot3(gm, &s->m.face_eps[3 * f]);
          const double w = refine_blend_weight(s->m.face_skew[f]);
          if (w < 1.0) phi = pl + w * (phi - pl);
        } else {
          phi = pl;
        }
      } else {
        phi = s->U_bnd[3 * b + k];
      }
      if (!isfinite(phi)) phi = 0.0;
      rhs[3 * o + k] = rhs[3 * o + k] - Fn * phi * s->inv_vol[o];
      if (nb >= 0) rhs[3 * nb + k] = rhs[3 * nb + k] + Fn * phi * s->inv_vol[nb];
    }
  }
}

/* ---- reconstruction schemes (QUICK, van Leer TVD, Koren TVD), vector SoA path
 * advection_rhs_vector_recon_soa (src/numerics/mod_advection.f90:1396-1635) and its device helpers
 * (:2334-2791). */

static double clamp_value(double x, double lo, double hi) { return fmin(fmax(x, lo), hi); }

/* soa_resolve_periodic_partner_dev (:2362-2414) */
static int periodic_partner(const om_solver *s, int f, double shift[3]) {
  shift[0] = shift[1] = shift[2] = 0.0;
  if (s->m.face_neighbour[f] >= 0 || bc_of_face(s, f) != BC_PERIODIC) return -1;
  const int b = f - s->nint, owner = s->m.face_owner[f];
  int pair_cell = resolve_periodic_nb(s, b), pair_face = -1; /* soa_resolve_periodic_nb_dev (:2334-2359) */
  const int pb = s->m.periodic_pair[b];
  if (pb >= 0) {
    pair_face = s->nint + pb;
    if (pair_cell < 0) pair_cell = s->m.face_owner[pair_face];
  }
  if (pair_cell < 0) return -1;
  double d_per;
  if (pair_face >= 0)
    d_per = s->m.face_dcf[f] + s->m.face_dcf[pair_face];
  else if (s->m.periodic_pair_dcf[b] > SMALL)
    d_per = s->m.face_dcf[f] + s->m.periodic_pair_dcf[b];
  else
    d_per = 2.0 * s->m.face_dcf[f];
  for (int i = 0; i < 3; ++i)
    shift[i] = (s->m.cell_centroid[3 * owner + i] + d_per * s->m.face_normal[3 * f + i]) -
               s->m.cell_centroid[3 * pair_cell + i];
  return pair_cell;
}

typedef struct {
  int up, down;
  double x_up[3], x_down[3], x_face[3], up_shift[3], dir_hat[3], s_face, s_down;
} recon_ctx;

static double norm3(const double *v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

/* soa_resolve_reconstruction_context_dev (:2417-2498) */
static int recon_context(const om_solver *s, int f, double Fn, recon_ctx *c) {
  memset(c, 0, sizeof(*c));
  for (int i = 0; i < 3; ++i) c->x_face[i] = s->m.face_centre[3 * f + i];
  const int owner = s->m.face_owner[f], nb = s->m.face_neighbour[f];
  if (nb >= 0) {
    if (Fn >= 0.0) {
      c->up = owner;
      c->down = nb;
    } else {
      c->up = nb;
      c->down = owner;
    }
    for (int i = 0; i < 3; ++i) {
      c->x_up[i] = s->m.cell_centroid[3 * c->up + i];
      c->x_down[i] = s->m.cell_centroid[3 * c->down + i];
    }
  } else {
    double sh[3];
    const int pc = periodic_partner(s, f, sh);
    if (pc < 0) return 0;
    if (Fn >= 0.0) {
      c->up = owner;
      c->down = pc;
      for (int i = 0; i < 3; ++i) {
        c->x_up[i] = s->m.cell_centroid[3 * owner + i];
        c->x_down[i] = s->m.cell_centroid[3 * pc + i] + sh[i];
      }
    } else {
      c->up = pc;
      c->down = owner;
      for (int i = 0; i < 3; ++i) {
        c->up_shift[i] = sh[i];
        c->x_up[i] = s->m.cell_centroid[3 * pc + i] + c->up_shift[i];
        c->x_down[i] = s->m.cell_centroid[3 * owner + i];
      }
    }
  }
  double dvec[3];
  for (int i = 0; i < 3; ++i) dvec[i] = c->x_face[i] - c->x_up[i];
  double dn = norm3(dvec);
  if (dn <= SMALL) {
    for (int i = 0; i < 3; ++i) dvec[i] = c->x_down[i] - c->x_up[i];
    dn = norm3(dvec);
  }
  if (dn <= SMALL) return 0;
  double tf[3], td[3];
  for (int i = 0; i < 3; ++i) {
    c->dir_hat[i] = dvec[i] / dn;
    tf[i] = c->x_face[i] - c->x_up[i];
    td[i] = c->x_down[i] - c->x_up[i];
  }
  c->s_face = dot3(tf, c->dir_hat);
  c->s_down = dot3(td, c->dir_hat);
  if (c->s_face <= SMALL || c->s_down <= c->s_face) {
    for (int i = 0; i < 3; ++i) dvec[i] = c->x_down[i] - c->x_up[i];
    dn = norm3(dvec);
    if (dn <= SMALL) return 0;
    for (int i = 0; i < 3; ++i) c->dir_hat[i] = dvec[i] / dn;
    c->s_face = dot3(tf, c->dir_hat);
    c->s_down = dot3(td, c->dir_hat);
  }
  if (c->s_down <= SMALL) return 0;
  c->s_face = fmin(fmax(c->s_face, 0.0), c->s_down);
  return 1;
}

/* the other cell across face fj of cell `cell` (owner / neighbour), -1 on a boundary face */
static int across(const om_solver *s, int fj, int cell) {
  int nb = s->m.face_owner[fj];
  if (nb == cell) nb = s->m.face_neighbour[fj];
  return nb;
}

/* soa_find_far_upwind_dev (:2501-2549) */
static int find_far_upwind(const om_solver *s, const recon_ctx *c) {
  int far_up = -1;
  double best_align = -1.0, best_perp2 = 1.0e308;
  for (int q = s->m.cell_face_offset[c->up]; q < s->m.cell_face_offset[c->up + 1]; ++q) {
    const int fj = s->m.cell_faces[q];
    int nb = across(s, fj, c->up);
    if (nb < 0 && fj >= s->nint && bc_of_face(s, fj) == BC_PERIODIC) nb = resolve_periodic_nb(s, fj - s->nint);
    if (nb == c->down || nb < 0) continue;
    double d[3];
    for (int i = 0; i < 3; ++i) d[i] = s->m.cell_centroid[3 * nb + i] + c->up_shift[i] - c->x_up[i];
    const double sd = dot3(d, c->dir_hat);
    if (sd >= -1.0e-12) continue;
    const double dist2 = fmax(d[0] * d[0] + d[1] * d[1] + d[2] * d[2], SMALL);
    const double align = -sd / sqrt(dist2);
    const double perp2 = fmax(0.0, dist2 - sd * sd);
    if (align > best_align + 1.0e-12 || (fabs(align - best_align) <= 1.0e-12 && perp2 < best_perp2 - 1.0e-12) ||
        (fabs(align - best_align) <= 1.0e-12 && fabs(perp2 - best_perp2) <= 1.0e-12 && nb < far_up)) {
      best_align = align;
      best_perp2 = perp2;
      far_up = nb;
    }
  }
  return far_up;
}

/* axial_delta (:1682-1686) */
static double axial_delta(double phi_far, double phi_up, const double *g, const double *dvec, double sv,
                          const double *dir) {
  double t[3];
  for (int i = 0; i < 3; ++i) t[i] = dvec[i] - sv * dir[i];
  return phi_far - phi_up - dot3(g, t);
}

/* soa_quick_reconstruction3_dev (:2632-2713) */
static void quick3(const om_solver *s, const recon_ctx *c, double out[3]) {
  const double *U = s->U;
  double phiU[3], phiD[3], b1[3], b2[3], ln[3], gup[3][3], dl[3];
  for (int k = 0; k < 3; ++k) {
    phiU[k] = U[3 * c->up + k];
    phiD[k] = U[3 * c->down + k];
    for (int i = 0; i < 3; ++i) gup[k][i] = s->grad_U[GI(c->up, i, k)];
  }
  double dvec[3];
  for (int i = 0; i < 3; ++i) dvec[i] = c->x_down[i] - c->x_up[i];
  double dist2 = fmax(dvec[0] * dvec[0] + dvec[1] * dvec[1] + dvec[2] * dvec[2], SMALL);
  double w = 1.0 / dist2;
  double a11 = w * c->s_down * c->s_down;
  double a12 = a11 * c->s_down;
  double a22 = a12 * c->s_down;
  double lden = a11;
  int nfit = 1;
  for (int k = 0; k < 3; ++k) {
    dl[k] = axial_delta(phiD[k], phiU[k], gup[k], dvec, c->s_down, c->dir_hat);
    b1[k] = w * c->s_down * dl[k];
    b2[k] = w * c->s_down * c->s_down * dl[k];
    ln[k] = b1[k];
  }
  for (int q = s->m.cell_face_offset[c->up]; q < s->m.cell_face_offset[c->up + 1]; ++q) {
    const int fj = s->m.cell_faces[q];
    const int nb = across(s, fj, c->up);
    if (nb < 0 || nb == c->down) continue;
    double xnb[3];
    for (int i = 0; i < 3; ++i) {
      xnb[i] = s->m.cell_centroid[3 * nb + i] + c->up_shift[i];
      dvec[i] = xnb[i] - c->x_up[i];
    }
    const double sv = dot3(dvec, c->dir_hat);
    if (sv >= -1.0e-12) continue;
    dist2 = fmax(dvec[0] * dvec[0] + dvec[1] * dvec[1] + dvec[2] * dvec[2], SMALL);
    if (sv * sv < 0.25 * dist2) continue;
    const double align = fmin(1.0, -sv / sqrt(dist2));
    w = (0.25 + align * align) / dist2;
    a11 = a11 + w * sv * sv;
    a12 = a12 + w * sv * sv * sv;
    a22 = a22 + w * sv * sv * sv * sv;
    lden = lden + w * sv * sv;
    for (int k = 0; k < 3; ++k) {
      dl[k] = axial_delta(U[3 * nb + k], phiU[k], gup[k], dvec, sv, c->dir_hat);
      b1[k] = b1[k] + w * sv * dl[k];
      b2[k] = b2[k] + w * sv * sv * dl[k];
      ln[k] = ln[k] + w * sv * dl[k];
    }
    nfit = nfit + 1;
  }
  const double det = a11 * a22 - a12 * a12;
  for (int k = 0; k < 3; ++k) {
    double a1, a2;
    if (nfit >= 2 && fabs(det) > SMALL) {
      a1 = (b1[k] * a22 - b2[k] * a12) / det;
      a2 = (-b1[k] * a12 + b2[k] * a11) / det;
    } else if (lden > SMALL) {
      a1 = ln[k] / lden;
      a2 = 0.0;
    } else {
      a1 = (phiD[k] - phiU[k]) / fmax(c->s_down, SMALL);
      a2 = 0.0;
    }
    out[k] = phiU[k] + a1 * c->s_face + a2 * c->s_face * c->s_face;
    out[k] = clamp_value(out[k], fmin(phiU[k], phiD[k]), fmax(phiU[k], phiD[k]));
  }
}

/* soa_vanleer_tvd_ctx_dev / soa_koren_tvd_ctx_dev (:2716-2786), one component */
static double tvd(const om_solver *s, const recon_ctx *c, int far_up, int k, int koren) {
  const double phiU = s->U[3 * c->up + k], phiD = s->U[3 * c->down + k];
  const double slope_dn = (phiD - phiU) / fmax(c->s_down, SMALL);
  double psi = 0.0;
  if (far_up >= 0 && fabs(slope_dn) > SMALL) {
    double t[3];
    for (int i = 0; i < 3; ++i) t[i] = s->m.cell_centroid[3 * far_up + i] + c->up_shift[i] - c->x_up[i];
    const double s_up = -dot3(t, c->dir_hat);
    if (s_up > SMALL) {
      const double slope_up = (phiU - s->U[3 * far_up + k]) / s_up;
      const double r = slope_up / slope_dn;
      if (koren) {
        psi = fmax(0.0, fmin(fmin(2.0 * r, (1.0 + 2.0 * r) / 3.0), 2.0));
      } else {
        psi = (r + fabs(r)) / (1.0 + fabs(r));
        psi = fmax(0.0, fmin(psi, 2.0));
      }
    }
  }
  return clamp_value(phiU + psi * slope_dn * c->s_face, fmin(phiU, phiD), fmax(phiU, phiD));
}

static void advection_recon(om_solver *s, double *rhs) {
  const int scheme = s->p.scheme;
  memset(rhs, 0, sizeof(double) * 3 * (size_t)s->nc);
  for (int f = 0; f < s->nf; ++f) {
    const int o = s->m.face_owner[f], nb = s->m.face_neighbour[f];
    int b = 0;
    double Fn;
    if (nb >= 0) {
      Fn = s->U_face[f];
    } else {
      b = f - s->nint;
      const int bc = bc_of_face(s, f);
      if (bc == BC_WALL_NOSLIP || bc == BC_WALL_MODEL || bc == BC_WALL_VELOCITY || bc == BC_SLIP ||
          bc == BC_SYMMETRY)
        Fn = 0.0;
      else
        Fn = s->U_face[f];
    }
    recon_ctx c;
    const int ok = recon_context(s, f, Fn, &c);
    int far_up = -1;
    if (ok && scheme != OM_SCHEME_QUICK) far_up = find_far_upwind(s, &c);
    if (!isfinite(Fn)) Fn = 0.0;
    double q[3] = {0.0, 0.0, 0.0};
    if (ok && scheme == OM_SCHEME_QUICK) quick3(s, &c, q);
    for (int k = 0; k < 3; ++k) {
      double phi;
      if (ok) {
        if (scheme == OM_SCHEME_QUICK)
          phi = q[k];
        else
          phi = tvd(s, &c, far_up, k, scheme == OM_SCHEME_KOREN);
        if (!isfinite(phi)) phi = (Fn >= 0.0 || nb < 0) ? s->U[3 * o + k] : s->U[3 * nb + k];
      } else {
        phi = s->U_bnd[3 * b + k];
      }
      if (!isfinite(phi)) phi = 0.0;
      rhs[3 * o + k] = rhs[3 * o + k] - Fn * phi * s->inv_vol[o];
      if (nb >= 0) rhs[3 * nb + k] = rhs[3 * nb + k] + Fn * phi * s->inv_vol[nb];
    }
  }
}

static double ortho_weight(double s_dot_dhat, double area) {
  return area * fmax(0.0, fmin(1.0, s_dot_dhat / (area + SMALL)));
}

/* viscous_rhs_vector (src/numerics/mod_viscous.f90:37-283) */
static void viscous_vector(om_solver *s, double *rhs) {
  memset(rhs, 0, sizeof(double) * 3 * (size_t)s->nc);
  const double nu = s->p.nu;
  for (int f = 0; f < s->nf; ++f) {
    const int o = s->m.face_owner[f], nb = s->m.face_neighbour[f];
    const double *sf = &s->m.face_sf[3 * f];
    const double *n_hat = &s->m.face_normal[3 * f];
    double st[3] = {0, 0, 0};
    int have = 1;
    if (nb >= 0) {
      const double gc = s->m.face_gc[f];
      const double nue = nu + gc * s->nu_sgs[o] + (1.0 - gc) * s->nu_sgs[nb];
      double gf[9], d_on[3], dn_perp[3];
      for (int q = 0; q < 9; ++q) gf[q] = gc * s->grad_U[9 * (size_t)o + q] + (1.0 - gc) * s->grad_U[9 * (size_t)nb + q];
      for (int i = 0; i < 3; ++i) d_on[i] = s->m.cell_centroid[3 * nb + i] - s->m.cell_centroid[3 * o + i];
      const double dlen = sqrt(d_on[0] * d_on[0] + d_on[1] * d_on[1] + d_on[2] * d_on[2]) + SMALL;
      const double sdd = dot3(sf, d_on) / dlen;
      for (int i = 0; i < 3; ++i) dn_perp[i] = sf[i] - (sdd / dlen) * d_on[i];
      const double wo = ortho_weight(sdd, s->m.face_area[f]);
      for (int k = 0; k < 3; ++k) {
        const double col = gf[3 * k] * dn_perp[0] + gf[3 * k + 1] * dn_perp[1] + gf[3 * k + 2] * dn_perp[2];
        const double row = gf[k] * sf[0] + gf[3 + k] * sf[1] + gf[6 + k] * sf[2];
        st[k] = nue * ((s->U[3 * nb + k] - s->U[3 * o + k]) * wo / dlen + col + row);
      }
    } else {
      const int b = f - s->nint;
      const int bc = bc_of_face(s, f);
      if (bc == BC_PERIODIC) {
        int nb_per = -1, pair_face = -1;
        const int pair = s->m.periodic_pair[b];
        if (pair >= 0) {
          pair_face = s->nint + pair;
          nb_per = s->m.face_owner[pair_face];
        }
        if (nb_per == o) nb_per = -1;
        if (nb_per >= 0) {
          double d_per, gcp;
          if (pair_face >= 0) {
            d_per = s->m.face_dcf[f] + s->m.face_dcf[pair_face];
            gcp = s->m.face_dcf[pair_face] / (d_per + SMALL);
          } else if (s->m.periodic_pair_dcf[b] > SMALL) {
            d_per = s->m.face_dcf[f] + s->m.periodic_pair_dcf[b];
            gcp = s->m.periodic_pair_dcf[b] / (d_per + SMALL);
          } else {
            d_per = 2.0 * s->m.face_dcf[f];
            gcp = 0.5;
          }
          const double nue = nu + gcp * s->nu_sgs[o] + (1.0 - gcp) * s->nu_sgs[nb_per];
          double d_on[3], dn_perp[3], gf[9];
          for (int i = 0; i < 3; ++i) d_on[i] = d_per * n_hat[i];
          const double dlen = d_per + SMALL;
          const double sdd = dot3(sf, d_on) / dlen;
          for (int i = 0; i < 3; ++i) dn_perp[i] = sf[i] - (sdd / dlen) * d_on[i];
          for (int q = 0; q < 9; ++q)
            gf[q] = gcp * s->grad_U[9 * (size_t)o + q] + (1.0 - gcp) * s->grad_U[9 * (size_t)nb_per + q];
          const double wo = ortho_weight(sdd, s->m.face_area[f]);
          for (int k = 0; k < 3; ++k) {
            const double col = gf[3 * k] * dn_perp[0] + gf[3 * k + 1] * dn_perp[1] + gf[3 * k + 2] * dn_perp[2];
            const double row = gf[k] * sf[0] + gf[3 + k] * sf[1] + gf[6 + k] * sf[2];
            st[k] = nue * ((s->U[3 * nb_per + k] - s->U[3 * o + k]) * wo / dlen + col + row);
          }
        } else {
          const double nue = nu + s->nu_sgs[o];
          for (int k = 0; k < 3; ++k)
            st[k] = nue * ((s->U_bnd[3 * b + k] - s->U[3 * o + k]) / (s->m.face_dcf[f] + SMALL)) * s->m.face_area[f];
        }
      } else if (bc == BC_WALL_MODEL) {
        have = 0;
      } else {
        const double nue = nu + s->nu_sgs[o];
        double d_on[3], dn_perp[3];
        for (int i = 0; i < 3; ++i) d_on[i] = s->m.face_centre[3 * f + i] - s->m.cell_centroid[3 * o + i];
        const double dlen = sqrt(d_on[0] * d_on[0] + d_on[1] * d_on[1] + d_on[2] * d_on[2]) + SMALL;
        const double sdd = dot3(sf, d_on) / dlen;
        for (int i = 0; i < 3; ++i) dn_perp[i] = sf[i] - (sdd / dlen) * d_on[i];
        const double *go = &s->grad_U[9 * (size_t)o];
        const double wo = ortho_weight(sdd, s->m.face_area[f]);
        for (int k = 0; k < 3; ++k) {
          const double col = go[3 * k] * dn_perp[0] + go[3 * k + 1] * dn_perp[1] + go[3 * k + 2] * dn_perp[2];
          const double row = go[k] * sf[0] + go[3 + k] * sf[1] + go[6 + k] * sf[2];
          st[k] = nue * ((s->U_bnd[3 * b + k] - s->U[3 * o + k]) * wo / dlen + col + row);
        }
      }
    }
    if (!have) continue;
    for (int k = 0; k < 3; ++k) {
      rhs[3 * o + k] = rhs[3 * o + k] + st[k] * s->inv_vol[o];
      if (nb >= 0) rhs[3 * nb + k] = rhs[3 * nb + k] - st[k] * s->inv_vol[nb];
    }
  }
}

/* ---- temperature transport */

/* face flux of the scalar paths (advection_rhs_scalar_{cds,recon}_soa, mod_advection.f90:363-383) */
static double scalar_face_flux(const om_solver *s, int f) {
  if (s->m.face_neighbour[f] >= 0) return s->U_face[f];
  const int bc = bc_of_face(s, f);
  if (bc == BC_WALL_NOSLIP || bc == BC_WALL_MODEL || bc == BC_WALL_VELOCITY || bc == BC_SLIP || bc == BC_SYMMETRY)
    return 0.0;
  return s->U_face[f];
}

/* advection_rhs_scalar_cds_soa (src/numerics/mod_advection.f90:267-485) */
static void advection_scalar_cds(om_solver *s, const double *phi, const double *pb, const double *gphi, double *rhs) {
  memset(rhs, 0, sizeof(double) * s->nc);
  for (int f = 0; f < s->nf; ++f) {
    const int o = s->m.face_owner[f], nb = s->m.face_neighbour[f];
    const int b = f - s->nint;
    double Fn = scalar_face_flux(s, f);
    int nbs = nb;
    if (nb < 0) nbs = (bc_of_face(s, f) == BC_PERIODIC) ? s->per_nb[b] : -1;
    double phif;
    if (nbs >= 0) {
      const double gc = (f >= s->nint) ? s->m.periodic_gc[b] : s->m.face_gc[f];
      const double pl = gc * phi[o] + (1.0 - gc) * phi[nbs];
      if (f < s->nint) {
        double gm[3];
        for (int i = 0; i < 3; ++i) gm[i] = gc * gphi[3 * o + i] + (1.0 - gc) * gphi[3 * nbs + i];
        phif = pl + dot3(gm, &s->m.face_eps[3 * f]);
        const double w = refine_blend_weight(s->m.face_skew[f]);
        if (w < 1.0) phif = pl + w * (phif - pl);
      } else {
        phif = pl;
      }
      if (!isfinite(phif)) phif = (Fn >= 0.0 || nb < 0) ? phi[o] : phi[nb];
    } else {
      phif = pb[b];
    }
    if (!isfinite(Fn)) Fn = 0.0;
    if (!isfinite(phif)) phif = 0.0;
    rhs[o] = rhs[o] - Fn * phif * s->inv_vol[o];
    if (nb >= 0) rhs[nb] = rhs[nb] + Fn * phif * s->inv_vol[nb];
  }
}

/* soa_quick_reconstruction_ctx_dev (src/numerics/mod_advection.f90:2552-2629), one scalar */
static double quick_scalar(const om_solver *s, const recon_ctx *c, const double *phi, const double *gphi) {
  const double pU = phi[c->up], pD = phi[c->down];
  const double *gup = &gphi[3 * (size_t)c->up];
  double a11 = 0.0, a12 = 0.0, a22 = 0.0, b1 = 0.0, b2 = 0.0, ln = 0.0, ld = 0.0;
  int nfit = 0;
  double dvec[3];
  for (int i = 0; i < 3; ++i) dvec[i] = c->x_down[i] - c->x_up[i];
  double dist2 = fmax(dvec[0] * dvec[0] + dvec[1] * dvec[1] + dvec[2] * dvec[2], SMALL);
  double w = 1.0 / dist2;
  double dl = axial_delta(pD, pU, gup, dvec, c->s_down, c->dir_hat);
  const double sd = c->s_down;
  a11 = a11 + w * sd * sd;
  a12 = a12 + w * sd * sd * sd;
  a22 = a22 + w * sd * sd * sd * sd;
  b1 = b1 + w * sd * dl;
  b2 = b2 + w * sd * sd * dl;
  ln = ln + w * sd * dl;
  ld = ld + w * sd * sd;
  nfit = nfit + 1;
  for (int q = s->m.cell_face_offset[c->up]; q < s->m.cell_face_offset[c->up + 1]; ++q) {
    const int fj = s->m.cell_faces[q];
    const int nb = across(s, fj, c->up);
    if (nb < 0 || nb == c->down) continue;
    for (int i = 0; i < 3; ++i) dvec[i] = (s->m.cell_centroid[3 * nb + i] + c->up_shift[i]) - c->x_up[i];
    const double sv = dot3(dvec, c->dir_hat);
    if (sv >= -1.0e-12) continue;
    dist2 = fmax(dvec[0] * dvec[0] + dvec[1] * dvec[1] + dvec[2] * dvec[2], SMALL);
    if (sv * sv < 0.25 * dist2) continue;
    const double align = fmin(1.0, -sv / sqrt(dist2));
    w = (0.25 + align * align) / dist2;
    dl = axial_delta(phi[nb], pU, gup, dvec, sv, c->dir_hat);
    a11 = a11 + w * sv * sv;
    a12 = a12 + w * sv * sv * sv;
    a22 = a22 + w * sv * sv * sv * sv;
    b1 = b1 + w * sv * dl;
    b2 = b2 + w * sv * sv * dl;
    ln = ln + w * sv * dl;
    ld = ld + w * sv * sv;
    nfit = nfit + 1;
  }
  const double det = a11 * a22 - a12 * a12;
  double a1, a2;
  if (nfit >= 2 && fabs(det) > SMALL) {
    a1 = (b1 * a22 - b2 * a12) / det;
    a2 = (-b1 * a12 + b2 * a11) / det;
  } else if (ld > SMALL) {
    a1 = ln / ld;
    a2 = 0.0;
  } else {
    a1 = (pD - pU) / fmax(c->s_down, SMALL);
    a2 = 0.0;
  }
  return clamp_value(pU + a1 * c->s_face + a2 * c->s_face * c->s_face, fmin(pU, pD), fmax(pU, pD));
}

/* TVD face value of one scalar field (soa_*_tvd_ctx_dev) */
static double tvd_scalar(const recon_ctx *c, const double *phi, const double *ctr, int far_up, int koren) {
  const double pU = phi[c->up], pD = phi[c->down];
  const double slope_dn = (pD - pU) / fmax(c->s_down, SMALL);
  double psi = 0.0;
  if (far_up >= 0 && fabs(slope_dn) > SMALL) {
    double t[3];
    for (int i = 0; i < 3; ++i) t[i] = ctr[3 * far_up + i] + c->up_shift[i] - c->x_up[i];
    const double s_up = -dot3(t, c->dir_hat);
    if (s_up > SMALL) {
      const double r = ((pU - phi[far_up]) / s_up) / slope_dn;
      if (koren) {
        psi = fmax(0.0, fmin(fmin(2.0 * r, (1.0 + 2.0 * r) / 3.0), 2.0));
      } else {
        psi = (r + fabs(r)) / (1.0 + fabs(r));
        psi = fmax(0.0, fmin(psi, 2.0));
      }
    }
  }
  return clamp_value(pU + psi * slope_dn * c->s_face, fmin(pU, pD), fmax(pU, pD));
}

/* advection_rhs_scalar_recon_soa (src/numerics/mod_advection.f90:488-675) */
static void advection_scalar_recon(om_solver *s, int scheme, const double *phi, const double *pb, const double *gphi,
                                   double *rhs) {
  memset(rhs, 0, sizeof(double) * s->nc);
  for (int f = 0; f < s->nf; ++f) {
    const int o = s->m.face_owner[f], nb = s->m.face_neighbour[f];
    const int b = f - s->nint;
    double Fn = scalar_face_flux(s, f);
    recon_ctx c;
    const int ok = recon_context(s, f, Fn, &c);
    double phif;
    if (ok) {
      if (scheme == OM_SCHEME_QUICK)
        phif = quick_scalar(s, &c, phi, gphi);
      else
        phif = tvd_scalar(&c, phi, s->m.cell_centroid, find_far_upwind(s, &c), scheme == OM_SCHEME_KOREN);
      if (!isfinite(phif)) phif = (Fn >= 0.0 || nb < 0) ? phi[o] : phi[nb];
    } else {
      phif = pb[b];
    }
    if (!isfinite(Fn)) Fn = 0.0;
    if (!isfinite(phif)) phif = 0.0;
    rhs[o] = rhs[o] - Fn * phif * s->inv_vol[o];
    if (nb >= 0) rhs[nb] = rhs[nb] + Fn * phif * s->inv_vol[nb];
  }
}

/* viscous_rhs_scalar (src/numerics/mod_viscous.f90:286-513) */
static void viscous_scalar(om_solver *s, const double *phi, const double *pb, const double *gphi, double alpha,
                           double prt, double *rhs) {
  memset(rhs, 0, sizeof(double) * s->nc);
  for (int f = 0; f < s->nf; ++f) {
    const int o = s->m.face_owner[f], nb = s->m.face_neighbour[f];
    const double *sf = &s->m.face_sf[3 * f];
    double flux, d_on[3], dnp[3], gf[3];
    if (nb >= 0) {
      const double gc = s->m.face_gc[f];
      const double ae = alpha + (gc * s->nu_sgs[o] + (1.0 - gc) * s->nu_sgs[nb]) / prt;
      for (int i = 0; i < 3; ++i) d_on[i] = s->m.cell_centroid[3 * nb + i] - s->m.cell_centroid[3 * o + i];
      const double dl = sqrt(d_on[0] * d_on[0] + d_on[1] * d_on[1] + d_on[2] * d_on[2]) + SMALL;
      const double sdd = dot3(sf, d_on) / dl;
      for (int i = 0; i < 3; ++i) {
        dnp[i] = sf[i] - (sdd / dl) * d_on[i];
        gf[i] = gc * gphi[3 * o + i] + (1.0 - gc) * gphi[3 * nb + i];
      }
      flux = ae * ((phi[nb] - phi[o]) * ortho_weight(sdd, s->m.face_area[f]) / dl + dot3(gf, dnp));
    } else {
      const int b = f - s->nint;
      const int nbp = (bc_of_face(s, f) == BC_PERIODIC) ? s->per_nb[b] : -1;
      if (nbp >= 0) {
        const int pb2 = s->m.periodic_pair[b];
        double d_per, gcp;
        if (pb2 >= 0) {
          const int pf = s->nint + pb2;
          d_per = s->m.face_dcf[f] + s->m.face_dcf[pf];
          gcp = s->m.face_dcf[pf] / (d_per + SMALL);
        } else if (s->m.periodic_pair_dcf[b] > SMALL) {
          d_per = s->m.face_dcf[f] + s->m.periodic_pair_dcf[b];
          gcp = s->m.periodic_pair_dcf[b] / (d_per + SMALL);
        } else {
          d_per = 2.0 * s->m.face_dcf[f];
          gcp = 0.5;
        }
        const double ae = alpha + (gcp * s->nu_sgs[o] + (1.0 - gcp) * s->nu_sgs[nbp]) / prt;
        for (int i = 0; i < 3; ++i) d_on[i] = d_per * s->m.face_normal[3 * f + i];
        const double dl = d_per + SMALL;
        const double sdd = dot3(sf, d_on) / dl;
        for (int i = 0; i < 3; ++i) {
          dnp[i] = sf[i] - (sdd / dl) * d_on[i];
          gf[i] = gcp * gphi[3 * o + i] + (1.0 - gcp) * gphi[3 * nbp + i];
        }
        flux = ae * ((phi[nbp] - phi[o]) * ortho_weight(sdd, s->m.face_area[f]) / dl + dot3(gf, dnp));
      } else {
        const double ae = alpha + s->nu_sgs[o] / prt;
        for (int i = 0; i < 3; ++i) d_on[i] = s->m.face_centre[3 * f + i] - s->m.cell_centroid[3 * o + i];
        const double dl = sqrt(d_on[0] * d_on[0] + d_on[1] * d_on[1] + d_on[2] * d_on[2]) + SMALL;
        const double sdd = dot3(sf, d_on) / dl;
        for (int i = 0; i < 3; ++i) dnp[i] = sf[i] - (sdd / dl) * d_on[i];
        flux = ae * ((pb[b] - phi[o]) * ortho_weight(sdd, s->m.face_area[f]) / dl + dot3(&gphi[3 * o], dnp));
      }
    }
    rhs[o] = rhs[o] + flux * s->inv_vol[o];
    if (nb >= 0) rhs[nb] = rhs[nb] - flux * s->inv_vol[nb];
  }
}

/* compute_explicit_rhs (src/solver/mod_fv_solver.f90:1617-1762) */
static void explicit_rhs(om_solver *s) {
  if (s->p.scheme == OM_SCHEME_KEP)
    advection_kep(s, s->conv);
  else if (s->p.scheme == OM_SCHEME_CDS)
    advection_cds(s, s->conv);
  else
    advection_recon(s, s->conv);
  viscous_vector(s, s->visc);
  const int bous = s->p.boussinesq;
  if (bous) {
    if (s->p.t_scheme == OM_SCHEME_CDS)
      advection_scalar_cds(s, s->T, s->T_bnd, s->grad_T, s->conv_T);
    else
      advection_scalar_recon(s, s->p.t_scheme, s->T, s->T_bnd, s->grad_T, s->conv_T);
    viscous_scalar(s, s->T, s->T_bnd, s->grad_T, s->p.alpha_T, s->p.Pr_t, s->visc_T);
  }
  double fb[3];
  body_force_now(s, fb);
  const double *om = s->p.omega_rot;
  for (int c = 0; c < s->nc; ++c) {
    for (int k = 0; k < 3; ++k) s->H[3 * c + k] = s->conv[3 * c + k] + s->visc[3 * c + k] + fb[k];
    if (s->p.use_rotation) {
      const double *u = &s->U[3 * c];
      s->H[3 * c] = s->H[3 * c] - 2.0 * (om[1] * u[2] - om[2] * u[1]);
      s->H[3 * c + 1] = s->H[3 * c + 1] - 2.0 * (om[2] * u[0] - om[0] * u[2]);
      s->H[3 * c + 2] = s->H[3 * c + 2] - 2.0 * (om[0] * u[1] - om[1] * u[0]);
    }
    if (bous)
      for (int k = 0; k < 3; ++k) s->H[3 * c + k] = s->H[3 * c + k] - s->p.beta * (s->T[c] - s->p.T_ref) * s->p.g[k];
    if (bous) s->H_T[c] = s->conv_T[c] + s->visc_T[c];
  }
  /* add_wall_stress_to_rhs (:2238-2275) */
  if (s->n_wm_patches > 0)
    for (int b = 0; b < s->nb; ++b) {
      const int f = s->nint + b;
      if (bc_of_face(s, f) != BC_WALL_MODEL) continue;
      const int o = s->m.face_owner[f];
      const double w = s->m.face_area[f] / s->m.cell_volume[o];
      for (int k = 0; k < 3; ++k) s->H[3 * o + k] = s->H[3 * o + k] + s->tau_w[3 * b + k] * w;
    }
  /* add_wall_heat_flux_to_rhs (:2279-2313): kinematic heat flux times area / volume */
  if (bous && s->p.wall_model == OM_WM_MOST && s->n_wm_patches > 0)
    for (int b = 0; b < s->nb; ++b) {
      const int f = s->nint + b;
      if (bc_of_face(s, f) != BC_WALL_MODEL) continue;
      const int o = s->m.face_owner[f];
      s->H_T[o] = s->H_T[o] + s->q_w[b] * s->m.face_area[f] / s->m.cell_volume[o];
    }
  /* passive scalars, one at a time (:1719-1743): advection by scalar_scheme, diffusion with the
   * scalar's diffusivity and turbulent Schmidt number */
  for (int k = 0; k < s->p.n_scalars; ++k) {
    if (s->p.scalar_scheme == OM_SCHEME_CDS)
      advection_scalar_cds(s, s->C[k], s->C_bnd[k], s->grad_C[k], s->conv_T);
    else
      advection_scalar_recon(s, s->p.scalar_scheme, s->C[k], s->C_bnd[k], s->grad_C[k], s->conv_T);
    viscous_scalar(s, s->C[k], s->C_bnd[k], s->grad_C[k], s->p.scalar_diffusivity[k], s->p.scalar_pr_t[k], s->visc_T);
    for (int c = 0; c < s->nc; ++c) s->H_C[k][c] = s->conv_T[c] + s->visc_T[c];
  }
}

/* ---------------------------------------------------------------- projection */

/* compute_face_fluxes_and_rhs (src/solver/mod_fv_solver.f90:2389-2752) */
static void face_fluxes_and_rhs(om_solver *s, double dt, int incr) {
  const int nc = s->nc;
  if (!s->skip_grad_p) {
    if (incr) {
      for (int c = 0; c < nc; ++c) s->p_tot[c] = s->phi_prev[c] + s->P[c];
      for (int b = 0; b < s->nb; ++b) s->p_tot_bnd[b] = s->phi_prev_bnd[b] + s->p_bnd[b];
      gradient_of_scalar(s, s->p_tot, s->p_tot_bnd, s->gp);
    } else {
      gradient_of_scalar(s, s->P, s->p_bnd, s->gp);
    }
  }
  double *rhs = s->rhs_p, *ff = s->face_flux;
  memset(rhs, 0, sizeof(double) * nc);
  memset(ff, 0, sizeof(double) * s->nf);
  memset(s->built, 0, s->nf);
  for (int idx = 0; idx < s->npf; ++idx) {
    if (s->ps_skip[idx]) continue;
    const int o = s->ps_owner[idx], nb = s->ps_nb[idx], fm = s->ps_fmesh[idx];
    double gcf;
    if (fm < s->nint) {
      gcf = s->m.face_gc[fm];
    } else {
      gcf = 0.5;
      const int b = fm - s->nint;
      const double d_o = s->m.face_dcf[fm];
      double d_nb = d_o;
      if (s->m.periodic_pair_dcf[b] > SMALL) d_nb = s->m.periodic_pair_dcf[b];
      if (d_nb <= SMALL && s->m.periodic_pair[b] >= 0) d_nb = s->m.face_dcf[s->nint + s->m.periodic_pair[b]];
      const double d_tot = d_o + d_nb;
      if (d_tot > SMALL) gcf = d_nb / d_tot;
    }
    double d_on[3], Uf[3];
    for (int i = 0; i < 3; ++i) d_on[i] = s->m.cell_centroid[3 * nb + i] - s->m.cell_centroid[3 * o + i];
    const double *sf = &s->m.face_sf[3 * fm];
    double Upl[3];
    for (int k = 0; k < 3; ++k) Upl[k] = gcf * s->U[3 * o + k] + (1.0 - gcf) * s->U[3 * nb + k];
    if (fm < s->nint) {
      const double wb = refine_blend_weight(s->m.face_skew[fm]);
      const int om = s->m.face_owner[fm], nm = s->m.face_neighbour[fm];
      const double g = s->m.face_gc[fm];
      const double *eps = &s->m.face_eps[3 * fm];
      double Ue[3];
      for (int j = 0; j < 3; ++j) {
        Ue[j] = g * s->U[3 * om + j] + (1.0 - g) * s->U[3 * nm + j];
        double gm[3];
        for (int i = 0; i < 3; ++i) gm[i] = g * s->grad_U[GI(om, i, j)] + (1.0 - g) * s->grad_U[GI(nm, i, j)];
        Ue[j] = Ue[j] + gm[0] * eps[0] + gm[1] * eps[1] + gm[2] * eps[2];
      }
      for (int k = 0; k < 3; ++k) Uf[k] = Upl[k] + wb * (Ue[k] - Upl[k]);
    } else {
      for (int k = 0; k < 3; ++k) Uf[k] = Upl[k];
    }
    double flux = Uf[0] * sf[0] + Uf[1] * sf[1] + Uf[2] * sf[2];
    const double coeff = s->ps_coeff[idx];
    double krc[3] = {0.0, 0.0, 0.0};
    if (!s->skip_nonortho && !s->p.rhie_chow_dual) {
      double kf[3];
      kfactor_limited(sf, d_on, s->p.min_cos_theta, kf);
      if (kf[0] * kf[0] + kf[1] * kf[1] + kf[2] * kf[2] >= NONORTHO_SIN2_TOL * (sf[0] * sf[0] + sf[1] * sf[1] + sf[2] * sf[2]))
        for (int i = 0; i < 3; ++i) krc[i] = kf[i];
    }
    /* nonortho_rhie_chow (src/numerics/mod_nonortho.f90:101-117) */
    double gm[3], smk[3];
    for (int i = 0; i < 3; ++i) {
      gm[i] = gcf * s->gp_rc[3 * o + i] + (1.0 - gcf) * s->gp_rc[3 * nb + i];
      smk[i] = sf[i] - krc[i];
    }
    const double rc = dot3(gm, smk) - coeff * (s->phi_prev[nb] - s->phi_prev[o]);
    flux = flux + dt * rc;
    ff[fm] = flux;
    s->built[fm] = 1;
    rhs[o] = rhs[o] - flux;
    rhs[nb] = rhs[nb] + flux;
  }
  for (int ib = 0; ib < s->n_nper_bf; ++ib) {
    const int b = s->nper_bf[ib], f = s->nint + b, o = s->m.face_owner[f];
    const int bc = bc_of_face(s, f);
    double flux;
    if (bc == BC_WALL_NOSLIP || bc == BC_WALL_MODEL || bc == BC_WALL_VELOCITY || bc == BC_SLIP || bc == BC_SYMMETRY)
      flux = 0.0;
    else
      flux = s->U_bnd[3 * b] * s->m.face_sf[3 * f] + s->U_bnd[3 * b + 1] * s->m.face_sf[3 * f + 1] +
             s->U_bnd[3 * b + 2] * s->m.face_sf[3 * f + 2];
    if (bc == BC_OUTLET && s->out_conv[s->m.face_patch[f]]) {
      /* outlet flux is set by continuity: the extrapolated cell velocity, not the convective
       * boundary state (mod_fv_solver.f90:2619-2627) */
      const double d[3] = {s->m.face_centre[3 * f] - s->m.cell_centroid[3 * o],
                           s->m.face_centre[3 * f + 1] - s->m.cell_centroid[3 * o + 1],
                           s->m.face_centre[3 * f + 2] - s->m.cell_centroid[3 * o + 2]};
      flux = 0.0;
      for (int k = 0; k < 3; ++k)
        flux = flux + (s->U[3 * o + k] + s->grad_U[GI(o, 0, k)] * d[0] + s->grad_U[GI(o, 1, k)] * d[1] +
                       s->grad_U[GI(o, 2, k)] * d[2]) * s->m.face_sf[3 * f + k];
    }
    if (bc == BC_OUTLET && s->p.cell_projection == OM_PROJ_EXACT) {
      /* EXACT: the compact old-pressure term of the outlet face (mod_fv_solver.f90:2629-2636) */
      const double *n = &s->m.face_normal[3 * f];
      const double d1 = s->m.face_centre[3 * f] - s->m.cell_centroid[3 * o];
      const double d2 = s->m.face_centre[3 * f + 1] - s->m.cell_centroid[3 * o + 1];
      const double d3 = s->m.face_centre[3 * f + 2] - s->m.cell_centroid[3 * o + 2];
      const double c_b = s->m.face_area[f] / fmax(fabs(d1 * n[0] + d2 * n[1] + d3 * n[2]), SMALL);
      flux = flux - dt * c_b * (s->outlet_pb[b] - s->phi_prev[o]);
    }
    ff[f] = flux;
    s->built[f] = 1;
    rhs[o] = rhs[o] - flux;
  }
  for (int ib = 0; ib < s->n_per_bf; ++ib) {
    const int b = s->per_bf[ib], f = s->nint + b, o = s->m.face_owner[f];
    if (s->built[f]) continue;
    const int pair = s->m.periodic_pair[b];
    if (pair >= 0 && s->built[s->nint + pair]) {
      ff[f] = -ff[s->nint + pair];
      s->built[f] = 1;
      continue;
    }
    const int nb = (pair >= 0) ? s->m.face_owner[s->nint + pair] : -1;
    const double *sf = &s->m.face_sf[3 * f];
    if (nb >= 0)
      ff[f] = 0.5 * ((s->U[3 * o] + s->U[3 * nb]) * sf[0] + (s->U[3 * o + 1] + s->U[3 * nb + 1]) * sf[1] +
                     (s->U[3 * o + 2] + s->U[3 * nb + 2]) * sf[2]);
    else
      ff[f] = s->U[3 * o] * sf[0] + s->U[3 * o + 1] * sf[1] + s->U[3 * o + 2] * sf[2];
    s->built[f] = 1;
  }
  const double rdt = 1.0 / fmax(dt, SMALL);
  for (int c = 0; c < nc; ++c) rhs[c] = rhs[c] * rdt;
  if (!s->skip_nonortho) {
    for (int idx = 0; idx < s->npf; ++idx) {
      if (s->ps_skip[idx] || !s->act_int[idx]) continue;
      const int o = s->ps_owner[idx], nb = s->ps_nb[idx], fm = s->ps_fmesh[idx];
      const double gcw = (fm >= s->nint) ? s->m.periodic_gc[fm - s->nint] : s->m.face_gc[fm];
      const double cr = cross_flux(&s->gp[3 * o], &s->gp[3 * nb], gcw, &s->k_int[3 * idx]);
      rhs[o] = rhs[o] + cr;
      rhs[nb] = rhs[nb] - cr;
    }
    for (int ib = 0; ib < s->n_nper_bf; ++ib) {
      const int b = s->nper_bf[ib], f = s->nint + b, o = s->m.face_owner[f];
      if (bc_of_face(s, f) != BC_OUTLET || !s->act_bnd[ib]) continue;
      rhs[o] = rhs[o] + dot3(&s->gp[3 * o], &s->k_bnd[3 * ib]);
    }
  }
}

/* correct_face_fluxes (src/solver/mod_fv_solver.f90:2756-2900) */
static void correct_face_fluxes(om_solver *s, double dt, int ptot) {
  double *ff = s->face_flux;
  for (int idx = 0; idx < s->npf; ++idx) {
    if (s->ps_skip[idx]) continue;
    const int o = s->ps_owner[idx], nb = s->ps_nb[idx], fm = s->ps_fmesh[idx];
    ff[fm] = ff[fm] - dt * s->ps_coeff[idx] * (s->P[nb] - s->P[o]);
    if (s->act_int[idx]) {
      const double gc = (fm >= s->nint) ? s->m.periodic_gc[fm - s->nint] : s->m.face_gc[fm];
      ff[fm] = ff[fm] - dt * cross_flux(&s->gp[3 * o], &s->gp[3 * nb], gc, &s->k_int[3 * idx]);
    }
  }
  for (int ib = 0; ib < s->n_nper_bf; ++ib) {
    const int b = s->nper_bf[ib], f = s->nint + b, o = s->m.face_owner[f];
    if (bc_of_face(s, f) != BC_OUTLET) continue;
    const double *n = &s->m.face_normal[3 * f];
    double d[3];
    for (int i = 0; i < 3; ++i) d[i] = s->m.face_centre[3 * f + i] - s->m.cell_centroid[3 * o + i];
    const double sn = fabs(d[0] * n[0] + d[1] * n[1] + d[2] * n[2]);
    const double coeff = s->m.face_area[f] / fmax(sn, SMALL);
    ff[f] = ff[f] - dt * coeff * ((ptot ? s->outlet_pb[b] : 0.0) - s->P[o]);
    if (s->act_bnd[ib]) ff[f] = ff[f] - dt * dot3(&s->gp[3 * o], &s->k_bnd[3 * ib]);
  }
  for (int ib = 0; ib < s->n_per_bf; ++ib) {
    const int b = s->per_bf[ib];
    const int pair = s->m.periodic_pair[b];
    if (pair < 0) continue;
    const int f = s->nint + b, pf = s->nint + pair;
    const int os = s->m.face_owner[f], op = s->m.face_owner[pf];
    if (os == op) continue;
    if (os > op) ff[f] = -ff[pf];
  }
}

/* correct_velocity_psolver_dual (src/solver/mod_fv_solver.f90:2927-3151) */
static void correct_velocity(om_solver *s, double dt, int with_cross, int ptot) {
  const int do_cross = with_cross && !s->skip_nonortho;
  double *gpc = s->gp_cell;
  memset(gpc, 0, sizeof(double) * 3 * (size_t)s->nc);
  for (int idx = 0; idx < s->npf; ++idx) {
    if (s->ps_skip[idx]) continue;
    const int o = s->ps_owner[idx], nb = s->ps_nb[idx], fm = s->ps_fmesh[idx];
    double g = s->ps_coeff[idx] * (s->P[nb] - s->P[o]);
    if (do_cross && s->act_int[idx]) {
      const double gc = (fm >= s->nint) ? s->m.periodic_gc[fm - s->nint] : s->m.face_gc[fm];
      g = g + cross_flux(&s->gp[3 * o], &s->gp[3 * nb], gc, &s->k_int[3 * idx]);
    }
    const double *n = &s->m.face_normal[3 * fm];
    for (int i = 0; i < 3; ++i) gpc[3 * o + i] = gpc[3 * o + i] + g * n[i];
    for (int i = 0; i < 3; ++i) gpc[3 * nb + i] = gpc[3 * nb + i] + g * n[i];
  }
  for (int ib = 0; ib < s->n_nper_bf; ++ib) {
    const int b = s->nper_bf[ib], f = s->nint + b, o = s->m.face_owner[f];
    if (bc_of_face(s, f) != BC_OUTLET) continue;
    const double *n = &s->m.face_normal[3 * f];
    double d[3];
    for (int i = 0; i < 3; ++i) d[i] = s->m.face_centre[3 * f + i] - s->m.cell_centroid[3 * o + i];
    const double sn = fabs(d[0] * n[0] + d[1] * n[1] + d[2] * n[2]);
    const double coeff = s->m.face_area[f] / fmax(sn, SMALL);
    double g = coeff * ((ptot ? s->outlet_pb[b] : 0.0) - s->P[o]);
    if (do_cross && s->act_bnd[ib]) g = g + dot3(&s->gp[3 * o], &s->k_bnd[3 * ib]);
    for (int i = 0; i < 3; ++i) gpc[3 * o + i] = gpc[3 * o + i] + g * n[i];
  }
  if (!do_cross) {
    double fb[3];
    body_force_now(s, fb);
    const double bb = s->p.boussinesq ? s->p.beta : 0.0;
    for (int ib = 0; ib < s->n_nper_bf; ++ib) {
      const int b = s->nper_bf[ib], f = s->nint + b, o = s->m.face_owner[f];
      if (bc_of_face(s, f) == BC_OUTLET) continue;
      const double *n = &s->m.face_normal[3 * f];
      const double tb = s->p.boussinesq ? s->T[o] : 0.0;
      const double d1 = fb[0] - bb * (tb - s->p.T_ref) * s->p.g[0], d2 = fb[1] - bb * (tb - s->p.T_ref) * s->p.g[1],
                   d3 = fb[2] - bb * (tb - s->p.T_ref) * s->p.g[2];
      const double fn = (d1 * n[0] + d2 * n[1] + d3 * n[2]) * s->m.face_area[f];
      for (int i = 0; i < 3; ++i) gpc[3 * o + i] = gpc[3 * o + i] + fn * n[i];
    }
  }
  for (int c = 0; c < s->nc; ++c) {
    const double *T = &s->Tinv[9 * (size_t)c]; /* row-major T[3i+j] */
    double g[3];
    for (int i = 0; i < 3; ++i) g[i] = T[3 * i] * gpc[3 * c] + T[3 * i + 1] * gpc[3 * c + 1] + T[3 * i + 2] * gpc[3 * c + 2];
    for (int i = 0; i < 3; ++i) {
      gpc[3 * c + i] = g[i];
      s->U[3 * c + i] = s->U[3 * c + i] - dt * g[i];
    }
  }
}

/* anchor_pressure_to_outlet (src/solver/mod_fv_solver.f90:3610-3639) */
static void anchor_pressure(om_solver *s) {
  if (s->n_outlet > 0) return;
  double pm = 0.0;
  for (int c = 0; c < s->nc; ++c) pm = pm + s->P[c] * s->m.cell_volume[c];
  const double shift = (s->V_total > 0.0) ? pm / s->V_total : 0.0;
  for (int c = 0; c < s->nc; ++c) s->P[c] = s->P[c] - shift;
}

/* ---------------------------------------------------------------- pressure solve */

static double vdot(int n, const double *a, const double *b) {
  double r = 0.0;
  for (int i = 0; i < n; ++i) r = r + a[i] * b[i];
  return r;
}

static void spmv(int n, int nf, const int32_t *own, const int32_t *nbr, const double *coeff, const double *diag,
                 const double *x, double *y) {
  for (int i = 0; i < n; ++i) y[i] = diag[i] * x[i];
  for (int f = 0; f < nf; ++f) {
    if (own[f] < 0 || nbr[f] < 0) continue;
    const double c = coeff[f];
    if (!isfinite(c) || c == 0.0) continue;
    y[own[f]] = y[own[f]] - c * x[nbr[f]];
    y[nbr[f]] = y[nbr[f]] - c * x[own[f]];
  }
}

static void remove_mean(int n, double *v) {
  double sm = 0.0;
  for (int i = 0; i < n; ++i) sm = sm + v[i];
  const double mean = sm / (double)n;
  for (int i = 0; i < n; ++i) v[i] = v[i] - mean;
}

/* Threads of the CG below (om_set_threads; 1 by default). With one thread the CG runs the
 * sequential face loops above; with more, the same Jacobi CG runs as OpenMP loops over rows of a
 * gathered (CSR) copy of the operator -- the same recurrence in another summation order. Only the
 * CPU baseline of the measurement asks for more than one thread. */
static int g_threads = 1;
void om_set_threads(int n) { g_threads = n > 1 ? n : 1; }
int om_get_threads(void) { return g_threads; }

static int cg_jacobi_omp(int n, int nfaces, const int32_t *owner, const int32_t *neighbour, const double *coeff,
                         const double *diag_in, const double *b_in, double *x, double rtol, double atol, int max_it,
                         int nullspace, int *its, double *resid);

int om_cg_jacobi(int n, int nfaces, const int32_t *owner, const int32_t *neighbour, const double *coeff,
                 const double *diag_in, const double *b_in, double *x, double rtol, double atol, int max_it,
                 int nullspace, int *its, double *resid) {
  if (g_threads > 1)
    return cg_jacobi_omp(n, nfaces, owner, neighbour, coeff, diag_in, b_in, x, rtol, atol, max_it, nullspace, its,
                         resid);
  double *diag = dz(n), *b = dz(n), *r = dz(n), *z = dz(n), *p = dz(n), *q = dz(n);
  for (int i = 0; i < n; ++i) {
    double d = diag_in[i];
    if (!isfinite(d) || fabs(d) < 1.0e-20) d = 1.0;
    diag[i] = d;
    b[i] = isfinite(b_in[i]) ? b_in[i] : 0.0;
    if (!isfinite(x[i])) x[i] = 0.0;
  }
  if (nullspace) remove_mean(n, b);
  const double bnorm = sqrt(vdot(n, b, b));
  const double ttol = fmax(rtol * bnorm, atol);
  spmv(n, nfaces, owner, neighbour, coeff, diag, x, q);
  for (int i = 0; i < n; ++i) r[i] = b[i] - q[i];
  double dp = sqrt(vdot(n, r, r));
  int it = 0, conv = (dp <= ttol);
  double beta = 0.0, betaold = 1.0;
  if (!conv) {
    for (int i = 0; i < n; ++i) z[i] = r[i] / diag[i];
    if (nullspace) remove_mean(n, z);
    beta = vdot(n, r, z);
    while (it < max_it) {
      if (it == 0) {
        memcpy(p, z, sizeof(double) * n);
      } else {
        const double bb = beta / betaold;
        for (int i = 0; i < n; ++i) p[i] = z[i] + bb * p[i];
      }
      spmv(n, nfaces, owner, neighbour, coeff, diag, p, q);
      const double dpi = vdot(n, p, q);
      const double a = beta / dpi;
      for (int i = 0; i < n; ++i) {
        x[i] = x[i] + a * p[i];
        r[i] = r[i] - a * q[i];
      }
      it++;
      dp = sqrt(vdot(n, r, r));
      if (dp <= ttol) {
        conv = 1;
        break;
      }
      for (int i = 0; i < n; ++i) z[i] = r[i] / diag[i];
      if (nullspace) remove_mean(n, z);
      betaold = beta;
      beta = vdot(n, r, z);
    }
  }
  /* true residual against the caller's rhs (KSPBuildResidual) */
  spmv(n, nfaces, owner, neighbour, coeff, diag, x, q);
  double rr = 0.0;
  for (int i = 0; i < n; ++i) {
    const double bi = isfinite(b_in[i]) ? b_in[i] : 0.0;
    rr = rr + (bi - q[i]) * (bi - q[i]);
  }
  *resid = sqrt(rr);
  *its = it;
  free(diag); free(b); free(r); free(z); free(p); free(q);
  return conv ? 0 : -1;
}

/* The Jacobi CG of om_cg_jacobi on g_threads OpenMP threads: rows gathered into CSR once per call. */
static int cg_jacobi_omp(int n, int nfaces, const int32_t *owner, const int32_t *neighbour, const double *coeff,
                         const double *diag_in, const double *b_in, double *x, double rtol, double atol, int max_it,
                         int nullspace, int *its, double *resid) {
  const int nt = g_threads;
  int *ptr = (int *)calloc((size_t)n + 1, sizeof(int));
  for (int f = 0; f < nfaces; ++f) {
    if (owner[f] < 0 || neighbour[f] < 0 || !isfinite(coeff[f]) || coeff[f] == 0.0) continue;
    ptr[owner[f] + 1]++;
    ptr[neighbour[f] + 1]++;
  }
  for (int i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
  int *col = (int *)malloc(sizeof(int) * (size_t)(ptr[n] > 0 ? ptr[n] : 1));
  double *val = (double *)malloc(sizeof(double) * (size_t)(ptr[n] > 0 ? ptr[n] : 1));
  int *at = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; ++i) at[i] = ptr[i];
  for (int f = 0; f < nfaces; ++f) {
    if (owner[f] < 0 || neighbour[f] < 0 || !isfinite(coeff[f]) || coeff[f] == 0.0) continue;
    col[at[owner[f]]] = neighbour[f];
    val[at[owner[f]]++] = coeff[f];
    col[at[neighbour[f]]] = owner[f];
    val[at[neighbour[f]]++] = coeff[f];
  }
  double *diag = dz(n), *b = dz(n), *r = dz(n), *z = dz(n), *p = dz(n), *q = dz(n);
  double sb = 0.0;
#pragma omp parallel for num_threads(nt) reduction(+ : sb)
  for (int i = 0; i < n; ++i) {
    double d = diag_in[i];
    if (!isfinite(d) || fabs(d) < 1.0e-20) d = 1.0;
    diag[i] = d;
    b[i] = isfinite(b_in[i]) ? b_in[i] : 0.0;
    if (!isfinite(x[i])) x[i] = 0.0;
    sb += b[i];
  }
  if (nullspace) {
    const double mb = sb / (double)n;
#pragma omp parallel for num_threads(nt)
    for (int i = 0; i < n; ++i) b[i] -= mb;
  }
#define OM_SPMV(xv, yv)                                                              \
  _Pragma("omp parallel for num_threads(nt)") for (int i = 0; i < n; ++i) {          \
    double y = diag[i] * (xv)[i];                                                    \
    for (int k = ptr[i]; k < ptr[i + 1]; ++k) y -= val[k] * (xv)[col[k]];            \
    (yv)[i] = y;                                                                     \
  }
  double bb = 0.0;
#pragma omp parallel for num_threads(nt) reduction(+ : bb)
  for (int i = 0; i < n; ++i) bb += b[i] * b[i];
  const double ttol = fmax(rtol * sqrt(bb), atol);
  OM_SPMV(x, q)
  double rr = 0.0;
#pragma omp parallel for num_threads(nt) reduction(+ : rr)
  for (int i = 0; i < n; ++i) {
    r[i] = b[i] - q[i];
    rr += r[i] * r[i];
  }
  int it = 0, conv = sqrt(rr) <= ttol;
  double beta = 0.0, betaold = 1.0;
  while (!conv && it < max_it) {
    double rz = 0.0, sz = 0.0, sr = 0.0;
#pragma omp parallel for num_threads(nt) reduction(+ : rz, sz, sr)
    for (int i = 0; i < n; ++i) {
      z[i] = r[i] / diag[i];
      rz += r[i] * z[i];
      sz += z[i];
      sr += r[i];
    }
    const double zm = nullspace ? sz / (double)n : 0.0;
    betaold = beta;
    beta = rz - zm * sr; /* r.(z - mean z) */
    const double bq = it == 0 ? 0.0 : beta / betaold;
#pragma omp parallel for num_threads(nt)
    for (int i = 0; i < n; ++i) p[i] = (z[i] - zm) + bq * p[i];
    double pq = 0.0;
    OM_SPMV(p, q)
#pragma omp parallel for num_threads(nt) reduction(+ : pq)
    for (int i = 0; i < n; ++i) pq += p[i] * q[i];
    const double a = beta / pq;
    rr = 0.0;
#pragma omp parallel for num_threads(nt) reduction(+ : rr)
    for (int i = 0; i < n; ++i) {
      x[i] += a * p[i];
      r[i] -= a * q[i];
      rr += r[i] * r[i];
    }
    it++;
    conv = sqrt(rr) <= ttol;
  }
  OM_SPMV(x, q)
  double tr = 0.0;
#pragma omp parallel for num_threads(nt) reduction(+ : tr)
  for (int i = 0; i < n; ++i) {
    const double bi = isfinite(b_in[i]) ? b_in[i] : 0.0;
    tr += (bi - q[i]) * (bi - q[i]);
  }
#undef OM_SPMV
  *resid = sqrt(tr);
  *its = it;
  free(diag); free(b); free(r); free(z); free(p); free(q);
  free(ptr); free(col); free(val); free(at);
  return conv ? 0 : -1;
}

/* pressure_solver_solve (src/solver/mod_pressure_solver.f90:638-806) */
static int pressure_solve(om_solver *s, double atol, int ptot) {
  const int nc = s->nc;
  for (int c = 0; c < nc; ++c) s->cg_b[c] = s->rhs_p[c];
  if (ptot)
    for (int i = 0; i < s->n_dir; ++i) s->cg_b[s->dir_rows[i]] = s->cg_b[s->dir_rows[i]] + s->dir_vals[i];
  int its = 0;
  double res = 0.0;
  const int max_it = (s->p.p_max_iter > 2000) ? s->p.p_max_iter : 2000;
  /* skipped faces carry a zero coefficient */
  const int rc = om_cg_jacobi(nc, s->npf, s->ps_owner, s->ps_nb, s->ps_coeff, s->ps_diag, s->cg_b, s->P, s->p.p_tol, atol,
                              max_it, s->n_dir == 0, &its, &res);
  s->n_solves++;
  s->n_iters += its;
  s->last_iters = its;
  return rc;
}

/* ---------------------------------------------------------------- exact cell projection */

/* build_cell_projection (src/solver/mod_fv_solver.f90:3164-3221): the pressure faces that are not
 * skipped, with the mesh face's area vector and the interpolation weight (face gc, periodic gc on
 * periodic edges), then the outlet faces among the non-periodic boundary faces. */
static void build_cell_projection(om_solver *s) {
  s->cp_own = iz(s->npf); s->cp_nbr = iz(s->npf);
  s->cp_sf = dz(3 * (size_t)s->npf); s->cp_gc = dz(s->npf);
  int n = 0;
  for (int idx = 0; idx < s->npf; ++idx) {
    if (s->ps_skip[idx]) continue;
    const int fm = s->ps_fmesh[idx];
    s->cp_own[n] = s->ps_owner[idx];
    s->cp_nbr[n] = s->ps_nb[idx];
    for (int k = 0; k < 3; ++k) s->cp_sf[3 * n + k] = s->m.face_sf[3 * fm + k];
    s->cp_gc[n] = (fm >= s->nint) ? s->m.periodic_gc[fm - s->nint] : s->m.face_gc[fm];
    n++;
  }
  s->cp_n = n;
  int nout = 0;
  for (int ib = 0; ib < s->n_nper_bf; ++ib)
    if (bc_of_face(s, s->nint + s->nper_bf[ib]) == BC_OUTLET) nout++;
  s->cp_bcell = iz(nout);
  s->cp_bsf = dz(3 * (size_t)nout);
  nout = 0;
  for (int ib = 0; ib < s->n_nper_bf; ++ib) {
    const int f = s->nint + s->nper_bf[ib];
    if (bc_of_face(s, f) != BC_OUTLET) continue;
    s->cp_bcell[nout] = s->m.face_owner[f];
    for (int k = 0; k < 3; ++k) s->cp_bsf[3 * nout + k] = s->m.face_sf[3 * f + k];
    nout++;
  }
  s->cp_nout = nout;
}

/* cp_divergence (mod_cell_projection.f90:89-130): outflow sum of the gc-interpolated flux
 * (gc u_own + (1-gc) u_nbr).Sf, plus u.Sf on outlet faces. */
void om_proj_divergence(om_solver *s, const double *u, double *out) {
  for (int c = 0; c < s->nc; ++c) out[c] = 0.0;
  for (int f = 0; f < s->cp_n; ++f) {
    const int o = s->cp_own[f], n = s->cp_nbr[f];
    const double w = s->cp_gc[f], *sf = &s->cp_sf[3 * f];
    const double fl = (w * u[3 * o] + (1.0 - w) * u[3 * n]) * sf[0] + (w * u[3 * o + 1] + (1.0 - w) * u[3 * n + 1]) * sf[1] +
                      (w * u[3 * o + 2] + (1.0 - w) * u[3 * n + 2]) * sf[2];
    out[o] = out[o] + fl;
    out[n] = out[n] - fl;
  }
  for (int f = 0; f < s->cp_nout; ++f) {
    const int o = s->cp_bcell[f];
    const double *bsf = &s->cp_bsf[3 * f];
    const double fl = u[3 * o] * bsf[0] + u[3 * o + 1] * bsf[1] + u[3 * o + 2] * bsf[2];
    out[o] = out[o] + fl;
  }
}

/* cp_adjoint_gradient (:133-188): v = V^-1 B^T chi, zero Dirichlet chi on outlet faces. */
void om_proj_adjoint(om_solver *s, const double *chi, double *v) {
  memset(v, 0, sizeof(double) * 3 * (size_t)s->nc);
  for (int f = 0; f < s->cp_n; ++f) {
    const int o = s->cp_own[f], n = s->cp_nbr[f];
    const double w = s->cp_gc[f], *sf = &s->cp_sf[3 * f];
    const double d = chi[o] - chi[n];
    for (int k = 0; k < 3; ++k) v[3 * o + k] = v[3 * o + k] + w * d * sf[k];
    for (int k = 0; k < 3; ++k) v[3 * n + k] = v[3 * n + k] + (1.0 - w) * d * sf[k];
  }
  for (int f = 0; f < s->cp_nout; ++f) {
    const int o = s->cp_bcell[f];
    for (int k = 0; k < 3; ++k) v[3 * o + k] = v[3 * o + k] + chi[o] * s->cp_bsf[3 * f + k];
  }
  for (int c = 0; c < s->nc; ++c) {
    const double iv = 1.0 / s->m.cell_volume[c];
    for (int k = 0; k < 3; ++k) v[3 * c + k] = v[3 * c + k] * iv;
  }
}

/* cproj_precond of project_cells_exact (mod_fv_solver.f90:3276-3281): z = 0, then the pressure
 * solve of r at rtol 0.05, 50 iterations, atol off (pressure_solver_solve's 1e-50 default). */
static void cp_precond(om_solver *s, int kind, const double *r, double *z) {
  const int nc = s->nc;
  if (kind == 0) {
    memcpy(z, r, sizeof(double) * nc);
    return;
  }
  memset(z, 0, sizeof(double) * nc);
  int its = 0;
  double res = 0.0;
  om_cg_jacobi(nc, s->npf, s->ps_owner, s->ps_nb, s->ps_coeff, s->ps_diag, r, z, 0.05, 1.0e-50, 50, s->n_dir == 0,
               &its, &res);
}

/* cp_project (:192-303): flexible PCG on (B V^-1 B^T) chi = B u + bconst; stops when max|r|
 * falls to 1e-12 max|b| or, from the second iteration, when the last step's energy alpha^2 p.q is
 * below tol^2 of the accumulated energy; then u -= V^-1 B^T chi. */
int om_proj_project(om_solver *s, double *u, const double *bconst, int precond, double tol, int max_it) {
  const int nc = s->nc;
  double *b = dz(nc), *x = dz(nc), *r = dz(nc), *z = dz(nc), *p = dz(nc), *q = dz(nc), *rold = dz(nc), *chi = dz(nc);
  double *v = dz(3 * (size_t)nc);
  om_proj_divergence(s, u, b);
  double smax = 0.0;
  for (int c = 0; c < nc; ++c) {
    b[c] = b[c] + bconst[c];
    r[c] = b[c];
    smax = fmax(smax, fabs(r[c]));
  }
  int n_it = 0;
  const double bn = smax;
  if (bn > DBL_MIN) {
    cp_precond(s, precond, r, z);
    double s1 = 0.0;
    for (int c = 0; c < nc; ++c) {
      p[c] = z[c];
      s1 = s1 + r[c] * z[c];
    }
    double rz = s1, en_tot = 0.0;
    for (int it = 1; it <= max_it; ++it) {
      n_it = it;
      memcpy(chi, p, sizeof(double) * nc);
      om_proj_adjoint(s, chi, v);
      om_proj_divergence(s, v, q);
      s1 = 0.0;
      for (int c = 0; c < nc; ++c) s1 = s1 + p[c] * q[c];
      const double pq = s1;
      if (fabs(pq) <= DBL_MIN) break;
      const double alpha = rz / pq;
      const double en_new = alpha * alpha * pq;
      en_tot = en_tot + en_new;
      smax = 0.0;
      for (int c = 0; c < nc; ++c) {
        rold[c] = r[c];
        x[c] = x[c] + alpha * p[c];
        r[c] = r[c] - alpha * q[c];
        smax = fmax(smax, fabs(r[c]));
      }
      const double rn = smax;
      if (rn <= 1.0e-12 * bn || (it >= 2 && en_new <= tol * tol * en_tot)) break;
      cp_precond(s, precond, r, z);
      double t1 = 0.0, t2 = 0.0;
      for (int c = 0; c < nc; ++c) {
        t1 = t1 + (r[c] - rold[c]) * z[c];
        t2 = t2 + r[c] * z[c];
      }
      const double beta = fmax(0.0, t1 / rz);
      rz = t2;
      for (int c = 0; c < nc; ++c) p[c] = z[c] + beta * p[c];
    }
    memcpy(chi, x, sizeof(double) * nc);
    om_proj_adjoint(s, chi, v);
    for (int c = 0; c < nc; ++c)
      for (int k = 0; k < 3; ++k) u[3 * c + k] = u[3 * c + k] - v[3 * c + k];
  }
  free(b); free(x); free(r); free(z); free(p); free(q); free(rold); free(chi); free(v);
  return n_it;
}

/* project_cells_exact (mod_fv_solver.f90:3224-3283): fixed boundary outflow of inlet-type faces into
 * bconst, the outlet pressure as the boundary-face force -dt p_b Sf / V on the outlet cell, then
 * the projection with the pressure solve as preconditioner (100 iterations at most). */
static void project_cells_exact(om_solver *s, double dt) {
  const int nc = s->nc;
  double *bconst = dz(nc);
  for (int ib = 0; ib < s->n_nper_bf; ++ib) {
    const int b = s->nper_bf[ib], f = s->nint + b, o = s->m.face_owner[f];
    const int bc = bc_of_face(s, f);
    const double *sf = &s->m.face_sf[3 * f];
    if (bc == BC_WALL_NOSLIP || bc == BC_WALL_MODEL || bc == BC_WALL_VELOCITY || bc == BC_SLIP || bc == BC_SYMMETRY)
      continue;
    if (bc == BC_OUTLET) {
      const double fl = dt * s->outlet_pb[b] / s->m.cell_volume[o];
      for (int k = 0; k < 3; ++k) s->U[3 * o + k] = s->U[3 * o + k] - fl * sf[k];
    } else {
      const double *ub = &s->U_bnd[3 * b];
      bconst[o] = bconst[o] + (ub[0] * sf[0] + ub[1] * sf[1] + ub[2] * sf[2]);
    }
  }
  const int its = om_proj_project(s, s->U, bconst, 1, s->p.cell_projection_tol, 100);
  s->n_proj++;
  s->n_proj_its += its;
  free(bconst);
}

/* ---------------------------------------------------------------- time step */

/* solver_cfl_rates + solver_compute_dt (src/solver/mod_fv_solver.f90:3718-3899) */
static double compute_dt(om_solver *s) {
  double *fs = s->dt_flux, *ds = s->dt_diff;
  memset(fs, 0, sizeof(double) * s->nc);
  memset(ds, 0, sizeof(double) * s->nc);
  for (int f = 0; f < s->nf; ++f) {
    const int o = s->m.face_owner[f], nb = s->m.face_neighbour[f];
    const double *sf = &s->m.face_sf[3 * f];
    const double area = sqrt(sf[0] * sf[0] + sf[1] * sf[1] + sf[2] * sf[2]);
    const double uf = s->u_face_ready ? s->U_face[f] : (s->U[3 * o] * sf[0] + s->U[3 * o + 1] * sf[1] + s->U[3 * o + 2] * sf[2]);
    fs[o] = fs[o] + fabs(uf);
    if (nb >= 0) fs[nb] = fs[nb] + fabs(uf);
    const double w = area / fmax(s->m.face_dcf[f], SMALL);
    if (nb >= 0) {
      const double gc = s->m.face_gc[f];
      const double nue = s->p.nu + gc * fmax(0.0, s->nu_sgs[o]) + (1.0 - gc) * fmax(0.0, s->nu_sgs[nb]);
      ds[o] = ds[o] + nue * w;
      ds[nb] = ds[nb] + nue * w;
    } else if (bc_of_face(s, f) != BC_WALL_MODEL) {
      ds[o] = ds[o] + (s->p.nu + fmax(0.0, s->nu_sgs[o])) * w;
    }
  }
  double cfl = 0.0, dif = 0.0;
  for (int c = 0; c < s->nc; ++c) {
    cfl = fmax(cfl, fs[c] / (s->m.cell_volume[c] + SMALL));
    dif = fmax(dif, ds[c] / (s->m.cell_volume[c] + SMALL));
  }
  double dc = 1e308, dd = 1e308;
  if (cfl > SMALL) dc = s->p.cfl_max / cfl;
  if (dif > SMALL) dd = 1.0 / dif;
  double dt = fmin(dc, dd);
  if (!(dt > 0.0)) dt = s->p.dt;
  dt = fmin(dt, s->p.dt);
  const double dtmin = 1.0e-6 * s->p.dt;
  if (!(dt > 0.0)) dt = s->p.dt;
  else if (dt < dtmin) return -1.0;
  else dt = fmax(dt, dtmin);
  return dt;
}

static double atol_for(om_solver *s, double gdt) {
  double ub1 = 0.0, ub2 = 0.0;
  for (int c = 0; c < s->nc; ++c) {
    const double *u = &s->U[3 * c];
    ub1 = ub1 + sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) * s->m.cell_volume[c];
    ub2 = ub2 + s->m.cell_volume[c];
  }
  const double us = ub1 / fmax(ub2, SMALL);
  return fmax(1.0e-12 * fmax(us, 1.0) * s->V_min / gdt, s->p_atol_factor);
}

/* solver_project_once (src/solver/mod_fv_solver.f90:837-917) */
int om_project_once(om_solver *s, double dt) {
  const int nc = s->nc;
  if (dt <= 0.0) dt = s->p.dt;
  memset(s->rhs_p, 0, sizeof(double) * nc);
  memset(s->phi_prev, 0, sizeof(double) * nc);
  memset(s->phi_prev_bnd, 0, sizeof(double) * s->nb);
  memset(s->gp_rc, 0, sizeof(double) * 3 * (size_t)nc);
  refresh_bc(s);
  face_fluxes_and_rhs(s, dt, 0);
  int rc = 0;
  if (s->p.p_max_iter > 0) rc = pressure_solve(s, atol_for(s, dt), 1);
  anchor_pressure(s);
  correct_face_fluxes(s, dt, 1);
  memcpy(s->U_face, s->face_flux, sizeof(double) * s->nf);
  s->u_face_ready = 1;
  apply_pressure_bc(s, s->P, s->p_bnd, 0, NULL, 1.0, s->outlet_pb);
  if (s->p.cell_projection == OM_PROJ_EXACT)
    project_cells_exact(s, dt);
  else
    correct_velocity(s, dt, 1, 1);
  apply_velocity_bc(s, 0, NULL);
  if (s->p.rotational) {
    for (int c = 0; c < nc; ++c) s->P[c] = s->P[c] + (s->p.nu + s->nu_sgs[c]) * s->rhs_p[c] * dt * s->inv_vol[c];
    apply_pressure_bc(s, s->P, s->p_bnd, 0, NULL, 1.0, s->outlet_pb);
  }
  memcpy(s->p_old, s->P, sizeof(double) * nc);
  memcpy(s->phi_prev, s->P, sizeof(double) * nc);
  memcpy(s->phi_prev_bnd, s->p_bnd, sizeof(double) * s->nb);
  apply_pressure_bc(s, s->P, s->p_bnd, 0, NULL, 1.0, s->outlet_pb);
  compute_pressure_gradient(s);
  return rc;
}

/* advance_rk3w (src/solver/mod_fv_solver.f90:1180-1614) */
/* Stage-start closures of solver_rk3w_stage (src/solver/mod_fv_solver.f90): velocity gradients,
 * SGS viscosity (stability cap from dt_stab) and wall-model traction. */
static void stage_closures(om_solver *s, double dt_stab, int stage) {
  compute_all_gradients(s);
  if (s->p.sgs == OM_SGS_VREMAN)
    vreman(s, dt_stab);
  else if (s->p.sgs == OM_SGS_SMAGORINSKY)
    smagorinsky(s, dt_stab);
  else if (s->p.sgs == OM_SGS_AMD)
    amd(s, dt_stab);
  else if (s->p.sgs == OM_SGS_DYNAMIC)
    dynamic_smag(s, dt_stab);
  else
    memset(s->nu_sgs, 0, sizeof(double) * s->nc);
  if (s->p.wall_model == OM_WM_NONE)
    memset(s->tau_w, 0, sizeof(double) * 3 * (size_t)s->nb);
  else if (s->p.wall_model == OM_WM_MOST)
    most_wall(s, stage == 0);
  else
    wall_model_dt(s, stage == 0 ? s->wm_dt : 0.0);
}

int om_eval_closures(om_solver *s, double dt_stab) {
  refresh_bc(s);
  stage_closures(s, dt_stab, 1);
  return 0;
}

/* Single routines of src/numerics/mod_gradient.f90, mod_advection.f90, mod_viscous.f90 and
 * src/sgs/mod_vreman.f90 / mod_amd.f90 / mod_wall_model.f90, one call each, for the operator-level
 * tests. The momentum routines read the solver's U, U_bnd, U_face, grad_U and nu_sgs (set them
 * through om_array first); the scalar routines take their field, boundary values and gradient. */
void om_op_gradient_vector(om_solver *s, int gg, int iterations, double *out) {
  if (!gg) {
    ls_gradient_vector(s, s->U, s->U_bnd, out);
    return;
  }
  double *prev = dz(9 * (size_t)s->nc);
  gg_gradient_vector(s, s->U, s->U_bnd, NULL, out);
  for (int it = 2; it <= iterations; ++it) {
    memcpy(prev, out, sizeof(double) * 9 * (size_t)s->nc);
    gg_gradient_vector(s, s->U, s->U_bnd, prev, out);
  }
  free(prev);
}

void om_op_gradient_scalar(om_solver *s, int gg, int iterations, const double *phi, const double *pb, double *out) {
  if (!gg) {
    ls_gradient_scalar(s, phi, pb, out);
    return;
  }
  double *prev = dz(3 * (size_t)s->nc);
  gg_gradient_scalar(s, phi, pb, NULL, out);
  for (int it = 2; it <= iterations; ++it) {
    memcpy(prev, out, sizeof(double) * 3 * (size_t)s->nc);
    gg_gradient_scalar(s, phi, pb, prev, out);
  }
  free(prev);
}

void om_op_advection_vector(om_solver *s, int scheme, double *out) {
  const int keep = s->p.scheme;
  s->p.scheme = scheme;
  if (scheme == OM_SCHEME_KEP)
    advection_kep(s, out);
  else if (scheme == OM_SCHEME_CDS)
    advection_cds(s, out);
  else
    advection_recon(s, out);
  s->p.scheme = keep;
}

void om_op_viscous_vector(om_solver *s, double *out) { viscous_vector(s, out); }

void om_op_advection_scalar(om_solver *s, int scheme, const double *phi, const double *pb, const double *g,
                            double *out) {
  if (scheme == OM_SCHEME_CDS)
    advection_scalar_cds(s, phi, pb, g, out);
  else
    advection_scalar_recon(s, scheme, phi, pb, g, out);
}

void om_op_viscous_scalar(om_solver *s, const double *phi, const double *pb, const double *g, double alpha,
                          double prt, double *out) {
  viscous_scalar(s, phi, pb, g, alpha, prt, out);
}

/* model 1 Vreman (coef C_v), 2 Smagorinsky (C_s), 3 AMD (C, wall_gradient) on the solver's grad_U */
void om_op_sgs(om_solver *s, int model, double coef, int wall_gradient, double dt_stab, double *out) {
  const om_params keep = s->p;
  double *nus = s->nu_sgs;
  s->nu_sgs = out;
  s->p.amd_wall_gradient = wall_gradient;
  if (model == OM_SGS_VREMAN) {
    s->p.c_vreman = coef;
    vreman(s, dt_stab);
  } else if (model == OM_SGS_SMAGORINSKY) {
    s->p.cs_smag = coef;
    smagorinsky(s, dt_stab);
  } else if (model == OM_SGS_AMD) {
    s->p.c_amd = coef;
    amd(s, dt_stab);
  } else {
    dynamic_smag(s, dt_stab);
  }
  s->nu_sgs = nus;
  s->p = keep;
}

/* the log-law wall traction of every wall-model face from the solver's U; dt advances the
 * solver's time filter (when configured) as the first stage of a step does */
void om_op_wall_model(om_solver *s, double dt, double *out) {
  double *t = s->tau_w;
  s->tau_w = out;
  wall_model_dt(s, dt);
  s->tau_w = t;
}

double om_wall_tau(double u_par, double y_w, double nu) { return wall_tau(u_par, y_w, nu); }

static int advance_rk3w(om_solver *s, double dt) {
  const int nc = s->nc;
  memset(s->H, 0, sizeof(double) * 3 * (size_t)nc);
  memset(s->H0, 0, sizeof(double) * 3 * (size_t)nc);
  memset(s->H_T, 0, sizeof(double) * nc);
  memset(s->H_T0, 0, sizeof(double) * nc);
  for (int q = 0; q < s->p.n_scalars; ++q) {
    memset(s->H_C[q], 0, sizeof(double) * nc);
    memset(s->H_C0[q], 0, sizeof(double) * nc);
  }
  memset(s->rhs_p, 0, sizeof(double) * nc);
  const double time_n = s->time;
  int rc = 0;
  s->wm_dt = dt;
  for (int k = 0; k < 3; ++k) {
    const double gdt = RK3_G[k] * dt;
    s->time = time_n + RK3_C[k] * dt;
    refresh_bc(s);
    stage_closures(s, gdt, k);
    explicit_rhs(s);
    const double a = RK3_A[k], bb = RK3_B[k];
    outlet_advance_boundary(s, dt, a, bb);
    for (int c = 0; c < nc; ++c) {
      for (int q = 0; q < 3; ++q) s->U[3 * c + q] = s->U[3 * c + q] + dt * (a * s->H[3 * c + q] + bb * s->H0[3 * c + q]);
      if (s->p.boussinesq) s->T[c] = s->T[c] + dt * (a * s->H_T[c] + bb * s->H_T0[c]);
    }
    for (int q = 0; q < s->p.n_scalars; ++q)
      for (int c = 0; c < nc; ++c) s->C[q][c] = s->C[q][c] + dt * (a * s->H_C[q][c] + bb * s->H_C0[q][c]);
    update_outlet_pressure(s);
    velocity_bc_both(s);
    memcpy(s->p_old, s->P, sizeof(double) * nc);
    if (s->p.cell_projection == OM_PROJ_EXACT)
      memset(s->gp_cell, 0, sizeof(double) * 3 * (size_t)nc); /* the projection takes the whole correction */
    else
      correct_velocity(s, gdt, 0, 1);
    memcpy(s->gp_rc, s->gp_cell, sizeof(double) * 3 * (size_t)nc);
    velocity_bc_both(s);
    compute_all_gradients(s);
    memcpy(s->phi_prev, s->P, sizeof(double) * nc);
    memcpy(s->phi_prev_bnd, s->p_bnd, sizeof(double) * s->nb);
    {
      const double atol = atol_for(s, gdt);
      int nmax = s->n_corr < 3 ? s->n_corr : 3;
      double stag = 0.9;
      if (s->diag_interface) {
        nmax = s->n_corr < 8 ? s->n_corr : 8;
        stag = 0.99;
      }
      double dprev = 1e308, dphi_prev = 1e308;
      if (s->p.nonortho_check) {
        nmax = s->n_corr < 8 ? s->n_corr : 8;
        memset(s->phi_it, 0, sizeof(double) * nc);
      }
      for (int jno = 1; jno <= nmax; ++jno) {
        face_fluxes_and_rhs(s, gdt, jno > 1);
        double dcur = 0.0;
        for (int c = 0; c < nc; ++c) dcur = fmax(dcur, fabs(s->rhs_p[c]));
        if (jno == 1) memset(s->P, 0, sizeof(double) * nc);
        if (pressure_solve(s, atol, 0) != 0) rc = -2;
        apply_pressure_bc(s, s->P, s->p_bnd, 1, s->skip_grad_p ? NULL : s->gp, 0.0, NULL);
        apply_pressure_bc(s, s->P, s->p_bnd, 2, NULL, 0.0, NULL);
        if (s->p.nonortho_check) {
          double dd = 0.0, dpp = 0.0;
          for (int c = 0; c < nc; ++c) {
            dd = fmax(dd, fabs(s->P[c] - s->phi_it[c]));
            dpp = fmax(dpp, fabs(s->P[c]));
            s->phi_it[c] = s->P[c];
          }
          if (jno > 1 && (dd < 1.0e-3 * dpp || dd > dphi_prev * stag)) break;
          dphi_prev = dd;
        } else if (jno > 1 && (dcur < atol || dcur > dprev * stag)) {
          break;
        }
        dprev = dcur;
      }
    }
    correct_face_fluxes(s, gdt, 0);
    memcpy(s->U_face, s->face_flux, sizeof(double) * s->nf);
    s->u_face_ready = 1;
    if (s->p.cell_projection == OM_PROJ_EXACT)
      project_cells_exact(s, gdt);
    else
      correct_velocity(s, gdt, 1, 0);
    velocity_bc_both(s);
    if (s->p.rotational) {
      for (int c = 0; c < nc; ++c) s->P[c] = s->P[c] + (s->p.nu + s->nu_sgs[c]) * s->rhs_p[c] * gdt * s->inv_vol[c];
      apply_pressure_bc(s, s->P, s->p_bnd, 1, s->skip_grad_p ? NULL : s->gp, 0.0, s->outlet_pb);
      apply_pressure_bc(s, s->P, s->p_bnd, 2, NULL, 0.0, NULL);
    }
    for (int c = 0; c < nc; ++c) s->P[c] = s->P[c] + s->p_old[c];
    anchor_pressure(s);
    apply_pressure_bc(s, s->P, s->p_bnd, 1, s->skip_grad_p ? NULL : s->gp, 1.0, s->outlet_pb);
    apply_pressure_bc(s, s->P, s->p_bnd, 2, NULL, 1.0, NULL);
    compute_pressure_gradient(s);
    memcpy(s->H0, s->H, sizeof(double) * 3 * (size_t)nc);
    if (s->p.boussinesq) memcpy(s->H_T0, s->H_T, sizeof(double) * nc);
    for (int q = 0; q < s->p.n_scalars; ++q) memcpy(s->H_C0[q], s->H_C[q], sizeof(double) * nc);
  }
  s->time = time_n;
  return rc;
}

int om_advance(om_solver *s, double dt_req, double *dt_used) {
  double dt = dt_req;
  const double safe = compute_dt(s);
  if (safe < 0.0) return -3;
  if (dt > safe) dt = safe;
  if (s->p.mass_flux) {
    /* update_mass_flux_forcing (src/solver/mod_fv_solver.f90:960-983) */
    double m[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < s->nc; ++c)
      for (int k = 0; k < 3; ++k) m[k] = m[k] + s->U[3 * c + k] * s->m.cell_volume[c];
    for (int k = 0; k < 3; ++k) s->mff[k] = (s->p.U_bulk_target[k] - m[k] / fmax(s->V_total, SMALL)) / dt;
  }
  const int rc = advance_rk3w(s, dt);
  /* apply_scalar_bc: no temperature or scalars here */
  s->time = s->time + dt;
  if (dt_used) *dt_used = dt;
  return rc;
}

/* compute_stats_sums (src/stats/mod_stats.f90:271-336) + solver_divergence_metrics (mod_fv_solver.f90:3643-3714) */
void om_stats(om_solver *s, double out[8]) {
  double E = 0, V = 0, eps = 0, mom[3] = {0, 0, 0}, sgs = 0, tn = 0, ta = 0, umax = 0;
  for (int c = 0; c < s->nc; ++c) {
    const double vol = s->m.cell_volume[c];
    const double *u = &s->U[3 * c];
    const double *g = &s->grad_U[9 * (size_t)c];
    E = E + 0.5 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) * vol;
    V = V + vol;
    const double S11 = g[GI(0, 0, 0)], S22 = g[GI(0, 1, 1)], S33 = g[GI(0, 2, 2)];
    const double S12 = 0.5 * (g[GI(0, 1, 0)] + g[GI(0, 0, 1)]);
    const double S13 = 0.5 * (g[GI(0, 2, 0)] + g[GI(0, 0, 2)]);
    const double S23 = 0.5 * (g[GI(0, 2, 1)] + g[GI(0, 1, 2)]);
    const double ss = S11 * S11 + S22 * S22 + S33 * S33 + 2.0 * (S12 * S12 + S13 * S13 + S23 * S23);
    eps = eps + (s->p.nu + s->nu_sgs[c]) * ss * vol;
    for (int k = 0; k < 3; ++k) mom[k] = mom[k] + u[k] * vol;
    sgs = sgs + s->nu_sgs[c] * vol;
    const double um = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    if (um > umax) umax = um;
  }
  for (int p = 0; p < s->npatch; ++p) {
    if (s->m.patch_bc[p] != BC_WALL_MODEL) continue;
    for (int q = 0; q < s->m.patch_nfaces[p]; ++q) {
      const int b = s->m.patch_start[p] + q;
      const double *t = &s->tau_w[3 * b];
      tn = tn + sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]) * s->m.face_area[s->nint + b];
      ta = ta + s->m.face_area[s->nint + b];
    }
  }
  out[0] = E / V;
  out[1] = 2.0 * eps / V;
  out[2] = umax;
  {
    const double a = mom[0] / V, b = mom[1] / V, c = mom[2] / V;
    out[3] = sqrt(a * a + b * b + c * c);
  }
  out[4] = (ta > SMALL) ? tn / ta : 0.0;
  out[5] = sgs / V;
  /* face-flux divergence (face_flux_divergence, :3532-3579) */
  double *div = s->div_diag;
  memset(div, 0, sizeof(double) * s->nc);
  for (int idx = 0; idx < s->npf; ++idx) {
    if (s->ps_skip[idx]) continue;
    const int o = s->ps_owner[idx], nb = s->ps_nb[idx];
    const double fl = s->face_flux[s->ps_fmesh[idx]];
    div[o] = div[o] + fl * s->inv_vol[o];
    div[nb] = div[nb] - fl * s->inv_vol[nb];
  }
  for (int ib = 0; ib < s->n_nper_bf; ++ib) {
    const int b = s->nper_bf[ib], f = s->nint + b, o = s->m.face_owner[f];
    div[o] = div[o] + s->face_flux[f] * s->inv_vol[o];
  }
  double mx = 0.0, as = 0.0;
  for (int c = 0; c < s->nc; ++c) {
    mx = fmax(mx, fabs(div[c]));
    as = as + fabs(div[c]) * s->m.cell_volume[c];
  }
  out[6] = mx;
  out[7] = as / V;
}

/* cell_velocity_divergence inside solver_divergence_metrics (src/solver/mod_fv_solver.f90:3582-3607,
 * 3641-3713): velocity boundary values refreshed without gradients, then per face the
 * gc-interpolated (interior) or boundary velocity flux, owner +, neighbour -. */
void om_divergence_cell(om_solver *s, double out[2]) {
  apply_velocity_bc(s, 0, NULL);
  double *div = s->div_diag;
  memset(div, 0, sizeof(double) * s->nc);
  for (int f = 0; f < s->nf; ++f) {
    const int o = s->m.face_owner[f], nb = s->m.face_neighbour[f];
    double uf[3];
    if (nb >= 0) {
      const double gc = s->m.face_gc[f];
      for (int k = 0; k < 3; ++k) uf[k] = gc * s->U[3 * o + k] + (1.0 - gc) * s->U[3 * nb + k];
    } else {
      for (int k = 0; k < 3; ++k) uf[k] = s->U_bnd[3 * (f - s->nint) + k];
    }
    const double *sf = &s->m.face_sf[3 * f];
    const double flux = uf[0] * sf[0] + uf[1] * sf[1] + uf[2] * sf[2];
    div[o] = div[o] + flux * s->inv_vol[o];
    if (nb >= 0) div[nb] = div[nb] - flux * s->inv_vol[nb];
  }
  double mx = 0.0, as = 0.0, V = 0.0;
  for (int c = 0; c < s->nc; ++c) {
    mx = fmax(mx, fabs(div[c]));
    as = as + fabs(div[c]) * s->m.cell_volume[c];
    V = V + s->m.cell_volume[c];
  }
  out[0] = mx;
  out[1] = as / V;
}

/* ---------------------------------------------------------------- set-up */

/* invert_sym3 (src/solver/mod_fv_solver.f90:3466-3491), row-major output */
static void invert_sym3(const double *A, double *Ai) {
#define A_(i, j) A[3 * ((i)-1) + ((j)-1)]
  const double c11 = A_(2, 2) * A_(3, 3) - A_(2, 3) * A_(3, 2);
  const double c12 = A_(1, 3) * A_(3, 2) - A_(1, 2) * A_(3, 3);
  const double c13 = A_(1, 2) * A_(2, 3) - A_(1, 3) * A_(2, 2);
  const double c22 = A_(1, 1) * A_(3, 3) - A_(1, 3) * A_(3, 1);
  const double c23 = A_(1, 3) * A_(2, 1) - A_(1, 1) * A_(2, 3);
  const double c33 = A_(1, 1) * A_(2, 2) - A_(1, 2) * A_(2, 1);
  const double det = A_(1, 1) * c11 + A_(1, 2) * (A_(2, 3) * A_(3, 1) - A_(2, 1) * A_(3, 3)) +
                     A_(1, 3) * (A_(2, 1) * A_(3, 2) - A_(2, 2) * A_(3, 1));
  const double tr = A_(1, 1) + A_(2, 2) + A_(3, 3);
#undef A_
  const double t = fmax(tr, SMALL);
  if (fabs(det) <= 1.0e-12 * t * t * t) {
    for (int q = 0; q < 9; ++q) Ai[q] = 0.0;
    Ai[0] = Ai[4] = Ai[8] = 3.0 / t;
    return;
  }
  Ai[0] = c11 / det; Ai[1] = c12 / det; Ai[2] = c13 / det;
  Ai[3] = c12 / det; Ai[4] = c22 / det; Ai[5] = c23 / det;
  Ai[6] = c13 / det; Ai[7] = c23 / det; Ai[8] = c33 / det;
}

om_solver *om_create(const om_mesh *m, const om_params *p) {
  om_solver *s = (om_solver *)calloc(1, sizeof(om_solver));
  s->m = *m;
  s->p = *p;
  const int nc = m->n_cells, nf = m->n_faces, nint = m->n_int_faces, nb = m->n_bnd_faces;
  s->nc = nc; s->nf = nf; s->nint = nint; s->nb = nb; s->npatch = m->n_patches;
  s->U = dz(3 * (size_t)nc); s->P = dz(nc); s->p_old = dz(nc); s->nu_sgs = dz(nc);
  s->U_bnd = dz(3 * (size_t)nb); s->p_bnd = dz(nb); s->phi_prev = dz(nc); s->phi_prev_bnd = dz(nb);
  s->tau_w = dz(3 * (size_t)nb); s->grad_U = dz(9 * (size_t)nc); s->gU_prev = dz(9 * (size_t)nc);
  s->face_flux = dz(nf); s->U_face = dz(nf); s->inv_vol = dz(nc);
  s->H = dz(3 * (size_t)nc); s->H0 = dz(3 * (size_t)nc); s->conv = dz(3 * (size_t)nc); s->visc = dz(3 * (size_t)nc);
  s->T = dz(nc); s->T_bnd = dz(nb); s->grad_T = dz(3 * (size_t)nc); s->gT_prev = dz(3 * (size_t)nc);
  s->H_T = dz(nc); s->H_T0 = dz(nc); s->conv_T = dz(nc); s->visc_T = dz(nc);
  if (p->n_scalars < 0 || p->n_scalars > OM_MAX_SCALARS) {
    om_free(s);
    return NULL;
  }
  for (int q = 0; q < p->n_scalars; ++q) {
    s->C[q] = dz(nc); s->C_bnd[q] = dz(nb); s->grad_C[q] = dz(3 * (size_t)nc);
    s->H_C[q] = dz(nc); s->H_C0[q] = dz(nc);
  }
  s->gC_prev = dz(3 * (size_t)nc);
  for (int c = 0; c < nc; ++c) s->T[c] = p->T_ref;
  for (int b = 0; b < nb; ++b) s->T_bnd[b] = p->T_ref;
  s->rhs_p = dz(nc); s->div_w = dz(nc); s->phi_it = dz(nc); s->div_diag = dz(nc);
  s->gp = dz(3 * (size_t)nc); s->gp_rc = dz(3 * (size_t)nc); s->gp_cell = dz(3 * (size_t)nc);
  s->p_tot = dz(nc); s->p_tot_bnd = dz(nb); s->gp_prev = dz(3 * (size_t)nc);
  s->Tinv = dz(9 * (size_t)nc); s->outlet_pb = dz(nb); s->Gf = dz(9 * (size_t)nc);
  s->ws_y = dz(nb); s->ws_cell = iz(nb); s->per_nb = iz(nb);
  s->built = (char *)calloc(nf ? nf : 1, 1);
  s->cg_b = dz(nc);
  s->dt_flux = dz(nc); s->dt_diff = dz(nc);
  for (int c = 0; c < nc; ++c) s->inv_vol[c] = 1.0 / fmax(m->cell_volume[c], SMALL);
  for (int b = 0; b < nb; ++b) s->per_nb[b] = resolve_periodic_nb(s, b);
  /* outlet_pb and patch classification (:413-422, 1766-1850) */
  s->per_bf = iz(nb); s->nper_bf = iz(nb);
  for (int q = 0; q < s->npatch; ++q) {
    for (int k = 0; k < m->patch_nfaces[q]; ++k) {
      const int b = m->patch_start[q] + k;
      s->outlet_pb[b] = p->patch_p_val ? p->patch_p_val[q] : 0.0;
      if (m->patch_bc[q] == BC_PERIODIC)
        s->per_bf[s->n_per_bf++] = b;
      else
        s->nper_bf[s->n_nper_bf++] = b;
    }
    if (m->patch_bc[q] == BC_WALL_MODEL) s->n_wm_patches++;
    if (m->patch_bc[q] == BC_OUTLET) s->n_outlet += m->patch_nfaces[q];
  }
  s->diag_interface = 0;
  for (int f = 0; f < nint; ++f)
    if (m->face_skew[f] > REFINE_SKEW_TOL) s->diag_interface = 1;
  /* solver_init: apply_all_bc, initialize_velocity_from_bc (zero without an inlet) */
  for (int q = 0; q < s->npatch; ++q)
    if (m->patch_bc[q] == BC_INLET) {
      for (int c = 0; c < nc; ++c)
        for (int k = 0; k < 3; ++k) s->U[3 * c + k] = p->patch_U_val[3 * q + k];
      break;
    }
  /* pressure_solver_init (src/solver/mod_pressure_solver.f90:100-452) */
  int nedge = 0;
  for (int q = 0; q < s->npatch; ++q) {
    if (m->patch_bc[q] != BC_PERIODIC) continue;
    for (int k = 0; k < m->patch_nfaces[q]; ++k) {
      const int b = m->patch_start[q] + k, pair = m->periodic_pair[b];
      if (pair < 0) continue;
      const int nbc = m->face_owner[nint + pair], o = m->face_owner[nint + b];
      if (nbc == o || o >= nbc) continue;
      nedge++;
    }
  }
  s->npf = nint + nedge;
  const int npf = s->npf;
  s->ps_owner = iz(npf); s->ps_nb = iz(npf); s->ps_fmesh = iz(npf); s->ps_skip = iz(npf);
  s->ps_coeff = dz(npf); s->ps_diag = dz(nc);
  for (int f = 0; f < nint; ++f) {
    s->ps_owner[f] = m->face_owner[f];
    s->ps_nb[f] = m->face_neighbour[f];
    s->ps_fmesh[f] = f;
  }
  int idx = nint;
  for (int q = 0; q < s->npatch; ++q) {
    if (m->patch_bc[q] != BC_PERIODIC) continue;
    for (int k = 0; k < m->patch_nfaces[q]; ++k) {
      const int b = m->patch_start[q] + k, pair = m->periodic_pair[b];
      if (pair < 0) continue;
      const int fg = nint + b, nbc = m->face_owner[nint + pair], o = m->face_owner[fg];
      if (nbc == o || o >= nbc) continue;
      s->ps_owner[idx] = o;
      s->ps_nb[idx] = nbc;
      s->ps_fmesh[idx] = fg;
      double pdcf = m->periodic_pair_dcf[b];
      const int pf = nint + pair;
      if (pdcf <= SMALL) pdcf = m->face_dcf[pf];
      if (pdcf <= SMALL) pdcf = m->face_dcf[fg];
      const double d_per = m->face_dcf[fg] + pdcf;
      const double area = 0.5 * (m->face_area[fg] + m->face_area[pf]);
      const double coeff = area / (d_per + SMALL);
      s->ps_coeff[idx] = coeff;
      s->ps_diag[o] = s->ps_diag[o] + coeff;
      s->ps_diag[nbc] = s->ps_diag[nbc] + coeff;
      idx++;
    }
  }
  for (int f = 0; f < nint; ++f) {
    const int o = m->face_owner[f], nbc = m->face_neighbour[f];
    double d_on[3];
    for (int i = 0; i < 3; ++i) d_on[i] = m->cell_centroid[3 * nbc + i] - m->cell_centroid[3 * o + i];
    double d_ortho = dot3(d_on, &m->face_normal[3 * f]);
    if (p->min_cos_theta > 0.0 && d_ortho < p->min_cos_theta * m->face_dcf[f]) d_ortho = p->min_cos_theta * m->face_dcf[f];
    const double coeff = (d_ortho > SMALL) ? m->face_area[f] / d_ortho : m->face_area[f] / (m->face_dcf[f] + SMALL);
    if (!isfinite(coeff) || coeff <= SMALL) {
      s->ps_skip[f] = 1;
      continue;
    }
    s->ps_coeff[f] = coeff;
    s->ps_diag[o] = s->ps_diag[o] + coeff;
    s->ps_diag[nbc] = s->ps_diag[nbc] + coeff;
  }
  s->dir_rows = iz(nb);
  s->dir_vals = dz(nb);
  for (int q = 0; q < s->npatch; ++q) {
    if (m->patch_bc[q] != BC_OUTLET) continue;
    for (int k = 0; k < m->patch_nfaces[q]; ++k) {
      const int b = m->patch_start[q] + k, fg = nint + b, o = m->face_owner[fg];
      double d[3];
      for (int i = 0; i < 3; ++i) d[i] = m->face_centre[3 * fg + i] - m->cell_centroid[3 * o + i];
      double dp = fabs(dot3(d, &m->face_normal[3 * fg]));
      if (dp <= SMALL) dp = m->face_dcf[fg];
      if (p->min_cos_theta > 0.0) dp = fmax(dp, p->min_cos_theta * m->face_dcf[fg]);
      if (dp <= SMALL) continue;
      const double cb = m->face_area[fg] / dp;
      if (!isfinite(cb) || cb <= SMALL) continue;
      s->ps_diag[o] = s->ps_diag[o] + cb;
      s->dir_rows[s->n_dir] = o;
      s->dir_vals[s->n_dir] = cb * p->patch_p_val[q];
      s->n_dir++;
    }
  }
  for (int c = 0; c < nc; ++c)
    if (!isfinite(s->ps_diag[c]) || s->ps_diag[c] <= SMALL) s->ps_diag[c] = 1.0;
  /* V_min / V_total / atol floor (:431-459) */
  s->V_min = 1e308;
  s->V_total = 0.0;
  for (int c = 0; c < nc; ++c) {
    s->V_min = fmin(s->V_min, m->cell_volume[c]);
    s->V_total = s->V_total + m->cell_volume[c];
  }
  s->p_atol_factor = 1.0e-15;
  /* detect_nonortho_skip + build_nonortho_cache (:1864-1996) */
  s->k_int = dz(3 * (size_t)npf);
  s->act_int = (char *)calloc(npf ? npf : 1, 1);
  s->k_bnd = dz(3 * (size_t)s->n_nper_bf);
  s->act_bnd = (char *)calloc(s->n_nper_bf ? s->n_nper_bf : 1, 1);
  int nact = 0, ngradbc = 0;
  for (int i = 0; i < npf; ++i) {
    if (s->ps_skip[i]) continue;
    const int o = s->ps_owner[i], nbc = s->ps_nb[i], fm = s->ps_fmesh[i];
    double d_on[3], kf[3];
    for (int q = 0; q < 3; ++q) d_on[q] = m->cell_centroid[3 * nbc + q] - m->cell_centroid[3 * o + q];
    const double *sf = &m->face_sf[3 * fm];
    kfactor_limited(sf, d_on, p->min_cos_theta, kf);
    if (kf[0] * kf[0] + kf[1] * kf[1] + kf[2] * kf[2] >= NONORTHO_SIN2_TOL * (sf[0] * sf[0] + sf[1] * sf[1] + sf[2] * sf[2])) {
      nact++;
      for (int q = 0; q < 3; ++q) s->k_int[3 * i + q] = kf[q];
      s->act_int[i] = 1;
    }
  }
  for (int ib = 0; ib < s->n_nper_bf; ++ib) {
    const int b = s->nper_bf[ib], fg = nint + b, o = m->face_owner[fg];
    if (bc_of_face(s, fg) != BC_OUTLET) continue;
    double d_on[3], kf[3];
    for (int q = 0; q < 3; ++q) d_on[q] = m->face_centre[3 * fg + q] - m->cell_centroid[3 * o + q];
    const double *sf = &m->face_sf[3 * fg];
    kfactor_limited(sf, d_on, p->min_cos_theta, kf);
    if (kf[0] * kf[0] + kf[1] * kf[1] + kf[2] * kf[2] >= NONORTHO_SIN2_TOL * (sf[0] * sf[0] + sf[1] * sf[1] + sf[2] * sf[2])) {
      nact++;
      for (int q = 0; q < 3; ++q) s->k_bnd[3 * ib + q] = kf[q];
      s->act_bnd[ib] = 1;
    }
  }
  for (int q = 0; q < s->npatch; ++q) {
    const int bc = m->patch_bc[q];
    if (!(bc == BC_OUTLET || bc == BC_PERIODIC || bc == BC_WALL_NOSLIP || bc == BC_WALL_VELOCITY || bc == BC_WALL_MODEL ||
          bc == BC_SLIP || bc == BC_SYMMETRY))
      ngradbc++;
  }
  s->skip_nonortho = (nact == 0);
  s->skip_grad_p = s->skip_nonortho && (ngradbc == 0);
  if (s->skip_nonortho) {
    memset(s->k_int, 0, sizeof(double) * 3 * (size_t)npf);
    memset(s->act_int, 0, npf);
    memset(s->k_bnd, 0, sizeof(double) * 3 * (size_t)s->n_nper_bf);
    memset(s->act_bnd, 0, s->n_nper_bf);
  }
  s->n_corr = p->n_nonortho_corr;
  if (s->n_corr <= 0) s->n_corr = s->skip_nonortho ? 1 : 2;
  /* build_reconstruct_operator (:3410-3462) */
  {
    double *T = dz(9 * (size_t)nc);
    for (int i = 0; i < npf; ++i) {
      if (s->ps_skip[i]) continue;
      const int o = s->ps_owner[i], nbc = s->ps_nb[i], fm = s->ps_fmesh[i];
      const double a = m->face_area[fm];
      const double *n = &m->face_normal[3 * fm];
      for (int jj = 0; jj < 3; ++jj)
        for (int ii = 0; ii < 3; ++ii) {
          T[9 * (size_t)o + 3 * ii + jj] = T[9 * (size_t)o + 3 * ii + jj] + a * n[ii] * n[jj];
          T[9 * (size_t)nbc + 3 * ii + jj] = T[9 * (size_t)nbc + 3 * ii + jj] + a * n[ii] * n[jj];
        }
    }
    for (int ib = 0; ib < s->n_nper_bf; ++ib) {
      const int fg = nint + s->nper_bf[ib], o = m->face_owner[fg];
      const double a = m->face_area[fg];
      const double *n = &m->face_normal[3 * fg];
      for (int jj = 0; jj < 3; ++jj)
        for (int ii = 0; ii < 3; ++ii) T[9 * (size_t)o + 3 * ii + jj] = T[9 * (size_t)o + 3 * ii + jj] + a * n[ii] * n[jj];
    }
    for (int c = 0; c < nc; ++c) invert_sym3(&T[9 * (size_t)c], &s->Tinv[9 * (size_t)c]);
    free(T);
  }
  if (p->sgs != OM_SGS_NONE) build_filter_tensor(s);
  /* wall-model sampling cache (find_sample_cell with h_wm <= 0, mod_wall_model.f90:124-173) */
  for (int b = 0; b < nb; ++b) {
    const int fg = nint + b, o = m->face_owner[fg];
    s->ws_cell[b] = o;
    const double *n = &m->face_normal[3 * fg];
    double d[3];
    for (int i = 0; i < 3; ++i) d[i] = m->cell_centroid[3 * o + i] - m->face_centre[3 * fg + i];
    s->ws_y[b] = fabs(-(d[0] * n[0]) - d[1] * n[1] - d[2] * n[2]) + SMALL;
  }
  s->q_w = dz(nb);
  s->L_bnd = dz(nb);
  s->L_cfg = p->L_obukhov;
  /* the log-law wall model samples h_wm above the wall (find_sample_cell); MOST keeps the
   * wall-adjacent cell (src/abl/mod_abl_wall.f90:101-103) */
  if (p->wall_model == OM_WM_LOGLAW && p->wall_model_height > 0.0)
    for (int b = 0; b < nb; ++b)
      if (bc_of_face(s, nint + b) == BC_WALL_MODEL) find_sample_cell(s, b, p->wall_model_height, &s->ws_cell[b], &s->ws_y[b]);
  s->wm_filt = dz(3 * (size_t)nb);
  /* convective outlet / backflow traction (ensure_bc_config_soa, src/bc/mod_bc.f90:119-148) */
  s->out_conv = iz(s->npatch);
  s->out_beta = dz(s->npatch);
  s->out_uc = dz(s->npatch);
  s->out_uref = dz(s->npatch);
  s->out_gprev = dz(3 * (size_t)nb);
  for (int q = 0; q < s->npatch; ++q) {
    s->out_conv[q] = p->patch_outlet_conv ? (p->patch_outlet_conv[q] != 0) : 0;
    s->out_beta[q] = p->patch_backflow_beta ? p->patch_backflow_beta[q] : 0.0;
    if (m->patch_bc[q] == BC_OUTLET && (s->out_conv[q] || s->out_beta[q] > 0.0)) s->out_active = 1;
  }
  /* patch_roughness (src/abl/mod_abl_wall.f90:249-262, input parser :524-529): a positive patch
   * z0_m and a non-zero patch z0_h override the global ones; a non-positive z0_h is z0_m e^-kB^-1 */
  s->z0m_p = dz(s->npatch);
  s->z0h_p = dz(s->npatch);
  for (int q = 0; q < s->npatch; ++q) {
    double zm = (p->patch_z0_m && p->patch_z0_m[q] > 0.0) ? p->patch_z0_m[q] : p->most_z0_m;
    if (zm <= 0.0) zm = p->most_z0_m;
    zm = fmax(zm, SMALL);
    double zh = (p->patch_z0_h && p->patch_z0_h[q] != 0.0) ? p->patch_z0_h[q] : p->most_z0_h;
    if (zh <= 0.0) zh = zm * exp(-KB_INV);
    s->z0m_p[q] = zm;
    s->z0h_p[q] = fmax(zh, SMALL);
  }
  if (p->wall_model == OM_WM_MOST) {
    /* the first cell must lie above each patch's roughness layer (mod_fv_solver.f90:736-777) */
    for (int b = 0; b < nb; ++b)
      if (bc_of_face(s, nint + b) == BC_WALL_MODEL && s->ws_y[b] / s->z0m_p[m->face_patch[nint + b]] <= 1.0) {
        om_free(s);
        return NULL;
      }
  }
  build_cell_projection(s);
  om_apply_all_bc(s);
  return s;
}

void om_apply_all_bc(om_solver *s) {
  apply_velocity_bc(s, 0, s->grad_U);
  apply_pressure_bc(s, s->P, s->p_bnd, 0, s->skip_grad_p ? NULL : s->gp, 1.0, s->outlet_pb);
  apply_temperature_bc(s, 0, s->grad_T);
  for (int k = 0; k < s->p.n_scalars; ++k) apply_scalar_bc(s, k, 0);
}

void om_free(om_solver *s) {
  if (!s) return;
  free(s->q_w);
  free(s->L_bnd);
  double *dd[] = {s->U, s->P, s->p_old, s->nu_sgs, s->U_bnd, s->p_bnd, s->phi_prev, s->phi_prev_bnd, s->tau_w,
                  s->grad_U, s->gU_prev, s->face_flux, s->U_face, s->inv_vol, s->H, s->H0, s->conv, s->visc,
                  s->rhs_p, s->div_w, s->phi_it, s->gp, s->gp_rc, s->gp_cell, s->p_tot, s->p_tot_bnd, s->gp_prev,
                  s->Tinv, s->outlet_pb, s->Gf, s->ws_y, s->div_diag, s->ps_coeff, s->ps_diag, s->dir_vals,
                  s->k_int, s->k_bnd, s->cg_b, s->dt_flux, s->dt_diff, s->T, s->T_bnd, s->grad_T, s->gT_prev,
                  s->H_T, s->H_T0, s->conv_T, s->visc_T};
  for (size_t i = 0; i < sizeof(dd) / sizeof(dd[0]); ++i) free(dd[i]);
  int *ii[] = {s->per_nb, s->per_bf, s->nper_bf, s->ws_cell, s->ps_owner, s->ps_nb, s->ps_fmesh, s->ps_skip, s->dir_rows};
  for (size_t i = 0; i < sizeof(ii) / sizeof(ii[0]); ++i) free(ii[i]);
  free(s->wm_filt); free(s->out_conv); free(s->out_beta); free(s->out_uc); free(s->out_uref); free(s->out_gprev);
  free(s->z0m_p); free(s->z0h_p);
  free(s->built);
  free(s->act_int);
  free(s->act_bnd);
  free(s->cp_own); free(s->cp_nbr); free(s->cp_bcell); free(s->cp_sf); free(s->cp_gc); free(s->cp_bsf);
  for (int q = 0; q < OM_MAX_SCALARS; ++q) {
    free(s->C[q]); free(s->C_bnd[q]); free(s->grad_C[q]); free(s->H_C[q]); free(s->H_C0[q]);
  }
  free(s->gC_prev);
  free(s);
}

double *om_array(om_solver *s, const char *name, int64_t *len) {
  const int64_t nc = s->nc, nb = s->nb, nf = s->nf;
  struct {
    const char *n;
    double *p;
    int64_t l;
  } t[] = {{"U", s->U, 3 * nc}, {"p", s->P, nc}, {"p_bnd", s->p_bnd, nb}, {"U_bnd", s->U_bnd, 3 * nb},
           {"nu_sgs", s->nu_sgs, nc}, {"grad_U", s->grad_U, 9 * nc}, {"U_face", s->U_face, nf},
           {"face_flux", s->face_flux, nf}, {"tau_w_bnd", s->tau_w, 3 * nb}, {"w_gp", s->gp, 3 * nc},
           {"rhs_p", s->rhs_p, nc}, {"H", s->H, 3 * nc}, {"Gf", s->Gf, 9 * nc}, {"Tinv", s->Tinv, 9 * nc},
           {"ps_coeff", s->ps_coeff, s->npf}, {"ps_diag", s->ps_diag, nc}, {"T", s->T, nc}, {"T_bnd", s->T_bnd, nb},
           {"grad_T", s->grad_T, 3 * nc}, {"q_w_bnd", s->q_w, nb}, {"L_bnd", s->L_bnd, nb},
           {"outlet_pb", s->outlet_pb, nb}};
  for (size_t i = 0; i < sizeof(t) / sizeof(t[0]); ++i)
    if (strcmp(t[i].n, name) == 0) {
      if (len) *len = t[i].l;
      return t[i].p;
    }
  for (int k = 0; k < s->p.n_scalars; ++k) {
    char a[32], b[32], g[32];
    snprintf(a, sizeof a, "C%d", k);
    snprintf(b, sizeof b, "C%d_bnd", k);
    snprintf(g, sizeof g, "grad_C%d", k);
    if (strcmp(name, a) == 0) { if (len) *len = nc; return s->C[k]; }
    if (strcmp(name, b) == 0) { if (len) *len = nb; return s->C_bnd[k]; }
    if (strcmp(name, g) == 0) { if (len) *len = 3 * nc; return s->grad_C[k]; }
  }
  if (len) *len = 0;
  return NULL;
}

int32_t *om_int_array(om_solver *s, const char *name, int64_t *len) {
  if (strcmp(name, "ps_owner") == 0) { *len = s->npf; return s->ps_owner; }
  if (strcmp(name, "ps_nb") == 0) { *len = s->npf; return s->ps_nb; }
  if (strcmp(name, "ps_skip") == 0) { *len = s->npf; return s->ps_skip; }
  *len = 0;
  return NULL;
}

void om_kfactor_limited(const double sf[3], const double d_on[3], double min_cos, double k[3]) {
  kfactor_limited(sf, d_on, min_cos, k);
}

double om_time(const om_solver *s) { return s->time; }

void om_proj_counters(const om_solver *s, int64_t out[2]) {
  out[0] = s->n_proj;
  out[1] = s->n_proj_its;
}

void om_counters(const om_solver *s, int64_t out[3]) {
  out[0] = s->n_solves;
  out[1] = s->n_iters;
  out[2] = s->last_iters;
}

int om_nonortho_info(const om_solver *s, int out[3]) {
  out[0] = s->skip_nonortho;
  out[1] = s->skip_grad_p;
  out[2] = s->n_corr;
  return 0;
}
