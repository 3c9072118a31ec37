// RK3W fractional-step driver over the hot-path kernels, templated on the execution policy.
//
// Set-up (host, once): builds the two per-cell face lists (mesh-face order for the transport terms,
// pressure-operator order for the projection terms), the LS stencil in ELL form, the pressure
// operator with its periodic edges, the non-orthogonal k-vectors, the reconstruction operator and
// the Vreman filter tensor, exactly as solver_init / pressure_solver_init build them
// (src/solver/mod_fv_solver.f90:302-507, src/solver/mod_pressure_solver.f90:100-452,
// src/sgs/mod_vreman.f90:168-264), and copies everything to the execution space once.
//
// Stepping: advance() follows solver_advance + advance_rk3w (src/solver/mod_fv_solver.f90:920-957,
// 1180-1614) call for call; every array stays resident between steps (to_device semantics,
// :3286-3316); the host reads back only the scalars that steer the step (CFL dt, corrector exit,
// CG convergence).
#pragma once

#include <cstdio>
#include <cstdlib>

#include <algorithm>
#include <cfloat>
#include <string>
#include <memory>
#include <vector>

#include "amg_impl.h"
#include "cg_impl.h"
#include "dopamine_hip.h"
#include "kernels.h"

namespace dm {

struct SolverScalars {
  double divmax, ub1, ub2, atol, pmean, shift, cfl, dif;
  double st[12];
  double vmin, vtot, gdt, pad;
  double pj[4];  // exact cell projection: reductions of the current iteration
  // exact cell projection: the flexible PCG's scalars, kept and updated on the device
  double pj_rz, pj_en_tot, pj_en_new, pj_bn, pj_alpha, pj_beta;
  int pj_done, pj_pad;
};

namespace sk {
struct DivFin {
  SolverScalars* s;
  DM_HD void operator()(const Red<1>& t) const { s->divmax = t.v[0]; }
};
template <int K>
struct ProjFin {  // exact cell projection: store K reduction results from pj[at]
  SolverScalars* s;
  int at;
  DM_HD void operator()(const Red<K>& t) const {
    for (int i = 0; i < K; ++i) s->pj[at + i] = t.v[i];
  }
};
// exact cell projection: b = B u + bconst, |b|_inf (the relative test's scale)
struct ProjBnFin {
  SolverScalars* s;
  DM_HD void operator()(const Red<1>& t) const {
    s->pj[0] = t.v[0];
    s->pj_bn = t.v[0];
    s->pj_done = 0;
    s->pj_en_tot = 0.0;
  }
};
// r.z of the first direction
struct ProjRzFin {
  SolverScalars* s;
  DM_HD void operator()(const Red<1>& t) const { s->pj_rz = t.v[0]; }
};
// p.q: alpha = (r.z) / (p.q) and the energy norm of the step; a vanishing p.q ends the loop
struct ProjAlphaFin {
  SolverScalars* s;
  DM_HD void operator()(const Red<1>& t) const {
    const double pq = t.v[0];
    if (fabs(pq) <= DBL_MIN) {
      s->pj_done = 1;
      return;
    }
    s->pj_alpha = s->pj_rz / pq;
    s->pj_en_new = s->pj_alpha * s->pj_alpha * pq;
    s->pj_en_tot = s->pj_en_tot + s->pj_en_new;
  }
};
// max|r| after the update: stop at 1e-12 |b|_inf, or (from the second iteration) when the step's
// energy falls below tol^2 of the total
struct ProjStopFin {
  SolverScalars* s;
  int it;
  double tol;
  DM_HD void operator()(const Red<1>& t) const {
    if (t.v[0] <= 1.0e-12 * s->pj_bn || (it >= 2 && s->pj_en_new <= tol * tol * s->pj_en_tot)) s->pj_done = 1;
  }
};
// flexible beta = max(0, (r - r_old).z / (r.z)_old) and the new r.z
struct ProjBetaFin {
  SolverScalars* s;
  DM_HD void operator()(const Red<2>& t) const {
    s->pj_beta = fmax(0.0, t.v[0] / s->pj_rz);
    s->pj_rz = t.v[1];
  }
};
struct AtolFin {
  SolverScalars* s;
  double gdt;
  DM_HD void operator()(const Red<2>& t) const {
    s->ub1 = t.v[0];
    s->ub2 = t.v[1];
    const double us = t.v[0] / fmax(t.v[1], SMALL);
    s->atol = fmax(1.0e-12 * fmax(us, 1.0) * s->vmin / gdt, 1.0e-15);
  }
};
struct AnchorFin {
  SolverScalars* s;
  int apply;
  DM_HD void operator()(const Red<1>& t) const {
    s->pmean = t.v[0];
    s->shift = (apply && s->vtot > 0.0) ? t.v[0] / s->vtot : 0.0;
  }
};
struct CflFin {
  SolverScalars* s;
  DM_HD void operator()(const Red<2>& t) const {
    s->cfl = t.v[0];
    s->dif = t.v[1];
  }
};
struct StatsFin7 {
  SolverScalars* s;
  DM_HD void operator()(const Red<7>& t) const {
    for (int k = 0; k < 7; ++k) s->st[k] = t.v[k];
  }
};
struct StatsFinMax {
  SolverScalars* s;
  DM_HD void operator()(const Red<1>& t) const { s->st[7] = t.v[0]; }
};
struct StatsFinWall {
  SolverScalars* s;
  DM_HD void operator()(const Red<2>& t) const {
    s->st[8] = t.v[0];
    s->st[9] = t.v[1];
  }
};
struct StatsFinDiv {
  SolverScalars* s;
  int slot;
  DM_HD void operator()(const Red<1>& t) const { s->st[slot] = t.v[0]; }
};
struct Copy {
  double* d;
  const double* s;
  DM_HD void operator()(int i) const { d[i] = s[i]; }
};
}  // namespace sk

static const double RK3_A[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
static const double RK3_B[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0};
static const double RK3_G[3] = {8.0 / 15.0, 2.0 / 15.0, 1.0 / 3.0};
static const double RK3_C[3] = {0.0, 8.0 / 15.0, 2.0 / 3.0};

// nonortho_kfactor_limited (src/numerics/mod_nonortho.f90:65-98)
inline void kfactor_limited(const double sf[3], const double d_on[3], double min_cos, double k[3]) {
  const double d_on_len = std::sqrt(d_on[0] * d_on[0] + d_on[1] * d_on[1] + d_on[2] * d_on[2]);
  if (d_on_len < 1.0e-30) {
    k[0] = k[1] = k[2] = 0.0;
    return;
  }
  const double inv_len = 1.0 / d_on_len;
  const double sdd = (sf[0] * d_on[0] + sf[1] * d_on[1] + sf[2] * d_on[2]) * inv_len;
  double d_len = sdd * d_on_len;
  const double sf_sq = sf[0] * sf[0] + sf[1] * sf[1] + sf[2] * sf[2];
  if (min_cos > 0.0) d_len = std::fmax(d_len, min_cos * std::sqrt(sf_sq) * d_on_len);
  if (d_len > 1.0e-30)
    for (int i = 0; i < 3; ++i) k[i] = sf[i] - (sf_sq / d_len) * d_on[i];
  else
    for (int i = 0; i < 3; ++i) k[i] = sf[i] - (sdd * inv_len) * d_on[i];
}

template <class X>
class SolverT {
 public:
  X x;
  DevMesh M{};
  dopamine_params prm{};
  int nc = 0, nf = 0, nint = 0, nb = 0, npf = 0, npatch = 0;
  int has_skipped = 0;
  int skip_nonortho = 0, skip_grad_p = 0, n_corr = 1, diag_interface = 0, n_outlet = 0, n_wm = 0, n_dir = 0;
  double time = 0.0;
  int u_face_ready = 0;
  int64_t n_solves = 0, n_iters = 0, last_iters = 0, n_sweeps = 0, n_retries = 0;
  std::string err;  // text of the last failed step (empty after a successful one)
  // exact cell projection (cell_projection = 1): work vectors, projections and iterations so far
  int cproj = 0;
  double *pj_b = nullptr, *pj_x = nullptr, *pj_r = nullptr, *pj_z = nullptr, *pj_p = nullptr, *pj_q = nullptr,
         *pj_ro = nullptr, *pj_bc = nullptr, *pj_v = nullptr;
  int64_t n_proj = 0, n_proj_its = 0;

  // fields (execution space)
  double *U = nullptr, *U2 = nullptr, *P = nullptr, *p_old = nullptr, *phi_prev = nullptr, *p_bnd = nullptr;
  double *phi_prev_bnd = nullptr, *U_bnd = nullptr, *nu_sgs = nullptr, *grad = nullptr, *gprev = nullptr;
  double *gp = nullptr, *gp_rc = nullptr, *gp_prev = nullptr, *H = nullptr, *H0 = nullptr, *rhs = nullptr;
  double *ff = nullptr, *UF = nullptr, *tau = nullptr, *outlet_pb = nullptr, *patchP = nullptr, *cg_b = nullptr;
  double *phi_it = nullptr;
  SolverScalars* sc = nullptr;
  EllOp A;
  EllOp32 A32;
  double* phiF = nullptr;  // reconstruction schemes: 3 face values per face
  // temperature (Boussinesq)
  int bous = 0;
  double *T = nullptr, *T2 = nullptr, *T_bnd = nullptr, *gT = nullptr, *gT_prev = nullptr, *HT = nullptr,
         *HT0 = nullptr, *phiT = nullptr, *patchT = nullptr;
  // passive scalars: per scalar field, predictor target, boundary values, gradient, H and H_prev;
  // shared face values and gradient scratch; per-patch BC arrays
  int ns = 0;
  double *Cs[4] = {}, *Cs2[4] = {}, *Cb[4] = {}, *gC[4] = {}, *HC[4] = {}, *HC0[4] = {};
  double *phiC = nullptr, *gCprev = nullptr;
  double* phiCk[4] = {};  // each scalar's face values, for the fused transport of T and the scalars
  double* faceT = nullptr;  // the fused transport's interior-face fluxes, 2 per field and face
  TvdCtx* tvd_ctx = nullptr;  // the fused transport's TVD face contexts (both flux signs)
  int* s_mode = nullptr;
  double *s_val = nullptr, *s_flux = nullptr, *s_om = nullptr, *s_ph = nullptr;
  // MOST wall model: heat flux and Obukhov length per boundary face, the configuration's L
  double *q_w = nullptr, *L_bnd = nullptr, *L_cfg = nullptr;
  // convective outlet / backflow traction: per patch (U_c, U_ref) of the last stage, per boundary
  // face the previous stage's tendency; the outlet patches that need the stage statistics
  double *out_sp = nullptr, *out_gprev = nullptr;
  int out_ready = 0;
  std::vector<int> out_q, out_start, out_n;
  // log-law wall model time filter: filtered velocity per boundary face, mode (0 off, 1 fixed,
  // 2 adaptive), whether it is seeded
  double* wm_filt = nullptr;
  int wm_fmode = 0, wm_filt_ready = 0;
  // dynamic Smagorinsky work arrays: packed and filtered values, filtered velocity and its
  // gradient, L.M / M.M before and after the second filter
  double *dyn_fb = nullptr, *dyn_fh = nullptr, *dyn_uh = nullptr, *dyn_gh = nullptr, *dyn_lm = nullptr,
         *dyn_lmh = nullptr;
  int n_amdw = 0;
  int *amdw_cell = nullptr, *amdw_off = nullptr, *amdw_face = nullptr;
  int* tbc_cell = nullptr;  // owners of the boundary faces whose temperature extrapolates with grad T
  int n_tbc = 0;
  // relative residual reduction of one FP32 sweep (DOPAMINE_MIXED_TAU overrides, for tuning)
  // Multigrid set-up: AmgCfg's defaults; DOPAMINE_AMG_THETA / _THETA0 / _ALPHA / _COARSE override
  // the strength thresholds (all levels / the first), the coarse-correction scaling and the
  // densely solved size.
  static AmgCfg amg_cfg() {
    AmgCfg c;
    if (const char* e = std::getenv("DOPAMINE_AMG_THETA")) c.theta = std::atof(e);
    if (const char* e = std::getenv("DOPAMINE_AMG_THETA0")) c.theta0 = std::atof(e);
    if (const char* e = std::getenv("DOPAMINE_AMG_ALPHA")) c.alpha = std::atof(e);
    if (const char* e = std::getenv("DOPAMINE_AMG_ALPHA_AGGR")) c.alpha_aggr = std::atof(e);
    if (const char* e = std::getenv("DOPAMINE_AMG_OMEGA")) c.omega = std::atof(e);
    if (const char* e = std::getenv("DOPAMINE_AMG_COARSE")) c.coarse_max = std::max(1, std::atoi(e));
    return c;
  }
  static double mixed_tau() {
    const char* e = std::getenv("DOPAMINE_MIXED_TAU");
    const double t = e ? std::atof(e) : 1.0e-4;
    return (t > 0.0 && t < 1.0) ? t : 1.0e-4;
  }
  CgT<X> cg;
  std::unique_ptr<PressurePolicy> policy;  // learned preconditioner (device only); null: CG passes
  std::unique_ptr<AmgPc<X>> amg;           // aggregation multigrid (precond 2); null: none
  static constexpr int PRECOND_MULTIGRID = 2;
  std::vector<void*> owned;
  std::vector<int> dir_rows_h;
  std::vector<double> dir_vals_h;
  int* dir_rows = nullptr;
  double* dir_vals = nullptr;

  template <class T>
  T* up(const std::vector<T>& v) {
    T* p = x.template alloc<T>(v.size());
    x.upload(p, v.data(), v.size());
    owned.push_back(p);
    return p;
  }
  template <class T>
  T* fresh(size_t n) {
    T* p = x.template alloc<T>(n);
    owned.push_back(p);
    return p;
  }

  ~SolverT() {
    policy.reset();
    cg.destroy();
    for (void* p : owned) x.release(p);
  }

  void create(const dopamine_mesh_desc& d, const dopamine_params& p) {
    prm = p;
    nc = d.n_cells;
    nf = d.n_faces;
    nint = d.n_int_faces;
    nb = d.n_bnd_faces;
    npatch = d.n_patches;
    validate(d);
    std::vector<double> patchP_h(std::max(1, npatch), 0.0), patchU_h(3 * size_t(std::max(1, npatch)), 0.0);
    for (int q = 0; q < npatch; ++q) {
      patchP_h[q] = p.patch_p_val ? p.patch_p_val[q] : 0.0;
      for (int k = 0; k < 3; ++k) patchU_h[3 * q + k] = p.patch_U_val ? p.patch_U_val[3 * q + k] : 0.0;
    }
    // ---- boundary faces
    std::vector<int> bbc(std::max(1, nb)), bpatch(std::max(1, nb)), bown(std::max(1, nb)), pnb(std::max(1, nb), -1),
        pair(std::max(1, nb), -1);
    std::vector<double> pgc(std::max(1, nb)), pdcf(std::max(1, nb)), cb(std::max(1, nb), 0.0), wsy(std::max(1, nb), 0.0);
    std::vector<int> wsc(std::max(1, nb));
    for (int b = 0; b < nb; ++b) {
      const int f = nint + b, pid = d.face_patch[f];
      bpatch[b] = (pid >= 0 && pid < npatch) ? pid : 0;
      bbc[b] = (pid >= 0 && pid < npatch) ? d.patch_bc[pid] : BC_ZEROGRADIENT;
      bown[b] = d.face_owner[f];
      pair[b] = d.periodic_pair[b];
      pgc[b] = d.periodic_gc[b];
      pdcf[b] = d.periodic_pair_dcf[b];
      // periodic partner through the pair (resolve_periodic_nb, src/bc/mod_bc.f90:822-859)
      if (pair[b] >= 0) {
        const int nbc = d.face_owner[nint + pair[b]];
        if (nbc >= 0 && nbc != bown[b]) pnb[b] = nbc;
      }
      // Dirichlet-face coefficient A / |(x_f - x_o).n| (src/solver/mod_fv_solver.f90:2851-2857)
      const int o = bown[b];
      const double* n = &d.face_normal[3 * size_t(f)];
      double dv[3];
      for (int i = 0; i < 3; ++i) dv[i] = d.face_centre[3 * size_t(f) + i] - d.cell_centroid[3 * size_t(o) + i];
      const double sn = std::fabs(dv[0] * n[0] + dv[1] * n[1] + dv[2] * n[2]);
      cb[b] = d.face_area[f] / std::fmax(sn, SMALL);
      // wall-model sampling cell and height with h_wm <= 0 (find_sample_cell, mod_wall_model.f90:124-173)
      wsc[b] = o;
      double dd[3];
      for (int i = 0; i < 3; ++i) dd[i] = d.cell_centroid[3 * size_t(o) + i] - d.face_centre[3 * size_t(f) + i];
      wsy[b] = std::fabs(-(dd[0] * n[0]) - dd[1] * n[1] - dd[2] * n[2]) + SMALL;
      if (bbc[b] == BC_OUTLET) n_outlet++;
    }
    for (int q = 0; q < npatch; ++q)
      if (d.patch_bc[q] == BC_WALL_MODEL) n_wm++;
    // the log-law wall model samples wall_model_height above the wall (find_sample_cell,
    // src/sgs/mod_wall_model.f90:126-173); MOST keeps the wall-adjacent cell
    if (p.wall_model == 0 && p.wall_model_height > 0.0)
      for (int b = 0; b < nb; ++b)
        if (bbc[b] == BC_WALL_MODEL) sample_walk(d, b, p.wall_model_height, &wsc[b], &wsy[b]);
    // per-patch roughness (patch_roughness, src/abl/mod_abl_wall.f90:249-262): a positive patch
    // z0_m and a non-zero patch z0_h override the global values; a non-positive z0_h is z0_m e^-2
    std::vector<double> z0m_h(std::max(1, npatch)), z0h_h(std::max(1, npatch)), beta_h(std::max(1, npatch), 0.0);
    std::vector<int> conv_h(std::max(1, npatch), 0);
    for (int q = 0; q < npatch; ++q) {
      double zm = (p.patch_z0_m && p.patch_z0_m[q] > 0.0) ? p.patch_z0_m[q] : p.most_z0_m;
      if (zm <= 0.0) zm = p.most_z0_m;
      zm = std::fmax(zm, SMALL);
      double zh = (p.patch_z0_h && p.patch_z0_h[q] != 0.0) ? p.patch_z0_h[q] : p.most_z0_h;
      if (zh <= 0.0) zh = zm * std::exp(-2.0);
      z0m_h[q] = zm;
      z0h_h[q] = std::fmax(zh, SMALL);
      conv_h[q] = p.patch_outlet_conv ? (p.patch_outlet_conv[q] != 0) : 0;
      beta_h[q] = p.patch_backflow_beta ? p.patch_backflow_beta[q] : 0.0;
      if (d.patch_bc[q] == BC_OUTLET && (conv_h[q] || beta_h[q] > 0.0) && d.patch_nfaces[q] > 0) {
        out_q.push_back(q);
        out_start.push_back(d.patch_start[q]);
        out_n.push_back(d.patch_nfaces[q]);
      }
    }
    if (p.wall_model == 2) {
      // the first cell must lie above each patch's roughness layer (mod_fv_solver.f90:736-777)
      for (int b = 0; b < nb; ++b)
        if (bbc[b] == BC_WALL_MODEL && wsy[b] / z0m_h[bpatch[b]] <= 1.0)
          throw std::runtime_error("wall model: first cell inside the roughness layer (y_w / z0_m <= 1)");
    }
    wm_fmode = p.wall_model_filter_tf_fac > SMALL ? 2 : (p.wall_model_filter_time > SMALL ? 1 : 0);
    // refinement-interface faces (build_interface_diagnostics, mod_fv_solver.f90:685-718): more
    // correctors and a looser stagnation test
    for (int f = 0; f < nint; ++f)
      if (d.face_skew[f] > 0.15) diag_interface = 1;

    // ---- mesh-face list (cell_faces CSR order)
    std::vector<int> nA(nc);
    int wA = 0;
    for (int c = 0; c < nc; ++c) {
      nA[c] = d.cell_face_offset[c + 1] - d.cell_face_offset[c];
      wA = std::max(wA, nA[c]);
    }
    std::vector<int> aface(size_t(wA) * nc, 0), aoth(size_t(wA) * nc, 0);
    for (int c = 0; c < nc; ++c)
      for (int k = 0; k < nA[c]; ++k) {
        const int f = d.cell_faces[d.cell_face_offset[c] + k];
        const int o = d.face_owner[f], nbc = d.face_neighbour[f];
        const int side = (nbc >= 0 && nbc == c && o != c) ? 1 : 0;
        aface[size_t(k) * nc + c] = 2 * f + side;
        aoth[size_t(k) * nc + c] = (nbc >= 0) ? (side ? o : nbc) : -1 - (f - nint);
      }

    // ---- pressure operator (pressure_solver_init)
    std::vector<int> pown, pnbr, pfm;
    std::vector<double> pcoef, diag(nc, 0.0);
    pown.reserve(nint + nb);
    for (int f = 0; f < nint; ++f) {
      pown.push_back(d.face_owner[f]);
      pnbr.push_back(d.face_neighbour[f]);
      pfm.push_back(f);
    }
    pcoef.assign(nint, 0.0);
    for (int q = 0; q < npatch; ++q) {
      if (d.patch_bc[q] != BC_PERIODIC) continue;
      for (int k = 0; k < d.patch_nfaces[q]; ++k) {
        const int b = d.patch_start[q] + k;
        if (pair[b] < 0) continue;
        const int fg = nint + b, o = bown[b], nbc = d.face_owner[nint + pair[b]];
        if (nbc == o || o >= nbc) continue;  // canonical side: lower cell id owns the edge
        const int pf = nint + pair[b];
        double pd = d.periodic_pair_dcf[b];
        if (pd <= SMALL) pd = d.face_dcf[pf];
        if (pd <= SMALL) pd = d.face_dcf[fg];
        const double dper = d.face_dcf[fg] + pd;
        const double area = 0.5 * (d.face_area[fg] + d.face_area[pf]);
        const double coeff = area / (dper + SMALL);
        pown.push_back(o);
        pnbr.push_back(nbc);
        pfm.push_back(fg);
        pcoef.push_back(coeff);
        diag[o] = diag[o] + coeff;
        diag[nbc] = diag[nbc] + coeff;
      }
    }
    npf = int(pown.size());
    for (int f = 0; f < nint; ++f) {
      const int o = d.face_owner[f], nbc = d.face_neighbour[f];
      double don[3];
      for (int i = 0; i < 3; ++i) don[i] = d.cell_centroid[3 * size_t(nbc) + i] - d.cell_centroid[3 * size_t(o) + i];
      const double* n = &d.face_normal[3 * size_t(f)];
      double dort = don[0] * n[0] + don[1] * n[1] + don[2] * n[2];
      if (p.min_cos_theta > 0.0 && dort < p.min_cos_theta * d.face_dcf[f]) dort = p.min_cos_theta * d.face_dcf[f];
      const double coeff = (dort > SMALL) ? d.face_area[f] / dort : d.face_area[f] / (d.face_dcf[f] + SMALL);
      if (!std::isfinite(coeff) || coeff <= SMALL) {
        pown[f] = -1;  // skipped face
        has_skipped = 1;
        continue;
      }
      pcoef[f] = coeff;
      diag[o] = diag[o] + coeff;
      diag[nbc] = diag[nbc] + coeff;
    }
    for (int q = 0; q < npatch; ++q) {
      if (d.patch_bc[q] != BC_OUTLET) continue;
      for (int k = 0; k < d.patch_nfaces[q]; ++k) {
        const int b = d.patch_start[q] + k, fg = nint + b, o = bown[b];
        double dv[3];
        for (int i = 0; i < 3; ++i) dv[i] = d.face_centre[3 * size_t(fg) + i] - d.cell_centroid[3 * size_t(o) + i];
        const double* n = &d.face_normal[3 * size_t(fg)];
        double dp = std::fabs(dv[0] * n[0] + dv[1] * n[1] + dv[2] * n[2]);
        if (dp <= SMALL) dp = d.face_dcf[fg];
        if (p.min_cos_theta > 0.0) dp = std::fmax(dp, p.min_cos_theta * d.face_dcf[fg]);
        if (dp <= SMALL) continue;
        const double cbv = d.face_area[fg] / dp;
        if (!std::isfinite(cbv) || cbv <= SMALL) continue;
        diag[o] = diag[o] + cbv;
        dir_rows_h.push_back(o);
        dir_vals_h.push_back(cbv * patchP_h[q]);
      }
    }
    n_dir = int(dir_rows_h.size());
    for (int c = 0; c < nc; ++c)
      if (!std::isfinite(diag[c]) || diag[c] <= SMALL) diag[c] = 1.0;
    // backend repair (pressure_backend_petsc.c:304-309)
    for (int c = 0; c < nc; ++c)
      if (!std::isfinite(diag[c]) || std::fabs(diag[c]) < 1.0e-20) diag[c] = 1.0;

    // ---- interpolation weights per operator face
    std::vector<double> rcgc(std::max(1, npf)), xgc(std::max(1, npf));
    for (int i = 0; i < npf; ++i) {
      const int fm = pfm[i];
      if (fm < nint) {
        rcgc[i] = d.face_gc[fm];
        xgc[i] = d.face_gc[fm];
      } else {
        const int b = fm - nint;
        double g = 0.5;
        const double d_o = d.face_dcf[fm];
        double d_n = d_o;
        if (d.periodic_pair_dcf[b] > SMALL) d_n = d.periodic_pair_dcf[b];
        if (d_n <= SMALL && pair[b] >= 0) d_n = d.face_dcf[nint + pair[b]];
        const double dt = d_o + d_n;
        if (dt > SMALL) g = d_n / dt;
        rcgc[i] = g;
        xgc[i] = d.periodic_gc[b];
      }
    }

    // ---- non-orthogonal k-vectors (detect_nonortho_skip + build_nonortho_cache)
    std::vector<double> kint(3 * size_t(std::max(1, npf)), 0.0), kbnd(3 * size_t(std::max(1, nb)), 0.0);
    int nact = 0, ngradbc = 0;
    for (int i = 0; i < npf; ++i) {
      if (pown[i] < 0) continue;
      const int o = pown[i], nbc = pnbr[i], fm = pfm[i];
      double don[3], kf[3];
      for (int q = 0; q < 3; ++q) don[q] = d.cell_centroid[3 * size_t(nbc) + q] - d.cell_centroid[3 * size_t(o) + q];
      const double* sf = &d.face_sf[3 * size_t(fm)];
      kfactor_limited(sf, don, p.min_cos_theta, kf);
      if (kf[0] * kf[0] + kf[1] * kf[1] + kf[2] * kf[2] >= 1.0e-10 * (sf[0] * sf[0] + sf[1] * sf[1] + sf[2] * sf[2])) {
        nact++;
        for (int q = 0; q < 3; ++q) kint[3 * size_t(i) + q] = kf[q];
      }
    }
    for (int b = 0; b < nb; ++b) {
      if (bbc[b] != BC_OUTLET) continue;
      const int fg = nint + b, o = bown[b];
      double don[3], kf[3];
      for (int q = 0; q < 3; ++q) don[q] = d.face_centre[3 * size_t(fg) + q] - d.cell_centroid[3 * size_t(o) + q];
      const double* sf = &d.face_sf[3 * size_t(fg)];
      kfactor_limited(sf, don, p.min_cos_theta, kf);
      if (kf[0] * kf[0] + kf[1] * kf[1] + kf[2] * kf[2] >= 1.0e-10 * (sf[0] * sf[0] + sf[1] * sf[1] + sf[2] * sf[2])) {
        nact++;
        for (int q = 0; q < 3; ++q) kbnd[3 * size_t(b) + q] = kf[q];
      }
    }
    for (int q = 0; q < npatch; ++q) {
      const int bc = d.patch_bc[q];
      if (!(bc == BC_OUTLET || bc == BC_PERIODIC || bc == BC_WALL_NOSLIP || bc == BC_WALL_VELOCITY ||
            bc == BC_WALL_MODEL || bc == BC_SLIP || bc == BC_SYMMETRY))
        ngradbc++;
    }
    skip_nonortho = (nact == 0);
    skip_grad_p = skip_nonortho && (ngradbc == 0);
    if (skip_nonortho) {
      std::fill(kint.begin(), kint.end(), 0.0);
      std::fill(kbnd.begin(), kbnd.end(), 0.0);
    }
    n_corr = p.n_nonortho_corr;
    if (n_corr <= 0) n_corr = skip_nonortho ? 1 : 2;

    // ---- projection-face list: operator faces in operator order, then non-periodic boundary faces
    std::vector<int> nB(nc, 0);
    for (int i = 0; i < npf; ++i)
      if (pown[i] >= 0) {
        nB[pown[i]]++;
        nB[pnbr[i]]++;
      }
    for (int b = 0; b < nb; ++b)
      if (bbc[b] != BC_PERIODIC) nB[bown[b]]++;
    int wB = 0;
    for (int c = 0; c < nc; ++c) wB = std::max(wB, nB[c]);
    std::vector<int> bface(size_t(std::max(1, wB)) * nc, 0), both(size_t(std::max(1, wB)) * nc, 0),
        bidx(size_t(std::max(1, wB)) * nc, -1), fillB(nc, 0);
    for (int i = 0; i < npf; ++i) {
      if (pown[i] < 0) continue;
      const int o = pown[i], nbc = pnbr[i], fm = pfm[i];
      size_t s = size_t(fillB[o]++) * nc + o;
      bface[s] = 2 * fm;
      both[s] = nbc;
      bidx[s] = i;
      s = size_t(fillB[nbc]++) * nc + nbc;
      bface[s] = 2 * fm + 1;
      both[s] = o;
      bidx[s] = i;
    }
    for (int b = 0; b < nb; ++b) {
      if (bbc[b] == BC_PERIODIC) continue;
      const int o = bown[b];
      const size_t s = size_t(fillB[o]++) * nc + o;
      bface[s] = 2 * (nint + b);
      both[s] = -1 - b;
      bidx[s] = -1;
    }
    // ---- CG operator rows: the operator entries of the projection list, same order
    int wP = 0;
    std::vector<int> nP(nc, 0);
    for (int i = 0; i < npf; ++i)
      if (pown[i] >= 0 && std::isfinite(pcoef[i]) && pcoef[i] != 0.0) {
        nP[pown[i]]++;
        nP[pnbr[i]]++;
      }
    for (int c = 0; c < nc; ++c) wP = std::max(wP, nP[c]);
    std::vector<int> anbr(size_t(std::max(1, wP)) * nc, -1), fillP(nc, 0);
    std::vector<double> acoef(size_t(std::max(1, wP)) * nc, 0.0);
    for (int i = 0; i < npf; ++i) {
      if (pown[i] < 0 || !std::isfinite(pcoef[i]) || pcoef[i] == 0.0) continue;
      const int o = pown[i], nbc = pnbr[i];
      size_t s = size_t(fillP[o]++) * nc + o;
      anbr[s] = nbc;
      acoef[s] = pcoef[i];
      s = size_t(fillP[nbc]++) * nc + nbc;
      anbr[s] = o;
      acoef[s] = pcoef[i];
    }

    // ---- LS stencil in ELL form (boundary pseudo-neighbour b encoded as -1-b)
    std::vector<int> nL(nc);
    int wL = 0;
    for (int c = 0; c < nc; ++c) {
      nL[c] = d.ls_offset[c + 1] - d.ls_offset[c];
      wL = std::max(wL, nL[c]);
    }
    std::vector<int> lidx(size_t(std::max(1, wL)) * nc, 0);
    std::vector<double> lw(3 * size_t(std::max(1, wL)) * nc, 0.0);
    for (int c = 0; c < nc; ++c)
      for (int k = 0; k < nL[c]; ++k) {
        const int e = d.ls_offset[c] + k;
        const int cell = d.ls_cell[e];
        int enc;
        if (cell >= 0) {
          enc = cell;
        } else {
          const int f = d.ls_face[e];
          const int b = (f >= 0) ? f - nint : -1;
          enc = (b >= 0 && b < nb) ? -1 - b : c;  // invalid entry: zero difference
        }
        lidx[size_t(k) * nc + c] = enc;
        for (int i = 0; i < 3; ++i) lw[(3 * size_t(k) + i) * nc + c] = d.ls_weight[3 * size_t(e) + i];
      }

    // ---- reconstruction operator T^-1 (build_reconstruct_operator + invert_sym3)
    std::vector<double> Tinv(9 * size_t(nc), 0.0);
    {
      std::vector<double> T(9 * size_t(nc), 0.0);
      auto addnn = [&](int c, double a, const double* n) {
        for (int jj = 0; jj < 3; ++jj)
          for (int ii = 0; ii < 3; ++ii) T[9 * size_t(c) + 3 * ii + jj] = T[9 * size_t(c) + 3 * ii + jj] + a * n[ii] * n[jj];
      };
      for (int i = 0; i < npf; ++i) {
        if (pown[i] < 0) continue;
        const int fm = pfm[i];
        addnn(pown[i], d.face_area[fm], &d.face_normal[3 * size_t(fm)]);
        addnn(pnbr[i], d.face_area[fm], &d.face_normal[3 * size_t(fm)]);
      }
      for (int b = 0; b < nb; ++b)
        if (bbc[b] != BC_PERIODIC) addnn(bown[b], d.face_area[nint + b], &d.face_normal[3 * size_t(nint + b)]);
      for (int c = 0; c < nc; ++c) invert_sym3(&T[9 * size_t(c)], &Tinv[9 * size_t(c)]);
    }
    // ---- Vreman filter tensor (build_filter_tensor + smooth_filter_tensor)
    std::vector<double> Gf(9 * size_t(nc), 0.0);
    if (p.sgs != 0) build_filter_tensor(d, Gf);

    // ---- upload
    std::vector<double> iv(nc), wbl(nf, 1.0);
    for (int c = 0; c < nc; ++c) iv[c] = 1.0 / std::fmax(d.cell_volume[c], SMALL);
    for (int f = 0; f < nint; ++f) wbl[f] = refine_blend_weight(d.face_skew[f]);
    auto vec_d = [](const double* p0, size_t n) { return std::vector<double>(p0, p0 + n); };
    M.nc = nc; M.nf = nf; M.nint = nint; M.nb = nb; M.npf = npf; M.wA = wA; M.wB = wB; M.wL = wL;
    M.vol = up(vec_d(d.cell_volume, nc));
    M.iv = up(iv);
    M.ctr = up(vec_d(d.cell_centroid, 3 * size_t(nc)));
    M.nA = up(nA); M.aface = up(aface); M.aoth = up(aoth);
    M.nB = up(nB); M.bface = up(bface); M.both = up(both); M.bidx = up(bidx);
    M.nL = up(nL); M.lidx = up(lidx); M.lw = up(lw);
    M.sf = up(vec_d(d.face_sf, 3 * size_t(nf)));
    M.gc = up(vec_d(d.face_gc, nf));
    M.area = up(vec_d(d.face_area, nf));
    M.nrm = up(vec_d(d.face_normal, 3 * size_t(nf)));
    M.eps = up(vec_d(d.face_eps, 3 * size_t(nf)));
    M.wbl = up(wbl);
    M.fctr = up(vec_d(d.face_centre, 3 * size_t(nf)));
    M.dcf = up(vec_d(d.face_dcf, nf));
    {
      std::vector<int> fo(d.face_owner, d.face_owner + nf), fn(d.face_neighbour, d.face_neighbour + nf);
      M.fown = up(fo);
      M.fnb = up(fn);
    }
    M.bbc = up(bbc); M.bpatch = up(bpatch); M.bown = up(bown); M.pnb = up(pnb); M.pair = up(pair);
    M.pgc = up(pgc); M.pdcf = up(pdcf); M.cb = up(cb); M.kbnd = up(kbnd); M.wsy = up(wsy); M.wsc = up(wsc);
    M.pconv = up(conv_h); M.pbeta = up(beta_h); M.pz0m = up(z0m_h); M.pz0h = up(z0h_h);
    out_sp = up(std::vector<double>(2 * size_t(std::max(1, npatch)), 0.0));
    out_gprev = up(std::vector<double>(3 * size_t(std::max(1, nb)), 0.0));
    wm_filt = up(std::vector<double>(3 * size_t(std::max(1, nb)), 0.0));
    if (p.sgs == 4) {
      dyn_fb = fresh<double>(15 * size_t(nc));
      dyn_fh = fresh<double>(15 * size_t(nc));
      dyn_uh = fresh<double>(3 * size_t(nc));
      dyn_gh = fresh<double>(9 * size_t(nc));
      dyn_lm = fresh<double>(2 * size_t(nc));
      dyn_lmh = fresh<double>(2 * size_t(nc));
    }
    M.pcoef = up(pcoef.empty() ? std::vector<double>(1, 0.0) : pcoef);
    M.rcgc = up(rcgc); M.xgc = up(xgc); M.kint = up(kint);
    M.pown = up(pown.empty() ? std::vector<int>(1, -1) : pown);
    M.pnbr = up(pnbr.empty() ? std::vector<int>(1, -1) : pnbr);
    M.pfm = up(pfm.empty() ? std::vector<int>(1, 0) : pfm);
    M.Tinv = up(Tinv); M.Gf = up(Gf);
    M.patchU = up(patchU_h);
    patchP = up(patchP_h);
    A.n = nc;
    A.w = wP;
    A.nbr = up(anbr);
    A.coef = up(acoef);
    A.diag = up(diag);
    if (p.precision == 1) {  // FP32 copy of the operator for the mixed-precision sweeps
      std::vector<float> c32(acoef.begin(), acoef.end()), d32(diag.size());
      for (size_t i = 0; i < diag.size(); ++i) {
        double d = diag[i];
        if (!std::isfinite(d) || std::fabs(d) < 1.0e-20) d = 1.0;
        d32[i] = float(d);
      }
      A32.n = nc;
      A32.w = wP;
      A32.nbr = A.nbr;
      A32.coef = up(c32);
      A32.diag = up(d32);
      std::vector<uint32_t> pr;  // the compact copy the batched FP32 direction pass streams
      std::vector<float> oc;
      if (off16_enabled() && ell_pack32(anbr.data(), acoef.data(), nc, wP, pr, oc, A.obase)) {
        A.opair = up(pr);
        A.ocoef = up(oc);
      }
    }
    if (n_dir) {
      dir_rows = up(dir_rows_h);
      dir_vals = up(dir_vals_h);
    }
    {  // AMD wall cells: cells owning wall-model faces, their faces in boundary order
      std::vector<int> first(nc, -1), cells, offs(1, 0), faces;
      std::vector<std::vector<int>> per;
      for (int b = 0; b < nb; ++b) {
        if (bbc[b] != BC_WALL_MODEL) continue;
        const int o = bown[b];
        if (first[o] < 0) {
          first[o] = int(cells.size());
          cells.push_back(o);
          per.emplace_back();
        }
        per[first[o]].push_back(b);
      }
      for (auto& v : per) {
        faces.insert(faces.end(), v.begin(), v.end());
        offs.push_back(int(faces.size()));
      }
      n_amdw = int(cells.size());
      if (n_amdw) {
        amdw_cell = up(cells);
        amdw_off = up(offs);
        amdw_face = up(faces);
      }
    }
    // cells whose temperature gradient a boundary value reads (BcTemperature extrapolates on every
    // face that is not an inlet, a fixed temperature or periodic), each once
    {
      std::vector<char> mark(nc, 0);
      std::vector<int> cells;
      for (int b = 0; b < nb; ++b) {
        const int bc = bbc[b];
        if (bc == BC_INLET || bc == BC_FIXEDTEMP || bc == BC_PERIODIC || mark[bown[b]]) continue;
        mark[bown[b]] = 1;
        cells.push_back(bown[b]);
      }
      std::sort(cells.begin(), cells.end());
      n_tbc = int(cells.size());
      if (n_tbc) tbc_cell = up(cells);
    }
    // outlet pressure per boundary face (patch p_val; no backflow traction)
    std::vector<double> opb(std::max(1, nb), 0.0);
    for (int b = 0; b < nb; ++b) opb[b] = patchP_h[bpatch[b]];
    outlet_pb = up(opb);

    U = fresh<double>(3 * size_t(nc)); U2 = fresh<double>(3 * size_t(nc)); P = fresh<double>(nc);
    p_old = fresh<double>(nc); phi_prev = fresh<double>(nc); phi_it = fresh<double>(nc);
    p_bnd = fresh<double>(std::max(1, nb)); phi_prev_bnd = fresh<double>(std::max(1, nb));
    U_bnd = fresh<double>(3 * size_t(std::max(1, nb))); tau = fresh<double>(3 * size_t(std::max(1, nb)));
    nu_sgs = fresh<double>(nc); grad = fresh<double>(9 * size_t(nc)); gprev = fresh<double>(9 * size_t(nc));
    gp = fresh<double>(3 * size_t(nc)); gp_rc = fresh<double>(3 * size_t(nc)); gp_prev = fresh<double>(3 * size_t(nc));
    H = fresh<double>(3 * size_t(nc)); H0 = fresh<double>(3 * size_t(nc)); rhs = fresh<double>(nc);
    cg_b = fresh<double>(nc);
    ff = fresh<double>(nf); UF = fresh<double>(nf);
    sc = fresh<SolverScalars>(1);
    SolverScalars hs{};
    hs.vmin = 1e308;
    for (int c = 0; c < nc; ++c) {
      hs.vmin = std::fmin(hs.vmin, d.cell_volume[c]);
      hs.vtot = hs.vtot + d.cell_volume[c];
    }
    x.upload(sc, &hs, 1);
    // initialize_velocity_from_bc: U = first inlet's velocity (zero without an inlet)
    for (int q = 0; q < npatch; ++q)
      if (d.patch_bc[q] == BC_INLET) {
        std::vector<double> u0(3 * size_t(nc));
        for (int c = 0; c < nc; ++c)
          for (int k = 0; k < 3; ++k) u0[3 * size_t(c) + k] = patchU_h[3 * q + k];
        x.upload(U, u0.data(), u0.size());
        break;
      }
    if (p.scheme == SCHEME_QUICK || p.scheme == SCHEME_VANLEER || p.scheme == SCHEME_KOREN)
      phiF = fresh<double>(3 * size_t(nf));
    if (p.wall_model == 2) {
      q_w = fresh<double>(std::max(1, nb));
      L_bnd = fresh<double>(std::max(1, nb));
      L_cfg = fresh<double>(1);
      x.upload(L_cfg, &p.L_obukhov, 1);
    }
    bous = p.boussinesq != 0;
    if (bous) {
      T = fresh<double>(nc); T2 = fresh<double>(nc); T_bnd = fresh<double>(std::max(1, nb));
      gT = fresh<double>(3 * size_t(nc)); gT_prev = fresh<double>(3 * size_t(nc));
      HT = fresh<double>(nc); HT0 = fresh<double>(nc);
      if (p.t_scheme != SCHEME_CDS) phiT = fresh<double>(nf);
      std::vector<double> tref_c(nc, p.T_ref), tref_b(std::max(1, nb), p.T_ref), pt(std::max(1, npatch), p.T_ref);
      for (int q = 0; q < npatch; ++q) pt[q] = p.patch_T_val ? p.patch_T_val[q] : p.T_ref;
      x.upload(T, tref_c.data(), tref_c.size());
      x.upload(T_bnd, tref_b.data(), tref_b.size());
      patchT = up(pt);
    }
    ns = p.n_scalars;
    if (ns) {
      for (int k = 0; k < ns; ++k) {
        Cs[k] = fresh<double>(nc); Cs2[k] = fresh<double>(nc); Cb[k] = fresh<double>(std::max(1, nb));
        gC[k] = fresh<double>(3 * size_t(nc)); HC[k] = fresh<double>(nc); HC0[k] = fresh<double>(nc);
      }
      gCprev = fresh<double>(3 * size_t(nc));
      if (p.scalar_scheme != SCHEME_CDS) {
        for (int k = 0; k < ns; ++k) phiCk[k] = fresh<double>(nf);
        phiC = phiCk[0];
      }
      const size_t m = 4 * size_t(std::max(1, npatch));
      std::vector<int> hm(m, 3);
      std::vector<double> hv(m, 0.0), hf(m, 0.0), ho(m, 0.0), hp(m, 0.0);
      for (int q = 0; q < 4 * npatch; ++q) {
        if (p.scalar_bc_mode) hm[q] = p.scalar_bc_mode[q];
        if (p.scalar_bc_value) hv[q] = p.scalar_bc_value[q];
        if (p.scalar_bc_flux) hf[q] = p.scalar_bc_flux[q];
        if (p.scalar_bc_omega) ho[q] = p.scalar_bc_omega[q];
        if (p.scalar_bc_phase) hp[q] = p.scalar_bc_phase[q];
      }
      s_mode = up(hm); s_val = up(hv); s_flux = up(hf); s_om = up(ho); s_ph = up(hp);
    }
    if (fused_scalars()) {
      faceT = fresh<double>(2 * size_t((bous ? 1 : 0) + ns) * size_t(std::max(1, nint)));
      const int sch = bous ? p.t_scheme : p.scalar_scheme;
      if (sch == SCHEME_VANLEER || sch == SCHEME_KOREN) {
        tvd_ctx = fresh<TvdCtx>(2 * size_t(nf));
        x.for_each(nf, TvdCtxBuild<SCHEME_VANLEER>{ReconFace<SCHEME_VANLEER>{M, nullptr, nullptr, nullptr, nullptr,
                                                                             nullptr},
                                                   tvd_ctx});
      }
    }
    cproj = p.cell_projection == 1;
    if (cproj) {
      pj_b = fresh<double>(nc); pj_x = fresh<double>(nc); pj_r = fresh<double>(nc); pj_z = fresh<double>(nc);
      pj_p = fresh<double>(nc); pj_q = fresh<double>(nc); pj_ro = fresh<double>(nc); pj_bc = fresh<double>(nc);
      pj_v = fresh<double>(3 * size_t(nc));
    }
    cg.init(&x, nc);
    cg.set_single_reduction(p.cg_variant == 1);
    if (p.precision == 1) cg.set_mixed(A32, mixed_tau());
    if (p.precond == PRECOND_MULTIGRID) {  // aggregation multigrid in the flexible CG
      AmgOp op;
      op.n = nc;
      op.w = std::max(1, wP);
      op.nbr = anbr;
      op.coef = acoef;
      op.diag = diag;
      amg.reset(new AmgPc<X>());
      amg->build(&x, op, A.nbr, amg_cfg());
      if (p.precision == 1)
        cg.set_pc32(amg.get());
      else
        cg.set_fcg(amg.get());
    }
    apply_all_bc();
    x.sync();
  }

  void validate(const dopamine_mesh_desc& d) {
    if (nc <= 0 || nf <= 0 || nint < 0 || nb < 0 || nint + nb != nf) throw std::runtime_error("inconsistent mesh sizes");
    for (int f = 0; f < nf; ++f) {
      const int o = d.face_owner[f], n = d.face_neighbour[f];
      if (o < 0 || o >= nc) throw std::runtime_error("face owner out of range");
      if (f < nint && (n < 0 || n >= nc)) throw std::runtime_error("interior face neighbour out of range");
      if (f >= nint && n >= 0) throw std::runtime_error("boundary face with a neighbour");
      const int pid = d.face_patch[f];
      if (f >= nint && (pid < 0 || pid >= npatch)) throw std::runtime_error("boundary face without a patch");
    }
    for (int b = 0; b < nb; ++b)
      if (d.periodic_pair[b] < -1 || d.periodic_pair[b] >= nb) throw std::runtime_error("periodic pair out of range");
    for (int c = 0; c <= nc; ++c) {
      if (c < nc && d.cell_face_offset[c + 1] < d.cell_face_offset[c]) throw std::runtime_error("bad cell_face_offset");
      if (d.ls_offset[c] < 0) throw std::runtime_error("bad ls_offset");
    }
    for (int i = 0; i < d.cell_face_offset[nc]; ++i)
      if (d.cell_faces[i] < 0 || d.cell_faces[i] >= nf) throw std::runtime_error("cell_faces out of range");
    for (int i = 0; i < d.ls_offset[nc]; ++i)
      if (d.ls_cell[i] >= nc || d.ls_face[i] >= nf) throw std::runtime_error("LS stencil entry out of range");
  }

  static double refine_blend_weight(double skew) { return dm::refine_blend_weight(skew); }

  // invert_sym3 (src/solver/mod_fv_solver.f90:3466-3491), row-major in and out
  static void invert_sym3(const double* A_, double* Ai) {
    auto a = [&](int i, int j) { return A_[3 * (i - 1) + (j - 1)]; };
    const double c11 = a(2, 2) * a(3, 3) - a(2, 3) * a(3, 2);
    const double c12 = a(1, 3) * a(3, 2) - a(1, 2) * a(3, 3);
    const double c13 = a(1, 2) * a(2, 3) - a(1, 3) * a(2, 2);
    const double c22 = a(1, 1) * a(3, 3) - a(1, 3) * a(3, 1);
    const double c23 = a(1, 3) * a(2, 1) - a(1, 1) * a(2, 3);
    const double c33 = a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1);
    const double det = a(1, 1) * c11 + a(1, 2) * (a(2, 3) * a(3, 1) - a(2, 1) * a(3, 3)) +
                       a(1, 3) * (a(2, 1) * a(3, 2) - a(2, 2) * a(3, 1));
    const double tr = a(1, 1) + a(2, 2) + a(3, 3);
    const double t = std::fmax(tr, SMALL);
    if (std::fabs(det) <= 1.0e-12 * t * t * t) {
      for (int q = 0; q < 9; ++q) Ai[q] = 0.0;
      Ai[0] = Ai[4] = Ai[8] = 3.0 / t;
      return;
    }
    Ai[0] = c11 / det; Ai[1] = c12 / det; Ai[2] = c13 / det;
    Ai[3] = c12 / det; Ai[4] = c22 / det; Ai[5] = c23 / det;
    Ai[6] = c13 / det; Ai[7] = c23 / det; Ai[8] = c33 / det;
  }

  void build_filter_tensor(const dopamine_mesh_desc& d, std::vector<double>& G) {
    for (int c = 0; c < nc; ++c) {
      const double V = d.cell_volume[c];
      double g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (int q = d.cell_face_offset[c]; q < d.cell_face_offset[c + 1]; ++q) {
        const int f = d.cell_faces[q];
        double dx[3];
        for (int i = 0; i < 3; ++i) dx[i] = d.face_centre[3 * size_t(f) + i] - d.cell_centroid[3 * size_t(c) + i];
        const double a = d.face_area[f];
        for (int n = 0; n < 3; ++n)
          for (int m = 0; m < 3; ++m) g[3 * n + m] = g[3 * n + m] + a * dx[m] * dx[n];
      }
      const double det = det3_colmajor(g);
      const double t = (g[0] + g[4] + g[8]) / 3.0;
      const double tr3 = t * t * t;
      double* Gc = &G[9 * size_t(c)];
      if (tr3 > 0.0 && det > 1.0e-12 * tr3) {
        const double s = std::pow(V * V / det, 1.0 / 3.0);
        for (int q = 0; q < 9; ++q) Gc[q] = s * g[q];
      } else {
        for (int q = 0; q < 9; ++q) Gc[q] = 0.0;
        Gc[0] = Gc[4] = Gc[8] = std::pow(V, 2.0 / 3.0);
      }
    }
    std::vector<double> gnb(9 * size_t(nc)), ws(nc);
    for (int pass = 0; pass < 2; ++pass) {
      std::fill(gnb.begin(), gnb.end(), 0.0);
      std::fill(ws.begin(), ws.end(), 0.0);
      for (int f = 0; f < nint; ++f) {
        const int o = d.face_owner[f], n = d.face_neighbour[f];
        const double a = d.face_area[f];
        for (int q = 0; q < 9; ++q) gnb[9 * size_t(o) + q] = gnb[9 * size_t(o) + q] + a * G[9 * size_t(n) + q];
        ws[o] = ws[o] + a;
        for (int q = 0; q < 9; ++q) gnb[9 * size_t(n) + q] = gnb[9 * size_t(n) + q] + a * G[9 * size_t(o) + q];
        ws[n] = ws[n] + a;
      }
      for (int c = 0; c < nc; ++c)
        if (ws[c] > SMALL)
          for (int q = 0; q < 9; ++q)
            G[9 * size_t(c) + q] = (1.0 - 0.5) * G[9 * size_t(c) + q] + 0.5 * gnb[9 * size_t(c) + q] / ws[c];
    }
  }

  // ---------------------------------------------------------------- boundary passes

  void bc_velocity(int filter, const double* gU) { x.for_each(nb, BcVelocity{M, U, gU, U_bnd, filter, out_ready}); }

  // find_sample_cell (src/sgs/mod_wall_model.f90:126-173): from the wall face's owner, step to
  // the face neighbour farthest from the wall along the inward normal until the distance reaches
  // h_wm (at most 20 steps, interior faces only)
  static void sample_walk(const dopamine_mesh_desc& d, int b, double h_wm, int* cell, double* y) {
    const int fg = d.n_int_faces + b;
    const double* n = &d.face_normal[3 * size_t(fg)];
    const double* xw = &d.face_centre[3 * size_t(fg)];
    const double n_in[3] = {-n[0], -n[1], -n[2]};
    int c = d.face_owner[fg];
    auto inward = [&](int cc) {
      double s = 0.0;
      for (int i = 0; i < 3; ++i) s = s + (d.cell_centroid[3 * size_t(cc) + i] - xw[i]) * n_in[i];
      return s;
    };
    *cell = c;
    *y = std::fabs(inward(c)) + SMALL;
    for (int step = 1; step <= 20; ++step) {
      if (*y >= h_wm) break;
      int best = -1;
      double best_dist = *y;
      for (int q = d.cell_face_offset[c]; q < d.cell_face_offset[c + 1]; ++q) {
        const int fj = d.cell_faces[q];
        if (fj < 0 || fj >= d.n_faces || d.face_neighbour[fj] < 0) continue;
        const int other = d.face_owner[fj] == c ? d.face_neighbour[fj] : d.face_owner[fj];
        const double dist = inward(other);
        if (dist > best_dist) {
          best_dist = dist;
          best = other;
        }
      }
      if (best < 0) break;
      c = best;
      *cell = best;
      *y = best_dist;
    }
  }

  // outlet_advance_boundary (src/bc/mod_bc.f90:681-772): each active outlet patch's stage-start
  // statistics, then one RK3W stage of the convective faces' boundary velocity
  void outlet_advance(double dt, double a, double bk) {
    if (out_q.empty()) return;
    for (size_t i = 0; i < out_q.size(); ++i)
      x.template reduce<3, false>(out_n[i], OutletStat{M, U, out_start[i]}, OutletStatFin{out_sp, out_q[i]});
    x.for_each(nb, OutletAdvance{M, U, grad, out_sp, U_bnd, out_gprev, dt, a, bk});
    out_ready = 1;
  }
  // apply_wall_model_bc with the time filter (src/sgs/mod_wall_model.f90:176-313); the first
  // stage of a step hands the filter dt, later stages 0, so it advances once per step
  void wall_model(double dt_use) {
    WallModel w{M, U, tau, prm.nu};
    if (wm_fmode) {
      w.uf = wm_filt;
      w.fmode = wm_fmode;
      w.seed = wm_filt_ready ? 0 : 1;
      w.w = wm_fmode == 1 ? dt_use / (prm.wall_model_filter_time + dt_use) : 0.0;
      w.dt_use = dt_use;
      w.tf = prm.wall_model_filter_tf_fac;
    }
    x.for_each(nb, w);
    if (wm_fmode) wm_filt_ready = 1;
  }
  // update_outlet_pressure (src/bc/mod_bc.f90:789-818) from the predicted velocity
  void outlet_pressure() {
    if (!out_q.empty()) x.for_each(nb, OutletPressure{M, U, out_sp, patchP, outlet_pb});
  }
  void bc_pressure(int filter, const double* g, double scale, const double* pob) {
    x.for_each(nb, BcPressure{M, P, nullptr, g, pob, patchP, p_bnd, scale, filter});
  }
  void bc_temperature(int filter, const double* g) {
    if (bous) x.for_each(nb, BcTemperature{M, T, g, patchT, T_bnd, filter});
  }
  void bc_scalars(int filter) {
    for (int k = 0; k < ns; ++k)
      x.for_each(nb, BcScalar{M, Cs[k], nu_sgs, Cb[k], s_mode, s_val, s_flux, s_om, s_ph, prm.scalar_diffusivity[k],
                              prm.Pr_t, time, k, filter});
  }
  void apply_all_bc() {
    bc_velocity(0, grad);
    bc_pressure(0, skip_grad_p ? nullptr : gp, 1.0, outlet_pb);
    bc_temperature(0, gT);
    bc_scalars(0);
  }
  // refresh_halos_and_bc (single rank: non-periodic pass, then periodic pass)
  void refresh_bc() {
    for (int pass = 1; pass <= 2; ++pass) {
      bc_velocity(pass, grad);
      bc_pressure(pass, skip_grad_p ? nullptr : gp, 1.0, outlet_pb);
      bc_temperature(pass, gT);
      bc_scalars(pass);
    }
  }
  void velocity_bc_both() {
    bc_velocity(1, grad);
    bc_velocity(2, nullptr);
  }

  // ---------------------------------------------------------------- gradients

  // compute_all_gradients (velocity), with the SGS viscosity of the stage fused in when sgs_dt > 0
  template <int SCHEME>
  void explicit_rhs(double dt, int k) {
    double fb[3];
    body_force_now(fb);
    x.for_each(nc, ExplicitRhs<SCHEME>{M, U, U_bnd, UF, grad, nu_sgs, tau, H0, H, U2, prm.nu, fb[0], fb[1], fb[2], dt,
                                       RK3_A[k], RK3_B[k], n_wm > 0, phiF, bous ? T : nullptr, prm.beta, prm.T_ref,
                                       prm.g[0], prm.g[1], prm.g[2], prm.use_rotation ? 1 : 0, prm.omega_rot[0],
                                       prm.omega_rot[1], prm.omega_rot[2]});
  }
  // body_force_now (src/solver/mod_fv_solver.f90:2904-2923): constant or mass-flux force plus the
  // harmonic forcing at the current (stage) time
  double mff[3] = {0.0, 0.0, 0.0};
  void body_force_now(double f[3]) const {
    for (int k = 0; k < 3; ++k) f[k] = prm.mass_flux ? mff[k] : prm.f_const[k];
    if (prm.use_harmonic) {
      const double* d = prm.forcing_dir;
      const double dn = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      double hd[3] = {1.0, 0.0, 0.0};
      if (dn > SMALL)
        for (int k = 0; k < 3; ++k) hd[k] = d[k] / dn;
      const double hs = prm.forcing_amp * std::cos(prm.forcing_omega * time + prm.forcing_phase);
      for (int k = 0; k < 3; ++k) f[k] = f[k] + hs * hd[k];
    }
  }
  // temperature face values (reconstruction schemes), then H_T and the T predictor into T2
  void explicit_rhs_T(double dt, int k) {
    const ReconFace<SCHEME_QUICK> rq{M, U, U_bnd, UF, grad, nullptr};
    const ReconFace<SCHEME_VANLEER> rv{M, U, U_bnd, UF, grad, nullptr};
    const ReconFace<SCHEME_KOREN> rk{M, U, U_bnd, UF, grad, nullptr};
    switch (prm.t_scheme) {
      case SCHEME_QUICK: x.for_each(nf, ReconFaceScalar<SCHEME_QUICK>{rq, T, T_bnd, gT, phiT}); break;
      case SCHEME_VANLEER: x.for_each(nf, ReconFaceScalar<SCHEME_VANLEER>{rv, T, T_bnd, gT, phiT}); break;
      case SCHEME_KOREN: x.for_each(nf, ReconFaceScalar<SCHEME_KOREN>{rk, T, T_bnd, gT, phiT}); break;
      default: break;
    }
    x.for_each(nc, ExplicitRhsT{M, T, T_bnd, gT, UF, nu_sgs, HT0, phiT, HT, T2, prm.alpha_T, prm.Pr_t, dt, RK3_A[k],
                                RK3_B[k], prm.t_scheme == SCHEME_CDS ? 1 : 0,
                                (prm.wall_model == 2 && n_wm) ? q_w : nullptr});
  }

  // Temperature and passive scalars that share a scheme run as one multi-field pass (face values,
  // then H and the predictor); the results are those of the per-field kernels.
  bool fused_scalars() const {
    return (bous ? 1 : 0) + ns >= 2 && (!bous || ns == 0 || prm.t_scheme == prm.scalar_scheme);
  }
  void explicit_rhs_TC(double dt, int k) {
    const double *fv[RHS_MAX_FIELDS] = {}, *fb[RHS_MAX_FIELDS] = {}, *fg[RHS_MAX_FIELDS] = {},
                 *f0[RHS_MAX_FIELDS] = {}, *fq[RHS_MAX_FIELDS] = {};
    double *fphi[RHS_MAX_FIELDS] = {}, *fh[RHS_MAX_FIELDS] = {}, *fn[RHS_MAX_FIELDS] = {};
    double fa[RHS_MAX_FIELDS] = {}, fp[RHS_MAX_FIELDS] = {};
    int n = 0;
    if (bous) {
      fv[n] = T; fb[n] = T_bnd; fg[n] = gT; f0[n] = HT0; fphi[n] = phiT; fh[n] = HT; fn[n] = T2;
      fq[n] = (prm.wall_model == 2 && n_wm) ? q_w : nullptr;
      fa[n] = prm.alpha_T; fp[n] = prm.Pr_t;
      ++n;
    }
    for (int q = 0; q < ns; ++q, ++n) {
      fv[n] = Cs[q]; fb[n] = Cb[q]; fg[n] = gC[q]; f0[n] = HC0[q]; fphi[n] = phiCk[q]; fh[n] = HC[q]; fn[n] = Cs2[q];
      fq[n] = nullptr;
      fa[n] = prm.scalar_diffusivity[q]; fp[n] = prm.scalar_pr_t[q];
    }
    const int scheme = bous ? prm.t_scheme : prm.scalar_scheme;
    const int cds = scheme == SCHEME_CDS ? 1 : 0;
    const auto run = [&](auto nfc) {
      constexpr int NF = decltype(nfc)::value;
      // face half: interior-face fluxes into faceT (and the boundary face values)
      const auto faces = [&](auto sch) {
        constexpr int S = decltype(sch)::value;
        FaceRhsTN<S, NF> r{{ReconFaceScalar<S>{ReconFace<S>{M, U, U_bnd, UF, grad, nullptr}, fv[0], fb[0], fg[0], fphi[0]},
                            {}, {}, {}, {}, tvd_ctx},
                           nu_sgs, cds, {}, {}, faceT};
        for (int i = 0; i < RHS_MAX_FIELDS; ++i) {
          r.rec.phi[i] = fv[i];
          r.rec.pb[i] = fb[i];
          r.rec.gphi[i] = fg[i];
          r.rec.phiF[i] = fphi[i];
          r.alpha[i] = fa[i];
          r.prt[i] = fp[i];
        }
        x.for_each(cds ? nint : nf, r);
      };
      if (scheme == SCHEME_QUICK) faces(std::integral_constant<int, SCHEME_QUICK>{});
      else if (scheme == SCHEME_KOREN) faces(std::integral_constant<int, SCHEME_KOREN>{});
      else faces(std::integral_constant<int, SCHEME_VANLEER>{});  // van Leer, or CDS (no reconstruction)
      ExplicitRhsTN<NF> e{M, UF, nu_sgs, dt, RK3_A[k], RK3_B[k], cds, faceT};
      for (int i = 0; i < RHS_MAX_FIELDS; ++i) {
        e.T[i] = fv[i]; e.Tb[i] = fb[i]; e.gT[i] = fg[i]; e.HT0[i] = f0[i]; e.phiF[i] = fphi[i]; e.qw[i] = fq[i];
        e.HT[i] = fh[i]; e.Tnew[i] = fn[i]; e.alpha[i] = fa[i]; e.prt[i] = fp[i];
      }
      x.for_each(nc, e);
    };
    switch (n) {
      case 2: run(std::integral_constant<int, 2>{}); break;
      case 3: run(std::integral_constant<int, 3>{}); break;
      case 4: run(std::integral_constant<int, 4>{}); break;
      default: run(std::integral_constant<int, 5>{}); break;
    }
  }

  // H_C and the predictor of each passive scalar into Cs2 (compute_explicit_rhs, :1719-1743)
  void explicit_rhs_C(double dt, int k) {
    const ReconFace<SCHEME_QUICK> rq{M, U, U_bnd, UF, grad, nullptr};
    const ReconFace<SCHEME_VANLEER> rv{M, U, U_bnd, UF, grad, nullptr};
    const ReconFace<SCHEME_KOREN> rk{M, U, U_bnd, UF, grad, nullptr};
    for (int q = 0; q < ns; ++q) {
      switch (prm.scalar_scheme) {
        case SCHEME_QUICK: x.for_each(nf, ReconFaceScalar<SCHEME_QUICK>{rq, Cs[q], Cb[q], gC[q], phiC}); break;
        case SCHEME_VANLEER: x.for_each(nf, ReconFaceScalar<SCHEME_VANLEER>{rv, Cs[q], Cb[q], gC[q], phiC}); break;
        case SCHEME_KOREN: x.for_each(nf, ReconFaceScalar<SCHEME_KOREN>{rk, Cs[q], Cb[q], gC[q], phiC}); break;
        default: break;
      }
      x.for_each(nc, ExplicitRhsT{M, Cs[q], Cb[q], gC[q], UF, nu_sgs, HC0[q], phiC, HC[q], Cs2[q],
                                  prm.scalar_diffusivity[q], prm.scalar_pr_t[q], dt, RK3_A[k], RK3_B[k],
                                  prm.scalar_scheme == SCHEME_CDS ? 1 : 0, nullptr});
    }
  }

  // dynamic Smagorinsky (sgs 4) on the current U, grad and U_bnd
  void dynamic_sgs(double dt_stab) {
    x.for_each(nc, DynPack{U, grad, dyn_fb});
    x.for_each(nc, DynFilter<15>{M, dyn_fb, dyn_fh, dyn_uh});
    x.for_each(nc, GgGradVec{M, dyn_uh, U_bnd, nullptr, dyn_gh});
    x.for_each(nc, DynLM{M, dyn_fh, dyn_gh, dyn_lm});
    x.for_each(nc, DynFilter<2>{M, dyn_lm, dyn_lmh, nullptr});
    x.for_each(nc, DynNu{M, dyn_lmh, grad, nu_sgs, dt_stab});
  }

  // bc_only (least squares, no SGS): the velocity gradient everywhere and the temperature gradient
  // only where a boundary value reads it before the next full call; the mid-stage call of the
  // first two stages, whose scalar gradients nothing else reads before the next stage recomputes
  // them (the last stage computes every field, so the state after a step is complete).
  void gradients(bool with_sgs, double sgs_dt, bool bc_only = false) {
    const bool dyn = with_sgs && prm.sgs == 4;
    const int sgs = (with_sgs && !dyn) ? prm.sgs : -1;
    if (prm.gradient == 0 && bc_only && !with_sgs) {
      x.tbegin(K_LS_GRAD);
      x.for_each(nc, LsGradAll<0>{M, U, U_bnd, grad, nu_sgs, -1, 0.0, 0.0, 0.0, 0.0, {}, {}, {}});
      x.tend(K_LS_GRAD);
      if (bous && n_tbc) x.for_each(n_tbc, LsGradScaAt{LsGradSca{M, T, nullptr, T_bnd, nullptr, gT}, tbc_cell});
      return;
    }
    if (prm.gradient == 0) {  // least squares: every field in one pass (LsGradAll)
      const double* sa[LS_MAX_SCALARS] = {};
      const double* sb[LS_MAX_SCALARS] = {};
      double* sg[LS_MAX_SCALARS] = {};
      int nsf = 0;
      if (bous) {
        sa[nsf] = T;
        sb[nsf] = T_bnd;
        sg[nsf++] = gT;
      }
      for (int k = 0; k < ns; ++k) {
        sa[nsf] = Cs[k];
        sb[nsf] = Cb[k];
        sg[nsf++] = gC[k];
      }
      x.tbegin(K_LS_GRAD);
      const auto run = [&](auto n) {
        constexpr int N = decltype(n)::value;
        LsGradAll<N> f{M, U, U_bnd, grad, nu_sgs, sgs, prm.c_vreman, prm.cs_smag, prm.c_amd, sgs_dt, {}, {}, {}};
        for (int i = 0; i < LS_MAX_SCALARS; ++i) {
          f.sa[i] = sa[i];
          f.sb[i] = sb[i];
          f.sg[i] = sg[i];
        }
        x.for_each(nc, f);
      };
      switch (nsf) {
        case 0: run(std::integral_constant<int, 0>{}); break;
        case 1: run(std::integral_constant<int, 1>{}); break;
        case 2: run(std::integral_constant<int, 2>{}); break;
        case 3: run(std::integral_constant<int, 3>{}); break;
        case 4: run(std::integral_constant<int, 4>{}); break;
        default: run(std::integral_constant<int, 5>{}); break;
      }
      x.tend(K_LS_GRAD);
      if (with_sgs && prm.sgs == 3 && prm.amd_wall_gradient && n_amdw > 0)
        x.for_each(n_amdw, AmdWall{M, amdw_cell, amdw_off, amdw_face, U, grad, nu_sgs, prm.nu, prm.c_amd, sgs_dt});
      if (dyn) dynamic_sgs(sgs_dt);
      return;
    }
    if (bous) {  // compute_all_gradients: the temperature gradient alongside the velocity's
      if (prm.gradient == 0) {
        x.for_each(nc, LsGradSca{M, T, nullptr, T_bnd, nullptr, gT});
      } else {
        x.for_each(nc, GgGradSca{M, T, nullptr, T_bnd, nullptr, nullptr, gT});
        for (int it = 2; it <= prm.gg_iterations; ++it) {
          x.copy(gT_prev, gT, 3 * size_t(nc));
          x.for_each(nc, GgGradSca{M, T, nullptr, T_bnd, nullptr, gT_prev, gT});
        }
      }
    }
    for (int k = 0; k < ns; ++k) {  // the scalars' gradients (LS, or GG with its sweeps)
      if (prm.gradient == 0) {
        x.for_each(nc, LsGradSca{M, Cs[k], nullptr, Cb[k], nullptr, gC[k]});
      } else {
        x.for_each(nc, GgGradSca{M, Cs[k], nullptr, Cb[k], nullptr, nullptr, gC[k]});
        for (int it = 2; it <= prm.gg_iterations; ++it) {
          x.copy(gCprev, gC[k], 3 * size_t(nc));
          x.for_each(nc, GgGradSca{M, Cs[k], nullptr, Cb[k], nullptr, gCprev, gC[k]});
        }
      }
    }
    if (prm.gradient == 0) {
      x.tbegin(K_LS_GRAD);
      x.for_each(nc, LsGradVec{M, U, U_bnd, grad, nu_sgs, sgs, prm.c_vreman, prm.cs_smag, prm.c_amd, sgs_dt});
      x.tend(K_LS_GRAD);
    } else {
      x.for_each(nc, GgGradVec{M, U, U_bnd, nullptr, grad});
      for (int it = 2; it <= prm.gg_iterations; ++it) {
        x.copy(gprev, grad, 9 * size_t(nc));
        x.for_each(nc, GgGradVec{M, U, U_bnd, gprev, grad});
      }
      if (dyn)
        dynamic_sgs(sgs_dt);
      else if (with_sgs)
        x.for_each(nc, SgsCell{M, grad, nu_sgs, prm.sgs, prm.c_vreman, prm.cs_smag, prm.c_amd, sgs_dt});
    }
    if (with_sgs && prm.sgs == 3 && prm.amd_wall_gradient && n_amdw > 0)
      x.for_each(n_amdw, AmdWall{M, amdw_cell, amdw_off, amdw_face, U, grad, nu_sgs, prm.nu, prm.c_amd, sgs_dt});
  }
  // gradient_of_scalar of a (+ a2) with boundary values pb (+ pb2) into gp
  void scalar_gradient(const double* a, const double* a2, const double* pb, const double* pb2) {
    if (prm.gradient == 0) {
      x.for_each(nc, LsGradSca{M, a, a2, pb, pb2, gp});
    } else {
      x.for_each(nc, GgGradSca{M, a, a2, pb, pb2, nullptr, gp});
      for (int it = 2; it <= prm.gg_iterations; ++it) {
        x.copy(gp_prev, gp, 3 * size_t(nc));
        x.for_each(nc, GgGradSca{M, a, a2, pb, pb2, gp_prev, gp});
      }
    }
  }
  void pressure_gradient() {
    if (!skip_grad_p) scalar_gradient(P, nullptr, p_bnd, nullptr);
  }

  // ---------------------------------------------------------------- projection pieces

  void face_fluxes_and_rhs(double dt, bool incr) {
    if (has_skipped) x.zero(ff, nf);
    if (!skip_grad_p) {
      if (incr)
        scalar_gradient(phi_prev, P, phi_prev_bnd, p_bnd);
      else
        scalar_gradient(P, nullptr, p_bnd, nullptr);
    }
    const int rck = (!skip_nonortho && !prm.rhie_chow_dual) ? 1 : 0;
    x.for_each(npf, FaceFluxRC{M, U, grad, gp_rc, phi_prev, ff, dt, rck});
    x.template reduce<1, true>(nc, FluxRhs{M, U, U_bnd, grad, phi_prev, gp, ff, rhs, dt, skip_nonortho ? 0 : 1,
                                           outlet_pb, cproj},
                               sk::DivFin{sc});
    x.for_each(nb, PeriodicAntisym{M, ff});
  }

  void correct_face_fluxes(double dt, bool ptot) {
    x.for_each(npf + nb, CorrectFaceFlux{M, P, gp, outlet_pb, ff, dt, ptot ? 1 : 0, skip_nonortho ? 0 : 1});
    x.for_each(nb, PeriodicAntisym{M, ff});
  }

  void correct_velocity(double dt, bool with_cross, bool ptot, double* gpc_out) {
    const int do_cross = (with_cross && !skip_nonortho) ? 1 : 0;
    double fb[3];
    body_force_now(fb);
    x.for_each(nc, CorrectVelocity{M, P, gp, outlet_pb, U, gpc_out, dt, fb[0], fb[1], fb[2],
                                   do_cross, ptot ? 1 : 0, bous ? T : nullptr, prm.beta, prm.T_ref, prm.g[0],
                                   prm.g[1], prm.g[2]});
  }

  int pressure_solve(bool ptot, const double* atol_dev, double atol) {
    x.copy(cg_b, rhs, nc);
    if (ptot && n_dir) {
      // Dirichlet rhs increments c_b p_val; rows can repeat, so apply them in order on the host side
      std::vector<double> hb(nc);
      x.download(hb.data(), cg_b, nc);
      for (int i = 0; i < n_dir; ++i) hb[dir_rows_h[i]] = hb[dir_rows_h[i]] + dir_vals_h[i];
      x.upload(cg_b, hb.data(), nc);
    }
    const int maxit = prm.p_max_iter > 2000 ? prm.p_max_iter : 2000;
    int its = 0, total = 0, retried = 0;
    double res = 0.0;
    const bool ns = n_dir == 0;
    int sweeps = 0;
    const int rc = solve_with_retry(
        maxit, prm.p_tol,
        [&](int* it, double* r) {
          const int q = policy ? policy->solve(A, cg_b, P, prm.p_tol, atol, atol_dev, maxit, ns, it, r)
                               : cg.solve(A, cg_b, P, prm.p_tol, atol, atol_dev, maxit, ns, it, r);
          sweeps += cg.last_sweeps();
          return q;
        },
        [&](int mi, int* it, double* r) {
          cg.set_plain(true);  // the retry runs the Jacobi passes, whatever preconditioner is set
          const int q = cg.solve(A, cg_b, P, prm.p_tol, atol, atol_dev, mi, ns, it, r);
          cg.set_plain(false);
          sweeps += cg.last_sweeps();
          return q;
        },
        [&] { x.zero(P, nc); }, [&] { return cg.last_atol(); }, &its, &res, &total, &retried);
    n_solves++;
    n_iters += total;
    last_iters = its;
    n_sweeps += sweeps;
    n_retries += retried;
    if (rc != 0) {
      err = "pressure solve failed: " + std::to_string(its) + " iterations of the retry from x = 0 (limit " +
            std::to_string(4 * maxit) + ") left the true residual at " + std::to_string(res) + " (rtol " +
            std::to_string(prm.p_tol) + ", atol " + std::to_string(cg.last_atol()) + "); returns " +
            std::to_string(rc);
    }
    return rc;
  }

  // project_cells_exact (src/solver/mod_fv_solver.f90:3224-3283) with cp_project
  // (src/numerics/mod_cell_projection.f90:192-303): flexible PCG on (B V^-1 B^T) chi = B u + bconst,
  // preconditioned by the pressure solve (z = 0, rtol 0.05, 50 iterations, no absolute floor),
  // at most 100 iterations; then u -= V^-1 B^T chi.
  double pj_get(int i) {
    SolverScalars h;
    x.download(&h, sc, 1);
    return h.pj[i];
  }
  void pj_precond() {
    x.zero(pj_z, nc);
    int its = 0;
    double res = 0.0;
    if (policy)
      policy->precondition(A, pj_r, pj_z, 0.05, 50, n_dir == 0, &its, &res);
    else
      cg.solve(A, pj_r, pj_z, 0.05, 1.0e-50, nullptr, 50, n_dir == 0, &its, &res);
  }
  // The loop's scalars (r.z, alpha, beta, the energy norms, the stop flag) live on the device: the
  // reductions' final kernels update them and the passes read them, so the host reads one flag per
  // iteration, to decide whether to run the preconditioner solve.
  int pj_done_flag() {
    SolverScalars h;
    x.download(&h, sc, 1);
    return h.pj_done;
  }
  void project_exact(double dt) {
    x.for_each(nc, ProjBoundary{M, U_bnd, outlet_pb, U, pj_bc, dt});
    x.template reduce<1, true>(nc, ProjInit{M, U, pj_bc, pj_b, pj_r}, sk::ProjBnFin{sc});
    const double bn = pj_get(0);
    int n_it = 0;
    if (bn > DBL_MIN) {
      x.zero(pj_x, nc);
      pj_precond();
      x.template reduce<1, false>(nc, ProjFirstDir{pj_r, pj_z, pj_p}, sk::ProjRzFin{sc});
      const double tol = prm.cell_projection_tol;
      const int* done = &sc->pj_done;
      for (int it = 1; it <= 100; ++it) {
        n_it = it;
        x.for_each(nc, ProjAdjoint{M, pj_p, pj_v});
        x.template reduce<1, false>(nc, ProjApply{M, pj_v, pj_p, pj_q}, sk::ProjAlphaFin{sc});
        x.template reduce<1, true>(nc, ProjUpdate{pj_p, pj_q, pj_x, pj_r, pj_ro, &sc->pj_alpha},
                                   sk::ProjStopFin{sc, it, tol}, done);
        if (pj_done_flag()) break;
        pj_precond();
        x.template reduce<2, false>(nc, ProjBeta{pj_r, pj_ro, pj_z}, sk::ProjBetaFin{sc});
        x.for_each(nc, ProjDir{pj_z, pj_p, &sc->pj_beta});
      }
      x.for_each(nc, ProjCorrect{M, pj_x, U});
    }
    n_proj++;
    n_proj_its += n_it;
  }

  void anchor_pressure() {
    // mean removal is folded into RotAccumulate's reduction; this variant serves project_once
    x.template reduce<1, false>(nc, RotAccumulate{M, P, zeros(), nu_sgs, rhs, 0.0, 0.0, 0}, sk::AnchorFin{sc, n_outlet == 0});
    x.for_each(nc, ShiftP{P, &sc->shift});
  }
  const double* zeros() {
    if (!zero_) zero_ = fresh<double>(nc);
    return zero_;
  }
  double* zero_ = nullptr;

  // Divergence tolerance from the volume-weighted mean speed (:1424-1437), left in the execution
  // space where the CG reads it.
  void atol_for(double gdt) { x.template reduce<2, false>(nc, BulkSpeed{M, U}, sk::AtolFin{sc, gdt}); }

  // ---------------------------------------------------------------- public steps

  int project_once(double dt) {
    err.clear();
    if (dt <= 0.0) dt = prm.dt;
    x.zero(rhs, nc);
    x.zero(phi_prev, nc);
    x.zero(phi_prev_bnd, std::max(1, nb));
    x.zero(gp_rc, 3 * size_t(nc));
    refresh_bc();
    face_fluxes_and_rhs(dt, false);
    int rc = 0;
    if (prm.p_max_iter > 0) {
      atol_for(dt);
      rc = pressure_solve(true, &sc->atol, 0.0);
    }
    anchor_pressure();
    correct_face_fluxes(dt, true);
    std::swap(UF, ff);
    u_face_ready = 1;
    bc_pressure(0, nullptr, 1.0, outlet_pb);
    if (cproj)
      project_exact(dt);
    else
      correct_velocity(dt, true, true, nullptr);
    bc_velocity(0, nullptr);
    if (prm.rotational) {
      x.template reduce<1, false>(nc, RotAccumulate{M, P, zeros(), nu_sgs, rhs, prm.nu, dt, 1}, sk::AnchorFin{sc, 0});
      bc_pressure(0, nullptr, 1.0, outlet_pb);
    }
    x.copy(p_old, P, nc);
    x.copy(phi_prev, P, nc);
    x.copy(phi_prev_bnd, p_bnd, std::max(1, nb));
    bc_pressure(0, nullptr, 1.0, outlet_pb);
    pressure_gradient();
    x.sync();
    return rc;
  }

  // max over the cells of |sum of the projected face fluxes| (m^3/s): the discrete divergence
  // integrated over the cell, the per-cell measure of the projection's mass error
  double cell_net_flux_max() {
    x.template reduce<1, true>(nc, FluxDivergence{M, UF, 1}, sk::ProjFin<1>{sc, 0});
    SolverScalars h;
    x.download(&h, sc, 1);
    return h.pj[0];
  }

  // solver_state_is_finite (src/solver/mod_fv_solver.f90:3925-3945): U and p hold no NaN or Inf
  bool state_is_finite() {
    x.template reduce<1, false>(nc, FiniteSum{U, P}, sk::ProjFin<1>{sc, 0});
    SolverScalars h;
    x.download(&h, sc, 1);
    const double v = h.pj[0];
    return std::fabs(v) <= DBL_MAX;
  }

  double compute_dt() {
    x.template reduce<2, true>(nc, CflRates{M, U, UF, nu_sgs, prm.nu, u_face_ready}, sk::CflFin{sc});
    SolverScalars h;
    x.download(&h, sc, 1);
    double dc = 1e308, dd = 1e308;
    if (h.cfl > SMALL) dc = prm.cfl_max / h.cfl;
    if (h.dif > SMALL) dd = 1.0 / h.dif;
    double dt = std::fmin(dc, dd);
    if (!(dt > 0.0)) dt = prm.dt;
    dt = std::fmin(dt, prm.dt);
    const double dtmin = 1.0e-6 * prm.dt;
    if (!(dt > 0.0)) dt = prm.dt;
    else if (dt < dtmin) return -1.0;
    else dt = std::fmax(dt, dtmin);
    return dt;
  }

  int advance(double dt_req, double* dt_used) {
    err.clear();
    double dt = dt_req;
    const double safe = compute_dt();
    if (safe < 0.0) {
      err = "time step below 1e-6 dt by the CFL/diffusion limit";
      return -3;
    }
    if (dt > safe) dt = safe;
    if (prm.mass_flux) {
      x.template reduce<3, false>(nc, Momentum{M, U}, sk::ProjFin<3>{sc, 0});
      SolverScalars h;
      x.download(&h, sc, 1);
      for (int k = 0; k < 3; ++k) mff[k] = (prm.U_bulk_target[k] - h.pj[k] / std::fmax(h.vtot, SMALL)) / dt;
    }
    const int rc = advance_rk3w(dt);
    time = time + dt;
    if (dt_used) *dt_used = dt;
    return rc;
  }

  int advance_rk3w(double dt) {
    x.zero(H, 3 * size_t(nc));
    x.zero(H0, 3 * size_t(nc));
    if (bous) {
      x.zero(HT, nc);
      x.zero(HT0, nc);
    }
    for (int q = 0; q < ns; ++q) {
      x.zero(HC[q], nc);
      x.zero(HC0[q], nc);
    }
    x.zero(rhs, nc);
    const double time_n = time;
    int rc = 0;
    const int scheme = prm.scheme;
    for (int k = 0; k < 3; ++k) {
      const double gdt = RK3_G[k] * dt;
      time = time_n + RK3_C[k] * dt;
      refresh_bc();
      // the MOST wall model (latency-bound, few faces) reads only U, T and its own state: it runs on
      // the second stream while the gradients stream on the main one; the explicit RHS joins it
      const bool most = n_wm && prm.wall_model == 2;
      if (most) {
        x.side_begin();
        x.template reduce<2, false>(nb, MostWall{M, U, bous ? T : nullptr, prm.T_ref, prm.T_surface,
                                                L_cfg, L_bnd, tau, q_w},
                                    MostLFin{L_cfg, (prm.compute_L && k == 0) ? 1 : 0});
        x.side_end();
      }
      gradients(true, gdt);
      if (prm.wall_model == 1)
        x.zero(tau, 3 * size_t(std::max(1, nb)));
      else if (most)
        x.join();
      else if (n_wm)
        wall_model(k == 0 ? dt : 0.0);
      if (scheme == SCHEME_QUICK) x.for_each(nf, ReconFace<SCHEME_QUICK>{M, U, U_bnd, UF, grad, phiF});
      if (scheme == SCHEME_VANLEER) x.for_each(nf, ReconFace<SCHEME_VANLEER>{M, U, U_bnd, UF, grad, phiF});
      if (scheme == SCHEME_KOREN) x.for_each(nf, ReconFace<SCHEME_KOREN>{M, U, U_bnd, UF, grad, phiF});
      x.tbegin(K_EXPLICIT);
      switch (scheme) {
        case SCHEME_KEP: explicit_rhs<SCHEME_KEP>(dt, k); break;
        case SCHEME_CDS: explicit_rhs<SCHEME_CDS>(dt, k); break;
        case SCHEME_QUICK: explicit_rhs<SCHEME_QUICK>(dt, k); break;
        case SCHEME_VANLEER: explicit_rhs<SCHEME_VANLEER>(dt, k); break;
        default: explicit_rhs<SCHEME_KOREN>(dt, k); break;
      }
      x.tend(K_EXPLICIT);
      if (fused_scalars()) {
        explicit_rhs_TC(dt, k);
      } else {
        if (bous) explicit_rhs_T(dt, k);
        if (ns) explicit_rhs_C(dt, k);
      }
      outlet_advance(dt, RK3_A[k], RK3_B[k]);  // from the stage-start velocity, before the predictor lands
      std::swap(U, U2);
      if (bous) std::swap(T, T2);
      for (int q = 0; q < ns; ++q) std::swap(Cs[q], Cs2[q]);
      outlet_pressure();
      velocity_bc_both();
      x.copy(p_old, P, nc);
      if (cproj)
        x.zero(gp_rc, 3 * size_t(nc));  // the exact projection takes the whole correction
      else
        correct_velocity(gdt, false, true, gp_rc);
      velocity_bc_both();
      gradients(false, 0.0, k < 2);
      x.copy(phi_prev, P, nc);
      x.copy(phi_prev_bnd, p_bnd, std::max(1, nb));
      {
        atol_for(gdt);
        int nmax = std::min(n_corr, diag_interface ? 8 : 3);
        const double stag = diag_interface ? 0.99 : 0.9;
        double dprev = 1e308, dphi_prev = 1e308;
        if (prm.nonortho_check) {
          nmax = std::min(n_corr, 8);
          x.zero(phi_it, nc);
        }
        for (int jno = 1; jno <= nmax; ++jno) {
          face_fluxes_and_rhs(gdt, jno > 1);
          if (jno == 1) x.zero(P, nc);
          if (pressure_solve(false, &sc->atol, 0.0) != 0) rc = -2;
          bc_pressure(1, skip_grad_p ? nullptr : gp, 0.0, nullptr);
          bc_pressure(2, nullptr, 0.0, nullptr);
          if (nmax == 1) break;
          SolverScalars h;
          x.download(&h, sc, 1);
          const double dcur = h.divmax;
          if (prm.nonortho_check) {
            const double dd = phi_change();
            if (jno > 1 && (dd < 1.0e-3 * last_pmax_ || dd > dphi_prev * stag)) break;
            dphi_prev = dd;
          } else if (jno > 1 && (dcur < h.atol || dcur > dprev * stag)) {
            break;
          }
          dprev = dcur;
        }
      }
      correct_face_fluxes(gdt, false);
      std::swap(UF, ff);
      u_face_ready = 1;
      if (cproj)
        project_exact(gdt);
      else
        correct_velocity(gdt, true, false, nullptr);
      velocity_bc_both();
      // rotational correction + accumulation + mean removal (the intermediate increment BC is
      // overwritten before anything reads it)
      x.template reduce<1, false>(nc, RotAccumulate{M, P, p_old, nu_sgs, rhs, prm.nu, gdt, prm.rotational},
                                  sk::AnchorFin{sc, n_outlet == 0});
      x.for_each(nc, ShiftP{P, &sc->shift});
      bc_pressure(1, skip_grad_p ? nullptr : gp, 1.0, outlet_pb);
      bc_pressure(2, nullptr, 1.0, nullptr);
      pressure_gradient();
      std::swap(H, H0);
      if (bous) std::swap(HT, HT0);
      for (int q = 0; q < ns; ++q) std::swap(HC[q], HC0[q]);
    }
    time = time_n;
    return rc;
  }

  // max|p - phi_it| and max|p|, then phi_it = p (nonortho_check exit test, :1495-1509)
  double last_pmax_ = 0.0;
  double phi_change() {
    std::vector<double> hp(nc), hi(nc);
    x.download(hp.data(), P, nc);
    x.download(hi.data(), phi_it, nc);
    double dd = 0.0, pm = 0.0;
    for (int c = 0; c < nc; ++c) {
      dd = std::fmax(dd, std::fabs(hp[c] - hi[c]));
      pm = std::fmax(pm, std::fabs(hp[c]));
    }
    x.copy(phi_it, P, nc);
    last_pmax_ = pm;
    return dd;
  }

  void stats(double out[8]) {
    x.template reduce<7, false>(nc, StatsCells{M, U, grad, nu_sgs, prm.nu}, sk::StatsFin7{sc});
    x.template reduce<1, true>(nc, StatsMax{U}, sk::StatsFinMax{sc});
    x.template reduce<2, false>(nb, StatsWall{M, tau}, sk::StatsFinWall{sc});
    x.template reduce<1, true>(nc, FluxDivergence{M, UF, 0}, sk::StatsFinDiv{sc, 10});
    x.template reduce<1, false>(nc, FluxDivergence{M, UF, 1}, sk::StatsFinDiv{sc, 11});
    SolverScalars h;
    x.download(&h, sc, 1);
    const double V = h.st[1];
    out[0] = h.st[0] / V;
    out[1] = 2.0 * h.st[2] / V;
    out[2] = h.st[7];
    const double a = h.st[3] / V, b = h.st[4] / V, c = h.st[5] / V;
    out[3] = std::sqrt(a * a + b * b + c * c);
    out[4] = (h.st[9] > SMALL) ? h.st[8] / h.st[9] : 0.0;
    out[5] = h.st[6] / V;
    out[6] = h.st[10];
    out[7] = h.st[11] / V;
  }

  // ---------------------------------------------------------------- checkpoint / restart
  // write_checkpoint / read_checkpoint (src/io/mod_checkpoint_io.f90:15-283): per-rank file
  // dir/processorR/checkpoint_XXXXXXXX.bin of Fortran unformatted sequential records (a 4-byte
  // length before and after each payload): header(8) int32 {nc, n_scalars, history = 0, step,
  // has_C, 0, has U_face, wall mask (1 tau_w, 2 q_w, 4 L)}, time, global ids (int64, 1..nc),
  // U (3, nc), p, T (T_ref when no temperature is solved), nu_sgs, C (n_scalars, nc) when there are
  // passive scalars, then U_face (only once the face flux is set: u_face_ready) and the wall-model
  // records the mask names. Rank 0 also writes dir/checkpoint_latest with the step.
  static void put_record(std::FILE* f, const void* p, size_t bytes) {
    if (bytes > 0x7fffffffull) throw std::runtime_error("checkpoint record above 2 GiB");
    const int32_t m = int32_t(bytes);
    if (std::fwrite(&m, 4, 1, f) != 1 || (bytes && std::fwrite(p, 1, bytes, f) != bytes) || std::fwrite(&m, 4, 1, f) != 1)
      throw std::runtime_error("checkpoint write failed");
  }
  static void get_record(std::FILE* f, void* p, size_t bytes) {
    int32_t m0 = 0, m1 = 0;
    if (std::fread(&m0, 4, 1, f) != 1) throw std::runtime_error("checkpoint: truncated file");
    if (size_t(m0) != bytes) throw std::runtime_error("checkpoint: record length does not match the mesh");
    if ((bytes && std::fread(p, 1, bytes, f) != bytes) || std::fread(&m1, 4, 1, f) != 1 || m1 != m0)
      throw std::runtime_error("checkpoint: corrupt record");
  }
  static std::string ckpt_name(const char* dir, int step, int rank) {
    char b[64];
    std::snprintf(b, sizeof(b), "/processor%d/checkpoint_%08d.bin", rank, step);
    return std::string(dir) + b;
  }
  int write_checkpoint(const char* dir, int step, int rank) {
    const std::string pdir = std::string(dir) + "/processor" + std::to_string(rank);
    if (std::system(("mkdir -p '" + pdir + "'").c_str()) != 0) throw std::runtime_error("cannot create " + pdir);
    std::FILE* f = std::fopen(ckpt_name(dir, step, rank).c_str(), "wb");
    if (!f) throw std::runtime_error("cannot open " + ckpt_name(dir, step, rank));
    const int nbb = std::max(1, nb);
    // header[7]: trailing wall records, 1 tau_w, 2 q_w, 4 L, 8 the wall-stress filter state
    const bool filt = wm_fmode && wm_filt_ready;
    int32_t header[8] = {nc, ns, 0, step, ns > 0 ? 1 : 0, 0, u_face_ready ? 1 : 0,
                         1 + (q_w ? 2 + 4 : 0) + (filt ? 8 : 0)};
    std::vector<double> h(3 * size_t(std::max(nc, nf)));
    std::vector<int64_t> gid(nc);
    for (int c = 0; c < nc; ++c) gid[c] = c + 1;
    try {
      put_record(f, header, sizeof(header));
      put_record(f, &time, 8);
      put_record(f, gid.data(), 8 * size_t(nc));
      x.download(h.data(), U, 3 * size_t(nc));
      put_record(f, h.data(), 24 * size_t(nc));
      x.download(h.data(), P, nc);
      put_record(f, h.data(), 8 * size_t(nc));
      if (bous) {
        x.download(h.data(), T, nc);
      } else {
        std::fill(h.begin(), h.begin() + nc, prm.T_ref);
      }
      put_record(f, h.data(), 8 * size_t(nc));
      x.download(h.data(), nu_sgs, nc);
      put_record(f, h.data(), 8 * size_t(nc));
      if (ns > 0) {  // C(1:ns, 1:nc): the scalar index fastest
        std::vector<double> hc(size_t(ns) * nc), one(nc);
        for (int q = 0; q < ns; ++q) {
          x.download(one.data(), Cs[q], nc);
          for (int c = 0; c < nc; ++c) hc[size_t(c) * ns + q] = one[c];
        }
        put_record(f, hc.data(), 8 * size_t(ns) * nc);
      }
      if (u_face_ready) {
        x.download(h.data(), UF, nf);
        put_record(f, h.data(), 8 * size_t(nf));
      }
      std::vector<double> hb(3 * size_t(nbb));
      x.download(hb.data(), tau, 3 * size_t(nbb));
      put_record(f, hb.data(), 24 * size_t(nb));
      if (q_w) {
        x.download(hb.data(), q_w, nbb);
        put_record(f, hb.data(), 8 * size_t(nb));
        x.download(hb.data(), L_bnd, nbb);
        put_record(f, hb.data(), 8 * size_t(nb));
      }
      if (filt) {
        x.download(hb.data(), wm_filt, 3 * size_t(nbb));
        put_record(f, hb.data(), 24 * size_t(nb));
      }
    } catch (...) {
      std::fclose(f);
      throw;
    }
    std::fclose(f);
    if (rank == 0) {
      std::FILE* m = std::fopen((std::string(dir) + "/checkpoint_latest").c_str(), "w");
      if (!m) throw std::runtime_error("cannot write checkpoint_latest");
      std::fprintf(m, "%d\n", step);
      std::fclose(m);
    }
    return 0;
  }
  // Returns the checkpoint's step. Restores the fields, then rebuilds boundary values and
  // gradients as solver_refresh_after_restart does (mod_fv_solver.f90:601-609).
  int read_checkpoint(const char* dir, int step, int rank) {
    std::FILE* f = std::fopen(ckpt_name(dir, step, rank).c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open " + ckpt_name(dir, step, rank));
    int32_t header[8];
    const int nbb = std::max(1, nb);
    std::vector<double> h(3 * size_t(std::max(nc, nf)));
    try {
      get_record(f, header, sizeof(header));
      if (header[0] != nc) throw std::runtime_error("checkpoint: n_cells does not match the mesh");
      if (header[1] != ns)
        throw std::runtime_error("checkpoint: n_scalars " + std::to_string(header[1]) + " does not match the solver's " +
                                 std::to_string(ns));
      get_record(f, &time, 8);
      std::vector<int64_t> gid(nc);
      get_record(f, gid.data(), 8 * size_t(nc));
      for (int c = 0; c < nc; ++c)
        if (gid[c] != c + 1) throw std::runtime_error("checkpoint: global_id mismatch (different partition)");
      get_record(f, h.data(), 24 * size_t(nc));
      x.upload(U, h.data(), 3 * size_t(nc));
      get_record(f, h.data(), 8 * size_t(nc));
      x.upload(P, h.data(), nc);
      get_record(f, h.data(), 8 * size_t(nc));
      if (bous) x.upload(T, h.data(), nc);
      get_record(f, h.data(), 8 * size_t(nc));
      x.upload(nu_sgs, h.data(), nc);
      if (header[1] > 0) {
        std::vector<double> hc(size_t(ns) * nc), one(nc);
        get_record(f, hc.data(), 8 * size_t(ns) * nc);
        for (int q = 0; q < ns; ++q) {
          for (int c = 0; c < nc; ++c) one[c] = hc[size_t(c) * ns + q];
          x.upload(Cs[q], one.data(), nc);
        }
      }
      if (header[6] == 1) {
        get_record(f, h.data(), 8 * size_t(nf));
        x.upload(UF, h.data(), nf);
        u_face_ready = 1;  // mod_checkpoint_io.f90:226-228
      }
      std::vector<double> hb(3 * size_t(nbb));
      if (header[7] & 1) {
        get_record(f, hb.data(), 24 * size_t(nb));
        x.upload(tau, hb.data(), 3 * size_t(nb));
      }
      for (int bit : {2, 4}) {
        if (!(header[7] & bit)) continue;
        get_record(f, hb.data(), 8 * size_t(nb));
        double* dst = bit == 2 ? q_w : L_bnd;
        if (dst) x.upload(dst, hb.data(), nb);
      }
      if (header[7] & 8) {  // the filter state is adopted as seeded (wall_model_set_filter_state)
        get_record(f, hb.data(), 24 * size_t(nb));
        x.upload(wm_filt, hb.data(), 3 * size_t(nb));
        wm_filt_ready = 1;
      }
      if (header[7] & 16) throw std::runtime_error("checkpoint: inflow-eddy records are not supported");
    } catch (...) {
      std::fclose(f);
      throw;
    }
    std::fclose(f);
    refresh_bc();
    gradients(false, 0.0);
    pressure_gradient();
    refresh_bc();
    gradients(false, 0.0);
    x.sync();
    return header[3];
  }

  // The cell-centred divergence of solver_divergence_metrics (mod_fv_solver.f90:3641-3713): the
  // velocity boundary values refreshed without gradients first, as there. out = {max |div|,
  // volume-averaged |div|}.
  void cell_divergence(double out[2]) {
    bc_velocity(0, nullptr);
    x.template reduce<1, true>(nc, CellDivergence{M, U, U_bnd, 0}, sk::StatsFinDiv{sc, 10});
    x.template reduce<1, false>(nc, CellDivergence{M, U, U_bnd, 1}, sk::StatsFinDiv{sc, 11});
    x.template reduce<7, false>(nc, StatsCells{M, U, grad, nu_sgs, prm.nu}, sk::StatsFin7{sc});
    SolverScalars h;
    x.download(&h, sc, 1);
    out[0] = h.st[10];
    out[1] = h.st[11] / h.st[1];
  }

  // One row of {prefix}_stats.csv (stats_sample, src/stats/mod_stats.f90:161-252): step, time, E,
  // eps, E/E0, eps_ref (= eps), Umax, face-flux and cell divergence (max, volume mean), Ub,
  // tauw_mean, Cf_bulk, nu_sgs_mean, nu_sgs/nu, dt; E0 <= 0 on entry takes this row's E. The
  // header is written when the file is new (append = 0 or the file is absent).
  int stats_csv(const char* path, int step, double dt, double* E0, int append) {
    double s8[8], cd[2];
    stats(s8);
    cell_divergence(cd);
    std::FILE* fp = append ? std::fopen(path, "r") : nullptr;
    const bool fresh = fp == nullptr;
    if (fp) std::fclose(fp);
    fp = std::fopen(path, fresh ? "w" : "a");
    if (!fp) return -1;
    if (fresh)
      std::fprintf(fp, "step,time,E,eps,E_over_E0,eps_ref,Umax,max_abs_divU,volavg_abs_divU,max_abs_divU_cell,"
                       "volavg_abs_divU_cell,Ub,tauw_mean,Cf_bulk,nu_sgs_mean,nu_sgs_over_nu,dt\n");
    if (!(*E0 > SMALL)) *E0 = s8[0];
    const double E_norm = *E0 > SMALL ? s8[0] / *E0 : 1.0;
    const double cf = s8[4] / std::fmax(0.5 * s8[3] * s8[3], SMALL);
    const double sgs = s8[5] / std::fmax(prm.nu, SMALL);
    const double v[16] = {time, s8[0], s8[1], E_norm, s8[1], s8[2], s8[6], s8[7], cd[0], cd[1],
                          s8[3], s8[4], cf, s8[5], sgs, dt};
    std::fprintf(fp, "%d", step);
    for (double q : v) std::fprintf(fp, ",%18.9E", q);  // Fortran ES18.9
    std::fprintf(fp, "\n");
    std::fclose(fp);
    return 0;
  }
};

}  // namespace dm
