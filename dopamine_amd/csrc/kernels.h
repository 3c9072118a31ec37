// Kernel bodies of the RK3W hot path, as functors (one call = one cell or one boundary face).
//
// Every face term is evaluated in gather form: a cell walks its own face list and adds each
// face's contribution with the sign of its side, in the same order as the reference's
// face-scatter loop would add it (ascending face order for the mesh-face loops, pressure-operator
// order for the projection loops). No atomics; each output is written once by its owner.
#pragma once

#include "dm_common.h"

namespace dm {

// grad(i,j) = d u_j / d x_i stored like the reference grad_U(3,3,nc): [9c + 3j + i]
#define DM_G(c, i, j) (9 * size_t(c) + 3 * (j) + (i))

struct DevMesh {
  int nc, nf, nint, nb, npf, wA, wB, wL;
  // cells
  const double *vol, *iv, *ctr;                   // ctr: 3 per cell
  const int *nA, *aface, *aoth;                   // mesh-face list, ELL wA x nc (slot-major)
  const int *nB, *bface, *both, *bidx;            // projection-face list, ELL wB x nc
  const int *nL, *lidx;                           // LS stencil, ELL wL x nc
  const double* lw;                               // LS weights: plane (3k+i) of nc values
  // faces
  const double *sf, *gc, *area, *nrm, *eps, *wbl, *fctr, *dcf;
  const int *fown, *fnb;                          // face owner / neighbour (-1 on boundary faces)
  // boundary faces
  const int *bbc, *bpatch, *bown, *pnb, *pair;
  const double *pgc, *pdcf, *cb, *kbnd, *wsy;
  const int* wsc;
  // pressure operator faces (interior + canonical periodic edges)
  const double *pcoef, *rcgc, *xgc, *kint;
  const int *pown, *pnbr, *pfm;
  // per-cell static operators
  const double *Tinv, *Gf;
  // patch values
  const double *patchU;
  // per patch: convective outlet (1), backflow-traction coefficient, MOST roughness lengths
  const int* pconv = nullptr;
  const double *pbeta = nullptr, *pz0m = nullptr, *pz0h = nullptr;
};

DM_INLINE double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// refine_blend_weight (src/mesh/mod_mesh.f90:491-500)
DM_INLINE double refine_blend_weight(double skew) {
  double t = (skew - 0.15) / (0.25 - 0.15);
  t = fmax(0.0, fmin(1.0, t));
  return 1.0 - t * t * (3.0 - 2.0 * t);
}

// ortho_weight (src/numerics/mod_viscous.f90:30-34)
DM_INLINE double ortho_weight(double s_dot_dhat, double area) {
  return area * fmax(0.0, fmin(1.0, s_dot_dhat / (area + SMALL)));
}

// nonortho_cross_flux (src/numerics/mod_nonortho.f90:120-128)
DM_INLINE double cross_flux(const double* go, const double* gn, double gc, const double* k) {
  double v[3];
  for (int i = 0; i < 3; ++i) v[i] = gc * go[i] + (1.0 - gc) * gn[i];
  return dot3(v, k);
}

// ------------------------------------------------------------------ boundary conditions

// apply_velocity_bc (src/bc/mod_bc.f90:196-379), one boundary face. filter: 0 all, 1 non-periodic,
// 2 periodic only. gU may be null (zeroth-order extrapolation).
struct BcVelocity {
  DevMesh M;
  const double *U, *gU;
  double* Ub;
  int filter;
  int conv_ready = 0;  // a convective outlet's boundary velocity is a state once a stage advanced it
  DM_HD void operator()(int b) const {
    const int f = M.nint + b, bc = M.bbc[b];
    if (filter == 2 && bc != BC_PERIODIC) return;
    if (filter == 1 && bc == BC_PERIODIC) return;
    const int c = M.bown[b];
    const double* n = &M.nrm[3 * size_t(f)];
    const double* uo = &U[3 * size_t(c)];
    double* ub = &Ub[3 * size_t(b)];
    switch (bc) {
      case BC_INLET: {
        const double* uv = &M.patchU[3 * size_t(M.bpatch[b])];
        ub[0] = uv[0]; ub[1] = uv[1]; ub[2] = uv[2];
        break;
      }
      case BC_OUTLET:
      case BC_ZEROGRADIENT: {
        // a convective outlet face keeps its RK state; only its start-up value is extrapolated
        // (src/bc/mod_bc.f90:304-311)
        if (bc == BC_OUTLET && conv_ready && M.pconv[M.bpatch[b]]) break;
        double u[3] = {uo[0], uo[1], uo[2]};
        if (gU) {
          const double d[3] = {M.fctr[3 * size_t(f)] - M.ctr[3 * size_t(c)], M.fctr[3 * size_t(f) + 1] - M.ctr[3 * size_t(c) + 1],
                               M.fctr[3 * size_t(f) + 2] - M.ctr[3 * size_t(c) + 2]};
          for (int k = 0; k < 3; ++k)
            u[k] = u[k] + gU[DM_G(c, 0, k)] * d[0] + gU[DM_G(c, 1, k)] * d[1] + gU[DM_G(c, 2, k)] * d[2];
        }
        ub[0] = u[0]; ub[1] = u[1]; ub[2] = u[2];
        break;
      }
      case BC_WALL_NOSLIP:
      case BC_WALL_MODEL:
        ub[0] = 0.0; ub[1] = 0.0; ub[2] = 0.0;
        break;
      case BC_WALL_VELOCITY: {
        const double* uv = &M.patchU[3 * size_t(M.bpatch[b])];
        const double un = uv[0] * n[0] + uv[1] * n[1] + uv[2] * n[2];
        for (int k = 0; k < 3; ++k) ub[k] = uv[k] - un * n[k];
        break;
      }
      case BC_PERIODIC: {
        double u[3] = {uo[0], uo[1], uo[2]};
        const int nb = M.pnb[b];
        if (nb >= 0) {
          const double w = M.pgc[b];
          for (int k = 0; k < 3; ++k) u[k] = w * u[k] + (1.0 - w) * U[3 * size_t(nb) + k];
        }
        ub[0] = u[0]; ub[1] = u[1]; ub[2] = u[2];
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
};

// apply_pressure_bc (src/bc/mod_bc.f90:383-486). phi = a + b (b null: phi = a) lets the same body
// serve the total pressure phi_prev + increment. gp may be null; pob null -> patch p_val (scale 0
// for an increment).
struct BcPressure {
  DevMesh M;
  const double *pa, *pb_add, *gp, *pob, *patchP;
  double* pbnd;
  double outlet_scale;
  int filter;
  DM_HD double val(int c) const { return pb_add ? pa[c] + pb_add[c] : pa[c]; }
  DM_HD void operator()(int b) const {
    const int f = M.nint + b, bc = M.bbc[b];
    if (filter == 2 && bc != BC_PERIODIC) return;
    if (filter == 1 && bc == BC_PERIODIC) return;
    const int o = M.bown[b];
    switch (bc) {
      case BC_OUTLET:
        pbnd[b] = outlet_scale * (pob ? pob[b] : patchP[M.bpatch[b]]);
        break;
      case BC_PERIODIC: {
        const int nb = M.pnb[b];
        pbnd[b] = (nb >= 0) ? M.pgc[b] * val(o) + (1.0 - M.pgc[b]) * val(nb) : val(o);
        break;
      }
      case BC_WALL_NOSLIP:
      case BC_WALL_VELOCITY:
      case BC_WALL_MODEL:
      case BC_SLIP:
      case BC_SYMMETRY: {
        // Kim-Moin Neumann with the default a_wall = 0
        const double* n = &M.nrm[3 * size_t(f)];
        const double an = 0.0 * n[0] + 0.0 * n[1] + 0.0 * n[2];
        double d[3];
        for (int i = 0; i < 3; ++i) d[i] = M.fctr[3 * size_t(f) + i] - M.ctr[3 * size_t(o) + i];
        const double d_ortho = fabs(d[0] * n[0] + d[1] * n[1] + d[2] * n[2]);
        pbnd[b] = val(o) - an * d_ortho;
        break;
      }
      default:
        if (gp) {
          double d[3];
          for (int i = 0; i < 3; ++i) d[i] = M.fctr[3 * size_t(f) + i] - M.ctr[3 * size_t(o) + i];
          pbnd[b] = val(o) + gp[3 * size_t(o)] * d[0] + gp[3 * size_t(o) + 1] * d[1] + gp[3 * size_t(o) + 2] * d[2];
        } else {
          pbnd[b] = val(o);
        }
    }
  }
};

// ------------------------------------------------------------------ gradients and SGS

// Vreman (src/sgs/mod_vreman.f90:50-153) and constant Smagorinsky (:295-364) for one cell.
DM_INLINE double det3_colmajor(const double* A) {
  return A[0] * (A[4] * A[8] - A[7] * A[5]) - A[3] * (A[1] * A[8] - A[7] * A[2]) + A[6] * (A[1] * A[5] - A[4] * A[2]);
}

DM_INLINE double vreman_nu(const double* al, const double* G, double c_v, double vol, double dt_stab) {
  double beta[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i <= j; ++i) {
      double b = 0.0;
      for (int n = 0; n < 3; ++n)
        for (int m = 0; m < 3; ++m) b = b + G[3 * n + m] * al[3 * i + m] * al[3 * j + n];
      beta[3 * j + i] = b;
    }
  double asq = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) asq = asq + al[3 * j + i] * al[3 * j + i];
  const double Bb = beta[0] * beta[4] - beta[3] * beta[3] + beta[0] * beta[8] - beta[6] * beta[6] +
                    beta[4] * beta[8] - beta[7] * beta[7];
  const double nu = (asq > SMALL && Bb > 0.0) ? c_v * sqrt(Bb / asq) : 0.0;
  const double cap = 0.4 * pow(vol, 2.0 / 3.0) / fmax(dt_stab, SMALL);
  return fmin(nu, cap);
}

DM_INLINE double smagorinsky_nu(const double* g, const double* G, double cs, double dt_stab) {
  double ss = 0.0;
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) {
      const double s = 0.5 * (g[3 * j + i] + g[3 * i + j]);
      ss = ss + s * s;
    }
  const double Sn = sqrt(2.0 * ss);
  const double delta = fmax(pow(fmax(det3_colmajor(G), 0.0), 1.0 / 6.0), SMALL);
  const double csd = cs * delta;
  const double nu = csd * csd * Sn;
  const double cap = 0.4 * delta * delta / fmax(dt_stab, SMALL);
  return fmin(nu, cap);
}

// amd_point_viscosity (src/sgs/mod_amd.f90:31-66) with the stability cap: al[3j+i] = alpha(i,j),
// G[3j+i] = G(i,j).
DM_INLINE double amd_nu(const double* al, const double* G, double c_amd, double vol, double dt_stab) {
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
  const double nu = (asq > 1.0e-20 && num > 1.0e-12 * trg * asq * sqrt(asq)) ? c_amd * num / asq : 0.0;
  const double cap = 0.4 * pow(vol, 2.0 / 3.0) / fmax(dt_stab, SMALL);
  return fmin(nu, cap);
}

DM_INLINE double sgs_nu(int sgs, const double* g, const double* G, double c_v, double cs, double c_amd, double vol,
                        double dt_stab) {
  if (sgs == 1) return vreman_nu(g, G, c_v, vol, dt_stab);
  if (sgs == 2) return smagorinsky_nu(g, G, cs, dt_stab);
  if (sgs == 3) return amd_nu(g, G, c_amd, vol, dt_stab);
  return 0.0;
}

// ls_gradient_vector (src/numerics/mod_gradient.f90:425-556), optionally followed by the SGS
// viscosity of the same cell (sgs: 0 none, 1 Vreman, 2 Smagorinsky, 3 AMD, -1 leave nu_sgs alone).
struct LsGradVec {
  DevMesh M;
  const double *U, *Ub;
  double *grad, *nu_sgs;
  int sgs;
  double c_v, cs, c_amd, dt_stab;
  DM_HD void operator()(int c) const {
    const int nc = M.nc;
    const double u0 = U[3 * size_t(c)], u1 = U[3 * size_t(c) + 1], u2 = U[3 * size_t(c) + 2];
    double g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int cnt = M.nL[c];
    for (int k = 0; k < cnt; ++k) {
      const size_t s = size_t(k) * nc + c;
      const int j = M.lidx[s];
      const double* un = (j >= 0) ? &U[3 * size_t(j)] : &Ub[3 * size_t(-1 - j)];
      const double du[3] = {un[0] - u0, un[1] - u1, un[2] - u2};
      const double w0 = M.lw[(3 * size_t(k)) * nc + c], w1 = M.lw[(3 * size_t(k) + 1) * nc + c],
                   w2 = M.lw[(3 * size_t(k) + 2) * nc + c];
      for (int q = 0; q < 3; ++q) {
        g[3 * q + 0] = g[3 * q + 0] + w0 * du[q];
        g[3 * q + 1] = g[3 * q + 1] + w1 * du[q];
        g[3 * q + 2] = g[3 * q + 2] + w2 * du[q];
      }
    }
    for (int q = 0; q < 9; ++q) grad[9 * size_t(c) + q] = g[q];
    if (sgs >= 0) nu_sgs[c] = sgs_nu(sgs, g, &M.Gf[9 * size_t(c)], c_v, cs, c_amd, M.vol[c], dt_stab);
  }
};

// compute_all_gradients with least squares (src/solver/mod_fv_solver.f90:1003-1077): the velocity
// gradient (+ SGS viscosity, as LsGradVec) and NSF scalar gradients (temperature, passive scalars,
// as LsGradSca) in one pass over the stencil, so its indices and weights are read once. Each field
// accumulates in the same order as its own kernel: the results are those kernels' bit for bit.
constexpr int LS_MAX_SCALARS = 5;
template <int NSF>
struct LsGradAll {
  // the velocity-only form fits five wavefronts per SIMD without spilling (2% faster measured);
  // the forms with scalars would spill
  static constexpr int kWaves = NSF == 0 ? 5 : 0;
  DevMesh M;
  const double *U, *Ub;
  double *grad, *nu_sgs;
  int sgs;
  double c_v, cs, c_amd, dt_stab;
  const double* sa[LS_MAX_SCALARS];  // scalar fields, their boundary values, their gradients
  const double* sb[LS_MAX_SCALARS];
  double* sg[LS_MAX_SCALARS];
  DM_HD void operator()(int c) const {
    const int nc = M.nc;
    const double u0 = U[3 * size_t(c)], u1 = U[3 * size_t(c) + 1], u2 = U[3 * size_t(c) + 2];
    double g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double pc[NSF > 0 ? NSF : 1], gs[NSF > 0 ? NSF : 1][3];
#pragma unroll
    for (int i = 0; i < NSF; ++i) {
      pc[i] = sa[i][c];
      gs[i][0] = gs[i][1] = gs[i][2] = 0.0;
    }
    const int cnt = M.nL[c];
    for (int k = 0; k < cnt; ++k) {
      const size_t s = size_t(k) * nc + c;
      const int j = M.lidx[s];
      const double* un = (j >= 0) ? &U[3 * size_t(j)] : &Ub[3 * size_t(-1 - j)];
      const double du[3] = {un[0] - u0, un[1] - u1, un[2] - u2};
      const double w0 = M.lw[(3 * size_t(k)) * nc + c], w1 = M.lw[(3 * size_t(k) + 1) * nc + c],
                   w2 = M.lw[(3 * size_t(k) + 2) * nc + c];
      for (int q = 0; q < 3; ++q) {
        g[3 * q + 0] = g[3 * q + 0] + w0 * du[q];
        g[3 * q + 1] = g[3 * q + 1] + w1 * du[q];
        g[3 * q + 2] = g[3 * q + 2] + w2 * du[q];
      }
#pragma unroll
      for (int i = 0; i < NSF; ++i) {
        const double d = ((j >= 0) ? sa[i][j] : sb[i][-1 - j]) - pc[i];
        gs[i][0] = gs[i][0] + w0 * d;
        gs[i][1] = gs[i][1] + w1 * d;
        gs[i][2] = gs[i][2] + w2 * d;
      }
    }
    for (int q = 0; q < 9; ++q) grad[9 * size_t(c) + q] = g[q];
    if (sgs >= 0) nu_sgs[c] = sgs_nu(sgs, g, &M.Gf[9 * size_t(c)], c_v, cs, c_amd, M.vol[c], dt_stab);
#pragma unroll
    for (int i = 0; i < NSF; ++i)
      for (int q = 0; q < 3; ++q) sg[i][3 * size_t(c) + q] = gs[i][q];
  }
};

// ls_gradient_scalar (src/numerics/mod_gradient.f90:307-422) of phi = a (+ a2) with boundary
// values pb (+ pb2).
struct LsGradSca {
  DevMesh M;
  const double *a, *a2, *pb, *pb2;
  double* g;
  DM_HD double cv(int c) const { return a2 ? a[c] + a2[c] : a[c]; }
  DM_HD double bv(int b) const { return pb2 ? pb[b] + pb2[b] : pb[b]; }
  DM_HD void operator()(int c) const {
    const int nc = M.nc;
    const double pc = cv(c);
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    const int cnt = M.nL[c];
    for (int k = 0; k < cnt; ++k) {
      const size_t s = size_t(k) * nc + c;
      const int j = M.lidx[s];
      const double d = ((j >= 0) ? cv(j) : bv(-1 - j)) - pc;
      g0 = g0 + M.lw[(3 * size_t(k)) * nc + c] * d;
      g1 = g1 + M.lw[(3 * size_t(k) + 1) * nc + c] * d;
      g2 = g2 + M.lw[(3 * size_t(k) + 2) * nc + c] * d;
    }
    g[3 * size_t(c)] = g0;
    g[3 * size_t(c) + 1] = g1;
    g[3 * size_t(c) + 2] = g2;
  }
};

// LsGradSca on a list of cells only.
struct LsGradScaAt {
  LsGradSca g;
  const int* cells;
  DM_HD void operator()(int i) const { g(cells[i]); }
};

// green_gauss_gradient_vector (src/numerics/mod_gradient.f90:172-304), gprev null on the first sweep.
struct GgGradVec {
  DevMesh M;
  const double *U, *Ub, *gprev;
  double* grad;
  DM_HD void operator()(int c) const {
    const int nc = M.nc;
    double g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int cnt = M.nA[c];
    for (int k = 0; k < cnt; ++k) {
      const size_t s = size_t(k) * nc + c;
      const int e = M.aface[s], f = e >> 1, side = e & 1, oth = M.aoth[s];
      const double* sf = &M.sf[3 * size_t(f)];
      double Uf[3];
      if (oth >= 0) {
        const int o = side ? oth : c, nb = side ? c : oth;
        const double gc = M.gc[f];
        if (gprev) {
          const double* ep = &M.eps[3 * size_t(f)];
          for (int q = 0; q < 3; ++q) {
            double gm[3];
            for (int i = 0; i < 3; ++i) gm[i] = gc * gprev[DM_G(o, i, q)] + (1.0 - gc) * gprev[DM_G(nb, i, q)];
            Uf[q] = gc * U[3 * size_t(o) + q] + (1.0 - gc) * U[3 * size_t(nb) + q] + dot3(gm, ep);
          }
          const double w = M.wbl[f];
          if (w < 1.0)
            for (int q = 0; q < 3; ++q) {
              const double pl = gc * U[3 * size_t(o) + q] + (1.0 - gc) * U[3 * size_t(nb) + q];
              Uf[q] = pl + w * (Uf[q] - pl);
            }
        } else {
          for (int q = 0; q < 3; ++q) Uf[q] = gc * U[3 * size_t(o) + q] + (1.0 - gc) * U[3 * size_t(nb) + q];
        }
      } else {
        const int b = -1 - oth;
        for (int q = 0; q < 3; ++q) Uf[q] = Ub[3 * size_t(b) + q];
      }
      for (int q = 0; q < 3; ++q)
        for (int i = 0; i < 3; ++i) g[3 * q + i] = side ? g[3 * q + i] - Uf[q] * sf[i] : g[3 * q + i] + Uf[q] * sf[i];
    }
    if (M.vol[c] > SMALL) {
      const double iv = 1.0 / M.vol[c];
      for (int q = 0; q < 9; ++q) g[q] = g[q] * iv;
    }
    for (int q = 0; q < 9; ++q) grad[9 * size_t(c) + q] = g[q];
  }
};

// green_gauss_gradient_scalar (src/numerics/mod_gradient.f90:38-168) of phi = a (+ a2).
struct GgGradSca {
  DevMesh M;
  const double *a, *a2, *pb, *pb2, *gprev;
  double* g;
  DM_HD double cv(int c) const { return a2 ? a[c] + a2[c] : a[c]; }
  DM_HD void operator()(int c) const {
    const int nc = M.nc;
    double r[3] = {0, 0, 0};
    const int cnt = M.nA[c];
    for (int k = 0; k < cnt; ++k) {
      const size_t s = size_t(k) * nc + c;
      const int e = M.aface[s], f = e >> 1, side = e & 1, oth = M.aoth[s];
      const double* sf = &M.sf[3 * size_t(f)];
      double pf;
      if (oth >= 0) {
        const int o = side ? oth : c, nb = side ? c : oth;
        const double gc = M.gc[f];
        if (gprev) {
          double gm[3];
          for (int i = 0; i < 3; ++i) gm[i] = gc * gprev[3 * size_t(o) + i] + (1.0 - gc) * gprev[3 * size_t(nb) + i];
          const double pl = gc * cv(o) + (1.0 - gc) * cv(nb);
          pf = pl + dot3(gm, &M.eps[3 * size_t(f)]);
          const double w = M.wbl[f];
          if (w < 1.0) pf = pl + w * (pf - pl);
        } else {
          pf = gc * cv(o) + (1.0 - gc) * cv(nb);
        }
      } else {
        const int b = -1 - oth;
        pf = pb2 ? pb[b] + pb2[b] : pb[b];
      }
      for (int i = 0; i < 3; ++i) r[i] = side ? r[i] - pf * sf[i] : r[i] + pf * sf[i];
    }
    if (M.vol[c] > SMALL) {
      const double iv = 1.0 / M.vol[c];
      for (int i = 0; i < 3; ++i) r[i] = r[i] * iv;
    }
    for (int i = 0; i < 3; ++i) g[3 * size_t(c) + i] = r[i];
  }
};

// SGS viscosity alone (after a Green-Gauss gradient).
struct SgsCell {
  DevMesh M;
  const double* grad;
  double* nu_sgs;
  int sgs;
  double c_v, cs, c_amd, dt_stab;
  DM_HD void operator()(int c) const {
    nu_sgs[c] = sgs_nu(sgs, &grad[9 * size_t(c)], &M.Gf[9 * size_t(c)], c_v, cs, c_amd, M.vol[c], dt_stab);
  }
};

// ------------------------------------------------------------------ dynamic Smagorinsky
// dynamic_smagorinsky_sgs (src/sgs/mod_dynamic_smagorinsky.f90:39-186): Germano-Lilly with a
// compact top-hat test filter, Cs^2 = <L_ij M_ij> / <M_ij M_ij> clipped at 0, ALPHA_SQ = 4.

// u, the six u_i u_j and the six |S| S_ij of a cell, 15 values per cell
struct DynPack {
  const double *U, *grad;
  double* fb;
  DM_HD void operator()(int c) const {
    const double* u = &U[3 * size_t(c)];
    const double* g = &grad[9 * size_t(c)];
    double* f = &fb[15 * size_t(c)];
    const double S11 = g[0], S22 = g[4], S33 = g[8];
    const double S12 = 0.5 * (g[3] + g[1]), S13 = 0.5 * (g[6] + g[2]), S23 = 0.5 * (g[7] + g[5]);
    const double Sn = sqrt(2.0 * (S11 * S11 + S22 * S22 + S33 * S33 + 2.0 * (S12 * S12 + S13 * S13 + S23 * S23)));
    f[0] = u[0]; f[1] = u[1]; f[2] = u[2];
    f[3] = u[0] * u[0]; f[4] = u[0] * u[1]; f[5] = u[0] * u[2];
    f[6] = u[1] * u[1]; f[7] = u[1] * u[2]; f[8] = u[2] * u[2];
    f[9] = Sn * S11; f[10] = Sn * S12; f[11] = Sn * S13; f[12] = Sn * S22; f[13] = Sn * S23; f[14] = Sn * S33;
  }
};

// test_filter (:238-304): the volume-weighted mean of a cell and its face neighbours (periodic
// faces through their partner cell) over NC packed values; u3 (optional) gets values 0..2 as a
// 3-per-cell vector
template <int NC>
struct DynFilter {
  DevMesh M;
  const double* in;
  double *out, *u3;
  DM_HD void operator()(int c) const {
    const int nc = M.nc;
    double acc[NC];
    double wsum = M.vol[c];
#pragma unroll
    for (int k = 0; k < NC; ++k) acc[k] = wsum * in[NC * size_t(c) + k];
    const int cnt = M.nA[c];
    for (int q = 0; q < cnt; ++q) {
      const int oth = M.aoth[size_t(q) * nc + c];
      const int nb = oth >= 0 ? oth : M.pnb[-1 - oth];
      if (nb < 0 || nb == c) continue;
      const double w = M.vol[nb];
      wsum = wsum + w;
#pragma unroll
      for (int k = 0; k < NC; ++k) acc[k] = acc[k] + w * in[NC * size_t(nb) + k];
    }
#pragma unroll
    for (int k = 0; k < NC; ++k) out[NC * size_t(c) + k] = acc[k] / fmax(wsum, SMALL);
    if (u3)
      for (int k = 0; k < 3; ++k) u3[3 * size_t(c) + k] = out[NC * size_t(c) + k];
  }
};

// germano_lilly_terms (:203-236): L_ij M_ij and M_ij M_ij of a cell from the filtered values, the
// gradient of the filtered velocity and Delta = det(G)^(1/6)
struct DynLM {
  DevMesh M;
  const double *fh, *gh;
  double* lm;
  DM_HD void operator()(int c) const {
    const double* h = &fh[15 * size_t(c)];
    const double* g = &gh[9 * size_t(c)];
    double Sh[9], ss = 0.0;
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) Sh[3 * j + i] = 0.5 * (g[3 * j + i] + g[3 * i + j]);
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) ss = ss + Sh[3 * j + i] * Sh[3 * j + i];
    const double Shn = sqrt(2.0 * ss);
    const double L11 = h[3] - h[0] * h[0], L12 = h[4] - h[0] * h[1], L13 = h[5] - h[0] * h[2];
    const double L22 = h[6] - h[1] * h[1], L23 = h[7] - h[1] * h[2], L33 = h[8] - h[2] * h[2];
    const double d = fmax(pow(fmax(det3_colmajor(&M.Gf[9 * size_t(c)]), 0.0), 1.0 / 6.0), SMALL);
    const double d2 = d * d;
    const double M11 = d2 * (2.0 * (h[9] - 4.0 * Shn * Sh[0])), M12 = d2 * (2.0 * (h[10] - 4.0 * Shn * Sh[3]));
    const double M13 = d2 * (2.0 * (h[11] - 4.0 * Shn * Sh[6])), M22 = d2 * (2.0 * (h[12] - 4.0 * Shn * Sh[4]));
    const double M23 = d2 * (2.0 * (h[13] - 4.0 * Shn * Sh[7])), M33 = d2 * (2.0 * (h[14] - 4.0 * Shn * Sh[8]));
    lm[2 * size_t(c)] = L11 * M11 + L22 * M22 + L33 * M33 + 2.0 * (L12 * M12 + L13 * M13 + L23 * M23);
    lm[2 * size_t(c) + 1] = M11 * M11 + M22 * M22 + M33 * M33 + 2.0 * (M12 * M12 + M13 * M13 + M23 * M23);
  }
};

// nu_t = max(0, <LM> / <MM>) Delta^2 |S|, capped by 0.4 Delta^2 / dt_stab
struct DynNu {
  DevMesh M;
  const double *lmh, *grad;
  double* nu_sgs;
  double dt_stab;
  DM_HD void operator()(int c) const {
    const double* g = &grad[9 * size_t(c)];
    double ss = 0.0;
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) {
        const double sij = 0.5 * (g[3 * j + i] + g[3 * i + j]);
        ss = ss + sij * sij;
      }
    const double d = fmax(pow(fmax(det3_colmajor(&M.Gf[9 * size_t(c)]), 0.0), 1.0 / 6.0), SMALL);
    const double nu = fmax(lmh[2 * size_t(c)] / fmax(lmh[2 * size_t(c) + 1], SMALL), 0.0) * (d * d) * sqrt(2.0 * ss);
    nu_sgs[c] = fmin(nu, 0.4 * (d * d) / fmax(dt_stab, SMALL));
  }
};

// ------------------------------------------------------------------ wall model

// Reichardt law and Newton solve for u_tau (src/sgs/mod_wall_model.f90:50-115).
DM_INLINE double wall_utau(double u_par, double y_w, double nu) {
  const double KAPPA = 0.41, BLOG = 5.2, A1I = 1.0 / 11.0, A2 = 0.33;
  const double cbuf = BLOG - log(KAPPA) / KAPPA;
  if (u_par <= SMALL || y_w <= SMALL) return 0.0;
  double ut = fmax(0.01 * u_par, sqrt(nu * u_par / (y_w + SMALL)));
  const double ywn = y_w * (1.0 / (nu + SMALL));
  for (int it = 0; it < 50; ++it) {
    const double yp = ut * ywn;
    const double e1 = exp(-yp * A1I), e2 = exp(-A2 * yp);
    const double up = log(1.0 + KAPPA * yp) / KAPPA + cbuf * (1.0 - e1 - yp * A1I * e2);
    const double dup = 1.0 / (1.0 + KAPPA * yp) + cbuf * (A1I * e1 - A1I * e2 + yp * A1I * A2 * e2);
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

DM_INLINE double wall_tau_mag(double u_par, double y_w, double nu) {
  const double ut = wall_utau(u_par, y_w, nu);
  return ut * ut;
}

// AMD in wall-model cells (src/sgs/mod_amd.f90:142-199): the cell's alpha with the wall-parallel
// part of its wall-normal derivative replaced by u_tau / (kappa y), for each wall-model face of the
// cell in boundary order; the last valid face's value stands. One wall cell per item.
struct AmdWall {
  DevMesh M;
  const int *cells, *off, *bfaces;
  const double *U, *grad;
  double* nu_sgs;
  double nu, c_amd, dt_stab;
  DM_HD void operator()(int i) const {
    const int o = cells[i];
    for (int q = off[i]; q < off[i + 1]; ++q) {
      const int b = bfaces[q], f = M.nint + b;
      double nrm[3], dx[3];
      for (int k = 0; k < 3; ++k) {
        nrm[k] = -M.nrm[3 * size_t(f) + k];
        dx[k] = M.fctr[3 * size_t(f) + k] - M.ctr[3 * size_t(o) + k];
      }
      const double y_w = fabs(dot3(dx, nrm));
      const double* u = &U[3 * size_t(o)];
      const double un = dot3(u, nrm);
      double ut[3];
      for (int k = 0; k < 3; ++k) ut[k] = u[k] - un * nrm[k];
      const double ut_mag = sqrt(ut[0] * ut[0] + ut[1] * ut[1] + ut[2] * ut[2]);
      if (ut_mag <= 1.0e-14 || y_w <= SMALL) continue;
      const double g_wall = wall_utau(ut_mag, y_w, nu) / (0.41 * y_w);
      double al[9], d_old[3], d_t[3];
      for (int k = 0; k < 9; ++k) al[k] = grad[9 * size_t(o) + k];
      for (int j = 0; j < 3; ++j) d_old[j] = nrm[0] * al[3 * j + 0] + nrm[1] * al[3 * j + 1] + nrm[2] * al[3 * j + 2];
      const double dn = dot3(d_old, nrm);
      for (int j = 0; j < 3; ++j) d_t[j] = g_wall * ut[j] / ut_mag - (d_old[j] - dn * nrm[j]);
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) al[3 * j + k] = al[3 * j + k] + nrm[k] * d_t[j];
      nu_sgs[o] = amd_nu(al, &M.Gf[9 * size_t(o)], c_amd, M.vol[o], dt_stab);
    }
  }
};

// apply_wall_model_bc (src/sgs/mod_wall_model.f90:176-313), one boundary face: the wall-parallel
// velocity of the sampling cell, optionally time-filtered (fmode 1: u_f += w (u - u_f) with
// w = dt / (T + dt); fmode 2: eps = min(1, dt kappa u_tau(|u_f|) / (f y)), eps = 0 when dt = 0;
// seed: the first filtered call copies u), then the Reichardt law.
struct WallModel {
  DevMesh M;
  const double* U;
  double* tau;
  double nu;
  double* uf = nullptr;  // filtered velocity, 3 per boundary face
  int fmode = 0, seed = 0;
  double w = 0.0, dt_use = 0.0, tf = 0.0;
  DM_HD void operator()(int b) const {
    if (M.bbc[b] != BC_WALL_MODEL) return;
    const int f = M.nint + b;
    const double* n = &M.nrm[3 * size_t(f)];
    const double* u = &U[3 * size_t(M.wsc[b])];
    const double y = M.wsy[b];
    const double un = u[0] * n[0] + u[1] * n[1] + u[2] * n[2];
    double up[3] = {u[0] - un * n[0], u[1] - un * n[1], u[2] - un * n[2]};
    if (fmode) {
      double* q = &uf[3 * size_t(b)];
      if (seed) {
        for (int k = 0; k < 3; ++k) q[k] = up[k];
      } else {
        double eps = w;
        if (fmode == 2) {
          eps = 0.0;
          if (dt_use > 0.0) {
            const double prev = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
            const double ut = wall_utau(prev, y, nu);
            eps = 1.0;
            if (ut > SMALL) eps = fmin(1.0, dt_use * 0.41 * ut / (tf * y));  // kappa = 0.41
          }
        }
        for (int k = 0; k < 3; ++k) q[k] = q[k] + eps * (up[k] - q[k]);
      }
      for (int k = 0; k < 3; ++k) up[k] = q[k];
    }
    const double mag = sqrt(up[0] * up[0] + up[1] * up[1] + up[2] * up[2]);
    const double t = wall_tau_mag(mag, y, nu);
    const double scl = -t / (mag + SMALL);
    for (int k = 0; k < 3; ++k) tau[3 * size_t(b) + k] = scl * up[k];
  }
};

// ------------------------------------------------------------------ convective outlet
// outlet_advance_boundary / update_outlet_pressure (src/bc/mod_bc.f90:681-818).

// area sums of u.n, |u| and the area over one outlet patch's faces (boundary faces start..)
struct OutletStat {
  DevMesh M;
  const double* U;
  int start;
  DM_HD Red<3> operator()(int i) const {
    const int b = start + i, f = M.nint + b, o = M.bown[b];
    const double* u = &U[3 * size_t(o)];
    const double a = M.area[f];
    Red<3> r;
    r.v[0] = a * dot3(u, &M.nrm[3 * size_t(f)]);
    r.v[1] = a * sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    r.v[2] = a;
    return r;
  }
};
// sp[2q] = max(mean u.n, 0) (the convection speed), sp[2q+1] = mean |u| (the reference speed)
struct OutletStatFin {
  double* sp;
  int q;
  DM_HD void operator()(const Red<3>& s) const {
    if (s.v[2] <= 0.0) return;
    sp[2 * q + 1] = s.v[1] / s.v[2];
    sp[2 * q] = fmax(s.v[0] / s.v[2], 0.0);
  }
};

// outlet_dn_u (src/bc/mod_bc.f90:775-786)
DM_INLINE double outlet_dn_u(double ub, double uo, double gdd, double gt, double len_d, double nd) {
  return nd * (2.0 * (ub - uo) / len_d - gdd) + gt;
}

// one RK3W stage of du_b/dt = -U_c n.grad u on a convective outlet face
struct OutletAdvance {
  DevMesh M;
  const double *U, *grad, *sp;
  double *Ub, *gprev;
  double dt, a, bk;
  DM_HD void operator()(int b) const {
    if (M.bbc[b] != BC_OUTLET) return;
    const int q = M.bpatch[b];
    if (!M.pconv[q]) return;
    const int f = M.nint + b, o = M.bown[b];
    double dl[3];
    for (int i = 0; i < 3; ++i) dl[i] = M.fctr[3 * size_t(f) + i] - M.ctr[3 * size_t(o) + i];
    const double len_d = fmax(sqrt(dot3(dl, dl)), SMALL);
    for (int i = 0; i < 3; ++i) dl[i] = dl[i] / len_d;
    const double* n = &M.nrm[3 * size_t(f)];
    const double nd = dot3(n, dl);
    double tn[3];
    for (int i = 0; i < 3; ++i) tn[i] = n[i] - nd * dl[i];
    const double uc = sp[2 * q];
    for (int k = 0; k < 3; ++k) {
      const double gdd = grad[DM_G(o, 0, k)] * dl[0] + grad[DM_G(o, 1, k)] * dl[1] + grad[DM_G(o, 2, k)] * dl[2];
      const double gt = grad[DM_G(o, 0, k)] * tn[0] + grad[DM_G(o, 1, k)] * tn[1] + grad[DM_G(o, 2, k)] * tn[2];
      const size_t j = 3 * size_t(b) + k;
      const double g = -uc * outlet_dn_u(Ub[j], U[3 * size_t(o) + k], gdd, gt, len_d, nd);
      Ub[j] = Ub[j] + dt * (a * g + bk * gprev[j]);
      gprev[j] = g;
    }
  }
};

// backflow traction p_b = p_val - beta |u|^2 S0 / 2, S0 = (1 - tanh(u.n / (U_ref 0.05))) / 2
struct OutletPressure {
  DevMesh M;
  const double *U, *sp, *patchP;
  double* pob;
  DM_HD void operator()(int b) const {
    if (M.bbc[b] != BC_OUTLET) return;
    const int q = M.bpatch[b];
    const double beta = M.pbeta[q];
    if (beta <= 0.0) return;
    const int f = M.nint + b, o = M.bown[b];
    const double* u = &U[3 * size_t(o)];
    const double un = dot3(u, &M.nrm[3 * size_t(f)]);
    const double u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const double ur = fmax(sp[2 * q + 1], SMALL);
    const double s0 = 0.5 * (1.0 - tanh(un / (ur * 0.05)));
    pob[b] = patchP[q] - beta * 0.5 * u2 * s0;
  }
};

// ------------------------------------------------------------------ MOST surface layer
// src/abl/mod_abl_wall.f90 and src/abl/mod_abl_profiles.f90 (Businger-Dyer corrections).
constexpr double MOST_KAPPA = 0.41, MOST_G = 9.81, MOST_L_CLIP = 5000.0;

// abl_psi_m (mod_abl_profiles.f90:51-68): stable -5 zeta, unstable Paulson (1970)
DM_INLINE double most_psi_m(double zeta) {
  if (fabs(zeta) < 1.0e-6) return 0.0;
  if (zeta > 0.0) return -5.0 * zeta;
  const double x = pow(1.0 - 16.0 * zeta, 0.25);
  return 2.0 * log((1.0 + x) / 2.0) + log((1.0 + x * x) / 2.0) - 2.0 * atan(x) + 1.5707963267948966;
}
// abl_psi_h (:71-85)
DM_INLINE double most_psi_h(double zeta) {
  if (fabs(zeta) < 1.0e-6) return 0.0;
  if (zeta > 0.0) return -5.0 * zeta;
  return 2.0 * log((1.0 + sqrt(1.0 - 16.0 * zeta)) / 2.0);
}
// most_residuals (mod_abl_wall.f90:260-284)
// (log(y / z0m) and log(y / z0h), fixed along the iteration, come in as lym and lyh)
DM_INLINE void most_residuals(double us, double ts, double L, double u_par, double T_o, double T_s, double y,
                              double z0m, double z0h, double lym, double lyh, double* Fu, double* FT) {
  double dm = lym - most_psi_m(y / L) + most_psi_m(z0m / L);
  double dh = lyh - most_psi_h(y / L) + most_psi_h(z0h / L);
  if (fabs(dm) < SMALL) dm = copysign(SMALL, dm);
  if (fabs(dh) < SMALL) dh = copysign(SMALL, dh);
  *Fu = us - MOST_KAPPA * u_par / dm;
  *FT = ts - MOST_KAPPA * (T_o - T_s) / dh;
}
// L_from (:226-240)
DM_INLINE double most_L_from(double us, double ts, double T_ref) {
  if (!(fabs(ts) > SMALL)) return MOST_L_CLIP;
  double L = us * us * T_ref / (MOST_KAPPA * MOST_G * ts);
  L = copysign(fmax(fabs(L), 0.001), L);
  return fabs(L) > MOST_L_CLIP ? copysign(MOST_L_CLIP, L) : L;
}
// most_iterate (:155-223): relaxed Newton (factor 0.5) on (Fu, FT), finite-difference Jacobian
// through L(u*, T*), at most 20 iterations, stop when |Fu| and |FT| < 1e-8.
DM_INLINE void most_iterate(double u_par, double T_o, double T_s, double y, double z0m, double z0h, double L_init,
                            double* us_out, double* ts_out, double* L_out) {
  const double lym = log(y / z0m), lyh = log(y / z0h);
  double us = fmax(SMALL, MOST_KAPPA * u_par / fmax(SMALL, lym));
  double ts = 0.1 * (T_o - T_s);
  double L = fabs(L_init) < SMALL ? MOST_L_CLIP : L_init;
  for (int it = 0; it < 20; ++it) {
    L = copysign(fmax(fabs(L), 0.001), L);
    if (fabs(L) > MOST_L_CLIP) L = copysign(MOST_L_CLIP, L);
    double Fu, FT, Fp, Tp;
    most_residuals(us, ts, L, u_par, T_o, T_s, y, z0m, z0h, lym, lyh, &Fu, &FT);
    if (fabs(Fu) < 1.0e-8 && fabs(FT) < 1.0e-8) break;
    const double du = fmax(1.0e-6, 1.0e-6 * fabs(us)), dT = fmax(1.0e-6, 1.0e-6 * fabs(ts));
    most_residuals(us + du, ts, most_L_from(us + du, ts, T_o), u_par, T_o, T_s, y, z0m, z0h, lym, lyh, &Fp, &Tp);
    const double a11 = (Fp - Fu) / du, a21 = (Tp - FT) / du;
    most_residuals(us, ts + dT, most_L_from(us, ts + dT, T_o), u_par, T_o, T_s, y, z0m, z0h, lym, lyh, &Fp, &Tp);
    const double a12 = (Fp - Fu) / dT, a22 = (Tp - FT) / dT;
    const double den = a11 * a22 - a12 * a21;
    if (fabs(den) < SMALL) break;
    const double dus = -(a22 * Fu - a12 * FT) / den;
    const double dts = -(a11 * FT - a21 * Fu) / den;
    us = us + 0.5 * dus;
    ts = ts + 0.5 * dts;
    us = fmax(SMALL, fmin(us, 0.5 * u_par));
    L = most_L_from(us, ts, T_o);
  }
  *us_out = us;
  *ts_out = ts;
  *L_out = L;
}

// abl_wall_model (mod_abl_wall.f90:64-153), one boundary face: the MOST fluxes from the owner
// cell, tau_w = -u*^2 u_par/|u_par|, q_w = -u* T*, the face's L kept for its next solve. Partials
// (L, 1) for the domain-mean L.
struct MostWall {
  DevMesh M;
  const double *U, *T;
  double T_const, Ts;
  const double* Lcfg;
  double *Lb, *tau, *qw;
  DM_HD Red<2> operator()(int b) const {
    Red<2> r;
    r.v[0] = r.v[1] = 0.0;
    if (M.bbc[b] != BC_WALL_MODEL) return r;
    const int f = M.nint + b, o = M.bown[b];
    const double* n = &M.nrm[3 * size_t(f)];
    const double* u = &U[3 * size_t(o)];
    const double T_o = T ? T[o] : T_const;
    const double un = u[0] * n[0] + u[1] * n[1] + u[2] * n[2];
    double up[3] = {u[0] - un * n[0], u[1] - un * n[1], u[2] - un * n[2]};
    const double mag = sqrt(up[0] * up[0] + up[1] * up[1] + up[2] * up[2]);
    for (int k = 0; k < 3; ++k) up[k] = mag > SMALL ? up[k] / mag : 0.0;
    double Lg = Lb[b];
    if (fabs(Lg) < SMALL) Lg = *Lcfg;
    double us, ts, L;
    const int q = M.bpatch[b];
    most_iterate(mag, T_o, Ts, M.wsy[b], M.pz0m[q], M.pz0h[q], Lg, &us, &ts, &L);
    Lb[b] = L;
    for (int k = 0; k < 3; ++k) tau[3 * size_t(b) + k] = -us * us * up[k];
    qw[b] = -us * ts;
    r.v[0] = L;
    r.v[1] = 1.0;
    return r;
  }
};
struct MostLFin {
  double* Lcfg;
  int update;
  DM_HD void operator()(const Red<2>& s) const {
    if (update && s.v[1] > 0.0) *Lcfg = s.v[0] / s.v[1];
  }
};

// ------------------------------------------------------------------ reconstruction schemes

// QUICK / van Leer TVD / Koren TVD face values of the vector SoA path
// (advection_rhs_vector_recon_soa, src/numerics/mod_advection.f90:1396-1635, helpers :2334-2791):
// upwind/downwind context along the face, far-upwind search over the upwind cell's faces, the
// limited reconstruction and its non-finite fallbacks. One face per item; the three components of
// the face value go to phiF, which the explicit-RHS kernel gathers.
enum { SCHEME_VANLEER = 1, SCHEME_KOREN = 2, SCHEME_CDS = 3, SCHEME_KEP = 4, SCHEME_QUICK = 5 };

struct ReconCtx {
  int up, down;
  double x_up[3], x_down[3], up_shift[3], dir_hat[3], s_face, s_down;
};

// What a TVD face value needs of the face's upwind context, which depends on the mesh and on the
// sign of the face flux only: up / down cells, the far-upwind cell, the face's and the downwind
// cell's distance from the upwind cell along the upwind direction and the far-upwind cell's
// distance behind it (tvd1's s_up). ok = 0: no context (the boundary value is the face value).
struct TvdCtx {
  int up, down, far, ok;
  double s_face, s_down, s_up;
};

template <int SCHEME>
struct ReconFace {
  DevMesh M;
  const double *U, *Ub, *UF, *grad;
  double* phiF;

  DM_HD static double clampv(double x, double lo, double hi) { return fmin(fmax(x, lo), hi); }
  DM_HD static double nrm3(const double* v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

  // soa_resolve_periodic_partner_dev (:2362-2414)
  DM_HD int partner(int f, double* shift) const {
    shift[0] = shift[1] = shift[2] = 0.0;
    const int b = f - M.nint;
    if (M.bbc[b] != BC_PERIODIC) return -1;
    const int owner = M.bown[b];
    int pc = M.pnb[b], pf = -1;
    if (M.pair[b] >= 0) {
      pf = M.nint + M.pair[b];
      if (pc < 0) pc = M.bown[M.pair[b]];
    }
    if (pc < 0) return -1;
    double d_per;
    if (pf >= 0) d_per = M.dcf[f] + M.dcf[pf];
    else if (M.pdcf[b] > SMALL) d_per = M.dcf[f] + M.pdcf[b];
    else d_per = 2.0 * M.dcf[f];
    for (int i = 0; i < 3; ++i)
      shift[i] = (M.ctr[3 * size_t(owner) + i] + d_per * M.nrm[3 * size_t(f) + i]) - M.ctr[3 * size_t(pc) + i];
    return pc;
  }
  // soa_resolve_reconstruction_context_dev (:2417-2498)
  DM_HD bool context(int f, double Fn, ReconCtx& c) const {
    c.up = c.down = 0;
    c.s_face = c.s_down = 0.0;
    for (int i = 0; i < 3; ++i) c.x_up[i] = c.x_down[i] = c.up_shift[i] = c.dir_hat[i] = 0.0;
    const double* xf = &M.fctr[3 * size_t(f)];
    const int owner = M.fown[f], nb = M.fnb[f];
    if (nb >= 0) {
      c.up = Fn >= 0.0 ? owner : nb;
      c.down = Fn >= 0.0 ? nb : owner;
      for (int i = 0; i < 3; ++i) {
        c.x_up[i] = M.ctr[3 * size_t(c.up) + i];
        c.x_down[i] = M.ctr[3 * size_t(c.down) + i];
      }
    } else {
      double sh[3];
      const int pc = partner(f, sh);
      if (pc < 0) return false;
      if (Fn >= 0.0) {
        c.up = owner;
        c.down = pc;
        for (int i = 0; i < 3; ++i) {
          c.x_up[i] = M.ctr[3 * size_t(owner) + i];
          c.x_down[i] = M.ctr[3 * size_t(pc) + i] + sh[i];
        }
      } else {
        c.up = pc;
        c.down = owner;
        for (int i = 0; i < 3; ++i) {
          c.up_shift[i] = sh[i];
          c.x_up[i] = M.ctr[3 * size_t(pc) + i] + c.up_shift[i];
          c.x_down[i] = M.ctr[3 * size_t(owner) + i];
        }
      }
    }
    double dv[3], tf[3], td[3];
    for (int i = 0; i < 3; ++i) dv[i] = xf[i] - c.x_up[i];
    double dn = nrm3(dv);
    if (dn <= SMALL) {
      for (int i = 0; i < 3; ++i) dv[i] = c.x_down[i] - c.x_up[i];
      dn = nrm3(dv);
    }
    if (dn <= SMALL) return false;
    for (int i = 0; i < 3; ++i) {
      c.dir_hat[i] = dv[i] / dn;
      tf[i] = xf[i] - c.x_up[i];
      td[i] = c.x_down[i] - c.x_up[i];
    }
    c.s_face = dot3(tf, c.dir_hat);
    c.s_down = dot3(td, c.dir_hat);
    if (c.s_face <= SMALL || c.s_down <= c.s_face) {
      for (int i = 0; i < 3; ++i) dv[i] = c.x_down[i] - c.x_up[i];
      dn = nrm3(dv);
      if (dn <= SMALL) return false;
      for (int i = 0; i < 3; ++i) c.dir_hat[i] = dv[i] / dn;
      c.s_face = dot3(tf, c.dir_hat);
      c.s_down = dot3(td, c.dir_hat);
    }
    if (c.s_down <= SMALL) return false;
    c.s_face = fmin(fmax(c.s_face, 0.0), c.s_down);
    return true;
  }
  // soa_find_far_upwind_dev (:2501-2549), the upwind cell's faces in its own face order
  DM_HD int far_upwind(const ReconCtx& c) const {
    const int nc = M.nc;
    int far_up = -1;
    double best_align = -1.0, best_perp2 = 1.0e308;
    const int cnt = M.nA[c.up];
    for (int k = 0; k < cnt; ++k) {
      const int oth = M.aoth[size_t(k) * nc + c.up];
      int nb = oth;
      if (oth < 0) {
        const int b = -1 - oth;
        nb = (M.bbc[b] == BC_PERIODIC) ? M.pnb[b] : -1;
      }
      if (nb == c.down || nb < 0) continue;
      double d[3];
      for (int i = 0; i < 3; ++i) d[i] = M.ctr[3 * size_t(nb) + i] + c.up_shift[i] - c.x_up[i];
      const double sd = dot3(d, c.dir_hat);
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
  DM_HD static double axial(double pf, double pu, const double* g, const double* dv, double sv, const double* dir) {
    double t[3];
    for (int i = 0; i < 3; ++i) t[i] = dv[i] - sv * dir[i];
    return pf - pu - dot3(g, t);
  }
  // soa_quick_reconstruction3_dev (:2632-2713)
  DM_HD void quick3(const ReconCtx& c, double* out) const {
    const int nc = M.nc;
    double pU[3], pD[3], b1[3], b2[3], ln[3], gup[3][3], dl[3];
    for (int k = 0; k < 3; ++k) {
      pU[k] = U[3 * size_t(c.up) + k];
      pD[k] = U[3 * size_t(c.down) + k];
      for (int i = 0; i < 3; ++i) gup[k][i] = grad[DM_G(c.up, i, k)];
    }
    double dv[3];
    for (int i = 0; i < 3; ++i) dv[i] = c.x_down[i] - c.x_up[i];
    double dist2 = fmax(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2], SMALL);
    double w = 1.0 / dist2;
    double a11 = w * c.s_down * c.s_down;
    double a12 = a11 * c.s_down;
    double a22 = a12 * c.s_down;
    double lden = a11;
    int nfit = 1;
    for (int k = 0; k < 3; ++k) {
      dl[k] = axial(pD[k], pU[k], gup[k], dv, c.s_down, c.dir_hat);
      b1[k] = w * c.s_down * dl[k];
      b2[k] = w * c.s_down * c.s_down * dl[k];
      ln[k] = b1[k];
    }
    const int cnt = M.nA[c.up];
    for (int q = 0; q < cnt; ++q) {
      const int nb = M.aoth[size_t(q) * nc + c.up];
      if (nb < 0 || nb == c.down) continue;
      for (int i = 0; i < 3; ++i) dv[i] = (M.ctr[3 * size_t(nb) + i] + c.up_shift[i]) - c.x_up[i];
      const double sv = dot3(dv, c.dir_hat);
      if (sv >= -1.0e-12) continue;
      dist2 = fmax(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2], SMALL);
      if (sv * sv < 0.25 * dist2) continue;
      const double align = fmin(1.0, -sv / sqrt(dist2));
      w = (0.25 + align * align) / dist2;
      a11 = a11 + w * sv * sv;
      a12 = a12 + w * sv * sv * sv;
      a22 = a22 + w * sv * sv * sv * sv;
      lden = lden + w * sv * sv;
      for (int k = 0; k < 3; ++k) {
        dl[k] = axial(U[3 * size_t(nb) + k], pU[k], gup[k], dv, sv, c.dir_hat);
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
        a1 = (pD[k] - pU[k]) / fmax(c.s_down, SMALL);
        a2 = 0.0;
      }
      out[k] = clampv(pU[k] + a1 * c.s_face + a2 * c.s_face * c.s_face, fmin(pU[k], pD[k]), fmax(pU[k], pD[k]));
    }
  }
  // soa_vanleer_tvd_ctx_dev / soa_koren_tvd_ctx_dev (:2716-2786)
  DM_HD double tvd(const ReconCtx& c, int far_up, int k) const {
    const double pU = U[3 * size_t(c.up) + k], pD = U[3 * size_t(c.down) + k];
    const double slope_dn = (pD - pU) / fmax(c.s_down, SMALL);
    double psi = 0.0;
    if (far_up >= 0 && fabs(slope_dn) > SMALL) {
      double t[3];
      for (int i = 0; i < 3; ++i) t[i] = M.ctr[3 * size_t(far_up) + i] + c.up_shift[i] - c.x_up[i];
      const double s_up = -dot3(t, c.dir_hat);
      if (s_up > SMALL) {
        const double r = ((pU - U[3 * size_t(far_up) + k]) / s_up) / slope_dn;
        if (SCHEME == SCHEME_KOREN) {
          psi = fmax(0.0, fmin(fmin(2.0 * r, (1.0 + 2.0 * r) / 3.0), 2.0));
        } else {
          psi = (r + fabs(r)) / (1.0 + fabs(r));
          psi = fmax(0.0, fmin(psi, 2.0));
        }
      }
    }
    return clampv(pU + psi * slope_dn * c.s_face, fmin(pU, pD), fmax(pU, pD));
  }
  DM_HD void operator()(int f) const {
    const int o = M.fown[f], nb = M.fnb[f];
    double Fn;
    int b = 0;
    if (nb >= 0) {
      Fn = UF[f];
    } else {
      b = f - M.nint;
      Fn = is_solid(M.bbc[b]) ? 0.0 : UF[f];
    }
    ReconCtx c;
    const bool ok = context(f, Fn, c);
    int far_up = -1;
    if (ok && SCHEME != SCHEME_QUICK) far_up = far_upwind(c);
    if (!isfinite(Fn)) Fn = 0.0;
    double q[3] = {0.0, 0.0, 0.0};
    if (ok && SCHEME == SCHEME_QUICK) quick3(c, q);
    for (int k = 0; k < 3; ++k) {
      double phi;
      if (ok) {
        phi = (SCHEME == SCHEME_QUICK) ? q[k] : tvd(c, far_up, k);
        if (!isfinite(phi)) phi = (Fn >= 0.0 || nb < 0) ? U[3 * size_t(o) + k] : U[3 * size_t(nb) + k];
      } else {
        phi = Ub[3 * size_t(b) + k];
      }
      if (!isfinite(phi)) phi = 0.0;
      phiF[3 * size_t(f) + k] = phi;
    }
  }
};

// ------------------------------------------------------------------ temperature

// apply_temperature_bc (src/bc/mod_bc.f90:489-580), one boundary face. filter: 0 all,
// 1 non-periodic, 2 periodic only. gT may be null (zero-gradient extrapolation).
struct BcTemperature {
  DevMesh M;
  const double *T, *gT, *patchT;
  double* Tb;
  int filter;
  DM_HD void operator()(int b) const {
    const int f = M.nint + b, bc = M.bbc[b];
    if (filter == 2 && bc != BC_PERIODIC) return;
    if (filter == 1 && bc == BC_PERIODIC) return;
    const int o = M.bown[b];
    if (bc == BC_INLET || bc == BC_FIXEDTEMP) {
      Tb[b] = patchT[M.bpatch[b]];
    } else if (bc == BC_PERIODIC) {
      const int nb = M.pnb[b];
      Tb[b] = (nb >= 0) ? M.pgc[b] * T[o] + (1.0 - M.pgc[b]) * T[nb] : T[o];
    } else if (gT) {
      double d[3];
      for (int i = 0; i < 3; ++i) d[i] = M.fctr[3 * size_t(f) + i] - M.ctr[3 * size_t(o) + i];
      Tb[b] = T[o] + gT[3 * size_t(o)] * d[0] + gT[3 * size_t(o) + 1] * d[1] + gT[3 * size_t(o) + 2] * d[2];
    } else {
      Tb[b] = T[o];
    }
  }
};

// apply_passive_scalar_bc (src/bc/mod_bc.f90:582-679) for one boundary face and scalar k:
// periodic faces interpolate with the partner cell; otherwise the patch's mode: Dirichlet value,
// Neumann flux q (times cos(omega t + phase) when omega != 0) as C_o + q d_cf / (D + nu_t / Pr_t),
// else zero gradient. Per-patch arrays are indexed [4 patch + k].
struct BcScalar {
  DevMesh M;
  const double *C, *nus;
  double* Cb;
  const int* mode;
  const double *val, *flux, *om, *ph;
  double D, prt, time;
  int k, filter;
  DM_HD void operator()(int b) const {
    const int f = M.nint + b, bc = M.bbc[b];
    if (filter == 2 && bc != BC_PERIODIC) return;
    if (filter == 1 && bc == BC_PERIODIC) return;
    const int o = M.bown[b];
    if (bc == BC_PERIODIC) {
      const int nb = M.pnb[b];
      Cb[b] = (nb >= 0) ? M.pgc[b] * C[o] + (1.0 - M.pgc[b]) * C[nb] : C[o];
      return;
    }
    const int q = 4 * M.bpatch[b] + k;
    if (mode[q] == 1) {
      Cb[b] = val[q];
    } else if (mode[q] == 2) {
      double fl = flux[q];
      if (fabs(om[q]) > SMALL) fl = fl * cos(om[q] * time + ph[q]);
      const double af = D + nus[o] / prt;
      const double dcf = fmax(M.dcf[f], SMALL);
      Cb[b] = (af > SMALL) ? C[o] + fl * dcf / af : C[o];
    } else {
      Cb[b] = C[o];
    }
  }
};

// Face values of a scalar by QUICK / van Leer / Koren (advection_rhs_scalar_recon_soa,
// src/numerics/mod_advection.f90:488-675): the context and far-upwind search of ReconFace, the
// single-field QUICK fit (soa_quick_reconstruction_ctx_dev, :2552-2629).
template <int SCHEME>
struct ReconFaceScalar {
  ReconFace<SCHEME> R;  // geometry helpers (its U/grad are unused)
  const double *phi, *pb, *gphi;
  double* phiF;
  DM_HD double quick1(const ReconCtx& c) const {
    const DevMesh& M = R.M;
    const int nc = M.nc;
    const double pU = phi[c.up], pD = phi[c.down];
    const double* gup = &gphi[3 * size_t(c.up)];
    double a11 = 0.0, a12 = 0.0, a22 = 0.0, b1 = 0.0, b2 = 0.0, ln = 0.0, ld = 0.0;
    int nfit = 0;
    double dv[3];
    for (int i = 0; i < 3; ++i) dv[i] = c.x_down[i] - c.x_up[i];
    double dist2 = fmax(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2], SMALL);
    double w = 1.0 / dist2;
    double dl = ReconFace<SCHEME>::axial(pD, pU, gup, dv, c.s_down, c.dir_hat);
    const double sd = c.s_down;
    a11 = a11 + w * sd * sd;
    a12 = a12 + w * sd * sd * sd;
    a22 = a22 + w * sd * sd * sd * sd;
    b1 = b1 + w * sd * dl;
    b2 = b2 + w * sd * sd * dl;
    ln = ln + w * sd * dl;
    ld = ld + w * sd * sd;
    nfit = nfit + 1;
    const int cnt = M.nA[c.up];
    for (int q = 0; q < cnt; ++q) {
      const int nb = M.aoth[size_t(q) * nc + c.up];
      if (nb < 0 || nb == c.down) continue;
      for (int i = 0; i < 3; ++i) dv[i] = (M.ctr[3 * size_t(nb) + i] + c.up_shift[i]) - c.x_up[i];
      const double sv = dot3(dv, c.dir_hat);
      if (sv >= -1.0e-12) continue;
      dist2 = fmax(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2], SMALL);
      if (sv * sv < 0.25 * dist2) continue;
      const double align = fmin(1.0, -sv / sqrt(dist2));
      w = (0.25 + align * align) / dist2;
      dl = ReconFace<SCHEME>::axial(phi[nb], pU, gup, dv, sv, c.dir_hat);
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
      a1 = (pD - pU) / fmax(c.s_down, SMALL);
      a2 = 0.0;
    }
    return ReconFace<SCHEME>::clampv(pU + a1 * c.s_face + a2 * c.s_face * c.s_face, fmin(pU, pD), fmax(pU, pD));
  }
  // tvd1 from a precomputed context (the same arithmetic on the same distances)
  DM_HD double tvd1p(const TvdCtx& t) const {
    const double pU = phi[t.up], pD = phi[t.down];
    const double slope_dn = (pD - pU) / fmax(t.s_down, SMALL);
    double psi = 0.0;
    if (t.far >= 0 && fabs(slope_dn) > SMALL) {
      if (t.s_up > SMALL) {
        const double r = ((pU - phi[t.far]) / t.s_up) / slope_dn;
        if (SCHEME == SCHEME_KOREN) {
          psi = fmax(0.0, fmin(fmin(2.0 * r, (1.0 + 2.0 * r) / 3.0), 2.0));
        } else {
          psi = (r + fabs(r)) / (1.0 + fabs(r));
          psi = fmax(0.0, fmin(psi, 2.0));
        }
      }
    }
    return ReconFace<SCHEME>::clampv(pU + psi * slope_dn * t.s_face, fmin(pU, pD), fmax(pU, pD));
  }
  DM_HD double tvd1(const ReconCtx& c, int far_up) const {
    const DevMesh& M = R.M;
    const double pU = phi[c.up], pD = phi[c.down];
    const double slope_dn = (pD - pU) / fmax(c.s_down, SMALL);
    double psi = 0.0;
    if (far_up >= 0 && fabs(slope_dn) > SMALL) {
      double t[3];
      for (int i = 0; i < 3; ++i) t[i] = M.ctr[3 * size_t(far_up) + i] + c.up_shift[i] - c.x_up[i];
      const double s_up = -dot3(t, c.dir_hat);
      if (s_up > SMALL) {
        const double r = ((pU - phi[far_up]) / s_up) / slope_dn;
        if (SCHEME == SCHEME_KOREN) {
          psi = fmax(0.0, fmin(fmin(2.0 * r, (1.0 + 2.0 * r) / 3.0), 2.0));
        } else {
          psi = (r + fabs(r)) / (1.0 + fabs(r));
          psi = fmax(0.0, fmin(psi, 2.0));
        }
      }
    }
    return ReconFace<SCHEME>::clampv(pU + psi * slope_dn * c.s_face, fmin(pU, pD), fmax(pU, pD));
  }
  DM_HD void operator()(int f) const {
    const DevMesh& M = R.M;
    const int o = M.fown[f], nb = M.fnb[f];
    double Fn;
    int b = 0;
    if (nb >= 0) {
      Fn = R.UF[f];
    } else {
      b = f - M.nint;
      Fn = is_solid(M.bbc[b]) ? 0.0 : R.UF[f];
    }
    ReconCtx c;
    double v;
    if (R.context(f, Fn, c)) {
      v = (SCHEME == SCHEME_QUICK) ? quick1(c) : tvd1(c, R.far_upwind(c));
      if (!isfinite(v)) v = (Fn >= 0.0 || nb < 0) ? phi[o] : phi[nb];  // before Fn is sanitised
    } else {
      v = pb[b];
    }
    if (!isfinite(v)) v = 0.0;
    phiF[f] = v;
  }
};

// The face values of NF scalar fields that share a reconstruction scheme (temperature and the
// passive scalars) in one pass: the upwind context and far-upwind cell depend on the face flux
// alone, so they are found once and each field is reconstructed as ReconFaceScalar does it.
constexpr int RHS_MAX_FIELDS = 5;
template <int SCHEME, int NF>
struct ReconFaceScalarN {
  ReconFaceScalar<SCHEME> base;  // field 0; the others swap in their arrays
  const double* phi[RHS_MAX_FIELDS];
  const double* pb[RHS_MAX_FIELDS];
  const double* gphi[RHS_MAX_FIELDS];
  double* phiF[RHS_MAX_FIELDS];
  const TvdCtx* tctx = nullptr;  // TVD schemes: the precomputed face contexts (TvdCtxBuild)
  DM_HD void values(int f, double* out) const {
    const DevMesh& M = base.R.M;
    const int o = M.fown[f], nb = M.fnb[f];
    double Fn;
    int b = 0;
    if (nb >= 0) {
      Fn = base.R.UF[f];
    } else {
      b = f - M.nint;
      Fn = is_solid(M.bbc[b]) ? 0.0 : base.R.UF[f];
    }
    if (SCHEME != SCHEME_QUICK && tctx) {
      const TvdCtx t = tctx[2 * size_t(f) + (Fn >= 0.0 ? 0 : 1)];
#pragma unroll
      for (int i = 0; i < NF; ++i) {
        ReconFaceScalar<SCHEME> g = base;
        g.phi = phi[i];
        double v;
        if (t.ok) {
          v = g.tvd1p(t);
          if (!isfinite(v)) v = (Fn >= 0.0 || nb < 0) ? phi[i][o] : phi[i][nb];
        } else {
          v = pb[i][b];
        }
        if (!isfinite(v)) v = 0.0;
        out[i] = v;
      }
      return;
    }
    ReconCtx c;
    const bool ok = base.R.context(f, Fn, c);
    const int far_up = (ok && SCHEME != SCHEME_QUICK) ? base.R.far_upwind(c) : -1;
#pragma unroll
    for (int i = 0; i < NF; ++i) {
      ReconFaceScalar<SCHEME> g = base;
      g.phi = phi[i];
      g.pb = pb[i];
      g.gphi = gphi[i];
      double v;
      if (ok) {
        v = (SCHEME == SCHEME_QUICK) ? g.quick1(c) : g.tvd1(c, far_up);
        if (!isfinite(v)) v = (Fn >= 0.0 || nb < 0) ? phi[i][o] : phi[i][nb];
      } else {
        v = pb[i][b];
      }
      if (!isfinite(v)) v = 0.0;
      out[i] = v;
    }
  }
  DM_HD void operator()(int f) const {
    double v[NF];
    values(f, v);
#pragma unroll
    for (int i = 0; i < NF; ++i) phiF[i][f] = v[i];
  }
};

// The TVD face contexts of every face for both flux signs (ctx[2 f]: F >= 0, ctx[2 f + 1]:
// F < 0), from ReconFace's context and far-upwind search; built once per mesh.
template <int SCHEME>
struct TvdCtxBuild {
  ReconFace<SCHEME> R;
  TvdCtx* ctx;
  DM_HD void operator()(int f) const {
    const DevMesh& M = R.M;
    for (int sg = 0; sg < 2; ++sg) {
      TvdCtx t{0, 0, -1, 0, 0.0, 0.0, 0.0};
      ReconCtx c;
      if (R.context(f, sg == 0 ? 1.0 : -1.0, c)) {
        t.ok = 1;
        t.up = c.up;
        t.down = c.down;
        t.s_face = c.s_face;
        t.s_down = c.s_down;
        t.far = R.far_upwind(c);
        if (t.far >= 0) {
          double d[3];
          for (int i = 0; i < 3; ++i) d[i] = M.ctr[3 * size_t(t.far) + i] + c.up_shift[i] - c.x_up[i];
          t.s_up = -dot3(d, c.dir_hat);
        }
      }
      ctx[2 * size_t(f) + sg] = t;
    }
  }
};

// Scalar transport of NF fields that share a scheme, face half: for each interior face the
// convective flux F phi_f and the diffusive flux of every field, as ExplicitRhsTN's per-cell sums
// use them (computed once per face instead of once from each side), into FT: plane 2q the
// convective, plane 2q + 1 the diffusive flux of field q, nint values each. Boundary faces of a
// reconstruction scheme get their face values phiF, as ReconFaceScalarN writes them; with CDS
// (cds = 1) the launch covers the interior faces only.
template <int SCHEME, int NF>
struct FaceRhsTN {
  ReconFaceScalarN<SCHEME, NF> rec;  // face values of the reconstruction scheme (cds = 0)
  const double* nus;
  int cds;
  double alpha[RHS_MAX_FIELDS], prt[RHS_MAX_FIELDS];
  double* FT;
  DM_HD void operator()(int f) const {
    const DevMesh& M = rec.base.R.M;
    const int nb = M.fnb[f];
    if (nb < 0) {
      if (!cds) rec(f);
      return;
    }
    const int o = M.fown[f];
    const size_t ni = size_t(M.nint);
    double pf[NF];
    if (!cds) rec.values(f, pf);
    const double* sf = &M.sf[3 * size_t(f)];
    const double gc = M.gc[f];
    double d_on[3], dnp[3];
    for (int i = 0; i < 3; ++i) d_on[i] = M.ctr[3 * size_t(nb) + i] - M.ctr[3 * size_t(o) + i];
    const double dl = sqrt(d_on[0] * d_on[0] + d_on[1] * d_on[1] + d_on[2] * d_on[2]) + SMALL;
    const double sdd = dot3(sf, d_on) / dl;
    for (int i = 0; i < 3; ++i) dnp[i] = sf[i] - (sdd / dl) * d_on[i];
    const double ow = ortho_weight(sdd, M.area[f]);
    const double nuf = gc * nus[o] + (1.0 - gc) * nus[nb];
    const double Fn0 = rec.base.R.UF[f];
#pragma unroll
    for (int q = 0; q < NF; ++q) {
      const double* Tq = rec.phi[q];
      const double* gq = rec.gphi[q];
      double Fn = Fn0;
      double phif;
      if (cds) {
        const double pl = gc * Tq[o] + (1.0 - gc) * Tq[nb];
        double gm[3];
        for (int i = 0; i < 3; ++i) gm[i] = gc * gq[3 * size_t(o) + i] + (1.0 - gc) * gq[3 * size_t(nb) + i];
        phif = pl + dot3(gm, &M.eps[3 * size_t(f)]);
        const double w = M.wbl[f];
        if (w < 1.0) phif = pl + w * (phif - pl);
        if (!isfinite(phif)) phif = (Fn >= 0.0) ? Tq[o] : Tq[nb];
        if (!isfinite(Fn)) Fn = 0.0;
        if (!isfinite(phif)) phif = 0.0;
      } else {
        if (!isfinite(Fn)) Fn = 0.0;
        phif = pf[q];
      }
      FT[2 * q * ni + f] = Fn * phif;
      const double ae = alpha[q] + nuf / prt[q];
      double gf[3];
      for (int i = 0; i < 3; ++i) gf[i] = gc * gq[3 * size_t(o) + i] + (1.0 - gc) * gq[3 * size_t(nb) + i];
      FT[(2 * q + 1) * ni + f] = ae * ((Tq[nb] - Tq[o]) * ow / dl + dot3(gf, dnp));
    }
  }
};

// Scalar transport of NF fields that share a scheme, cell half: each field's sums as ExplicitRhsT
// forms them (same expressions, same order), the interior faces' fluxes read from FaceRhsTN's FT,
// the boundary faces computed here (one side only).
template <int NF>
struct ExplicitRhsTN {
  DevMesh M;
  const double *UF, *nus;
  double dt, a, bcoef;
  int cds;
  const double* FT;  // interior-face fluxes from FaceRhsTN
  const double *T[RHS_MAX_FIELDS], *Tb[RHS_MAX_FIELDS], *gT[RHS_MAX_FIELDS], *HT0[RHS_MAX_FIELDS],
      *phiF[RHS_MAX_FIELDS], *qw[RHS_MAX_FIELDS];
  double *HT[RHS_MAX_FIELDS], *Tnew[RHS_MAX_FIELDS];
  double alpha[RHS_MAX_FIELDS], prt[RHS_MAX_FIELDS];
  DM_HD void operator()(int c) const {
    const int nc = M.nc;
    const double iv = M.iv[c];
    double cv[NF], vs[NF];
#pragma unroll
    for (int i = 0; i < NF; ++i) cv[i] = vs[i] = 0.0;
    const int cnt = M.nA[c];
    for (int k = 0; k < cnt; ++k) {
      const size_t s = size_t(k) * nc + c;
      const int e = M.aface[s], f = e >> 1, side = e & 1, oth = M.aoth[s];
      const double* sf = &M.sf[3 * size_t(f)];
      if (oth >= 0) {
        const size_t ni = size_t(M.nint);
#pragma unroll
        for (int q = 0; q < NF; ++q) {
          const double cf = FT[2 * q * ni + f], vf = FT[(2 * q + 1) * ni + f];
          cv[q] = side ? cv[q] + cf * iv : cv[q] - cf * iv;
          vs[q] = side ? vs[q] - vf * iv : vs[q] + vf * iv;
        }
      } else {
        const int b = -1 - oth, bc = M.bbc[b];
        const int nbp = (bc == BC_PERIODIC) ? M.pnb[b] : -1;
        const double Fn0 = is_solid(bc) ? 0.0 : UF[f];
        double d_per = 0.0, gcp = 0.0;
        double d_on[3], dnp[3], dl;
        if (nbp >= 0) {
          if (M.pair[b] >= 0) {
            const int pf = M.nint + M.pair[b];
            d_per = M.dcf[f] + M.dcf[pf];
            gcp = M.dcf[pf] / (d_per + SMALL);
          } else if (M.pdcf[b] > SMALL) {
            d_per = M.dcf[f] + M.pdcf[b];
            gcp = M.pdcf[b] / (d_per + SMALL);
          } else {
            d_per = 2.0 * M.dcf[f];
            gcp = 0.5;
          }
          for (int i = 0; i < 3; ++i) d_on[i] = d_per * M.nrm[3 * size_t(f) + i];
          dl = d_per + SMALL;
        } else {
          for (int i = 0; i < 3; ++i) d_on[i] = M.fctr[3 * size_t(f) + i] - M.ctr[3 * size_t(c) + i];
          dl = sqrt(d_on[0] * d_on[0] + d_on[1] * d_on[1] + d_on[2] * d_on[2]) + SMALL;
        }
        const double sdd = dot3(sf, d_on) / dl;
        for (int i = 0; i < 3; ++i) dnp[i] = sf[i] - (sdd / dl) * d_on[i];
        const double ow = ortho_weight(sdd, M.area[f]);
#pragma unroll
        for (int q = 0; q < NF; ++q) {
          const double* Tq = T[q];
          const double* gq = gT[q];
          double Fn = Fn0;
          double phif;
          if (cds) {
            if (nbp >= 0) {
              const double g = M.pgc[b];
              phif = g * Tq[c] + (1.0 - g) * Tq[nbp];
              if (!isfinite(phif)) phif = Tq[c];
            } else {
              phif = Tb[q][b];
            }
            if (!isfinite(Fn)) Fn = 0.0;
            if (!isfinite(phif)) phif = 0.0;
          } else {
            if (!isfinite(Fn)) Fn = 0.0;
            phif = phiF[q][f];
          }
          cv[q] = cv[q] - Fn * phif * iv;
          double flux;
          if (nbp >= 0) {
            const double ae = alpha[q] + (gcp * nus[c] + (1.0 - gcp) * nus[nbp]) / prt[q];
            double gf[3];
            for (int i = 0; i < 3; ++i) gf[i] = gcp * gq[3 * size_t(c) + i] + (1.0 - gcp) * gq[3 * size_t(nbp) + i];
            flux = ae * ((Tq[nbp] - Tq[c]) * ow / dl + dot3(gf, dnp));
          } else {
            const double ae = alpha[q] + nus[c] / prt[q];
            flux = ae * ((Tb[q][b] - Tq[c]) * ow / dl + dot3(&gq[3 * size_t(c)], dnp));
          }
          vs[q] = vs[q] + flux * iv;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < NF; ++q) {
      double h = cv[q] + vs[q];
      if (qw[q])
        for (int k = 0; k < cnt; ++k) {
          const size_t s = size_t(k) * nc + c;
          const int oth = M.aoth[s];
          if (oth >= 0 || M.bbc[-1 - oth] != BC_WALL_MODEL) continue;
          const int f = M.aface[s] >> 1;
          h = h + qw[q][-1 - oth] * M.area[f] / M.vol[c];
        }
      HT[q][c] = h;
      Tnew[q][c] = T[q][c] + dt * (a * h + bcoef * HT0[q][c]);
    }
  }
};

// Temperature right-hand side and predictor for one cell: advection_rhs_scalar (CDS:
// src/numerics/mod_advection.f90:267-485; reconstruction schemes: face values phiF) +
// viscous_rhs_scalar (src/numerics/mod_viscous.f90:286-513), H_T = conv + visc,
// T* = T + dt (alpha H_T + beta H_T,prev) (src/solver/mod_fv_solver.f90:1326-1331).
struct ExplicitRhsT {
  DevMesh M;
  const double *T, *Tb, *gT, *UF, *nus, *HT0, *phiF;
  double *HT, *Tnew;
  double alpha, prt, dt, a, bcoef;
  int cds;
  const double* qw;  // MOST wall heat flux per boundary face, or null
  int part = 0;      // 0: H_T and the predictor; 1: only the advection term into HT; 2: only the diffusion term
  DM_HD void operator()(int c) const {
    const int nc = M.nc;
    const double iv = M.iv[c];
    double cv = 0.0, vs = 0.0;
    const int cnt = M.nA[c];
    for (int k = 0; k < cnt; ++k) {
      const size_t s = size_t(k) * nc + c;
      const int e = M.aface[s], f = e >> 1, side = e & 1, oth = M.aoth[s];
      const double* sf = &M.sf[3 * size_t(f)];
      if (oth >= 0) {
        const int o = side ? oth : c, nb = side ? c : oth;
        double Fn = UF[f];
        double phif;
        if (cds) {
          const double gcf = M.gc[f];
          const double pl = gcf * T[o] + (1.0 - gcf) * T[nb];
          double gm[3];
          for (int i = 0; i < 3; ++i) gm[i] = gcf * gT[3 * size_t(o) + i] + (1.0 - gcf) * gT[3 * size_t(nb) + i];
          phif = pl + dot3(gm, &M.eps[3 * size_t(f)]);
          const double w = M.wbl[f];
          if (w < 1.0) phif = pl + w * (phif - pl);
          if (!isfinite(phif)) phif = (Fn >= 0.0) ? T[o] : T[nb];
          if (!isfinite(Fn)) Fn = 0.0;
          if (!isfinite(phif)) phif = 0.0;
        } else {
          if (!isfinite(Fn)) Fn = 0.0;
          phif = phiF[f];
        }
        cv = side ? cv + Fn * phif * iv : cv - Fn * phif * iv;
        // diffusion
        const double gc = M.gc[f];
        const double ae = alpha + (gc * nus[o] + (1.0 - gc) * nus[nb]) / prt;
        double d_on[3], dnp[3], gf[3];
        for (int i = 0; i < 3; ++i) d_on[i] = M.ctr[3 * size_t(nb) + i] - M.ctr[3 * size_t(o) + i];
        const double dl = sqrt(d_on[0] * d_on[0] + d_on[1] * d_on[1] + d_on[2] * d_on[2]) + SMALL;
        const double sdd = dot3(sf, d_on) / dl;
        for (int i = 0; i < 3; ++i) {
          dnp[i] = sf[i] - (sdd / dl) * d_on[i];
          gf[i] = gc * gT[3 * size_t(o) + i] + (1.0 - gc) * gT[3 * size_t(nb) + i];
        }
        const double flux = ae * ((T[nb] - T[o]) * ortho_weight(sdd, M.area[f]) / dl + dot3(gf, dnp));
        vs = side ? vs - flux * iv : vs + flux * iv;
      } else {
        const int b = -1 - oth, bc = M.bbc[b];
        const int nbp = (bc == BC_PERIODIC) ? M.pnb[b] : -1;
        double Fn = is_solid(bc) ? 0.0 : UF[f];
        double phif;
        if (cds) {
          if (nbp >= 0) {
            const double g = M.pgc[b];
            phif = g * T[c] + (1.0 - g) * T[nbp];
            if (!isfinite(phif)) phif = T[c];
          } else {
            phif = Tb[b];
          }
          if (!isfinite(Fn)) Fn = 0.0;
          if (!isfinite(phif)) phif = 0.0;
        } else {
          if (!isfinite(Fn)) Fn = 0.0;
          phif = phiF[f];
        }
        cv = cv - Fn * phif * iv;
        double flux;
        if (nbp >= 0) {
          double d_per, gcp;
          if (M.pair[b] >= 0) {
            const int pf = M.nint + M.pair[b];
            d_per = M.dcf[f] + M.dcf[pf];
            gcp = M.dcf[pf] / (d_per + SMALL);
          } else if (M.pdcf[b] > SMALL) {
            d_per = M.dcf[f] + M.pdcf[b];
            gcp = M.pdcf[b] / (d_per + SMALL);
          } else {
            d_per = 2.0 * M.dcf[f];
            gcp = 0.5;
          }
          const double ae = alpha + (gcp * nus[c] + (1.0 - gcp) * nus[nbp]) / prt;
          double d_on[3], dnp[3], gf[3];
          for (int i = 0; i < 3; ++i) d_on[i] = d_per * M.nrm[3 * size_t(f) + i];
          const double dl = d_per + SMALL;
          const double sdd = dot3(sf, d_on) / dl;
          for (int i = 0; i < 3; ++i) {
            dnp[i] = sf[i] - (sdd / dl) * d_on[i];
            gf[i] = gcp * gT[3 * size_t(c) + i] + (1.0 - gcp) * gT[3 * size_t(nbp) + i];
          }
          flux = ae * ((T[nbp] - T[c]) * ortho_weight(sdd, M.area[f]) / dl + dot3(gf, dnp));
        } else {
          const double ae = alpha + nus[c] / prt;
          double d_on[3], dnp[3];
          for (int i = 0; i < 3; ++i) d_on[i] = M.fctr[3 * size_t(f) + i] - M.ctr[3 * size_t(c) + i];
          const double dl = sqrt(d_on[0] * d_on[0] + d_on[1] * d_on[1] + d_on[2] * d_on[2]) + SMALL;
          const double sdd = dot3(sf, d_on) / dl;
          for (int i = 0; i < 3; ++i) dnp[i] = sf[i] - (sdd / dl) * d_on[i];
          flux = ae * ((Tb[b] - T[c]) * ortho_weight(sdd, M.area[f]) / dl + dot3(&gT[3 * size_t(c)], dnp));
        }
        vs = vs + flux * iv;
      }
    }
    if (part) {  // advection_rhs_scalar / viscous_rhs_scalar on their own
      HT[c] = part == 1 ? cv : vs;
      return;
    }
    double h = cv + vs;
    // add_wall_heat_flux_to_rhs (src/solver/mod_fv_solver.f90:2279-2313): q_w A / V
    if (qw)
      for (int k = 0; k < cnt; ++k) {
        const size_t s = size_t(k) * nc + c;
        const int oth = M.aoth[s];
        if (oth >= 0 || M.bbc[-1 - oth] != BC_WALL_MODEL) continue;
        const int f = M.aface[s] >> 1;
        h = h + qw[-1 - oth] * M.area[f] / M.vol[c];
      }
    HT[c] = h;
    Tnew[c] = T[c] + dt * (a * h + bcoef * HT0[c]);
  }
};

// ------------------------------------------------------------------ explicit right-hand side

// advection_rhs_vector_{kep,cds}_soa (src/numerics/mod_advection.f90:1013-1392) + viscous_rhs_vector
// (src/numerics/mod_viscous.f90:37-283) + compute_explicit_rhs assembly with body force and wall
// traction (src/solver/mod_fv_solver.f90:1617-1711, 2238-2275) + the RK3W predictor
// u* = u + dt (alpha H + beta H_prev) (:1326-1331), for one cell. Writes H and u* (to Unew).
template <int SCHEME>
struct ExplicitRhs {
  DevMesh M;
  const double *U, *Ub, *UF, *grad, *nus, *tau, *H0;
  double *H, *Unew;
  double nu, f0, f1, f2, dt, a, bcoef;
  int has_wm;
  const double* phiF;  // reconstruction schemes: face values from ReconFace
  const double* T;     // Boussinesq buoyancy -bb (T - T_ref) g (bb = 0: off)
  double bb, Tref, g0, g1, g2;
  int rot;             // Coriolis -2 om x U (src/solver/mod_fv_solver.f90:1694-1698)
  double om0, om1, om2;
  int part = 0;        // 0: H and the predictor; 1: only the advection term into H; 2: only the viscous term
  DM_HD void operator()(int c) const {
    const int nc = M.nc;
    const double iv = M.iv[c];
    const double* Uc = &U[3 * size_t(c)];
    double cv[3] = {0, 0, 0}, vs[3] = {0, 0, 0};
    double dw = 0.0;
    const int cnt = M.nA[c];
    for (int k = 0; k < cnt; ++k) {
      const size_t s = size_t(k) * nc + c;
      const int e = M.aface[s], f = e >> 1, side = e & 1, oth = M.aoth[s];
      const double* sf = &M.sf[3 * size_t(f)];
      if (oth >= 0) {
        const int o = side ? oth : c, nb = side ? c : oth;
        const double* Uo = &U[3 * size_t(o)];
        const double* Un = &U[3 * size_t(nb)];
        const double Fn = UF[f];
        // convective flux
        if (SCHEME == 4) {
          dw = side ? dw - Fn : dw + Fn;
          for (int q = 0; q < 3; ++q) {
            const double phi = 0.5 * (Uo[q] + Un[q]);
            cv[q] = side ? cv[q] + Fn * phi * iv : cv[q] - Fn * phi * iv;
          }
        } else if (SCHEME != 3) {
          const double Fr = isfinite(Fn) ? Fn : 0.0;
          for (int q = 0; q < 3; ++q) {
            const double phi = phiF[3 * size_t(f) + q];
            cv[q] = side ? cv[q] + Fr * phi * iv : cv[q] - Fr * phi * iv;
          }
        } else {
          const double gcf = M.gc[f];
          const double* ep = &M.eps[3 * size_t(f)];
          const double w = M.wbl[f];
          for (int q = 0; q < 3; ++q) {
            const double pl = gcf * Uo[q] + (1.0 - gcf) * Un[q];
            double gm[3];
            for (int i = 0; i < 3; ++i) gm[i] = gcf * grad[DM_G(o, i, q)] + (1.0 - gcf) * grad[DM_G(nb, i, q)];
            double phi = pl + dot3(gm, ep);
            if (w < 1.0) phi = pl + w * (phi - pl);
            if (!isfinite(phi)) phi = 0.0;
            cv[q] = side ? cv[q] + Fn * phi * iv : cv[q] - Fn * phi * iv;
          }
        }
        // viscous stress
        const double gc = M.gc[f];
        const double nue = nu + gc * nus[o] + (1.0 - gc) * nus[nb];
        double gf[9];
        for (int q = 0; q < 9; ++q) gf[q] = gc * grad[9 * size_t(o) + q] + (1.0 - gc) * grad[9 * size_t(nb) + q];
        double d_on[3], dnp[3];
        for (int i = 0; i < 3; ++i) d_on[i] = M.ctr[3 * size_t(nb) + i] - M.ctr[3 * size_t(o) + i];
        const double dlen = sqrt(d_on[0] * d_on[0] + d_on[1] * d_on[1] + d_on[2] * d_on[2]) + SMALL;
        const double sdd = dot3(sf, d_on) / dlen;
        for (int i = 0; i < 3; ++i) dnp[i] = sf[i] - (sdd / dlen) * d_on[i];
        const double wo = ortho_weight(sdd, M.area[f]);
        for (int q = 0; q < 3; ++q) {
          const double col = gf[3 * q] * dnp[0] + gf[3 * q + 1] * dnp[1] + gf[3 * q + 2] * dnp[2];
          const double row = gf[q] * sf[0] + gf[3 + q] * sf[1] + gf[6 + q] * sf[2];
          const double st = nue * ((Un[q] - Uo[q]) * wo / dlen + col + row);
          vs[q] = side ? vs[q] - st * iv : vs[q] + st * iv;
        }
      } else {
        const int b = -1 - oth;
        const int bc = M.bbc[b];
        const double* ub = &Ub[3 * size_t(b)];
        const int nbp = (bc == BC_PERIODIC) ? M.pnb[b] : -1;
        // convective flux (owner side)
        double Fn;
        if (bc == BC_PERIODIC || bc == BC_OUTLET) Fn = UF[f];
        else if (is_solid(bc)) Fn = 0.0;
        else if (SCHEME != 3 && SCHEME != 4) Fn = UF[f];
        else Fn = ub[0] * sf[0] + ub[1] * sf[1] + ub[2] * sf[2];
        if (SCHEME != 3 && SCHEME != 4) {
          if (!isfinite(Fn)) Fn = 0.0;
          for (int q = 0; q < 3; ++q) cv[q] = cv[q] - Fn * phiF[3 * size_t(f) + q] * iv;
        } else if (SCHEME == 4) {
          dw = dw + Fn;
          for (int q = 0; q < 3; ++q) {
            const double phi = (nbp >= 0) ? 0.5 * (Uc[q] + U[3 * size_t(nbp) + q]) : ub[q];
            cv[q] = cv[q] - Fn * phi * iv;
          }
        } else {
          for (int q = 0; q < 3; ++q) {
            double phi;
            if (nbp >= 0) {
              const double g = M.pgc[b];
              phi = g * Uc[q] + (1.0 - g) * U[3 * size_t(nbp) + q];
            } else {
              phi = ub[q];
            }
            if (!isfinite(phi)) phi = 0.0;
            cv[q] = cv[q] - Fn * phi * iv;
          }
        }
        // viscous stress (owner side)
        if (bc == BC_PERIODIC) {
          if (nbp >= 0) {
            const int pf = M.nint + M.pair[b];
            const double d_per = M.dcf[f] + M.dcf[pf];
            const double gcp = M.dcf[pf] / (d_per + SMALL);
            const double nue = nu + gcp * nus[c] + (1.0 - gcp) * nus[nbp];
            const double* nh = &M.nrm[3 * size_t(f)];
            double d_on[3], dnp[3], gf[9];
            for (int i = 0; i < 3; ++i) d_on[i] = d_per * nh[i];
            const double dlen = d_per + SMALL;
            const double sdd = dot3(sf, d_on) / dlen;
            for (int i = 0; i < 3; ++i) dnp[i] = sf[i] - (sdd / dlen) * d_on[i];
            for (int q = 0; q < 9; ++q) gf[q] = gcp * grad[9 * size_t(c) + q] + (1.0 - gcp) * grad[9 * size_t(nbp) + q];
            const double wo = ortho_weight(sdd, M.area[f]);
            for (int q = 0; q < 3; ++q) {
              const double col = gf[3 * q] * dnp[0] + gf[3 * q + 1] * dnp[1] + gf[3 * q + 2] * dnp[2];
              const double row = gf[q] * sf[0] + gf[3 + q] * sf[1] + gf[6 + q] * sf[2];
              const double st = nue * ((U[3 * size_t(nbp) + q] - Uc[q]) * wo / dlen + col + row);
              vs[q] = vs[q] + st * iv;
            }
          } else {
            const double nue = nu + nus[c];
            for (int q = 0; q < 3; ++q) {
              const double st = nue * ((ub[q] - Uc[q]) / (M.dcf[f] + SMALL)) * M.area[f];
              vs[q] = vs[q] + st * iv;
            }
          }
        } else if (bc != BC_WALL_MODEL) {
          const double nue = nu + nus[c];
          double d_on[3], dnp[3];
          for (int i = 0; i < 3; ++i) d_on[i] = M.fctr[3 * size_t(f) + i] - M.ctr[3 * size_t(c) + i];
          const double dlen = sqrt(d_on[0] * d_on[0] + d_on[1] * d_on[1] + d_on[2] * d_on[2]) + SMALL;
          const double sdd = dot3(sf, d_on) / dlen;
          for (int i = 0; i < 3; ++i) dnp[i] = sf[i] - (sdd / dlen) * d_on[i];
          const double* go = &grad[9 * size_t(c)];
          const double wo = ortho_weight(sdd, M.area[f]);
          for (int q = 0; q < 3; ++q) {
            const double col = go[3 * q] * dnp[0] + go[3 * q + 1] * dnp[1] + go[3 * q + 2] * dnp[2];
            const double row = go[q] * sf[0] + go[3 + q] * sf[1] + go[6 + q] * sf[2];
            const double st = nue * ((ub[q] - Uc[q]) * wo / dlen + col + row);
            vs[q] = vs[q] + st * iv;
          }
        }
      }
    }
    if (SCHEME == 4)
      for (int q = 0; q < 3; ++q) cv[q] = cv[q] + 0.5 * Uc[q] * dw * iv;
    if (part) {  // advection_rhs_vector / viscous_rhs_vector on their own
      for (int q = 0; q < 3; ++q) H[3 * size_t(c) + q] = part == 1 ? cv[q] : vs[q];
      return;
    }
    double h[3] = {cv[0] + vs[0] + f0, cv[1] + vs[1] + f1, cv[2] + vs[2] + f2};
    if (rot) {
      h[0] = h[0] - 2.0 * (om1 * Uc[2] - om2 * Uc[1]);
      h[1] = h[1] - 2.0 * (om2 * Uc[0] - om0 * Uc[2]);
      h[2] = h[2] - 2.0 * (om0 * Uc[1] - om1 * Uc[0]);
    }
    if (T) {
      const double dT = T[c] - Tref;
      h[0] = h[0] - bb * dT * g0;
      h[1] = h[1] - bb * dT * g1;
      h[2] = h[2] - bb * dT * g2;
    }
    if (has_wm) {
      for (int k = 0; k < cnt; ++k) {
        const size_t s = size_t(k) * nc + c;
        const int oth = M.aoth[s];
        if (oth >= 0) continue;
        const int b = -1 - oth;
        if (M.bbc[b] != BC_WALL_MODEL) continue;
        const int f = M.aface[s] >> 1;
        const double w = M.area[f] / M.vol[c];
        for (int q = 0; q < 3; ++q) h[q] = h[q] + tau[3 * size_t(b) + q] * w;
      }
    }
    for (int q = 0; q < 3; ++q) {
      H[3 * size_t(c) + q] = h[q];
      Unew[3 * size_t(c) + q] = Uc[q] + dt * (a * h[q] + bcoef * H0[3 * size_t(c) + q]);
    }
  }
};

// ------------------------------------------------------------------ projection

// compute_face_fluxes_and_rhs (src/solver/mod_fv_solver.f90:2389-2752), first half: the
// provisional flux of each pressure-operator face (interior faces and canonical periodic edges)
// with its Rhie-Chow term, computed once per face (one lane per face) into ff[fm].
struct FaceFluxRC {
  DevMesh M;
  const double *U, *grad, *gprc, *phi;
  double* ff;
  double dt;
  int rck;  // Rhie-Chow subtracts the k-vector (non-orthogonal mesh, rhie_chow_dual off)
  DM_HD void operator()(int i) const {
    const int o = M.pown[i];
    if (o < 0) return;  // skipped face (no finite operator coefficient)
    const int nb = M.pnbr[i], fm = M.pfm[i];
    const double* sf = &M.sf[3 * size_t(fm)];
    const double gcf = M.rcgc[i];
    double Upl[3], Uf[3];
    for (int q = 0; q < 3; ++q) Upl[q] = gcf * U[3 * size_t(o) + q] + (1.0 - gcf) * U[3 * size_t(nb) + q];
    if (fm < M.nint) {
      const double wb = M.wbl[fm], g = M.gc[fm];
      const double* ep = &M.eps[3 * size_t(fm)];
      for (int j = 0; j < 3; ++j) {
        double ue = g * U[3 * size_t(o) + j] + (1.0 - g) * U[3 * size_t(nb) + j];
        double gm[3];
        for (int k = 0; k < 3; ++k) gm[k] = g * grad[DM_G(o, k, j)] + (1.0 - g) * grad[DM_G(nb, k, j)];
        ue = ue + gm[0] * ep[0] + gm[1] * ep[1] + gm[2] * ep[2];
        Uf[j] = Upl[j] + wb * (ue - Upl[j]);
      }
    } else {
      Uf[0] = Upl[0]; Uf[1] = Upl[1]; Uf[2] = Upl[2];
    }
    double flux = Uf[0] * sf[0] + Uf[1] * sf[1] + Uf[2] * sf[2];
    double gm[3], smk[3];
    const double* kr = &M.kint[3 * size_t(i)];
    for (int k = 0; k < 3; ++k) {
      gm[k] = gcf * gprc[3 * size_t(o) + k] + (1.0 - gcf) * gprc[3 * size_t(nb) + k];
      smk[k] = sf[k] - (rck ? kr[k] : 0.0);
    }
    const double rc = dot3(gm, smk) - M.pcoef[i] * (phi[nb] - phi[o]);
    ff[fm] = flux + dt * rc;
  }
};

// Second half, one cell: the boundary-face fluxes (written here), rhs_p = -(1/dt) sum q over the
// cell's list in its order (operator-face fluxes read from ff, as FaceFluxRC left them) + deferred
// non-orthogonal cross terms. Returns |rhs_p| for the max reduction.
struct FluxRhs {
  DevMesh M;
  const double *U, *Ub, *grad, *phi, *gp;
  double *ff, *rhs;
  double dt;
  int nonortho;  // deferred cross terms on (mesh not orthogonal)
  const double* pob;  // cell_projection EXACT: outlet faces carry -dt c_b (p_b - phi_o) (:2629-2636)
  int cproj;
  DM_HD Red<1> operator()(int c) const {
    const int nc = M.nc;
    double acc = 0.0;
    const int cnt = M.nB[c];
    for (int k = 0; k < cnt; ++k) {
      const size_t s = size_t(k) * nc + c;
      const int e = M.bface[s], fm = e >> 1, side = e & 1, oth = M.both[s], idx = M.bidx[s];
      const double* sf = &M.sf[3 * size_t(fm)];
      if (idx >= 0) {
        const double flux = ff[fm];
        acc = side ? acc + flux : acc - flux;
      } else {
        const int b = -1 - oth;
        const int bc = M.bbc[b];
        const double* ub = &Ub[3 * size_t(b)];
        double flux = is_solid(bc) ? 0.0 : ub[0] * sf[0] + ub[1] * sf[1] + ub[2] * sf[2];
        if (bc == BC_OUTLET && M.pconv[M.bpatch[b]]) {
          // continuity sets the outlet flux: the extrapolated cell velocity, not the convective
          // boundary state (src/solver/mod_fv_solver.f90:2619-2627)
          const int f = M.nint + b;
          const double d[3] = {M.fctr[3 * size_t(f)] - M.ctr[3 * size_t(c)],
                               M.fctr[3 * size_t(f) + 1] - M.ctr[3 * size_t(c) + 1],
                               M.fctr[3 * size_t(f) + 2] - M.ctr[3 * size_t(c) + 2]};
          flux = 0.0;
          for (int k = 0; k < 3; ++k)
            flux = flux + (U[3 * size_t(c) + k] + grad[DM_G(c, 0, k)] * d[0] + grad[DM_G(c, 1, k)] * d[1] +
                           grad[DM_G(c, 2, k)] * d[2]) * sf[k];
        }
        if (cproj && bc == BC_OUTLET) flux = flux - dt * M.cb[b] * (pob[b] - phi[c]);
        ff[fm] = flux;
        acc = acc - flux;
      }
    }
    double r = acc * (1.0 / fmax(dt, SMALL));
    if (nonortho) {
      for (int k = 0; k < cnt; ++k) {
        const size_t s = size_t(k) * nc + c;
        const int idx = M.bidx[s];
        if (idx < 0) continue;
        const int side = M.bface[s] & 1, oth = M.both[s];
        const int o = side ? oth : c, nb = side ? c : oth;
        const double cr = cross_flux(&gp[3 * size_t(o)], &gp[3 * size_t(nb)], M.xgc[idx], &M.kint[3 * size_t(idx)]);
        r = side ? r - cr : r + cr;
      }
      for (int k = 0; k < cnt; ++k) {
        const size_t s = size_t(k) * nc + c;
        if (M.bidx[s] >= 0) continue;
        const int b = -1 - M.both[s];
        if (M.bbc[b] != BC_OUTLET) continue;
        r = r + dot3(&gp[3 * size_t(c)], &M.kbnd[3 * size_t(b)]);
      }
    }
    rhs[c] = r;
    Red<1> out;
    out.v[0] = fabs(r);
    return out;
  }
};

// ------------------------------------------------------------------ exact cell projection
// cell_projection = 'EXACT' (src/numerics/mod_cell_projection.f90:89-303 on the operators of
// build_cell_projection, src/solver/mod_fv_solver.f90:3164-3221): the projection faces are the
// pressure-operator faces of the cell's list (interior faces and canonical periodic edges, with the
// face gc or the periodic gc), then its outlet faces. Every sum runs over the cell's list in that
// order, which is the order in which the reference's sequential face loops reach the cell.

// (B u)_c: outflow sum of (gc u_own + (1-gc) u_nbr).Sf, plus u_c.Sf on outlet faces (cp_divergence).
DM_HD inline double proj_div(const DevMesh& M, const double* u, int c) {
  const int nc = M.nc, cnt = M.nB[c];
  double out = 0.0;
  for (int k = 0; k < cnt; ++k) {
    const size_t s = size_t(k) * nc + c;
    const int idx = M.bidx[s];
    if (idx < 0) continue;
    const int e = M.bface[s], fm = e >> 1, side = e & 1, oth = M.both[s];
    const size_t o = size_t(side ? oth : c), n = size_t(side ? c : oth);
    const double w = M.xgc[idx];
    const double* sf = &M.sf[3 * size_t(fm)];
    const double fl = (w * u[3 * o] + (1.0 - w) * u[3 * n]) * sf[0] + (w * u[3 * o + 1] + (1.0 - w) * u[3 * n + 1]) * sf[1] +
                      (w * u[3 * o + 2] + (1.0 - w) * u[3 * n + 2]) * sf[2];
    out = side ? out - fl : out + fl;
  }
  for (int k = 0; k < cnt; ++k) {
    const size_t s = size_t(k) * nc + c;
    if (M.bidx[s] >= 0) continue;
    const int b = -1 - M.both[s];
    if (M.bbc[b] != BC_OUTLET) continue;
    const double* sf = &M.sf[3 * size_t(M.nint + b)];
    const double* uc = &u[3 * size_t(c)];
    out = out + (uc[0] * sf[0] + uc[1] * sf[1] + uc[2] * sf[2]);
  }
  return out;
}

// V_c (V^-1 B^T chi)_c before the 1/V scaling (cp_adjoint_gradient): own side gc d Sf, neighbour
// side (1-gc) d Sf with d = chi_own - chi_nbr, chi_c Sf on outlet faces.
DM_HD inline void proj_adj(const DevMesh& M, const double* chi, int c, double a[3]) {
  const int nc = M.nc, cnt = M.nB[c];
  a[0] = a[1] = a[2] = 0.0;
  for (int k = 0; k < cnt; ++k) {
    const size_t s = size_t(k) * nc + c;
    const int idx = M.bidx[s];
    if (idx < 0) continue;
    const int e = M.bface[s], fm = e >> 1, side = e & 1, oth = M.both[s];
    const int o = side ? oth : c, n = side ? c : oth;
    const double w = side ? 1.0 - M.xgc[idx] : M.xgc[idx];
    const double d = chi[o] - chi[n];
    const double* sf = &M.sf[3 * size_t(fm)];
    for (int i = 0; i < 3; ++i) a[i] = a[i] + w * d * sf[i];
  }
  for (int k = 0; k < cnt; ++k) {
    const size_t s = size_t(k) * nc + c;
    if (M.bidx[s] >= 0) continue;
    const int b = -1 - M.both[s];
    if (M.bbc[b] != BC_OUTLET) continue;
    const double* sf = &M.sf[3 * size_t(M.nint + b)];
    for (int i = 0; i < 3; ++i) a[i] = a[i] + chi[c] * sf[i];
  }
}

// project_cells_exact's boundary pass (src/solver/mod_fv_solver.f90:3240-3270) for one cell: the
// fixed outflow of inlet-type faces into bconst, the outlet pressure as the force -dt p_b Sf / V.
struct ProjBoundary {
  DevMesh M;
  const double *Ub, *pob;
  double *U, *bconst;
  double dt;
  DM_HD void operator()(int c) const {
    const int nc = M.nc, cnt = M.nB[c];
    double bcn = 0.0;
    for (int k = 0; k < cnt; ++k) {
      const size_t s = size_t(k) * nc + c;
      if (M.bidx[s] >= 0) continue;
      const int b = -1 - M.both[s], bc = M.bbc[b];
      if (is_solid(bc)) continue;
      const double* sf = &M.sf[3 * size_t(M.nint + b)];
      if (bc == BC_OUTLET) {
        const double fl = dt * pob[b] / M.vol[c];
        for (int i = 0; i < 3; ++i) U[3 * size_t(c) + i] = U[3 * size_t(c) + i] - fl * sf[i];
      } else {
        const double* ub = &Ub[3 * size_t(b)];
        bcn = bcn + (ub[0] * sf[0] + ub[1] * sf[1] + ub[2] * sf[2]);
      }
    }
    bconst[c] = bcn;
  }
};

// b = B u + bconst, r = b; returns |r| for max|r|.
struct ProjInit {
  DevMesh M;
  const double *U, *bconst;
  double *b, *r;
  DM_HD Red<1> operator()(int c) const {
    const double v = proj_div(M, U, c) + bconst[c];
    b[c] = v;
    r[c] = v;
    Red<1> o;
    o.v[0] = fabs(v);
    return o;
  }
};

// p = z; returns r z.
struct ProjFirstDir {
  const double *r, *z;
  double* p;
  DM_HD Red<1> operator()(int c) const {
    p[c] = z[c];
    Red<1> o;
    o.v[0] = r[c] * z[c];
    return o;
  }
};

// v = V^-1 B^T p (chi = p).
struct ProjAdjoint {
  DevMesh M;
  const double* chi;
  double* v;
  DM_HD void operator()(int c) const {
    double a[3];
    proj_adj(M, chi, c, a);
    const double iv = 1.0 / M.vol[c];
    for (int i = 0; i < 3; ++i) v[3 * size_t(c) + i] = a[i] * iv;
  }
};

// q = B v; returns p q.
struct ProjApply {
  DevMesh M;
  const double *v, *p;
  double* q;
  DM_HD Red<1> operator()(int c) const {
    const double qc = proj_div(M, v, c);
    q[c] = qc;
    Red<1> o;
    o.v[0] = p[c] * qc;
    return o;
  }
};

// r_old = r; x += alpha p; r -= alpha q; returns |r| for max|r| (alpha read from the device).
struct ProjUpdate {
  const double *p, *q;
  double *x, *r, *rold;
  const double* alpha_dev;
  DM_HD Red<1> operator()(int c) const {
    const double alpha = *alpha_dev;
    rold[c] = r[c];
    x[c] = x[c] + alpha * p[c];
    const double rc = r[c] - alpha * q[c];
    r[c] = rc;
    Red<1> o;
    o.v[0] = fabs(rc);
    return o;
  }
};

// returns ((r - r_old) z, r z) for the flexible beta.
struct ProjBeta {
  const double *r, *rold, *z;
  DM_HD Red<2> operator()(int c) const {
    Red<2> o;
    o.v[0] = (r[c] - rold[c]) * z[c];
    o.v[1] = r[c] * z[c];
    return o;
  }
};

// p = z + beta p (beta read from the device).
struct ProjDir {
  const double* z;
  double* p;
  const double* beta_dev;
  DM_HD void operator()(int c) const { p[c] = z[c] + *beta_dev * p[c]; }
};

// u -= V^-1 B^T x.
struct ProjCorrect {
  DevMesh M;
  const double* chi;
  double* U;
  DM_HD void operator()(int c) const {
    double a[3];
    proj_adj(M, chi, c, a);
    const double iv = 1.0 / M.vol[c];
    for (int i = 0; i < 3; ++i) U[3 * size_t(c) + i] = U[3 * size_t(c) + i] - a[i] * iv;
  }
};

// Non-canonical periodic faces take the negated flux of their canonical partner
// (src/solver/mod_fv_solver.f90:2644-2672, 2871-2897). One boundary face.
struct PeriodicAntisym {
  DevMesh M;
  double* ff;
  DM_HD void operator()(int b) const {
    if (M.bbc[b] != BC_PERIODIC) return;
    const int pr = M.pair[b];
    if (pr < 0) return;
    const int f = M.nint + b, pf = M.nint + pr;
    const int os = M.bown[b], op = M.bown[pr];
    if (os == op) return;
    if (os > op) ff[f] = -ff[pf];
  }
};

// correct_face_fluxes (src/solver/mod_fv_solver.f90:2756-2870): items [0, npf) are pressure-operator
// faces, items [npf, npf+nb) boundary faces (outlets only).
struct CorrectFaceFlux {
  DevMesh M;
  const double *P, *gp, *pob;
  double* ff;
  double dt;
  int ptot, nonortho;
  DM_HD void operator()(int i) const {
    if (i < M.npf) {
      const int o = M.pown[i], nb = M.pnbr[i], fm = M.pfm[i];
      if (o < 0) return;  // skipped face
      ff[fm] = ff[fm] - dt * M.pcoef[i] * (P[nb] - P[o]);
      if (nonortho)
        ff[fm] = ff[fm] - dt * cross_flux(&gp[3 * size_t(o)], &gp[3 * size_t(nb)], M.xgc[i], &M.kint[3 * size_t(i)]);
      return;
    }
    const int b = i - M.npf;
    if (M.bbc[b] != BC_OUTLET) return;
    const int f = M.nint + b, o = M.bown[b];
    ff[f] = ff[f] - dt * M.cb[b] * ((ptot ? pob[b] : 0.0) - P[o]);
    if (nonortho) ff[f] = ff[f] - dt * dot3(&gp[3 * size_t(o)], &M.kbnd[3 * size_t(b)]);
  }
};

// correct_velocity_psolver_dual (src/solver/mod_fv_solver.f90:2927-3151) for one cell.
struct CorrectVelocity {
  DevMesh M;
  const double *P, *gp, *pob;
  double *U, *gpc_out;
  double dt, fb0, fb1, fb2;
  int do_cross, ptot;
  const double* T;  // Boussinesq: the wall term's acceleration carries -bb (T_o - T_ref) g
  double bb, Tref, g0, g1, g2;
  DM_HD void operator()(int c) const {
    const int nc = M.nc;
    double g3[3] = {0, 0, 0};
    const int cnt = M.nB[c];
    for (int k = 0; k < cnt; ++k) {
      const size_t s = size_t(k) * nc + c;
      const int idx = M.bidx[s];
      if (idx < 0) continue;
      const int e = M.bface[s], fm = e >> 1, side = e & 1, oth = M.both[s];
      const int o = side ? oth : c, nb = side ? c : oth;
      double g = M.pcoef[idx] * (P[nb] - P[o]);
      if (do_cross) g = g + cross_flux(&gp[3 * size_t(o)], &gp[3 * size_t(nb)], M.xgc[idx], &M.kint[3 * size_t(idx)]);
      const double* n = &M.nrm[3 * size_t(fm)];
      for (int i = 0; i < 3; ++i) g3[i] = g3[i] + g * n[i];
    }
    for (int k = 0; k < cnt; ++k) {  // Dirichlet-pressure (outlet) faces
      const size_t s = size_t(k) * nc + c;
      if (M.bidx[s] >= 0) continue;
      const int b = -1 - M.both[s];
      if (M.bbc[b] != BC_OUTLET) continue;
      const int f = M.nint + b;
      double g = M.cb[b] * ((ptot ? pob[b] : 0.0) - P[c]);
      if (do_cross) g = g + dot3(&gp[3 * size_t(c)], &M.kbnd[3 * size_t(b)]);
      const double* n = &M.nrm[3 * size_t(f)];
      for (int i = 0; i < 3; ++i) g3[i] = g3[i] + g * n[i];
    }
    if (!do_cross) {  // wall / slip / inlet faces carry dp/dn = f.n
      for (int k = 0; k < cnt; ++k) {
        const size_t s = size_t(k) * nc + c;
        if (M.bidx[s] >= 0) continue;
        const int b = -1 - M.both[s];
        if (M.bbc[b] == BC_OUTLET) continue;
        const int f = M.nint + b;
        const double* n = &M.nrm[3 * size_t(f)];
        const double tb = T ? T[c] : 0.0;
        const double d1 = fb0 - bb * (tb - Tref) * g0, d2 = fb1 - bb * (tb - Tref) * g1,
                     d3 = fb2 - bb * (tb - Tref) * g2;
        const double fn = (d1 * n[0] + d2 * n[1] + d3 * n[2]) * M.area[f];
        for (int i = 0; i < 3; ++i) g3[i] = g3[i] + fn * n[i];
      }
    }
    const double* T = &M.Tinv[9 * size_t(c)];
    double g[3];
    for (int i = 0; i < 3; ++i) g[i] = T[3 * i] * g3[0] + T[3 * i + 1] * g3[1] + T[3 * i + 2] * g3[2];
    for (int i = 0; i < 3; ++i) {
      if (gpc_out) gpc_out[3 * size_t(c) + i] = g[i];
      U[3 * size_t(c) + i] = U[3 * size_t(c) + i] - dt * g[i];
    }
  }
};

// ------------------------------------------------------------------ reductions

// Volume-weighted mean speed for the divergence tolerance (src/solver/mod_fv_solver.f90:1424-1437).
// Momentum sum U V for the constant mass flux forcing (update_mass_flux_forcing,
// src/solver/mod_fv_solver.f90:960-983).
struct Momentum {
  DevMesh M;
  const double* U;
  DM_HD Red<3> operator()(int c) const {
    Red<3> r;
    for (int k = 0; k < 3; ++k) r.v[k] = U[3 * size_t(c) + k] * M.vol[c];
    return r;
  }
};

// sum of U1 + U2 + U3 + p over the cells; a NaN or Inf anywhere makes it non-finite
// (solver_state_is_finite, src/solver/mod_fv_solver.f90:3925-3945)
struct FiniteSum {
  const double *U, *p;
  DM_HD Red<1> operator()(int c) const {
    Red<1> r;
    r.v[0] = U[3 * size_t(c)] + U[3 * size_t(c) + 1] + U[3 * size_t(c) + 2] + p[c];
    return r;
  }
};

struct BulkSpeed {
  DevMesh M;
  const double* U;
  DM_HD Red<2> operator()(int c) const {
    const double* u = &U[3 * size_t(c)];
    Red<2> r;
    r.v[0] = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) * M.vol[c];
    r.v[1] = M.vol[c];
    return r;
  }
};

// Rotational correction p += (nu + nu_t) rhs_p gk_dt / V, then p += p_old (src/solver/mod_fv_solver.f90:
// 1549-1571); returns p V for the mean removal of anchor_pressure_to_outlet (:3610-3639).
struct RotAccumulate {
  DevMesh M;
  double* P;
  const double *p_old, *nus, *rhs;
  double nu, gdt;
  int rotational;
  DM_HD Red<1> operator()(int c) const {
    double p = P[c];
    if (rotational) p = p + (nu + nus[c]) * rhs[c] * gdt * M.iv[c];
    p = p + p_old[c];
    P[c] = p;
    Red<1> r;
    r.v[0] = p * M.vol[c];
    return r;
  }
};

struct ShiftP {
  double* P;
  const double* shift;  // device scalar
  DM_HD void operator()(int c) const { P[c] = P[c] - *shift; }
};

// solver_cfl_rates (src/solver/mod_fv_solver.f90:3718-3862) for one cell.
struct CflRates {
  DevMesh M;
  const double *U, *UF, *nus;
  double nu;
  int have_uface;
  DM_HD Red<2> operator()(int c) const {
    const int nc = M.nc;
    double fs = 0.0, ds = 0.0;
    const int cnt = M.nA[c];
    for (int k = 0; k < cnt; ++k) {
      const size_t s = size_t(k) * nc + c;
      const int e = M.aface[s], f = e >> 1, side = e & 1, oth = M.aoth[s];
      const double* sf = &M.sf[3 * size_t(f)];
      const double area = sqrt(sf[0] * sf[0] + sf[1] * sf[1] + sf[2] * sf[2]);
      const int o = (oth >= 0 && side) ? oth : c;
      const double uf = have_uface ? UF[f] : (U[3 * size_t(o)] * sf[0] + U[3 * size_t(o) + 1] * sf[1] + U[3 * size_t(o) + 2] * sf[2]);
      fs = fs + fabs(uf);
      const double w = area / fmax(M.dcf[f], SMALL);
      if (oth >= 0) {
        const int nb = side ? c : oth;
        const double gc = M.gc[f];
        const double nue = nu + gc * fmax(0.0, nus[o]) + (1.0 - gc) * fmax(0.0, nus[nb]);
        ds = ds + nue * w;
      } else if (M.bbc[-1 - oth] != BC_WALL_MODEL) {
        ds = ds + (nu + fmax(0.0, nus[c])) * w;
      }
    }
    Red<2> r;
    r.v[0] = fs / (M.vol[c] + SMALL);
    r.v[1] = ds / (M.vol[c] + SMALL);
    return r;
  }
};

// compute_stats_sums (src/stats/mod_stats.f90:271-336), cell part: E, V, eps, momentum, nu_t.
struct StatsCells {
  DevMesh M;
  const double *U, *grad, *nus;
  double nu;
  DM_HD Red<7> operator()(int c) const {
    const double vol = M.vol[c];
    const double* u = &U[3 * size_t(c)];
    const double* g = &grad[9 * size_t(c)];
    const double S11 = g[0], S22 = g[4], S33 = g[8];
    const double S12 = 0.5 * (g[1] + g[3]), S13 = 0.5 * (g[2] + g[6]), S23 = 0.5 * (g[5] + g[7]);
    const double ss = S11 * S11 + S22 * S22 + S33 * S33 + 2.0 * (S12 * S12 + S13 * S13 + S23 * S23);
    Red<7> r;
    r.v[0] = 0.5 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) * vol;
    r.v[1] = vol;
    r.v[2] = (nu + nus[c]) * ss * vol;
    r.v[3] = u[0] * vol;
    r.v[4] = u[1] * vol;
    r.v[5] = u[2] * vol;
    r.v[6] = nus[c] * vol;
    return r;
  }
};

struct StatsMax {
  const double* U;
  DM_HD Red<1> operator()(int c) const {
    const double* u = &U[3 * size_t(c)];
    Red<1> r;
    r.v[0] = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    return r;
  }
};

struct StatsWall {
  DevMesh M;
  const double* tau;
  DM_HD Red<2> operator()(int b) const {
    Red<2> r;
    r.v[0] = 0.0;
    r.v[1] = 0.0;
    if (M.bbc[b] != BC_WALL_MODEL) return r;
    const double* t = &tau[3 * size_t(b)];
    const double a = M.area[M.nint + b];
    r.v[0] = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]) * a;
    r.v[1] = a;
    return r;
  }
};

// face_flux_divergence (src/solver/mod_fv_solver.f90:3532-3579) for one cell: {|div|, |div| V}.
// |face-flux divergence| of a cell; weighted = 1 multiplies it by the cell volume (the max and the
// volume integral are two reductions).
struct FluxDivergence {
  DevMesh M;
  const double* ff;
  int weighted;
  DM_HD Red<1> operator()(int c) const {
    const int nc = M.nc;
    double d = 0.0;
    const int cnt = M.nB[c];
    for (int k = 0; k < cnt; ++k) {
      const size_t s = size_t(k) * nc + c;
      const int e = M.bface[s], fm = e >> 1, side = e & 1;
      const double fl = ff[fm];
      if (M.bidx[s] >= 0)
        d = side ? d - fl * M.iv[c] : d + fl * M.iv[c];
    }
    for (int k = 0; k < cnt; ++k) {
      const size_t s = size_t(k) * nc + c;
      if (M.bidx[s] >= 0) continue;
      d = d + ff[M.bface[s] >> 1] * M.iv[c];
    }
    Red<1> r;
    r.v[0] = weighted ? fabs(d) * M.vol[c] : fabs(d);
    return r;
  }
};

// cell_velocity_divergence (src/solver/mod_fv_solver.f90:3582-3607): div u from the
// gc-interpolated cell velocities (boundary faces: U_bnd), owner +, neighbour -, per cell in
// face order; |div| or |div| V.
struct CellDivergence {
  DevMesh M;
  const double *U, *Ub;
  int weighted;
  DM_HD Red<1> operator()(int c) const {
    const int nc = M.nc;
    double d = 0.0;
    const int cnt = M.nA[c];
    for (int k = 0; k < cnt; ++k) {
      const size_t s = size_t(k) * nc + c;
      const int e = M.aface[s], f = e >> 1, side = e & 1, oth = M.aoth[s];
      double uf[3];
      if (oth >= 0) {
        const int o = side ? oth : c, nb = side ? c : oth;
        const double gc = M.gc[f];
        for (int i = 0; i < 3; ++i) uf[i] = gc * U[3 * size_t(o) + i] + (1.0 - gc) * U[3 * size_t(nb) + i];
      } else {
        for (int i = 0; i < 3; ++i) uf[i] = Ub[3 * size_t(-1 - oth) + i];
      }
      const double* sf = &M.sf[3 * size_t(f)];
      const double flux = uf[0] * sf[0] + uf[1] * sf[1] + uf[2] * sf[2];
      d = side ? d - flux * M.iv[c] : d + flux * M.iv[c];
    }
    Red<1> r;
    r.v[0] = weighted ? fabs(d) * M.vol[c] : fabs(d);
    return r;
  }
};

}  // namespace dm
