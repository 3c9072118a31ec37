// Host-check build of the product solver: the same kernel bodies and driver as
// libdopamine_hip.so (dopamine_amd/csrc/solver_impl.h), executed by HostExec on the CPU, so the
// CPU test suite can compare the product's indexing and operation order with the oracle without
// a GPU. Test infrastructure only; the product library never contains HostExec.
#include <algorithm>
#include <string>

#include "solver_impl.h"

namespace {
using Solver = dm::SolverT<dm::HostExec>;
thread_local std::string g_error;

double* field(Solver* s, const std::string& name, int64_t* n) {
  const int64_t nc = s->nc, nb = std::max(1, s->nb), nf = s->nf;
  if (name == "U") { *n = 3 * nc; return s->U; }
  if (name == "p") { *n = nc; return s->P; }
  if (name == "nu_sgs") { *n = nc; return s->nu_sgs; }
  if (name == "U_bnd") { *n = 3 * nb; return s->U_bnd; }
  if (name == "p_bnd") { *n = nb; return s->p_bnd; }
  if (name == "grad_U") { *n = 9 * nc; return s->grad; }
  if (name == "U_face") { *n = nf; return s->UF; }
  if (name == "tau_w_bnd") { *n = 3 * nb; return s->tau; }
  if (name == "rhs_p") { *n = nc; return s->rhs; }
  if (name == "w_gp") { *n = 3 * nc; return s->gp; }
  if (name == "outlet_pb") { *n = nb; return s->outlet_pb; }
  if (name == "wm_filt") { *n = 3 * nb; return s->wm_filt; }
  if (s->q_w && name == "q_w_bnd") { *n = nb; return s->q_w; }
  if (s->q_w && name == "L_bnd") { *n = nb; return s->L_bnd; }
  if (s->bous && name == "T") { *n = nc; return s->T; }
  if (s->bous && name == "T_bnd") { *n = nb; return s->T_bnd; }
  if (s->bous && name == "grad_T") { *n = 3 * nc; return s->gT; }
  for (int k = 0; k < s->ns; ++k) {
    const std::string c = "C" + std::to_string(k);
    if (name == c) { *n = nc; return s->Cs[k]; }
    if (name == c + "_bnd") { *n = nb; return s->Cb[k]; }
  }
  *n = 0;
  return nullptr;
}
}  // namespace

extern "C" {
const char* hc_last_error(void) { return g_error.c_str(); }

void* hc_solver_create(const dopamine_mesh_desc* m, const dopamine_params* p) {
  try {
    auto* s = new Solver();
    s->create(*m, *p);
    return s;
  } catch (const std::exception& e) {
    g_error = e.what();
    return nullptr;
  }
}
void hc_solver_destroy(void* h) { delete static_cast<Solver*>(h); }
int hc_solver_stats_csv(void* h, const char* path, int step, double dt, double* E0, int append) {
  return static_cast<Solver*>(h)->stats_csv(path, step, dt, E0, append);
}
void hc_solver_cell_divergence(void* h, double out[2]) { static_cast<Solver*>(h)->cell_divergence(out); }
int hc_solver_write_checkpoint(void* h, const char* dir, int step, int rank) {
  try {
    return static_cast<Solver*>(h)->write_checkpoint(dir, step, rank);
  } catch (const std::exception& e) {
    g_error = e.what();
    return -100;
  }
}
int hc_solver_read_checkpoint(void* h, const char* dir, int step, int rank) {
  try {
    return static_cast<Solver*>(h)->read_checkpoint(dir, step, rank);
  } catch (const std::exception& e) {
    g_error = e.what();
    return -100;
  }
}
// binary16 storage of the multigrid couplings: round to nearest even and back
void hc_coef_roundtrip(const float* in, uint16_t* h, float* out, int n) {
  for (int i = 0; i < n; ++i) {
    h[i] = dm::coef_h(in[i]);
    out[i] = dm::coef_f(h[i]);
  }
}
double* hc_solver_field(void* h, const char* name, int64_t* n) { return field(static_cast<Solver*>(h), name, n); }
void hc_solver_apply_all_bc(void* h) { static_cast<Solver*>(h)->apply_all_bc(); }
int hc_solver_project_once(void* h, double dt) { return static_cast<Solver*>(h)->project_once(dt); }
int hc_solver_advance(void* h, double dt, double* used) { return static_cast<Solver*>(h)->advance(dt, used); }
void hc_solver_stats(void* h, double out[8]) { static_cast<Solver*>(h)->stats(out); }
double hc_solver_time(void* h) { return static_cast<Solver*>(h)->time; }
void hc_solver_counters(void* h, int64_t out[4]) {
  auto* s = static_cast<Solver*>(h);
  out[0] = s->n_solves; out[1] = s->n_iters; out[2] = s->last_iters; out[3] = s->n_sweeps;
}
void hc_kfactor_limited(const double sf[3], const double d[3], double min_cos, double k[3]) {
  dm::kfactor_limited(sf, d, min_cos, k);
}
void hc_solver_proj_counters(void* h, int64_t out[2]) {
  out[0] = static_cast<Solver*>(h)->n_proj;
  out[1] = static_cast<Solver*>(h)->n_proj_its;
}
int hc_solver_state_is_finite(void* h) { return static_cast<Solver*>(h)->state_is_finite() ? 1 : 0; }
int64_t hc_solver_retries(void* h) { return static_cast<Solver*>(h)->n_retries; }
const char* hc_solver_error(void* h) { return static_cast<Solver*>(h)->err.c_str(); }
// The product's CG with the reference's retry (solve_with_retry) on an ELL operator, with any
// iteration limit (the solver path clamps it to >= 2000), so the retry, soft-accept and failure
// branches can be driven on small systems.
int hc_cg_solve_retry(int n, int w, const int* nbr, const double* coef, const double* diag, const double* b,
                      double* xv, double rtol, double atol, int maxit, int nullspace, int single, int mixed, int* its,
                      double* resid, int* total, int* retried) {
  dm::HostExec ex;
  dm::CgT<dm::HostExec> cg;
  cg.init(&ex, n);
  cg.set_single_reduction(single != 0);
  if (mixed) cg.set_mixed(dm::EllOp32{}, 1.0e-4);
  dm::EllOp A;
  A.n = n;
  A.w = w;
  A.nbr = nbr;
  A.coef = coef;
  A.diag = diag;
  const bool ns = nullspace != 0;
  auto once = [&](int mi, int* it, double* r) { return cg.solve(A, b, xv, rtol, atol, nullptr, mi, ns, it, r); };
  const int rc = dm::solve_with_retry(
      maxit, rtol, [&](int* it, double* r) { return once(maxit, it, r); }, once,
      [&] { std::fill(xv, xv + n, 0.0); }, [&] { return cg.last_atol(); }, its, resid, total, retried);
  cg.destroy();
  return rc;
}
// The product's flexible CG with the aggregation multigrid (AmgPc) on an ELL operator: FP64
// flexible CG (mixed = 0) or FP32 sweeps in FP64 refinement (mixed = 1). cfg = {theta, omega,
// alpha, coarse_max}; info receives {levels, sizes of up to 14 levels...}.
// The 16-bit offset form of an ELL neighbour list (dm::ell_off16): returns the padded width W
// (0: does not fit); base receives 8 slot bases, d and src W * n entries each (room for 8 * n).
int hc_ell_off16(int n, int w, const int* nbr, int* base, int16_t* d, int* src) {
  dm::EllOff16 o;
  if (!dm::ell_off16(nbr, n, w, o)) return 0;
  std::copy(o.base, o.base + 8, base);
  std::copy(o.d.begin(), o.d.end(), d);
  std::copy(o.src.begin(), o.src.end(), src);
  return o.W;
}
int hc_amg_solve(int n, int w, const int* nbr, const double* coef, const double* diag, const double* b, double* xv,
                 double rtol, double atol, int maxit, int nullspace, int mixed, const double cfg[5], int* its,
                 double* resid, int* info) {
  try {
    dm::HostExec ex;
    dm::AmgOp op;
    op.n = n;
    op.w = w;
    op.nbr.assign(nbr, nbr + size_t(w) * n);
    op.coef.assign(coef, coef + size_t(w) * n);
    op.diag.assign(diag, diag + n);
    dm::AmgCfg c;
    c.theta = cfg[0];
    c.omega = cfg[1];
    c.alpha = cfg[2];
    c.coarse_max = int(cfg[3]);
    c.theta0 = cfg[4];  // < 0: no aggressive first coarsening
    dm::AmgPc<dm::HostExec> amg;
    amg.build(&ex, op, nbr, c);
    info[0] = amg.levels();
    for (int l = 0; l < amg.levels() && l < 14; ++l) info[1 + l] = amg.sizes()[l];
    dm::CgT<dm::HostExec> cg;
    cg.init(&ex, n);
    if (mixed) {
      cg.set_mixed(dm::EllOp32{}, 1.0e-4);
      cg.set_pc32(&amg);
    } else {
      cg.set_fcg(&amg);
    }
    dm::EllOp A;
    A.n = n;
    A.w = w;
    A.nbr = nbr;
    A.coef = coef;
    A.diag = diag;
    std::vector<uint32_t> pr;  // the direction pass's compact operator, as the solver attaches it
    std::vector<float> oc;
    if (mixed && dm::off16_enabled() && dm::ell_pack32(nbr, coef, n, w, pr, oc, A.obase)) {
      A.opair = pr.data();
      A.ocoef = oc.data();
    }
    const int rc = cg.solve(A, b, xv, rtol, atol, nullptr, maxit, nullspace != 0, its, resid);
    info[15] = cg.last_sweeps();
    cg.destroy();
    return rc;
  } catch (const std::exception& e) {
    g_error = e.what();
    return -100;
  }
}
// One V-cycle z = M r of the aggregation multigrid (FP32 inside, FP64 in and out) and its
// hierarchy: agg_out receives the level-0 aggregate of every cell (n entries).
int hc_amg_vcycle(int n, int w, const int* nbr, const double* coef, const double* diag, const double* r, double* z,
                  const double cfg[5], int* agg_out, int* info) {
  try {
    dm::HostExec ex;
    dm::AmgOp op;
    op.n = n;
    op.w = w;
    op.nbr.assign(nbr, nbr + size_t(w) * n);
    op.coef.assign(coef, coef + size_t(w) * n);
    op.diag.assign(diag, diag + n);
    dm::AmgCfg c;
    c.theta = cfg[0];
    c.omega = cfg[1];
    c.alpha = cfg[2];
    c.coarse_max = int(cfg[3]);
    c.theta0 = cfg[4];
    std::vector<int> agg;
    dm::amg_aggregate(op, c.theta, agg);
    std::copy(agg.begin(), agg.end(), agg_out);
    dm::AmgPc<dm::HostExec> amg;
    amg.build(&ex, op, nbr, c);
    info[0] = amg.levels();
    for (int l = 0; l < amg.levels() && l < 14; ++l) info[1 + l] = amg.sizes()[l];
    amg.vcycle(r, z);
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return -100;
  }
}
void hc_solver_nonortho_info(void* h, int32_t out[3]) {
  auto* s = static_cast<Solver*>(h);
  out[0] = s->skip_nonortho; out[1] = s->skip_grad_p; out[2] = s->n_corr;
}
}
