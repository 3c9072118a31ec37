// C ABI of the HIP solver (include/dopamine_hip.h): SolverT over DeviceExec on gfx950.
#include <string>

#include "dm_device.h"
#include "mpnn_dev.h"
#include "solver_impl.h"

namespace {
thread_local std::string g_error;
using Solver = dm::SolverT<dm::DeviceExec>;

template <class F>
int guard(F&& f) {
  try {
    g_error.clear();
    return f();
  } catch (const std::exception& e) {
    g_error = e.what();
    return -100;
  }
}

struct FieldRef {
  double* p;
  int64_t n;
};

FieldRef field(Solver* s, const std::string& name) {
  const int64_t nc = s->nc, nb = std::max(1, s->nb), nf = s->nf;
  if (name == "U") return {s->U, 3 * nc};
  if (name == "p") return {s->P, nc};
  if (name == "nu_sgs") return {s->nu_sgs, nc};
  if (name == "U_bnd") return {s->U_bnd, 3 * nb};
  if (name == "p_bnd") return {s->p_bnd, nb};
  if (name == "grad_U") return {s->grad, 9 * nc};
  if (name == "U_face") return {s->UF, nf};
  if (name == "tau_w_bnd") return {s->tau, 3 * nb};
  if (name == "rhs_p") return {s->rhs, nc};
  if (name == "H") return {s->H0, 3 * nc};  // after a step the last stage's H sits in H0
  if (name == "w_gp") return {s->gp, 3 * nc};
  if (name == "outlet_pb") return {s->outlet_pb, nb};
  if (name == "wm_filt") return {s->wm_filt, 3 * nb};
  if (s->q_w) {
    if (name == "q_w_bnd") return {s->q_w, nb};
    if (name == "L_bnd") return {s->L_bnd, nb};
  }
  if (s->bous) {
    if (name == "T") return {s->T, nc};
    if (name == "T_bnd") return {s->T_bnd, nb};
    if (name == "grad_T") return {s->gT, 3 * nc};
  }
  for (int k = 0; k < s->ns; ++k) {
    const std::string c = "C" + std::to_string(k);
    if (name == c) return {s->Cs[k], nc};
    if (name == c + "_bnd") return {s->Cb[k], nb};
    if (name == "grad_" + c) return {s->gC[k], 3 * nc};
  }
  throw std::runtime_error("unknown field " + name);
}
}  // namespace

extern "C" {

const char* dopamine_last_error(void) { return g_error.c_str(); }

int dopamine_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

void* dopamine_solver_create(const dopamine_mesh_desc* m, const dopamine_params* p) {
  Solver* s = nullptr;
  const int rc = guard([&] {
    if (!m || !p) throw std::runtime_error("null mesh or params");
    if (dopamine_device_count() <= 0) throw std::runtime_error("no HIP device available");
    if (p->wall_model < 0 || p->wall_model > 2)
      throw std::runtime_error("wall_model must be 0 (log-law), 1 (none) or 2 (MOST)");
    if (p->scheme < 1 || p->scheme > 5)
      throw std::runtime_error("advection scheme must be VANLEER_TVD (1), KOREN_TVD (2), CDS (3), KEP (4) or QUICK (5)");
    if (p->precision != 0 && p->precision != 1) throw std::runtime_error("precision must be 0 (FP64) or 1 (mixed)");
    if (p->cg_variant != 0 && p->cg_variant != 1) throw std::runtime_error("cg_variant must be 0 or 1");
    if (p->boussinesq && (p->t_scheme < 1 || p->t_scheme > 5 || p->t_scheme == 4))
      throw std::runtime_error("temperature scheme must be VANLEER_TVD (1), KOREN_TVD (2), CDS (3) or QUICK (5)");
    if (p->precond < 0 || p->precond > 2)
      throw std::runtime_error("precond must be 0 (Jacobi), 1 (learned) or 2 (aggregation multigrid)");
    if (p->precond == 1 && p->precision == 1)
      throw std::runtime_error("the learned preconditioner runs in the FP64 flexible CG only: precision 1 with "
                               "precond 1 is not supported");
    if (p->mass_flux && (p->f_const[0] != 0.0 || p->f_const[1] != 0.0 || p->f_const[2] != 0.0))
      throw std::runtime_error("f_const must be 0 with constant mass flux forcing");
    if (p->n_scalars < 0 || p->n_scalars > 4) throw std::runtime_error("n_scalars must be 0..4");
    if (p->n_scalars && p->scalar_scheme != dm::SCHEME_CDS && p->scalar_scheme != dm::SCHEME_QUICK &&
        p->scalar_scheme != dm::SCHEME_VANLEER && p->scalar_scheme != dm::SCHEME_KOREN)
      throw std::runtime_error("scalar_scheme must be CDS, QUICK, VANLEER_TVD or KOREN_TVD");
    if (p->cell_projection != 0 && p->cell_projection != 1)
      throw std::runtime_error("cell_projection must be 0 (LS) or 1 (EXACT)");
    if (p->cell_projection == 1 && !(p->cell_projection_tol > 0.0))
      throw std::runtime_error("cell_projection_tol must be positive");
    if (p->precond == 1 && (p->learn_collect < 1 || p->learn_capacity < 1 || p->learn_iters < 0 ||
                            p->learn_batch < 1 || !(p->learn_lr > 0.0)))
      throw std::runtime_error("learned preconditioner: learn_collect, learn_capacity, learn_batch >= 1, "
                               "learn_iters >= 0, learn_lr > 0 required");
    s = new Solver();
    s->x.init(std::max(m->n_cells, m->n_faces + m->n_bnd_faces));
    s->create(*m, *p);
    if (p->precond == 1) {
      dm::mp::LearnCfg cfg;
      cfg.collect = p->learn_collect;
      cfg.capacity = p->learn_capacity;
      cfg.iters = p->learn_iters;
      cfg.batch = p->learn_batch;
      cfg.lr = p->learn_lr;
      cfg.seed = uint64_t(p->learn_seed);
      s->policy.reset(new dm::mp::LearnedPolicy(&s->x, &s->cg, s->A, cfg));
      s->x.sync();
    }
    return 0;
  });
  if (rc != 0) {
    delete s;
    return nullptr;
  }
  return s;
}

void dopamine_solver_destroy(void* h) {
  auto* s = static_cast<Solver*>(h);
  if (!s) return;
  try {
    s->x.sync();
  } catch (...) {
  }
  dm::DeviceExec x = s->x;
  delete s;
  x.destroy();
}

int dopamine_solver_set_field(void* h, const char* name, const double* host) {
  return guard([&] {
    auto* s = static_cast<Solver*>(h);
    FieldRef f = field(s, name);
    s->x.upload(f.p, host, size_t(f.n));
    const std::string n(name);
    if (n == "U" || n == "T" || (n.size() == 2 && n[0] == 'C')) s->apply_all_bc();
    s->x.sync();
    return 0;
  });
}

int dopamine_solver_get_field(void* h, const char* name, double* host, int64_t len) {
  return guard([&] {
    auto* s = static_cast<Solver*>(h);
    FieldRef f = field(s, name);
    if (len < f.n) throw std::runtime_error("output buffer too small for " + std::string(name));
    s->x.download(host, f.p, size_t(f.n));
    return 0;
  });
}

int dopamine_solver_project_once(void* h, double dt) {
  return guard([&] {
    auto* s = static_cast<Solver*>(h);
    const int rc = s->project_once(dt);
    if (rc != 0) g_error = s->err.empty() ? "project_once failed (" + std::to_string(rc) + ")" : s->err;
    return rc;
  });
}

int dopamine_solver_advance(void* h, double dt_req, double* dt_used) {
  return guard([&] {
    auto* s = static_cast<Solver*>(h);
    const int rc = s->advance(dt_req, dt_used);
    if (rc != 0) g_error = s->err.empty() ? "advance failed (" + std::to_string(rc) + ")" : s->err;
    return rc;
  });
}

int dopamine_solver_cell_net_flux(void* h, double* out) {
  return guard([&] {
    *out = static_cast<Solver*>(h)->cell_net_flux_max();
    return 0;
  });
}

int dopamine_solver_state_is_finite(void* h) {
  int ok = 0;
  const int rc = guard([&] {
    ok = static_cast<Solver*>(h)->state_is_finite() ? 1 : 0;
    return 0;
  });
  return rc != 0 ? rc : ok;
}

double dopamine_solver_compute_dt(void* h) {
  double dt = -1.0;
  guard([&] {
    dt = static_cast<Solver*>(h)->compute_dt();
    return 0;
  });
  return dt;
}

int dopamine_solver_stats(void* h, double out[8]) {
  return guard([&] {
    static_cast<Solver*>(h)->stats(out);
    return 0;
  });
}

double dopamine_solver_time(void* h) { return static_cast<Solver*>(h)->time; }

int dopamine_solver_counters(void* h, int64_t out[4]) {
  auto* s = static_cast<Solver*>(h);
  out[0] = s->n_solves;
  out[1] = s->n_iters;
  out[2] = s->last_iters;
  out[3] = s->n_sweeps;
  return 0;
}

int64_t dopamine_solver_retries(void* h) { return static_cast<Solver*>(h)->n_retries; }

int dopamine_solver_nonortho_info(void* h, int32_t out[3]) {
  auto* s = static_cast<Solver*>(h);
  out[0] = s->skip_nonortho;
  out[1] = s->skip_grad_p;
  out[2] = s->n_corr;
  return 0;
}

int dopamine_solver_kernel_timing(void* h, int kernel, int enable) {
  return guard([&] {
    static_cast<Solver*>(h)->x.timing(kernel, enable != 0);
    return 0;
  });
}

int dopamine_solver_kernel_time(void* h, int kernel, int64_t* launches, double* total_ms) {
  return guard([&] {
    auto* s = static_cast<Solver*>(h);
    if (kernel < 0 || kernel >= dm::K_NCLASS) throw std::runtime_error("unknown kernel class");
    s->x.sync();
    *launches = s->x.kt->launches[kernel];
    *total_ms = s->x.kt->total_ms[kernel];
    return 0;
  });
}

int dopamine_solver_write_checkpoint(void* h, const char* dir, int step, int rank) {
  return guard([&] {
    if (!dir) throw std::runtime_error("null directory");
    return static_cast<Solver*>(h)->write_checkpoint(dir, step, rank);
  });
}

int dopamine_solver_read_checkpoint(void* h, const char* dir, int step, int rank) {
  return guard([&] {
    if (!dir) throw std::runtime_error("null directory");
    return static_cast<Solver*>(h)->read_checkpoint(dir, step, rank);
  });
}

int dopamine_solver_cell_divergence(void* h, double out[2]) {
  return guard([&] {
    static_cast<Solver*>(h)->cell_divergence(out);
    return 0;
  });
}

int dopamine_solver_stats_csv(void* h, const char* path, int step, double dt, double* E0, int append) {
  return guard([&] {
    if (!path || !E0) throw std::runtime_error("null path or E0");
    if (static_cast<Solver*>(h)->stats_csv(path, step, dt, E0, append) != 0)
      throw std::runtime_error(std::string("cannot open ") + path);
    return 0;
  });
}

int dopamine_solver_learned_info(void* h, double out[8]) {
  return guard([&] {
    auto* lp = dynamic_cast<dm::mp::LearnedPolicy*>(static_cast<Solver*>(h)->policy.get());
    if (!lp) return -1;
    const double v[8] = {double(lp->phase),        double(lp->collected),    double(lp->pc().samples()),
                         double(lp->learned_solves), double(lp->learned_its), double(lp->fallbacks),
                         lp->ref_its,                lp->train_ms};
    for (int i = 0; i < 8; ++i) out[i] = v[i];
    return 0;
  });
}

int dopamine_solver_learned_precond_info(void* h, int64_t out[2]) {
  return guard([&] {
    auto* lp = dynamic_cast<dm::mp::LearnedPolicy*>(static_cast<Solver*>(h)->policy.get());
    if (!lp) return -1;
    out[0] = lp->precond_solves;
    out[1] = lp->precond_its;
    return 0;
  });
}

int dopamine_solver_multigrid_info(void* h, int64_t* out, int max) {
  return guard([&] {
    auto* s = static_cast<Solver*>(h);
    if (!s->amg) throw std::runtime_error("no multigrid hierarchy (precond is not multigrid)");
    const auto& sz = s->amg->sizes();
    const auto& nz = s->amg->nnz();
    const auto& st = s->amg->stored();
    const int nl = int(sz.size());
    if (max < 2 + 3 * nl) throw std::runtime_error("multigrid_info: output holds fewer than 2 + 3 levels entries");
    out[0] = nl;
    out[1] = s->amg->aggressive() ? 1 : 0;
    for (int l = 0; l < nl; ++l) out[2 + l] = sz[size_t(l)];
    for (int l = 0; l < nl; ++l) out[2 + nl + l] = l == 0 ? 0 : (size_t(l - 1) < nz.size() ? nz[size_t(l - 1)] : 0);
    for (int l = 0; l < nl; ++l) out[2 + 2 * nl + l] = l == 0 ? 0 : (size_t(l - 1) < st.size() ? st[size_t(l - 1)] : 0);
    return 0;
  });
}

int dopamine_solver_projection_info(void* h, int64_t out[2]) {
  return guard([&] {
    auto* s = static_cast<Solver*>(h);
    out[0] = s->n_proj;
    out[1] = s->n_proj_its;
    return 0;
  });
}

int dopamine_solver_learned_loss(void* h, double* out, int max) {
  return guard([&] {
    auto* lp = dynamic_cast<dm::mp::LearnedPolicy*>(static_cast<Solver*>(h)->policy.get());
    if (!lp) return -1;
    const int m = std::min<int>(max, int(lp->loss_hist.size()));
    for (int i = 0; i < m; ++i) out[i] = lp->loss_hist[i];
    return m;
  });
}

int dopamine_solver_sync(void* h) {
  return guard([&] {
    static_cast<Solver*>(h)->x.sync();
    return 0;
  });
}

}  // extern "C"

// ---------------------------------------------------------------- measurement utility

namespace {
__global__ void __launch_bounds__(256) k_hbm_copy(const double* __restrict__ a, double* __restrict__ b, int64_t n) {
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) b[i] = a[i];
}
}  // namespace

extern "C" int dopamine_hbm_copy_probe(int64_t n, int reps, double* avg_ms) {
  return guard([&] {
    if (n <= 0 || reps <= 0) throw std::runtime_error("hbm copy probe: n and reps must be positive");
    double *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    DM_CHECK(hipMalloc(&a, size_t(n) * sizeof(double)));
    DM_CHECK(hipMalloc(&b, size_t(n) * sizeof(double)));
    DM_CHECK(hipStreamCreate(&st));
    DM_CHECK(hipEventCreate(&e0));
    DM_CHECK(hipEventCreate(&e1));
    DM_CHECK(hipMemsetAsync(a, 0, size_t(n) * sizeof(double), st));
    const int grid = 1024;
    hipLaunchKernelGGL(k_hbm_copy, dim3(grid), dim3(256), 0, st, a, b, n);  // warm
    float total = 0.0f;
    for (int r = 0; r < reps; ++r) {
      float ms = 0.0f;
      hipExtLaunchKernelGGL(k_hbm_copy, dim3(grid), dim3(256), 0, st, e0, e1, 0, (const double*)a, b, n);
      DM_CHECK(hipEventSynchronize(e1));
      DM_CHECK(hipEventElapsedTime(&ms, e0, e1));
      total += ms;
    }
    *avg_ms = double(total) / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
  });
}

// ---------------------------------------------------------------- operators on their own
// The numerics and SGS routines of the reference's Fortran modules as single calls on the
// solver's mesh (the Fortran drop-in modules bind them): host arrays in and out in the
// reference's layout, copied to scratch device buffers, one kernel launch per routine.
namespace {
struct OpBufs {
  dm::DeviceExec& x;
  std::vector<void*> ptrs;
  explicit OpBufs(dm::DeviceExec& e) : x(e) {}
  ~OpBufs() {
    try {
      x.sync();
    } catch (...) {
    }
    for (void* p : ptrs) x.release(p);
  }
  template <class T>
  T* in(const T* h, size_t n) {
    T* d = x.alloc<T>(n);
    ptrs.push_back(d);
    if (h) x.upload(d, h, n);
    return d;
  }
  double* zeros(size_t n) { return in<double>(nullptr, n); }
  template <class T>
  void out(T* h, const T* d, size_t n) {
    x.download(h, d, n);
  }
};
struct Sizes {
  size_t nc, nf, nb;
};
Sizes sizes(Solver* s) { return {size_t(s->nc), size_t(s->nf), size_t(std::max(1, s->nb))}; }
Solver* solver_of(void* h) {
  if (!h) throw std::runtime_error("null solver handle");
  return static_cast<Solver*>(h);
}
}  // namespace

extern "C" {

int dopamine_op_sizes(void* h, int64_t out[3]) {
  return guard([&] {
    Solver* s = solver_of(h);
    out[0] = s->nc;
    out[1] = s->nf;
    out[2] = s->nb;
    return 0;
  });
}

int dopamine_op_ls_gradient_vector(void* h, const double* U, const double* U_bnd, double* grad_U) {
  return guard([&] {
    Solver* s = solver_of(h);
    const Sizes z = sizes(s);
    OpBufs b(s->x);
    double* g = b.zeros(9 * z.nc);
    s->x.for_each(s->nc, dm::LsGradVec{s->M, b.in(U, 3 * z.nc), b.in(U_bnd, 3 * z.nb), g, nullptr, -1, 0.0, 0.0, 0.0,
                                       0.0});
    b.out(grad_U, g, 9 * z.nc);
    return 0;
  });
}

int dopamine_op_ls_gradient_scalar(void* h, const double* phi, const double* phi_bnd, double* grad) {
  return guard([&] {
    Solver* s = solver_of(h);
    const Sizes z = sizes(s);
    OpBufs b(s->x);
    double* g = b.zeros(3 * z.nc);
    s->x.for_each(s->nc, dm::LsGradSca{s->M, b.in(phi, z.nc), nullptr, b.in(phi_bnd, z.nb), nullptr, g});
    b.out(grad, g, 3 * z.nc);
    return 0;
  });
}

int dopamine_op_gg_gradient_vector(void* h, const double* U, const double* U_bnd, const double* grad_prev,
                                   double* grad_U) {
  return guard([&] {
    Solver* s = solver_of(h);
    const Sizes z = sizes(s);
    OpBufs b(s->x);
    double* g = b.zeros(9 * z.nc);
    s->x.for_each(s->nc, dm::GgGradVec{s->M, b.in(U, 3 * z.nc), b.in(U_bnd, 3 * z.nb),
                                       grad_prev ? b.in(grad_prev, 9 * z.nc) : nullptr, g});
    b.out(grad_U, g, 9 * z.nc);
    return 0;
  });
}

int dopamine_op_gg_gradient_scalar(void* h, const double* phi, const double* phi_bnd, const double* grad_prev,
                                   double* grad) {
  return guard([&] {
    Solver* s = solver_of(h);
    const Sizes z = sizes(s);
    OpBufs b(s->x);
    double* g = b.zeros(3 * z.nc);
    s->x.for_each(s->nc, dm::GgGradSca{s->M, b.in(phi, z.nc), nullptr, b.in(phi_bnd, z.nb), nullptr,
                                       grad_prev ? b.in(grad_prev, 3 * z.nc) : nullptr, g});
    b.out(grad, g, 3 * z.nc);
    return 0;
  });
}

int dopamine_op_sgs_viscosity(void* h, int model, const double* U, const double* grad_U, double coef, double nu,
                              int wall_gradient, double dt_stab, double* nu_sgs) {
  return guard([&] {
    Solver* s = solver_of(h);
    if (model < 1 || model > 3) throw std::runtime_error("sgs model must be 1 (Vreman), 2 (Smagorinsky) or 3 (AMD)");
    const Sizes z = sizes(s);
    OpBufs b(s->x);
    double* nus = b.zeros(z.nc);
    const double* g = b.in(grad_U, 9 * z.nc);
    s->x.for_each(s->nc, dm::SgsCell{s->M, g, nus, model, model == 1 ? coef : 0.0, model == 2 ? coef : 0.0,
                                     model == 3 ? coef : 0.0, dt_stab});
    if (model == 3 && wall_gradient && s->n_amdw > 0)
      s->x.for_each(s->n_amdw, dm::AmdWall{s->M, s->amdw_cell, s->amdw_off, s->amdw_face, b.in(U, 3 * z.nc), g, nus,
                                           nu, coef, dt_stab});
    b.out(nu_sgs, nus, z.nc);
    return 0;
  });
}

int dopamine_op_dynamic_smagorinsky(void* h, const double* U, const double* grad_U, const double* U_bnd,
                                    double dt_stab, double* nu_sgs) {
  return guard([&] {
    Solver* s = solver_of(h);
    if (!s->M.Gf) throw std::runtime_error("dynamic Smagorinsky needs the filter-width tensor: create the solver with an SGS model");
    const Sizes z = sizes(s);
    OpBufs b(s->x);
    const double* u = b.in(U, 3 * z.nc);
    const double* g = b.in(grad_U, 9 * z.nc);
    const double* ub = b.in(U_bnd, 3 * z.nb);
    double *fb = b.zeros(15 * z.nc), *fh = b.zeros(15 * z.nc), *uh = b.zeros(3 * z.nc), *gh = b.zeros(9 * z.nc);
    double *lm = b.zeros(2 * z.nc), *lmh = b.zeros(2 * z.nc), *nus = b.zeros(z.nc);
    s->x.for_each(s->nc, dm::DynPack{u, g, fb});
    s->x.for_each(s->nc, dm::DynFilter<15>{s->M, fb, fh, uh});
    s->x.for_each(s->nc, dm::GgGradVec{s->M, uh, ub, nullptr, gh});
    s->x.for_each(s->nc, dm::DynLM{s->M, fh, gh, lm});
    s->x.for_each(s->nc, dm::DynFilter<2>{s->M, lm, lmh, nullptr});
    s->x.for_each(s->nc, dm::DynNu{s->M, lmh, g, nus, dt_stab});
    b.out(nu_sgs, nus, z.nc);
    return 0;
  });
}

int dopamine_op_wall_model(void* h, const double* U, double nu, double h_wm, double t_filt, double dt,
                           double tf_fac, double* tau_w_bnd) {
  return guard([&] {
    Solver* s = solver_of(h);
    const double hs = s->prm.wall_model_height;
    if ((h_wm > 0.0 || hs > 0.0) && h_wm != hs)
      throw std::runtime_error("h_wm must equal the solver's wall_model_height (" + std::to_string(hs) +
                               "): the sampling cells are found at create");
    const Sizes z = sizes(s);
    OpBufs b(s->x);
    double* t = b.zeros(3 * z.nb);
    const int fmode = tf_fac > dm::SMALL ? 2 : (t_filt > dm::SMALL ? 1 : 0);
    dm::WallModel w{s->M, b.in(U, 3 * z.nc), t, nu};
    if (fmode) {
      w.uf = s->wm_filt;
      w.fmode = fmode;
      w.seed = s->wm_filt_ready ? 0 : 1;
      w.w = fmode == 1 ? dt / (t_filt + dt) : 0.0;
      w.dt_use = dt;
      w.tf = tf_fac;
    }
    if (s->nb > 0) s->x.for_each(s->nb, w);
    if (fmode) s->wm_filt_ready = 1;
    b.out(tau_w_bnd, t, 3 * z.nb);
    return 0;
  });
}

// part 1: advection_rhs_vector, part 2: viscous_rhs_vector
static int op_rhs_vector(void* h, int part, int scheme, double nu, const double* U, const double* U_bnd,
                  const double* face_flux, const double* nu_sgs, const double* grad_U, double* rhs) {
  return guard([&] {
    Solver* s = solver_of(h);
    if (scheme < 1 || scheme > 5) throw std::runtime_error("unknown advection scheme");
    const Sizes z = sizes(s);
    OpBufs b(s->x);
    const double* u = b.in(U, 3 * z.nc);
    const double* ub = b.in(U_bnd, 3 * z.nb);
    const double* uf = face_flux ? b.in(face_flux, z.nf) : b.zeros(z.nf);
    const double* nus = nu_sgs ? b.in(nu_sgs, z.nc) : b.zeros(z.nc);
    const double* g = grad_U ? b.in(grad_U, 9 * z.nc) : b.zeros(9 * z.nc);
    double* H = b.zeros(3 * z.nc);
    double* Un = b.zeros(3 * z.nc);
    double* phiF = (part == 1 && scheme != dm::SCHEME_CDS && scheme != dm::SCHEME_KEP) ? b.zeros(3 * z.nf) : nullptr;
    const auto run = [&](auto sch) {
      constexpr int S = decltype(sch)::value;
      if (phiF) s->x.for_each(s->nf, dm::ReconFace<S>{s->M, u, ub, uf, g, phiF});
      dm::ExplicitRhs<S> f{s->M, u, ub, uf, g, nus, nullptr, H, H, Un, part == 2 ? nu : 0.0, 0.0, 0.0, 0.0, 0.0,
                           0.0, 0.0, 0, phiF, nullptr, 0.0, 0.0, 0.0, 0.0, 0.0, 0, 0.0, 0.0, 0.0, part};
      s->x.for_each(s->nc, f);
    };
    switch (part == 2 ? int(dm::SCHEME_CDS) : scheme) {
      case dm::SCHEME_VANLEER: run(std::integral_constant<int, dm::SCHEME_VANLEER>{}); break;
      case dm::SCHEME_KOREN: run(std::integral_constant<int, dm::SCHEME_KOREN>{}); break;
      case dm::SCHEME_CDS: run(std::integral_constant<int, dm::SCHEME_CDS>{}); break;
      case dm::SCHEME_KEP: run(std::integral_constant<int, dm::SCHEME_KEP>{}); break;
      default: run(std::integral_constant<int, dm::SCHEME_QUICK>{}); break;
    }
    std::vector<double> hv(3 * z.nc);
    b.out(hv.data(), H, 3 * z.nc);
    if (part == 1)
      for (size_t i = 0; i < 3 * z.nc; ++i) rhs[i] = hv[i];  // advection overwrites (mod_advection.f90:724)
    else
      for (size_t i = 0; i < 3 * z.nc; ++i) rhs[i] = rhs[i] + hv[i];  // viscous adds (mod_viscous.f90:37-283)
    return 0;
  });
}

int dopamine_op_advection_rhs_vector(void* h, int scheme, const double* U, const double* U_bnd,
                                     const double* face_flux, const double* grad_U, double* rhs) {
  return op_rhs_vector(h, 1, scheme, 0.0, U, U_bnd, face_flux, nullptr, grad_U, rhs);
}

int dopamine_op_viscous_rhs_vector(void* h, double nu, const double* U, const double* nu_sgs, const double* grad_U,
                                   const double* U_bnd, double* rhs) {
  return op_rhs_vector(h, 2, dm::SCHEME_CDS, nu, U, U_bnd, nullptr, nu_sgs, grad_U, rhs);
}

// part 1: advection_rhs_scalar, part 2: viscous_rhs_scalar
static int op_rhs_scalar(void* h, int part, int scheme, const double* phi, const double* U, const double* phi_bnd,
                  const double* U_bnd, const double* face_flux, const double* grad_phi, double alpha,
                  const double* nu_sgs, double pr_t, double* rhs) {
  return guard([&] {
    Solver* s = solver_of(h);
    if (scheme < 1 || scheme > 5 || scheme == dm::SCHEME_KEP)
      throw std::runtime_error("scalar advection scheme must be VANLEER_TVD, KOREN_TVD, CDS or QUICK");
    if (part == 2 && !(pr_t > 0.0)) throw std::runtime_error("Pr_t must be positive");
    const Sizes z = sizes(s);
    OpBufs b(s->x);
    const double* t = b.in(phi, z.nc);
    const double* tb = b.in(phi_bnd, z.nb);
    const double* u = U ? b.in(U, 3 * z.nc) : b.zeros(3 * z.nc);
    const double* ub = U_bnd ? b.in(U_bnd, 3 * z.nb) : b.zeros(3 * z.nb);
    const double* uf = face_flux ? b.in(face_flux, z.nf) : b.zeros(z.nf);
    const double* gt = grad_phi ? b.in(grad_phi, 3 * z.nc) : b.zeros(3 * z.nc);
    const double* nus = nu_sgs ? b.in(nu_sgs, z.nc) : b.zeros(z.nc);
    double* gv = b.zeros(9 * z.nc);  // the velocity gradient the reconstruction context does not use here
    double* H = b.zeros(z.nc);
    double* Tn = b.zeros(z.nc);
    const bool cds = part == 2 || scheme == dm::SCHEME_CDS;
    double* phiT = cds ? nullptr : b.zeros(z.nf);
    if (!cds) {
      const dm::ReconFace<dm::SCHEME_QUICK> rq{s->M, u, ub, uf, gv, nullptr};
      const dm::ReconFace<dm::SCHEME_VANLEER> rv{s->M, u, ub, uf, gv, nullptr};
      const dm::ReconFace<dm::SCHEME_KOREN> rk{s->M, u, ub, uf, gv, nullptr};
      if (scheme == dm::SCHEME_QUICK) s->x.for_each(s->nf, dm::ReconFaceScalar<dm::SCHEME_QUICK>{rq, t, tb, gt, phiT});
      if (scheme == dm::SCHEME_VANLEER)
        s->x.for_each(s->nf, dm::ReconFaceScalar<dm::SCHEME_VANLEER>{rv, t, tb, gt, phiT});
      if (scheme == dm::SCHEME_KOREN) s->x.for_each(s->nf, dm::ReconFaceScalar<dm::SCHEME_KOREN>{rk, t, tb, gt, phiT});
    }
    dm::ExplicitRhsT f{s->M, t, tb, gt, uf, nus, H, phiT, H, Tn, part == 2 ? alpha : 0.0, part == 2 ? pr_t : 1.0, 0.0,
                       0.0, 0.0, cds ? 1 : 0, nullptr, part};
    s->x.for_each(s->nc, f);
    std::vector<double> hv(z.nc);
    b.out(hv.data(), H, z.nc);
    if (part == 1)
      for (size_t i = 0; i < z.nc; ++i) rhs[i] = hv[i];
    else
      for (size_t i = 0; i < z.nc; ++i) rhs[i] = rhs[i] + hv[i];
    return 0;
  });
}

int dopamine_op_advection_rhs_scalar(void* h, int scheme, const double* phi, const double* U, const double* phi_bnd,
                                     const double* U_bnd, const double* face_flux, const double* grad_phi,
                                     double* rhs) {
  return op_rhs_scalar(h, 1, scheme, phi, U, phi_bnd, U_bnd, face_flux, grad_phi, 0.0, nullptr, 1.0, rhs);
}

int dopamine_op_viscous_rhs_scalar(void* h, const double* phi, double alpha_mol, const double* nu_sgs, double pr_t,
                                   const double* grad_phi, const double* phi_bnd, double* rhs) {
  return op_rhs_scalar(h, 2, dm::SCHEME_CDS, phi, nullptr, phi_bnd, nullptr, nullptr, grad_phi, alpha_mol, nu_sgs,
                       pr_t, rhs);
}

}  // extern "C"
