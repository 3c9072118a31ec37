// Pressure-backend ABI (include/fvm_pressure_backend.h): the symbols of the reference's
// src/solver/pressure_backend_petsc.c, served by the device-resident Jacobi CG on gfx950.
#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "amg_impl.h"
#include "cg_impl.h"
#include "dm_device.h"
#include "fvm_pressure_backend.h"
#include "mpnn_dev.h"

namespace {

std::string g_solver = "boomeramg";
std::vector<std::string> g_options;
int g_local_rank = -1;
int g_io_on_device = 0;
int g_initialized = 0;
int g_last_info[4] = {0, 0, 0, 0};  // first solve's rc and iterations, retried, total iterations

// Cached solve context: rebuilt when (n_local, n_global, n_faces, reference row) change
// (pressure_backend_petsc.c:687-715).
struct Ctx {
  int n_local = 0, n_global = 0, n_faces = 0, ref_row = -1;
  dm::DeviceExec x;
  dm::CgT<dm::DeviceExec> cg;
  dm::EllOp A;
  dm::EllOp32 A32;
  int* nbr = nullptr;
  double *coef = nullptr, *diag = nullptr, *b = nullptr, *sol = nullptr;
  float *coef32 = nullptr, *diag32 = nullptr;
  uint32_t* opair = nullptr;  // the direction pass's 16-bit offsets and their FP32 coefficients
  float* ocoef = nullptr;
  bool mixed = false;
  std::string selector;
  std::unique_ptr<dm::mp::LearnedPolicy> learned;  // selector "fcg-learned"
  std::unique_ptr<dm::AmgPc<dm::DeviceExec>> amg;  // multigrid selectors
  ~Ctx() {
    learned.reset();
    amg.reset();
    cg.destroy();
    for (void* p : {(void*)nbr, (void*)coef, (void*)diag, (void*)b, (void*)sol, (void*)coef32, (void*)diag32, (void*)opair,
                     (void*)ocoef})
      x.release(p);
    x.destroy();
  }
};
std::unique_ptr<Ctx> g_ctx;

// Selectors: the reference's algebraic-multigrid names ("boomeramg", its default, and "gamg") run
// the flexible CG with this backend's aggregation multigrid; "gmres" and "jacobi" the
// Jacobi-preconditioned CG; "mixed" FP32 Jacobi sweeps inside FP64 refinement, "mixed-multigrid"
// FP32 multigrid-preconditioned sweeps inside FP64 refinement; "fcg-learned" the learned
// preconditioner in flexible CG. Any other name fails the solve with -11 instead of silently
// running something else.
bool known_selector(const std::string& s) {
  return s == "boomeramg" || s == "gamg" || s == "gmres" || s == "jacobi" || s == "mixed" || s == "fcg-learned" ||
         s == "mixed-multigrid";
}
bool multigrid_selector(const std::string& s) { return s == "boomeramg" || s == "gamg" || s == "mixed-multigrid"; }
bool mixed_selector(const std::string& s) { return s == "mixed" || s == "mixed-multigrid"; }

int build_ctx(int n, int n_global, int n_faces, const int* l2g, const int* owner, const int* nb_local,
              const int* nb_global, const int* skip, const double* diagonal, const double* coeff, int ref_row) {
  if (n_global != n) return -46;  // one rank: every row is local
  for (int i = 0; i < n; ++i)
    if (l2g[i] != i + 1) return -45;  // compact ids on one rank are the identity
  std::vector<std::vector<std::pair<int, double>>> rows(n);
  for (int f = 0; f < n_faces; ++f) {
    if (skip[f] != 0 || nb_global[f] <= 0) continue;
    const int o = owner[f] - 1;
    if (o < 0 || o >= n) continue;
    const double c = coeff[f];
    if (!std::isfinite(c) || c == 0.0) continue;
    const int g = nb_global[f] - 1;
    if (g < 0 || g >= n) return -16;
    rows[o].push_back({g, c});
    if (nb_local[f] > 0) {
      const int nbl = nb_local[f] - 1;
      if (nbl < 0 || nbl >= n) return -17;
      rows[nbl].push_back({o, c});
    }
  }
  int w = 0;
  for (auto& r : rows) w = std::max<int>(w, int(r.size()));
  std::vector<int> nbr(size_t(std::max(1, w)) * n, -1);
  std::vector<double> cf(size_t(std::max(1, w)) * n, 0.0), dg(n);
  for (int i = 0; i < n; ++i) {
    for (size_t k = 0; k < rows[i].size(); ++k) {
      nbr[k * n + i] = rows[i][k].first;
      cf[k * n + i] = rows[i][k].second;
    }
    double d = diagonal[i];
    if (!std::isfinite(d) || std::fabs(d) < 1.0e-20) d = 1.0;
    dg[i] = d;
  }
  auto ctx = std::make_unique<Ctx>();
  ctx->x.init(n);
  ctx->nbr = ctx->x.alloc<int>(nbr.size());
  ctx->coef = ctx->x.alloc<double>(cf.size());
  ctx->diag = ctx->x.alloc<double>(n);
  ctx->b = ctx->x.alloc<double>(n);
  ctx->sol = ctx->x.alloc<double>(n);
  ctx->x.upload(ctx->nbr, nbr.data(), nbr.size());
  ctx->x.upload(ctx->coef, cf.data(), cf.size());
  ctx->x.upload(ctx->diag, dg.data(), dg.size());
  ctx->A.n = n;
  ctx->A.w = w;
  ctx->A.nbr = ctx->nbr;
  ctx->A.coef = ctx->coef;
  ctx->A.diag = ctx->diag;
  ctx->cg.init(&ctx->x, n);
  if (mixed_selector(g_solver)) {  // FP32 operator copy for the mixed-precision sweeps
    std::vector<float> c32(cf.begin(), cf.end()), d32(dg.begin(), dg.end());
    ctx->coef32 = ctx->x.alloc<float>(c32.size());
    ctx->diag32 = ctx->x.alloc<float>(n);
    ctx->x.upload(ctx->coef32, c32.data(), c32.size());
    ctx->x.upload(ctx->diag32, d32.data(), d32.size());
    ctx->A32.n = n;
    ctx->A32.w = w;
    ctx->A32.nbr = ctx->nbr;
    ctx->A32.coef = ctx->coef32;
    ctx->A32.diag = ctx->diag32;
    ctx->cg.set_mixed(ctx->A32, 1.0e-4);
    ctx->mixed = true;
    std::vector<uint32_t> pr;
    std::vector<float> oc;
    if (dm::off16_enabled() && dm::ell_pack32(nbr.data(), cf.data(), n, w, pr, oc, ctx->A.obase)) {
      ctx->opair = ctx->x.alloc<uint32_t>(pr.size());
      ctx->ocoef = ctx->x.alloc<float>(oc.size());
      ctx->x.upload(ctx->opair, pr.data(), pr.size());
      ctx->x.upload(ctx->ocoef, oc.data(), oc.size());
      ctx->A.opair = ctx->opair;
      ctx->A.ocoef = ctx->ocoef;
    }
  }
  if (multigrid_selector(g_solver)) {
    dm::AmgOp op;
    op.n = n;
    op.w = std::max(1, w);
    op.nbr = nbr;
    op.coef = cf;
    op.diag = dg;
    ctx->amg.reset(new dm::AmgPc<dm::DeviceExec>());
    ctx->amg->build(&ctx->x, op, ctx->nbr, dm::AmgCfg{});
    if (ctx->mixed)
      ctx->cg.set_pc32(ctx->amg.get());
    else
      ctx->cg.set_fcg(ctx->amg.get());
  }
  if (g_solver == "fcg-learned") {
    // options: -learn_collect N, -learn_capacity N, -learn_iters N, -learn_batch N, -learn_lr X,
    // -learn_seed N (the solver's precond = 1 parameters)
    dm::mp::LearnCfg cfg;
    for (const auto& o : g_options) {
      const auto val = [&](const char* key) -> const char* {
        const size_t k = std::string(key).size();
        return o.rfind(key, 0) == 0 && o.size() > k ? o.c_str() + k : nullptr;
      };
      if (const char* v = val("-learn_collect ")) cfg.collect = std::max(1, std::atoi(v));
      if (const char* v = val("-learn_capacity ")) cfg.capacity = std::max(1, std::atoi(v));
      if (const char* v = val("-learn_iters ")) cfg.iters = std::max(0, std::atoi(v));
      if (const char* v = val("-learn_batch ")) cfg.batch = std::max(1, std::atoi(v));
      if (const char* v = val("-learn_lr ")) cfg.lr = std::atof(v) > 0.0 ? std::atof(v) : cfg.lr;
      if (const char* v = val("-learn_seed ")) cfg.seed = uint64_t(std::atoi(v));
    }
    ctx->learned.reset(new dm::mp::LearnedPolicy(&ctx->x, &ctx->cg, ctx->A, cfg));
  }
  ctx->selector = g_solver;
  ctx->n_local = n;
  ctx->n_global = n_global;
  ctx->n_faces = n_faces;
  ctx->ref_row = ref_row;
  g_ctx = std::move(ctx);
  return 0;
}

// ---- the allocation-free variant (fvm_petsc_solve_pressure_noalloc)
//
// Its device memory is this file-scope array, one copy per GPU, sized for operators of at most
// NOALLOC_MAX_CELLS rows and NOALLOC_MAX_WIDTH off-diagonal entries per row: the ELL operator
// (neighbours, coefficients, diagonal), then the CG's work vectors and partials. Host code never
// uses the array's name as an address: scratch_ptr() looks the device address up once per device
// (hipGetSymbolAddress) and everything else uses that pointer. The ELL rows are assembled in
// host arrays of static storage. Every call on a GPU reuses the same memory, so calls that overlap
// (on one GPU, or from several host threads) race; each call writes every word it reads.
constexpr int NOALLOC_MAX_CELLS = 131072;
constexpr int NOALLOC_MAX_WIDTH = 8;
constexpr int NOALLOC_MAX_DEVICES = 64;
constexpr size_t noalloc_al(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t NOALLOC_NBR_BYTES = noalloc_al(sizeof(int) * NOALLOC_MAX_WIDTH * size_t(NOALLOC_MAX_CELLS));
constexpr size_t NOALLOC_COEF_BYTES = noalloc_al(sizeof(double) * NOALLOC_MAX_WIDTH * size_t(NOALLOC_MAX_CELLS));
constexpr size_t NOALLOC_DIAG_BYTES = noalloc_al(sizeof(double) * size_t(NOALLOC_MAX_CELLS));
constexpr size_t NOALLOC_OP_BYTES = NOALLOC_NBR_BYTES + NOALLOC_COEF_BYTES + NOALLOC_DIAG_BYTES;
// the CG's work vectors and partials for the largest operator (the pass grid never exceeds PGRID_MAX)
constexpr size_t NOALLOC_CG_BYTES = 6 * noalloc_al(sizeof(double) * size_t(NOALLOC_MAX_CELLS)) +
                                    noalloc_al(sizeof(dm::CgScalars)) +
                                    noalloc_al(sizeof(double) * 17 * size_t(dm::PGRID_MAX));
constexpr size_t NOALLOC_BYTES = NOALLOC_OP_BYTES + NOALLOC_CG_BYTES;

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

unsigned char* g_scratch_dev[NOALLOC_MAX_DEVICES] = {};
int h_cnt[NOALLOC_MAX_CELLS];
int h_nbr[NOALLOC_MAX_WIDTH * NOALLOC_MAX_CELLS];
double h_coef[NOALLOC_MAX_WIDTH * NOALLOC_MAX_CELLS];
double h_diag[NOALLOC_MAX_CELLS];
double h_part[dm::PGRID_MAX];
char g_noalloc_error[256] = "";

// Device address of this GPU's scratch array (looked up on the first call on each device).
unsigned char* scratch_ptr() {
  int dev = 0;
  DM_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) throw std::runtime_error("device index beyond the scratch table");
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    DM_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_dev[dev] = static_cast<unsigned char*>(p);
  }
  return g_scratch_dev[dev];
}

}  // namespace

extern "C" {

const char* dopamine_backend_last_error(void) { return g_noalloc_error; }

int fvm_petsc_solve_pressure_noalloc(int n_cells_local, int n_cells_global, int n_faces, const int* local_to_global,
                                     const int* face_owner, const int* face_neighbor_local,
                                     const int* face_neighbor_global, const int* skip_face, const double* diagonal,
                                     const double* rhs, const double* face_coeff, int reference_cell,
                                     int max_iterations, double rtol, double atol, const int* dirichlet_rows,
                                     const double* dirichlet_values, int n_dirichlet_rows, double* solution,
                                     int* iteration_count, double* residual_norm) {
  (void)dirichlet_rows;
  (void)dirichlet_values;
  (void)reference_cell;
  g_noalloc_error[0] = '\0';
  const auto fail = [&](int code, const char* what) {
    std::snprintf(g_noalloc_error, sizeof(g_noalloc_error), "%s", what);
    return code;
  };
  try {
    const int n = n_cells_local;
    if (n <= 0) return fail(-10, "no cells");
    if (n > NOALLOC_MAX_CELLS)
      return fail(-61, "fvm_petsc_solve_pressure_noalloc: more cells than its scratch holds (131072)");
    if (n_cells_global != n) return fail(-46, "one rank: every row is local");
    for (int i = 0; i < n; ++i)
      if (local_to_global[i] != i + 1) return fail(-45, "compact ids on one rank are the identity");
    // ELL rows in face order, owner side then neighbour side (the slot order of the cached path)
    const auto keep = [&](int f) {
      const double c = face_coeff[f];
      return skip_face[f] == 0 && face_neighbor_global[f] > 0 && face_owner[f] >= 1 && face_owner[f] <= n &&
             std::isfinite(c) && c != 0.0;
    };
    for (int i = 0; i < n; ++i) h_cnt[i] = 0;
    for (int f = 0; f < n_faces; ++f) {
      if (!keep(f)) continue;
      const int o = face_owner[f] - 1, g = face_neighbor_global[f] - 1;
      if (g < 0 || g >= n) return fail(-16, "face neighbour out of range");
      h_cnt[o]++;
      if (face_neighbor_local[f] > 0) {
        const int nl = face_neighbor_local[f] - 1;
        if (nl < 0 || nl >= n) return fail(-17, "face neighbour out of range");
        h_cnt[nl]++;
      }
    }
    int w = 1;
    for (int i = 0; i < n; ++i) w = std::max(w, h_cnt[i]);
    if (w > NOALLOC_MAX_WIDTH)
      return fail(-62, "fvm_petsc_solve_pressure_noalloc: a row with more than 8 neighbours");
    for (int i = 0; i < w * n; ++i) {
      h_nbr[i] = -1;
      h_coef[i] = 0.0;
    }
    for (int i = 0; i < n; ++i) h_cnt[i] = 0;
    for (int f = 0; f < n_faces; ++f) {
      if (!keep(f)) continue;
      const int o = face_owner[f] - 1, g = face_neighbor_global[f] - 1;
      const double c = face_coeff[f];
      h_nbr[size_t(h_cnt[o]) * n + o] = g;
      h_coef[size_t(h_cnt[o]++) * n + o] = c;
      if (face_neighbor_local[f] > 0) {
        const int nl = face_neighbor_local[f] - 1;
        h_nbr[size_t(h_cnt[nl]) * n + nl] = o;
        h_coef[size_t(h_cnt[nl]++) * n + nl] = c;
      }
    }
    for (int i = 0; i < n; ++i) {
      double d = diagonal[i];
      if (!std::isfinite(d) || std::fabs(d) < 1.0e-20) d = 1.0;
      h_diag[i] = d;
    }
    unsigned char* base = scratch_ptr();
    dm::DeviceExec ex;  // the null stream; no partials buffer, no timer: nothing allocated
    ex.stream = nullptr;
    int* d_nbr = reinterpret_cast<int*>(base);
    double* d_coef = reinterpret_cast<double*>(base + NOALLOC_NBR_BYTES);
    double* d_diag = reinterpret_cast<double*>(base + NOALLOC_NBR_BYTES + NOALLOC_COEF_BYTES);
    ex.upload(d_nbr, h_nbr, size_t(w) * n);
    ex.upload(d_coef, h_coef, size_t(w) * n);
    ex.upload(d_diag, h_diag, size_t(n));
    dm::CgT<dm::DeviceExec> cg;
    if (dm::CgT<dm::DeviceExec>::external_bytes(n, ex.pass_grid(n)) > NOALLOC_CG_BYTES)
      return fail(-63, "scratch too small for the CG workspace");
    cg.init_external(&ex, n, base + NOALLOC_OP_BYTES, h_part);
    dm::EllOp A;
    A.n = n;
    A.w = w;
    A.nbr = d_nbr;
    A.coef = d_coef;
    A.diag = d_diag;
    int maxit = max_iterations > 2000 ? max_iterations : 2000;
    for (const auto& o : g_options)
      if (o.rfind("-ksp_max_it", 0) == 0) {
        const int v = std::atoi(o.c_str() + 11);
        maxit = v > 0 ? v : (max_iterations > 2000 ? max_iterations : 2000);
      }
    const bool ns = n_dirichlet_rows <= 0;
    int its = 0, total = 0, retried = 0;
    double res = 0.0;
    auto once = [&](int mi, int* it, double* r) { return cg.solve(A, rhs, solution, rtol, atol, nullptr, mi, ns, it, r); };
    const int rc = dm::solve_with_retry(
        maxit, rtol, [&](int* it, double* r) { return once(maxit, it, r); }, once,
        [&] { ex.zero(solution, size_t(n)); }, [&] { return atol; }, &its, &res, &total, &retried);
    ex.sync();
    *iteration_count = its;
    *residual_norm = res;
    return rc;
  } catch (const std::exception& e) {  // a HIP error reported by the runtime
    std::snprintf(g_noalloc_error, sizeof(g_noalloc_error), "%s", e.what());
    return -33;
  }
}


int fvm_petsc_backend_available(void) {
  int n = 0;
  return (hipGetDeviceCount(&n) == hipSuccess && n > 0) ? 1 : 0;
}

int fvm_petsc_backend_initialize(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return -3;
  if (g_local_rank >= 0 && hipSetDevice(g_local_rank % n) != hipSuccess) return -3;
  g_initialized = 1;
  return 0;
}

int fvm_petsc_backend_finalize(void) {
  try {
    g_ctx.reset();
  } catch (...) {
    return -20;
  }
  g_initialized = 0;
  return 0;
}

void fvm_petsc_set_pressure_solver(const char* name) {
  if (name) g_solver = name;
}

void fvm_petsc_insert_option(const char* opt) {
  if (opt && opt[0]) g_options.emplace_back(opt);
}

void fvm_petsc_set_device_local_rank(int local_rank) { g_local_rank = local_rank; }

int fvm_petsc_device_io_supported(void) { return 1; }

void fvm_petsc_set_io_on_device(int on) { g_io_on_device = on; }

int fvm_petsc_solve_pressure(int n_cells_local, int n_cells_global, int n_faces, const int* local_to_global,
                             const int* face_owner, const int* face_neighbor_local, const int* face_neighbor_global,
                             const int* skip_face, const double* diagonal, const double* rhs,
                             const double* face_coeff, int reference_cell, int max_iterations, double rtol,
                             double atol, const int* dirichlet_rows, const double* dirichlet_values,
                             int n_dirichlet_rows, double* solution, int* iteration_count, double* residual_norm) {
  (void)dirichlet_rows;
  (void)dirichlet_values;
  try {
    if (n_cells_local <= 0) return -10;
    if (!known_selector(g_solver)) return -11;
    int ref_row = reference_cell - 1;
    if (ref_row < 0 || ref_row >= n_cells_global) ref_row = 0;
    if (!g_ctx || g_ctx->n_local != n_cells_local || g_ctx->n_global != n_cells_global || g_ctx->n_faces != n_faces ||
        g_ctx->ref_row != ref_row || g_ctx->selector != g_solver) {
      g_ctx.reset();
      const int rc = build_ctx(n_cells_local, n_cells_global, n_faces, local_to_global, face_owner,
                               face_neighbor_local, face_neighbor_global, skip_face, diagonal, face_coeff, ref_row);
      if (rc != 0) return rc;
    }
    Ctx& c = *g_ctx;
    const int n = n_cells_local;
    const double* b = rhs;
    double* x = solution;
    if (!g_io_on_device) {
      c.x.upload(c.b, rhs, n);
      c.x.upload(c.sol, solution, n);
      b = c.b;
      x = c.sol;
    }
    int maxit = max_iterations > 2000 ? max_iterations : 2000;
    bool single = false;
    for (const auto& o : g_options) {
      if (o.rfind("-ksp_max_it", 0) == 0) {  // a later "-ksp_max_it 0" restores the default
        const int v = std::atoi(o.c_str() + 11);
        maxit = v > 0 ? v : (max_iterations > 2000 ? max_iterations : 2000);
      }
      // PETSc's single-synchronisation CG types select the one-pass recurrence
      if (o.rfind("-ksp_type", 0) == 0)
        single = o.find("pipecg") != std::string::npos || o.find("groppcg") != std::string::npos;
    }
    c.cg.set_single_reduction(single);
    const bool ns = n_dirichlet_rows <= 0;
    int its = 0, total = 0, retried = 0;
    double res = 0.0;
    // a solve stopped at its limit: once more from x = 0 with 4x the limit, then the soft accept
    // or -(1000 + KSP_DIVERGED_ITS) (pressure_backend_petsc.c:878-937)
    const int rc = dm::solve_with_retry(
        maxit, rtol,
        [&](int* it, double* r) {
          const int q = c.learned ? c.learned->solve(c.A, b, x, rtol, atol, nullptr, maxit, ns, it, r)
                                  : c.cg.solve(c.A, b, x, rtol, atol, nullptr, maxit, ns, it, r);
          g_last_info[0] = q;
          g_last_info[1] = *it;
          return q;
        },
        [&](int mi, int* it, double* r) {
          c.cg.set_plain(true);  // the retry runs the Jacobi passes
          const int q = c.cg.solve(c.A, b, x, rtol, atol, nullptr, mi, ns, it, r);
          c.cg.set_plain(false);
          return q;
        },
        [&] { c.x.zero(x, size_t(n)); }, [&] { return atol; }, &its, &res, &total, &retried);
    g_last_info[2] = retried;
    g_last_info[3] = total;
    if (!g_io_on_device) c.x.download(solution, c.sol, n);
    c.x.sync();
    *iteration_count = its;
    *residual_norm = res;
    return rc;
  } catch (const std::exception&) {
    return -33;
  }
}

void dopamine_backend_last_solve(int out[4]) {
  for (int i = 0; i < 4; ++i) out[i] = g_last_info[i];
}

int dopamine_mpi_gpu_aware(void) {
  const char* env = std::getenv("DOPAMINE_GPU_AWARE_MPI");
  if (env != nullptr && env[0] != '\0') return env[0] != '0';
  return 0;
}

void dopamine_nvtx_push(const char* name) { roctxRangePushA(name); }
void dopamine_nvtx_pop(void) { roctxRangePop(); }

}  // extern "C"
