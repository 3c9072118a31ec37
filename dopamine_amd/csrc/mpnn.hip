// C ABI of the learned preconditioner on its own operator (include/dopamine_mpnn.h): the network's
// forward and backward passes, Adam, training on captured residuals, and Jacobi CG / flexible CG
// solves with it, on the current HIP device.
#include <string>
#include <vector>

#include "dm_device.h"
#include "dopamine_mpnn.h"
#include "mpnn_dev.h"

namespace {
thread_local std::string g_error;

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

struct Ctx {
  dm::DeviceExec x;
  dm::EllOp A;
  dm::CgT<dm::DeviceExec> cg;
  dm::mp::LearnedPc pc;
  std::vector<void*> owned;
  double *b = nullptr, *v = nullptr, *w = nullptr;
  int n = 0;

  template <class T>
  T* up(const T* h, size_t m) {
    T* p = x.alloc<T>(m);
    x.upload(p, h, m);
    owned.push_back(p);
    return p;
  }
  ~Ctx() {
    pc.destroy();
    cg.destroy();
    for (void* p : owned) x.release(p);
    x.destroy();
  }
};
}  // namespace

extern "C" {

const char* dopamine_mpnn_last_error(void) { return g_error.c_str(); }

int dopamine_mpnn_nparams(void) { return dm::mp::NPARAM; }

void* dopamine_mpnn_create(int n, int w, const int32_t* nbr, const double* coef, const double* diag, int capacity,
                           int seed) {
  Ctx* c = nullptr;
  const int rc = guard([&] {
    if (n <= 0 || w < 1 || !nbr || !coef || !diag) throw std::runtime_error("empty operator");
    int dev = 0;
    if (hipGetDeviceCount(&dev) != hipSuccess || dev <= 0) throw std::runtime_error("no HIP device available");
    for (size_t i = 0; i < size_t(w) * n; ++i)
      if (nbr[i] < -1 || nbr[i] >= n) throw std::runtime_error("neighbour index out of range");
    for (int i = 0; i < n; ++i)
      if (!(diag[i] > 0.0)) throw std::runtime_error("diagonal must be positive");
    c = new Ctx();
    c->n = n;
    c->x.init(n);
    c->A.n = n;
    c->A.w = w;
    c->A.nbr = c->up(nbr, size_t(w) * n);
    c->A.coef = c->up(coef, size_t(w) * n);
    c->A.diag = c->up(diag, size_t(n));
    c->b = c->x.alloc<double>(n);
    c->v = c->x.alloc<double>(n);
    c->w = c->x.alloc<double>(n);
    c->owned.push_back(c->b);
    c->owned.push_back(c->v);
    c->owned.push_back(c->w);
    c->cg.init(&c->x, n);
    c->cg.set_single_reduction(true);
    c->pc.init(&c->x, c->A, capacity, uint64_t(seed));
    c->x.sync();
    return 0;
  });
  if (rc != 0) {
    delete c;
    return nullptr;
  }
  return c;
}

void dopamine_mpnn_free(void* h) { delete static_cast<Ctx*>(h); }

int dopamine_mpnn_params(void* h, float* params, int set) {
  return guard([&] {
    auto* c = static_cast<Ctx*>(h);
    if (set)
      c->pc.set_params(params);
    else
      c->pc.get_params(params);
    c->x.sync();
    return 0;
  });
}

int dopamine_mpnn_forward(void* h, const double* r, double* z, uint16_t* acts) {
  return guard([&] {
    auto* c = static_cast<Ctx*>(h);
    c->x.upload(c->v, r, size_t(c->n));
    c->pc.forward(c->v, c->w, acts != nullptr);
    c->x.download(z, c->w, size_t(c->n));
    if (acts) c->pc.activations(acts);
    return 0;
  });
}

int dopamine_mpnn_loss_grad(void* h, const double* r, double* loss, double* grad) {
  return guard([&] {
    auto* c = static_cast<Ctx*>(h);
    c->x.upload(c->v, r, size_t(c->n));
    c->pc.loss_grad(c->v, false, 0);
    if (loss) *loss = c->pc.loss(0);
    if (grad) c->pc.gradient(grad);
    return 0;
  });
}

int dopamine_mpnn_loss_grad_acts(void* h, const double* r, double* loss, double* grad, uint16_t* acts) {
  return guard([&] {
    auto* c = static_cast<Ctx*>(h);
    c->x.upload(c->v, r, size_t(c->n));
    c->pc.loss_grad(c->v, false, 0);
    if (loss) *loss = c->pc.loss(0);
    if (grad) c->pc.gradient(grad);
    if (acts) c->pc.activations(acts);
    return 0;
  });
}

int dopamine_mpnn_adam(void* h, double lr, int count) {
  return guard([&] {
    auto* c = static_cast<Ctx*>(h);
    c->pc.adam(lr, count < 1 ? 1 : count);
    c->x.sync();
    return 0;
  });
}

int dopamine_mpnn_add_sample(void* h, const double* r) {
  return guard([&] {
    auto* c = static_cast<Ctx*>(h);
    c->x.upload(c->v, r, size_t(c->n));
    c->pc.add_sample(c->v);
    c->x.sync();
    return c->pc.samples();
  });
}

int dopamine_mpnn_train(void* h, int iters, int batch, double lr, double* loss_hist) {
  return guard([&] {
    auto* c = static_cast<Ctx*>(h);
    const int rc = c->pc.train(iters, batch, lr, loss_hist);
    c->x.sync();
    if (rc != 0) throw std::runtime_error("no residual samples to train on");
    return 0;
  });
}

int dopamine_mpnn_solve(void* h, int mode, const double* b, double* x, double rtol, double atol, int maxit,
                        int nullspace, int* its, double* resid) {
  return guard([&] {
    auto* c = static_cast<Ctx*>(h);
    if (mode < 0 || mode > 2) throw std::runtime_error("mode must be 0 (Jacobi), 1 (learned) or 2 (Jacobi, capture)");
    c->x.upload(c->b, b, size_t(c->n));
    c->x.upload(c->v, x, size_t(c->n));
    if (mode == 1) c->cg.set_fcg(&c->pc);
    if (mode == 2) {
      c->pc.capturing = true;
      c->cg.set_capture(&c->pc);
    }
    const int rc = c->cg.solve(c->A, c->b, c->v, rtol, atol, nullptr, maxit, nullspace != 0, its, resid);
    c->cg.set_fcg(nullptr);
    c->cg.set_capture(nullptr);
    c->pc.capturing = false;
    c->x.download(x, c->v, size_t(c->n));
    return rc;
  });
}

int dopamine_mpnn_samples(void* h) { return static_cast<Ctx*>(h)->pc.samples(); }

}  // extern "C"
