// Jacobi-preconditioned conjugate gradient for the pressure Poisson operator.
//
// Replaces the reference's PETSc KSPCG solve (src/solver/pressure_backend_petsc.c:648-1034) with
// the same recurrence (Hestenes-Stiefel CG, left Jacobi preconditioning), the same stopping test
// ||r||_2 <= max(rtol*||b'||_2, atol) on the unpreconditioned residual, the constant null space
// removed from b and from every preconditioned residual when no Dirichlet row exists, and the
// true residual ||b - A x||_2 reported at the end. Two fused passes per iteration; convergence is
// decided on the device (later passes become no-ops), so the host only polls between batches.
//
// Operator storage: ELLPACK rows, slot-major (slot k of every row contiguous) so that the 64 lanes
// of a wavefront read consecutive rows. A_ii = diag_i, A_ij = -coef for each listed j.
#pragma once

#include "dm_common.h"

#include <algorithm>

namespace dm {

struct CgScalars {
  double bmean, ttol, rr, rz, rz_old, pq, zmean, resid;
  double rtol, atol;
  int its, done, maxit, first, nullspace, atol_from_dev;
};

struct EllOp {
  int n = 0, w = 0;
  const int* nbr = nullptr;      // w*n, -1 = empty slot
  const double* coef = nullptr;  // w*n
  const double* diag = nullptr;  // n (already repaired: non-finite or tiny -> 1)
};

namespace cgk {

struct SumB {
  const double* b;
  double *bw, *x;
  DM_HD Red<1> operator()(int c) const {
    double bi = b[c];
    if (!isfinite(bi)) bi = 0.0;
    bw[c] = bi;
    if (!isfinite(x[c])) x[c] = 0.0;
    Red<1> r;
    r.v[0] = bi;
    return r;
  }
};
struct SumBFin {
  CgScalars* sc;
  int n;
  const double* atol_dev;
  DM_HD void operator()(const Red<1>& t) const {
    sc->bmean = sc->nullspace ? t.v[0] / double(n) : 0.0;
    if (atol_dev) sc->atol = *atol_dev;
  }
};

struct Init {
  EllOp A;
  double *bw, *r;
  const double* x;
  const CgScalars* sc;
  DM_HD Red<2> operator()(int c) const {
    const int n = A.n;
    const double bp = bw[c] - sc->bmean;
    bw[c] = bp;
    double ax = A.diag[c] * x[c];
    for (int k = 0; k < A.w; ++k) {
      const int j = A.nbr[size_t(k) * n + c];
      if (j < 0) continue;
      ax = ax - A.coef[size_t(k) * n + c] * x[j];
    }
    const double rc = bp - ax;
    r[c] = rc;
    Red<2> o;
    o.v[0] = bp * bp;
    o.v[1] = rc * rc;
    return o;
  }
};
struct InitFin {
  CgScalars* sc;
  DM_HD void operator()(const Red<2>& t) const {
    sc->ttol = fmax(sc->rtol * sqrt(t.v[0]), sc->atol);
    sc->rr = t.v[1];
    sc->its = 0;
    sc->first = 1;
    sc->done = (sqrt(t.v[1]) <= sc->ttol) ? 1 : 0;
  }
};

struct Z {
  const double *diag, *r;
  double* z;
  DM_HD Red<3> operator()(int c) const {
    const double rc = r[c], zc = rc / diag[c];
    z[c] = zc;
    Red<3> o;
    o.v[0] = rc * zc;
    o.v[1] = zc;
    o.v[2] = rc;
    return o;
  }
};
struct ZFin {
  CgScalars* sc;
  int n;
  DM_HD void operator()(const Red<3>& t) const {
    sc->zmean = sc->nullspace ? t.v[1] / double(n) : 0.0;
    sc->rz = t.v[0] - sc->zmean * t.v[2];
  }
};

// p = (z - zmean) + beta p_old for this row and, recomputed on the fly, for its neighbours;
// q = A p; partial p.q.
struct Dir {
  EllOp A;
  const double *z, *pold;
  double *pnew, *q;
  const CgScalars* sc;
  DM_HD Red<1> operator()(int c) const {
    const int n = A.n;
    const int first = sc->first;
    const double zm = sc->zmean;
    const double beta = first ? 0.0 : sc->rz / sc->rz_old;
    const double pc = first ? (z[c] - zm) : (z[c] - zm) + beta * pold[c];
    double qc = A.diag[c] * pc;
    for (int k = 0; k < A.w; ++k) {
      const int j = A.nbr[size_t(k) * n + c];
      if (j < 0) continue;
      const double pj = first ? (z[j] - zm) : (z[j] - zm) + beta * pold[j];
      qc = qc - A.coef[size_t(k) * n + c] * pj;
    }
    pnew[c] = pc;
    q[c] = qc;
    Red<1> o;
    o.v[0] = pc * qc;
    return o;
  }
};
struct DirFin {
  CgScalars* sc;
  DM_HD void operator()(const Red<1>& t) const { sc->pq = t.v[0]; }
};

// x += alpha p, r -= alpha q, ||r||, z = D^-1 r and the next r.z (KSPCG order: count, residual
// norm, test, then the next preconditioner apply).
struct Upd {
  const double *diag, *p, *q;
  double *x, *r, *z;
  const CgScalars* sc;
  DM_HD Red<4> operator()(int c) const {
    const double a = sc->rz / sc->pq;
    x[c] = x[c] + a * p[c];
    const double rc = r[c] - a * q[c];
    r[c] = rc;
    const double zc = rc / diag[c];
    z[c] = zc;
    Red<4> o;
    o.v[0] = rc * rc;
    o.v[1] = rc * zc;
    o.v[2] = zc;
    o.v[3] = rc;
    return o;
  }
};
struct UpdFin {
  CgScalars* sc;
  int n;
  DM_HD void operator()(const Red<4>& t) const {
    sc->its = sc->its + 1;
    sc->rr = t.v[0];
    if (sqrt(t.v[0]) <= sc->ttol)
      sc->done = 1;
    else if (sc->its >= sc->maxit)
      sc->done = 2;
    sc->zmean = sc->nullspace ? t.v[2] / double(n) : 0.0;
    sc->rz_old = sc->rz;
    sc->rz = t.v[1] - sc->zmean * t.v[3];
    sc->first = 0;
  }
};

struct Resid {
  EllOp A;
  const double *b, *x;
  DM_HD Red<1> operator()(int c) const {
    const int n = A.n;
    double ax = A.diag[c] * x[c];
    for (int k = 0; k < A.w; ++k) {
      const int j = A.nbr[size_t(k) * n + c];
      if (j < 0) continue;
      ax = ax - A.coef[size_t(k) * n + c] * x[j];
    }
    double bi = b[c];
    if (!isfinite(bi)) bi = 0.0;
    Red<1> o;
    o.v[0] = (bi - ax) * (bi - ax);
    return o;
  }
};
struct ResidFin {
  CgScalars* sc;
  DM_HD void operator()(const Red<1>& t) const { sc->resid = sqrt(t.v[0]); }
};

}  // namespace cgk

template <class X>
class CgT {
 public:
  void init(X* x, int n) {
    x_ = x;
    n_ = n;
    bw_ = x->template alloc<double>(n);
    r_ = x->template alloc<double>(n);
    z_ = x->template alloc<double>(n);
    pa_ = x->template alloc<double>(n);
    pb_ = x->template alloc<double>(n);
    q_ = x->template alloc<double>(n);
    sc_ = x->template alloc<CgScalars>(1);
  }
  void destroy() {
    if (!x_) return;
    for (void* p : {(void*)bw_, (void*)r_, (void*)z_, (void*)pa_, (void*)pb_, (void*)q_, (void*)sc_}) x_->release(p);
    x_ = nullptr;
  }
  // Solve A x = b (pointers in the execution space). atol_dev, when not null, points to an
  // absolute tolerance computed in the execution space (overrides atol). Returns 0 if converged.
  int solve(const EllOp& A, const double* b, double* x, double rtol, double atol, const double* atol_dev, int maxit,
            bool nullspace, int* its, double* resid) {
    using namespace cgk;
    X& xx = *x_;
    const int n = A.n;
    CgScalars h{};
    h.rtol = rtol;
    h.atol = atol;
    h.maxit = maxit;
    h.nullspace = nullspace ? 1 : 0;
    xx.upload(sc_, &h, 1);
    xx.template reduce<1, false>(n, SumB{b, bw_, x}, SumBFin{sc_, n, atol_dev});
    xx.template reduce<2, false>(n, Init{A, bw_, r_, x, sc_}, InitFin{sc_});
    xx.template reduce<3, false>(n, Z{A.diag, r_, z_}, ZFin{sc_, n}, &sc_->done);
    int launched = 0;
    int batch = std::max(4, std::min(maxit, predict_ - 1));
    double* pold = pa_;
    double* pnew = pb_;
    for (;;) {
      for (int i = 0; i < batch; ++i) {
        xx.template reduce<1, false>(n, Dir{A, z_, pold, pnew, q_, sc_}, DirFin{sc_}, &sc_->done);
        xx.template reduce<4, false>(n, Upd{A.diag, pnew, q_, x, r_, z_, sc_}, UpdFin{sc_, n}, &sc_->done);
        std::swap(pold, pnew);
      }
      launched += batch;
      xx.download(&h, sc_, 1);
      if (h.done || launched > maxit + 1) break;
      batch = std::max(2, std::min(16, (h.its + 7) / 8));
    }
    xx.template reduce<1, false>(n, Resid{A, b, x}, ResidFin{sc_});
    xx.download(&h, sc_, 1);
    last_its_ = h.its;
    predict_ = std::max(2, h.its);
    if (its) *its = h.its;
    if (resid) *resid = h.resid;
    return h.done == 1 ? 0 : -1;
  }
  int last_its() const { return last_its_; }

 private:
  X* x_ = nullptr;
  int n_ = 0;
  double *bw_ = nullptr, *r_ = nullptr, *z_ = nullptr, *pa_ = nullptr, *pb_ = nullptr, *q_ = nullptr;
  CgScalars* sc_ = nullptr;
  int last_its_ = 0, predict_ = 24;
};

}  // namespace dm
