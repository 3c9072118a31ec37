// Jacobi-preconditioned conjugate gradient for the pressure Poisson operator.
//
// Replaces the reference's PETSc KSPCG solve (src/solver/pressure_backend_petsc.c:648-1034) with
// the same recurrence (Hestenes-Stiefel CG, left Jacobi preconditioning), the same stopping test
// ||r||_2 <= max(rtol*||b'||_2, atol) on the unpreconditioned residual, the constant null space
// removed from b and from every preconditioned residual when no Dirichlet row exists, and the
// true residual ||b - A x||_2 reported at the end. Two fused passes per iteration; convergence is
// decided on the device (later passes become no-ops), so the host only polls between batches.
// Every dot product is a fixed-order sum, so a solve is bitwise reproducible.
//
// Operator storage: ELLPACK rows, slot-major (slot k of every row contiguous) so that the 64 lanes
// of a wavefront read consecutive rows. A_ii = diag_i, A_ij = -coef for each listed j.
#pragma once

#include "dm_common.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <stdexcept>
#include <type_traits>
#include <vector>

namespace dm {

struct CgScalars {
  double ttol, rr, rz_hist[2], resid;
  double rtol, atol;
  double gamma_hist[2], alpha_hist[2];  // single-reduction variant
  double inner_tol2;                    // mixed precision: inner stop on ||r_in||^2
  int its, done, maxit, nullspace;
  int inner_done, inner_its;            // mixed precision: inner solve state
  double alpha32;                       // FP32 flexible CG: the last update's alpha
};

struct EllOp {
  int n = 0, w = 0;
  const int* nbr = nullptr;      // w*n, -1 = empty slot
  const double* coef = nullptr;  // w*n
  const double* diag = nullptr;  // n (already repaired: non-finite or tiny -> 1)
  // optional compact copy for the batched FP32 passes (see EllOff16): 16-bit offsets two to a
  // word [W/2][n], the FP32 coefficients in the re-seated slot order [W][n], the slot bases
  const uint32_t* opair = nullptr;
  const float* ocoef = nullptr;
  int obase[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

// Neighbour offsets of an ELL operator in 16 bits, for the batched passes that stream the whole
// operator every iteration (an int32 index is a third of their bytes). The listed neighbours of
// each cell are re-seated so that one slot holds offsets j - c of similar size across the cells:
// the cells with the most neighbours give each slot its typical offset (the median of their k-th
// smallest offset; on a box mesh in natural order -plane, -row, -1, +1, +row, +plane), and every
// cell places its sorted offsets into the slots, in order, at the least total distance from the
// typical ones. Slot k of cell c then stores j - c - base[k] as a signed 16-bit value; an empty
// slot stores the offset 0 (the cell itself) and carries a zero coefficient. src[k * n + c] is
// the original slot (-1: empty). The re-seating changes only the order of a row's terms. Returns
// false (the int32 lists are used) when the padded width W exceeds 8 or a slot's offsets span
// more than 65,536 values (long periodic wraps, unbanded numberings).
struct EllOff16 {
  int W = 0;
  int base[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  std::vector<int16_t> d;  // [W][n]
  std::vector<int> src;    // [W][n]
};
inline bool ell_off16(const int* nbr, int n, int w, EllOff16& o) {
  const int W = w <= 4 ? 4 : w <= 6 ? 6 : w <= 8 ? 8 : 0;
  if (W == 0 || n <= 0) return false;
  o.W = W;
  o.src.assign(size_t(W) * n, -1);
  // each cell's listed neighbours as (offset, original slot), ascending
  std::vector<std::pair<int, int>> ent(size_t(w) * n);
  std::vector<int> cnt(n, 0);
  int mmax = 0;
  for (int c = 0; c < n; ++c) {
    std::pair<int, int>* e = &ent[size_t(c) * w];
    for (int k = 0; k < w; ++k) {
      const int j = nbr[size_t(k) * n + c];
      if (j >= 0) e[cnt[c]++] = {j - c, k};
    }
    std::sort(e, e + cnt[c]);
    mmax = std::max(mmax, cnt[c]);
  }
  if (mmax == 0) return false;
  double typ[8];
  for (int k = 0; k < mmax; ++k) {
    std::vector<int> v;
    for (int c = 0; c < n; ++c)
      if (cnt[c] == mmax) v.push_back(ent[size_t(c) * w + k].first);
    std::nth_element(v.begin(), v.begin() + v.size() / 2, v.end());
    typ[k] = double(v[v.size() / 2]);
  }
  // least-distance monotone placement of m sorted offsets into the mmax slots (dynamic program)
  double cost[9][9];
  int from[9][9];
  for (int c = 0; c < n; ++c) {
    const int m = cnt[c];
    const std::pair<int, int>* e = &ent[size_t(c) * w];
    for (int i = 0; i <= m; ++i)
      for (int k = 0; k <= mmax; ++k) cost[i][k] = 1e300;
    for (int k = 0; k <= mmax; ++k) cost[0][k] = 0.0;
    for (int i = 1; i <= m; ++i)
      for (int k = i; k <= mmax; ++k) {
        const double place = cost[i - 1][k - 1] + std::fabs(double(e[i - 1].first) - typ[k - 1]);
        const double skip = cost[i][k - 1];
        cost[i][k] = place <= skip ? place : skip;
        from[i][k] = place <= skip ? 1 : 0;
      }
    for (int i = m, k = mmax; i > 0; --k)
      if (from[i][k]) {
        o.src[size_t(k - 1) * n + c] = e[i - 1].second;
        --i;
      }
  }
  auto off = [&](int k, int c) {
    const int s = o.src[size_t(k) * n + c];
    return s < 0 ? 0 : nbr[size_t(s) * n + c] - c;
  };
  for (int k = 0; k < W; ++k) {
    int lo = off(k, 0), hi = lo;
    for (int c = 1; c < n; ++c) {
      const int v = off(k, c);
      lo = std::min(lo, v);
      hi = std::max(hi, v);
    }
    if (int64_t(hi) - int64_t(lo) > 65535) return false;
    o.base[k] = lo + 32768;
  }
  o.d.resize(size_t(W) * n);
  for (int k = 0; k < W; ++k)
    for (int c = 0; c < n; ++c) o.d[size_t(k) * n + c] = int16_t(off(k, c) - o.base[k]);
  return true;
}
DM_INLINE int off16_nbr(int c, int base, uint32_t bits16) { return c + base + int(int16_t(uint16_t(bits16))); }
// The direction pass's compact operator: offsets two to a word [W/2][n] and the FP32
// coefficients in the re-seated order [W][n]; false when the offsets do not fit (ell_off16).
inline bool ell_pack32(const int* nbr, const double* coef, int n, int w, std::vector<uint32_t>& pr,
                       std::vector<float>& oc, int base[8]) {
  EllOff16 o;
  if (!ell_off16(nbr, n, w, o)) return false;
  const size_t W = size_t(o.W);
  pr.assign(W / 2 * n, 0u);
  oc.assign(W * n, 0.0f);
  for (size_t k = 0; k < W; ++k)
    for (int c = 0; c < n; ++c) {
      const size_t q = k * n + c;
      const int s = o.src[q];
      oc[q] = s < 0 ? 0.0f : float(coef[size_t(s) * n + c]);
      pr[k / 2 * n + c] |= uint32_t(uint16_t(o.d[q])) << (16 * (k & 1));
    }
  for (int k = 0; k < 8; ++k) base[k] = o.base[k];
  return true;
}
// DOPAMINE_ELL_OFF16=0 keeps the int32 lists everywhere (for comparison)
inline bool off16_enabled() {
  const char* e = std::getenv("DOPAMINE_ELL_OFF16");
  return !(e && e[0] == '0');
}

// Packed per-cell state of the single-reduction recurrence: one 32-byte (16-byte in FP32) load
// fetches everything a neighbour contributes.
struct alignas(32) St4 {
  double r, s, w, p;
};
struct alignas(16) St4f {
  float r, s, w, p;
};

// Operator in the form the single-reduction passes use: chat = coef / diag(column) and
// dinv = 1 / diag, so A D^-1 r = r_c - sum chat r_j needs no division.
struct JacobiScaled {
  int n = 0, w = 0;
  const int* nbr = nullptr;
  const double* chat = nullptr;
  const double* dinv = nullptr;
  const float* chat32 = nullptr;
  const float* dinv32 = nullptr;
};

namespace cgk {
struct Scale {
  int n;
  const int* nbr;
  const double *coef, *diag;
  double *chat, *dinv;
  float *chat32, *dinv32;  // may be null
  int w;
  DM_HD void operator()(int c) const {
    const double di = 1.0 / diag[c];
    dinv[c] = di;
    if (dinv32) dinv32[c] = float(di);
    for (int k = 0; k < w; ++k) {
      const size_t q = size_t(k) * n + c;
      const int j = nbr[q];
      const double v = j >= 0 ? coef[q] / diag[j] : 0.0;
      chat[q] = v;
      if (chat32) chat32[q] = float(v);
    }
  }
};
}  // namespace cgk

// Calls fn with the slot count of the batched (branch-free) passes for an operator of width w
// (0: the generic per-slot loop).
template <class Fn>
inline void with_width(int w, Fn&& fn) {
  if (w <= 4)
    fn(std::integral_constant<int, 4>{});
  else if (w <= 6)
    fn(std::integral_constant<int, 6>{});
  else if (w <= 8)
    fn(std::integral_constant<int, 8>{});
  else
    fn(std::integral_constant<int, 0>{});
}

namespace cgk {
struct To32 {
  int n, w;
  const double *coef, *diag;
  float *coef32, *diag32;
  DM_HD void operator()(int c) const {
    diag32[c] = float(diag[c]);
    for (int k = 0; k < w; ++k) coef32[size_t(k) * n + c] = float(coef[size_t(k) * n + c]);
  }
};
}  // namespace cgk

// FP32 copy of the operator for the mixed-precision inner solves (same rows and slots).
struct EllOp32 {
  int n = 0, w = 0;
  const int* nbr = nullptr;
  const float* coef = nullptr;
  const float* diag = nullptr;
  const uint32_t* opair = nullptr;  // as EllOp: the compact copy, or null
  const float* ocoef = nullptr;
  int obase[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

// Chained passes of one solve. Each reads the totals of the previous pass's partials (identical
// in every block), derives its scalars in setup(), and returns per-cell contributions to its own
// partials. `leader` is one thread of the grid: it publishes scalars for later passes and the
// host. A pass that finds the solve finished returns false from setup() in every block.
namespace cgk {

struct NoScal {};

// b' = finite(b), x scrubbed; partial: sum b'
struct SumB {
  const double* b;
  double *bw, *x;
  using Scal = NoScal;
  DM_HD bool setup(const Red<1>&, bool, Scal&) const { return true; }
  DM_HD Red<1> operator()(int c, const Scal&) const {
    double bi = b[c];
    if (!isfinite(bi)) bi = 0.0;
    bw[c] = bi;
    if (!isfinite(x[c])) x[c] = 0.0;
    Red<1> r;
    r.v[0] = bi;
    return r;
  }
};

// b' -= mean(b') on a singular operator; r = b' - A x; partials: b'.b', r.r
struct Init {
  EllOp A;
  double *bw, *r;
  const double* x;
  CgScalars* sc;
  const double* atol_dev;
  struct Scal {
    double bmean;
  };
  DM_HD bool setup(const Red<1>& t, bool leader, Scal& s) const {
    s.bmean = sc->nullspace ? t.v[0] / double(A.n) : 0.0;
    if (leader && atol_dev) sc->atol = *atol_dev;
    return true;
  }
  DM_HD Red<2> operator()(int c, const Scal& s) const {
    const int n = A.n;
    const double bp = bw[c] - s.bmean;
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

// ttol = max(rtol ||b'||, atol); z = D^-1 r; partials: r.r, r.z, sum z, sum r
struct Z0 {
  const double *diag, *r;
  double* z;
  CgScalars* sc;
  using Scal = NoScal;
  DM_HD bool setup(const Red<2>& t, bool leader, Scal&) const {
    if (leader) sc->ttol = fmax(sc->rtol * sqrt(t.v[0]), sc->atol);
    return true;
  }
  DM_HD Red<4> operator()(int c, const Scal&) const {
    const double rc = r[c], zc = rc / diag[c];
    z[c] = zc;
    Red<4> o;
    o.v[0] = rc * rc;
    o.v[1] = rc * zc;
    o.v[2] = zc;
    o.v[3] = rc;
    return o;
  }
};

// Iteration k, first pass. From the residual partials of the last update: the convergence test
// ||r_k|| <= ttol (KSPCG tests after each update, and once before the first), the projected
// r.z = r.z - mean(z) sum r, and beta = (r.z)_k / (r.z)_(k-1). Then p = (z - mean z) + beta p_old
// for this row and, recomputed on the fly, for its neighbours; q = A p; partial p.q.
struct Dir {
  EllOp A;
  const double *z, *pold;
  double *pnew, *q;
  CgScalars* sc;
  int k;
  struct Scal {
    double zm, beta;
  };
  DM_HD bool setup(const Red<4>& t, bool leader, Scal& s) const {
    if (sc->done) return false;
    const double rr = t.v[0];
    s.zm = sc->nullspace ? t.v[2] / double(A.n) : 0.0;
    const double rz = t.v[1] - s.zm * t.v[3];
    const int fin = (sqrt(rr) <= sc->ttol) ? 1 : (k >= sc->maxit ? 2 : 0);
    if (leader) {
      sc->its = k;
      sc->rr = rr;
      if (fin) sc->done = fin;
      else sc->rz_hist[k & 1] = rz;
    }
    if (fin) return false;
    s.beta = k == 0 ? 0.0 : rz / sc->rz_hist[(k + 1) & 1];
    return true;
  }
  DM_HD Red<1> operator()(int c, const Scal& s) const {
    const int n = A.n;
    const bool first = k == 0;
    const double pc = first ? (z[c] - s.zm) : (z[c] - s.zm) + s.beta * pold[c];
    double qc = A.diag[c] * pc;
    for (int kk = 0; kk < A.w; ++kk) {
      const int j = A.nbr[size_t(kk) * n + c];
      if (j < 0) continue;
      const double pj = first ? (z[j] - s.zm) : (z[j] - s.zm) + s.beta * pold[j];
      qc = qc - A.coef[size_t(kk) * n + c] * pj;
    }
    pnew[c] = pc;
    q[c] = qc;
    Red<1> o;
    o.v[0] = pc * qc;
    return o;
  }
};

// Iteration k, second pass: alpha = (r.z)_k / p.q; x += alpha p, r -= alpha q, z = D^-1 r;
// partials r.r, r.z, sum z, sum r for the next test.
struct Upd {
  const double *diag, *p, *q;
  double *x, *r, *z;
  CgScalars* sc;
  int k;
  struct Scal {
    double a;
  };
  DM_HD bool setup(const Red<1>& t, bool, Scal& s) const {
    if (sc->done) return false;
    s.a = sc->rz_hist[k & 1] / t.v[0];
    return true;
  }
  DM_HD Red<4> operator()(int c, const Scal& s) const {
    x[c] = x[c] + s.a * p[c];
    const double rc = r[c] - s.a * q[c];
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

// true residual ||b - A x||^2 (partials summed on the host)
struct Resid {
  EllOp A;
  const double *b, *x;
  using Scal = NoScal;
  DM_HD bool setup(const Red<1>&, bool, Scal&) const { return true; }
  DM_HD Red<1> operator()(int c, const Scal&) const {
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

// ---- single-reduction variant (Chronopoulos & Gear, J. Comput. Appl. Math. 25 (1989) 153-168):
// the same Krylov iterates as Hestenes-Stiefel CG in exact arithmetic, with s = A p carried by
// recurrence, so one pass per iteration: gamma = (r,u), delta = (w,u) with u = D^-1 r and
// w = A u come out of the same reduction. On a singular operator u = D^-1 r - mean(D^-1 r); since
// A 1 = 0, w = A D^-1 r and the mean enters only through the scalars:
// gamma = (r, D^-1 r) - mean * sum r, delta = (w, D^-1 r) - mean * sum w.

// w0 = A D^-1 r0 and ttol; packed state {r0, 0, w0, 0}; partials: r.r, r.u, w.u, sum u, sum r,
// sum w (u = D^-1 r; A D^-1 r = r_c - sum chat r_j)
struct W0 {
  JacobiScaled J;
  const double* r;
  St4* st;
  CgScalars* sc;
  using Scal = NoScal;
  DM_HD bool setup(const Red<2>& t, bool leader, Scal&) const {
    if (leader) sc->ttol = fmax(sc->rtol * sqrt(t.v[0]), sc->atol);
    return true;
  }
  DM_HD Red<6> operator()(int c, const Scal&) const {
    const int n = J.n;
    const double rc = r[c], uc = rc * J.dinv[c];
    double wc = rc;
    for (int k = 0; k < J.w; ++k) {
      const size_t q = size_t(k) * n + c;
      const int j = J.nbr[q];
      if (j < 0) continue;
      wc = wc - J.chat[q] * r[j];
    }
    st[c] = St4{rc, 0.0, wc, 0.0};
    Red<6> o;
    o.v[0] = rc * rc;
    o.v[1] = rc * uc;
    o.v[2] = wc * uc;
    o.v[3] = uc;
    o.v[4] = rc;
    o.v[5] = wc;
    return o;
  }
};

// Iteration k: test ||r_k||; beta_k = gamma_k / gamma_(k-1), alpha_k = gamma_k / (delta_k -
// beta_k gamma_k / alpha_(k-1)); p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s,
// w' = A D^-1 r' with the neighbours' r' recomputed on the fly from their old r, s, w.
// The old packed state is read from one buffer and the new one written to the other.
// The batched form (WC > 0, used on the device) reads WC neighbour slots branch-free: a slot past
// the operator's width or without a neighbour reads the cell's own state and has a zero
// coefficient (chat is zero on empty slots), which subtracts an exact zero. Each thread takes U
// cells at a time, so all their loads are issued before the first result is needed. At k = 0 the
// state's s and p are zero (W0 writes them), so beta = 0 gives the first iteration's values.
#ifndef DM_CG_CELLS_PER_THREAD
#define DM_CG_CELLS_PER_THREAD 1
#endif
constexpr int CG_U = DM_CG_CELLS_PER_THREAD;

template <int WC>
struct FusedT {
  static constexpr int kSlots = WC;
  JacobiScaled J;
  const St4* st0;
  St4* st1;
  double* x;
  CgScalars* sc;
  int k;
  struct Scal {
    double um, beta, alpha;
  };
  DM_HD bool setup(const Red<6>& t, bool leader, Scal& s) const {
    if (sc->done) return false;
    const double rr = t.v[0];
    s.um = sc->nullspace ? t.v[3] / double(J.n) : 0.0;
    const double gamma = t.v[1] - s.um * t.v[4];
    const double delta = t.v[2] - s.um * t.v[5];
    const int fin = (sqrt(rr) <= sc->ttol) ? 1 : (k >= sc->maxit ? 2 : 0);
    if (fin) {
      if (leader) {
        sc->its = k;
        sc->rr = rr;
        sc->done = fin;
      }
      return false;
    }
    if (k == 0) {
      s.beta = 0.0;
      s.alpha = gamma / delta;
    } else {
      s.beta = gamma / sc->gamma_hist[(k + 1) & 1];
      s.alpha = gamma / (delta - s.beta * gamma / sc->alpha_hist[(k + 1) & 1]);
    }
    if (leader) {
      sc->its = k;
      sc->rr = rr;
      sc->gamma_hist[k & 1] = gamma;
      sc->alpha_hist[k & 1] = s.alpha;
    }
    return true;
  }
  DM_HD static Red<6> sums(double rc, double un, double wc) {
    Red<6> out;
    out.v[0] = rc * rc;
    out.v[1] = rc * un;
    out.v[2] = wc * un;
    out.v[3] = un;
    out.v[4] = rc;
    out.v[5] = wc;
    return out;
  }
  DM_HD Red<6> operator()(int c, const Scal& s) const {
    const int n = J.n;
    const St4 o = st0[c];
    const double di = J.dinv[c];
    const double uc = o.r * di - s.um;
    const double pc = uc + s.beta * o.p;
    const double sc_ = o.w + s.beta * o.s;
    x[c] = x[c] + s.alpha * pc;
    const double rc = o.r - s.alpha * sc_;
    double wc = rc;
    for (int kk = 0; kk < J.w; ++kk) {
      const size_t q = size_t(kk) * n + c;
      const int j = J.nbr[q];
      if (j < 0) continue;
      const St4 nj = st0[j];
      wc = wc - J.chat[q] * (nj.r - s.alpha * (nj.w + s.beta * nj.s));
    }
    st1[c] = St4{rc, sc_, wc, pc};
    return sums(rc, rc * di, wc);
  }
  DM_HD Red<6> run(int lo, int hi, int t, int nt, const Scal& s) const {
    const int n = J.n, w = J.w;
    Red<6> acc = red_identity<6, false>();
    for (int c0 = lo + t; c0 < hi; c0 += CG_U * nt) {
      int cc[CG_U];
      St4 o[CG_U];
      double di[CG_U], xo[CG_U], ch[CG_U][WC], g[CG_U][WC];
      int jj[CG_U][WC];
#pragma unroll
      for (int u = 0; u < CG_U; ++u) {
        cc[u] = (c0 + u * nt < hi) ? c0 + u * nt : c0;
        o[u] = st0[cc[u]];
        di[u] = J.dinv[cc[u]];
        xo[u] = x[cc[u]];
#pragma unroll
        for (int kk = 0; kk < WC; ++kk) {
          const bool live = kk < w;
          const size_t q = size_t(live ? kk : 0) * n + cc[u];
          const int j = J.nbr[q];
          const double h = J.chat[q];
          jj[u][kk] = (live && j >= 0) ? j : cc[u];
          ch[u][kk] = live ? h : 0.0;
        }
      }
#pragma unroll
      for (int u = 0; u < CG_U; ++u)
#pragma unroll
        for (int kk = 0; kk < WC; ++kk) {
          const St4 nj = st0[jj[u][kk]];
          g[u][kk] = nj.r - s.alpha * (nj.w + s.beta * nj.s);
        }
#pragma unroll
      for (int u = 0; u < CG_U; ++u) {
        const double uc = o[u].r * di[u] - s.um;
        const double pc = uc + s.beta * o[u].p;
        const double sc_ = o[u].w + s.beta * o[u].s;
        const double rc = o[u].r - s.alpha * sc_;
        double wc = rc;
#pragma unroll
        for (int kk = 0; kk < WC; ++kk) wc = wc - ch[u][kk] * g[u][kk];
        if (c0 + u * nt < hi) {
          x[cc[u]] = xo[u] + s.alpha * pc;
          st1[cc[u]] = St4{rc, sc_, wc, pc};
          const Red<6> v = sums(rc, rc * di[u], wc);
#pragma unroll
          for (int q = 0; q < 6; ++q) acc.v[q] = acc.v[q] + v.v[q];
        }
      }
    }
    return acc;
  }
};
using Fused = FusedT<0>;

// ---- mixed precision: FP64 iterative refinement around FP32 single-reduction CG sweeps.
// The outer loop keeps x and the true residual r = b' - A x in FP64 (FP64 operator); each sweep
// solves A e = r in FP32 (FP32 operator and state, FP64 dot-product accumulation) to a relative
// reduction tau, accumulating the correction e in FP32; the next sweep's start folds it into the
// FP64 x (x += e), and the residual pass reads x + e meanwhile.

// r = b' - A (x + e) in FP64; partials (0, r.r) in the layout Init produces
struct IrResid {
  EllOp A;
  const double *bw, *x;
  const float* e;
  double* r;
  CgScalars* sc;
  using Scal = NoScal;
  DM_HD bool setup(const Red<1>&, bool leader, Scal&) const {
    if (sc->done) return false;
    if (leader) sc->its = sc->its + sc->inner_its;
    return true;
  }
  DM_HD Red<2> operator()(int c, const Scal&) const {
    const int n = A.n;
    double ax = A.diag[c] * (x[c] + double(e[c]));
    for (int k = 0; k < A.w; ++k) {
      const int j = A.nbr[size_t(k) * n + c];
      if (j < 0) continue;
      ax = ax - A.coef[size_t(k) * n + c] * (x[j] + double(e[j]));
    }
    const double rc = bw[c] - ax;
    r[c] = rc;
    Red<2> o;
    o.v[0] = 0.0;
    o.v[1] = rc * rc;
    return o;
  }
};

// x += e, e = 0 (the FP32 correction of a finished sweep)
struct IrFold {
  double* x;
  float* e;
  DM_HD void operator()(int c) const {
    x[c] = x[c] + double(e[c]);
    e[c] = 0.0f;
  }
};

// Start of a sweep: the outer test ||r|| <= ttol on the FP64 residual, then the FP32 packed state
// {r, 0, w, 0} with w = A D^-1 r, partials as W0. The first sweep also fixes ttol from ||b'||.
struct IrStart {
  JacobiScaled J;
  const double* r64;
  St4f* st;
  double* x;
  float* e;
  CgScalars* sc;
  int sweep, max_sweeps;
  double tau;
  using Scal = NoScal;
  DM_HD bool setup(const Red<2>& t, bool leader, Scal&) const {
    if (sc->done) return false;
    const double ttol = sweep == 0 ? fmax(sc->rtol * sqrt(t.v[0]), sc->atol) : sc->ttol;
    const double rr = t.v[1];
    const int fin = (sqrt(rr) <= ttol) ? 1 : ((sweep >= max_sweeps || sc->its >= sc->maxit) ? 2 : 0);
    if (leader) {
      if (sweep == 0) sc->ttol = ttol;
      sc->rr = rr;
      sc->inner_done = 0;
      sc->inner_its = 0;
      // a sweep stops at tau ||r|| or at half the outer tolerance, whichever is larger
      sc->inner_tol2 = fmax(tau * tau * rr, 0.25 * ttol * ttol);
      if (fin) sc->done = fin;
    }
    return fin == 0;
  }
  DM_HD Red<6> operator()(int c, const Scal&) const {
    const int n = J.n;
    const float rc = float(r64[c]);
    const float uc = rc * J.dinv32[c];
    float wc = rc;
    for (int k = 0; k < J.w; ++k) {
      const size_t q = size_t(k) * n + c;
      const int j = J.nbr[q];
      if (j < 0) continue;
      wc = wc - J.chat32[q] * float(r64[j]);
    }
    st[c] = St4f{rc, 0.0f, wc, 0.0f};
    x[c] = x[c] + double(e[c]);
    e[c] = 0.0f;
    Red<6> o;
    o.v[0] = double(rc) * double(rc);
    o.v[1] = double(rc) * double(uc);
    o.v[2] = double(wc) * double(uc);
    o.v[3] = double(uc);
    o.v[4] = double(rc);
    o.v[5] = double(wc);
    return o;
  }
};

// One FP32 single-reduction iteration of a sweep (as FusedT); the correction e accumulates in FP32.
template <int WC>
struct Fused32T {
  static constexpr int kSlots = WC;
  JacobiScaled J;
  const St4f* st0;
  St4f* st1;
  float* e;  // the sweep's correction
  CgScalars* sc;
  int k;
  struct Scal {
    double um, beta, alpha;
  };
  DM_HD bool setup(const Red<6>& t, bool leader, Scal& s) const {
    if (sc->done || sc->inner_done) return false;
    const double rr = t.v[0];
    s.um = sc->nullspace ? t.v[3] / double(J.n) : 0.0;
    const double gamma = t.v[1] - s.um * t.v[4];
    const double delta = t.v[2] - s.um * t.v[5];
    if (k > 0 && (rr <= sc->inner_tol2 || sc->its + k >= sc->maxit)) {
      if (leader) {
        sc->inner_done = 1;
        sc->inner_its = k;
      }
      return false;
    }
    if (k == 0) {
      s.beta = 0.0;
      s.alpha = gamma / delta;
    } else {
      s.beta = gamma / sc->gamma_hist[(k + 1) & 1];
      s.alpha = gamma / (delta - s.beta * gamma / sc->alpha_hist[(k + 1) & 1]);
    }
    if (leader) {
      sc->inner_its = k;
      sc->gamma_hist[k & 1] = gamma;
      sc->alpha_hist[k & 1] = s.alpha;
    }
    return true;
  }
  DM_HD static Red<6> sums(float rc, float un, float wc) {
    Red<6> out;
    out.v[0] = double(rc) * double(rc);
    out.v[1] = double(rc) * double(un);
    out.v[2] = double(wc) * double(un);
    out.v[3] = double(un);
    out.v[4] = double(rc);
    out.v[5] = double(wc);
    return out;
  }
  DM_HD Red<6> operator()(int c, const Scal& s) const {
    const int n = J.n;
    const float beta = float(s.beta), alpha = float(s.alpha), um = float(s.um);
    const St4f o = st0[c];
    const float di = J.dinv32[c];
    const float uc = o.r * di - um;
    const float pc = uc + beta * o.p;
    const float sc_ = o.w + beta * o.s;
    e[c] = e[c] + alpha * pc;
    const float rc = o.r - alpha * sc_;
    float wc = rc;
    for (int kk = 0; kk < J.w; ++kk) {
      const size_t q = size_t(kk) * n + c;
      const int j = J.nbr[q];
      if (j < 0) continue;
      const St4f nj = st0[j];
      wc = wc - J.chat32[q] * (nj.r - alpha * (nj.w + beta * nj.s));
    }
    st1[c] = St4f{rc, sc_, wc, pc};
    return sums(rc, rc * di, wc);
  }
  DM_HD Red<6> run(int lo, int hi, int t, int nt, const Scal& s) const {
    const int n = J.n, w = J.w;
    const float beta = float(s.beta), alpha = float(s.alpha), um = float(s.um);
    Red<6> acc = red_identity<6, false>();
    for (int c0 = lo + t; c0 < hi; c0 += CG_U * nt) {
      int cc[CG_U];
      St4f o[CG_U];
      float di[CG_U], ch[CG_U][WC], g[CG_U][WC];
      float eo[CG_U];
      int jj[CG_U][WC];
#pragma unroll
      for (int u = 0; u < CG_U; ++u) {
        cc[u] = (c0 + u * nt < hi) ? c0 + u * nt : c0;
        o[u] = st0[cc[u]];
        di[u] = J.dinv32[cc[u]];
        eo[u] = e[cc[u]];
#pragma unroll
        for (int kk = 0; kk < WC; ++kk) {
          const bool live = kk < w;
          const size_t q = size_t(live ? kk : 0) * n + cc[u];
          const int j = J.nbr[q];
          const float h = J.chat32[q];
          jj[u][kk] = (live && j >= 0) ? j : cc[u];
          ch[u][kk] = live ? h : 0.0f;
        }
      }
#pragma unroll
      for (int u = 0; u < CG_U; ++u)
#pragma unroll
        for (int kk = 0; kk < WC; ++kk) {
          const St4f nj = st0[jj[u][kk]];
          g[u][kk] = nj.r - alpha * (nj.w + beta * nj.s);
        }
#pragma unroll
      for (int u = 0; u < CG_U; ++u) {
        const float uc = o[u].r * di[u] - um;
        const float pc = uc + beta * o[u].p;
        const float sc_ = o[u].w + beta * o[u].s;
        const float rc = o[u].r - alpha * sc_;
        float wc = rc;
#pragma unroll
        for (int kk = 0; kk < WC; ++kk) wc = wc - ch[u][kk] * g[u][kk];
        if (c0 + u * nt < hi) {
          e[cc[u]] = eo[u] + alpha * pc;
          st1[cc[u]] = St4f{rc, sc_, wc, pc};
          const Red<6> v = sums(rc, rc * di[u], wc);
#pragma unroll
          for (int q = 0; q < 6; ++q) acc.v[q] = acc.v[q] + v.v[q];
        }
      }
    }
    return acc;
  }
};
using Fused32 = Fused32T<0>;

}  // namespace cgk

// ---- flexible CG around a preconditioner applied elsewhere (the learned one, device only).
// beta = max(0, z.(r - r_old) / (z.r)_old) (src/solver/mod_cell_projection.f90:273-281 uses the
// same flexible beta), alpha = z.r / p.A p; the stopping test is the unpreconditioned
// ||r|| <= max(rtol ||b'||, atol), as in the Jacobi passes.

// z = M(r) as a chain of passes: the first sums the residual partials rr_part (rr_g blocks) of the
// last update, records its = k, tests ||r_k|| <= ttol and k >= maxit (setting done and stopping
// its grid like FusedT), the last leaves z.r, z.r_old, sum z, sum r, sum r_old in out_part
// (out_grid blocks).
struct FcgPc {
  virtual ~FcgPc() = default;
  virtual void apply(int k, const double* r, const double* rold, double* z, CgScalars* sc, const double* rr_part,
                     int rr_g, double* out_part) = 0;
  virtual int out_grid(int n) const = 0;
};

// Preconditioner of the FP32 sweeps of mixed precision: apply32(k, r, v, z, zstride, ...) (z[c * zstride]) with the
// contract of FcgPc on FP32 vectors — the five partials z.r, z.v, sum z, sum r, sum v, where the
// FP32 flexible CG passes v = q = A p of the last step (it forms z.r_old as z.r + alpha z.q) —
// and the sweep's stopping test (rr <= inner_tol2 or the iteration limit, setting inner_done) in
// its first pass.
struct FcgPc32 {
  virtual ~FcgPc32() = default;
  virtual void apply32(int k, const float* r, const float* v, float* z, int zstride, CgScalars* sc,
                       const double* rr_part, int rr_g, double* out_part) = 0;
  virtual int out_grid32(int n) const = 0;
};

// Chooses, solve by solve, how a pressure system is solved (the learned preconditioner's
// collection, training and fallback policy); same contract as CgT::solve.
struct PressurePolicy {
  virtual ~PressurePolicy() = default;
  virtual int solve(const EllOp& A, const double* b, double* x, double rtol, double atol, const double* atol_dev,
                    int maxit, bool nullspace, int* its, double* resid) = 0;
  // A loose inner solve used as a preconditioner (the EXACT cell projection's): the policy's
  // current preconditioner, without its bookkeeping (no samples, no fallbacks, no regression count).
  virtual int precondition(const EllOp& A, const double* b, double* x, double rtol, int maxit, bool nullspace,
                           int* its, double* resid) = 0;
};

// Receives the residual of a running Jacobi CG solve at its batch boundaries (r[c * stride]).
struct CgCapture {
  virtual ~CgCapture() = default;
  virtual void residual(int k, const double* r, int stride) = 0;
};

namespace cgk {
// ttol from the partials (b'.b', r.r) of Init; r_old = r; partial r.r
struct FcgStart {
  const double* r;
  double* rold;
  CgScalars* sc;
  using Scal = NoScal;
  DM_HD bool setup(const Red<2>& t, bool leader, Scal&) const {
    if (leader) sc->ttol = fmax(sc->rtol * sqrt(t.v[0]), sc->atol);
    return true;
  }
  DM_HD Red<1> operator()(int c, const Scal&) const {
    const double rc = r[c];
    rold[c] = rc;
    Red<1> o;
    o.v[0] = rc * rc;
    return o;
  }
};

// z -= mean(z) on a singular operator; beta; p = z + beta p_old; q = A p (neighbours' p formed on
// the fly); partial p.q. The device form (WC > 0) reads WC operator slots branch-free, as FusedT.
template <int WC>
struct FcgDirT {
  static constexpr int kSlots = WC;
  EllOp A;
  const double *z, *pold;
  double *pnew, *q;
  CgScalars* sc;
  int k;
  struct Scal {
    double zm, beta;
  };
  DM_HD bool setup(const Red<5>& t, bool leader, Scal& s) const {
    if (sc->done) return false;
    s.zm = sc->nullspace ? t.v[2] / double(A.n) : 0.0;
    const double zr = t.v[0] - s.zm * t.v[3];
    const double zro = t.v[1] - s.zm * t.v[4];
    s.beta = k == 0 ? 0.0 : fmax(0.0, (zr - zro) / sc->rz_hist[(k + 1) & 1]);
    if (leader) sc->rz_hist[k & 1] = zr;
    return true;
  }
  DM_HD Red<1> operator()(int c, const Scal& s) const {
    const int n = A.n;
    const bool first = k == 0;
    const double pc = first ? z[c] - s.zm : (z[c] - s.zm) + s.beta * pold[c];
    double qc = A.diag[c] * pc;
    for (int kk = 0; kk < A.w; ++kk) {
      const size_t q_ = size_t(kk) * n + c;
      const int j = A.nbr[q_];
      if (j < 0) continue;
      const double pj = first ? z[j] - s.zm : (z[j] - s.zm) + s.beta * pold[j];
      qc = qc - A.coef[q_] * pj;
    }
    pnew[c] = pc;
    q[c] = qc;
    Red<1> o;
    o.v[0] = pc * qc;
    return o;
  }
  DM_HD Red<1> run(int lo, int hi, int t, int nt, const Scal& s) const {
    const int n = A.n, w = A.w;
    const bool first = k == 0;
    Red<1> acc = red_identity<1, false>();
    for (int c = lo + t; c < hi; c += nt) {
      int jj[WC > 0 ? WC : 1];
      double cf[WC > 0 ? WC : 1], pj[WC > 0 ? WC : 1];
#pragma unroll
      for (int kk = 0; kk < WC; ++kk) {
        const bool live = kk < w;
        const size_t q_ = size_t(live ? kk : 0) * n + c;
        const int j = A.nbr[q_];
        const double h = A.coef[q_];
        jj[kk] = (live && j >= 0) ? j : c;
        cf[kk] = (live && j >= 0) ? h : 0.0;
      }
#pragma unroll
      for (int kk = 0; kk < WC; ++kk)
        pj[kk] = first ? z[jj[kk]] - s.zm : (z[jj[kk]] - s.zm) + s.beta * pold[jj[kk]];
      const double pc = first ? z[c] - s.zm : (z[c] - s.zm) + s.beta * pold[c];
      double qc = A.diag[c] * pc;
#pragma unroll
      for (int kk = 0; kk < WC; ++kk) qc = qc - cf[kk] * pj[kk];
      pnew[c] = pc;
      q[c] = qc;
      acc.v[0] = acc.v[0] + pc * qc;
    }
    return acc;
  }
};
using FcgDir = FcgDirT<0>;

// alpha = (z.r) / (p.q); x += alpha p; r_old = r; r -= alpha q; partial r.r
struct FcgUpd {
  double *x, *r, *rold;
  const double *p, *q;
  CgScalars* sc;
  int k;
  struct Scal {
    double alpha;
  };
  DM_HD bool setup(const Red<1>& t, bool, Scal& s) const {
    if (sc->done) return false;
    s.alpha = sc->rz_hist[k & 1] / t.v[0];
    return true;
  }
  DM_HD Red<1> operator()(int c, const Scal& s) const {
    x[c] = x[c] + s.alpha * p[c];
    const double rc = r[c];
    rold[c] = rc;
    const double rn = rc - s.alpha * q[c];
    r[c] = rn;
    Red<1> o;
    o.v[0] = rn * rn;
    return o;
  }
};

// ---- the FP32 flexible CG of a mixed-precision sweep (vectors FP32, dot products FP64)

// Sweep start: the outer FP64 test on r64 (as IrStart), then r = float(r64), r_old = r, x += e,
// e = 0; partial r.r of the FP32 residual.
struct FcgStart32 {
  const double* r64;
  float* r;
  double* x;
  float* e;
  CgScalars* sc;
  int sweep, max_sweeps;
  double tau;
  using Scal = NoScal;
  DM_HD bool setup(const Red<2>& t, bool leader, Scal&) const {
    if (sc->done) return false;
    const double ttol = sweep == 0 ? fmax(sc->rtol * sqrt(t.v[0]), sc->atol) : sc->ttol;
    const double rr = t.v[1];
    const int fin = (sqrt(rr) <= ttol) ? 1 : ((sweep >= max_sweeps || sc->its >= sc->maxit) ? 2 : 0);
    if (leader) {
      if (sweep == 0) sc->ttol = ttol;
      sc->rr = rr;
      sc->inner_done = 0;
      sc->inner_its = 0;
      sc->inner_tol2 = fmax(tau * tau * rr, 0.25 * ttol * ttol);
      if (fin) sc->done = fin;
    }
    return fin == 0;
  }
  DM_HD Red<1> operator()(int c, const Scal&) const {
    const float rc = float(r64[c]);
    r[c] = rc;
    x[c] = x[c] + double(e[c]);
    e[c] = 0.0f;
    Red<1> o;
    o.v[0] = double(rc) * double(rc);
    return o;
  }
};

// z -= mean(z) on a singular operator; beta = max(0, z.(r - r_old) / (z.r)_old);
// p = z + beta p_old; q = A p (FP32 operator); partial p.q. z and the directions are plain FP32
// arrays (the new direction and q are whole-word stores). The partials come as z.r, z.q, sum z,
// sum r, sum q with q of the last step: r_old = r + alpha q, so z.r_old = z.r + alpha z.q and
// sum r_old = sum r + alpha sum q (no r_old stored, and no cancellation in z.(r - r_old)).
template <int WC>
struct FcgDir32T {
  static constexpr int kSlots = WC;
  EllOp32 A;
  const float *z, *pold;
  float* pnew;
  float* q;
  CgScalars* sc;
  int k;
  struct Scal {
    double zm, beta;
  };
  DM_HD bool setup(const Red<5>& t, bool leader, Scal& s) const {
    if (sc->done || sc->inner_done) return false;
    s.zm = sc->nullspace ? t.v[2] / double(A.n) : 0.0;
    const double zr = t.v[0] - s.zm * t.v[3];
    const double al = sc->alpha32;
    const double zro = (t.v[0] + al * t.v[1]) - s.zm * (t.v[3] + al * t.v[4]);
    s.beta = k == 0 ? 0.0 : fmax(0.0, (zr - zro) / sc->rz_hist[(k + 1) & 1]);
    if (leader) sc->rz_hist[k & 1] = zr;
    return true;
  }
  DM_HD static float dir(float zv, float pv, float zm, float beta, bool first) {
    return first ? zv - zm : (zv - zm) + beta * pv;
  }
  DM_HD Red<1> operator()(int c, const Scal& s) const {
    const int n = A.n;
    const float zm = float(s.zm), beta = float(s.beta);
    const bool first = k == 0;
    const float pc = dir(z[c], first ? 0.0f : pold[c], zm, beta, first);
    float qc = A.diag[c] * pc;
    for (int kk = 0; kk < A.w; ++kk) {
      const size_t q_ = size_t(kk) * n + c;
      const int j = A.nbr[q_];
      if (j < 0) continue;
      qc = qc - A.coef[q_] * dir(z[j], first ? 0.0f : pold[j], zm, beta, first);
    }
    pnew[c] = pc;
    q[c] = qc;
    Red<1> o;
    o.v[0] = double(pc) * double(qc);
    return o;
  }
  // batched form: all WC slot loads, then all neighbour gathers, branch-free (an empty slot reads
  // the cell itself with a zero coefficient)
  DM_HD Red<1> run(int lo, int hi, int t, int nt, const Scal& s) const {
    const int n = A.n, w = A.w;
    const float zm = float(s.zm), beta = float(s.beta);
    const bool first = k == 0;
    Red<1> acc = red_identity<1, false>();
    for (int c = lo + t; c < hi; c += nt) {
      int jj[WC > 0 ? WC : 1];
      float cf[WC > 0 ? WC : 1];
      float zj[WC > 0 ? WC : 1], pj[WC > 0 ? WC : 1];
      if (A.opair) {  // 16-bit offsets two to a word, every slot live (EllOff16)
#pragma unroll
        for (int m = 0; m < WC / 2; ++m) {
          const uint32_t pr = A.opair[size_t(m) * n + c];
          jj[2 * m] = off16_nbr(c, A.obase[2 * m], pr & 0xffffu);
          jj[2 * m + 1] = off16_nbr(c, A.obase[2 * m + 1], pr >> 16);
        }
#pragma unroll
        for (int kk = 0; kk < WC; ++kk) cf[kk] = A.ocoef[size_t(kk) * n + c];
      } else {
#pragma unroll
        for (int kk = 0; kk < WC; ++kk) {
          const bool live = kk < w;
          const size_t q_ = size_t(live ? kk : 0) * n + c;
          const int j = A.nbr[q_];
          const float h = A.coef[q_];
          jj[kk] = (live && j >= 0) ? j : c;
          cf[kk] = (live && j >= 0) ? h : 0.0f;
        }
      }
#pragma unroll
      for (int kk = 0; kk < WC; ++kk) {
        zj[kk] = z[jj[kk]];
        pj[kk] = first ? 0.0f : pold[jj[kk]];
      }
      const float pc = dir(z[c], first ? 0.0f : pold[c], zm, beta, first);
      float qc = A.diag[c] * pc;
#pragma unroll
      for (int kk = 0; kk < WC; ++kk) qc = qc - cf[kk] * dir(zj[kk], pj[kk], zm, beta, first);
      pnew[c] = pc;
      q[c] = qc;
      acc.v[0] = acc.v[0] + double(pc) * double(qc);
    }
    return acc;
  }
};
using FcgDir32 = FcgDir32T<0>;

// alpha = (z.r) / (p.q) (kept for the next direction's beta); e += alpha p; r -= alpha q;
// partial r.r
struct FcgUpd32 {
  float *e, *r;
  const float* p;  // the direction
  const float* q;
  CgScalars* sc;
  int k;
  struct Scal {
    double alpha;
  };
  DM_HD bool setup(const Red<1>& t, bool leader, Scal& s) const {
    if (sc->done || sc->inner_done) return false;
    s.alpha = sc->rz_hist[k & 1] / t.v[0];
    if (leader) sc->alpha32 = s.alpha;
    return true;
  }
  DM_HD Red<1> operator()(int c, const Scal& s) const {
    const float a = float(s.alpha);
    e[c] = e[c] + a * p[c];
    const float rc = r[c];
    const float rn = rc - a * q[c];
    r[c] = rn;
    Red<1> o;
    o.v[0] = double(rn) * double(rn);
    return o;
  }
};

}  // namespace cgk

// The reference's handling of a pressure solve that stops at its iteration limit
// (src/solver/pressure_backend_petsc.c:878-937): the solve is repeated once from x = 0 with four
// times the limit (the reference switches its preconditioner to block Jacobi for that retry; here
// the retry is the Jacobi-preconditioned CG); if the retry stops at its limit too, the result is
// accepted when the true residual norm is below 1e4 rtol or 1e3 atol (absolute comparisons, as
// the reference makes them), otherwise the solve fails with -(1000 + KSP_DIVERGED_ITS).
// first(its, resid) and retry(maxit, its, resid) return 0 when converged; zero_x() sets x = 0;
// atol() is the absolute tolerance the solves used. *its is the last solve's iteration count (as
// the reference reports it), *total_its the count over both solves.
constexpr int KSP_DIVERGED_ITS = -3;
template <class First, class Retry, class ZeroX, class Atol>
int solve_with_retry(int maxit, double rtol, First&& first, Retry&& retry, ZeroX&& zero_x, Atol&& atol, int* its,
                     double* resid, int* total_its, int* retried) {
  *retried = 0;
  const int rc1 = first(its, resid);
  *total_its = *its;
  if (rc1 == 0) return 0;
  *retried = 1;
  zero_x();
  const int rc2 = retry(4 * maxit, its, resid);
  *total_its += *its;
  if (rc2 == 0) return 0;
  if (*resid < rtol * 1.0e4 || *resid < atol() * 1.0e3) return 0;
  return -(1000 + KSP_DIVERGED_ITS);
}

template <class X>
class CgT {
 public:
  void init(X* x, int n) {
    x_ = x;
    n_ = n;
    g_ = x->pass_grid(n);
    bw_ = x->template alloc<double>(n);
    r_ = x->template alloc<double>(n);
    z_ = x->template alloc<double>(n);
    pa_ = x->template alloc<double>(n);
    pb_ = x->template alloc<double>(n);
    q_ = x->template alloc<double>(n);
    sc_ = x->template alloc<CgScalars>(1);
    part_ = x->template alloc<double>(size_t(PARTS) * g_);
    hpart_.resize(g_);
    hp_ = hpart_.data();
  }
  // Bytes of caller memory that init_external carves the work vectors and partials from (a
  // multiple of 256; each piece 256-aligned).
  static size_t external_bytes(int n, int g) {
    const auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    return 6 * al(sizeof(double) * size_t(n)) + al(sizeof(CgScalars)) + al(sizeof(double) * PARTS * size_t(g));
  }
  // The two-pass Jacobi CG on memory the caller owns: the work vectors and partials are carved
  // from mem (external_bytes(n, pass_grid(n)) bytes in the execution space), the host copy of the
  // residual partials goes to host_part (pass_grid(n) doubles). Nothing is allocated, here or in
  // solve(); destroy() releases nothing. Only the plain Jacobi passes run on such a workspace.
  void init_external(X* x, int n, void* mem, double* host_part) {
    x_ = x;
    n_ = n;
    g_ = x->pass_grid(n);
    external_ = true;
    char* m = static_cast<char*>(mem);
    const auto take = [&](size_t b) {
      char* p = m;
      m += (b + 255) / 256 * 256;
      return p;
    };
    bw_ = reinterpret_cast<double*>(take(sizeof(double) * size_t(n)));
    r_ = reinterpret_cast<double*>(take(sizeof(double) * size_t(n)));
    z_ = reinterpret_cast<double*>(take(sizeof(double) * size_t(n)));
    pa_ = reinterpret_cast<double*>(take(sizeof(double) * size_t(n)));
    pb_ = reinterpret_cast<double*>(take(sizeof(double) * size_t(n)));
    q_ = reinterpret_cast<double*>(take(sizeof(double) * size_t(n)));
    sc_ = reinterpret_cast<CgScalars*>(take(sizeof(CgScalars)));
    part_ = reinterpret_cast<double*>(take(sizeof(double) * PARTS * size_t(g_)));
    hp_ = host_part;
  }
  // Hestenes-Stiefel two passes per iteration (PETSc KSPCG's recurrence) or the single-reduction
  // variant, one pass per iteration.
  void set_single_reduction(bool on) { single_ = on; }
  // Flexible CG with the preconditioner pc (null: back to the Jacobi passes).
  void set_fcg(FcgPc* pc) {
    fcg_ = pc;
    if (pc && !ro_) ro_ = x_->template alloc<double>(n_);
    if (pc && fpart_g_ < pc->out_grid(n_)) {
      if (fpart_) x_->release(fpart_);
      fpart_g_ = pc->out_grid(n_);
      fpart_ = x_->template alloc<double>(size_t(5) * fpart_g_);
    }
  }
  bool fcg() const { return fcg_ != nullptr; }
  // Mixed precision with a preconditioner: each FP32 sweep is a flexible CG with pc (null: the
  // Jacobi single-reduction sweeps).
  void set_pc32(FcgPc32* pc) {
    pc32_ = pc;
    if (!pc) return;
    for (float** v : {&r32_, &q32_})
      if (!*v) *v = x_->template alloc<float>(n_);
    for (float** v : {&z32_, &p32_[0], &p32_[1]})
      if (!*v) *v = x_->template alloc<float>(n_);
    if (fpart32_g_ < pc->out_grid32(n_)) {
      if (fpart32_) x_->release(fpart32_);
      fpart32_g_ = pc->out_grid32(n_);
      fpart32_ = x_->template alloc<double>(size_t(5) * fpart32_g_);
    }
  }
  // The next solves run the plain Jacobi passes, whatever preconditioner is set (the retry of a
  // failed solve).
  void set_plain(bool on) { plain_ = on; }
  // Residual snapshots of the Jacobi passes at their batch boundaries (null: none).
  void set_capture(CgCapture* c) { capture_ = c; }
  // Mixed precision: FP64 iterative refinement around FP32 sweeps on A32 (same rows as the FP64
  // operator) that each reduce the residual by tau.
  void set_mixed(const EllOp32& A32, double tau) {
    (void)A32;  // the FP32 operator is derived from the FP64 one on the device
    tau_ = tau;
    mixed_ = true;
  }
  int last_sweeps() const { return last_sweeps_; }
  void destroy() {
    if (!x_) return;
    if (external_) {
      x_ = nullptr;
      return;
    }
    for (void* p : {(void*)bw_, (void*)r_, (void*)z_, (void*)pa_, (void*)pb_, (void*)q_, (void*)sc_, (void*)part_,
                    (void*)st_[0], (void*)st_[1], (void*)st32_[0], (void*)st32_[1], (void*)chat_, (void*)dinv_,
                    (void*)chat32_, (void*)dinv32_, (void*)ro_, (void*)fpart_, (void*)e32_, (void*)r32_,
                    (void*)z32_, (void*)p32_[0], (void*)p32_[1], (void*)q32_, (void*)fpart32_,
                    (void*)coef32_, (void*)diag32_})
      if (p) x_->release(p);
    r32_ = q32_ = coef32_ = diag32_ = nullptr;
    z32_ = p32_[0] = p32_[1] = nullptr;
    fpart32_ = nullptr;
    fpart32_g_ = 0;
    pc32_ = nullptr;
    st_[0] = st_[1] = nullptr;
    st32_[0] = st32_[1] = nullptr;
    e32_ = nullptr;
    chat_ = dinv_ = nullptr;
    chat32_ = dinv32_ = nullptr;
    ro_ = fpart_ = nullptr;
    fpart_g_ = 0;
    fcg_ = nullptr;
    x_ = nullptr;
  }
  // Solve A x = b (pointers in the execution space). atol_dev, when not null, points to an
  // absolute tolerance computed in the execution space (overrides atol). Returns 0 if converged.
  int solve(const EllOp& A, const double* b, double* x, double rtol, double atol, const double* atol_dev, int maxit,
            bool nullspace, int* its, double* resid) {
    using namespace cgk;
    X& xx = *x_;
    const int n = A.n;
    if (n > n_) throw std::runtime_error("CG operator larger than its workspace");
    const int g = xx.pass_grid(n);
    double* P0 = part_;              // 1 x g
    double* P1 = part_ + g;          // 2 x g
    double* D = part_ + 3 * size_t(g);   // 1 x g
    double* R = part_ + 4 * size_t(g);   // 1 x g
    double* U[3] = {part_ + 5 * size_t(g), part_ + 9 * size_t(g), part_ + 13 * size_t(g)};  // 4 x g each
    CgScalars h{};
    h.rtol = rtol;
    h.atol = atol;
    h.maxit = maxit;
    h.nullspace = nullspace ? 1 : 0;
    xx.upload(sc_, &h, 1);
    if (external_ && (single_ || mixed_ || fcg_ || pc32_))
      throw std::runtime_error("an external CG workspace runs the two-pass Jacobi CG only");
    JacobiScaled J;
    if (single_ || mixed_) J = scaled(A);
    xx.template pass<1, 1>(n, SumB{b, bw_, x}, nullptr, 0, P0);
    xx.template pass<1, 2>(n, Init{A, bw_, r_, x, sc_, atol_dev}, P0, g, P1);
    int k = 0;
    int batch = std::max(4, std::min(maxit + 1, predict_));
    if (capture_) batch = std::min(batch, 16);  // residual snapshots along the whole solve
    last_sweeps_ = 0;
    FcgPc* fcg = plain_ ? nullptr : fcg_;
    FcgPc32* pc32 = plain_ ? nullptr : pc32_;
    if (fcg) {
      double* RR = U[0];
      const int gt = fcg->out_grid(n);
      xx.template pass<2, 1>(n, FcgStart{r_, ro_, sc_}, P1, g, RR);
      double* pold = pa_;
      double* pnew = pb_;
      batch = std::max(4, std::min(maxit + 1, predict_fcg_));
      for (;;) {
        for (int i = 0; i < batch; ++i, ++k) {
          fcg->apply(k, r_, ro_, z_, sc_, RR, g, fpart_);  // times its hidden layer as K_CG_DIR
          xx.tbegin(K_CG_UPD);
          by_width(A.w, [&](auto wc) {
            xx.template pass<5, 1>(n, FcgDirT<decltype(wc)::value>{A, z_, pold, pnew, q_, sc_, k}, fpart_, gt, D);
          });
          xx.template pass<1, 1>(n, FcgUpd{x, r_, ro_, pnew, q_, sc_, k}, D, g, RR);
          xx.tend(K_CG_UPD);
          std::swap(pold, pnew);
        }
        xx.download(&h, sc_, 1);
        if (h.done || k > maxit + 1) break;
        batch = std::max(2, std::min(16, (h.its + 7) / 8));
      }
    } else if (mixed_ && pc32) {
      // FP64 refinement around FP32 flexible-CG sweeps preconditioned by pc32
      const EllOp32 A32 = op32(A);
      double* RR = U[0];
      const int gt = pc32->out_grid32(n);
      const double* in = P1;
      int total = 0;
      xx.zero(e32_, size_t(n));
      for (int sweep = 0;; ++sweep) {
        xx.template pass<2, 1>(n, FcgStart32{r_, r32_, x, e32_, sc_, sweep, MAX_SWEEPS, tau_}, in, g, RR);
        int kin = 0;
        int ibatch = std::max(4, std::min(maxit + 1, predict_sw_[std::min(sweep, MAX_SWEEPS)]));
        for (;;) {
          for (int i = 0; i < ibatch; ++i, ++kin) {
            const float* pold = p32_[kin & 1];
            float* pnew = p32_[(kin + 1) & 1];
            pc32->apply32(kin, r32_, q32_, z32_, 1, sc_, RR, g, fpart32_);
            xx.tbegin(K_CG_DIR);
            by_width(A32.w, [&](auto wc) {
              xx.template pass<5, 1>(n, FcgDir32T<decltype(wc)::value>{A32, z32_, pold, pnew, q32_, sc_, kin}, fpart32_,
                                     gt, D);
            });
            xx.tend(K_CG_DIR);
            xx.tbegin(K_CG_UPD);
            xx.template pass<1, 1>(n, FcgUpd32{e32_, r32_, pnew, q32_, sc_, kin}, D, g, RR);
            xx.tend(K_CG_UPD);
          }
          xx.download(&h, sc_, 1);
          if (h.done || h.inner_done || kin > maxit + 1) break;
          ibatch = std::max(2, std::min(16, (h.inner_its + 3) / 4));
        }
        for (int kc : {int(K_CG_DIR), int(K_CG_UPD), int(K_MG_DOWN), int(K_MG_UP)}) xx.tcommit(kc, h.inner_its);
        predict_sw_[std::min(sweep, MAX_SWEEPS)] = std::max(1, h.inner_its + 1);
        if (h.done) break;
        total += h.inner_its;
        last_sweeps_ = sweep + 1;
        h.its = total;
        xx.template pass<1, 2>(n, IrResid{A, bw_, x, e32_, r_, sc_}, nullptr, 0, P1);
        in = P1;
        if (sweep + 1 > MAX_SWEEPS) break;
      }
      xx.for_each(n, IrFold{x, e32_});
      h.its = total;
    } else if (mixed_) {
      double* F[2] = {U[0], U[0] + 6 * size_t(g)};
      const double* in = P1;
      int total = 0;
      xx.zero(e32_, size_t(n));
      for (int sweep = 0;; ++sweep) {
        xx.template pass<2, 6>(n, IrStart{J, r_, st32_[0], x, e32_, sc_, sweep, MAX_SWEEPS, tau_}, in, g, F[0]);
        int kin = 0;
        // sized from the same sweep of the previous solve; later batches a quarter of the
        // iterations so far (a pass after the stop costs a few microseconds, a host poll more)
        int ibatch = std::max(8, std::min(maxit + 1, predict_sw_[std::min(sweep, MAX_SWEEPS)]));
        for (;;) {
          for (int i = 0; i < ibatch; ++i, ++kin) {
            const int a = kin & 1, b2 = a ^ 1;
            xx.tbegin(K_CG_DIR);
            by_width(J.w, [&](auto wc) {
              xx.template pass<6, 6>(n, Fused32T<decltype(wc)::value>{J, st32_[a], st32_[b2], e32_, sc_, kin}, F[a],
                                     g, F[b2]);
            });
            xx.tend(K_CG_DIR);
          }
          xx.download(&h, sc_, 1);
          if (h.done || h.inner_done || kin > maxit + 1) break;
          ibatch = std::max(8, std::min(64, (h.inner_its + 3) / 4));
        }
        xx.tcommit(K_CG_DIR, h.inner_its);
        predict_sw_[std::min(sweep, MAX_SWEEPS)] = std::max(1, h.inner_its + 1);
        if (h.done) break;
        total += h.inner_its;
        last_sweeps_ = sweep + 1;
        h.its = total;
        // IrResid adds the sweep's inner iterations to the device-side count
        xx.template pass<1, 2>(n, IrResid{A, bw_, x, e32_, r_, sc_}, nullptr, 0, P1);
        in = P1;
        if (sweep + 1 > MAX_SWEEPS) break;
      }
      xx.for_each(n, IrFold{x, e32_});  // the last sweep's correction, when no sweep start followed
      h.its = total;
    } else if (single_) {
      double* F[2] = {U[0], U[0] + 6 * size_t(g)};  // 6 x g each, over the U blocks
      xx.template pass<2, 6>(n, W0{J, r_, st_[0], sc_}, P1, g, F[0]);
      for (;;) {
        for (int i = 0; i < batch; ++i, ++k) {
          const int a = k & 1, b2 = a ^ 1;
          xx.tbegin(K_CG_DIR);
          by_width(J.w, [&](auto wc) {
            xx.template pass<6, 6>(n, FusedT<decltype(wc)::value>{J, st_[a], st_[b2], x, sc_, k}, F[a], g, F[b2]);
          });
          xx.tend(K_CG_DIR);
        }
        xx.download(&h, sc_, 1);
        if (h.done || k > maxit + 1) break;
        if (capture_) capture_->residual(k, &st_[k & 1]->r, 4);
        batch = std::max(2, std::min(capture_ ? 16 : 32, (h.its + 7) / 8));
      }
    } else {
      xx.template pass<2, 4>(n, Z0{A.diag, r_, z_, sc_}, P1, g, U[2]);
      double* pold = pa_;
      double* pnew = pb_;
      for (;;) {
        for (int i = 0; i < batch; ++i, ++k) {
          xx.tbegin(K_CG_DIR);
          xx.template pass<4, 1>(n, Dir{A, z_, pold, pnew, q_, sc_, k}, U[(k + 2) % 3], g, D);
          xx.tend(K_CG_DIR);
          xx.tbegin(K_CG_UPD);
          xx.template pass<1, 4>(n, Upd{A.diag, pnew, q_, x, r_, z_, sc_, k}, D, g, U[k % 3]);
          xx.tend(K_CG_UPD);
          std::swap(pold, pnew);
        }
        xx.download(&h, sc_, 1);
        if (h.done || k > maxit + 1) break;
        if (capture_) capture_->residual(k, r_, 1);
        batch = std::max(2, std::min(capture_ ? 8 : 16, (h.its + 7) / 8));
      }
    }
    xx.template pass<1, 1>(n, Resid{A, b, x}, nullptr, 0, R);
    xx.download(hp_, R, size_t(g));
    xx.download(&h, sc_, 1);
    double rr = 0.0;
    for (int q = 0; q < g; ++q) rr = rr + hp_[q];
    if (!mixed_ || pc32)
      for (int kc : {int(K_CG_DIR), int(K_CG_UPD), int(K_MG_DOWN), int(K_MG_UP)}) xx.tcommit(kc, h.its);
    last_its_ = h.its;
    last_ttol_ = h.ttol;
    last_atol_ = h.atol;
    (fcg ? predict_fcg_ : predict_) = std::max(2, h.its + 1);
    if (its) *its = h.its;
    if (resid) *resid = std::sqrt(rr);
    return h.done == 1 ? 0 : -1;
  }
  static constexpr int MAX_SWEEPS = 40;

 private:
  // Calls fn with the slot count of the batched passes for an operator of width w (0: generic).
  template <class Fn>
  static void by_width(int w, Fn&& fn) {
    with_width(w, fn);
  }

  // chat = coef / diag(column), dinv = 1 / diag (and FP32 copies when mixed), built on the device
  // once per operator and kept while the operator's arrays stay the same.
  JacobiScaled scaled(const EllOp& A) {
    X& xx = *x_;
    const size_t m = size_t(std::max(1, A.w)) * A.n;
    if (!chat_ || scaled_of_ != A.coef || scaled_w_ != A.w || (mixed_ && !chat32_)) {
      for (void* q : {(void*)chat_, (void*)dinv_, (void*)chat32_, (void*)dinv32_})
        if (q) xx.release(q);
      chat_ = xx.template alloc<double>(m);
      dinv_ = xx.template alloc<double>(A.n);
      chat32_ = mixed_ ? xx.template alloc<float>(m) : nullptr;
      dinv32_ = mixed_ ? xx.template alloc<float>(A.n) : nullptr;
      xx.for_each(A.n, cgk::Scale{A.n, A.nbr, A.coef, A.diag, chat_, dinv_, chat32_, dinv32_, A.w});
      scaled_of_ = A.coef;
      scaled_w_ = A.w;
    }
    if (single_ && !st_[0]) {
      st_[0] = xx.template alloc<St4>(n_);
      st_[1] = xx.template alloc<St4>(n_);
    }
    if (mixed_ && !st32_[0]) {
      st32_[0] = xx.template alloc<St4f>(n_);
      st32_[1] = xx.template alloc<St4f>(n_);
      e32_ = xx.template alloc<float>(n_);
    }
    JacobiScaled J;
    J.n = A.n;
    J.w = A.w;
    J.nbr = A.nbr;
    J.chat = chat_;
    J.dinv = dinv_;
    J.chat32 = chat32_;
    J.dinv32 = dinv32_;
    return J;
  }
  // FP32 copy of the operator (coefficients and diagonal) for the preconditioned FP32 sweeps
  EllOp32 op32(const EllOp& A) {
    X& xx = *x_;
    const size_t m = size_t(std::max(1, A.w)) * A.n;
    if (!coef32_ || op32_of_ != A.coef || op32_w_ != A.w) {
      if (coef32_) xx.release(coef32_);
      if (diag32_) xx.release(diag32_);
      coef32_ = xx.template alloc<float>(m);
      diag32_ = xx.template alloc<float>(A.n);
      xx.for_each(A.n, cgk::To32{A.n, A.w, A.coef, A.diag, coef32_, diag32_});
      op32_of_ = A.coef;
      op32_w_ = A.w;
    }
    EllOp32 o;
    o.n = A.n;
    o.w = A.w;
    o.nbr = A.nbr;
    o.coef = coef32_;
    o.diag = diag32_;
    o.opair = A.opair;
    o.ocoef = A.ocoef;
    for (int k = 0; k < 8; ++k) o.obase[k] = A.obase[k];
    return o;
  }
  float *coef32_ = nullptr, *diag32_ = nullptr;
  const double* op32_of_ = nullptr;
  int op32_w_ = -1;
  St4* st_[2] = {nullptr, nullptr};
  St4f* st32_[2] = {nullptr, nullptr};
  float* e32_ = nullptr;  // mixed precision: the sweep's FP32 correction
  double *chat_ = nullptr, *dinv_ = nullptr;
  float *chat32_ = nullptr, *dinv32_ = nullptr;
  const double* scaled_of_ = nullptr;
  int scaled_w_ = -1;

 public:
  int last_its() const { return last_its_; }
  double last_ttol() const { return last_ttol_; }
  double last_atol() const { return last_atol_; }

 private:
  static constexpr int PARTS = 17;  // partial quantities per block: P0 1, P1 2, D 1, R 1, U 3 x 4
  X* x_ = nullptr;
  int n_ = 0, g_ = 1;
  double *bw_ = nullptr, *r_ = nullptr, *z_ = nullptr, *pa_ = nullptr, *pb_ = nullptr, *q_ = nullptr;
  bool single_ = false, mixed_ = false;
  FcgPc* fcg_ = nullptr;
  FcgPc32* pc32_ = nullptr;
  bool plain_ = false;
  float *r32_ = nullptr, *q32_ = nullptr;
  float* z32_ = nullptr;                   // FP32 flexible CG: preconditioned residual
  float* p32_[2] = {nullptr, nullptr};     // and its directions (previous, new)
  double* fpart32_ = nullptr;
  int fpart32_g_ = 0;
  CgCapture* capture_ = nullptr;
  double *ro_ = nullptr, *fpart_ = nullptr;
  int fpart_g_ = 0;
  double tau_ = 1e-3;
  int predict_sw_[MAX_SWEEPS + 1] = {};
  int last_sweeps_ = 0;
  double* part_ = nullptr;
  CgScalars* sc_ = nullptr;
  std::vector<double> hpart_;
  double* hp_ = nullptr;   // host copy of the residual partials (hpart_ or the caller's array)
  bool external_ = false;  // work vectors carved from caller memory (init_external)
  int last_its_ = 0, predict_ = 24;
  double last_ttol_ = 0.0, last_atol_ = 0.0;
  int predict_fcg_ = 24;
};

}  // namespace dm
