// Aggregation multigrid preconditioner (FP32) for the pressure Poisson operator.
//
// The reference preconditions its pressure CG with hypre BoomerAMG or PETSc GAMG
// (src/solver/pressure_backend_petsc.c:375-465), 21-46 CG iterations per solve on its TGV
// benchmark (docs/usage.md:193), where Jacobi CG needs hundreds to thousands. This is the
// iteration-cutting preconditioner of this build, written for the GPU:
//
// - set-up on the host, once per operator: greedy aggregation by strong connections (Vanek,
//   Mandel & Brezina, Computing 56 (1996) 179-196: phase 1 takes a cell and its strongly connected
//   neighbours when none of them is aggregated yet, phase 2 attaches the rest to the aggregate of
//   their strongest aggregated neighbour, phase 3 groups what is left), piecewise-constant
//   prolongation P, Galerkin coarse operators A_c = P^T A P (for constant P: sums of the fine
//   coefficients between and inside aggregates, so every coarse operator is again a symmetric
//   M-matrix in the same ELL form), down to at most COARSE_MAX cells, whose inverse (with a
//   rank-one shift when the operator is singular) is formed densely;
// - when the strength threshold leaves the first level poorly coarsened (under 3.5 cells per
//   aggregate: strongly anisotropic cells give line aggregates), the first aggregation is redone
//   with a small threshold (every coupling strong: compact 3-D aggregates of about 9 cells) and
//   the cycle uses the larger over-correction alpha = 2 (AmgCfg::theta0, min_ratio0, alpha_aggr);
// - one application is a V(1,1) cycle: per level a damped-Jacobi pre-smoothing step from zero
//   fused with the residual, a gather restriction over each aggregate's children, the dense solve
//   at the bottom, and the prolongation (scaled by alpha, the over-correction that unsmoothed
//   aggregation needs) fused with a damped-Jacobi post-smoothing step. Pre- and post-smoothing are
//   the same operator, so the cycle is a fixed symmetric linear map;
// - everything is FP32 (operator values, vectors); the flexible CG around it keeps its dot
//   products in FP64 and, in mixed precision, sits inside FP64 iterative refinement.
//
// Stored coefficients: cs_ij = coef_ij / diag_j (the pre-smoothing residual
// r = b - A (w D^-1 b) = (1 - w) b + w sum_j cs_ij b_j needs no neighbour diagonals) and
// cr_ij = coef_ij / diag_i (the post-smoothing step y = x + w (D^-1 b - x + sum_j cr_ij x_j)).
#pragma once

#include "cg_impl.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace dm {

// ---------------------------------------------------------------- host set-up

// One level's operator, FP64, ELL slot-major: A_ii = diag_i, A_ij = -coef for each listed j.
struct AmgOp {
  int n = 0, w = 0;
  std::vector<int> nbr;
  std::vector<double> coef, diag;
};

// Greedy aggregation; returns the aggregate count and agg[i] for every cell. j is strongly
// connected to i when coef_ij >= theta sqrt(diag_i diag_j). Deterministic (cell order).
inline int amg_aggregate(const AmgOp& A, double theta, std::vector<int>& agg) {
  const int n = A.n, w = A.w;
  agg.assign(n, -1);
  auto strong = [&](int i, int k, int* j) {
    const size_t q = size_t(k) * n + i;
    *j = A.nbr[q];
    return *j >= 0 && *j != i && A.coef[q] >= theta * std::sqrt(A.diag[i] * A.diag[*j]);
  };
  int na = 0;
  // phase 1: a cell whose strong neighbourhood is entirely free roots a new aggregate
  for (int i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    bool free_nbhd = true;
    for (int k = 0; k < w && free_nbhd; ++k) {
      int j;
      if (strong(i, k, &j) && agg[j] >= 0) free_nbhd = false;
    }
    if (!free_nbhd) continue;
    agg[i] = na;
    for (int k = 0; k < w; ++k) {
      int j;
      if (strong(i, k, &j)) agg[j] = na;
    }
    ++na;
  }
  // phase 2: join the aggregate of the strongest aggregated strong neighbour (decided on the
  // phase-1 state, applied afterwards)
  std::vector<int> join(n, -1);
  for (int i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    double best = -1.0;
    for (int k = 0; k < w; ++k) {
      int j;
      if (!strong(i, k, &j) || agg[j] < 0) continue;
      const double c = A.coef[size_t(k) * n + i];
      if (c > best) {
        best = c;
        join[i] = agg[j];
      }
    }
  }
  for (int i = 0; i < n; ++i)
    if (agg[i] < 0 && join[i] >= 0) agg[i] = join[i];
  // phase 3: what is left groups with its free neighbours
  for (int i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    agg[i] = na;
    for (int k = 0; k < w; ++k) {
      const int j = A.nbr[size_t(k) * n + i];
      if (j >= 0 && agg[j] < 0) agg[j] = na;
    }
    ++na;
  }
  return na;
}

// Children of each aggregate as a CSR list (cells in increasing order).
inline void amg_children(const std::vector<int>& agg, int na, std::vector<int>& ptr, std::vector<int>& idx) {
  ptr.assign(size_t(na) + 1, 0);
  for (int a : agg) ptr[size_t(a) + 1]++;
  for (int I = 0; I < na; ++I) ptr[size_t(I) + 1] += ptr[I];
  idx.assign(agg.size(), 0);
  std::vector<int> at(ptr.begin(), ptr.end() - 1);
  for (int i = 0; i < int(agg.size()); ++i) idx[at[agg[i]]++] = i;
}

// A_c = P^T A P for piecewise-constant P: diag_I = sum_{i in I} diag_i - sum of coef_ij with both
// ends in I (each inside pair is listed from both rows, as the double sum counts it);
// coef_IJ = sum of coef_ij with i in I, j in J.
inline AmgOp amg_galerkin(const AmgOp& F, const std::vector<int>& agg, int na, const std::vector<int>& ptr,
                          const std::vector<int>& idx) {
  AmgOp C;
  C.n = na;
  C.diag.assign(na, 0.0);
  std::vector<std::vector<std::pair<int, double>>> rows(na);
  std::vector<std::pair<int, double>> tmp;
  for (int I = 0; I < na; ++I) {
    tmp.clear();
    double d = 0.0;
    for (int q = ptr[I]; q < ptr[I + 1]; ++q) {
      const int i = idx[q];
      d = d + F.diag[i];
      for (int k = 0; k < F.w; ++k) {
        const size_t s = size_t(k) * F.n + i;
        const int j = F.nbr[s];
        if (j < 0) continue;
        const int J = agg[j];
        if (J == I)
          d = d - F.coef[s];
        else
          tmp.push_back({J, F.coef[s]});
      }
    }
    std::stable_sort(tmp.begin(), tmp.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    auto& row = rows[I];
    for (const auto& e : tmp) {
      if (!row.empty() && row.back().first == e.first)
        row.back().second = row.back().second + e.second;
      else
        row.push_back(e);
    }
    C.diag[I] = d;
  }
  C.w = 0;
  for (const auto& r : rows) C.w = std::max<int>(C.w, int(r.size()));
  const int w = std::max(1, C.w);
  C.w = w;
  C.nbr.assign(size_t(w) * na, -1);
  C.coef.assign(size_t(w) * na, 0.0);
  for (int I = 0; I < na; ++I)
    for (size_t k = 0; k < rows[I].size(); ++k) {
      C.nbr[k * na + I] = rows[I][k].first;
      C.coef[k * na + I] = rows[I][k].second;
    }
  return C;
}

// Dense inverse of the coarsest operator (column-major, FP32 out). Each connected component whose
// rows all sum to zero (no Dirichlet row: its constant is in the null space) is shifted by
// s 1 1^T over its cells, s = mean diagonal / component size; for a right-hand side with zero sum
// over the component the shifted inverse returns the pseudo-inverse solution. Cholesky in FP64.
inline void amg_dense_inverse(const AmgOp& C, std::vector<float>& inv, bool* singular) {
  const int n = C.n;
  std::vector<double> a(size_t(n) * n, 0.0), rs(n);
  double dsum = 0.0;
  for (int i = 0; i < n; ++i) {
    a[size_t(i) * n + i] = C.diag[i];
    rs[i] = C.diag[i];
    for (int k = 0; k < C.w; ++k) {
      const int j = C.nbr[size_t(k) * n + i];
      if (j < 0) continue;
      a[size_t(i) * n + j] -= C.coef[size_t(k) * n + i];
      rs[i] -= C.coef[size_t(k) * n + i];
    }
    dsum += std::fabs(C.diag[i]);
  }
  const double dmean = dsum > 0.0 ? dsum / n : 1.0;
  // connected components (breadth first over the ELL lists)
  std::vector<int> comp(n, -1), queue;
  int ncomp = 0;
  *singular = false;
  for (int s0 = 0; s0 < n; ++s0) {
    if (comp[s0] >= 0) continue;
    queue.assign(1, s0);
    comp[s0] = ncomp;
    for (size_t h = 0; h < queue.size(); ++h) {
      const int i = queue[h];
      for (int k = 0; k < C.w; ++k) {
        const int j = C.nbr[size_t(k) * n + i];
        if (j >= 0 && comp[j] < 0) {
          comp[j] = ncomp;
          queue.push_back(j);
        }
      }
    }
    bool sing = true;
    for (int i : queue) sing = sing && std::fabs(rs[i]) <= 1.0e-10 * std::fabs(C.diag[i]);
    if (sing) {
      *singular = true;
      const double sh = dmean / double(queue.size());
      for (int i : queue)
        for (int j : queue) a[size_t(i) * n + j] += sh;
    }
    ++ncomp;
  }
  inv.assign(size_t(n) * n, 0.0f);
  // Cholesky a = L L^T (lower, in place)
  for (int j = 0; j < n; ++j) {
    double d = a[size_t(j) * n + j];
    for (int k = 0; k < j; ++k) d -= a[size_t(j) * n + k] * a[size_t(j) * n + k];
    if (!(d > 0.0)) throw std::runtime_error("multigrid: coarsest operator is not positive definite");
    const double l = std::sqrt(d);
    a[size_t(j) * n + j] = l;
    for (int i = j + 1; i < n; ++i) {
      double v = a[size_t(i) * n + j];
      for (int k = 0; k < j; ++k) v -= a[size_t(i) * n + k] * a[size_t(j) * n + k];
      a[size_t(i) * n + j] = v / l;
    }
  }
  // columns of the inverse: solve L L^T x = e_c
  std::vector<double> y(n);
  for (int c = 0; c < n; ++c) {
    for (int i = 0; i < n; ++i) {
      double v = (i == c) ? 1.0 : 0.0;
      for (int k = 0; k < i; ++k) v -= a[size_t(i) * n + k] * y[k];
      y[i] = v / a[size_t(i) * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
      double v = y[i];
      for (int k = i + 1; k < n; ++k) v -= a[size_t(k) * n + i] * y[k];
      y[i] = v / a[size_t(i) * n + i];
    }
    for (int i = 0; i < n; ++i) inv[size_t(c) * n + i] = float(y[i]);  // column c
  }
}

struct AmgCfg {
  double theta = 0.08;       // strength threshold
  double omega = 0.8;        // Jacobi damping
  double alpha = 1.8;        // coarse-correction scaling
  // Aggressive first coarsening: when theta leaves the first level with fewer than min_ratio0
  // fine cells per aggregate (strongly anisotropic cells whose weak couplings all fall below
  // theta, so the aggregates are thin lines), the first level is aggregated again with theta0
  // and the whole cycle uses alpha_aggr, the larger over-correction that bigger unsmoothed
  // aggregates need. theta0 < 0 turns this off.
  double theta0 = 0.02;
  double min_ratio0 = 3.5;
  double alpha_aggr = 2.4;
  int coarse_max = 1500;     // cells of the densely solved level
  int max_levels = 16;
};

// ---------------------------------------------------------------- device kernels

// One level in the execution space. Level 0 keeps the operator's ELL lists (one thread per cell,
// the solver's own neighbour array).
// Coarse levels store their rows, and every level the children lists of the next coarser level's
// cells, as sliced ELL: slices of AMG_SL rows (one wavefront), each as wide as its longest row,
// entry k of row r at sptr[r / AMG_SL] + k AMG_SL + r % AMG_SL. One lane per row then reads a
// contiguous 256-byte line per slot with no reduction across lanes. Row padding points at the row
// itself with zero coefficients; children padding is -1.
constexpr int AMG_SL = 64;
constexpr int AMG_UN = 4;  // slice widths are multiples of this: each lane keeps 4 independent loads in flight

// Slot of every CSR entry in the sliced-ELL layout; sptr gets the slice offsets. Returns the size.
inline size_t amg_sell_layout(int n, const std::vector<int>& rptr, std::vector<int>& sptr, std::vector<int>& pos) {
  const int nsl = (n + AMG_SL - 1) / AMG_SL;
  sptr.assign(size_t(nsl) + 1, 0);
  for (int sl = 0; sl < nsl; ++sl) {
    int w = 0;
    for (int r = sl * AMG_SL; r < std::min(n, (sl + 1) * AMG_SL); ++r) w = std::max(w, rptr[r + 1] - rptr[r]);
    w = (w + AMG_UN - 1) / AMG_UN * AMG_UN;
    sptr[size_t(sl) + 1] = sptr[sl] + w * AMG_SL;
  }
  pos.assign(size_t(rptr[n]), 0);
  for (int r = 0; r < n; ++r)
    for (int q = rptr[r]; q < rptr[r + 1]; ++q) pos[q] = sptr[r / AMG_SL] + (q - rptr[r]) * AMG_SL + r % AMG_SL;
  return size_t(sptr[nsl]);
}

// The smoothers' scaled couplings cs = coef / diag_j, cr = coef / diag_i lie in [0, 1] and are
// stored as IEEE binary16 (AmgCoef): half the bytes of the smoothing passes' coefficient stream.
// Their 2^-11 relative rounding only perturbs the preconditioner, which the flexible CG and the
// FP64 refinement around it absorb.
using AmgCoef = uint16_t;
DM_INLINE float coef_f(AmgCoef h) {
#if defined(__HIP_DEVICE_COMPILE__)
  return float(__builtin_bit_cast(_Float16, h));
#else
  const uint32_t sign = uint32_t(h >> 15) << 31, e = (h >> 10) & 0x1f, m = h & 0x3ffu;
  uint32_t bits;
  if (e == 0) {
    if (m == 0) {
      bits = sign;
    } else {  // subnormal: m 2^-24
      const float v = float(m) * 5.9604644775390625e-8f;
      std::memcpy(&bits, &v, 4);
      bits |= sign;
    }
  } else if (e == 31) {
    bits = sign | 0x7f800000u | (m << 13);
  } else {
    bits = sign | ((e + 112u) << 23) | (m << 13);
  }
  float f;
  std::memcpy(&f, &bits, 4);
  return f;
#endif
}
// round to nearest even (set-up, on the host)
inline AmgCoef coef_h(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u) return AmgCoef(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
  if (ax >= 0x477ff000u) return AmgCoef(sign | 0x7c00u);  // rounds beyond the largest half: inf
  if (ax < 0x38800000u) {  // below the smallest normal half: subnormal or zero
    if (ax < 0x33000000u) return AmgCoef(sign);
    const uint32_t e = ax >> 23, m = (ax & 0x7fffffu) | 0x800000u;
    const uint32_t shift = 126 - e;  // value = m 2^(e - 150); half subnormal unit 2^-24
    const uint32_t q = m >> (shift - 1 + 1), rem = m & ((1u << shift) - 1), half = 1u << (shift - 1);
    uint32_t r = q;
    if (rem > half || (rem == half && (q & 1u))) ++r;
    return AmgCoef(sign | r);
  }
  const uint32_t e = (ax >> 23) - 112, m = ax & 0x7fffffu;
  uint32_t r = (e << 10) | (m >> 13);
  const uint32_t rem = m & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (r & 1u))) ++r;
  return AmgCoef(sign | r);
}

struct AmgLevelDev {
  int n = 0, w = 0;                              // cells; ELL width (level 0 only)
  const int* nbr = nullptr;                      // level 0: ELL neighbours [slot][cell]
  const int *sptr = nullptr, *col = nullptr;     // coarse levels: sliced-ELL rows
  const AmgCoef *cs = nullptr, *cr = nullptr;
  // level 0, when its neighbour offsets fit 16 bits (EllOff16): each slot's coupling and offset
  // in one word (binary16 coupling in the low half, offset - obase[slot] in the high half)
  const uint32_t *hs = nullptr, *hr = nullptr;
  int obase[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const float* dinv = nullptr;
  const int *kptr = nullptr, *kidx = nullptr;    // children (cells of this level) of the next coarser level's cells, sliced ELL
  float *b = nullptr, *x = nullptr, *r = nullptr;
  // level 0 with three or more levels: its cells' aggregates and level 1's post-smoothed values,
  // prolongated by a streaming pass (Prolong0)
  const int* agg = nullptr;
  const float* yc = nullptr;
  float* y = nullptr;  // level 1: the post-smoothed values the finest level reads
};

namespace amgk {

// How the first and last kernels of a cycle take part in the flexible CG around it:
// OUTER: the FP64 flexible CG's convergence test ||r_k|| <= ttol, k >= maxit (sets done);
// INNER: an FP32 sweep of mixed precision, stopping at rr <= inner_tol2 or the iteration limit
// (sets inner_done).
enum Mode { OUTER = 0, INNER = 1 };

DM_INLINE bool gated(const CgScalars* sc, int mode) { return sc && (sc->done || (mode == INNER && sc->inner_done)); }

DM_INLINE bool cycle_test(CgScalars* sc, int mode, int k, double rr, bool leader) {
  if (gated(sc, mode)) return false;
  if (mode == OUTER) {
    const int fin = (sqrt(rr) <= sc->ttol) ? 1 : (k >= sc->maxit ? 2 : 0);
    if (leader) {
      sc->its = k;
      sc->rr = rr;
      if (fin) sc->done = fin;
    }
    return fin == 0;
  }
  if (k > 0 && (rr <= sc->inner_tol2 || sc->its + k >= sc->maxit)) {
    if (leader) {
      sc->inner_done = 1;
      sc->inner_its = k;
    }
    return false;
  }
  if (leader) sc->inner_its = k;
  return true;
}

// Sum over the G lanes of an aligned lane group (every lane gets the total). On the host the
// group's first lane does the whole row, so the value passes through.
template <int G>
DM_INLINE float group_sum(float v) {
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v = v + __shfl_xor(v, off, 64);
#endif
  return v;
}
// Lane of the group and its stride through a row: all G lanes on the device, the first lane alone
// on the host (the others return before touching memory).
template <int G>
DM_INLINE bool group_lane(int t, int* lane, int* stride) {
#if defined(__HIP_DEVICE_COMPILE__)
  *lane = t % G;
  *stride = G;
  return true;
#else
  *lane = 0;
  *stride = 1;
  return t % G == 0;
#endif
}

// pre-smoothing from zero and its residual, level 0 (ELL): x = w D^-1 b,
// r = (1 - w) b + w sum_j cs_ij b_j
template <class T>
DM_INLINE void down_cell(const AmgLevelDev& L, const T* b, float omega, int c) {
  const int n = L.n;
  const float bc = float(b[c]);
  float s = 0.0f;
  for (int k = 0; k < L.w; ++k) {
    const size_t q = size_t(k) * n + c;
    const int j = L.nbr[q];
    if (j < 0) continue;
    s = s + coef_f(L.cs[q]) * float(b[j]);
  }
  L.x[c] = omega * L.dinv[c] * bc;
  L.r[c] = (1.0f - omega) * bc + omega * s;
}

// post-smoothing, level 0 (ELL), x already holding the prolongated coarse correction:
// y = x + w (D^-1 b - x + sum_j cr_ij x_j)
template <class T>
DM_INLINE float up_cell(const AmgLevelDev& L, const T* b, float omega, float al, int c) {
  (void)al;
  const int n = L.n;
  const float xc = L.x[c];
  float s = 0.0f;
  for (int k = 0; k < L.w; ++k) {
    const size_t q = size_t(k) * n + c;
    const int j = L.nbr[q];
    if (j < 0) continue;
    s = s + coef_f(L.cr[q]) * L.x[j];
  }
  return xc + omega * (L.dinv[c] * float(b[c]) - xc + s);
}

// the prolongation of level 1's post-smoothed values into level 0 as a streaming pass over the
// fine cells: x += alpha y_c[agg] (coalesced reads of x and agg, y_c resident in L2), instead of
// each coarse cell scattering into its children
struct Prolong0 {
  AmgLevelDev L;
  const CgScalars* sc;
  int mode;
  float al;
  DM_HD void operator()(int c) const {
    if (gated(sc, mode)) return;
    L.x[c] = L.x[c] + al * L.yc[L.agg[c]];
  }
};

// Level-0 slot loads of the batched (branch-free) kernels: WC slots of cell c, an empty slot
// pointing at the cell itself with a zero coefficient.
template <int WC>
DM_INLINE void slots0(const AmgLevelDev& L, const AmgCoef* cf, int c, int* jj, float* h) {
#pragma unroll
  for (int kk = 0; kk < WC; ++kk) {
    const bool live = kk < L.w;
    const size_t q = size_t(live ? kk : 0) * L.n + c;
    const int j = L.nbr[q];
    const float v = coef_f(cf[q]);
    jj[kk] = (live && j >= 0) ? j : c;
    h[kk] = (live && j >= 0) ? v : 0.0f;
  }
}

// the same from the packed words: every slot live, an empty one the cell itself with coupling 0
template <int WC>
DM_INLINE void slots0p(const AmgLevelDev& L, const uint32_t* hw, int c, int* jj, float* h) {
#pragma unroll
  for (int kk = 0; kk < WC; ++kk) {
    const uint32_t v = hw[size_t(kk) * L.n + c];
    jj[kk] = off16_nbr(c, L.obase[kk], v >> 16);
    h[kk] = coef_f(AmgCoef(v & 0xffffu));
  }
}

// level 0 pre-smoothing as a chained pass: its setup is the convergence test of iteration k
template <class T, int WC = 0>
struct Down0 {
  static constexpr int kSlots = WC;
  AmgLevelDev L;
  const T* b;
  CgScalars* sc;
  int mode, k;
  float omega;
  using Scal = cgk::NoScal;
  DM_HD bool setup(const Red<1>& t, bool leader, Scal&) const { return cycle_test(sc, mode, k, t.v[0], leader); }
  DM_HD Red<1> operator()(int c, const Scal&) const {
    down_cell(L, b, omega, c);
    Red<1> o;
    o.v[0] = 0.0;
    return o;
  }
  DM_HD Red<1> run(int lo, int hi, int t, int nt, const Scal&) const {
    for (int c = lo + t; c < hi; c += nt) {
      int jj[WC > 0 ? WC : 1];
      float h[WC > 0 ? WC : 1], bj[WC > 0 ? WC : 1];
      if (L.hs)
        slots0p<WC>(L, L.hs, c, jj, h);
      else
        slots0<WC>(L, L.cs, c, jj, h);
#pragma unroll
      for (int kk = 0; kk < WC; ++kk) bj[kk] = float(b[jj[kk]]);
      const float bc = float(b[c]);
      float s = 0.0f;
#pragma unroll
      for (int kk = 0; kk < WC; ++kk) s = s + h[kk] * bj[kk];
      L.x[c] = omega * L.dinv[c] * bc;
      L.r[c] = (1.0f - omega) * bc + omega * s;
    }
    return red_identity<1, false>();
  }
};

// the same level-0 steps outside a solve (one V-cycle on its own)
template <class T>
struct Down0F {
  AmgLevelDev L;
  const T* b;
  float omega;
  DM_HD void operator()(int c) const { down_cell(L, b, omega, c); }
};
template <class T>
struct Up0F {
  AmgLevelDev L;
  const T* r;
  T* z;
  float omega, al;
  DM_HD void operator()(int c) const { z[c] = T(up_cell(L, r, omega, al, c)); }
};

// Row slot helpers of the sliced-ELL layout: the first slot of row r and the slice's width.
DM_INLINE int sell_base(const int* sptr, int r) { return sptr[r / AMG_SL] + r % AMG_SL; }
DM_INLINE int sell_width(const int* sptr, int r) { return (sptr[r / AMG_SL + 1] - sptr[r / AMG_SL]) / AMG_SL; }
// sum_k a_rk v[col_rk] over this lane's chunks of AMG_UN slots of row r (chunks lane, lane + st,
// ...), AMG_UN independent loads at a time
DM_INLINE float sell_row(const int* sptr, const int* col, const AmgCoef* a, const float* v, int r, int lane, int st) {
  const int q0 = sell_base(sptr, r), nch = sell_width(sptr, r) / AMG_UN;
  float s = 0.0f;
  for (int ch = lane; ch < nch; ch += st) {
    const int k = ch * AMG_UN;
    int j[AMG_UN];
    float h[AMG_UN], x[AMG_UN];
#pragma unroll
    for (int u = 0; u < AMG_UN; ++u) {
      j[u] = col[q0 + (k + u) * AMG_SL];
      h[u] = coef_f(a[q0 + (k + u) * AMG_SL]);
    }
#pragma unroll
    for (int u = 0; u < AMG_UN; ++u) x[u] = v[j[u]];
#pragma unroll
    for (int u = 0; u < AMG_UN; ++u) s = s + h[u] * x[u];
  }
  return s;
}

// Lanes per row: one on the large levels (every wavefront reads whole 256-byte slot lines), four
// on levels of fewer rows, where the serial slot loop would leave the level latency-bound.
constexpr int AMG_SMALL_ROWS = 1 << 18;

// b_coarse[I] = sum of the fine residual over aggregate I's cells (G lanes per aggregate)
template <int G>
struct Restrict {
  const int *kptr, *kidx;
  const float* rf;
  float* bc;
  const CgScalars* sc;
  int mode;
  DM_HD void operator()(int t) const {
    int lane, st;
    if (!group_lane<G>(t, &lane, &st)) return;
    if (gated(sc, mode)) return;
    const int I = t / G;
    const int q0 = sell_base(kptr, I), nch = sell_width(kptr, I) / AMG_UN;
    float s = 0.0f;
    for (int ch = lane; ch < nch; ch += st) {
      const int k = ch * AMG_UN;
      int i[AMG_UN];
      float v[AMG_UN];
#pragma unroll
      for (int u = 0; u < AMG_UN; ++u) i[u] = kidx[q0 + (k + u) * AMG_SL];
#pragma unroll
      for (int u = 0; u < AMG_UN; ++u) v[u] = i[u] >= 0 ? rf[i[u]] : 0.0f;
#pragma unroll
      for (int u = 0; u < AMG_UN; ++u) s = s + v[u];
    }
    s = group_sum<G>(s);
    if (lane == 0) bc[I] = s;
  }
};

// coarse-level pre-smoothing and residual (G lanes per row)
template <int G>
struct DownC {
  AmgLevelDev L;
  const CgScalars* sc;
  int mode;
  float omega;
  DM_HD void operator()(int t) const {
    int lane, st;
    if (!group_lane<G>(t, &lane, &st)) return;
    if (gated(sc, mode)) return;
    const int c = t / G;
    const float s = group_sum<G>(sell_row(L.sptr, L.col, L.cs, L.b, c, lane, st));
    if (lane == 0) {
      const float bc = L.b[c];
      L.x[c] = omega * L.dinv[c] * bc;
      L.r[c] = (1.0f - omega) * bc + omega * s;
    }
  }
};

// Coarse-level post-smoothing (G lanes per row), then the prolongation with over-correction into
// the next finer level: x_fine[i] += alpha y for the cell's children i (each fine cell has one
// parent, so the writes never collide).
template <int G>
struct UpC {
  AmgLevelDev L;
  const int *kptr, *kidx;  // this level's cells' children on the finer level
  float* xf;
  const CgScalars* sc;
  int mode;
  float omega, alpha;
  float* yout = nullptr;  // level 1 under a level 0 that prolongates itself: store y, no scatter
  DM_HD void operator()(int t) const {
    int lane, st;
    if (!group_lane<G>(t, &lane, &st)) return;
    if (gated(sc, mode)) return;
    const int c = t / G;
    const float s = group_sum<G>(sell_row(L.sptr, L.col, L.cr, L.x, c, lane, st));
    const float xc = L.x[c];
    const float y = xc + omega * (L.dinv[c] * L.b[c] - xc + s);
    if (yout) {
      if (lane == 0) yout[c] = y;
      return;
    }
    const float a = alpha * y;
    const int p0 = sell_base(kptr, c), nch = sell_width(kptr, c) / AMG_UN;
    for (int ch = lane; ch < nch; ch += st) {
      const int k = ch * AMG_UN;
      int i[AMG_UN];
      float v[AMG_UN];
#pragma unroll
      for (int u = 0; u < AMG_UN; ++u) i[u] = kidx[p0 + (k + u) * AMG_SL];
#pragma unroll
      for (int u = 0; u < AMG_UN; ++u) v[u] = i[u] >= 0 ? xf[i[u]] : 0.0f;
#pragma unroll
      for (int u = 0; u < AMG_UN; ++u)
        if (i[u] >= 0) xf[i[u]] = v[u] + a;
    }
  }
};

// Coarsest level: y = inv b (dense, symmetric, a wavefront per row), then the prolongation into
// the next finer level as in UpC.
constexpr int AMG_GC = 64;
struct Coarsest {
  int n;
  const float* inv;
  const float* b;
  const int *kptr, *kidx;
  float* xf;
  const CgScalars* sc;
  int mode;
  float alpha;
  DM_HD void operator()(int t) const {
    int lane, st;
    if (!group_lane<AMG_GC>(t, &lane, &st)) return;
    if (gated(sc, mode)) return;
    const int i = t / AMG_GC;
    float s = 0.0f;
    for (int j = lane; j < n; j += st) s = s + inv[size_t(i) * n + j] * b[j];
    s = group_sum<AMG_GC>(s);
    const float a = alpha * s;
    const int p0 = sell_base(kptr, i), pw = sell_width(kptr, i);
    for (int k = lane; k < pw; k += st) {
      const int c = kidx[p0 + k * AMG_SL];
      if (c >= 0) xf[c] = xf[c] + a;
    }
  }
};

// level 0 post-smoothing as a chained pass writing z and the flexible-CG partials
// z.r, z.v, sum z, sum r, sum v (v = r_old in the FP64 flexible CG, q of the last step in the FP32
// sweeps, which form z.r_old from it)
template <class T, int WC = 0>
struct Up0 {
  static constexpr int kSlots = WC;
  AmgLevelDev L;
  const T *r, *rold;  // rold: the v of the partials
  T* z;  // z[c * zs]
  int zs;
  const CgScalars* sc;
  int mode;
  float omega, al;
  using Scal = cgk::NoScal;
  DM_HD bool setup(const Red<1>&, bool, Scal&) const { return !gated(sc, mode); }
  DM_HD Red<5> operator()(int c, const Scal&) const {
    const float y = up_cell(L, r, omega, al, c);
    const T zc = T(y);
    z[size_t(c) * zs] = zc;
    const double rc = double(r[c]), ro = double(rold[c]), zd = double(zc);
    Red<5> o;
    o.v[0] = zd * rc;
    o.v[1] = zd * ro;
    o.v[2] = zd;
    o.v[3] = rc;
    o.v[4] = ro;
    return o;
  }
  DM_HD Red<5> run(int lo, int hi, int t, int nt, const Scal&) const {
    Red<5> acc = red_identity<5, false>();
    for (int c = lo + t; c < hi; c += nt) {
      int jj[WC > 0 ? WC : 1];
      float h[WC > 0 ? WC : 1], xj[WC > 0 ? WC : 1];
      if (L.hr)
        slots0p<WC>(L, L.hr, c, jj, h);
      else
        slots0<WC>(L, L.cr, c, jj, h);
#pragma unroll
      for (int kk = 0; kk < WC; ++kk) xj[kk] = L.x[jj[kk]];
      const float xc = L.x[c];
      const float bc = float(r[c]);
      float s = 0.0f;
#pragma unroll
      for (int kk = 0; kk < WC; ++kk) s = s + h[kk] * xj[kk];
      const T zc = T(xc + omega * (L.dinv[c] * bc - xc + s));
      z[size_t(c) * zs] = zc;
      const double rc = double(r[c]), ro = double(rold[c]), zd = double(zc);
      acc.v[0] = acc.v[0] + zd * rc;
      acc.v[1] = acc.v[1] + zd * ro;
      acc.v[2] = acc.v[2] + zd;
      acc.v[3] = acc.v[3] + rc;
      acc.v[4] = acc.v[4] + ro;
    }
    return acc;
  }
};

}  // namespace amgk

// The multigrid hierarchy in the execution space and its V-cycle, usable as the FP64 flexible
// CG's preconditioner (FcgPc) and as the FP32 sweeps' (FcgPc32).
template <class X>
class AmgPc : public FcgPc, public FcgPc32 {
 public:
  // A: the operator on the host (FP64); nbr_dev: its neighbour list already in the execution
  // space (shared by level 0).
  void build(X* x, const AmgOp& A, const int* nbr_dev, const AmgCfg& cfg) {
    x_ = x;
    cfg_ = cfg;
    alpha_ = cfg.alpha;
    aggressive_ = false;
    const int* agg0 = nullptr;
    AmgOp cur = A;
    for (int lev = 0;; ++lev) {
      AmgLevelDev L;
      L.n = cur.n;
      // cs_ij = coef_ij / diag_j, cr_ij = coef_ij / diag_i, dinv_i = 1 / diag_i; a cell whose
      // diagonal vanishes (an aggregate that is a whole connected component of a singular
      // operator) is left out of the smoothing (dinv = 0, no couplings)
      double dmax = 0.0;
      for (double d : cur.diag) dmax = std::max(dmax, std::fabs(d));
      const double dmin = 1.0e-12 * dmax;
      std::vector<float> di(cur.n, 0.0f);
      for (int c = 0; c < cur.n; ++c)
        if (cur.diag[c] > dmin) di[c] = float(1.0 / cur.diag[c]);
      auto scaled = [&](int c, size_t q, float* cs, float* cr) {
        const int j = cur.nbr[q];
        if (j < 0 || !(cur.diag[c] > dmin) || !(cur.diag[j] > dmin)) return false;
        *cs = float(cur.coef[q] / cur.diag[j]);
        *cr = float(cur.coef[q] / cur.diag[c]);
        return true;
      };
      if (lev == 0) {  // ELL, the solver's neighbour array
        L.w = cur.w;
        L.nbr = nbr_dev;
        const size_t m = size_t(std::max(1, cur.w)) * cur.n;
        std::vector<float> cs(m, 0.0f), cr(m, 0.0f);
        for (int c = 0; c < cur.n; ++c)
          for (int k = 0; k < cur.w; ++k) {
            const size_t q = size_t(k) * cur.n + c;
            scaled(c, q, &cs[q], &cr[q]);
          }
        const std::vector<AmgCoef> hcs = to_coef(cs), hcr = to_coef(cr);
        L.cs = up(hcs);
        L.cr = up(hcr);
        EllOff16 o16;
        if (off16_enabled() && ell_off16(cur.nbr.data(), cur.n, cur.w, o16)) {
          std::vector<uint32_t> ws(size_t(o16.W) * cur.n), wr(ws.size());
          for (size_t q = 0; q < ws.size(); ++q) {
            const int s = o16.src[q];
            const size_t q0 = size_t(s) * cur.n + q % size_t(cur.n);
            const uint32_t hi = uint32_t(uint16_t(o16.d[q])) << 16;
            ws[q] = hi | (s < 0 ? 0u : uint32_t(hcs[q0]));
            wr[q] = hi | (s < 0 ? 0u : uint32_t(hcr[q0]));
          }
          L.hs = up(ws);
          L.hr = up(wr);
          for (int k = 0; k < 8; ++k) L.obase[k] = o16.base[k];
        }
      } else {  // the non-empty entries, sliced ELL
        std::vector<int> rptr(size_t(cur.n) + 1, 0), col;
        std::vector<float> cs, cr;
        for (int c = 0; c < cur.n; ++c) {
          for (int k = 0; k < cur.w; ++k) {
            float a, b;
            if (!scaled(c, size_t(k) * cur.n + c, &a, &b)) continue;
            col.push_back(cur.nbr[size_t(k) * cur.n + c]);
            cs.push_back(a);
            cr.push_back(b);
          }
          rptr[size_t(c) + 1] = int(col.size());
        }
        std::vector<int> sptr, pos;
        const size_t m = amg_sell_layout(cur.n, rptr, sptr, pos);
        std::vector<int> scol(std::max<size_t>(1, m), 0);
        std::vector<float> scs(std::max<size_t>(1, m), 0.0f), scr(std::max<size_t>(1, m), 0.0f);
        for (int c = 0; c < cur.n; ++c) {  // padding: the row itself, zero coefficients
          const int sl = c / AMG_SL, w = (sptr[sl + 1] - sptr[sl]) / AMG_SL;
          for (int k = 0; k < w; ++k) scol[size_t(sptr[sl]) + size_t(k) * AMG_SL + c % AMG_SL] = c;
        }
        for (size_t q = 0; q < col.size(); ++q) {
          scol[pos[q]] = col[q];
          scs[pos[q]] = cs[q];
          scr[pos[q]] = cr[q];
        }
        L.sptr = up(sptr);
        L.col = up(scol);
        L.cs = up(to_coef(scs));
        L.cr = up(to_coef(scr));
        nnz_.push_back(int64_t(col.size()));
        stored_.push_back(int64_t(m));
      }
      L.dinv = up(di);
      L.b = lev == 0 ? nullptr : fresh<float>(cur.n);
      L.x = fresh<float>(cur.n);
      L.r = fresh<float>(cur.n);
      sizes_.push_back(cur.n);
      const bool last = (cur.n <= cfg.coarse_max && lev > 0) || lev + 1 >= cfg.max_levels;
      if (last) {
        if (cur.n > 4096) throw std::runtime_error("multigrid: coarsest level too large for the dense solve");
        std::vector<float> inv;
        bool sing = false;
        amg_dense_inverse(cur, inv, &sing);
        inv_ = up(inv);
        n_coarse_ = cur.n;
        singular_ = sing;
        levels_.push_back(L);
        break;
      }
      std::vector<int> agg;
      int na = amg_aggregate(cur, cfg.theta, agg);
      if (lev == 0 && cfg.theta0 >= 0.0 && double(cur.n) < cfg.min_ratio0 * double(na)) {
        na = amg_aggregate(cur, cfg.theta0, agg);
        alpha_ = cfg.alpha_aggr;
        aggressive_ = true;
      }
      if (na * 3 > cur.n * 2 && cur.n > 16)  // stalled coarsening: every connection counts as strong
        na = amg_aggregate(cur, 0.0, agg);
      if (lev == 0) agg0 = up(agg);
      std::vector<int> ptr, idx;
      amg_children(agg, na, ptr, idx);
      AmgOp next = amg_galerkin(cur, agg, na, ptr, idx);
      {
        std::vector<int> kptr, pos;
        const size_t m = amg_sell_layout(na, ptr, kptr, pos);
        std::vector<int> kidx(std::max<size_t>(1, m), -1);
        for (size_t q = 0; q < idx.size(); ++q) kidx[pos[q]] = idx[q];
        L.kptr = up(kptr);
        L.kidx = up(kidx);
      }
      levels_.push_back(L);
      cur = std::move(next);
    }
    if (levels_.size() >= 3) {  // the finest post-smoothing adds level 1's correction itself
      levels_[1].y = fresh<float>(size_t(levels_[1].n));
      levels_[0].agg = agg0;
      levels_[0].yc = levels_[1].y;
    }
    zpart_ = fresh<double>(size_t(x_->pass_grid(A.n)));
  }
  void destroy() {
    if (!x_) return;
    for (void* p : owned_) x_->release(p);
    owned_.clear();
    levels_.clear();
    x_ = nullptr;
  }
  ~AmgPc() override { destroy(); }

  int levels() const { return int(levels_.size()); }
  const std::vector<int>& sizes() const { return sizes_; }
  const std::vector<int64_t>& nnz() const { return nnz_; }
  const std::vector<int64_t>& stored() const { return stored_; }  // sliced-ELL slots, padding included
  bool singular() const { return singular_; }
  double alpha() const { return alpha_; }
  bool aggressive() const { return aggressive_; }
  const AmgCfg& cfg() const { return cfg_; }

  // ---- FcgPc (FP64 flexible CG)
  int out_grid(int n) const override { return x_->pass_grid(n); }
  void apply(int k, const double* r, const double* rold, double* z, CgScalars* sc, const double* rr_part, int rr_g,
             double* out_part) override {
    cycle<double>(k, r, rold, z, 1, sc, rr_part, rr_g, out_part, amgk::OUTER);
  }
  // ---- FcgPc32 (FP32 sweeps)
  int out_grid32(int n) const override { return x_->pass_grid(n); }
  void apply32(int k, const float* r, const float* v, float* z, int zstride, CgScalars* sc, const double* rr_part,
               int rr_g, double* out_part) override {
    cycle<float>(k, r, v, z, zstride, sc, rr_part, rr_g, out_part, amgk::INNER);
  }

  // One V-cycle z = M r outside a solve (tests): no gate, no convergence test.
  template <class T>
  void vcycle(const T* r, T* z) {
    cycle<T>(0, r, r, z, 1, nullptr, nullptr, 0, nullptr, amgk::OUTER);
  }

 private:
  template <class T>
  T* up(const std::vector<T>& v) {
    T* p = x_->template alloc<T>(std::max<size_t>(1, v.size()));
    x_->upload(p, v.data(), v.size());
    owned_.push_back(p);
    return p;
  }
  static std::vector<AmgCoef> to_coef(const std::vector<float>& v) {
    std::vector<AmgCoef> h(v.size());
    for (size_t i = 0; i < v.size(); ++i) h[i] = coef_h(v[i]);
    return h;
  }
  template <class T>
  T* fresh(size_t n) {
    T* p = x_->template alloc<T>(std::max<size_t>(1, n));
    owned_.push_back(p);
    return p;
  }

  template <class T>
  void cycle(int k, const T* r, const T* rold, T* z, int zs, CgScalars* sc, const double* rr_part, int rr_g,
             double* out_part, int mode) {
    X& xx = *x_;
    const float om = float(cfg_.omega), al = float(alpha_);
    const int nl = int(levels_.size());
    const AmgLevelDev& L0 = levels_[0];
    if (sc)
      with_width(L0.w, [&](auto wc) {
        xx.tbegin(K_MG_DOWN);
        xx.template pass<1, 1>(L0.n, amgk::Down0<T, decltype(wc)::value>{L0, r, sc, mode, k, om}, rr_part, rr_g,
                               zpart_);
        xx.tend(K_MG_DOWN);
      });
    else
      xx.for_each(L0.n, amgk::Down0F<T>{L0, r, om});
    for (int l = 1; l < nl; ++l) {
      const AmgLevelDev& F = levels_[l - 1];
      const AmgLevelDev& C = levels_[l];
      if (C.n >= amgk::AMG_SMALL_ROWS) {
        xx.for_each(C.n, amgk::Restrict<1>{F.kptr, F.kidx, F.r, C.b, sc, mode});
        if (l < nl - 1) xx.for_each(C.n, amgk::DownC<1>{C, sc, mode, om});
      } else {
        xx.for_each(C.n * 4, amgk::Restrict<4>{F.kptr, F.kidx, F.r, C.b, sc, mode});
        if (l < nl - 1) xx.for_each(C.n * 4, amgk::DownC<4>{C, sc, mode, om});
      }
    }
    {
      const AmgLevelDev& F = levels_[nl - 2];
      xx.for_each(n_coarse_ * amgk::AMG_GC,
                  amgk::Coarsest{n_coarse_, inv_, levels_[nl - 1].b, F.kptr, F.kidx, F.x, sc, mode, al});
    }
    for (int l = nl - 2; l >= 1; --l) {
      const AmgLevelDev& C = levels_[l];
      const AmgLevelDev& F = levels_[l - 1];
      float* yout = (l == 1 && L0.agg) ? C.y : nullptr;
      if (C.n >= amgk::AMG_SMALL_ROWS)
        xx.for_each(C.n, amgk::UpC<1>{C, F.kptr, F.kidx, F.x, sc, mode, om, al, yout});
      else
        xx.for_each(C.n * 4, amgk::UpC<4>{C, F.kptr, F.kidx, F.x, sc, mode, om, al, yout});
    }
    if (L0.agg) xx.for_each(L0.n, amgk::Prolong0{L0, sc, mode, al});
    if (sc)
      with_width(L0.w, [&](auto wc) {
        xx.tbegin(K_MG_UP);
        xx.template pass<1, 5>(L0.n, amgk::Up0<T, decltype(wc)::value>{L0, r, rold, z, zs, sc, mode, om, al}, nullptr,
                               0, out_part);
        xx.tend(K_MG_UP);
      });
    else
      xx.for_each(L0.n, amgk::Up0F<T>{L0, r, z, om, al});
  }

  X* x_ = nullptr;
  AmgCfg cfg_;
  std::vector<AmgLevelDev> levels_;
  std::vector<int> sizes_;
  std::vector<int64_t> nnz_, stored_;
  std::vector<void*> owned_;
  float* inv_ = nullptr;
  double alpha_ = 1.5;
  bool aggressive_ = false;
  int n_coarse_ = 0;
  bool singular_ = false;
  double* zpart_ = nullptr;
};

}  // namespace dm
