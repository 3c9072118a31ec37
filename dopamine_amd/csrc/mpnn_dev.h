// Learned preconditioner for the pressure solve: a 3-layer, 32-channel message-passing network
// over the pressure operator's face graph, applied inside flexible CG (device only, gfx950).
//
// Network (one residual r in, one correction z out):
//   features  x_c = [r_c / s, (r_c / s) dbar / d_c, log(d_c / dbar)],  s = RMS(r), dbar = mean(d)
//   edges     a_cj = coef_cj / sqrt(d_c d_j)  (symmetric)
//   layer l   h_(l+1),c = tanh(Wa_l h_l,c + Wb_l m_l,c + b_l),  m_l,c = sum_j a_cj h_l,j,  h_0 = x
//   readout   o_c = wo . h_3,c
//   output    z_c = (r_c + s o_c) / d_c     (o = 0 is the Jacobi preconditioner)
// The per-cell dense layers run on the matrix cores (v_mfma_f32_32x32x16_bf16): bfloat16 weights
// and bfloat16 layer inputs [h, m] (m summed in FP32, then rounded), FP32 accumulation, FP32
// bias and tanh; stored activations are bfloat16. Training minimises ||A z - r||^2 / ||r||^2
// over residuals captured from the solver's own Jacobi CG solves, with FP32 master weights, a
// backward pass in FP32 (FP64 accumulation of the weight gradients) and Adam.
#ifndef DOPAMINE_MPNN_DEV_H
#define DOPAMINE_MPNN_DEV_H

#include <cmath>
#include <cstdint>
#include <random>
#include <stdexcept>
#include <vector>

#include "cg_impl.h"
#include "dm_device.h"

#define DM_DEV __device__

namespace dm {
namespace mp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int NCH = 32;
// flat parameter vector: Wa0 (32x3), Wb0 (32x3), b0, Wa1 (32x32), Wb1, b1, Wa2, Wb2, b2, wo
constexpr int OFF_WA[3] = {0, 224, 2304};
constexpr int OFF_WB[3] = {96, 1248, 3328};
constexpr int OFF_B[3] = {192, 2272, 4352};
constexpr int OFF_WO = 4384;
constexpr int NPARAM = 4416;
constexpr int KIN[3] = {3, 32, 32};

constexpr int TBLOCK = 256;  // 4 waves, 32 cells per wave and tile
constexpr int TGRID_MAX = 512;   // output layer: its block partials feed the next pass
constexpr int TGRID_BIG = 2048;  // input and hidden layers

// tanh from its definition: tanh x = sign(x) (1 - e) / (1 + e), e = exp(-2|x|), with the Taylor
// series x - x^3/3 + 2x^5/15 below |x| = 1/16 (truncation below 4e-9 relative there).
__device__ inline float tanh_f(float x) {
  const float ax = fabsf(x);
  if (ax < 0.0625f) {
    const float x2 = x * x;
    return x * (1.0f - x2 * (1.0f / 3.0f - x2 * (2.0f / 15.0f)));
  }
  const float e = __builtin_amdgcn_exp2f(-2.8853900817779268f * ax);  // 2 log2(e) |x|
  return copysignf((1.0f - e) / (1.0f + e), x);
}

__device__ inline __bf16 bf(float v) { return static_cast<__bf16>(v); }
__device__ inline float fl(__bf16 v) { return static_cast<float>(v); }

// Graph and static features of the network on one operator (device arrays).
struct Graph {
  int n = 0, w = 0;
  const int* nbr = nullptr;      // [w][n], shared with the FP64 operator
  const float* a = nullptr;      // [w][n] edge weights, 0 on empty slots
  const float* q = nullptr;      // dbar / d
  const float* f2 = nullptr;     // log(d / dbar)
  const float* mx2 = nullptr;    // sum_j a_cj f2_j
  const double* dinv = nullptr;  // 1 / d
};

// Weights as the matrix cores read them.
struct Weights {
  const bf16x8* W0 = nullptr;  // [32 rows][2 chunks]: [Wa0 | Wb0 | 0]
  const bf16x8* W1 = nullptr;  // [32 rows][8 chunks]: [Wa1 | Wb1]
  const bf16x8* W2 = nullptr;
  const __bf16* wo = nullptr;  // [32]
  const float* bias = nullptr; // [3][32]
};

// Slot kk of cell c read branch-free: slots past the width or without a neighbour read cell c
// itself with weight 0.
template <int WC>
struct Slots {
  static constexpr int count(int w) { return WC > 0 ? WC : w; }
};

__device__ inline void slot(const Graph& G, int kk, int c, int& j, float& a) {
  const bool live = kk < G.w;
  const size_t q = size_t(live ? kk : 0) * G.n + c;
  const int nb = G.nbr[q];
  const float av = G.a[q];
  j = (live && nb >= 0) ? nb : c;
  a = live ? av : 0.0f;
}

// Epilogue shared by the layers: D[v] (v = 0..15) of lane (cell j, half) holds output channel
// i = 8 (v / 4) + 4 half + v % 4; bias, tanh, bfloat16.
__device__ inline void act(const f32x16& d, const float* bias, float* h) {
#pragma unroll
  for (int v = 0; v < 16; ++v) h[v] = fl(bf(tanh_f(d[v] + bias[v])));
}
__device__ inline void store_h(__bf16* out, int j, int half, const float* h) {
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    bf16x4 o;
#pragma unroll
    for (int e = 0; e < 4; ++e) o[e] = bf(h[4 * g + e]);
    *reinterpret_cast<bf16x4*>(out + size_t(j) * NCH + 8 * g + 4 * half) = o;
  }
}

struct ScalS {
  double s, sinv;
};

// A wave's operands that do not change from tile to tile: its lanes' weight chunks (K blocks),
// and the bias (and readout) values of the 16 output channels the lane holds.
template <int KB>
struct WReg {
  bf16x8 a[KB];
  float bias[16], wo[16];
};
template <int KB>
__device__ inline WReg<KB> load_w(const bf16x8* Wl, const float* bias, const __bf16* wo, int lane) {
  WReg<KB> r;
  const int half = lane >> 5, row = (lane & 31) * (2 * KB);
#pragma unroll
  for (int kb = 0; kb < KB; ++kb) r.a[kb] = Wl[row + kb * 2 + half];
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int i = 8 * (v >> 2) + 4 * half + (v & 3);
    r.bias[v] = bias[i];
    r.wo[v] = wo ? fl(wo[i]) : 0.0f;
  }
  return r;
}

// Layer 0: features of the residual, their aggregate, one MFMA (K = 16: 6 inputs and zeros).
// In the solve (sc != null) the setup is also the FCG convergence test of iteration k.
template <int WC>
struct L0 {
  Graph G;
  Weights W;
  const double* r;
  __bf16* hout;      // [n][32]
  __bf16* xsave;     // [n][16] layer inputs (training) or null
  double* s_out;     // s for the later layers
  CgScalars* sc;     // null outside a solve
  int k;
  using Scal = ScalS;
  using W_t = WReg<1>;
  DM_DEV W_t wload(int lane) const { return load_w<1>(W.W0, W.bias, nullptr, lane); }
  DM_DEV bool setup(const Red<1>& t, bool leader, Scal& s) const {
    const double rr = t.v[0];
    if (sc) {
      if (sc->done) return false;
      const int fin = (sqrt(rr) <= sc->ttol) ? 1 : (k >= sc->maxit ? 2 : 0);
      if (leader) {
        sc->its = k;
        sc->rr = rr;
        if (fin) sc->done = fin;
      }
      if (fin) return false;
    }
    s.s = sqrt(rr / double(G.n));
    s.sinv = s.s > 0.0 ? 1.0 / s.s : 0.0;
    if (leader) *s_out = s.s;
    return true;
  }
  DM_DEV void tile(int t, int lane, const Scal& s, const W_t& w) const {
    const int half = lane >> 5, j = t * 32 + (lane & 31);
    const bool ok = j < G.n;
    const int jc = ok ? j : G.n - 1;
    bf16x8 bx;
#pragma unroll
    for (int e = 0; e < 8; ++e) bx[e] = bf(0.0f);
    if (half == 0) {
      const float x0 = float(r[jc] * s.sinv);
      const float x1 = x0 * G.q[jc];
      float m0 = 0.0f, m1 = 0.0f;
      for (int kk = 0; kk < Slots<WC>::count(G.w); ++kk) {
        int nb;
        float av;
        slot(G, kk, jc, nb, av);
        const float y0 = float(r[nb] * s.sinv);
        m0 = m0 + av * y0;
        m1 = m1 + av * (y0 * G.q[nb]);
      }
      bx[0] = bf(x0);
      bx[1] = bf(x1);
      bx[2] = bf(G.f2[jc]);
      bx[3] = bf(m0);
      bx[4] = bf(m1);
      bx[5] = bf(G.mx2[jc]);
    }
    f32x16 d = {};
    d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w.a[0], bx, d, 0, 0, 0);
    float h[16];
    act(d, w.bias, h);
    if (ok) {
      store_h(hout, j, half, h);
      if (xsave) *reinterpret_cast<bf16x8*>(xsave + size_t(j) * 16 + 8 * half) = bx;
    }
  }
};

// The MFMA operands of layers 1 and 2 for lane (cell jc, half): own h channels [8 half, +8) and
// [16 + 8 half, +8), and the same channels of m = sum a h_j, each as one bfloat16x8.
template <int WC>
__device__ inline void mid_inputs(const Graph& G, const __bf16* hin, int jc, int half, bf16x8* b) {
  const __bf16* row = hin + size_t(jc) * NCH + 8 * half;
  b[0] = *reinterpret_cast<const bf16x8*>(row);
  b[1] = *reinterpret_cast<const bf16x8*>(row + 16);
  float m[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) m[e] = 0.0f;
  if constexpr (WC > 0) {
    // all slot indices, then all neighbour rows, in flight together; summed in slot order
    int nb[WC];
    float av[WC];
    bf16x8 g0[WC], g1[WC];
#pragma unroll
    for (int kk = 0; kk < WC; ++kk) slot(G, kk, jc, nb[kk], av[kk]);
#pragma unroll
    for (int kk = 0; kk < WC; ++kk) {
      const __bf16* nr = hin + size_t(nb[kk]) * NCH + 8 * half;
      g0[kk] = *reinterpret_cast<const bf16x8*>(nr);
      g1[kk] = *reinterpret_cast<const bf16x8*>(nr + 16);
    }
#pragma unroll
    for (int kk = 0; kk < WC; ++kk)
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        m[e] = m[e] + av[kk] * fl(g0[kk][e]);
        m[8 + e] = m[8 + e] + av[kk] * fl(g1[kk][e]);
      }
  } else {
    for (int kk = 0; kk < G.w; ++kk) {
      int nb;
      float av;
      slot(G, kk, jc, nb, av);
      const __bf16* nr = hin + size_t(nb) * NCH + 8 * half;
      const bf16x8 g0 = *reinterpret_cast<const bf16x8*>(nr);
      const bf16x8 g1 = *reinterpret_cast<const bf16x8*>(nr + 16);
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        m[e] = m[e] + av * fl(g0[e]);
        m[8 + e] = m[8 + e] + av * fl(g1[e]);
      }
    }
  }
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    b[2][e] = bf(m[e]);
    b[3][e] = bf(m[8 + e]);
  }
}

__device__ inline f32x16 mid_mfma(const WReg<4>& w, const bf16x8* b) {
  f32x16 d = {};
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w.a[kb], b[kb], d, 0, 0, 0);
  return d;
}

// Layer 1 (hidden to hidden).
template <int WC>
struct L1 {
  Graph G;
  Weights W;
  const __bf16* hin;
  __bf16* hout;
  __bf16* xsave;  // [n][64] or null
  const CgScalars* sc;
  using Scal = cgk::NoScal;
  using W_t = WReg<4>;
  DM_DEV W_t wload(int lane) const { return load_w<4>(W.W1, W.bias + NCH, nullptr, lane); }
  DM_DEV bool setup(const Red<1>&, bool, Scal&) const { return !(sc && sc->done); }
  DM_DEV void tile(int t, int lane, const Scal&, const W_t& w) const {
    const int half = lane >> 5, j = t * 32 + (lane & 31);
    const bool ok = j < G.n;
    const int jc = ok ? j : G.n - 1;
    bf16x8 b[4];
    mid_inputs<WC>(G, hin, jc, half, b);
    const f32x16 d = mid_mfma(w, b);
    float h[16];
    act(d, w.bias, h);
    if (ok) {
      store_h(hout, j, half, h);
      if (xsave)
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) *reinterpret_cast<bf16x8*>(xsave + size_t(j) * 64 + kb * 16 + 8 * half) = b[kb];
    }
  }
};

// Layer 2 with the readout: z = (r + s o) / d, and the FCG partials z.r, z.r_old, sum z, sum r,
// sum r_old.
template <int WC>
struct L2 {
  Graph G;
  Weights W;
  const __bf16* hin;
  const double *r, *rold;  // rold may be null (then r)
  double* z;
  const double* s_in;
  __bf16* xsave;   // [n][64] or null
  __bf16* h3save;  // [n][32] or null
  const CgScalars* sc;
  struct Scal {
    double s;
  };
  using W_t = WReg<4>;
  DM_DEV W_t wload(int lane) const { return load_w<4>(W.W2, W.bias + 2 * NCH, W.wo, lane); }
  DM_DEV bool setup(const Red<1>&, bool, Scal& s) const {
    if (sc && sc->done) return false;
    s.s = *s_in;
    return true;
  }
  DM_DEV Red<5> tile(int t, int lane, const Scal& s, const W_t& w) const {
    const int half = lane >> 5, j = t * 32 + (lane & 31);
    const bool ok = j < G.n;
    const int jc = ok ? j : G.n - 1;
    bf16x8 b[4];
    mid_inputs<WC>(G, hin, jc, half, b);
    const f32x16 d = mid_mfma(w, b);
    float h[16];
    act(d, w.bias, h);
    float op = 0.0f;
#pragma unroll
    for (int v = 0; v < 16; ++v) op = op + w.wo[v] * h[v];
    // the partner half's 16 channels; every lane takes part in the exchange (a + b == b + a, so
    // both halves hold the same o)
    const float o = op + __shfl_xor(op, 32);
    Red<5> acc = red_identity<5, false>();
    if (ok) {
      if (xsave)
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) *reinterpret_cast<bf16x8*>(xsave + size_t(j) * 64 + kb * 16 + 8 * half) = b[kb];
      if (h3save) store_h(h3save, j, half, h);
      if (half == 0) {
        const double rc = r[j], ro = rold ? rold[j] : rc;
        const double zc = (rc + s.s * double(o)) * G.dinv[j];
        z[j] = zc;
        acc.v[0] = zc * rc;
        acc.v[1] = zc * ro;
        acc.v[2] = zc;
        acc.v[3] = rc;
        acc.v[4] = ro;
      }
    }
    return acc;
  }
};

template <class F>
struct TileOut {
  static constexpr int value = 0;
};
template <int WC>
struct TileOut<L2<WC>> {
  static constexpr int value = 5;
};

// A chained pass over tiles of 32 cells per wave (see k_pass): sums the previous pass's partials
// (NIN quantities), lets the functor set up (or stop the grid), then each wave takes tiles of its
// block's contiguous range; NOUT partials per block.
template <int NIN, class F>
__global__ void __launch_bounds__(TBLOCK) k_tile(int ntiles, int tpb, F f, const double* in_part, int in_g,
                                                 double* out_part) {
  constexpr int NOUT = TileOut<F>::value;
  Red<NIN> tot = red_identity<NIN, false>();
  for (int q = threadIdx.x; q < in_g; q += TBLOCK)
#pragma unroll
    for (int k = 0; k < NIN; ++k) tot.v[k] = tot.v[k] + in_part[size_t(k) * in_g + q];
  tot = block_reduce<NIN, false, TBLOCK>(tot);
  const int b = xcd_block(gridDim.x);
  typename F::Scal sc;
  if (!f.setup(tot, b == 0 && threadIdx.x == 0, sc)) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int t0 = b * tpb, t1 = min(ntiles, t0 + tpb);
  const typename F::W_t w = f.wload(lane);
  if constexpr (NOUT > 0) {
    Red<NOUT> acc = red_identity<NOUT, false>();
    for (int t = t0 + wave; t < t1; t += TBLOCK / 64) {
      const Red<NOUT> v = f.tile(t, lane, sc, w);
#pragma unroll
      for (int k = 0; k < NOUT; ++k) acc.v[k] = acc.v[k] + v.v[k];
    }
    acc = block_reduce<NOUT, false, TBLOCK>(acc);
    if (threadIdx.x == 0)
#pragma unroll
      for (int k = 0; k < NOUT; ++k) out_part[size_t(k) * gridDim.x + b] = acc.v[k];
  } else {
    for (int t = t0 + wave; t < t1; t += TBLOCK / 64) f.tile(t, lane, sc, w);
  }
}

// ---- set-up kernels

// a = coef / sqrt(d_c d_j) (0 on empty slots), q = dbar / d, f2 = log(d / dbar), 1/d
struct GraphBuild {
  EllOp A;
  const double* dbar;
  float *a, *q, *f2;
  double* dinv;
  DM_DEV void operator()(int c) const {
    const int n = A.n;
    const double dc = A.diag[c], db = *dbar;
    for (int k = 0; k < A.w; ++k) {
      const size_t i = size_t(k) * n + c;
      const int j = A.nbr[i];
      a[i] = j >= 0 ? float(A.coef[i] / sqrt(dc * A.diag[j])) : 0.0f;
    }
    q[c] = float(db / dc);
    f2[c] = float(log(dc / db));
    dinv[c] = 1.0 / dc;
  }
};
struct GraphMx2 {
  int n, w;
  const int* nbr;
  const float *a, *f2;
  float* mx2;
  DM_DEV void operator()(int c) const {
    float m = 0.0f;
    for (int k = 0; k < w; ++k) {
      const size_t i = size_t(k) * n + c;
      const int j = nbr[i];
      if (j >= 0) m = m + a[i] * f2[j];
    }
    mx2[c] = m;
  }
};
struct DiagSum {
  const double* d;
  DM_DEV Red<1> operator()(int c) const {
    Red<1> o;
    o.v[0] = d[c];
    return o;
  }
};
struct MeanFin {
  double* out;
  int n;
  DM_DEV void operator()(const Red<1>& s) const { *out = s.v[0] / double(n); }
};

// master FP32 parameters -> matrix-core layout
struct Pack {
  const float* p;
  __bf16 *W0, *W1, *W2, *wo;
  float* bias;
  DM_DEV void operator()(int e) const {
    if (e < 32 * 16) {  // W0 row i: Wa0[i][0..2], Wb0[i][0..2], zeros
      const int i = e >> 4, k = e & 15;
      const float v = k < 3 ? p[OFF_WA[0] + i * 3 + k] : (k < 6 ? p[OFF_WB[0] + i * 3 + k - 3] : 0.0f);
      W0[e] = bf(v);
      return;
    }
    e -= 32 * 16;
    if (e < 2 * 32 * 64) {  // W1, W2 row i: Wa[i][0..31], Wb[i][0..31]
      const int l = 1 + e / (32 * 64), r = e % (32 * 64), i = r >> 6, k = r & 63;
      const float v = k < 32 ? p[OFF_WA[l] + i * 32 + k] : p[OFF_WB[l] + i * 32 + k - 32];
      (l == 1 ? W1 : W2)[r] = bf(v);
      return;
    }
    e -= 2 * 32 * 64;
    if (e < 32) {
      wo[e] = bf(p[OFF_WO + e]);
      return;
    }
    e -= 32;
    if (e < 96) bias[e] = p[OFF_B[e >> 5] + (e & 31)];
  }
  static constexpr int count() { return 32 * 16 + 2 * 32 * 64 + 32 + 96; }
};

// ---- training kernels

struct SqSum {
  const double* r;
  using Scal = cgk::NoScal;
  DM_HD bool setup(const Red<1>&, bool, Scal&) const { return true; }
  DM_HD Red<1> operator()(int c, const Scal&) const {
    Red<1> o;
    o.v[0] = r[c] * r[c];
    return o;
  }
};

// y = A z - r; partials y.y, r.r
struct LossY {
  EllOp A;
  const double *z, *r;
  double* y;
  using Scal = cgk::NoScal;
  DM_HD bool setup(const Red<1>&, bool, Scal&) const { return true; }
  DM_HD Red<2> operator()(int c, const Scal&) const {
    double v = A.diag[c] * z[c];
    for (int k = 0; k < A.w; ++k) {
      const size_t i = size_t(k) * A.n + c;
      const int j = A.nbr[i];
      if (j >= 0) v = v - A.coef[i] * z[j];
    }
    const double yc = v - r[c];
    y[c] = yc;
    Red<2> o;
    o.v[0] = yc * yc;
    o.v[1] = r[c] * r[c];
    return o;
  }
};

// loss = y.y / r.r; dL/do = 2 (A y) / r.r * s / d
struct GradO {
  EllOp A;
  const double *y, *s_in, *dinv;
  float* go;
  double* loss;
  struct Scal {
    double f;
  };
  DM_HD bool setup(const Red<2>& t, bool leader, Scal& s) const {
    if (leader) *loss = t.v[0] / t.v[1];
    s.f = 2.0 / t.v[1] * (*s_in);
    return true;
  }
  DM_HD Red<1> operator()(int c, const Scal& s) const {
    double v = A.diag[c] * y[c];
    for (int k = 0; k < A.w; ++k) {
      const size_t i = size_t(k) * A.n + c;
      const int j = A.nbr[i];
      if (j >= 0) v = v - A.coef[i] * y[j];
    }
    go[c] = float(v * s.f * dinv[c]);
    return Red<1>{};
  }
};

// top layer: ga = go wo (1 - h3^2)
struct GaTop {
  const float* go;
  const __bf16 *h3, *wo;
  float* ga;
  DM_DEV void operator()(int c) const {
    const float g = go[c];
    for (int i = 0; i < NCH; ++i) {
      const float h = fl(h3[size_t(c) * NCH + i]);
      ga[size_t(c) * NCH + i] = (g * fl(wo[i])) * (1.0f - h * h);
    }
  }
};

// t = ga Wb, u = ga Wa (rows of the bfloat16 weights, layer l = 1, 2)
struct BackTU {
  const float* ga;
  const __bf16* Wl;  // [32][64]
  float *t, *u;
  DM_DEV void operator()(int c) const {
    float g[NCH];
    for (int i = 0; i < NCH; ++i) g[i] = ga[size_t(c) * NCH + i];
    for (int k = 0; k < NCH; ++k) {
      float su = 0.0f, st = 0.0f;
      for (int i = 0; i < NCH; ++i) {
        su = su + g[i] * fl(Wl[i * 64 + k]);
        st = st + g[i] * fl(Wl[i * 64 + 32 + k]);
      }
      u[size_t(c) * NCH + k] = su;
      t[size_t(c) * NCH + k] = st;
    }
  }
};

// gh = u + sum_j a_cj t_j; ga of the layer below = gh (1 - h^2), h = that layer's output
// (the first 32 inputs saved by the layer above)
struct BackGh {
  Graph G;
  const float *t, *u;
  const __bf16* xs;  // [n][64]: h in the first 32
  float* ga;
  DM_DEV void operator()(int c) const {
    float gh[NCH];
    for (int k = 0; k < NCH; ++k) gh[k] = 0.0f;
    for (int kk = 0; kk < G.w; ++kk) {
      const size_t i = size_t(kk) * G.n + c;
      const int j = G.nbr[i];
      if (j < 0) continue;
      const float av = G.a[i];
      for (int k = 0; k < NCH; ++k) gh[k] = gh[k] + av * t[size_t(j) * NCH + k];
    }
    for (int k = 0; k < NCH; ++k) {
      const float h = fl(xs[size_t(c) * 64 + k]);
      ga[size_t(c) * NCH + k] = (u[size_t(c) * NCH + k] + gh[k]) * (1.0f - h * h);
    }
  }
};

// Weight gradients of layer l: G[i][k] = sum_c ga[c][i] X[c][k] (X = saved inputs, row stride
// XS, K used columns), bias sum_c ga[c][i], and for the top layer sum_c go_c h3[c][i]. Cells
// are staged 32 at a time in LDS; block partials in FP64, layout [block][NPARAM].
template <int K>
__global__ void __launch_bounds__(256) k_wgrad(int n, int chunk, const float* ga, const __bf16* X, int XS, int layer,
                                               const float* go, const __bf16* h3, double* part) {
  constexpr int NPT = K / 8;  // outputs per thread (32 K / 256)
  __shared__ float sga[32][NCH + 1];
  __shared__ float sx[32][K + 1];
  __shared__ float sgo[32];
  const int b = blockIdx.x, t = threadIdx.x;
  const int lo = b * chunk, hi = min(n, lo + chunk);
  double acc[NPT];
#pragma unroll
  for (int e = 0; e < NPT; ++e) acc[e] = 0.0;
  double bacc = 0.0, wacc = 0.0;
  const int o0 = t * NPT, oi = o0 / K, ok0 = o0 % K;
  for (int c0 = lo; c0 < hi; c0 += 32) {
    const int m = min(32, hi - c0);
    for (int e = t; e < 32 * NCH; e += 256) {
      const int cc = e / NCH, i = e % NCH;
      sga[cc][i] = cc < m ? ga[size_t(c0 + cc) * NCH + i] : 0.0f;
    }
    for (int e = t; e < 32 * K; e += 256) {
      const int cc = e / K, k = e % K;
      sx[cc][k] = cc < m ? fl(X[size_t(c0 + cc) * XS + k]) : 0.0f;
    }
    if (t < 32) sgo[t] = (go && t < m) ? go[c0 + t] : 0.0f;
    __syncthreads();
    for (int cc = 0; cc < m; ++cc) {
      const double gi = sga[cc][oi];
#pragma unroll
      for (int e = 0; e < NPT; ++e) acc[e] = acc[e] + gi * double(sx[cc][ok0 + e]);
    }
    if (t < 32) {
      for (int cc = 0; cc < m; ++cc) bacc = bacc + double(sga[cc][t]);
      if (go)
        for (int cc = 0; cc < m; ++cc) wacc = wacc + double(sgo[cc]) * double(fl(h3[size_t(c0 + cc) * NCH + t]));
    }
    __syncthreads();
  }
  double* pb = part + size_t(b) * NPARAM;
  const int kin = KIN[layer];
#pragma unroll
  for (int e = 0; e < NPT; ++e) {
    const int k = ok0 + e;
    if (k < kin)
      pb[OFF_WA[layer] + oi * kin + k] = acc[e];
    else if (k < 2 * kin)
      pb[OFF_WB[layer] + oi * kin + k - kin] = acc[e];
  }
  if (t < 32) {
    pb[OFF_B[layer] + t] = bacc;
    if (go) pb[OFF_WO + t] = wacc;
  }
}

// gsum[p] (+)= sum over blocks of part[b][p], in block order
struct GradReduce {
  const double* part;
  int nb;
  double* gsum;
  int accumulate;
  DM_DEV void operator()(int p) const {
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s = s + part[size_t(b) * NPARAM + p];
    gsum[p] = accumulate ? gsum[p] + s : s;
  }
};

// Adam (Kingma and Ba 2015, Algorithm 1) on the FP32 master parameters; g = gsum * scale.
struct AdamStep {
  float *p, *m, *v;
  const double* gsum;
  double scale, lr, b1, b2, eps, c1, c2;  // c1 = 1 - b1^t, c2 = 1 - b2^t
  DM_DEV void operator()(int i) const {
    const double g = gsum[i] * scale;
    const double mi = b1 * double(m[i]) + (1.0 - b1) * g;
    const double vi = b2 * double(v[i]) + (1.0 - b2) * g * g;
    m[i] = float(mi);
    v[i] = float(vi);
    p[i] = float(double(p[i]) - lr * (mi / c1) / (sqrt(vi / c2) + eps));
  }
};

// y[c] = x[c * stride]
struct Gather {
  const double* x;
  int stride;
  double* y;
  DM_DEV void operator()(int c) const { y[c] = x[size_t(c) * stride]; }
};

template <class Fn>
inline void by_width(int w, Fn&& fn) {
  if (w <= 4)
    fn(std::integral_constant<int, 4>{});
  else if (w <= 6)
    fn(std::integral_constant<int, 6>{});
  else if (w <= 8)
    fn(std::integral_constant<int, 8>{});
  else
    fn(std::integral_constant<int, 0>{});
}

// The network on one operator: weights, training state, sample store, and its application as
// the FCG preconditioner.
class LearnedPc : public FcgPc, public CgCapture {
 public:
  // A: the operator's device arrays (kept by the caller). capacity: residual samples kept.
  void init(DeviceExec* xx, const EllOp& A, int capacity, uint64_t seed) {
    x_ = xx;
    A_ = A;
    const int n = A.n, w = std::max(1, A.w);
    n_ = n;
    const size_t nw = size_t(w) * n;
    a_ = xx->alloc<float>(nw);
    q_ = xx->alloc<float>(n);
    f2_ = xx->alloc<float>(n);
    mx2_ = xx->alloc<float>(n);
    dinv_ = xx->alloc<double>(n);
    dbar_ = xx->alloc<double>(1);
    xx->reduce<1, false>(n, DiagSum{A.diag}, MeanFin{dbar_, n});
    xx->for_each(n, GraphBuild{A, dbar_, a_, q_, f2_, dinv_});
    xx->for_each(n, GraphMx2{n, A.w, A.nbr, a_, f2_, mx2_});
    G_.n = n;
    G_.w = A.w;
    G_.nbr = A.nbr;
    G_.a = a_;
    G_.q = q_;
    G_.f2 = f2_;
    G_.mx2 = mx2_;
    G_.dinv = dinv_;
    W0_ = xx->alloc<__bf16>(32 * 16);
    W1_ = xx->alloc<__bf16>(32 * 64);
    W2_ = xx->alloc<__bf16>(32 * 64);
    wo_ = xx->alloc<__bf16>(32);
    bias_ = xx->alloc<float>(96);
    W_.W0 = reinterpret_cast<const bf16x8*>(W0_);
    W_.W1 = reinterpret_cast<const bf16x8*>(W1_);
    W_.W2 = reinterpret_cast<const bf16x8*>(W2_);
    W_.wo = wo_;
    W_.bias = bias_;
    p_ = xx->alloc<float>(NPARAM);
    m_ = xx->alloc<float>(NPARAM);
    v_ = xx->alloc<float>(NPARAM);
    gsum_ = xx->alloc<double>(NPARAM);
    h1_ = xx->alloc<__bf16>(size_t(n) * NCH);
    h2_ = xx->alloc<__bf16>(size_t(n) * NCH);
    s_ = xx->alloc<double>(1);
    g_ = xx->pass_grid(n);
    rrp_ = xx->alloc<double>(size_t(2) * g_);
    zp_ = xx->alloc<double>(size_t(5) * out_grid(n));
    cap_ = std::max(0, capacity);
    if (cap_) samples_ = xx->alloc<double>(size_t(cap_) * n);
    rng_.seed(seed);
    seed_ = seed;
    init_params(seed);
  }
  void destroy() {
    if (!x_) return;
    for (void* q : {(void*)a_, (void*)q_, (void*)f2_, (void*)mx2_, (void*)dinv_, (void*)dbar_, (void*)W0_, (void*)W1_,
                    (void*)W2_, (void*)wo_, (void*)bias_, (void*)p_, (void*)m_, (void*)v_, (void*)gsum_, (void*)h1_,
                    (void*)h2_, (void*)s_, (void*)rrp_, (void*)zp_, (void*)samples_, (void*)x0s_, (void*)x1s_,
                    (void*)x2s_, (void*)h3s_, (void*)z_, (void*)y_, (void*)go_, (void*)ga_, (void*)gb_, (void*)t_,
                    (void*)u_, (void*)wpart_, (void*)lossd_, (void*)lpart_})
      if (q) x_->release(q);
    x_ = nullptr;
  }
  ~LearnedPc() override { destroy(); }

  // Glorot-style uniform initial weights (fixed seed), zero biases, zero readout: the untrained
  // network is the Jacobi preconditioner.
  void init_params(uint64_t seed) {
    std::mt19937_64 g(seed);
    std::vector<float> h(NPARAM, 0.0f);
    for (int l = 0; l < 3; ++l) {
      const double lim = std::sqrt(6.0 / double(KIN[l] + NCH));
      std::uniform_real_distribution<double> U(-lim, lim);
      for (int i = 0; i < NCH * KIN[l]; ++i) {
        h[OFF_WA[l] + i] = float(U(g));
        h[OFF_WB[l] + i] = float(U(g));
      }
    }
    set_params(h.data());
    x_->zero(m_, NPARAM);
    x_->zero(v_, NPARAM);
    adam_t_ = 0;
  }
  void set_params(const float* host) {
    x_->upload(p_, host, NPARAM);
    pack();
  }
  void get_params(float* host) { x_->download(host, p_, NPARAM); }

  // ---- FcgPc
  int out_grid(int n) const override {
    const int nt = (n + 31) / 32;
    return std::max(1, std::min(TGRID_MAX, (nt + 3) / 4));
  }
  void apply(int k, const double* r, const double* rold, double* z, CgScalars* sc, const double* rr_part, int rr_g,
             double* out_part) override {
    layers(r, rold, z, sc, k, rr_part, rr_g, out_part, false);
  }

  // ---- CgCapture: keep a copy of the residual (ring buffer of cap samples)
  bool capturing = false;
  int capture_every = 1;
  void residual(int k, const double* r, int stride) override {
    (void)k;
    if (!capturing || !cap_) return;
    if ((seen_++ % std::max(1, capture_every)) != 0) return;
    const int slot = nsamp_ < cap_ ? nsamp_ : int(seen_ % uint64_t(cap_));
    x_->for_each(n_, Gather{r, stride, samples_ + size_t(slot) * n_});
    if (nsamp_ < cap_) ++nsamp_;
  }
  int samples() const { return nsamp_; }
  void add_sample(const double* r_dev) {
    const bool c = capturing;
    capturing = true;
    const uint64_t s = seen_;
    seen_ = 0;
    const int ce = capture_every;
    capture_every = 1;
    residual(0, r_dev, 1);
    capturing = c;
    capture_every = ce;
    seen_ = s + 1;
  }

  // ---- forward on its own: z = M(r); train saves what the backward pass reads
  void forward(const double* r, double* z, bool train) {
    if (train) ensure_train();
    x_->template pass<1, 1>(n_, SqSum{r}, nullptr, 0, rrp_);
    layers(r, nullptr, z, nullptr, 0, rrp_, g_, zp_, train);
  }
  // activations of the last training forward, bfloat16 [3][n][32]
  void activations(uint16_t* host) {
    const size_t m = size_t(n_) * NCH;
    x_->download(reinterpret_cast<__bf16*>(host), h1_, m);
    x_->download(reinterpret_cast<__bf16*>(host) + m, h2_, m);
    x_->download(reinterpret_cast<__bf16*>(host) + 2 * m, h3s_, m);
  }
  // loss of residual r (device) and its weight gradient, added to (or stored in) gsum; the
  // loss goes to lossd_[slot]
  void loss_grad(const double* r, bool accumulate, int slot) {
    ensure_train();
    forward(r, z_, true);
    x_->template pass<1, 2>(n_, LossY{A_, z_, r, y_}, nullptr, 0, lpart_);
    x_->template pass<2, 1>(n_, GradO{A_, y_, s_, dinv_, go_, lossd_ + slot}, lpart_, g_, lpart_ + 2 * size_t(g_));
    x_->for_each(n_, GaTop{go_, h3s_, wo_, ga_});
    wgrad<64>(ga_, x2s_, 64, 2, go_, h3s_);
    x_->for_each(n_, BackTU{ga_, W2_, t_, u_});
    x_->for_each(n_, BackGh{G_, t_, u_, x2s_, gb_});
    wgrad<64>(gb_, x1s_, 64, 1, nullptr, nullptr);
    x_->for_each(n_, BackTU{gb_, W1_, t_, u_});
    x_->for_each(n_, BackGh{G_, t_, u_, x1s_, ga_});
    wgrad<8>(ga_, x0s_, 16, 0, nullptr, nullptr);
    x_->for_each(NPARAM, GradReduce{wpart_, wgrid(), gsum_, accumulate ? 1 : 0});
  }
  void gradient(double* host) { x_->download(host, gsum_, NPARAM); }
  double loss(int slot) {
    double v = 0.0;
    x_->download(&v, lossd_ + slot, 1);
    return v;
  }
  // Adam on the mean gradient of gsum over `count` samples
  void adam(double lr, int count) {
    ++adam_t_;
    const double b1 = 0.9, b2 = 0.999;
    const double c1 = 1.0 - std::pow(b1, double(adam_t_)), c2 = 1.0 - std::pow(b2, double(adam_t_));
    x_->for_each(NPARAM, AdamStep{p_, m_, v_, gsum_, 1.0 / double(count), lr, b1, b2, 1e-8, c1, c2});
    pack();
  }
  // iters Adam steps, each on `batch` residuals drawn from the sample store; loss_hist (host,
  // may be null) receives each step's mean loss.
  int train(int iters, int batch, double lr, double* loss_hist) {
    if (nsamp_ == 0) return -1;
    ensure_train();
    batch = std::max(1, batch);
    for (int it = 0; it < iters; ++it) {
      for (int b = 0; b < batch; ++b) {
        const int i = int(rng_() % uint64_t(nsamp_));
        loss_grad(samples_ + size_t(i) * n_, b > 0, (it % LOSS_SLOTS_STEPS) * batch_cap() + b % batch_cap());
      }
      adam(lr, batch);
      if (loss_hist && ((it + 1) % LOSS_SLOTS_STEPS == 0 || it + 1 == iters)) {
        const int done = (it % LOSS_SLOTS_STEPS) + 1;
        std::vector<double> all(size_t(done) * batch_cap());
        x_->download(all.data(), lossd_, all.size());
        for (int q = 0; q < done; ++q) {
          double s = 0.0;
          const int m = std::min(batch, batch_cap());
          for (int b = 0; b < m; ++b) s += all[size_t(q) * batch_cap() + b];
          loss_hist[it + 1 - done + q] = s / double(m);
        }
      }
    }
    trained_steps_ += iters;
    return 0;
  }
  int trained_steps() const { return trained_steps_; }

 private:
  static constexpr int LOSS_SLOTS_STEPS = 64;
  static int batch_cap() { return 16; }

  int wgrid() const { return std::max(1, std::min(256, (n_ + 255) / 256)); }
  template <int K>
  void wgrad(const float* ga, const __bf16* X, int XS, int layer, const float* go, const __bf16* h3) {
    const int gb = wgrid(), chunk = (n_ + gb - 1) / gb;
    x_->launch(k_wgrad<K>, gb, 256, n_, chunk, ga, X, XS, layer, go, h3, wpart_);
  }
  void pack() {
    x_->for_each(Pack::count(), Pack{p_, W0_, W1_, W2_, wo_, bias_});
  }
  void ensure_train() {
    if (z_) return;
    const size_t n = size_t(n_);
    x0s_ = x_->alloc<__bf16>(n * 16);
    x1s_ = x_->alloc<__bf16>(n * 64);
    x2s_ = x_->alloc<__bf16>(n * 64);
    h3s_ = x_->alloc<__bf16>(n * NCH);
    z_ = x_->alloc<double>(n);
    y_ = x_->alloc<double>(n);
    go_ = x_->alloc<float>(n);
    ga_ = x_->alloc<float>(n * NCH);
    gb_ = x_->alloc<float>(n * NCH);
    t_ = x_->alloc<float>(n * NCH);
    u_ = x_->alloc<float>(n * NCH);
    wpart_ = x_->alloc<double>(size_t(wgrid()) * NPARAM);
    lossd_ = x_->alloc<double>(size_t(LOSS_SLOTS_STEPS) * batch_cap());
    lpart_ = x_->alloc<double>(size_t(3) * g_);
  }
  void layers(const double* r, const double* rold, double* z, CgScalars* sc, int k, const double* rr_part, int rr_g,
              double* out_part, bool train) {
    const int nt = (n_ + 31) / 32, gt = out_grid(n_), tpb = (nt + gt - 1) / gt;
    const int gb = std::max(1, std::min(TGRID_BIG, (nt + 3) / 4)), tpbb = (nt + gb - 1) / gb;
    by_width(G_.w, [&](auto wc) {
      constexpr int WC = decltype(wc)::value;
      x_->launch(k_tile<1, L0<WC>>, gb, TBLOCK, nt, tpbb,
                 L0<WC>{G_, W_, r, h1_, train ? x0s_ : nullptr, s_, sc, k}, rr_part, rr_g, (double*)nullptr);
      x_->tbegin(K_CG_DIR);  // the hidden layer is the timed kernel of a preconditioned iteration
      x_->launch(k_tile<1, L1<WC>>, gb, TBLOCK, nt, tpbb, L1<WC>{G_, W_, h1_, h2_, train ? x1s_ : nullptr, sc},
                 (const double*)nullptr, 0, (double*)nullptr);
      x_->tend(K_CG_DIR);
      x_->launch(k_tile<1, L2<WC>>, gt, TBLOCK, nt, tpb,
                 L2<WC>{G_, W_, h2_, r, rold, z, s_, train ? x2s_ : nullptr, train ? h3s_ : nullptr, sc},
                 (const double*)nullptr, 0, out_part);
    });
  }

  DeviceExec* x_ = nullptr;
  EllOp A_;
  Graph G_;
  Weights W_;
  int n_ = 0, g_ = 1, cap_ = 0, nsamp_ = 0, adam_t_ = 0, trained_steps_ = 0;
  uint64_t seen_ = 0, seed_ = 0;
  std::mt19937_64 rng_;
  float *a_ = nullptr, *q_ = nullptr, *f2_ = nullptr, *mx2_ = nullptr, *bias_ = nullptr;
  double *dinv_ = nullptr, *dbar_ = nullptr, *s_ = nullptr, *rrp_ = nullptr, *zp_ = nullptr, *samples_ = nullptr;
  __bf16 *W0_ = nullptr, *W1_ = nullptr, *W2_ = nullptr, *wo_ = nullptr, *h1_ = nullptr, *h2_ = nullptr;
  float *p_ = nullptr, *m_ = nullptr, *v_ = nullptr;
  double* gsum_ = nullptr;
  // training buffers
  __bf16 *x0s_ = nullptr, *x1s_ = nullptr, *x2s_ = nullptr, *h3s_ = nullptr;
  double *z_ = nullptr, *y_ = nullptr, *wpart_ = nullptr, *lossd_ = nullptr, *lpart_ = nullptr;
  float *go_ = nullptr, *ga_ = nullptr, *gb_ = nullptr, *t_ = nullptr, *u_ = nullptr;
};

struct LearnCfg {
  int collect = 600;      // Jacobi solves whose residuals are captured before training
  int capacity = 128;     // residual samples kept
  int capture_every = 2;  // keep one residual in this many batch boundaries
  int iters = 300;        // Adam steps
  int batch = 4;          // residuals per step
  double lr = 3e-3;
  int max_regress = 3;    // consecutive regressions before the learned solve is switched off
  uint64_t seed = 1;
};

// Collect residuals from the first `collect` Jacobi CG solves, train, then solve with FCG and
// the network; a learned solve that does not converge, whose true residual exceeds the
// tolerance, or that needs more iterations than the Jacobi solves did, counts as a regression;
// a failed one is repeated with Jacobi CG from the same initial guess. After max_regress
// regressions in a row the solver stays with Jacobi CG.
class LearnedPolicy : public PressurePolicy {
 public:
  enum Phase { COLLECT = 0, USE = 1, OFF = 2 };
  LearnedPolicy(DeviceExec* xx, CgT<DeviceExec>* cg, const EllOp& A, const LearnCfg& cfg) : x_(xx), cg_(cg), cfg_(cfg) {
    pc_.init(xx, A, cfg.capacity, cfg.seed);
    pc_.capture_every = cfg.capture_every;
    x0_ = xx->alloc<double>(A.n);
  }
  ~LearnedPolicy() override {
    pc_.destroy();
    if (x0_) x_->release(x0_);
  }
  int solve(const EllOp& A, const double* b, double* x, double rtol, double atol, const double* atol_dev, int maxit,
            bool nullspace, int* its, double* resid) override {
    if (phase == COLLECT) {
      pc_.capturing = true;
      cg_->set_capture(&pc_);
      const int rc = cg_->solve(A, b, x, rtol, atol, atol_dev, maxit, nullspace, its, resid);
      cg_->set_capture(nullptr);
      pc_.capturing = false;
      note_jacobi(*its);
      if (++collected >= cfg_.collect && pc_.samples() > 0) train_now();
      return rc;
    }
    if (phase == OFF) {
      const int rc = cg_->solve(A, b, x, rtol, atol, atol_dev, maxit, nullspace, its, resid);
      note_jacobi(*its);
      return rc;
    }
    x_->copy(x0_, x, size_t(A.n));
    cg_->set_fcg(&pc_);
    int rc = cg_->solve(A, b, x, rtol, atol, atol_dev, maxit, nullspace, its, resid);
    cg_->set_fcg(nullptr);
    ++learned_solves;
    learned_its += *its;
    const bool failed = rc != 0 || !(*resid <= 2.0 * cg_->last_ttol());
    const bool slower = ref_its > 0.0 && double(*its) > ref_its;
    if (failed) {
      ++fallbacks;
      x_->copy(x, x0_, size_t(A.n));
      rc = cg_->solve(A, b, x, rtol, atol, atol_dev, maxit, nullspace, its, resid);
    }
    if (failed || slower) {
      if (++regress_run >= cfg_.max_regress) phase = OFF;
    } else {
      regress_run = 0;
    }
    return rc;
  }
  int precondition(const EllOp& A, const double* b, double* x, double rtol, int maxit, bool nullspace, int* its,
                   double* resid) override {
    if (phase != USE) return cg_->solve(A, b, x, rtol, 1.0e-50, nullptr, maxit, nullspace, its, resid);
    cg_->set_fcg(&pc_);
    const int rc = cg_->solve(A, b, x, rtol, 1.0e-50, nullptr, maxit, nullspace, its, resid);
    cg_->set_fcg(nullptr);
    ++precond_solves;
    precond_its += *its;
    return rc;
  }
  void train_now() {
    loss_hist.assign(size_t(std::max(0, cfg_.iters)), 0.0);
    hipEvent_t e0, e1;
    DM_CHECK(hipEventCreate(&e0));
    DM_CHECK(hipEventCreate(&e1));
    DM_CHECK(hipEventRecord(e0, x_->stream));
    pc_.train(cfg_.iters, cfg_.batch, cfg_.lr, loss_hist.data());
    DM_CHECK(hipEventRecord(e1, x_->stream));
    DM_CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    DM_CHECK(hipEventElapsedTime(&ms, e0, e1));
    train_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    phase = USE;
  }
  LearnedPc& pc() { return pc_; }

  Phase phase = COLLECT;
  int collected = 0, learned_solves = 0, fallbacks = 0, regress_run = 0;
  int64_t learned_its = 0;
  int64_t precond_solves = 0, precond_its = 0;  // precondition() calls that ran the learned FCG
  double ref_its = 0.0, train_ms = 0.0;
  std::vector<double> loss_hist;

 private:
  // reference iteration count: mean over the Jacobi solves seen so far
  void note_jacobi(int its) {
    ++jacobi_solves;
    ref_its += (double(its) - ref_its) / double(jacobi_solves);
  }
  DeviceExec* x_;
  CgT<DeviceExec>* cg_;
  LearnCfg cfg_;
  LearnedPc pc_;
  double* x0_ = nullptr;
  int jacobi_solves = 0;
};

}  // namespace mp
}  // namespace dm
#endif
