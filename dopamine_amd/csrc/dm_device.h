// Device execution policy: runs kernel-body functors as HIP kernels on gfx950.
#pragma once

#include "dm_common.h"

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <type_traits>
#include <vector>

namespace dm {

inline void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
#define DM_CHECK(x) ::dm::check((x), #x)

// Blocks are dealt round-robin to the 8 XCDs; remapping the block id gives each XCD a contiguous
// run of cells, so the neighbour gathers of adjacent cells hit the same L2.
__device__ inline int xcd_block(int nblk) {
  const int b = blockIdx.x;
  const int per = nblk >> 3;
  if (per == 0 || b >= (per << 3)) return b;
  return (b & 7) * per + (b >> 3);
}

template <int N, bool MAX>
__device__ inline Red<N> wave_reduce(Red<N> x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
      const double o = __shfl_down(x.v[k], off, 64);
      x.v[k] = MAX ? fmax(x.v[k], o) : x.v[k] + o;
    }
  }
  return x;
}

// Fixed-shape block reduction (wavefront shuffles, then the wavefront sums in order). Every
// thread returns the block total, and every block of any grid sums in the same order.
template <int N, bool MAX, int NT = BLOCK>
__device__ inline Red<N> block_reduce(Red<N> x) {
  __shared__ double sh[NT / 64][N];
  x = wave_reduce<N, MAX>(x);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < N; ++k) sh[w][k] = x.v[k];
  __syncthreads();
  Red<N> r;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double a = sh[0][k];
#pragma unroll
    for (int q = 1; q < NT / 64; ++q) a = MAX ? fmax(a, sh[q][k]) : a + sh[q][k];
    r.v[k] = a;
  }
  __syncthreads();
  return r;
}

template <class F>
__global__ void __launch_bounds__(BLOCK) k_for(int n, F f) {
  const int i = xcd_block(gridDim.x) * BLOCK + threadIdx.x;
  if (i < n) f(i);
}
// the same with a wavefronts-per-SIMD floor W for the register allocator (functors with a static
// kWaves: gather-latency-bound cell kernels whose registers would otherwise cap the occupancy)
template <class F, int W>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(W))) k_for_w(int n, F f) {
  const int i = xcd_block(gridDim.x) * BLOCK + threadIdx.x;
  if (i < n) f(i);
}
template <class F, class = void>
struct ForWaves {
  static constexpr int value = 0;
};
template <class F>
struct ForWaves<F, std::void_t<decltype(F::kWaves)>> {
  static constexpr int value = F::kWaves;
};

// Grid reductions are two launches: every block stores its partials (layout [quantity][block]),
// then one block sums them in block order and runs the finaliser. The kernel boundary orders the
// two, so no block needs a device-scope fence (on gfx950 that is an L2 write-back and invalidate
// per block). The result does not depend on block scheduling.
template <int N, bool MAX, class F>
__global__ void __launch_bounds__(BLOCK) k_partial(int n, F f, double* part, const int* gate) {
  if (gate && *gate) return;  // uniform across the grid
  const int b = xcd_block(gridDim.x);
  const int i = b * BLOCK + threadIdx.x;
  Red<N> v = red_identity<N, MAX>();
  if (i < n) v = f(i);
  v = block_reduce<N, MAX>(v);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < N; ++k) part[size_t(k) * gridDim.x + b] = v.v[k];
}

template <int N, bool MAX, class Fin>
__global__ void __launch_bounds__(BLOCK) k_final(int g, Fin fin, const double* part, const int* gate) {
  if (gate && *gate) return;
  Red<N> s = red_identity<N, MAX>();
  for (int q = threadIdx.x; q < g; q += BLOCK)
#pragma unroll
    for (int k = 0; k < N; ++k) s.v[k] = MAX ? fmax(s.v[k], part[size_t(k) * g + q]) : s.v[k] + part[size_t(k) * g + q];
  s = block_reduce<N, MAX>(s);
  if (threadIdx.x == 0) fin(s);
}

// Chained reduction passes (the CG's iterations): a pass first sums the previous pass's block
// partials -- every block redundantly, in the same order, so all blocks hold bit-identical totals
// without any cross-block synchronisation -- lets the functor derive its scalars (and stop the
// whole grid uniformly), then streams its cells (a contiguous chunk, or tiles interleaved with
// the other blocks of its XCD) and stores its own block partials for the next pass.
#ifndef DM_PBLOCK
#define DM_PBLOCK 256
#endif
#ifndef DM_PGRID_MAX
#define DM_PGRID_MAX 1024
#endif
constexpr int PBLOCK = DM_PBLOCK;
constexpr int PGRID_MAX = DM_PGRID_MAX;  // four blocks of four waves per CU


template <int NIN, int NOUT, class F>
__global__ void __launch_bounds__(PBLOCK) k_pass(int n, int chunk, F f, const double* in_part, int in_g,
                                                 double* out_part, int tile) {
  Red<NIN> tot = red_identity<NIN, false>();
  for (int q = threadIdx.x; q < in_g; q += PBLOCK)
#pragma unroll
    for (int k = 0; k < NIN; ++k) tot.v[k] = tot.v[k] + in_part[size_t(k) * in_g + q];
  tot = block_reduce<NIN, false, PBLOCK>(tot);
  const int b = xcd_block(gridDim.x);
  typename F::Scal sc;
  if (!f.setup(tot, b == 0 && threadIdx.x == 0, sc)) return;  // uniform across the grid
  Red<NOUT> acc = red_identity<NOUT, false>();
  const int lo = b * chunk, hi = min(n, lo + chunk);
  const int per = gridDim.x >> 3;
  if (tile > 0 && per > 0 && b < (per << 3)) {
    // the XCD's blocks sweep its run of cells together, tile by tile (block j takes tiles j,
    // j + per, ...), so the neighbours' records are still in that XCD's L2 when they are read
    const int x0 = (b / per) * per * chunk, x1 = min(n, x0 + per * chunk);
    for (int t0 = x0 + (b % per) * tile; t0 < x1; t0 += per * tile) {
      const int t1 = min(x1, t0 + tile);
      Red<NOUT> v;
      if constexpr (PassSlots<F>::value > 0) {
        v = f.run(t0, t1, int(threadIdx.x), PBLOCK, sc);
      } else {
        v = red_identity<NOUT, false>();
        for (int c = t0 + int(threadIdx.x); c < t1; c += PBLOCK) {
          const Red<NOUT> w = f(c, sc);
#pragma unroll
          for (int k = 0; k < NOUT; ++k) v.v[k] = v.v[k] + w.v[k];
        }
      }
#pragma unroll
      for (int k = 0; k < NOUT; ++k) acc.v[k] = acc.v[k] + v.v[k];
    }
  } else if constexpr (PassSlots<F>::value > 0) {
    acc = f.run(lo, hi, int(threadIdx.x), PBLOCK, sc);
  } else {
    for (int c = lo + int(threadIdx.x); c < hi; c += PBLOCK) {
      const Red<NOUT> v = f(c, sc);
#pragma unroll
      for (int k = 0; k < NOUT; ++k) acc.v[k] = acc.v[k] + v.v[k];
    }
  }
  acc = block_reduce<NOUT, false, PBLOCK>(acc);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NOUT; ++k) out_part[size_t(k) * gridDim.x + b] = acc.v[k];
}

template <class T>
struct ZeroF {
  T* p;
  DM_HD void operator()(int i) const { p[i] = T(0); }
};

// Device timing of selected kernel classes: a pair of HIP events on the launching stream around
// each launch, resolved whenever the stream has been synchronised (or the pool is full).
struct KernelTimer {
  static constexpr int POOL = 1024;  // event pairs
  bool on[K_NCLASS] = {};
  int64_t launches[K_NCLASS] = {};
  double total_ms[K_NCLASS] = {};
  hipEvent_t ev[2 * POOL] = {};
  int cls[POOL] = {};
  int used = 0;
  bool created = false;
  // Classes whose launches are kept only when committed: the CG's passes after convergence are
  // gated no-ops, so a solve commits just its first `iterations` launches.
  bool deferred[K_NCLASS] = {true, true, false, false, true, true};
  std::vector<float> pending[K_NCLASS];

  void create() {
    if (created) return;
    for (auto& e : ev) DM_CHECK(hipEventCreate(&e));
    created = true;
  }
  void destroy() {
    if (!created) return;
    for (auto& e : ev) (void)hipEventDestroy(e);
    created = false;
  }
  // all recorded events have completed (the stream was synchronised)
  void resolve() {
    for (int i = 0; i < used; ++i) {
      float ms = 0.0f;
      DM_CHECK(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
      if (deferred[cls[i]]) {
        pending[cls[i]].push_back(ms);
      } else {
        launches[cls[i]] += 1;
        total_ms[cls[i]] += double(ms);
      }
    }
    used = 0;
  }
  void commit(int c, int keep) {
    const int k = std::min<int>(keep, int(pending[c].size()));
    for (int i = 0; i < k; ++i) total_ms[c] += double(pending[c][i]);
    launches[c] += k;
    pending[c].clear();
  }
};

struct DeviceExec {
  static constexpr bool device = true;
  hipStream_t stream = nullptr;
  double* part = nullptr;    // reduction partials (8 per block, sized for the largest launch)
  int part_blocks = 0;
  KernelTimer* kt = nullptr;
  int open_ = -1;            // class of the launch being timed
  // chained passes: each XCD's blocks sweep its cells together in tiles of this many (k_pass);
  // 0: one contiguous chunk per block. DOPAMINE_PASS_TILE overrides (for tuning).
  int pass_tile = 256;

  void init(int max_n) {
    DM_CHECK(hipStreamCreate(&stream));  // default priority
    if (const char* e = std::getenv("DOPAMINE_PASS_TILE")) pass_tile = std::max(0, std::atoi(e)) / 64 * 64;
    part_blocks = nblocks(max_n > 0 ? max_n : 1);
    DM_CHECK(hipMalloc(&part, sizeof(double) * 8 * size_t(part_blocks)));
    kt = new KernelTimer();
  }
  void destroy() {
    if (kt) {
      kt->destroy();
      delete kt;
    }
    if (part) (void)hipFree(part);
    if (stream) (void)hipStreamDestroy(stream);
    if (part_side) (void)hipFree(part_side);
    if (side) (void)hipStreamDestroy(side);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    kt = nullptr;
    part = part_side = nullptr;
    stream = side = nullptr;
    ev_fork = ev_join = nullptr;
  }
  // A second stream (default priority, created on first use) for work independent of what the
  // main stream runs meanwhile: launches between side_begin() and side_end() go to it, after
  // everything queued on the main stream so far, with a partials buffer of their own; join()
  // makes the main stream wait for them.
  hipStream_t side = nullptr;
  double* part_side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  void side_begin() {
    if (!side) {
      DM_CHECK(hipStreamCreate(&side));  // default priority
      DM_CHECK(hipMalloc(&part_side, sizeof(double) * 8 * size_t(part_blocks)));
      DM_CHECK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
      DM_CHECK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    }
    DM_CHECK(hipEventRecord(ev_fork, stream));
    DM_CHECK(hipStreamWaitEvent(side, ev_fork, 0));
    std::swap(stream, side);
    std::swap(part, part_side);
  }
  void side_end() {
    std::swap(stream, side);
    std::swap(part, part_side);
    DM_CHECK(hipEventRecord(ev_join, side));
  }
  void join() { DM_CHECK(hipStreamWaitEvent(stream, ev_join, 0)); }
  void timing(int c, bool enable) {
    if (c < 0 || c >= K_NCLASS) throw std::runtime_error("unknown kernel class");
    if (enable) {
      kt->create();
      sync();
      kt->launches[c] = 0;
      kt->total_ms[c] = 0.0;
      kt->pending[c].clear();
    }
    kt->on[c] = enable;
  }
  // The next launch after tbegin(c) carries an event pair that the dispatch itself stamps at the
  // kernel's start and end (hipExtLaunchKernelGGL), so the duration excludes queueing gaps.
  void tbegin(int c) {
    if (!kt || !kt->on[c]) return;
    if (kt->used == KernelTimer::POOL) sync();
    open_ = c;
  }
  void tend(int c) {
    if (open_ == c) open_ = -1;  // nothing was launched
  }
  // keep the first `keep` timed launches of class c since the last commit (stream synchronised)
  void tcommit(int c, int keep) {
    if (kt && kt->on[c]) kt->commit(c, keep);
  }
  template <class K, class... A>
  void launch(K kernel, int grid, int block, const A&... args) {
    if (open_ >= 0) {
      const int i = kt->used++;
      kt->cls[i] = open_;
      open_ = -1;
      hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, stream, kt->ev[2 * i], kt->ev[2 * i + 1], 0, args...);
    } else {
      hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, stream, args...);
    }
  }
  template <class T>
  T* alloc(size_t n) {
    T* p = nullptr;
    DM_CHECK(hipMalloc(&p, (n ? n : 1) * sizeof(T)));
    DM_CHECK(hipMemsetAsync(p, 0, (n ? n : 1) * sizeof(T), stream));
    return p;
  }
  void release(void* p) { if (p) (void)hipFree(p); }
  template <class T>
  void upload(T* dst, const T* src, size_t n) {
    if (n) DM_CHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, stream));
    DM_CHECK(hipStreamSynchronize(stream));
  }
  template <class T>
  void download(T* dst, const T* src, size_t n) {
    if (n) DM_CHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, stream));
    sync();
  }
  template <class T>
  void copy(T* dst, const T* src, size_t n) {
    if (n) DM_CHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToDevice, stream));
  }
  // A fill kernel rather than hipMemsetAsync, so the step stays valid inside a captured graph.
  template <class T>
  void zero(T* dst, size_t n) {
    if (n) for_each(int(n), ZeroF<T>{dst});
  }
  template <class F>
  void for_each(int n, const F& f) {
    if (n <= 0) return;
    if constexpr (ForWaves<F>::value > 0)
      launch(k_for_w<F, ForWaves<F>::value>, nblocks(n), BLOCK, n, f);
    else
      launch(k_for<F>, nblocks(n), BLOCK, n, f);
  }
  template <int N, bool MAX, class F, class Fin>
  void reduce(int n, const F& f, const Fin& fin, const int* gate = nullptr) {
    static_assert(N <= 8, "partials hold 8 values per block");
    const int g = nblocks(n > 0 ? n : 1);
    if (g > part_blocks) throw std::runtime_error("reduction larger than the partials buffer");
    k_partial<N, MAX, F><<<g, BLOCK, 0, stream>>>(n, f, part, gate);
    k_final<N, MAX, Fin><<<1, BLOCK, 0, stream>>>(g, fin, part, gate);
  }
  // Geometry of the chained passes over n items: grid size and cells per block (a multiple of
  // the wavefront, so every wave's loads start aligned).
  static void pass_shape(int n, int* grid, int* chunk) {
    const int n1 = n > 0 ? n : 1;
    int g = std::min(PGRID_MAX, nblocks(n1, PBLOCK));
    int c = (n1 + g - 1) / g;
    c = (c + 63) / 64 * 64;
    *chunk = c;
    *grid = (n1 + c - 1) / c;
  }
  int pass_grid(int n) const {
    int g, c;
    pass_shape(n, &g, &c);
    return g;
  }
  // One chained pass (see k_pass): in_part holds in_g partials per input quantity, out_part
  // receives pass_grid(n) per output quantity.
  template <int NIN, int NOUT, class F>
  void pass(int n, const F& f, const double* in_part, int in_g, double* out_part) {
    int g, c;
    pass_shape(n, &g, &c);
    launch(k_pass<NIN, NOUT, F>, g, PBLOCK, n, c, f, in_part, in_g, out_part, pass_tile);
  }
  void sync() {
    DM_CHECK(hipStreamSynchronize(stream));
    if (kt && kt->used) kt->resolve();
  }
};

}  // namespace dm
