// Device execution policy: runs kernel-body functors as HIP kernels on gfx950.
#pragma once

#include "dm_common.h"

#include <hip/hip_runtime.h>

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

// Fixed-shape block reduction (wavefront shuffles, then the 4 wavefront sums in order).
template <int N, bool MAX>
__device__ inline Red<N> block_reduce(Red<N> x) {
  __shared__ double sh[BLOCK / 64][N];
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
    for (int q = 1; q < BLOCK / 64; ++q) a = MAX ? fmax(a, sh[q][k]) : a + sh[q][k];
    r.v[k] = a;
  }
  __syncthreads();
  return r;
}

// Each block writes one partial per quantity; the last block to take a ticket sums the partials
// in block order (agent-scope atomics keep them coherent across the XCDs' L2s) and returns true.
// Every addition happens in a fixed order, so the result does not depend on block scheduling.
template <int N, bool MAX>
__device__ inline bool grid_reduce(Red<N> x, double* partials, unsigned* ticket, Red<N>* total) {
  __shared__ bool last;
  __shared__ double fin[N];
  Red<N> b = block_reduce<N, MAX>(x);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < N; ++k)
      __hip_atomic_store(&partials[size_t(blockIdx.x) * N + k], b.v[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    const unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  Red<N> s = red_identity<N, MAX>();
  for (int q = threadIdx.x; q < int(gridDim.x); q += BLOCK)
#pragma unroll
    for (int k = 0; k < N; ++k) {
      const double pv = __hip_atomic_load(&partials[size_t(q) * N + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s.v[k] = MAX ? fmax(s.v[k], pv) : s.v[k] + pv;
    }
  s = block_reduce<N, MAX>(s);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < N; ++k) fin[k] = s.v[k];
    __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < N; ++k) total->v[k] = fin[k];
  return true;
}

template <class F>
__global__ void __launch_bounds__(BLOCK) k_for(int n, F f) {
  const int i = xcd_block(gridDim.x) * BLOCK + threadIdx.x;
  if (i < n) f(i);
}

template <int N, bool MAX, class F, class Fin>
__global__ void __launch_bounds__(BLOCK) k_reduce(int n, F f, Fin fin, double* part, unsigned* ticket, const int* gate) {
  if (gate && *gate) return;  // uniform across the grid
  const int i = xcd_block(gridDim.x) * BLOCK + threadIdx.x;
  Red<N> v = red_identity<N, MAX>();
  if (i < n) v = f(i);
  Red<N> t;
  if (grid_reduce<N, MAX>(v, part, ticket, &t) && threadIdx.x == 0) fin(t);
}

template <class T>
struct ZeroF {
  T* p;
  DM_HD void operator()(int i) const { p[i] = T(0); }
};

struct DeviceExec {
  static constexpr bool device = true;
  hipStream_t stream = nullptr;
  double* part = nullptr;    // reduction partials (4 per block, sized for the largest launch)
  unsigned* ticket = nullptr;
  int part_blocks = 0;

  void init(int max_n) {
    DM_CHECK(hipStreamCreate(&stream));  // default priority
    part_blocks = nblocks(max_n > 0 ? max_n : 1);
    DM_CHECK(hipMalloc(&part, sizeof(double) * 8 * size_t(part_blocks)));
    DM_CHECK(hipMalloc(&ticket, sizeof(unsigned)));
    DM_CHECK(hipMemsetAsync(ticket, 0, sizeof(unsigned), stream));
  }
  void destroy() {
    if (part) (void)hipFree(part);
    if (ticket) (void)hipFree(ticket);
    if (stream) (void)hipStreamDestroy(stream);
    part = nullptr;
    ticket = nullptr;
    stream = nullptr;
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
    DM_CHECK(hipStreamSynchronize(stream));
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
    k_for<F><<<nblocks(n), BLOCK, 0, stream>>>(n, f);
  }
  template <int N, bool MAX, class F, class Fin>
  void reduce(int n, const F& f, const Fin& fin, const int* gate = nullptr) {
    static_assert(N <= 8, "partials hold 8 values per block");
    const int g = nblocks(n > 0 ? n : 1);
    if (g > part_blocks) throw std::runtime_error("reduction larger than the partials buffer");
    k_reduce<N, MAX, F, Fin><<<g, BLOCK, 0, stream>>>(n, f, fin, part, ticket, gate);
  }
  void sync() { DM_CHECK(hipStreamSynchronize(stream)); }
};

}  // namespace dm
