// Shared definitions for the DOPAMINE hot-path kernels: constants, the reduction value type and
// the execution policies that run kernel bodies.
//
// Kernel bodies are functors with a __host__ __device__ call operator. The product library runs
// them through DeviceExec (HIP launches on gfx950). HostExec, compiled only into the test-side
// host-check library, runs the same bodies in a loop so their indexing can be checked on a
// machine without a GPU; it is never part of libdopamine_hip.so.
#pragma once

#include <cmath>
#include <cstdint>
#include <type_traits>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DM_HD __host__ __device__
#define DM_INLINE __host__ __device__ inline
#else
#define DM_HD
#define DM_INLINE inline
#endif

namespace dm {

#if !defined(__HIPCC__)
using std::copysign;
using std::exp;
using std::fabs;
using std::fmax;
using std::fmin;
using std::isfinite;
using std::log;
using std::pow;
using std::sqrt;
#endif

constexpr double SMALL = 1.0e-30;  // src/mod_parameters.f90:25
constexpr int BLOCK = 256;         // 4 wavefronts of 64 lanes

enum Bc : int {
  BC_INLET = 1, BC_OUTLET = 2, BC_WALL_NOSLIP = 3, BC_WALL_MODEL = 4, BC_SLIP = 5, BC_PERIODIC = 6,
  BC_SYMMETRY = 7, BC_FIXEDTEMP = 8, BC_ZEROGRADIENT = 9, BC_WALL_VELOCITY = 10, BC_ABL_INLET = 11
};

// Walls and slip planes carry no mass flux (mod_advection.f90:1154-1156).
DM_INLINE bool is_solid(int bc) {
  return bc == BC_WALL_NOSLIP || bc == BC_WALL_MODEL || bc == BC_WALL_VELOCITY || bc == BC_SLIP || bc == BC_SYMMETRY;
}

template <int N>
struct Red {
  double v[N > 0 ? N : 1];
};

template <int N, bool MAX>
DM_INLINE Red<N> red_identity() {
  Red<N> r;
  for (int k = 0; k < N; ++k) r.v[k] = MAX ? -1.0e308 : 0.0;
  return r;
}

inline int nblocks(int64_t n, int b = BLOCK) { return int((n + b - 1) / b); }

// A pass functor with a positive static kSlots streams its chunk itself (run(lo, hi, thread,
// threads, scalars) returns the thread's partials, accumulated in the same cell order as the
// generic loop); the others are called cell by cell.
template <class F, class = void>
struct PassSlots {
  static constexpr int value = 0;
};
template <class F>
struct PassSlots<F, std::void_t<decltype(F::kSlots)>> {
  static constexpr int value = F::kSlots;
};

// ------------------------------------------------------------------ host execution (tests only)
struct HostExec {
  static constexpr bool device = false;
  void init(int) {}
  void destroy() {}
  template <class T>
  T* alloc(size_t n) {
    T* p = static_cast<T*>(std::calloc(n ? n : 1, sizeof(T)));
    if (!p) throw std::runtime_error("host allocation failed");
    return p;
  }
  void release(void* p) { std::free(p); }
  template <class T>
  void upload(T* dst, const T* src, size_t n) { if (n) std::memcpy(dst, src, n * sizeof(T)); }
  template <class T>
  void download(T* dst, const T* src, size_t n) { if (n) std::memcpy(dst, src, n * sizeof(T)); }
  template <class T>
  void copy(T* dst, const T* src, size_t n) { if (n) std::memmove(dst, src, n * sizeof(T)); }
  template <class T>
  void zero(T* dst, size_t n) { if (n) std::memset(dst, 0, n * sizeof(T)); }
  template <class F>
  void for_each(int n, const F& f) {
    for (int i = 0; i < n; ++i) f(i);
  }
  template <int N, bool MAX, class F, class Fin>
  void reduce(int n, const F& f, const Fin& fin, const int* gate = nullptr) {
    if (gate && *gate) return;
    Red<N> s = red_identity<N, MAX>();
    for (int i = 0; i < n; ++i) {
      const Red<N> x = f(i);
      for (int k = 0; k < N; ++k) s.v[k] = MAX ? std::fmax(s.v[k], x.v[k]) : s.v[k] + x.v[k];
    }
    fin(s);
  }
  int pass_grid(int) const { return 1; }
  // chained pass with one block (see DeviceExec::pass)
  template <int NIN, int NOUT, class F>
  void pass(int n, const F& f, const double* in_part, int in_g, double* out_part) {
    Red<NIN> tot = red_identity<NIN, false>();
    for (int q = 0; q < in_g; ++q)
      for (int k = 0; k < NIN; ++k) tot.v[k] = tot.v[k] + in_part[size_t(k) * in_g + q];
    typename F::Scal sc;
    if (!f.setup(tot, true, sc)) return;
    Red<NOUT> acc = red_identity<NOUT, false>();
    if constexpr (PassSlots<F>::value > 0) {  // the batched form the device runs, one thread
      acc = f.run(0, n, 0, 1, sc);
    } else {
      for (int c = 0; c < n; ++c) {
        const Red<NOUT> v = f(c, sc);
        for (int k = 0; k < NOUT; ++k) acc.v[k] = acc.v[k] + v.v[k];
      }
    }
    for (int k = 0; k < NOUT; ++k) out_part[k] = acc.v[k];
  }
  void sync() {}
  // one sequence of work on the host: the second stream's work simply runs in order
  void side_begin() {}
  void side_end() {}
  void join() {}
  // kernel timing is a device feature; on the host these are no-ops
  void tbegin(int) {}
  void tend(int) {}
  void tcommit(int, int) {}
};

// Kernel classes with device timing (dopamine_solver_kernel_timing).
enum KernelClass {
  K_CG_DIR = 0, K_CG_UPD = 1, K_EXPLICIT = 2, K_LS_GRAD = 3, K_MG_DOWN = 4, K_MG_UP = 5, K_NCLASS = 6
};

}  // namespace dm
