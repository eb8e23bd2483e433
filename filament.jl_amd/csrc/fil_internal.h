// fil_internal.h -- shared types of libfilament_hip.so (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <stdint.h>

#include "../../include/filament_hip.h"

namespace fil {

// Complex value with the memory layout of float2 / double2.
template <typename R>
struct __align__(2 * sizeof(R)) cx {
  R x, y;
};

template <typename R> __host__ __device__ __forceinline__ cx<R> mk(R a, R b) { cx<R> c; c.x = a; c.y = b; return c; }
template <typename R> __host__ __device__ __forceinline__ cx<R> operator+(cx<R> a, cx<R> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename R> __host__ __device__ __forceinline__ cx<R> operator-(cx<R> a, cx<R> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename R> __host__ __device__ __forceinline__ cx<R> operator*(cx<R> a, cx<R> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename R> __host__ __device__ __forceinline__ cx<R> operator*(R s, cx<R> a) { return mk(s * a.x, s * a.y); }
template <typename R> __host__ __device__ __forceinline__ cx<R> conj(cx<R> a) { return mk(a.x, -a.y); }
// multiply by +i or -i
template <typename R> __host__ __device__ __forceinline__ cx<R> mul_i(cx<R> a) { return mk(-a.y, a.x); }
template <typename R> __host__ __device__ __forceinline__ cx<R> mul_mi(cx<R> a) { return mk(a.y, -a.x); }
template <typename R> __host__ __device__ __forceinline__ R norm2(cx<R> a) { return a.x * a.x + a.y * a.y; }

// Kernel timing without marker packets: a launcher of a timed phase starts its
// kernel through fil_launch, which attaches the armed (start, stop) event pair
// to the dispatch itself (hipExtLaunchKernel) and disarms it.  Event records
// between kernels would each leave a gap of several microseconds in the stream.
struct LaunchEvents {
  hipEvent_t start = nullptr, stop = nullptr;
};
extern thread_local LaunchEvents fil_launch_events;
template <typename F, typename... Args>
inline void fil_launch(F kernel, const dim3& grid, const dim3& block, uint32_t smem, hipStream_t st, Args... args) {
  if (fil_launch_events.start) {
    const LaunchEvents ev = fil_launch_events;
    fil_launch_events = LaunchEvents{};
    hipExtLaunchKernelGGL(kernel, grid, block, smem, st, ev.start, ev.stop, 0, args...);
  } else {
    hipLaunchKernelGGL(kernel, grid, block, smem, st, args...);
  }
}

// Plasma-equation parameters passed by value to the row kernel
// (photoionization_avalanche.jl:1-116, right-hand sides :119-205), in the working precision.
template <typename R>
struct PlasmaParams {
  int ncomp, ereal, kdep, avalanche;
  R rho0;
  R frhont[FIL_MAXCOMP], Rava[FIL_MAXCOMP], K[FIL_MAXCOMP];
  // log2-domain rate tables, uniform in log2(x):
  //   W = exp2(ly[i] + slope[i] (log2 I - lx0 - i dlx)), clamped to the end nodes
  int ntab[FIL_MAXCOMP];
  R lx0[FIL_MAXCOMP], inv_dlx[FIL_MAXCOMP], dlx[FIL_MAXCOMP];
  R xmin[FIL_MAXCOMP];              // linear x of the first node
  const R* ly[FIL_MAXCOMP];         // log2 y nodes (device)
  const R* slope[FIL_MAXCOMP];      // per-interval log-log slope (device)
};

struct Handle;

}  // namespace fil

struct fil_handle;
