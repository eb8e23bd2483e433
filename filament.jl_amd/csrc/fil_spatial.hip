// fil_spatial.hip -- the spatial geometries R and XY (a monochromatic beam,
// no time axis, no plasma) on gfx950.
//
// Models.zstep for R and XY (Models.jl:75-121 without the plasma and the t
// transforms):
//   nonlinear RK step  dE/dz = func!(E)   (nonlinear_propagators.jl:38-74)
//     paraxial:      i QZ Rnl g 3/4 |E|^2 E           (cubic.jl:49-54; nonlinear_propagators.jl:56-57)
//     non-paraxial:  T^-1[ i QZ Kg T[ Rnl g 3/4 |E|^2 E ] ]  (nonlinear_propagators.jl:58-63)
//   linear step        E <- T^-1[ exp(i KZ dz) Kg T[E] ]  (spectral.jl:16-26)
//   field filter       E <- g E                          (Guards.jl:199-215)
// with g the field guard (R: Rguard(r); XY: Xguard(x) Yguard(y)), Kg the
// spectral guard, T the space transform: the quasi-discrete Hankel transform
// for R (HankelTransforms plan, Fields.jl:38-42), the 2-D FFT for XY
// (plan_fft!, Fields.jl:131).
//
// R: the Hankel pair folds like the RT one, T^-1 diag(P) T =
// diag(1/RJ) Tm diag(P) Tm diag(RJ) = M2 diag(P) M1, and each transform is a
// matrix-vector product (one wave per output row; Nr^2 matrix reads per
// product, the bound is HBM).  The paraxial nonlinear step runs first as an
// elementwise kernel; the field filter is the second product's epilogue.
// XY: the state E[ix + Nx iy] (the reference's column-major layout) takes
// three passes per step: x lines (paraxial nonlinear step fused in front of
// the forward FFT), y columns (forward FFT, times P, inverse FFT; WB adjacent
// columns per workgroup so every row read is a full 64-byte segment), x lines
// (inverse FFT, field filter).  1/(Nx Ny) is folded into P.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "fil_fft.h"
#include "fil_host.h"
#include "fil_internal.h"
#include "fil_kernels.h"

using namespace fil;

namespace {

// ---------------------------------------------------------------------------
// Nonlinear right-hand side of one point (func!, nonlinear_propagators.jl:38-67):
//   F = 3/4 |U|^2 U         calc_cubic_nothg_spatial (cubic.jl:49-54)
//   F = g F                 apply_field_filter! (:50)
//   dE = Rnl F              update_dE! (:53), one response
//   dE = i QZ dE            paraxial (:55-57)
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ cx<R> sp_response(cx<R> u, R g, R rnl) {
  const cx<R> F = ((R)0.75 * norm2(u)) * u;
  return rnl * (g * F);
}
template <typename R>
__device__ __forceinline__ cx<R> sp_rhs_par(cx<R> u, R g, R rnl, cx<R> qz) {
  return mul_i(qz) * sp_response(u, g, rnl);
}

// One RK step of the paraxial right-hand side at one point: Kutta's third-order
// method (ALG 3) or the classic fourth-order one (ALG 4), the ODEIntegrators
// algorithms RK3 and RK4 (nonlinear_propagators.jl:29-32,70-74).
template <typename R, int ALG>
__device__ __forceinline__ cx<R> sp_rk_point(cx<R> e, R g, R rnl, cx<R> qz, R h) {
  const cx<R> k1 = sp_rhs_par(e, g, rnl, qz);
  const cx<R> k2 = sp_rhs_par(e + ((R)0.5 * h) * k1, g, rnl, qz);
  if constexpr (ALG == 3) {
    const cx<R> k3 = sp_rhs_par(e + h * (k2 + k2 - k1), g, rnl, qz);
    const R c16 = (R)1 / (R)6, c23 = (R)2 / (R)3;
    return e + h * (c16 * k1 + c23 * k2 + c16 * k3);
  } else {
    const cx<R> k3 = sp_rhs_par(e + ((R)0.5 * h) * k2, g, rnl, qz);
    const cx<R> k4 = sp_rhs_par(e + h * k3, g, rnl, qz);
    return e + (h / (R)6) * (k1 + k2 + k2 + k3 + k3 + k4);
  }
}

// field guard of point q: R Rg[q]; XY Xg[ix] Yg[iy]
template <typename R>
__device__ __forceinline__ R sp_guard(const R* Rg, const R* Xg, const R* Yg, int Nx, size_t q) {
  return Rg ? Rg[q] : Xg[q % Nx] * Yg[q / Nx];
}

template <typename R, int ALG>
__global__ void sp_nl_kernel(cx<R>* __restrict__ E, size_t n, const R* __restrict__ Rg, const R* __restrict__ Xg,
                             const R* __restrict__ Yg, int Nx, R rnl, const cx<R>* __restrict__ QZ, R h) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  E[q] = sp_rk_point<R, ALG>(E[q], sp_guard(Rg, Xg, Yg, Nx, q), rnl, QZ[q], h);
}

// Non-paraxial step: stage s (1-based) of the RK writes the untransformed
// right-hand side Rnl g 3/4 |U|^2 U of its stage field U into K[s-1];
// sp_np_combine_kernel adds the weighted stages after their transforms.
template <typename R, int ALG>
__global__ void sp_np_stage_kernel(const cx<R>* __restrict__ E, cx<R>* K0, cx<R>* K1, cx<R>* K2, cx<R>* K3, int s,
                                   size_t n, const R* __restrict__ Rg, const R* __restrict__ Xg,
                                   const R* __restrict__ Yg, int Nx, R rnl, R h) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  cx<R> u = E[q];
  if (s == 2) u = u + ((R)0.5 * h) * K0[q];
  if (s == 3) u = (ALG == 3) ? u + h * (K1[q] + K1[q] - K0[q]) : u + ((R)0.5 * h) * K1[q];
  if (s == 4) u = u + h * K2[q];
  cx<R>* out = s == 1 ? K0 : (s == 2 ? K1 : (s == 3 ? K2 : K3));
  out[q] = sp_response(u, sp_guard(Rg, Xg, Yg, Nx, q), rnl);
}
template <typename R, int ALG>
__global__ void sp_np_combine_kernel(cx<R>* __restrict__ E, const cx<R>* __restrict__ K0, const cx<R>* __restrict__ K1,
                                     const cx<R>* __restrict__ K2, const cx<R>* __restrict__ K3, size_t n, R h) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  if constexpr (ALG == 3) {
    const R c16 = (R)1 / (R)6, c23 = (R)2 / (R)3;
    E[q] = E[q] + h * (c16 * K0[q] + c23 * K1[q] + c16 * K2[q]);
  } else {
    E[q] = E[q] + (h / (R)6) * (K0[q] + K1[q] + K1[q] + K2[q] + K2[q] + K3[q]);
  }
}

// P = exp(i KZ dz) Sg norm in FP64, rounded to R; Sg = Kg[q] (R) or
// KX[ix] KY[iy] (XY).
template <typename R>
__global__ void sp_prop_kernel(const double2* __restrict__ KZ, size_t n, const double* __restrict__ Kg,
                               const double* __restrict__ KX, const double* __restrict__ KY, int Nx, double norm,
                               double dz, cx<R>* __restrict__ P) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const double2 kz = KZ[q];
  const double g = (Kg ? Kg[q] : KX[q % Nx] * KY[q / Nx]) * norm;
  const double amp = exp(-kz.y * dz) * g;
  double s, c;
  sincos(kz.x * dz, &s, &c);
  P[q] = mk((R)(amp * c), (R)(amp * s));
}

// out[k] = F[k] (P[k] sum_j M[k][j] in[j]); one wave per row k, lanes over j.
// P and F may be null.
template <typename R>
__global__ __launch_bounds__(256) void sp_matvec_kernel(const R* __restrict__ M, const cx<R>* __restrict__ in,
                                                        cx<R>* __restrict__ out, int n, const cx<R>* __restrict__ P,
                                                        const R* __restrict__ F) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= n) return;
  const R* m = M + (size_t)row * n;
  R ar = 0, ai = 0;
  for (int j = lane; j < n; j += 64) {
    const R a = m[j];
    const cx<R> v = in[j];
    ar += a * v.x;
    ai += a * v.y;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    ar += __shfl_xor(ar, off, 64);
    ai += __shfl_xor(ai, off, 64);
  }
  if (lane == 0) {
    cx<R> v = mk(ar, ai);
    if (P) v = v * P[row];
    if (F) v = F[row] * v;
    out[row] = v;
  }
}

// ---------------------------------------------------------------------------
// XY transforms
// ---------------------------------------------------------------------------
template <int N> constexpr int sp_line_threads() { return N / 8 < 64 ? 64 : (N / 8 > 1024 ? 1024 : N / 8); }

// One x line (iy = blockIdx.x, N = Nx contiguous points) per workgroup.
// FWD: (optional paraxial RK step of every point, then) forward FFT (exp(-i));
// else inverse FFT (exp(+i), unnormalised), then (optional) field filter.
template <typename R, int N, bool FWD, int ALG>
__global__ __launch_bounds__(sp_line_threads<N>()) void sp_xline_kernel(cx<R>* __restrict__ E,
                                                                        const cx<R>* __restrict__ tw,
                                                                        const R* __restrict__ Xg,
                                                                        const R* __restrict__ Yg, R rnl,
                                                                        const cx<R>* __restrict__ QZ, R h) {
  constexpr int NTH = sp_line_threads<N>();
  extern __shared__ __align__(16) unsigned char smem[];
  cx<R>* Z = reinterpret_cast<cx<R>*>(smem);
  const int tid = threadIdx.x;
  const size_t base = (size_t)blockIdx.x * N;
  {
    // all of the thread's loads in flight before its first use
    constexpr int PER = (N + NTH - 1) / NTH;
    constexpr bool FULL = N % NTH == 0;      // else lines shorter than the workgroup
    const cx<R> zero = mk((R)0, (R)0);
    cx<R> v[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int t = tid + u * NTH;
      v[u] = (FULL || t < N) ? E[base + t] : zero;
    }
    if constexpr (FWD && ALG > 0) {
      cx<R> qz[PER];
#pragma unroll
      for (int u = 0; u < PER; u++) {
        const int t = tid + u * NTH;
        qz[u] = (FULL || t < N) ? QZ[base + t] : zero;
      }
#pragma unroll
      for (int u = 0; u < PER; u++) {
        const int t = tid + u * NTH;
        if (FULL || t < N) v[u] = sp_rk_point<R, ALG>(v[u], Xg[t] * Yg[blockIdx.x], rnl, qz[u], h);
      }
    }
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int t = tid + u * NTH;
      if (FULL || t < N) Z[lds_index(t)] = v[u];
    }
  }
  __syncthreads();
  fft_lds<R, N, NTH, FWD ? -1 : +1>(Z, tw, tid);
  for (int t = tid; t < N; t += NTH) {
    cx<R> v = Z[lds_index(t)];
    if constexpr (!FWD) {
      if (Xg) v = Yg[blockIdx.x] * (Xg[t] * v);   // kernel! F[i, j] * A[i] * B[j] (Guards.jl:263-278)
    }
    E[base + t] = v;
  }
}

// WB adjacent y columns per workgroup (ix0 = blockIdx.x WB): forward FFT along
// y, times P, inverse FFT along y.  Column c's N points sit at Z[c CS + ...].
template <typename R, int N> constexpr int sp_col_width() {
  int wb = 64 / (int)sizeof(cx<R>);
  while (wb > 1 && (size_t)wb * (lds_size(N) + 1) * sizeof(cx<R>) > 147456) wb /= 2;
  return wb;
}
template <typename R, int N> constexpr int sp_col_threads() {
  return N / 8 < 1 ? 1 : (N / 8 > 1024 / sp_col_width<R, N>() ? 1024 / sp_col_width<R, N>() : N / 8);
}
template <typename R, int N>
__global__ __launch_bounds__((sp_col_width<R, N>() * sp_col_threads<R, N>())) void sp_ycol_kernel(
    cx<R>* __restrict__ E, const cx<R>* __restrict__ tw, const cx<R>* __restrict__ P, int Nx) {
  constexpr int WB = sp_col_width<R, N>();
  constexpr int NTHC = sp_col_threads<R, N>();
  constexpr int NT = WB * NTHC;
  constexpr int CS = lds_size(N) + 1;   // +1: the columns start in different banks
  extern __shared__ __align__(16) unsigned char smem[];
  cx<R>* Z = reinterpret_cast<cx<R>*>(smem);
  const int tid = threadIdx.x, col = tid / NTHC, ctid = tid % NTHC;
  const size_t ix0 = (size_t)blockIdx.x * WB;
  // Each thread moves PER elements of one column (NT is a multiple of WB);
  // each batch of loads is in flight together before its first use.
  constexpr int PER = N * WB / NT;
  static_assert(PER * NT == N * WB, "whole elements per thread");
  constexpr int EST = NT / WB;
  const int c = tid % WB, e0 = tid / WB;
  {
    cx<R> v[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) v[k] = E[(size_t)(e0 + k * EST) * Nx + ix0 + c];
#pragma unroll
    for (int k = 0; k < PER; k++) Z[c * CS + lds_index(e0 + k * EST)] = v[k];
  }
  __syncthreads();
  cx<R>* Zc = Z + col * CS;
  fft_lds<R, N, NTHC, -1>(Zc, tw, ctid);
  {
    cx<R> p[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) p[k] = P[(size_t)(e0 + k * EST) * Nx + ix0 + c];
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int e = e0 + k * EST;
      Z[c * CS + lds_index(e)] = Z[c * CS + lds_index(e)] * p[k];
    }
  }
  __syncthreads();
  fft_lds<R, N, NTHC, +1>(Zc, tw, ctid);
  {
    // the indices pass through an empty asm so that the store addresses are
    // formed here rather than kept live across the transforms
    int cb = c, eb = e0;
    asm volatile("" : "+v"(cb), "+v"(eb));
    cx<R> v[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) v[k] = Z[cb * CS + lds_index(eb + k * EST)];
#pragma unroll
    for (int k = 0; k < PER; k++) E[(size_t)(eb + k * EST) * Nx + ix0 + cb] = v[k];
  }
}

#define FIL_DISPATCH_SP(N_, ...)                                  \
  switch (N_) {                                                   \
    case 16: { constexpr int NN = 16; __VA_ARGS__; } break;       \
    case 32: { constexpr int NN = 32; __VA_ARGS__; } break;       \
    case 64: { constexpr int NN = 64; __VA_ARGS__; } break;       \
    case 128: { constexpr int NN = 128; __VA_ARGS__; } break;     \
    case 256: { constexpr int NN = 256; __VA_ARGS__; } break;     \
    case 512: { constexpr int NN = 512; __VA_ARGS__; } break;     \
    case 1024: { constexpr int NN = 1024; __VA_ARGS__; } break;   \
    case 2048: { constexpr int NN = 2048; __VA_ARGS__; } break;   \
    case 4096: { constexpr int NN = 4096; __VA_ARGS__; } break;   \
    default: return hipErrorInvalidValue;                         \
  }

template <typename R, int N, bool FWD, int ALG>
hipError_t launch_xline_n(cx<R>* E, const cx<R>* tw, const R* Xg, const R* Yg, R rnl, const cx<R>* QZ, R h, int Ny,
                          hipStream_t st) {
  const size_t sm = (size_t)lds_size(N) * sizeof(cx<R>);
  auto k = sp_xline_kernel<R, N, FWD, ALG>;
  hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(Ny), dim3(sp_line_threads<N>()), sm, st, E, tw, Xg, Yg, rnl, QZ, h);
  return hipGetLastError();
}
template <typename R, int FWD_ALG>
hipError_t launch_xline(bool fwd, cx<R>* E, const cx<R>* tw, const R* Xg, const R* Yg, R rnl, const cx<R>* QZ, R h,
                        int Nx, int Ny, hipStream_t st) {
  hipError_t e;
  if (fwd) {
    FIL_DISPATCH_SP(Nx, e = (launch_xline_n<R, NN, true, FWD_ALG>(E, tw, Xg, Yg, rnl, QZ, h, Ny, st)));
  } else {
    FIL_DISPATCH_SP(Nx, e = (launch_xline_n<R, NN, false, 0>(E, tw, Xg, Yg, rnl, QZ, h, Ny, st)));
  }
  return e;
}
template <typename R, int N>
hipError_t launch_ycol_n(cx<R>* E, const cx<R>* tw, const cx<R>* P, int Nx, hipStream_t st) {
  constexpr int WB = sp_col_width<R, N>();
  if (Nx % WB) return hipErrorInvalidValue;
  const size_t sm = (size_t)WB * (lds_size(N) + 1) * sizeof(cx<R>);
  auto k = sp_ycol_kernel<R, N>;
  hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3(Nx / WB), dim3(WB * sp_col_threads<R, N>()), sm, st, E, tw, P, Nx);
  return hipGetLastError();
}
template <typename R>
hipError_t launch_ycol(cx<R>* E, const cx<R>* tw, const cx<R>* P, int Nx, int Ny, hipStream_t st) {
  hipError_t e;
  FIL_DISPATCH_SP(Ny, e = (launch_ycol_n<R, NN>(E, tw, P, Nx, st)));
  return e;
}

// ---------------------------------------------------------------------------
// Analysis: I = |E|^2 (working precision), per block the maximum with its
// first index, a NaN flag and sum(I w) in FP64 (w = r dr for R, 1 for XY).
// ---------------------------------------------------------------------------
constexpr int SP_STAT_BLOCKS = 256;
template <typename R>
__global__ __launch_bounds__(256) void sp_stats_kernel(const cx<R>* __restrict__ E, size_t n,
                                                       const double* __restrict__ wgt, R* __restrict__ I,
                                                       double* __restrict__ pmax, long long* __restrict__ pidx,
                                                       double* __restrict__ psum, int* __restrict__ pnan) {
  __shared__ R smax[256];
  __shared__ long long sidx[256];
  __shared__ double ssum[256];
  __shared__ int snan[256];
  const int tid = threadIdx.x;
  R best = -std::numeric_limits<R>::infinity();
  long long bi = -1;
  double sum = 0;
  int nan = 0;
  for (size_t q = (size_t)blockIdx.x * 256 + tid; q < n; q += (size_t)gridDim.x * 256) {
    const R v = norm2(E[q]);
    I[q] = v;
    if (v != v) nan = 1;
    if (v > best) { best = v; bi = (long long)q; }   // q increases: the first maximum stays
    sum += wgt ? (double)v * wgt[q] : (double)v;
  }
  smax[tid] = best; sidx[tid] = bi; ssum[tid] = sum; snan[tid] = nan;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) {
      const R o = smax[tid + s];
      const long long oi = sidx[tid + s];
      if (o > smax[tid] || (o == smax[tid] && oi >= 0 && (sidx[tid] < 0 || oi < sidx[tid]))) {
        smax[tid] = o; sidx[tid] = oi;
      }
      ssum[tid] += ssum[tid + s];
      snan[tid] |= snan[tid + s];
    }
    __syncthreads();
  }
  if (tid == 0) {
    pmax[blockIdx.x] = (double)smax[0];
    pidx[blockIdx.x] = sidx[0];
    psum[blockIdx.x] = ssum[0];
    pnan[blockIdx.x] = snan[0];
  }
}

template <typename R>
unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

// ---------------------------------------------------------------------------
// Handle
// ---------------------------------------------------------------------------
struct fil_spatial {
  fil_spatial_desc d{};
  bool f64 = false, xy = false;
  int n = 0;
  hipStream_t st = nullptr;
  DevBuf E, W, M1, M2, KZ, Kg, KX, KY, Rg, Xg, Yg, QZ, P, Pnl, K[4], twx, twy;
  DevBuf I, wgt, pmax, pidx, psum, pnan;
  double rnl = 0, P_dz = std::numeric_limits<double>::quiet_NaN();
  bool tables = false, field = false;
  std::vector<double> r, x, y;   // host grids for the analyzer's radii
  double dx = 0, dy = 0;
};

namespace {

int sp_set_device(const fil_spatial* h) {
  HIPCHK(hipSetDevice(h->d.device));
  return 0;
}

// Julia's radius (FieldAnalyzers.jl:357-380) on samples y of working precision
// R: level = max(y) exp(-1) in R, the first and the last node at or above it.
template <typename R>
double radius_r(const double* x, const R* y, int n) {
  R ymax = -std::numeric_limits<R>::infinity();
  for (int i = 0; i < n; i++) if (y[i] > ymax || y[i] != y[i]) ymax = y[i];
  const R lev = ymax * (R)std::exp(-1.0);
  double radl = 0, radr = 0;
  for (int i = 0; i < n; i++) if (y[i] >= lev) { radl = x[i]; break; }
  for (int i = n - 1; i >= 0; i--) if (y[i] >= lev) { radr = x[i]; break; }
  return (std::fabs(radl) + std::fabs(radr)) / 2;
}

template <typename R>
int sp_linear(fil_spatial* h, double dz) {
  const size_t n = h->n;
  if (!(dz == h->P_dz)) {
    const double norm = h->xy ? 1.0 / ((double)h->d.Nx * h->d.Ny) : 1.0;
    hipLaunchKernelGGL(sp_prop_kernel<R>, dim3(blocks_for<R>(n)), dim3(256), 0, h->st, h->KZ.as<double2>(), n,
                       h->xy ? nullptr : h->Kg.as<double>(), h->KX.as<double>(), h->KY.as<double>(), h->d.Nx, norm,
                       dz, h->P.as<cx<R>>());
    HIPCHK(hipGetLastError());
    h->P_dz = dz;
  }
  return 0;
}

// T^-1[ Pm .* T[A] ] in place (R: through W; XY: the three passes), with the
// field filter at the end when filter.
template <typename R>
int sp_transform(fil_spatial* h, cx<R>* A, const cx<R>* Pm, bool filter, bool nl_fused, R h_dz) {
  if (!h->xy) {
    const int n = h->n;
    hipLaunchKernelGGL(sp_matvec_kernel<R>, dim3((n + 3) / 4), dim3(256), 0, h->st, h->M1.as<R>(), A, h->W.as<cx<R>>(),
                       n, Pm, (const R*)nullptr);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(sp_matvec_kernel<R>, dim3((n + 3) / 4), dim3(256), 0, h->st, h->M2.as<R>(), h->W.as<cx<R>>(), A,
                       n, (const cx<R>*)nullptr, filter ? h->Rg.as<R>() : (const R*)nullptr);
    HIPCHK(hipGetLastError());
    return 0;
  }
  const int Nx = h->d.Nx, Ny = h->d.Ny;
  const R rnl = (R)h->rnl;
  const R* Xg = h->Xg.as<R>();
  const R* Yg = h->Yg.as<R>();
  const cx<R>* QZ = h->QZ.as<cx<R>>();
  if (nl_fused && h->d.alg == FIL_ALG_RK4) {
    HIPCHK((launch_xline<R, 4>(true, A, h->twx.as<cx<R>>(), Xg, Yg, rnl, QZ, h_dz, Nx, Ny, h->st)));
  } else if (nl_fused) {
    HIPCHK((launch_xline<R, 3>(true, A, h->twx.as<cx<R>>(), Xg, Yg, rnl, QZ, h_dz, Nx, Ny, h->st)));
  } else {
    HIPCHK((launch_xline<R, 0>(true, A, h->twx.as<cx<R>>(), Xg, Yg, rnl, QZ, h_dz, Nx, Ny, h->st)));
  }
  HIPCHK(launch_ycol<R>(A, h->twy.as<cx<R>>(), Pm, Nx, Ny, h->st));
  HIPCHK((launch_xline<R, 0>(false, A, h->twx.as<cx<R>>(), filter ? Xg : nullptr, filter ? Yg : nullptr, rnl, QZ, h_dz,
                             Nx, Ny, h->st)));
  return 0;
}

template <typename R, int ALG>
int sp_np_step(fil_spatial* h, R hdz) {
  const size_t n = h->n;
  const R* Rg = h->xy ? nullptr : h->Rg.as<R>();
  const R* Xg = h->Xg.as<R>();
  const R* Yg = h->Yg.as<R>();
  cx<R>* K[4];
  for (int j = 0; j < 4; j++) K[j] = h->K[j].as<cx<R>>();
  for (int s = 1; s <= ALG; s++) {
    hipLaunchKernelGGL((sp_np_stage_kernel<R, ALG>), dim3(blocks_for<R>(n)), dim3(256), 0, h->st, h->E.as<cx<R>>(),
                       K[0], K[1], K[2], K[3], s, n, Rg, Xg, Yg, h->d.Nx, (R)h->rnl, hdz);
    HIPCHK(hipGetLastError());
    if (int rc = sp_transform<R>(h, K[s - 1], h->Pnl.as<cx<R>>(), false, false, hdz)) return rc;
  }
  hipLaunchKernelGGL((sp_np_combine_kernel<R, ALG>), dim3(blocks_for<R>(n)), dim3(256), 0, h->st, h->E.as<cx<R>>(),
                     K[0], K[1], K[2], K[3], n, hdz);
  HIPCHK(hipGetLastError());
  return 0;
}

template <typename R>
int sp_zstep(fil_spatial* h, double dz) {
  const R hdz = (R)dz;
  bool fused = false;
  if (h->d.nonlinearity) {
    if (h->d.nonparaxial) {
      int rc = h->d.alg == FIL_ALG_RK4 ? sp_np_step<R, 4>(h, hdz) : sp_np_step<R, 3>(h, hdz);
      if (rc) return rc;
    } else if (h->xy) {
      fused = true;   // in front of the forward x transform
    } else {
      const size_t n = h->n;
      if (h->d.alg == FIL_ALG_RK4)
        hipLaunchKernelGGL((sp_nl_kernel<R, 4>), dim3(blocks_for<R>(n)), dim3(256), 0, h->st, h->E.as<cx<R>>(), n,
                           h->Rg.as<R>(), (const R*)nullptr, (const R*)nullptr, 0, (R)h->rnl, h->QZ.as<cx<R>>(), hdz);
      else
        hipLaunchKernelGGL((sp_nl_kernel<R, 3>), dim3(blocks_for<R>(n)), dim3(256), 0, h->st, h->E.as<cx<R>>(), n,
                           h->Rg.as<R>(), (const R*)nullptr, (const R*)nullptr, 0, (R)h->rnl, h->QZ.as<cx<R>>(), hdz);
      HIPCHK(hipGetLastError());
    }
  }
  if (int rc = sp_linear<R>(h, dz)) return rc;
  return sp_transform<R>(h, h->E.as<cx<R>>(), h->P.as<cx<R>>(), true, fused, hdz);
}

template <typename R>
int sp_analyze(fil_spatial* h, double z, fil_diag_spatial* d, double* Iout, double* Izx, double* Izy) {
  const size_t n = h->n;
  hipLaunchKernelGGL(sp_stats_kernel<R>, dim3(SP_STAT_BLOCKS), dim3(256), 0, h->st, h->E.as<cx<R>>(), n,
                     h->xy ? (const double*)nullptr : h->wgt.as<double>(), h->I.as<R>(), h->pmax.as<double>(),
                     h->pidx.as<long long>(), h->psum.as<double>(), h->pnan.as<int>());
  HIPCHK(hipGetLastError());
  std::vector<double> pmax(SP_STAT_BLOCKS), psum(SP_STAT_BLOCKS);
  std::vector<long long> pidx(SP_STAT_BLOCKS);
  std::vector<int> pnan(SP_STAT_BLOCKS);
  HIPCHK(hipMemcpyAsync(pmax.data(), h->pmax.p, SP_STAT_BLOCKS * sizeof(double), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipMemcpyAsync(pidx.data(), h->pidx.p, SP_STAT_BLOCKS * sizeof(long long), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipMemcpyAsync(psum.data(), h->psum.p, SP_STAT_BLOCKS * sizeof(double), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipMemcpyAsync(pnan.data(), h->pnan.p, SP_STAT_BLOCKS * sizeof(int), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  double Imax = -INFINITY, sum = 0;
  long long idx = -1;
  bool nan = false;
  for (int b = 0; b < SP_STAT_BLOCKS; b++) {
    sum += psum[b];
    nan = nan || pnan[b];
    if (pidx[b] < 0) continue;
    if (pmax[b] > Imax || (pmax[b] == Imax && pidx[b] < idx)) { Imax = pmax[b]; idx = pidx[b]; }
  }
  if (idx < 0) idx = 0;
  d->z = z;
  d->Imax = nan ? std::numeric_limits<double>::quiet_NaN() : Imax;
  d->rfil = d->ax = d->ay = 0;
  const R* Id = h->I.as<R>();
  auto get = [&](std::vector<R>& v, size_t off, size_t count, size_t stride) -> int {
    v.resize(count);
    if (stride == 1) {
      HIPCHK(hipMemcpy(v.data(), Id + off, count * sizeof(R), hipMemcpyDeviceToHost));
    } else {
      HIPCHK(hipMemcpy2D(v.data(), sizeof(R), Id + off, stride * sizeof(R), sizeof(R), count, hipMemcpyDeviceToHost));
    }
    return 0;
  };
  std::vector<R> line;
  if (!h->xy) {
    d->imax = idx;
    d->jmax = 0;
    if (int rc = get(line, 0, n, 1)) return rc;
    d->rfil = 2 * radius_r<R>(h->r.data(), line.data(), h->n);   // FieldAnalyzers.jl:64
    d->P = 2 * kPi * sum;                                          // :66
    if (Iout) for (size_t q = 0; q < n; q++) Iout[q] = (double)line[q];
    return 0;
  }
  const int Nx = h->d.Nx, Ny = h->d.Ny;
  d->imax = idx % Nx;
  d->jmax = idx / Nx;
  if (int rc = get(line, (size_t)d->jmax * Nx, Nx, 1)) return rc;   // I[:, jmax]
  d->ax = radius_r<R>(h->x.data(), line.data(), Nx);
  if (int rc = get(line, (size_t)d->imax, Ny, Nx)) return rc;       // I[imax, :]
  d->ay = radius_r<R>(h->y.data(), line.data(), Ny);
  d->P = sum * h->dx * h->dy;                                        // FieldAnalyzers.jl:262
  if (Izx) {
    if (int rc = get(line, (size_t)(Ny / 2 - 1) * Nx, Nx, 1)) return rc;
    for (int i = 0; i < Nx; i++) Izx[i] = (double)line[i];
  }
  if (Izy) {
    if (int rc = get(line, (size_t)(Nx / 2 - 1), Ny, Nx)) return rc;
    for (int j = 0; j < Ny; j++) Izy[j] = (double)line[j];
  }
  if (Iout) {
    if (int rc = get(line, 0, n, 1)) return rc;
    for (size_t q = 0; q < n; q++) Iout[q] = (double)line[q];
  }
  return 0;
}

template <typename R>
int upload_real(DevBuf& b, const double* src, size_t n) {
  std::vector<R> v(n);
  for (size_t i = 0; i < n; i++) v[i] = (R)src[i];
  HIPCHK(upload(b, v.data(), n * sizeof(R)));
  return 0;
}
template <typename R>
int upload_cplx(DevBuf& b, const double* src, size_t n) {
  std::vector<R> v(2 * n);
  for (size_t i = 0; i < 2 * n; i++) v[i] = (R)src[i];
  HIPCHK(upload(b, v.data(), n * sizeof(cx<R>)));
  return 0;
}

template <typename R>
int sp_set_tables(fil_spatial* h, const fil_spatial_tables* t) {
  const size_t n = h->n;
  if (!t->KZ || !t->QZ) return fail(FIL_EINVAL, "fil_spatial_set_tables: KZ and QZ are required");
  HIPCHK(upload(h->KZ, t->KZ, n * 2 * sizeof(double)));
  if (int rc = upload_cplx<R>(h->QZ, t->QZ, n)) return rc;
  h->rnl = t->Rnl;
  std::vector<double> pnl(2 * n);
  if (!h->xy) {
    const int N = h->d.Nr;
    if (!t->Kguard || !t->Rguard || !t->dht_T || !t->dht_RJ || !t->dht_JV || !t->r || !t->dr)
      return fail(FIL_EINVAL, "fil_spatial_set_tables: R needs Kguard, Rguard, dht_T, dht_RJ, dht_JV, r, dr");
    HIPCHK(upload(h->Kg, t->Kguard, n * sizeof(double)));
    if (int rc = upload_real<R>(h->Rg, t->Rguard, n)) return rc;
    // M1 = Tm diag(RJ), M2 = diag(1/RJ) Tm, row-major (Tm is symmetric; the
    // caller's column-major array is read as Tm[i + N j])
    std::vector<double> m1((size_t)N * N), m2((size_t)N * N);
    for (int k = 0; k < N; k++)
      for (int j = 0; j < N; j++) {
        const double tkj = t->dht_T[(size_t)k + (size_t)N * j];
        m1[(size_t)k * N + j] = tkj * t->dht_RJ[j];
        m2[(size_t)k * N + j] = tkj / t->dht_RJ[k];
      }
    if (int rc = upload_real<R>(h->M1, m1.data(), m1.size())) return rc;
    if (int rc = upload_real<R>(h->M2, m2.data(), m2.size())) return rc;
    std::vector<double> w(n);
    for (size_t i = 0; i < n; i++) w[i] = t->r[i] * t->dr[i];
    HIPCHK(upload(h->wgt, w.data(), n * sizeof(double)));
    h->r.assign(t->r, t->r + n);
    for (size_t q = 0; q < n; q++) {   // i QZ Kguard
      pnl[2 * q] = -t->QZ[2 * q + 1] * t->Kguard[q];
      pnl[2 * q + 1] = t->QZ[2 * q] * t->Kguard[q];
    }
  } else {
    const int Nx = h->d.Nx, Ny = h->d.Ny;
    if (!t->KXguard || !t->KYguard || !t->Xguard || !t->Yguard || !t->x || !t->y)
      return fail(FIL_EINVAL, "fil_spatial_set_tables: XY needs KXguard, KYguard, Xguard, Yguard, x, y");
    HIPCHK(upload(h->KX, t->KXguard, Nx * sizeof(double)));
    HIPCHK(upload(h->KY, t->KYguard, Ny * sizeof(double)));
    if (int rc = upload_real<R>(h->Xg, t->Xguard, Nx)) return rc;
    if (int rc = upload_real<R>(h->Yg, t->Yguard, Ny)) return rc;
    if (int rc = build_twiddles(h->twx, Nx, sizeof(R) == 8, false)) return rc;
    if (int rc = build_twiddles(h->twy, Ny, sizeof(R) == 8, false)) return rc;
    h->x.assign(t->x, t->x + Nx);
    h->y.assign(t->y, t->y + Ny);
    h->dx = t->dx;
    h->dy = t->dy;
    const double norm = 1.0 / ((double)Nx * Ny);
    for (size_t q = 0; q < n; q++) {   // i QZ KX KY / (Nx Ny)
      const double g = t->KXguard[q % Nx] * t->KYguard[q / Nx] * norm;
      pnl[2 * q] = -t->QZ[2 * q + 1] * g;
      pnl[2 * q + 1] = t->QZ[2 * q] * g;
    }
  }
  if (h->d.nonparaxial) {
    if (int rc = upload_cplx<R>(h->Pnl, pnl.data(), n)) return rc;
  }
  h->P_dz = std::numeric_limits<double>::quiet_NaN();
  h->tables = true;
  return 0;
}

int sp_check(const fil_spatial* h, bool need_field) {
  if (!h) return fail(FIL_EINVAL, "null spatial handle");
  if (!h->tables) return fail(FIL_ESTATE, "spatial tables not set");
  if (need_field && !h->field) return fail(FIL_ESTATE, "spatial field not uploaded");
  return sp_set_device(h);
}

bool pow2_in(int v, int lo, int hi) { return v >= lo && v <= hi && (v & (v - 1)) == 0; }

}  // namespace

extern "C" {

int fil_spatial_create(const fil_spatial_desc* d, fil_spatial** out) {
  if (!d || !out) return fail(FIL_EINVAL, "fil_spatial_create: null argument");
  *out = nullptr;
  if (d->abi_version != FIL_ABI_VERSION) return fail(FIL_EINVAL, "fil_spatial_create: ABI version mismatch");
  if (d->geometry != FIL_GEOM_R && d->geometry != FIL_GEOM_XY)
    return fail(FIL_EINVAL, "fil_spatial_create: geometry must be FIL_GEOM_R or FIL_GEOM_XY");
  if (d->precision != FIL_FP32 && d->precision != FIL_FP64)
    return fail(FIL_EINVAL, "fil_spatial_create: precision must be FIL_FP32 or FIL_FP64");
  if (d->alg != FIL_ALG_RK3 && d->alg != FIL_ALG_RK4)
    return fail(FIL_EINVAL, "fil_spatial_create: alg must be FIL_ALG_RK3 or FIL_ALG_RK4");
  const bool xy = d->geometry == FIL_GEOM_XY;
  if (!xy && (d->Nr < 2 || d->Nr > 16384)) return fail(FIL_EINVAL, "fil_spatial_create: Nr must be in [2, 16384]");
  if (xy && (!pow2_in(d->Nx, 16, 4096) || !pow2_in(d->Ny, 16, 4096)))
    return fail(FIL_EINVAL, "fil_spatial_create: Nx and Ny must be powers of two in [16, 4096]");
  fil_spatial* h = new fil_spatial();
  h->d = *d;
  h->xy = xy;
  h->f64 = d->precision == FIL_FP64;
  h->n = xy ? d->Nx * d->Ny : d->Nr;
  auto bail = [&](int rc) { delete h; return rc; };
  if (hipSetDevice(d->device) != hipSuccess) return bail(fail(FIL_EDEVICE, "fil_spatial_create: hipSetDevice"));
  if (hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(FIL_EDEVICE, "fil_spatial_create: hipStreamCreate"));
  const size_t cb = (h->f64 ? 16 : 8) * (size_t)h->n, rb = (h->f64 ? 8 : 4) * (size_t)h->n;
  bool ok = h->E.alloc(cb) == hipSuccess && h->P.alloc(cb) == hipSuccess && h->I.alloc(rb) == hipSuccess &&
            h->pmax.alloc(SP_STAT_BLOCKS * sizeof(double)) == hipSuccess &&
            h->pidx.alloc(SP_STAT_BLOCKS * sizeof(long long)) == hipSuccess &&
            h->psum.alloc(SP_STAT_BLOCKS * sizeof(double)) == hipSuccess &&
            h->pnan.alloc(SP_STAT_BLOCKS * sizeof(int)) == hipSuccess;
  if (ok && !xy) ok = h->W.alloc(cb) == hipSuccess;
  if (ok && d->nonlinearity && d->nonparaxial)
    for (int j = 0; j < d->alg && ok; j++) ok = h->K[j].alloc(cb) == hipSuccess;
  if (!ok) return bail(fail(FIL_ENOMEM, "fil_spatial_create: device allocation"));
  if (hipMemset(h->E.p, 0, cb) != hipSuccess) return bail(fail(FIL_EDEVICE, "fil_spatial_create: hipMemset"));
  *out = h;
  return 0;
}

int fil_spatial_set_tables(fil_spatial* h, const fil_spatial_tables* t) {
  if (!h || !t) return fail(FIL_EINVAL, "fil_spatial_set_tables: null argument");
  if (int rc = sp_set_device(h)) return rc;
  HIPCHK(hipStreamSynchronize(h->st));
  return h->f64 ? sp_set_tables<double>(h, t) : sp_set_tables<float>(h, t);
}

int fil_spatial_upload_field(fil_spatial* h, const double* E) {
  if (!h || !E) return fail(FIL_EINVAL, "fil_spatial_upload_field: null argument");
  if (int rc = sp_set_device(h)) return rc;
  HIPCHK(hipStreamSynchronize(h->st));
  int rc = h->f64 ? upload_cplx<double>(h->E, E, h->n) : upload_cplx<float>(h->E, E, h->n);
  if (rc == 0) h->field = true;
  return rc;
}

int fil_spatial_download_field(fil_spatial* h, double* E) {
  if (!h || !E) return fail(FIL_EINVAL, "fil_spatial_download_field: null argument");
  if (int rc = sp_set_device(h)) return rc;
  const size_t n2 = 2 * (size_t)h->n;
  HIPCHK(hipStreamSynchronize(h->st));
  if (h->f64) {
    HIPCHK(hipMemcpy(E, h->E.p, n2 * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    std::vector<float> v(n2);
    HIPCHK(hipMemcpy(v.data(), h->E.p, n2 * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n2; i++) E[i] = v[i];
  }
  return 0;
}

int fil_spatial_zstep(fil_spatial* h, double z, double dz) {
  (void)z;   // the spatial responses do not depend on z
  if (int rc = sp_check(h, true)) return rc;
  if (!(dz >= 0) || !std::isfinite(dz)) return fail(FIL_EINVAL, "fil_spatial_zstep: dz must be finite and >= 0");
  return h->f64 ? sp_zstep<double>(h, dz) : sp_zstep<float>(h, dz);
}

int fil_spatial_analyze(fil_spatial* h, double z, fil_diag_spatial* d, double* I, double* Izx, double* Izy) {
  if (!d) return fail(FIL_EINVAL, "fil_spatial_analyze: null diag");
  if (int rc = sp_check(h, true)) return rc;
  return h->f64 ? sp_analyze<double>(h, z, d, I, Izx, Izy) : sp_analyze<float>(h, z, d, I, Izx, Izy);
}

int fil_spatial_synchronize(fil_spatial* h) {
  if (!h) return fail(FIL_EINVAL, "null spatial handle");
  if (int rc = sp_set_device(h)) return rc;
  HIPCHK(hipStreamSynchronize(h->st));
  return 0;
}

int fil_spatial_destroy(fil_spatial* h) {
  if (!h) return 0;
  (void)hipSetDevice(h->d.device);
  if (h->st) {
    (void)hipStreamSynchronize(h->st);
    (void)hipStreamDestroy(h->st);
  }
  delete h;
  return 0;
}

}  // extern "C"
