// fil_xyt.hip -- the linear step of the XYT geometry on gfx950.
//
// propagate!(E, LP::LinearPropagatorSpectral, dz) (linear_propagators/spectral.jl:16-26)
// with the 2-D FFT space plan of Field XYT (Fields.jl:152, plan_fft!(E, [1, 2])):
//   E <- IFFT_xy[ P .* FFT_xy[E] ],  P = exp(i KZ dz) KXguard KYguard Wguard,
// forward_transform_space! = P * E (unnormalised, exp(-i k x)),
// inverse_transform_space! = P \ E (normalised) (Models.jl:135-148).
//
// The state is the half spectrum S[(ix + Nx iy)][w], w < Nt/2 contiguous.  The
// 2-D transform of every w column is done in three passes over HBM:
//   x pass      FFT along x of every (iy, w) line       (stride Nt/2)
//   y pass      FFT along y, multiply by P, inverse FFT along y (stride Nx Nt/2)
//   x pass      inverse FFT along x
// Each workgroup owns one line (iy or ix) and WB consecutive w columns: it
// reads WB contiguous complex values per grid point (full cache lines), runs
// WB transforms in LDS, one group of N/8 threads per transform, and writes
// them back.  1/(Nx Ny) of the inverse transform is folded into P.
//
// P depends on kx^2 and ky^2 (and on |kx|, |ky| through the guards), which the
// FFT frequency grid makes equal at ix and Nx - ix, iy and Ny - iy: P is stored
// for ix <= Nx/2, iy <= Ny/2 only (a quarter: 135 MB in FP32 at 512^3 instead
// of 537 MB).  In the y pass one load serves iy and Ny - iy, and the lines ix
// and Nx - ix are given to blocks nwb apart in launch order (the same XCD
// when nwb is a multiple of 8), so the second reader of a table row finds it
// in L2.
#include "fil_internal.h"
#include "fil_fft.h"
#include "fil_kernels.h"

namespace fil {

namespace {

// complex FP64 helpers for the propagator
struct cd { double x, y; };
__device__ __forceinline__ cd cmul(cd a, cd b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ cd cdiv(cd a, cd b) {
  const double d = b.x * b.x + b.y * b.y;
  return {(a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d};
}
// principal square root; a zero imaginary part counts as +0 (the reference
// adds 0im before sqrt, spectral.jl:182)
__device__ __forceinline__ cd csqrt(cd z) {
  const double r = hypot(z.x, z.y);
  const double re = sqrt(0.5 * (r + z.x));
  const double im = sqrt(0.5 * fmax(r - z.x, 0.0));
  return {re, (z.y < 0.0) ? -im : im};
}

}  // namespace

// P[(ix + Px iy) Nh + w] for w < Nh, ix < Px, iy < Py (Kfunc spectral.jl:160-184,
// KZ XYT :131-156, spectral filter Guards.jl:250-261 with kernel! :301-318).
template <typename R>
__global__ void xyt_propagator_kernel(XytProp pp, double dz_arg, const double* __restrict__ dz_dev,
                                      cx<R>* __restrict__ P) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n = (size_t)pp.Px * pp.Py * pp.Nh;
  if (q >= n) return;
  const double dz = dz_dev ? *dz_dev : dz_arg;
  const int w = (int)(q % pp.Nh);
  const size_t row = q / pp.Nh;
  const int ix = (int)(row % pp.Px), iy = (int)(row / pp.Px);
  const double kt2 = pp.kt2x[ix] + pp.kt2y[iy];
  const cd beta = {pp.beta[2 * w], pp.beta[2 * w + 1]};
  cd K;
  if (pp.kparaxial) {
    if (beta.x == 0.0 && beta.y == 0.0) K = {0.0, 0.0};
    else {
      const cd d = cdiv({kt2, 0.0}, {2.0 * beta.x, 2.0 * beta.y});
      K = {beta.x - d.x, beta.y - d.y};
    }
  } else {
    const cd b2 = cmul(beta, beta);
    K = csqrt({b2.x - kt2, b2.y});
  }
  const double kzr = (K.x - pp.wvf[w]) * pp.zu, kzi = K.y * pp.zu;
  // exp(i KZ dz)
  const double mag = exp(-kzi * dz);
  double sn, cs;
  sincos(kzr * dz, &sn, &cs);
  const double g = pp.KX[ix] * pp.KY[iy] * (pp.guard3d_bug ? 1.0 : pp.Wg[w]) * pp.norm;
  P[q] = mk((R)(g * mag * cs), (R)(g * mag * sn));
}

template <typename R>
hipError_t launch_xyt_propagator(const XytProp& pp, double dz, cx<R>* P, hipStream_t st, const double* dz_dev) {
  const size_t n = (size_t)pp.Px * pp.Py * pp.Nh;
  hipLaunchKernelGGL(xyt_propagator_kernel<R>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, pp, dz, dz_dev,
                     P);
  return hipGetLastError();
}

// Q[(ix + Px iy) Nh + w] for w < Nh, ix < Px, iy < Py: the non-paraxial
// prefactor of func! (nonlinear_propagators.jl:58-63; QZfunc XYT :184-204,
// Qfunc_nonparaxial :232-242) over the paraxial QZ(w) (Qfunc_paraxial
// :221-230) that the stage right-hand side already carries, with func!'s
// spectral filter (:61) and the inverse transform's 1/(Nx Ny).  mu0 mu w^2
// zu / Eu cancels in the ratio: Q = beta / kz, kz = sqrt(beta^2 - kt^2 + 0im);
// 0 where beta = 0 (QZ(w) = 0) or kz = 0 (the reference's zero there).
template <typename R>
__global__ void xyt_qz_kernel(XytProp pp, cx<R>* __restrict__ Q) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n = (size_t)pp.Px * pp.Py * pp.Nh;
  if (q >= n) return;
  const int w = (int)(q % pp.Nh);
  const size_t row = q / pp.Nh;
  const int ix = (int)(row % pp.Px), iy = (int)(row / pp.Px);
  const double kt2 = pp.kt2x[ix] + pp.kt2y[iy];
  const cd beta = {pp.beta[2 * w], pp.beta[2 * w + 1]};
  cd v = {0.0, 0.0};
  if (beta.x != 0.0 || beta.y != 0.0) {
    const cd b2 = cmul(beta, beta);
    const cd kz = csqrt({b2.x - kt2, b2.y});
    if (kz.x != 0.0 || kz.y != 0.0) v = cdiv(beta, kz);
  }
  const double g = pp.KX[ix] * pp.KY[iy] * (pp.guard3d_bug ? 1.0 : pp.Wg[w]) * pp.norm;
  Q[q] = mk((R)(g * v.x), (R)(g * v.y));
}

template <typename R>
hipError_t launch_xyt_qz(const XytProp& pp, cx<R>* Q, hipStream_t st) {
  const size_t n = (size_t)pp.Px * pp.Py * pp.Nh;
  hipLaunchKernelGGL(xyt_qz_kernel<R>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, pp, Q);
  return hipGetLastError();
}

// Transforms of length N along a strided dimension: WB columns per workgroup.
// MODE 0: forward (exp(-i)), 1: inverse (exp(+i), unnormalised),
// 2: forward, multiply by P (the stored part, xyt_propagator_kernel), inverse.
template <int N> constexpr int xyt_col_threads() { return N / FIL_FFT_MAXRAD >= 2 ? N / FIL_FFT_MAXRAD : 2; }
template <typename R, int N> constexpr int xyt_cols() {
  return (sizeof(R) == 4) ? (N <= 512 ? 16 : 8) : (N <= 512 ? 8 : 4);
}
template <int N> constexpr int xyt_col_stride() { return lds_size(N) + 1; }   // +1: columns start in different banks

// Line of the j-th group of nwb blocks in the y pass with xsym: 0, Nx/2, 1,
// Nx - 1, 2, Nx - 2, ... (ix and Nx - ix read the same rows of P).
__device__ __forceinline__ long xyt_paired_line(long j, int Nx) {
  if (j < 2) return j == 0 ? 0 : Nx / 2;
  const long k = j / 2;
  return (j & 1) ? Nx - k : k;
}

template <typename R, int N, int MODE>
__global__ __launch_bounds__((xyt_cols<R, N>() * xyt_col_threads<N>())) void xyt_fft_kernel(
    cx<R>* S, const cx<R>* __restrict__ tw, const cx<R>* __restrict__ P, XytProp pp, long line_stride,
    long elem_stride, int Nh) {
  constexpr int WB = xyt_cols<R, N>();
  constexpr int NTHC = xyt_col_threads<N>();
  constexpr int NT = WB * NTHC;
  constexpr int CS = xyt_col_stride<N>();
  extern __shared__ __align__(16) unsigned char smem[];
  cx<R>* Z = reinterpret_cast<cx<R>*>(smem);
  const int tid = threadIdx.x, col = tid / NTHC, ctid = tid % NTHC;
  const int nwb = Nh / WB;
  const long line = (MODE == 2 && pp.xsym) ? xyt_paired_line(blockIdx.x / nwb, pp.Nx) : (long)(blockIdx.x / nwb);
  const int w0 = (blockIdx.x % nwb) * WB;
  const size_t base = (size_t)(line * line_stride + w0);
  // Each thread moves PER elements of one column (NT is a multiple of WB).
  // All its loads are issued before its first LDS write: a load followed by
  // its write in the same loop iteration would wait out each load's latency
  // in turn.
  constexpr int PER = N * WB / NT;
  static_assert(PER * NT == N * WB, "whole elements per thread");
  const int c = tid % WB, e0 = tid / WB;
  constexpr int EST = NT / WB;                    // element step between a thread's elements
  {
    cx<R> v[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) v[k] = S[base + (size_t)(e0 + k * EST) * elem_stride + c];
#pragma unroll
    for (int k = 0; k < PER; k++) Z[c * CS + lds_index(e0 + k * EST)] = v[k];
  }
  __syncthreads();
  cx<R>* Zc = Z + col * CS;
  if constexpr (MODE == 1) {
    fft_lds<R, N, NTHC, +1>(Zc, tw, ctid);
  } else {
    fft_lds<R, N, NTHC, -1>(Zc, tw, ctid);
  }
  if constexpr (MODE == 2) {
    // This thread's propagator values, all loads issued before the first
    // multiply (held across the forward transform they would cost the second
    // workgroup per CU its registers).  The line is ix, the elements iy: P row
    // min(ix, Nx - ix) (xsym) or ix, plus Px iy; with ysym the value of
    // iy <= Ny/2 also serves Ny - iy.
    constexpr int PLIM = (N / 2 + 1) * WB;
    constexpr int PPER = (PLIM + NT - 1) / NT;
    const long ixp = (pp.xsym && line > pp.Nx / 2) ? pp.Nx - line : line;
    const cx<R>* Pl = P + (size_t)ixp * Nh + w0 + c;
    const size_t pstride = (size_t)pp.Px * Nh;
    cx<R>* Zt = Z + c * CS;
    if (pp.ysym) {
      cx<R> pv[PPER];
#pragma unroll
      for (int k = 0; k < PPER; k++) {
        const int e = e0 + k * EST;
        if (e <= N / 2) pv[k] = Pl[(size_t)e * pstride];
      }
#pragma unroll
      for (int k = 0; k < PPER; k++) {
        const int e = e0 + k * EST;
        if (e <= N / 2) {
          Zt[lds_index(e)] = Zt[lds_index(e)] * pv[k];
          if (e != 0 && e != N / 2) Zt[lds_index(N - e)] = Zt[lds_index(N - e)] * pv[k];
        }
      }
    } else {
      // tables that are not even in ky (not the FFT grid's): one element at a time
      for (int k = 0; k < PER; k++) {
        const int e = e0 + k * EST;
        Zt[lds_index(e)] = Zt[lds_index(e)] * Pl[(size_t)e * pstride];
      }
    }
    __syncthreads();
    fft_lds<R, N, NTHC, +1>(Zc, tw, ctid);
  }
  {
    // the element indices pass through an empty asm so that the store
    // addresses are formed here rather than kept live across the transforms
    int cb = c, eb = e0;
    asm volatile("" : "+v"(cb), "+v"(eb));
    cx<R> v[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) v[k] = Z[cb * CS + lds_index(eb + k * EST)];
#pragma unroll
    for (int k = 0; k < PER; k++) S[base + (size_t)(eb + k * EST) * elem_stride + cb] = v[k];
  }
}

#define FIL_DISPATCH_XY(N_, ...)                                  \
  switch (N_) {                                                   \
    case 16: { constexpr int NN = 16; __VA_ARGS__; } break;       \
    case 32: { constexpr int NN = 32; __VA_ARGS__; } break;       \
    case 64: { constexpr int NN = 64; __VA_ARGS__; } break;       \
    case 128: { constexpr int NN = 128; __VA_ARGS__; } break;     \
    case 256: { constexpr int NN = 256; __VA_ARGS__; } break;     \
    case 512: { constexpr int NN = 512; __VA_ARGS__; } break;     \
    case 1024: { constexpr int NN = 1024; __VA_ARGS__; } break;   \
    default: return hipErrorInvalidValue;                         \
  }

template <typename R, int N, int MODE>
static hipError_t launch_xyt_fft_n(cx<R>* S, const cx<R>* tw, const cx<R>* P, const XytProp& pp, long nlines,
                                   long line_stride, long elem_stride, int Nh, hipStream_t st) {
  constexpr int WB = xyt_cols<R, N>();
  constexpr int NT = WB * xyt_col_threads<N>();
  if (Nh % WB) return hipErrorInvalidValue;
  const size_t sm = (size_t)WB * xyt_col_stride<N>() * sizeof(cx<R>);
  auto k = xyt_fft_kernel<R, N, MODE>;
  hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k, dim3((unsigned)(nlines * (Nh / WB))), dim3(NT), sm, st, S, tw, P, pp, line_stride,
                     elem_stride, Nh);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_xyt_linear(cx<R>* S, const cx<R>* P, const XytProp& pp, const cx<R>* twx, const cx<R>* twy,
                             hipStream_t st) {
  const int Nx = pp.Nx, Ny = pp.Ny, Nh = pp.Nh;
  const long nh = Nh;
  // the stored part of P must match the symmetry flags
  if (pp.Px != (pp.xsym ? Nx / 2 + 1 : Nx) || pp.Py != (pp.ysym ? Ny / 2 + 1 : Ny)) return hipErrorInvalidValue;
  hipError_t e;
  // forward along x: lines iy, elements ix
  FIL_DISPATCH_XY(Nx, e = (launch_xyt_fft_n<R, NN, 0>(S, twx, nullptr, pp, Ny, (long)Nx * nh, nh, Nh, st)));
  if (e != hipSuccess) return e;
  // along y: forward, propagator, inverse; lines ix, elements iy
  FIL_DISPATCH_XY(Ny, e = (launch_xyt_fft_n<R, NN, 2>(S, twy, P, pp, Nx, nh, (long)Nx * nh, Nh, st)));
  if (e != hipSuccess) return e;
  // inverse along x
  FIL_DISPATCH_XY(Nx, e = (launch_xyt_fft_n<R, NN, 1>(S, twx, nullptr, pp, Ny, (long)Nx * nh, nh, Nh, st)));
  return e;
}

template hipError_t launch_xyt_qz<float>(const XytProp&, cx<float>*, hipStream_t);
template hipError_t launch_xyt_qz<double>(const XytProp&, cx<double>*, hipStream_t);
template hipError_t launch_xyt_propagator<float>(const XytProp&, double, cx<float>*, hipStream_t, const double*);
template hipError_t launch_xyt_propagator<double>(const XytProp&, double, cx<double>*, hipStream_t, const double*);
template hipError_t launch_xyt_linear<float>(cx<float>*, const cx<float>*, const XytProp&, const cx<float>*,
                                             const cx<float>*, hipStream_t);
template hipError_t launch_xyt_linear<double>(cx<double>*, const cx<double>*, const XytProp&, const cx<double>*,
                                              const cx<double>*, hipStream_t);

}  // namespace fil
