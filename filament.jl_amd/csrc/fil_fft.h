// fil_fft.h -- workgroup-wide FFT of one t-row held in LDS (gfx950).
//
// Replaces the cuFFT plan along t of the reference (plan_fft!(E, [2]),
// src/modules/Fields.jl:113; used at Models.jl:162-171, 180-183).
//
// Stockham decimation-in-frequency with radix-FIL_FFT_MAXRAD passes (a last
// smaller-radix pass when needed): at N = 4096 and radix 8 four passes, each
// one LDS read and one LDS write of the row.  Every pass reads its operands into
// registers, barriers, then writes the outputs back into the same array, so
// one buffer suffices.  The row is stored with one pad element per 16
// (lds_index(i) = i + i/16), which spreads the strided outputs of the first
// passes and the contiguous chunks of the plasma scan over the LDS banks.
// probe_fft_padding.hip times the N = 4096 transform alone for pad periods 8
// to 128: 32 is the fastest there (fewer bank conflicts), but in the row
// kernel with the FP64 plasma scan it costs two VGPRs over the 128 that keep
// two workgroups per CU, and the plain row kernel gains nothing from it.
// No normalisation is applied.
#pragma once

#include "fil_internal.h"

namespace fil {

// Largest radix of the Stockham passes.  With N/8 threads per row every
// thread owns one radix-8 butterfly per pass.
#define FIL_FFT_MAXRAD 8

// one pad element per 2^FIL_LDS_PAD_SHIFT elements (a build-time constant so
// that probe_fft_padding can build the other periods)
#ifndef FIL_LDS_PAD_SHIFT
#define FIL_LDS_PAD_SHIFT 4
#endif
__host__ __device__ constexpr int lds_index(int i) { return i + (i >> FIL_LDS_PAD_SHIFT); }
__host__ __device__ constexpr int lds_size(int n) { return n + (n >> FIL_LDS_PAD_SHIFT); }

// cos, sin of 2 pi k / 16, k = 0..7
__device__ __forceinline__ double c16(int k) {
  constexpr double c[8] = {1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173,
                           0.0, -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613};
  return c[k];
}
__device__ __forceinline__ double s16(int k) {
  constexpr double s[8] = {0.0, 0.38268343236508977173, 0.70710678118654752440, 0.92387953251128675613,
                           1.0, 0.92387953251128675613, 0.70710678118654752440, 0.38268343236508977173};
  return s[k];
}

// In-register DFT of size RAD (2, 4, 8, 16): a[k] <- sum_j a[j] exp(SIGN 2 pi i jk / RAD).
// Radix-2 decimation in time, recursively.
template <typename T, int RAD, int SIGN>
struct Dft {
  static __device__ __forceinline__ void run(cx<T>* a) {
    constexpr int H = RAD / 2;
    cx<T> e[H], o[H];
#pragma unroll
    for (int j = 0; j < H; j++) { e[j] = a[2 * j]; o[j] = a[2 * j + 1]; }
    Dft<T, H, SIGN>::run(e);
    Dft<T, H, SIGN>::run(o);
#pragma unroll
    for (int k = 0; k < H; k++) {
      const cx<T> t = twiddle16(k * (16 / RAD), o[k]);
      a[k] = e[k] + t;
      a[k + H] = e[k] - t;
    }
  }
  // o * exp(SIGN 2 pi i m / 16); m = 0 and m = 4 (a factor SIGN i) need no multiply
  static __device__ __forceinline__ cx<T> twiddle16(int m, cx<T> o) {
    if (m == 0) return o;
    if (m == 4) return (SIGN > 0) ? mul_i(o) : mul_mi(o);
    return mk((T)c16(m), (T)(SIGN * s16(m))) * o;
  }
  // the same DFT when a[j] = 0 for j >= RAD/2 (the upper half of the input is
  // not read): both half-size DFTs see only their lower half
  static __device__ __forceinline__ void run_half(cx<T>* a) {
    constexpr int H = RAD / 2;
    cx<T> e[H], o[H];
#pragma unroll
    for (int j = 0; j < H; j++) {
      e[j] = (j < H / 2) ? a[2 * j] : mk((T)0, (T)0);
      o[j] = (j < H / 2) ? a[2 * j + 1] : mk((T)0, (T)0);
    }
    Dft<T, H, SIGN>::run_half(e);
    Dft<T, H, SIGN>::run_half(o);
#pragma unroll
    for (int k = 0; k < H; k++) {
      const cx<T> t = twiddle16(k * (16 / RAD), o[k]);
      a[k] = e[k] + t;
      a[k + H] = e[k] - t;
    }
  }
};
template <typename T, int SIGN>
struct Dft<T, 2, SIGN> {
  static __device__ __forceinline__ void run(cx<T>* a) {
    cx<T> t = a[0];
    a[0] = t + a[1];
    a[1] = t - a[1];
  }
  static __device__ __forceinline__ void run_half(cx<T>* a) { a[1] = a[0]; }
};
template <typename T, int SIGN>
struct Dft<T, 4, SIGN> {
  static __device__ __forceinline__ void run(cx<T>* a) {
    cx<T> apc = a[0] + a[2], amc = a[0] - a[2], bpd = a[1] + a[3], bmd = a[1] - a[3];
    cx<T> jb = (SIGN > 0) ? mul_i(bmd) : mul_mi(bmd);
    a[0] = apc + bpd;
    a[1] = amc + jb;
    a[2] = apc - bpd;
    a[3] = amc - jb;
  }
  // a[2] = a[3] = 0
  static __device__ __forceinline__ void run_half(cx<T>* a) {
    const cx<T> jb = (SIGN > 0) ? mul_i(a[1]) : mul_mi(a[1]);
    const cx<T> a0 = a[0], a1 = a[1];
    a[0] = a0 + a1;
    a[1] = a0 + jb;
    a[2] = a0 - a1;
    a[3] = a0 - jb;
  }
};

// Twiddle table layout (built on the host, fil_set_tables): for every pass
// that has twiddles (stride S < N/RAD) a block of P x FIL_TW_STRIDE entries,
// P = N/(S RAD), entry [p * FIL_TW_STRIDE + k] = exp(-2 pi i p S k / N),
// k = 0..RAD-1 (k = 0 is 1 and unused).  A thread's twiddles are contiguous, so
// it reads them with a few wide loads at constant offsets from one address.
#define FIL_TW_STRIDE FIL_FFT_MAXRAD
__host__ __device__ constexpr int fft_radix(int N, int S) { return (N / S >= FIL_FFT_MAXRAD) ? FIL_FFT_MAXRAD : N / S; }
__host__ __device__ constexpr int fft_tw_offset(int N, int S) {
  int off = 0;
  for (int s = 1; s < S; s *= fft_radix(N, s)) {
    const int rad = fft_radix(N, s);
    if (s < N / rad) off += (N / (s * rad)) * FIL_TW_STRIDE;
  }
  return off;
}
__host__ __device__ constexpr int fft_tw_size(int N) { return fft_tw_offset(N, N); }
// Extended table of the row transforms: after the N-point table the
// N/2-point table (a real signal of length N transformed as N/2 complex
// points), then the N/2 post-processing factors exp(+2 pi i w / N), w < N/2.
__host__ __device__ constexpr int fft_tw_half_offset(int N) { return fft_tw_size(N); }
__host__ __device__ constexpr int fft_tw_post_offset(int N) { return fft_tw_size(N) + fft_tw_size(N / 2); }
__host__ __device__ constexpr int fft_tw_ext_size(int N) { return fft_tw_post_offset(N) + N / 2; }

// LDS offset of element i0 + c for a compile-time c that is a multiple of a
// power of two C with i0 mod C + c's carry staying below the pad period
// (every access of the passes below): lds_index(i0 + c) = lds_index(i0) + lds_step(c).
__host__ __device__ constexpr int lds_step(int c) { return c + (c >> FIL_LDS_PAD_SHIFT); }

// One Stockham pass of radix RAD at stride S over N points.
// Inputs  x[q + S (p + k N/(S RAD))] = x[b + k N/RAD],  b = p S + q;
// outputs y[q + S (RAD p + k)] = DFT_RAD(...)[k] * w^(p S k),  w = exp(SIGN 2 pi i / N).
// LDS addresses are one padded base per thread plus compile-time offsets:
// for reads lds_index(b + k NB) = lds_index(b) + lds_step(k NB) when NB is a
// multiple of the pad period; for writes lds_index(w0 + S k) = lds_index(w0) + lds_step(S k) with
// w0 = q + S RAD p, because S, RAD and the period are powers of two and q < S
// keeps the low bits of w0 + S k from carrying (checked for every pass of
// N = 16..8192 and periods 8..256).
// NAT: the outputs go to z in the natural (unpadded) order, z[i] = y[i]; only
// for the last pass of a transform whose reader wants that layout.
template <typename T, int N, int NTH, int SIGN, int RAD, int S, bool NAT = false>
__device__ __forceinline__ void stockham_pass(cx<T>* z, const cx<T>* __restrict__ tw, int tid) {
  constexpr int NB = N / RAD;
  constexpr int BPT = (NB + NTH - 1) / NTH;
  constexpr bool TW = S < N / RAD;     // the last pass has p == 0: no twiddles
  // The LDS addresses of a pass depend only on the thread index; without this
  // opaque copy the compiler hoists them out of every FFT of the kernel and
  // keeps them all live, which costs ~50 VGPRs and halves the occupancy.
  asm volatile("" : "+v"(tid));
  cx<T> a[BPT][RAD];
  cx<T> w[BPT][TW ? RAD : 1];
  // twiddle loads first, so their latency overlaps the LDS reads and the barrier
  if constexpr (TW) {
    const cx<T>* twp = tw + fft_tw_offset(N, S);
#pragma unroll
    for (int u = 0; u < BPT; u++) {
      const int b = tid + u * NTH;
      if (NB % NTH == 0 || b < NB) {
        const cx<T>* t = twp + (b / S) * FIL_TW_STRIDE;
#pragma unroll
        for (int k = 1; k < RAD; k++) w[u][k] = t[k];
      }
    }
  }
#pragma unroll
  for (int u = 0; u < BPT; u++) {
    const int b = tid + u * NTH;
    if (NB % NTH == 0 || b < NB) {
      if constexpr (NB % (1 << FIL_LDS_PAD_SHIFT) == 0) {
        const cx<T>* zr = z + lds_index(b);
#pragma unroll
        for (int k = 0; k < RAD; k++) a[u][k] = zr[lds_step(k * NB)];
      } else {   // short transforms: the read stride does not keep the padding aligned
#pragma unroll
        for (int k = 0; k < RAD; k++) a[u][k] = z[lds_index(b + k * NB)];
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < BPT; u++) {
    const int b = tid + u * NTH;
    if (NB % NTH == 0 || b < NB) {
      Dft<T, RAD, SIGN>::run(a[u]);
      const int q = b % S, p = b / S;
      if constexpr (TW) {
#pragma unroll
        for (int k = 1; k < RAD; k++) a[u][k] = a[u][k] * ((SIGN > 0) ? conj(w[u][k]) : w[u][k]);
      }
      cx<T>* zw = z + (NAT ? q + S * RAD * p : lds_index(q + S * RAD * p));
#pragma unroll
      for (int k = 0; k < RAD; k++) zw[NAT ? S * k : lds_step(S * k)] = a[u][k];
    }
  }
  __syncthreads();
}

template <typename T, int N, int NTH, int SIGN, int S, bool NAT = false>
struct FftPasses {
  static __device__ __forceinline__ void run(cx<T>* z, const cx<T>* tw, int tid) {
    constexpr int REM = N / S;
    if constexpr (REM >= FIL_FFT_MAXRAD) {
      stockham_pass<T, N, NTH, SIGN, FIL_FFT_MAXRAD, S, NAT && REM == FIL_FFT_MAXRAD>(z, tw, tid);
      FftPasses<T, N, NTH, SIGN, S * FIL_FFT_MAXRAD, NAT>::run(z, tw, tid);
    } else if constexpr (REM > 1) {
      stockham_pass<T, N, NTH, SIGN, REM, S, NAT>(z, tw, tid);
    }
  }
};

// SIGN = -1: y_k = sum_j x_j exp(-2 pi i jk/N)   (the reference's P * E)
// SIGN = +1: y_k = sum_j x_j exp(+2 pi i jk/N)   (N times the reference's P \ E)
// z is stored with lds_index(); tw: the per-pass twiddle table (global memory).
// tid: the thread's index among the NTH threads working on this transform
// (several transforms may share a workgroup; every pass ends in a
// workgroup-wide barrier, so all threads of the workgroup must call it).
// NAT_OUT: the result is left in natural order (z[i], no padding) instead.
template <typename T, int N, int NTH, int SIGN, bool NAT_OUT = false>
__device__ __forceinline__ void fft_lds(cx<T>* z, const cx<T>* __restrict__ tw, int tid) {
  static_assert((N & (N - 1)) == 0 && N >= 16, "N must be a power of two >= 16");
  FftPasses<T, N, NTH, SIGN, 1, NAT_OUT>::run(z, tw, tid);
}
template <typename T, int N, int NTH, int SIGN, bool NAT_OUT = false>
__device__ __forceinline__ void fft_lds(cx<T>* z, const cx<T>* __restrict__ tw) {
  fft_lds<T, N, NTH, SIGN, NAT_OUT>(z, tw, (int)threadIdx.x);
}

// ---------------------------------------------------------------------------
// Transforms whose input or output stays in registers (NTH = N/8, N >= 512).
// With N/8 threads the first pass (radix 8, stride 1) reads x[tid + k N/8]
// and the last pass writes y[tid + j N/8], j = 0..7: exactly the elements
// t = tid + j NTH that the elementwise code of the row kernel owns.  Taking
// the input from registers saves the LDS store + barrier + load in front of
// the first pass; leaving the output in registers saves the store + barrier +
// load behind the last one.
// ---------------------------------------------------------------------------
template <int N, int NTH>
__host__ __device__ constexpr bool fft_regs_ok() { return NTH * 8 == N && N >= 512; }

// stride of the last pass
__host__ __device__ constexpr int fft_last_stride(int N) {
  int s = 1;
  while (s * fft_radix(N, s) < N) s *= fft_radix(N, s);
  return s;
}

// First pass from registers: in[k] = x[tid + k N/8]; HALF: x[i] = 0 for
// i >= N/2 (in[4..7] are not read).  The caller guarantees that no thread
// still reads z; ends with a barrier.
template <typename T, int N, int NTH, int SIGN, bool HALF>
__device__ __forceinline__ void fft_first_pass_regs(const cx<T>* in, cx<T>* z, const cx<T>* __restrict__ tw, int tid) {
  static_assert(fft_regs_ok<N, NTH>(), "register passes need NTH = N/8");
  asm volatile("" : "+v"(tid));
  cx<T> a[8], w[8];
  const cx<T>* t = tw + fft_tw_offset(N, 1) + tid * FIL_TW_STRIDE;
#pragma unroll
  for (int k = 1; k < 8; k++) w[k] = t[k];
#pragma unroll
  for (int k = 0; k < 8; k++) a[k] = (HALF && k >= 4) ? mk((T)0, (T)0) : in[k];
  if constexpr (HALF) Dft<T, 8, SIGN>::run_half(a);
  else Dft<T, 8, SIGN>::run(a);
#pragma unroll
  for (int k = 1; k < 8; k++) a[k] = a[k] * ((SIGN > 0) ? conj(w[k]) : w[k]);
  cx<T>* zw = z + lds_index(8 * tid);
#pragma unroll
  for (int k = 0; k < 8; k++) zw[lds_step(k)] = a[k];
  __syncthreads();
}

// the passes between the first (stride 1) and the last
template <typename T, int N, int NTH, int SIGN, int S>
struct FftMidPasses {
  static __device__ __forceinline__ void run(cx<T>* z, const cx<T>* tw, int tid) {
    if constexpr (S < fft_last_stride(N)) {
      stockham_pass<T, N, NTH, SIGN, fft_radix(N, S), S>(z, tw, tid);
      FftMidPasses<T, N, NTH, SIGN, S * fft_radix(N, S)>::run(z, tw, tid);
    }
  }
};

// Last pass into registers: out[j] = y[tid + j N/8].  Reads z, barriers (z is
// free afterwards), then transforms.
template <typename T, int N, int NTH, int SIGN>
__device__ __forceinline__ void fft_last_pass_regs(const cx<T>* z, cx<T>* out, int tid) {
  static_assert(fft_regs_ok<N, NTH>(), "register passes need NTH = N/8");
  constexpr int S = fft_last_stride(N);
  constexpr int RAD = N / S;
  constexpr int NB = N / RAD;          // = S
  constexpr int BPT = NB / NTH;
  asm volatile("" : "+v"(tid));
  cx<T> a[BPT][RAD];
#pragma unroll
  for (int u = 0; u < BPT; u++) {
    const cx<T>* zr = z + lds_index(tid + u * NTH);
#pragma unroll
    for (int k = 0; k < RAD; k++) a[u][k] = zr[lds_step(k * NB)];
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < BPT; u++) {
    Dft<T, RAD, SIGN>::run(a[u]);
    // output q + S k with q = tid + u NTH: element tid + NTH (u + BPT k)
#pragma unroll
    for (int k = 0; k < RAD; k++) out[u + BPT * k] = a[u][k];
  }
}

// registers -> z (stored like fft_lds; NAT_OUT: unpadded natural order)
template <typename T, int N, int NTH, int SIGN, bool HALF, bool NAT_OUT = false>
__device__ __forceinline__ void fft_regs_to_lds(const cx<T>* in, cx<T>* z, const cx<T>* __restrict__ tw, int tid) {
  fft_first_pass_regs<T, N, NTH, SIGN, HALF>(in, z, tw, tid);
  FftMidPasses<T, N, NTH, SIGN, 8>::run(z, tw, tid);
  constexpr int S = fft_last_stride(N);
  stockham_pass<T, N, NTH, SIGN, N / S, S, NAT_OUT>(z, tw, tid);
}
// registers -> registers, z as the exchange buffer (free on return)
template <typename T, int N, int NTH, int SIGN, bool HALF>
__device__ __forceinline__ void fft_regs_to_regs(const cx<T>* in, cx<T>* out, cx<T>* z, const cx<T>* __restrict__ tw,
                                                 int tid) {
  fft_first_pass_regs<T, N, NTH, SIGN, HALF>(in, z, tw, tid);
  FftMidPasses<T, N, NTH, SIGN, 8>::run(z, tw, tid);
  fft_last_pass_regs<T, N, NTH, SIGN>(z, out, tid);
}

}  // namespace fil
