// fil_fft.h -- workgroup-wide FFT of one t-row held in LDS (gfx950).
//
// Replaces the cuFFT plan along t of the reference (plan_fft!(E, [2]),
// src/modules/Fields.jl:353; used at Models.jl:162-171, 180-183).
// Stockham decimation-in-frequency, radix-4 passes plus one radix-2 pass when
// log2(N) is odd.  Every pass reads its operands into registers, barriers,
// then writes the outputs back into the same LDS array, so one N-point
// buffer suffices.  No normalisation is applied.
#pragma once

#include "fil_internal.h"

namespace fil {

// SIGN = -1: y_k = sum_j x_j exp(-2 pi i jk/N)   (the reference's P * E)
// SIGN = +1: y_k = sum_j x_j exp(+2 pi i jk/N)   (N times the reference's P \ E)
// tw: exp(-2 pi i j / N), j < N, in global memory (cached).
template <typename R, int N, int NTH, int SIGN>
__device__ __forceinline__ void fft_lds(cx<R>* z, const cx<R>* __restrict__ tw) {
  static_assert((N & (N - 1)) == 0 && N >= 4, "N must be a power of two >= 4");
  constexpr int Q = N / 4;
  constexpr int BPT = (Q + NTH - 1) / NTH;
  const int tid = threadIdx.x;
  int s = 1;
#pragma unroll
  for (int n = N; n >= 4; n >>= 2) {
    cx<R> a[BPT][4];
#pragma unroll
    for (int u = 0; u < BPT; u++) {
      const int b = tid + u * NTH;
      if (Q % NTH == 0 || b < Q) {
#pragma unroll
        for (int k = 0; k < 4; k++) a[u][k] = z[b + k * Q];
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < BPT; u++) {
      const int b = tid + u * NTH;
      if (Q % NTH == 0 || b < Q) {
        const int q = b & (s - 1);
        const int p = b / s;
        cx<R> apc = a[u][0] + a[u][2], amc = a[u][0] - a[u][2];
        cx<R> bpd = a[u][1] + a[u][3], bmd = a[u][1] - a[u][3];
        cx<R> jb = (SIGN > 0) ? mul_i(bmd) : mul_mi(bmd);
        cx<R> y0 = apc + bpd, y1 = amc + jb, y2 = apc - bpd, y3 = amc - jb;
        const int e = p * s;
        if (e != 0) {
          cx<R> w1 = tw[e], w2 = tw[2 * e], w3 = tw[3 * e];
          if (SIGN > 0) { w1 = conj(w1); w2 = conj(w2); w3 = conj(w3); }
          y1 = y1 * w1; y2 = y2 * w2; y3 = y3 * w3;
        }
        const int o = q + 4 * s * p;
        z[o] = y0;
        z[o + s] = y1;
        z[o + 2 * s] = y2;
        z[o + 3 * s] = y3;
      }
    }
    __syncthreads();
    s <<= 2;
  }
  if (s == N / 2) {   // one radix-2 pass left (log2 N odd); in place per thread
    constexpr int H = N / 2;
#pragma unroll
    for (int u = 0; u < (H + NTH - 1) / NTH; u++) {
      const int b = tid + u * NTH;
      if (H % NTH == 0 || b < H) {
        cx<R> a0 = z[b], a1 = z[b + H];
        z[b] = a0 + a1;
        z[b + H] = a0 - a1;
      }
    }
    __syncthreads();
  }
}

}  // namespace fil
