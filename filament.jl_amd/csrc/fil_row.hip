// fil_row.hip -- the row phase of one z-step: everything that is local to one
// r row of the field, fused into one workgroup per row on gfx950.
//
// Per row (src/modules/Models.jl:75-121 restated for one r):
//   (pending field filter of the previous step: t-transform back, multiply by
//    Rguard(r)*Tguard(t), forward again -- Models.jl:109-119, Guards.jl:265-280)
//   plasma density along t                      plasma_equations.jl:27-72
//   RK3 in z of the nonlinear spectral RHS       nonlinear_propagators.jl:38-74
//     (responses, field filter, analytic spectrum, update_dE!, i*QZ)
// The row's spectrum stays in LDS for the whole step; HBM sees one read and
// one write of the Nt/2 positive-frequency bins.
#include "fil_internal.h"
#include "fil_fft.h"
#include "fil_kernels.h"
#include "fil_rowdev.h"

namespace fil {

// ---------------------------------------------------------------------------
// The row kernel.  One workgroup per r row; LDS: Z (N complex), rho, kd (N R).
// ---------------------------------------------------------------------------
template <typename R, typename A, int N, int NTH, bool RAMAN>
__global__ __launch_bounds__(NTH) void row_kernel(RowArgs<R> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  cx<R>* Z = reinterpret_cast<cx<R>*>(smem);
  R* rho = reinterpret_cast<R*>(Z + lds_size(N));
  R* kd = rho + rk_size<R>(N);
  A* red = reinterpret_cast<A*>(kd + rk_size<R>(N));
  constexpr int NH = N / 2;
  constexpr int SP = NH / NTH;
  const int tid = threadIdx.x;
  const int i = blockIdx.x;
  cx<R>* Srow = a.S + (size_t)i * NH;
  const R g_r = a.Rg[i / a.batch];

  cx<R> S0[SP];
#pragma unroll
  for (int m = 0; m < SP; m++) S0[m] = Srow[tid + m * NTH];

  constexpr bool REGS = fft_regs_ok<N, NTH>();
  constexpr int TP = N / NTH;
  cx<R> e[TP];   // time-domain field at t = tid + u NTH
  // A corrector output still to be added (window rows of an FP32 single-pulse
  // handle): D is interpolated along r once per row into Dy [2][W] (LDS, in
  // the rho area, which the plasma scan fills only later; the transform's
  // barriers order these stores before the reads), then along t per sample.
  bool corr_row = false;
  if constexpr (sizeof(R) == 4) corr_row = a.corr_D != nullptr && i < a.corr_g.nr;
  float* Dy = reinterpret_cast<float*>(rho);
  if constexpr (sizeof(R) == 4) {
    if (corr_row) {
      const CorrGeom& g = a.corr_g;
      const float y = fminf(fmaxf(((float)i + 0.5f) / (float)g.pr - 0.5f, 0.0f), (float)(g.H - 1));
      const int y0 = (int)y, y1 = min(y0 + 1, g.H - 1);
      const float fy = y - (float)y0;
      for (int q = tid; q < 2 * g.W; q += NTH) {
        const float* P = a.corr_D + (q / g.W) * g.H * g.W + q % g.W;
        const float d0 = P[y0 * g.W], d1 = P[y1 * g.W];
        Dy[q] = d0 + fy * (d1 - d0);
      }
    }
  }
  auto corr_at = [&](int t, cx<R> v) {
    if constexpr (sizeof(R) == 4) {
      const CorrGeom& g = a.corr_g;
      if (corr_row && t >= g.t0 && t < g.t0 + g.nt) {
        const float x = fminf(fmaxf(((float)(t - g.t0) + 0.5f) / (float)g.pt - 0.5f, 0.0f), (float)(g.W - 1));
        const int x0 = (int)x, x1 = min(x0 + 1, g.W - 1);
        const float fx = x - (float)x0;
        const cx<float> d = mk(a.corr_scale * (Dy[x0] + fx * (Dy[x1] - Dy[x0])),
                               a.corr_scale * (Dy[g.W + x0] + fx * (Dy[g.W + x1] - Dy[g.W + x0])));
        v = v + d * conj(g.carrier[t]);
      }
    }
    return v;
  };
  if (a.apply_filter || corr_row) {
    // field filter deferred from the end of the previous zstep (and the
    // corrector's addition, which comes before it)
    const R inv = (R)1 / (R)N;
    const R gr = a.apply_filter ? g_r : (R)1;
    if constexpr (REGS) {
      fft_regs_to_regs<R, N, NTH, -1, true>(S0, e, Z, a.tw, tid);
#pragma unroll
      for (int u = 0; u < TP; u++) {
        const int t = tid + u * NTH;
        e[u] = (gr * (a.apply_filter ? a.Tg[t] : (R)1)) * corr_at(t, e[u]);
      }
      fft_regs_to_regs<R, N, NTH, +1, false>(e, e, Z, a.tw, tid);
#pragma unroll
      for (int m = 0; m < SP; m++) S0[m] = inv * e[m];
    } else {
      store_padded<R, N, NTH>(Z, S0);
      fft_lds<R, N, NTH, -1>(Z, a.tw);
      for (int t = tid; t < N; t += NTH)
        Z[lds_index(t)] = (gr * (a.apply_filter ? a.Tg[t] : (R)1)) * corr_at(t, Z[lds_index(t)]);
      __syncthreads();
      fft_lds<R, N, NTH, +1>(Z, a.tw);
#pragma unroll
      for (int m = 0; m < SP; m++) S0[m] = inv * Z[lds_index(tid + m * NTH)];
      __syncthreads();
    }
  }
  // time-domain field of the step's starting spectrum
  if constexpr (REGS) {
    fft_regs_to_lds<R, N, NTH, -1, true>(S0, Z, a.tw, tid);
  } else {
    store_padded<R, N, NTH>(Z, S0);
    fft_lds<R, N, NTH, -1>(Z, a.tw);
  }

  // stage 1 of the RK3 starts from this field; the plasma scan then takes Z
  // as its cache (its barriers order these reads before its writes)
#pragma unroll
  for (int u = 0; u < TP; u++) e[u] = Z[lds_index(tid + u * NTH)];
  if (a.plasma) {
    plasma_row<R, A, N, NTH, true>(Z, rho, kd, a.pp, a.dt, red);
    R mx = 0;
    for (int t = tid; t < N; t += NTH) mx = max(mx, rho[rk_index<R>(t)]);
    mx = block_max<R>(mx, reinterpret_cast<R*>(red));
    if (tid == 0) { a.rho_end[i] = (double)rho[rk_index<R>(N - 1)]; a.rho_max[i] = (double)mx; }
    if (a.rho_full) {
      for (int t = tid; t < N; t += NTH) {
        a.rho_full[(size_t)i * N + t] = rho[rk_index<R>(t)];
        a.kd_full[(size_t)i * N + t] = kd[rk_index<R>(t)];
      }
    }
  } else {
    for (int t = tid; t < N; t += NTH) { rho[rk_index<R>(t)] = 0; kd[rk_index<R>(t)] = 0; }
    __syncthreads();
  }

  if (!a.nonlinearity) {
    if (a.apply_filter || corr_row) {
#pragma unroll
      for (int m = 0; m < SP; m++) Srow[tid + m * NTH] = S0[m];
    }
    return;
  }

  // RK3 (Kutta) in z: k1 = f(S0); k2 = f(S0 + dz/2 k1); k3 = f(S0 - dz k1 + 2 dz k2)
  const R dz = a.dz_dev ? (R)*a.dz_dev : a.dz;
  cx<R> k1[SP], k2[SP], k[SP], U[SP];
  stage_rhs<R, N, NTH, RAMAN>(Z, e, rho, kd, a.tw, a.coef, a.hr, a.sc, g_r, a.Tg, a.responses, a.losses_ereal, k1);
#pragma unroll
  for (int m = 0; m < SP; m++) U[m] = S0[m] + ((R)0.5 * dz) * k1[m];
  stage_field<R, N, NTH>(U, e, Z, a.tw);
  stage_rhs<R, N, NTH, RAMAN>(Z, e, rho, kd, a.tw, a.coef, a.hr, a.sc, g_r, a.Tg, a.responses, a.losses_ereal, k2);
#pragma unroll
  for (int m = 0; m < SP; m++) U[m] = S0[m] + dz * (k2[m] + k2[m] - k1[m]);
  stage_field<R, N, NTH>(U, e, Z, a.tw);
  stage_rhs<R, N, NTH, RAMAN>(Z, e, rho, kd, a.tw, a.coef, a.hr, a.sc, g_r, a.Tg, a.responses, a.losses_ereal, k);
  const R c16 = (R)1 / (R)6, c23 = (R)2 / (R)3;
#pragma unroll
  for (int m = 0; m < SP; m++) {
    cx<R> acc = c16 * k1[m] + c23 * k2[m] + c16 * k[m];
    Srow[tid + m * NTH] = S0[m] + dz * acc;
  }
}

// ---------------------------------------------------------------------------
// Upload: time-domain rows [Nr][N] -> positive-frequency spectrum S [Nr][N/2]
// (forward_transform_time!, Models.jl:162-165).
// ---------------------------------------------------------------------------
template <typename R, int N, int NTH>
__global__ __launch_bounds__(NTH) void to_spectrum_kernel(const cx<R>* __restrict__ Et, cx<R>* S, const cx<R>* tw,
                                                          int rs, int ro) {
  extern __shared__ __align__(16) unsigned char smem[];
  cx<R>* Z = reinterpret_cast<cx<R>*>(smem);
  const int i = blockIdx.x, tid = threadIdx.x;
  {
    // all of the thread's loads in flight before its first LDS write
    cx<R> v[N / NTH];
#pragma unroll
    for (int u = 0; u < N / NTH; u++) v[u] = Et[(size_t)i * N + tid + u * NTH];
#pragma unroll
    for (int u = 0; u < N / NTH; u++) Z[lds_index(tid + u * NTH)] = v[u];
  }
  __syncthreads();
  fft_lds<R, N, NTH, +1>(Z, tw);
  const R inv = (R)1 / (R)N;
  for (int w = tid; w < N / 2; w += NTH) S[((size_t)i * rs + ro) * (N / 2) + w] = inv * Z[lds_index(w)];
}

// Download: spectrum -> time rows, with the pending field filter.
template <typename R, int N, int NTH>
__global__ __launch_bounds__(NTH) void to_time_kernel(const cx<R>* __restrict__ S, cx<R>* Et, const cx<R>* tw,
                                                      const R* Rg, const R* Tg, int apply_filter, int rs, int ro) {
  extern __shared__ __align__(16) unsigned char smem[];
  cx<R>* Z = reinterpret_cast<cx<R>*>(smem);
  const int i = blockIdx.x, tid = threadIdx.x;
  {
    cx<R> v[N / 2 / NTH];
#pragma unroll
    for (int m = 0; m < N / 2 / NTH; m++) v[m] = S[((size_t)i * rs + ro) * (N / 2) + tid + m * NTH];
#pragma unroll
    for (int m = 0; m < N / 2 / NTH; m++) {
      Z[lds_index(tid + m * NTH)] = v[m];
      Z[lds_index(N / 2 + tid + m * NTH)] = mk((R)0, (R)0);
    }
  }
  __syncthreads();
  fft_lds<R, N, NTH, -1>(Z, tw);
  const R g = apply_filter ? Rg[i] : (R)1;
  for (int t = tid; t < N; t += NTH) {
    cx<R> v = Z[lds_index(t)];
    if (apply_filter) v = (g * Tg[t]) * v;
    Et[(size_t)i * N + t] = v;
  }
}

// ---------------------------------------------------------------------------
// FieldAnalyzers.analyze! RT, per-row part (FieldAnalyzers.jl:166-215):
// Fr(r) = dt sum_t |E|^2; It[r][t] = |E|^2 (for Ft); the projected spectrum
// replaces the state (aspec2rspec!/rspec2aspec! round trip, :188-197);
// Sw[r][w] = |2 Nt dt rspec(w)|^2 for w <= Nt/2; Imax of the projected field.
// ---------------------------------------------------------------------------
template <typename R, int N, int NTH>
__global__ __launch_bounds__(NTH) void analyze_rows_kernel(cx<R>* S, const cx<R>* tw, const R* Rg, const R* Tg,
                                                           int apply_filter, R dt, double* Fr, R* It,
                                                           R* Sw, double* Imax_row, int rs, int ro, int gdiv) {
  extern __shared__ __align__(16) unsigned char smem[];
  cx<R>* Z = reinterpret_cast<cx<R>*>(smem);
  double* red = reinterpret_cast<double*>(Z + lds_size(N));
  const int i = blockIdx.x, tid = threadIdx.x;
  constexpr int NH = N / 2;
  cx<R>* Srow = S + ((size_t)i * rs + ro) * NH;
  {
    // all of the thread's loads in flight before its first LDS write
    cx<R> v[NH / NTH];
#pragma unroll
    for (int m = 0; m < NH / NTH; m++) v[m] = Srow[tid + m * NTH];
#pragma unroll
    for (int m = 0; m < NH / NTH; m++) {
      Z[lds_index(tid + m * NTH)] = v[m];
      Z[lds_index(NH + tid + m * NTH)] = mk((R)0, (R)0);
    }
  }
  __syncthreads();
  fft_lds<R, N, NTH, -1>(Z, tw);
  const R g = apply_filter ? Rg[i / gdiv] : (R)1;
  double fr = 0;
  for (int t = tid; t < N; t += NTH) {
    cx<R> v = Z[lds_index(t)];
    if (apply_filter) { v = (g * Tg[t]) * v; Z[lds_index(t)] = v; }
    double I = (double)norm2(v);
    fr += I;
    if (It) It[(size_t)i * N + t] = (R)I;
  }
  // block sum of fr
  const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) fr += __shfl_xor(fr, off, 64);
  if (lane == 0) red[wv] = fr;
  __syncthreads();
  if (tid == 0) { double s = 0; for (int w = 0; w < NTH / 64; w++) s += red[w]; Fr[i] = s * (double)dt; }
  __syncthreads();
  fft_lds<R, N, NTH, +1>(Z, tw);
  const R inv = (R)1 / (R)N;
  const double c = 2.0 * (double)N * (double)dt;
  for (int w = tid; Sw && w <= NH; w += NTH) {
    cx<R> v = inv * Z[lds_index(w)];
    double re = (double)v.x, im = (double)v.y;
    if (w >= 1 && w < NH) { re *= 0.5; im *= 0.5; }   // aspec2rspec
    Sw[(size_t)i * (NH + 1) + w] = (R)(c * c * (re * re + im * im));
  }
  __syncthreads();
  // projection: keep bins 0..N/2-1 as the new state
  cx<R> keep[NH / NTH];
#pragma unroll
  for (int m = 0; m < NH / NTH; m++) keep[m] = inv * Z[lds_index(tid + m * NTH)];
  __syncthreads();
#pragma unroll
  for (int m = 0; m < NH / NTH; m++) {
    Srow[tid + m * NTH] = keep[m];
    Z[lds_index(tid + m * NTH)] = keep[m];
    Z[lds_index(NH + tid + m * NTH)] = mk((R)0, (R)0);
  }
  __syncthreads();
  fft_lds<R, N, NTH, -1>(Z, tw);
  // maximum(abs2, E) propagates NaN (FieldAnalyzers.jl:201), so does this reduction
  double mx = 0;
  for (int t = tid; t < N; t += NTH) mx = nan_max(mx, (double)norm2(Z[lds_index(t)]));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = nan_max(mx, __shfl_xor(mx, off, 64));
  if (lane == 0) red[wv] = mx;
  __syncthreads();
  if (tid == 0) { double m = 0; for (int w = 0; w < NTH / 64; w++) m = nan_max(m, red[w]); Imax_row[i] = m; }
}

// Diagonal linear step of the T geometry: S[i][w] *= P[w] for every row.
template <typename R>
__global__ void scale_rows_kernel(cx<R>* S, const cx<R>* __restrict__ P, int Nh, size_t n) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) S[q] = S[q] * P[q % Nh];
}

template <typename R>
hipError_t launch_scale_rows(cx<R>* S, const cx<R>* P, int rows, int Nh, hipStream_t st) {
  const size_t n = (size_t)rows * Nh;
  fil_launch(scale_rows_kernel<R>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, S, P, Nh, n);
  return hipGetLastError();
}
template hipError_t launch_scale_rows<float>(cx<float>*, const cx<float>*, int, int, hipStream_t);
template hipError_t launch_scale_rows<double>(cx<double>*, const cx<double>*, int, int, hipStream_t);

// Parseval fluence of each row from its half-spectrum: out[i] = scale * sum_w |S[i][w]|^2.
template <typename R>
__global__ __launch_bounds__(256) void rowsum_abs2_kernel(const cx<R>* __restrict__ S, int Nh, double scale,
                                                          double* out) {
  __shared__ double red[4];
  const int i = blockIdx.x, tid = threadIdx.x;
  double s = 0;
  for (int w = tid; w < Nh; w += 256) s += (double)norm2(S[(size_t)i * Nh + w]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((tid & 63) == 0) red[tid >> 6] = s;
  __syncthreads();
  if (tid == 0) out[i] = scale * (red[0] + red[1] + red[2] + red[3]);
}

template <typename R>
hipError_t launch_rowsum_abs2(const cx<R>* S, int Nr, int Nh, double scale, double* out, hipStream_t st) {
  hipLaunchKernelGGL(rowsum_abs2_kernel<R>, dim3(Nr), dim3(256), 0, st, S, Nh, scale, out);
  return hipGetLastError();
}
template hipError_t launch_rowsum_abs2<float>(const cx<float>*, int, int, double, double*, hipStream_t);
template hipError_t launch_rowsum_abs2<double>(const cx<double>*, int, int, double, double*, hipStream_t);

// Column sums: out[j] = sum_i in[i][j] * wt[i] in FP64, in a fixed order.
template <typename T>
__global__ void colsum_part_kernel(const T* __restrict__ in, int Nr, int ncol, const double* __restrict__ wt,
                                   int rows, double* __restrict__ part) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ncol) return;
  const int i0 = blockIdx.y * rows, i1 = min(Nr, i0 + rows);
  double s = 0;
  // eight loads in flight, added in the same row order as one at a time
  int i = i0;
  for (; i + 8 <= i1; i += 8) {
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = in[(size_t)(i + u) * ncol + j];
#pragma unroll
    for (int u = 0; u < 8; u++) s += (double)v[u] * wt[i + u];
  }
  for (; i < i1; i++) s += (double)in[(size_t)i * ncol + j] * wt[i];
  part[(size_t)blockIdx.y * ncol + j] = s;
}

__global__ void colsum_final_kernel(const double* __restrict__ part, int nchunk, int ncol, double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ncol) return;
  double s = 0;
  // eight loads in flight, added in the same chunk order as one at a time
  int c = 0;
  for (; c + 8 <= nchunk; c += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) v[u] = part[(size_t)(c + u) * ncol + j];
#pragma unroll
    for (int u = 0; u < 8; u++) s += v[u];
  }
  for (; c < nchunk; c++) s += part[(size_t)c * ncol + j];
  out[j] = s;
}

// ---------------------------------------------------------------------------
// Learned corrector, field side.  Window rows i < g.nr.
// Features: time-domain field of the row, its plasma density (FP32 scan),
// then block averages over pt samples of A = E exp(i w0 t), |E|^2 and rho.
// The carrier exp(i w0 t_j) is a table rounded from FP64: a phase computed in
// FP32 from t_j would be off by ~1e-5 rad at |t| ~ 100 carrier periods.
// ---------------------------------------------------------------------------
template <int N, int NTH>
__global__ __launch_bounds__(NTH) void corr_features_kernel(const cx<float>* __restrict__ S, const cx<float>* tw,
                                                            PlasmaParams<float> pp, int plasma, CorrGeom g,
                                                            float* Frow) {
  extern __shared__ __align__(16) unsigned char smem[];
  cx<float>* Z = reinterpret_cast<cx<float>*>(smem);
  float* rho = reinterpret_cast<float*>(Z + lds_size(N));
  float* kd = rho + rk_size<float>(N);
  float* red = kd + rk_size<float>(N);
  constexpr int NH = N / 2, SP = NH / NTH;
  const int tid = threadIdx.x, i = blockIdx.x;
  cx<float> v[SP];
#pragma unroll
  for (int m = 0; m < SP; m++) v[m] = S[(size_t)i * NH + tid + m * NTH];
  if constexpr (fft_regs_ok<N, NTH>()) {
    fft_regs_to_lds<float, N, NTH, -1, true>(v, Z, tw, tid);
  } else {
    store_padded<float, N, NTH>(Z, v);
    fft_lds<float, N, NTH, -1>(Z, tw);
  }
  if (plasma) plasma_row<float, float, N, NTH>(Z, rho, kd, pp, g.dt, red);
  float* out = Frow + (size_t)i * 4 * g.W;
  const float inv = 1.0f / (float)g.pt;
  // tpb consecutive lanes share a block (a power of two dividing pt, at most
  // NTH / W): each sums pt / tpb consecutive samples, then the group adds its
  // partial sums by lane exchanges
  int tpb = 1;
  while (tpb < 64 && 2 * tpb * g.W <= NTH && g.pt % (2 * tpb) == 0) tpb *= 2;
  const int per = g.pt / tpb;
  for (int q = tid; q < g.W * tpb; q += NTH) {
    const int b = q / tpb, s0 = (q % tpb) * per;
    float sr = 0, si = 0, sI = 0, sn = 0;
    for (int s = s0; s < s0 + per; s++) {
      const int j = g.t0 + b * g.pt + s;
      const cx<float> e = Z[lds_index(j)];
      const cx<float> a = e * g.carrier[j];
      sr += a.x;
      si += a.y;
      sI += norm2(e);
      if (plasma) sn += rho[rk_index<float>(j)];
    }
    for (int off = 1; off < tpb; off <<= 1) {
      sr += __shfl_xor(sr, off, 64);
      si += __shfl_xor(si, off, 64);
      sI += __shfl_xor(sI, off, 64);
      sn += __shfl_xor(sn, off, 64);
    }
    if (q % tpb == 0) {
      out[b] = sr * inv * g.inv_as;
      out[g.W + b] = si * inv * g.inv_as;
      out[2 * g.W + b] = sI * inv * g.inv_as * g.inv_as;
      out[3 * g.W + b] = sn * inv * g.inv_rs;
    }
  }
}

// X[c][h][b] = mean over the pr rows of block h of Frow[row][c][b]
__global__ void corr_pool_rows_kernel(const float* __restrict__ Frow, CorrGeom g, float* X) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int HW = g.H * g.W;
  if (q >= 4 * HW) return;
  const int c = q / HW, rem = q % HW, h = rem / g.W, b = rem % g.W;
  float s = 0;
  for (int k = 0; k < g.pr; k++) s += Frow[((size_t)(h * g.pr + k) * 4 + c) * g.W + b];
  X[q] = s / (float)g.pr;
}

template <int N, int NTH>
__global__ __launch_bounds__(NTH) void corr_apply_kernel(cx<float>* S, const cx<float>* tw, CorrGeom g,
                                                         const float* __restrict__ D, float scale) {
  extern __shared__ __align__(16) unsigned char smem[];
  cx<float>* Z = reinterpret_cast<cx<float>*>(smem);
  constexpr int NH = N / 2, SP = NH / NTH;
  const int tid = threadIdx.x, i = blockIdx.x;
  cx<float> v[SP];
#pragma unroll
  for (int m = 0; m < SP; m++) v[m] = S[(size_t)i * NH + tid + m * NTH];
  auto corr = [&](int j) { return corr_value(g, D, scale, i, j); };
  const float inv = 1.0f / (float)N;
  if constexpr (fft_regs_ok<N, NTH>()) {
    constexpr int TP = N / NTH;
    cx<float> e[TP];
    fft_regs_to_regs<float, N, NTH, -1, true>(v, e, Z, tw, tid);
#pragma unroll
    for (int u = 0; u < TP; u++) {
      const int j = tid + u * NTH;
      if (j >= g.t0 && j < g.t0 + g.nt) e[u] = e[u] + corr(j);
    }
    fft_regs_to_regs<float, N, NTH, +1, false>(e, e, Z, tw, tid);
#pragma unroll
    for (int m = 0; m < SP; m++) S[(size_t)i * NH + tid + m * NTH] = inv * e[m];
  } else {
    store_padded<float, N, NTH>(Z, v);
    fft_lds<float, N, NTH, -1>(Z, tw);
    for (int j = g.t0 + tid; j < g.t0 + g.nt; j += NTH) Z[lds_index(j)] = Z[lds_index(j)] + corr(j);
    __syncthreads();
    fft_lds<float, N, NTH, +1>(Z, tw);
#pragma unroll
    for (int m = 0; m < SP; m++) S[(size_t)i * NH + tid + m * NTH] = inv * Z[lds_index(tid + m * NTH)];
  }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
// one radix-FIL_FFT_MAXRAD butterfly per thread per FFT pass
template <int N> constexpr int threads_for() {
  return N / FIL_FFT_MAXRAD < 64 ? 64 : (N / FIL_FFT_MAXRAD > 1024 ? 1024 : N / FIL_FFT_MAXRAD);
}

template <typename R, typename A>
static size_t row_smem(int N, int nth) {
  return (size_t)lds_size(N) * sizeof(cx<R>) + (size_t)rk_size<R>(N) * 2 * sizeof(R) + 2 * (nth / 64) * sizeof(A) + 64;
}

template <typename R, typename A, int N>
static hipError_t launch_row_n(const RowArgs<R>& a, int Nr, hipStream_t st) {
  constexpr int NTH = threads_for<N>();
  size_t sm = row_smem<R, A>(N, NTH);
  const bool raman = a.nonlinearity && (a.responses & (FIL_RESP_RAMAN_THG | FIL_RESP_RAMAN_NOTHG));
  if (raman && !a.hr) return hipErrorInvalidValue;
  auto k = raman ? row_kernel<R, A, N, NTH, true> : row_kernel<R, A, N, NTH, false>;
  hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
  if (e != hipSuccess) return e;
  fil_launch(k, dim3(Nr), dim3(NTH), sm, st, a);
  return hipGetLastError();
}

#define FIL_DISPATCH_N(N_, ...)                                  \
  switch (N_) {                                                  \
    case 256: { constexpr int NN = 256; __VA_ARGS__; } break;    \
    case 512: { constexpr int NN = 512; __VA_ARGS__; } break;    \
    case 1024: { constexpr int NN = 1024; __VA_ARGS__; } break;  \
    case 2048: { constexpr int NN = 2048; __VA_ARGS__; } break;  \
    case 4096: { constexpr int NN = 4096; __VA_ARGS__; } break;  \
    case 8192: { constexpr int NN = 8192; __VA_ARGS__; } break;  \
    default: return hipErrorInvalidValue;                        \
  }

template <typename R>
hipError_t launch_row(const RowArgs<R>& a, int Nr, int N, bool scan64, hipStream_t st) {
  if (scan64 || sizeof(R) == 8) {
    FIL_DISPATCH_N(N, return (launch_row_n<R, double, NN>(a, Nr, st)));
  } else {
    FIL_DISPATCH_N(N, return (launch_row_n<R, float, NN>(a, Nr, st)));
  }
  return hipSuccess;
}

template <typename R>
hipError_t launch_to_spectrum(const cx<R>* Et, cx<R>* S, const cx<R>* tw, int Nr, int N, int rs, int ro, hipStream_t st) {
  FIL_DISPATCH_N(N, {
    constexpr int NTH = threads_for<NN>();
    size_t sm = (size_t)lds_size(NN) * sizeof(cx<R>);
    auto k = to_spectrum_kernel<R, NN, NTH>;
    { hipError_t e0 = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm); if (e0 != hipSuccess) return e0; }
    hipLaunchKernelGGL(k, dim3(Nr), dim3(NTH), sm, st, Et, S, tw, rs, ro);
    return hipGetLastError();
  });
  return hipSuccess;
}

template <typename R>
hipError_t launch_to_time(const cx<R>* S, cx<R>* Et, const cx<R>* tw, const R* Rg, const R* Tg, int apply,
                          int Nr, int N, int rs, int ro, hipStream_t st) {
  FIL_DISPATCH_N(N, {
    constexpr int NTH = threads_for<NN>();
    size_t sm = (size_t)lds_size(NN) * sizeof(cx<R>);
    auto k = to_time_kernel<R, NN, NTH>;
    { hipError_t e0 = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm); if (e0 != hipSuccess) return e0; }
    hipLaunchKernelGGL(k, dim3(Nr), dim3(NTH), sm, st, S, Et, tw, Rg, Tg, apply, rs, ro);
    return hipGetLastError();
  });
  return hipSuccess;
}

template <typename R>
hipError_t launch_analyze_rows(cx<R>* S, const cx<R>* tw, const R* Rg, const R* Tg, int apply, R dt, double* Fr,
                               R* It, R* Sw, double* Imax_row, int Nr, int N, int rs, int ro, int gdiv,
                               hipStream_t st) {
  FIL_DISPATCH_N(N, {
    constexpr int NTH = threads_for<NN>();
    size_t sm = (size_t)lds_size(NN) * sizeof(cx<R>) + 16 * sizeof(double);
    auto k = analyze_rows_kernel<R, NN, NTH>;
    { hipError_t e0 = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm); if (e0 != hipSuccess) return e0; }
    hipLaunchKernelGGL(k, dim3(Nr), dim3(NTH), sm, st, S, tw, Rg, Tg, apply, dt, Fr, It, Sw, Imax_row, rs, ro, gdiv);
    return hipGetLastError();
  });
  return hipSuccess;
}

template <typename T>
hipError_t launch_colsum(const T* in, int Nr, int ncol, const double* wt, double* part, double* out,
                         hipStream_t st) {
  int nchunk = (Nr + 63) / 64;
  if (nchunk > COLSUM_MAXCHUNK) nchunk = COLSUM_MAXCHUNK;
  const int rows = (Nr + nchunk - 1) / nchunk;
  hipLaunchKernelGGL(colsum_part_kernel<T>, dim3((ncol + 255) / 256, nchunk), dim3(256), 0, st, in, Nr, ncol, wt, rows,
                     part);
  hipLaunchKernelGGL(colsum_final_kernel, dim3((ncol + 255) / 256), dim3(256), 0, st, part, nchunk, ncol, out);
  return hipGetLastError();
}

hipError_t launch_corr_features(const cx<float>* S, const cx<float>* tw, const PlasmaParams<float>& pp, int plasma,
                                const CorrGeom& g, int N, float* Frow, float* X, hipStream_t st) {
  FIL_DISPATCH_N(N, {
    constexpr int NTH = threads_for<NN>();
    size_t sm = row_smem<float, float>(NN, NTH);
    auto k = corr_features_kernel<NN, NTH>;
    { hipError_t e0 = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm); if (e0 != hipSuccess) return e0; }
    hipLaunchKernelGGL(k, dim3(g.nr), dim3(NTH), sm, st, S, tw, pp, plasma, g, Frow);
    hipError_t e1 = hipGetLastError();
    if (e1 != hipSuccess) return e1;
  });
  if (!X) return hipSuccess;
  const int n = 4 * g.H * g.W;
  hipLaunchKernelGGL(corr_pool_rows_kernel, dim3((n + 255) / 256), dim3(256), 0, st, Frow, g, X);
  return hipGetLastError();
}

hipError_t launch_corr_apply(cx<float>* S, const cx<float>* tw, const CorrGeom& g, int N, const float* D, float scale,
                             hipStream_t st) {
  FIL_DISPATCH_N(N, {
    constexpr int NTH = threads_for<NN>();
    size_t sm = (size_t)lds_size(NN) * sizeof(cx<float>);
    auto k = corr_apply_kernel<NN, NTH>;
    { hipError_t e0 = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm); if (e0 != hipSuccess) return e0; }
    hipLaunchKernelGGL(k, dim3(g.nr), dim3(NTH), sm, st, S, tw, g, D, scale);
    return hipGetLastError();
  });
  return hipSuccess;
}

// Max over rows with the analyzer's fold (start at 0, a NaN row never wins):
// per thread, across the wave by shuffles, then across the waves in LDS.
constexpr int ADAPT_THREADS = 1024;
__global__ __launch_bounds__(ADAPT_THREADS) void adapt_kernel(AdaptRule rule, AdaptState* state,
                                                              const double* Imax_row, const double* rhomax_row,
                                                              double* log) {
  __shared__ double red[2][ADAPT_THREADS / 64];
  if (state->done) return;
  double im = 0, rm = 0;
  for (int i = threadIdx.x; i < rule.nrows; i += ADAPT_THREADS) {
    im = nan_max(im, Imax_row[i]);
    if (rule.plasma) rm = nan_max(rm, rhomax_row[i]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    im = nan_max(im, __shfl_xor(im, o));
    rm = nan_max(rm, __shfl_xor(rm, o));
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = im;
    red[1][threadIdx.x >> 6] = rm;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double Imax = 0, rhomax = 0;
  for (int w = 0; w < ADAPT_THREADS / 64; w++) {
    Imax = nan_max(Imax, red[0][w]);
    rhomax = nan_max(rhomax, red[1][w]);
  }
  const int s = state->steps;
  log[4 * s + 2] = Imax;
  log[4 * s + 3] = rhomax;
  // Filament.jl:160-164: stop after the step whose analysis exceeds Istop;
  // Filament.jl:129: loop while z < zmax
  int done = 0;
  // a NaN analysis makes the reference's dz and z NaN, which ends `while z < zmax`
  if (s > 0 && Imax > rule.Istop) done = FIL_STOP_ISTOP;
  else if (Imax != Imax || rhomax != rhomax) done = FIL_STOP_NONFINITE;
  else if (!(state->z < rule.zmax)) done = FIL_STOP_ZMAX;
  else if (s >= rule.max_steps) done = FIL_STOP_MAXSTEPS;
  if (done) {
    state->done = done;
    return;
  }
  // AdaptiveSteps.jl:96-107, min(dz0, dzk, dzp) in FP64, then single precision
  double dz = rule.dz0;
  if (rule.nonlinearity) {
    const double pk = rule.phik * Imax, pp = rule.phip * rhomax;
    const double dzk = pk != 0 ? rule.phimax / pk : INFINITY;
    const double dzp = pp != 0 ? rule.phimax / pp : INFINITY;
    if (dzk < dz) dz = dzk;
    if (dzp < dz) dz = dzp;
  }
  const float dzf = (float)dz;
  const float zn = state->z + dzf;
  state->dz = (double)dzf;
  state->z = zn;
  state->steps = s + 1;
  log[4 * (s + 1) + 0] = (double)zn;
  log[4 * (s + 1) + 1] = (double)dzf;
}

hipError_t launch_adapt(const AdaptRule& rule, AdaptState* state, const double* Imax_row, const double* rhomax_row,
                        double* log, hipStream_t st) {
  hipLaunchKernelGGL(adapt_kernel, dim3(1), dim3(ADAPT_THREADS), 0, st, rule, state, Imax_row, rhomax_row, log);
  return hipGetLastError();
}

template hipError_t launch_row<float>(const RowArgs<float>&, int, int, bool, hipStream_t);
template hipError_t launch_row<double>(const RowArgs<double>&, int, int, bool, hipStream_t);
template hipError_t launch_to_spectrum<float>(const cx<float>*, cx<float>*, const cx<float>*, int, int, int, int, hipStream_t);
template hipError_t launch_to_spectrum<double>(const cx<double>*, cx<double>*, const cx<double>*, int, int, int, int, hipStream_t);
template hipError_t launch_to_time<float>(const cx<float>*, cx<float>*, const cx<float>*, const float*, const float*, int, int, int, int, int, hipStream_t);
template hipError_t launch_to_time<double>(const cx<double>*, cx<double>*, const cx<double>*, const double*, const double*, int, int, int, int, int, hipStream_t);
template hipError_t launch_analyze_rows<float>(cx<float>*, const cx<float>*, const float*, const float*, int, float, double*, float*, float*, double*, int, int, int, int, int, hipStream_t);
template hipError_t launch_analyze_rows<double>(cx<double>*, const cx<double>*, const double*, const double*, int, double, double*, double*, double*, double*, int, int, int, int, int, hipStream_t);
template hipError_t launch_colsum<float>(const float*, int, int, const double*, double*, double*, hipStream_t);
template hipError_t launch_colsum<double>(const double*, int, int, const double*, double*, double*, hipStream_t);

}  // namespace fil
