// fil_nonpar.hip -- the nonlinear step with the non-paraxial prefactor
// (QPARAXIAL = false), RT geometry.
//
// func! then ends with a space transform (nonlinear_propagators.jl:58-63):
//   dE <- H^-1 [ i QZ(k, w) Kguard Wguard H[ sum_resp Rnl A[F_resp] ] ],
//   QZ(k, w) = mu0 mu w^2 / (2 kz), kz = sqrt(beta^2 - k^2 + 0im)  (:232-242),
// which couples the rows, so the RK3 step cannot stay inside one workgroup per
// row.  Each stage runs as a launch of np_stage_kernel (the paraxial
// right-hand side i QZ(w) Rnl A[F] of every row, the row kernel's stage_rhs)
// followed by the two folded Hankel products with the table
// Q(k, w) = QZ(k, w) / QZ(w) Kguard Wguard in the first product's epilogue
// (the same M2 (Q .* (M1 x)) as the linear step), and the stage combinations
// of Kutta's RK3 run as one elementwise pass each.
#include "fil_internal.h"
#include "fil_fft.h"
#include "fil_kernels.h"
#include "fil_rowdev.h"

namespace fil {

// One RK stage's paraxial right-hand side for every row.  first: the stage
// starts from S (pending field filter applied and written back, plasma solved
// from the filtered field and kept in rho_g / kd_g); otherwise from U, with the
// plasma of the first stage.  Rout [rows][N/2] receives i QZ(w) Rnl A[F].
template <typename R, typename A, int N, int NTH, bool RAMAN>
__global__ __launch_bounds__(NTH) void np_stage_kernel(RowArgs<R> a, const cx<R>* __restrict__ U,
                                                       cx<R>* __restrict__ Rout, R* __restrict__ rho_g,
                                                       R* __restrict__ kd_g, int first) {
  extern __shared__ __align__(16) unsigned char smem[];
  cx<R>* Z = reinterpret_cast<cx<R>*>(smem);
  R* rho = reinterpret_cast<R*>(Z + lds_size(N));
  R* kd = rho + rk_size<R>(N);
  A* red = reinterpret_cast<A*>(kd + rk_size<R>(N));
  constexpr int NH = N / 2;
  constexpr int SP = NH / NTH;
  constexpr int TP = N / NTH;
  const int tid = threadIdx.x;
  const int i = blockIdx.x;
  const R g_r = a.Rg[i / a.batch];
  cx<R> e[TP];
  if (first) {
    cx<R>* Srow = a.S + (size_t)i * NH;
    cx<R> S0[SP];
#pragma unroll
    for (int m = 0; m < SP; m++) S0[m] = Srow[tid + m * NTH];
    if (a.apply_filter) {
      // field filter deferred from the end of the previous zstep (Models.jl:116-119)
      const R inv = (R)1 / (R)N;
      store_padded<R, N, NTH>(Z, S0);
      fft_lds<R, N, NTH, -1>(Z, a.tw);
      for (int t = tid; t < N; t += NTH) Z[lds_index(t)] = (g_r * a.Tg[t]) * Z[lds_index(t)];
      __syncthreads();
      fft_lds<R, N, NTH, +1>(Z, a.tw);
#pragma unroll
      for (int m = 0; m < SP; m++) {
        S0[m] = inv * Z[lds_index(tid + m * NTH)];
        Srow[tid + m * NTH] = S0[m];
      }
      __syncthreads();
    }
    store_padded<R, N, NTH>(Z, S0);
    fft_lds<R, N, NTH, -1>(Z, a.tw);
#pragma unroll
    for (int u = 0; u < TP; u++) e[u] = Z[lds_index(tid + u * NTH)];
    if (a.plasma) {
      plasma_row<R, A, N, NTH, true>(Z, rho, kd, a.pp, a.dt, red);
      R mx = 0;
      for (int t = tid; t < N; t += NTH) mx = max(mx, rho[rk_index<R>(t)]);
      mx = block_max<R>(mx, reinterpret_cast<R*>(red));
      if (tid == 0) { a.rho_end[i] = (double)rho[rk_index<R>(N - 1)]; a.rho_max[i] = (double)mx; }
      for (int t = tid; t < N; t += NTH) {
        rho_g[(size_t)i * N + t] = rho[rk_index<R>(t)];
        kd_g[(size_t)i * N + t] = kd[rk_index<R>(t)];
        if (a.rho_full) {
          a.rho_full[(size_t)i * N + t] = rho[rk_index<R>(t)];
          a.kd_full[(size_t)i * N + t] = kd[rk_index<R>(t)];
        }
      }
    } else {
      for (int t = tid; t < N; t += NTH) { rho[rk_index<R>(t)] = 0; kd[rk_index<R>(t)] = 0; }
      __syncthreads();
    }
  } else {
    cx<R> Us[SP];
#pragma unroll
    for (int m = 0; m < SP; m++) Us[m] = U[(size_t)i * NH + tid + m * NTH];
    stage_field<R, N, NTH>(Us, e, Z, a.tw);
    {
      // the row's density and kd from the first stage, all loads in flight
      // before the first LDS write
      R rv[TP], kv[TP];
#pragma unroll
      for (int u = 0; u < TP; u++) {
        const size_t q = (size_t)i * N + tid + u * NTH;
        rv[u] = a.plasma ? rho_g[q] : (R)0;
        kv[u] = a.plasma ? kd_g[q] : (R)0;
      }
#pragma unroll
      for (int u = 0; u < TP; u++) {
        rho[rk_index<R>(tid + u * NTH)] = rv[u];
        kd[rk_index<R>(tid + u * NTH)] = kv[u];
      }
    }
    __syncthreads();
  }
  cx<R> k[SP];
  stage_rhs<R, N, NTH, RAMAN>(Z, e, rho, kd, a.tw, a.coef, a.hr, a.sc, g_r, a.Tg, a.responses, a.losses_ereal, k);
#pragma unroll
  for (int m = 0; m < SP; m++) Rout[(size_t)i * NH + tid + m * NTH] = k[m];
}

// Kutta's RK3 combinations (nonlinear_propagators.jl:70-74, ODEIntegrators.RK3):
//   stage 1: out = S + dz/2 K1;  stage 2: out = S + dz (2 K2 - K1);
//   stage 3: out = S + dz (K1/6 + 2 K2/3 + K3/6)  (the row kernel's order).
template <typename R>
__global__ void rk3_combine_kernel(cx<R>* out, const cx<R>* S, const cx<R>* K1, const cx<R>* K2, const cx<R>* K3,
                                   R dz, const double* dz_dev, int stage, size_t n) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const R h = dz_dev ? (R)*dz_dev : dz;
  const cx<R> s = S[q];
  if (stage == 1) {
    out[q] = s + ((R)0.5 * h) * K1[q];
  } else if (stage == 2) {
    out[q] = s + h * (K2[q] + K2[q] - K1[q]);
  } else {
    const R c16 = (R)1 / (R)6, c23 = (R)2 / (R)3;
    out[q] = s + h * (c16 * K1[q] + c23 * K2[q] + c16 * K3[q]);
  }
}

template <int N> constexpr int np_threads() {
  return N / FIL_FFT_MAXRAD < 64 ? 64 : (N / FIL_FFT_MAXRAD > 1024 ? 1024 : N / FIL_FFT_MAXRAD);
}

template <typename R, typename A, int N>
static hipError_t launch_np_stage_n(const RowArgs<R>& a, const cx<R>* U, cx<R>* Rout, R* rho_g, R* kd_g, int first,
                                    int rows, hipStream_t st) {
  constexpr int NTH = np_threads<N>();
  const size_t sm = (size_t)lds_size(N) * sizeof(cx<R>) + (size_t)rk_size<R>(N) * 2 * sizeof(R) +
                    2 * (NTH / 64) * sizeof(A) + 64;
  const bool raman = a.responses & (FIL_RESP_RAMAN_THG | FIL_RESP_RAMAN_NOTHG);
  if (raman && !a.hr) return hipErrorInvalidValue;
  auto k = raman ? np_stage_kernel<R, A, N, NTH, true> : np_stage_kernel<R, A, N, NTH, false>;
  hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
  if (e != hipSuccess) return e;
  fil_launch(k, dim3(rows), dim3(NTH), sm, st, a, U, Rout, rho_g, kd_g, first);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_np_stage(const RowArgs<R>& a, const cx<R>* U, cx<R>* Rout, R* rho_g, R* kd_g, int first, int rows,
                           int N, bool scan64, hipStream_t st) {
#define FIL_NP_CASE(NN)                                                                         \
  case NN:                                                                                      \
    return (scan64 || sizeof(R) == 8) ? launch_np_stage_n<R, double, NN>(a, U, Rout, rho_g, kd_g, first, rows, st) \
                                      : launch_np_stage_n<R, float, NN>(a, U, Rout, rho_g, kd_g, first, rows, st);
  switch (N) {
    FIL_NP_CASE(256)
    FIL_NP_CASE(512)
    FIL_NP_CASE(1024)
    FIL_NP_CASE(2048)
    FIL_NP_CASE(4096)
    FIL_NP_CASE(8192)
    default: return hipErrorInvalidValue;
  }
#undef FIL_NP_CASE
}

template <typename R>
hipError_t launch_rk3_combine(cx<R>* out, const cx<R>* S, const cx<R>* K1, const cx<R>* K2, const cx<R>* K3, R dz,
                              const double* dz_dev, int stage, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(rk3_combine_kernel<R>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, out, S, K1, K2, K3,
                     dz, dz_dev, stage, n);
  return hipGetLastError();
}

template hipError_t launch_np_stage<float>(const RowArgs<float>&, const cx<float>*, cx<float>*, float*, float*, int,
                                           int, int, bool, hipStream_t);
template hipError_t launch_np_stage<double>(const RowArgs<double>&, const cx<double>*, cx<double>*, double*, double*,
                                            int, int, int, bool, hipStream_t);
template hipError_t launch_rk3_combine<float>(cx<float>*, const cx<float>*, const cx<float>*, const cx<float>*,
                                              const cx<float>*, float, const double*, int, size_t, hipStream_t);
template hipError_t launch_rk3_combine<double>(cx<double>*, const cx<double>*, const cx<double>*, const cx<double>*,
                                               const cx<double>*, double, const double*, int, size_t, hipStream_t);

}  // namespace fil
