// fil_kernels.h -- kernel argument blocks and host launchers.
#pragma once

#include "fil_internal.h"

namespace fil {

template <typename R>
struct RowArgs {
  cx<R>* S;                 // [Nr][Nt/2] positive-frequency spectrum (in/out)
  const cx<R>* tw;          // [Nt] exp(-2 pi i j/Nt)
  const R* Tg;              // [Nt] Tguard
  const R* Rg;              // [Nr] Rguard
  const cx<R>* coef;        // [3][Nt/2] i QZ Rnl d(w)/(2 Nt sc): cubic, free, losses
  R sc[3];                  // per-response scale applied before the packed FFTs
  R dz, dt;
  int apply_filter;         // field filter of the previous step still pending
  int nonlinearity, plasma, responses, losses_ereal;
  PlasmaParams<R> pp;
  double* rho_end;          // [Nr] rho(r, t_end)
  double* rho_max;          // [Nr] max_t rho(r, t)
  R* rho_full;              // optional [Nr][Nt]
  R* kd_full;               // optional [Nr][Nt]
};

template <typename R>
hipError_t launch_row(const RowArgs<R>& a, int Nr, int N, bool scan64, hipStream_t st);
template <typename R>
hipError_t launch_to_spectrum(const cx<R>* Et, cx<R>* S, const cx<R>* tw, int Nr, int N, hipStream_t st);
template <typename R>
hipError_t launch_to_time(const cx<R>* S, cx<R>* Et, const cx<R>* tw, const R* Rg, const R* Tg, int apply,
                          int Nr, int N, hipStream_t st);
template <typename R>
hipError_t launch_analyze_rows(cx<R>* S, const cx<R>* tw, const R* Rg, const R* Tg, int apply, R dt, double* Fr,
                               float* It, float* Sw, double* Imax_row, int Nr, int N, hipStream_t st);
hipError_t launch_colsum(const float* in, int Nr, int ncol, const double* rdr, double* out, hipStream_t st);
template <typename R>
hipError_t launch_rowsum_abs2(const cx<R>* S, int Nr, int Nh, double scale, double* out, hipStream_t st);

// Hankel-transform phase (fil_dht.hip)
// Propagator P[k][w] = exp(i KZ dz) Kguard Wguard computed in FP64.
template <typename R>
hipError_t launch_propagator(const double* KZ64, const double* KG64, double dz, cx<R>* P, int Nr, int Nh,
                             hipStream_t st);
// Out[m][w] = sum_n M[m][n] In[n][w]  (optionally Out *= P), complex In/Out [Nr][Nh], real M [Nr][Nr].
template <typename R>
hipError_t launch_dht_gemm(const R* M, const cx<R>* In, cx<R>* Out, const cx<R>* P, int Nr, int Nh,
                           hipStream_t st);
// split-BF16 3-term variant: M given as hi/lo bf16 planes.
hipError_t launch_dht_gemm_bf16x3(const uint16_t* Mhi, const uint16_t* Mlo, const cx<float>* In, cx<float>* Out,
                                  const cx<float>* P, int Nr, int Nh, hipStream_t st);

}  // namespace fil
