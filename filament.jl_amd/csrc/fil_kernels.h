// fil_kernels.h -- kernel argument blocks and host launchers.
#pragma once

#include "fil_internal.h"

namespace fil {

// Learned coarse-step corrector (fil_row.hip: features / apply; fil_corr.hip: network).
// Window rows [0, nr), t samples [t0, t0 + nt); pooled grid H = nr/pr by W = nt/pt.
struct CorrGeom {
  int nr, t0, nt, pr, pt, H, W;
  const cx<float>* carrier;  // [Nt] exp(i w0 t_j) (rounded from FP64): envelope A = E carrier
  float dt;        // t step (plasma scan)
  float inv_as;    // 1 / field scale
  float inv_rs;    // 1 / density scale
};

template <typename R>
struct RowArgs {
  cx<R>* S;                 // [Nr][Nt/2] positive-frequency spectrum (in/out)
  const cx<R>* tw;          // [Nt] exp(-2 pi i j/Nt)
  const R* Tg;              // [Nt] Tguard
  const R* Rg;              // [Nr] Rguard
  const cx<R>* coef;        // [4][Nt/2] i QZ Rnl d(w)/(2 Nt sc): cubic, free, losses, Raman
  const cx<R>* hr;          // [Nt] Raman response spectrum / Nt (null without Raman)
  R sc[4];                  // per-response scale applied before the packed FFTs
  R dz, dt;
  int apply_filter;         // field filter of the previous step still pending
  int nonlinearity, plasma, responses, losses_ereal;
  int batch;                // rows are (r, member) pairs, r slowest: Rg index = row / batch
  PlasmaParams<R> pp;
  double* rho_end;          // [Nr] rho(r, t_end)
  double* rho_max;          // [Nr] max_t rho(r, t)
  R* rho_full;              // optional [Nr][Nt]
  R* kd_full;               // optional [Nr][Nt]
  // pending corrector output (FP32 single-pulse handles; null: none): added
  // in time to the window rows before the field filter, as launch_corr_apply
  // would add it
  const float* corr_D;      // [2][H][W]
  CorrGeom corr_g;
  float corr_scale;
  const double* dz_dev;     // non-null: dz is read from device memory (the adaptive run)
};

template <typename R>
hipError_t launch_row(const RowArgs<R>& a, int Nr, int N, bool scan64, hipStream_t st);
// Non-paraxial nonlinear step (fil_nonpar.hip): one RK stage's paraxial
// right-hand side of every row into Rout (first: from S, with the pending
// filter and the plasma solve, whose rows go to rho_g / kd_g; else from U),
// and Kutta's combination for stage 1, 2 (out = next stage's input) or 3
// (out = the new state) over n elements.
template <typename R>
hipError_t launch_np_stage(const RowArgs<R>& a, const cx<R>* U, cx<R>* Rout, R* rho_g, R* kd_g, int first, int rows,
                           int N, bool scan64, hipStream_t st);
template <typename R>
hipError_t launch_rk3_combine(cx<R>* out, const cx<R>* S, const cx<R>* K1, const cx<R>* K2, const cx<R>* K3, R dz,
                              const double* dz_dev, int stage, size_t n, hipStream_t st);
// T geometry linear step: S[i][w] *= P[w], i < rows, w < Nh
template <typename R>
hipError_t launch_scale_rows(cx<R>* S, const cx<R>* P, int rows, int Nh, hipStream_t st);
// The per-member kernels below address spectrum row i (i < Nr) at S row i*rs + ro
// (rs = batch size, ro = member); Et, Rg and the outputs are indexed by i.
// Upload / download of one pulse: row i of Et is row (i rs + ro) of S (rs = batch,
// ro = member; 1, 0 for a single pulse).
template <typename R>
hipError_t launch_to_spectrum(const cx<R>* Et, cx<R>* S, const cx<R>* tw, int Nr, int N, int rs, int ro, hipStream_t st);
template <typename R>
hipError_t launch_to_time(const cx<R>* S, cx<R>* Et, const cx<R>* tw, const R* Rg, const R* Tg, int apply,
                          int Nr, int N, int rs, int ro, hipStream_t st);
// Analyzer rows (It and Sw may be null: only Fr, Imax_row and the projected S):
// output row i (< Nr) reads row (i rs + ro) of S with the r guard
// Rg[i / gdiv]: every row of a batch with (1, 0, batch), one member with
// (batch, member, 1).
template <typename R>
hipError_t launch_analyze_rows(cx<R>* S, const cx<R>* tw, const R* Rg, const R* Tg, int apply, R dt, double* Fr,
                               R* It, R* Sw, double* Imax_row, int Nr, int N, int rs, int ro, int gdiv,
                               hipStream_t st);
// out[j] = sum_i in[i][j] * wt[i] over Nr rows, FP64, in a fixed order (row chunks, then the
// chunks in order); part: scratch of COLSUM_MAXCHUNK * ncol doubles
constexpr int COLSUM_MAXCHUNK = 256;
template <typename T>
hipError_t launch_colsum(const T* in, int Nr, int ncol, const double* wt, double* part, double* out,
                         hipStream_t st);
template <typename R>
hipError_t launch_rowsum_abs2(const cx<R>* S, int Nr, int Nh, double scale, double* out, hipStream_t st);

// Features of the current state, X[4][H][W]: Re A, Im A (/a_s), |E|^2 (/a_s^2),
// rho (/rho_s), each averaged over its pr x pt block.  Frow: [nr][4][W], the
// per-row block averages; X null: Frow only (the pooling over rows is left to
// launch_conv_fwd_pooled).
hipError_t launch_corr_features(const cx<float>* S, const cx<float>* tw, const PlasmaParams<float>& pp, int plasma,
                                const CorrGeom& g, int N, float* Frow, float* X, hipStream_t st);
// S rows [0, nr): E += scale * D(r, t) exp(-i w0 t) with D[2][H][W] bilinearly
// interpolated from the pooled-block centres, in the t window.
hipError_t launch_corr_apply(cx<float>* S, const cx<float>* tw, const CorrGeom& g, int N, const float* D, float scale,
                             hipStream_t st);

// Network of the corrector: 4 layers of 3x3 convolutions (zero padding),
// 4 -> C -> C -> C -> 2 channels, ReLU between layers, BF16 weights, FP32
// accumulation.  Activations [B][ch][H][W] FP32.
constexpr int CORR_CH = 32;
constexpr int CORR_IN = 4;
constexpr int CORR_OUT = 2;
constexpr int CORR_LAYERS = 4;
__host__ __device__ constexpr int corr_cin(int l) { return l == 0 ? CORR_IN : CORR_CH; }
__host__ __device__ constexpr int corr_cout(int l) { return l == CORR_LAYERS - 1 ? CORR_OUT : CORR_CH; }
// offset of layer l's weights [cout][cin][9] in the flat parameter vector; its bias [cout] follows
__host__ __device__ constexpr int corr_woff(int l) {
  int o = 0;
  for (int k = 0; k < l; k++) o += corr_cout(k) * corr_cin(k) * 9 + corr_cout(k);
  return o;
}
__host__ __device__ constexpr int corr_nparam() { return corr_woff(CORR_LAYERS); }

// BF16 weights in the operand order of v_mfma_f32_16x16x32_bf16 (A[m][k]
// with m = output channel, k = tap * cin + ci, zero-padded to whole 16-row
// groups and 32-k chunks): fragment (g, c) of layer l holds, for lane L, the
// eight values A[16 g + L % 16][32 c + 8 (L / 16) + j], j = 0..7, so every
// lane loads its operand with one 16-byte read.
__host__ __device__ constexpr int corr_frag_ng(int l) { return (corr_cout(l) + 15) / 16; }
__host__ __device__ constexpr int corr_frag_nc(int l) { return (9 * corr_cin(l) + 31) / 32; }
__host__ __device__ constexpr int corr_frag_off(int l) {   // in fragments of 64 x 8 values
  int o = 0;
  for (int k = 0; k < l; k++) o += corr_frag_ng(k) * corr_frag_nc(k);
  return o;
}
__host__ __device__ constexpr int corr_frag_size() { return corr_frag_off(CORR_LAYERS) * 64 * 8; }   // bf16 values
// wfr <- the fragment layout of every layer's BF16 weights (wbf: the flat
// BF16 parameter copy; biases are not part of wfr)
hipError_t launch_conv_weight_layout(const uint16_t* wbf, uint16_t* wfr, hipStream_t st);

// Forward of the whole network for B samples in one launch:
//   act[l] = relu(conv_l(...)) for l = 0, 1, 2 (when act is not null, the
//   activations training needs), out = conv_3(act[2]) [B][2][H][W].
// Input: with pr > 0 the per-row features Frow [H pr][4][W] of one sample
// (B = 1), averaged over the pr rows of each pooled row as they are read;
// otherwise X [.][4][H][W], sample b taken from X + (idx ? idx[b] : b) 4HW.
// params: the flat FP32 parameter vector (the biases are read from it).
hipError_t launch_conv_net(const float* in, const int* idx, int pr, const uint16_t* wfr, const float* params,
                           float* const* act, float* out, int B, int H, int W, hipStream_t st);
// din = conv^T(dout) (* [act > 0] when act != null)
hipError_t launch_conv_bwd_data(int l, const float* dout, const uint16_t* wbf, const float* act, float* din, int B,
                                int H, int W, hipStream_t st);
// dW[cout][cin][9], db[cout] summed over the batch (deterministic order); part: scratch of
// CORR_WCHUNK_MAX * CORR_CH * (CORR_CH + 1) * 9 floats for the per-chunk partial sums
constexpr int CORR_WCHUNK_PIX = 4096;   // pixels per partial-sum chunk
constexpr int CORR_WCHUNK_MAX = 64;
hipError_t launch_conv_bwd_weight(int l, const float* dout, const float* in, const int* idx, float* dW, float* db,
                                  float* part, int B, int H, int W, hipStream_t st);
// dout = 2 (out - Y[idx[b]] * inv_ys) / (B*2*H*W); part[blk] = partial sums of squares
hipError_t launch_corr_loss(const float* out, const float* Y, const int* idx, float inv_ys, float* dout, double* part,
                            int B, int HW, int* nblk, hipStream_t st);
// Adam step on the flat parameter vector; refreshes the BF16 copy of the weights.
hipError_t launch_adam(float* w, const float* g, float* m, float* v, uint16_t* wbf, int n, float lr, float b1,
                       float b2, float eps, float c1, float c2, hipStream_t st);
hipError_t launch_round_bf16(const float* w, uint16_t* wbf, int n, hipStream_t st);
// part[blk] = partial sums of x^2 over x[0, n)
hipError_t launch_sum_squares(const float* x, int n, double* part, int* nblk, hipStream_t st);

// XYT linear step (fil_xyt.hip).  Propagator inputs in FP64, device pointers.
struct XytProp {
  int Nx, Ny, Nh;
  int kparaxial, guard3d_bug;
  // the x (y) factors of P are the same at ix and Nx - ix (iy and Ny - iy),
  // bit for bit, as the FFT frequency grid makes them: P is stored for
  // ix <= Nx/2 (iy <= Ny/2) only.  Px = Nx/2 + 1 or Nx, Py = Ny/2 + 1 or Ny.
  int xsym, ysym, Px, Py;
  const double* kt2x;   // [Nx] (kx kxu)^2 [1/m^2]
  const double* kt2y;   // [Ny]
  const double* KX;     // [Nx] KXguard
  const double* KY;     // [Ny] KYguard, times Wguard[iy] with guard3d_bug
  const double* beta;   // [Nh] complex beta(w) [1/m]
  const double* wvf;    // [Nh] w / vf [1/m]
  const double* Wg;     // [Nt] Wguard
  double zu;            // [m]
  double norm;          // 1 / (Nx Ny), the inverse transform's normalisation
};
// P[(ix + Px iy) Nh + w], ix < Px, iy < Py, for the step's dz (dz_dev: dz read from device memory)
template <typename R>
hipError_t launch_xyt_propagator(const XytProp& pp, double dz, cx<R>* P, hipStream_t st,
                                 const double* dz_dev = nullptr);
// Q[(ix + Px iy) Nh + w]: the non-paraxial nonlinear prefactor over the paraxial
// one, with the spectral filter and 1/(Nx Ny) (launch_xyt_linear applies it as P)
template <typename R>
hipError_t launch_xyt_qz(const XytProp& pp, cx<R>* Q, hipStream_t st);
// S <- IFFT_xy[P .* FFT_xy[S]] for every w column of S [(ix + Nx iy)][Nh]
template <typename R>
hipError_t launch_xyt_linear(cx<R>* S, const cx<R>* P, const XytProp& pp, const cx<R>* twx, const cx<R>* twy,
                             hipStream_t st);

// Hankel-transform phase (fil_dht.hip)
// Propagator P[k][w] = exp(i KZ dz) Kguard Wguard computed in FP64 (dz_dev: dz
// read from device memory instead).
template <typename R>
hipError_t launch_propagator(const double* KZ64, const double* KG64, double dz, cx<R>* P, int Nr, int Nh,
                             hipStream_t st, const double* dz_dev = nullptr);
// Out[m][w] = sum_n M[m][n] In[n][w]  (optionally Out *= P), complex In/Out [Nr][Nh], real M [Nr][Nr].
template <typename R>
// Nh: complex columns of In/Out (the batch's members side by side); the propagator P
// [Nr][Pw] is shared by the members: Out[m][w] *= P[m][w mod Pw].
hipError_t launch_dht_gemm(const R* M, const cx<R>* In, cx<R>* Out, const cx<R>* P, int Nr, int Nh, int Pw,
                           hipStream_t st, int tiling = 0);
// Tile shapes of the MFMA products (M rows per workgroup): FIL_TILING_RULE picks
// 128-row tiles when they still give >= 256 workgroups for one pulse, else 64;
// the other two force one shape (128 needs Nr % 128 == 0).  The k summation
// order of every output is the same for all shapes, so results do not depend
// on the choice.
// split-BF16 3-term variant, M given as hi/lo bf16 planes.  The first product
// writes X = P .* (M1 In) as four BF16 planes Xp [re_hi, re_lo, im_hi, im_lo][Nh][Nr]
// (the split of the FP32 X, k of the second product contiguous); the second
// reads them: Out = M2 X.
hipError_t launch_dht_gemm_bf16x3_to_planes(const uint16_t* Mhi, const uint16_t* Mlo, const cx<float>* In,
                                            uint16_t* Xp, const cx<float>* P, int Nr, int Nh, int Pw, hipStream_t st,
                                            int tiling = 0);
// (Pw: one pulse's columns, which picks the tile shape as in the first product)
hipError_t launch_dht_gemm_bf16x3_from_planes(const uint16_t* Mhi, const uint16_t* Mlo, const uint16_t* Xp,
                                              cx<float>* Out, int Nr, int Nh, int Pw, hipStream_t st, int tiling = 0);

// max that returns NaN when either argument is NaN (Julia's maximum does)
__device__ inline double nan_max(double a, double b) { return (a != a || b != b) ? a + b : fmax(a, b); }

// Adaptive z-step on the device (fil_row.hip): the step rule of AStepKerrPlasma
// (AdaptiveSteps.jl:96-107) and the loop tests of Filament.run
// (Filament.jl:129-164), evaluated after each step's analysis.
struct AdaptRule {
  double dz0, phimax, phik, phip, Istop;
  float zmax;             // z and zmax are single precision, as the RT run keeps them
  int nonlinearity, plasma;
  int nrows;              // analyzer rows reduced for Imax and rhomax
  int max_steps;          // steps the log holds
};
struct AdaptState {
  double dz;              // dz of the next step, read by the propagator and the row kernel
  float z;                // z after `steps` steps
  int steps;
  int done;               // 0 running, else an FIL_STOP_* reason
  int pad;
};
// One workgroup: reduces Imax_row (and rhomax_row with plasma) over the rows,
// logs them for step `steps`, then stops or sets dz, z and steps for the next
// step.  After `done` is set it does nothing.  log: [max_steps + 1][4] of
// (z, dz, Imax, rhomax).
hipError_t launch_adapt(const AdaptRule& rule, AdaptState* state, const double* Imax_row, const double* rhomax_row,
                        double* log, hipStream_t st);

}  // namespace fil
