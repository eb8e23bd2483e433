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

namespace fil {

// ---------------------------------------------------------------------------
// Rate of one plasma component: TabulatedFunctions.TFunction evaluated as
// log-log linear interpolation, clamped to the end nodes
// (photoionization_avalanche.jl:158,184,196).  `power` gives the local
// log-log slope (tfpower, :321).
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ R tf_eval(const PlasmaParams<R>& pp, int q, R I, R* power) {
  const int n = pp.ntab[q];
  int i;
  R lc;
  if (!(I > (R)pp.xmin[q])) { i = 0; lc = (R)pp.lx0[q]; }
  else {
    R l = (sizeof(R) == 8) ? (R)log2((double)I) : (R)__log2f((float)I);
    R u = (l - (R)pp.lx0[q]) * (R)pp.inv_dlx[q];
    if (u >= (R)(n - 1)) { i = n - 2; lc = (R)pp.lx0[q] + (R)(n - 1) * (R)pp.dlx[q]; }
    else { i = (int)u; if (i > n - 2) i = n - 2; lc = l; }
  }
  const R s = (R)pp.slope[q][i];
  if (power) *power = s;
  const R y = (R)pp.ly[q][i] + s * (lc - ((R)pp.lx0[q] + (R)i * (R)pp.dlx[q]));
  return (sizeof(R) == 8) ? (R)exp2((double)y) : (R)exp2f((float)y);
}

// bilinear interpolation of one pooled plane at pooled coordinates (y, x), clamped
__device__ __forceinline__ float bilinear(const float* P, int H, int W, float y, float x) {
  y = fminf(fmaxf(y, 0.0f), (float)(H - 1));
  x = fminf(fmaxf(x, 0.0f), (float)(W - 1));
  const int y0 = (int)y, x0 = (int)x;
  const int y1 = min(y0 + 1, H - 1), x1 = min(x0 + 1, W - 1);
  const float fy = y - (float)y0, fx = x - (float)x0;
  const float a = P[y0 * W + x0] + fx * (P[y0 * W + x1] - P[y0 * W + x0]);
  const float b = P[y1 * W + x0] + fx * (P[y1 * W + x1] - P[y1 * W + x0]);
  return a + fy * (b - a);
}

// Corrector output at row i, sample j of the window: scale (Dre + i Dim)
// exp(-i w0 t_j), D bilinearly interpolated from the pooled-block centres
// (block h covers rows [h pr, (h+1) pr), centre at row (h + 1/2) pr - 1/2).
__device__ __forceinline__ cx<float> corr_value(const CorrGeom& g, const float* __restrict__ D, float scale, int i,
                                                int j) {
  const float y = ((float)i + 0.5f) / (float)g.pr - 0.5f;
  const float x = ((float)(j - g.t0) + 0.5f) / (float)g.pt - 0.5f;
  const int HW = g.H * g.W;
  const cx<float> d = mk(scale * bilinear(D, g.H, g.W, y, x), scale * bilinear(D + HW, g.H, g.W, y, x));
  return d * conj(g.carrier[j]);
}

template <typename R>
__device__ __forceinline__ R fiarg(cx<R> e, int ereal) { return ereal ? e.x * e.x : norm2(e); }

// Gain of one Kutta RK3 step of d rho/dt = W(t)(f - rho) with stage rates
// W1, W2, W3: rho -> rho + h (f - rho),
// h = dt [W1/6 + 2/3 W2 (1 - dt W1/2) + 1/6 W3 (1 + dt W1 - 2 dt W2 + dt^2 W1 W2)].
template <typename A>
__device__ __forceinline__ A rk3_gain(A dt, A w1, A w2, A w3) {
  constexpr A c6 = (A)1 / (A)6, c23 = (A)2 / (A)3;
  const A dw1 = dt * w1, dw2 = dt * w2;
  return dt * (c6 * w1 + c23 * w2 * ((A)1 - (A)0.5 * dw1) + c6 * w3 * ((A)1 + dw1 - (A)2 * dw2 + dw1 * dw2));
}

template <typename R>
__device__ __forceinline__ R recip(R x) {
  if constexpr (sizeof(R) == 4) return __builtin_amdgcn_rcpf(x);
  else return (R)1 / x;
}

// Affine map rho -> a*rho + b, composed along t.  later o earlier.
template <typename A>
struct Aff { A a, b; };
template <typename A>
__device__ __forceinline__ Aff<A> compose(Aff<A> later, Aff<A> earlier) {
  Aff<A> r; r.a = later.a * earlier.a; r.b = later.a * earlier.b + later.b; return r;
}

// Exclusive block scan of affine maps in thread order.  red: LDS scratch of
// 2*(NTH/64) A values.
template <typename A, int NTH>
__device__ Aff<A> block_exclusive_scan(Aff<A> v, A* red) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  Aff<A> inc = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    A pa = __shfl_up(inc.a, off, 64), pb = __shfl_up(inc.b, off, 64);
    if (lane >= off) { Aff<A> p; p.a = pa; p.b = pb; inc = compose(inc, p); }
  }
  // exclusive within the wave
  A ea = __shfl_up(inc.a, 1, 64), eb = __shfl_up(inc.b, 1, 64);
  Aff<A> ex; ex.a = (lane == 0) ? (A)1 : ea; ex.b = (lane == 0) ? (A)0 : eb;
  constexpr int NW = NTH / 64;
  if (lane == 63) { red[2 * wv] = inc.a; red[2 * wv + 1] = inc.b; }
  __syncthreads();
  Aff<A> wp; wp.a = (A)1; wp.b = (A)0;
  for (int w = 0; w < wv && w < NW; w++) { Aff<A> t; t.a = red[2 * w]; t.b = red[2 * w + 1]; wp = compose(t, wp); }
  __syncthreads();
  return compose(ex, wp);
}

template <typename R>
__device__ __forceinline__ R block_max(R v, R* red) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  if (lane == 0) red[wv] = v;
  __syncthreads();
  R m = red[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = max(m, red[w]);
  __syncthreads();
  return m;
}

// rho and kdrho rows in LDS carry one pad element per 128 bytes: the scan
// writes them in per-thread chunks of N/NTH consecutive samples (a lane stride
// of 32 or 64 bytes), which would put a wave's stores into 4 (FP32) or 2
// (FP64) banks; padded, the lanes of a half-wave hit distinct banks.
template <typename R> __host__ __device__ constexpr int rk_index(int j) { return j + j / (128 / (int)sizeof(R)); }
template <typename R> __host__ __device__ constexpr int rk_size(int n) { return n + n / (128 / (int)sizeof(R)); }

// ---------------------------------------------------------------------------
// Plasma density of one row from the time-domain field Z (LDS).
// Writes rho[rk_index(t)], kd[rk_index(t)] (LDS).  Without avalanche every component obeys the
// linear ODE d rho/dt = W(t) (f - rho); one RK3 step of it (the reference's
// ODEIntegrators.RK3, stages at t_j, t_j + dt/2, t_j + dt with the field
// linearly interpolated) is exactly the affine map
//   rho_{j+1} = rho_j + h_j (f - rho_j),
//   h_j = dt [W1/6 + 2/3 W2 (1 - dt W1/2) + 1/6 W3 (1 + dt W1 - 2 dt W2 + dt^2 W1 W2)],
// so the sequential t loop becomes a block-wide prefix scan of affine maps,
// carried in A (FP64 for the fp64-scan form).  With avalanche the RHS is
// quadratic and one lane walks the row sequentially.
// ---------------------------------------------------------------------------
template <typename R, typename A, int N, int NTH>
__device__ void plasma_row(const cx<R>* Z, R* rho, R* kd, const PlasmaParams<R>& pp, R dt, A* red) {
  const int tid = threadIdx.x;
  const int nc = pp.ncomp;
  if (pp.avalanche) {
    if (tid == 0) {
      // sequential reference loop (plasma_equations.jl:60-69)
      R u[FIL_MAXCOMP], ut[FIL_MAXCOMP];
      for (int q = 0; q < nc; q++) { u[q] = (R)pp.rho0; ut[q] = (R)pp.rho0; }
      R s0 = 0; for (int q = 0; q < nc; q++) s0 += u[q];
      rho[rk_index<R>(0)] = s0;
      for (int j = 0; j < N - 1; j++) {
        cx<R> E0 = Z[lds_index(j)], E1 = Z[lds_index(j + 1)];
        cx<R> Em = (R)0.5 * (E0 + E1);
        R Ist[3] = {fiarg(E0, pp.ereal), fiarg(Em, pp.ereal), fiarg(E1, pp.ereal)};
        R E2st[3] = {E0.x * E0.x, Em.x * Em.x, E1.x * E1.x};
        R W[3][FIL_MAXCOMP];
        for (int st = 0; st < 3; st++)
          for (int q = 0; q < nc; q++) W[st][q] = tf_eval<R>(pp, q, Ist[st], nullptr);
        // avalanche integrator
        R k[3][FIL_MAXCOMP], tmp[FIL_MAXCOMP];
        for (int st = 0; st < 3; st++) {
          for (int q = 0; q < nc; q++) {
            if (st == 0) tmp[q] = u[q];
            else if (st == 1) tmp[q] = u[q] + dt * ((R)0.5 * k[0][q]);
            else tmp[q] = u[q] + dt * (-k[0][q] + (R)2 * k[1][q]);
          }
          R tot = 0; for (int q = 0; q < nc; q++) tot += tmp[q];
          for (int q = 0; q < nc; q++) {
            R f = (R)pp.frhont[q];
            k[st][q] = W[st][q] * (f - tmp[q]) + (R)pp.Rava[q] * E2st[st] * (f - tmp[q]) / f * tot;
          }
        }
        R s = 0;
        for (int q = 0; q < nc; q++) {
          u[q] = u[q] + dt * (k[0][q] / (R)6 + (R)2 * k[1][q] / (R)3 + k[2][q] / (R)6);
          s += u[q];
        }
        rho[rk_index<R>(j + 1)] = s;
        // avalanche-free integrator for kdrho
        for (int st = 0; st < 3; st++) {
          for (int q = 0; q < nc; q++) {
            if (st == 0) tmp[q] = ut[q];
            else if (st == 1) tmp[q] = ut[q] + dt * ((R)0.5 * k[0][q]);
            else tmp[q] = ut[q] + dt * (-k[0][q] + (R)2 * k[1][q]);
          }
          for (int q = 0; q < nc; q++) k[st][q] = W[st][q] * ((R)pp.frhont[q] - tmp[q]);
        }
        R kdv = 0;
        for (int q = 0; q < nc; q++) {
          ut[q] = ut[q] + dt * (k[0][q] / (R)6 + (R)2 * k[1][q] / (R)3 + k[2][q] / (R)6);
          R pw;
          R w1 = tf_eval<R>(pp, q, Ist[0], &pw);
          R K = pp.kdep ? pw : (R)pp.K[q];
          kdv += K * (w1 * ((R)pp.frhont[q] - ut[q]));
        }
        kd[rk_index<R>(j)] = kdv;
      }
      kd[rk_index<R>(N - 1)] = 0;
    }
    __syncthreads();
    return;
  }

  constexpr int C = N / NTH;   // contiguous chunk of t per thread
  const int j0 = tid * C;
  const A dtA = (A)dt;
  // One component: pass 1 leaves the gain h_j of every sample (rounded to R,
  // and used rounded by both passes, so the scan stays consistent) in rho[j+1]
  // and K W1(t_j) in kd[j], the slots pass 2 overwrites with rho and kdrho
  // (each thread owns the slots of its chunk in both passes); pass 2 then
  // evaluates no rates.
  const bool cache = (nc == 1);
  // pass 1: this thread's composed affine map for each component
  Aff<A> m[FIL_MAXCOMP];
#pragma unroll
  for (int q = 0; q < FIL_MAXCOMP; q++) { m[q].a = (A)1; m[q].b = (A)0; }
  R Wprev[FIL_MAXCOMP];
  R Pprev = 0;
  bool have_prev = false;
  for (int c = 0; c < C; c++) {
    const int j = j0 + c;
    if (j >= N - 1) break;
    cx<R> E0 = Z[lds_index(j)], E1 = Z[lds_index(j + 1)];
    cx<R> Em = (R)0.5 * (E0 + E1);
    R I0 = fiarg(E0, pp.ereal), Im = fiarg(Em, pp.ereal), I1 = fiarg(E1, pp.ereal);
    const bool slope = cache && pp.kdep;
#pragma unroll
    for (int q = 0; q < FIL_MAXCOMP; q++) {
      if (q >= nc) break;
      R pw1 = Pprev, pw3 = 0;
      R W1 = have_prev ? Wprev[q] : tf_eval<R>(pp, q, I0, slope ? &pw1 : nullptr);
      R W2 = tf_eval<R>(pp, q, Im, nullptr);
      R W3 = tf_eval<R>(pp, q, I1, slope ? &pw3 : nullptr);
      Wprev[q] = W3;
      Pprev = pw3;
      A w1 = W1, w2 = W2, w3 = W3;
      A h = rk3_gain<A>(dtA, w1, w2, w3);
      if (cache) {
        const R hr = (R)h;
        h = (A)hr;
        rho[rk_index<R>(j + 1)] = hr;
        kd[rk_index<R>(j)] = (pp.kdep ? pw1 : (R)pp.K[0]) * W1;
      }
      Aff<A> st; st.a = (A)1 - h; st.b = h * (A)pp.frhont[q];
      m[q] = compose(st, m[q]);
    }
    have_prev = true;
  }
  // block scan per component -> density at the start of this thread's chunk
  A r[FIL_MAXCOMP];
#pragma unroll
  for (int q = 0; q < FIL_MAXCOMP; q++) {
    if (q >= nc) break;
    Aff<A> ex = block_exclusive_scan<A, NTH>(m[q], red);
    r[q] = ex.a * (A)pp.rho0 + ex.b;
  }
  if (tid == 0) {
    R s0 = 0; for (int q = 0; q < nc; q++) s0 += (R)pp.rho0;
    rho[rk_index<R>(0)] = s0;
  }
  // pass 2: walk the chunk again
  if (cache) {
    const A f = (A)pp.frhont[0];
    A r0 = r[0];
    for (int c = 0; c < C; c++) {
      const int j = j0 + c;
      if (j >= N - 1) break;
      const A h = (A)rho[rk_index<R>(j + 1)];
      const R kw1 = kd[rk_index<R>(j)];
      r0 = ((A)1 - h) * r0 + h * f;
      rho[rk_index<R>(j + 1)] = (R)r0;
      kd[rk_index<R>(j)] = kw1 * (R)(f - r0);
    }
    if (tid == NTH - 1) kd[rk_index<R>(N - 1)] = 0;
    __syncthreads();
    return;
  }
  have_prev = false;
  for (int c = 0; c < C; c++) {
    const int j = j0 + c;
    if (j >= N - 1) break;
    cx<R> E0 = Z[lds_index(j)], E1 = Z[lds_index(j + 1)];
    cx<R> Em = (R)0.5 * (E0 + E1);
    R I0 = fiarg(E0, pp.ereal), Im = fiarg(Em, pp.ereal), I1 = fiarg(E1, pp.ereal);
    A tot = 0;
    R kdv = 0;
#pragma unroll
    for (int q = 0; q < FIL_MAXCOMP; q++) {
      if (q >= nc) break;
      R pw = 0;
      R W1 = tf_eval<R>(pp, q, I0, pp.kdep ? &pw : nullptr);
      R W2 = tf_eval<R>(pp, q, Im, nullptr);
      R W3 = tf_eval<R>(pp, q, I1, nullptr);
      A w1 = W1, w2 = W2, w3 = W3;
      const A h = rk3_gain<A>(dtA, w1, w2, w3);
      r[q] = ((A)1 - h) * r[q] + h * (A)pp.frhont[q];
      tot += r[q];
      R K = pp.kdep ? pw : (R)pp.K[q];
      kdv += K * (W1 * (R)((A)pp.frhont[q] - r[q]));
    }
    rho[rk_index<R>(j + 1)] = (R)tot;
    kd[rk_index<R>(j)] = kdv;
  }
  if (tid == NTH - 1) kd[rk_index<R>(N - 1)] = 0;
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Responses of one RK stage.  ev[u] holds the stage field in time at
// t = tid + u NTH (registers); on exit k[m] (spectral bins w = tid + NTH*m,
// w < N/2) holds func!'s dE.  Z (LDS) is the transforms' work buffer: no
// thread may still read it on entry, except the caller's own reads of the
// elements it passes in ev when the register passes are not used.
//   F_cubic = 3/4 |E|^2 Re E  or Re(E)^3          cubic.jl:36,44
//   F_free  = rho Re E                             current_free.jl:35
//   F_loss  = kdrho Re E / |E|^2 (0 below 1e-30)   current_losses.jl:36-44, :71-83
// each multiplied by Rguard(r) Tguard(t) (nonlinear_propagators.jl:50), taken
// to the analytic spectrum (Models.jl:180-183) and weighted by i QZ Rnl
// (nonlinear_propagators.jl:51,57).  Two real signals share one complex FFT:
// with Z = F_a + i F_b, 2 F_a^(w) = Z^(w) + conj Z^(N-w),
// 2 F_b^(w) = -i (Z^(w) - conj Z^(N-w)).  coef = i QZ Rnl d(w) / (2N s_q).
// Each signal enters the FFT pre-multiplied by its scale s_q (the typical
// size of |i QZ Rnl| for that response): the raw responses differ by up to
// ~20 orders of magnitude (3/4 |E|^2 Re E ~ 1 against rho Re E ~ 1e-16), and
// without the scaling the FFT rounding of the large one would swamp the
// small one, whose coefficient is correspondingly huge.
// ---------------------------------------------------------------------------
template <typename R, int N, int NTH, bool RAMAN>
__device__ __forceinline__ void stage_rhs(cx<R>* Z, const cx<R>* ev, const R* rho, const R* kd,
                                          const cx<R>* __restrict__ tw,
                                          const cx<R>* __restrict__ coef, const cx<R>* __restrict__ hr,
                                          const R* sc, R g_r, const R* __restrict__ Tg, int resp, int losses_ereal,
                                          cx<R>* k) {
  constexpr int TP = N / NTH;
  constexpr int NH = N / 2;
  constexpr int SP = NH / NTH;
  const int tid = threadIdx.x;
  const bool cub = resp & (FIL_RESP_CUBIC_THG | FIL_RESP_CUBIC_NOTHG);
  const bool thg = resp & FIL_RESP_CUBIC_THG;
  const bool fre = resp & FIL_RESP_FREE;
  const bool los = resp & FIL_RESP_LOSSES;
  const bool ram_thg = resp & FIL_RESP_RAMAN_THG;
  // the Raman variant keeps Re(E) and the Raman intensity of its samples
  // across the convolution; the kernel without Raman does not hold them
  constexpr int TR = RAMAN ? TP : 1;
  constexpr bool REGS = fft_regs_ok<N, NTH>();
  R Fl[TP], Re[TR], Ir[TR];
  cx<R> zin[TP];
#pragma unroll
  for (int u = 0; u < TP; u++) {
    const int t = tid + u * NTH;
    const cx<R> e = ev[u];
    const R re = e.x, I = norm2(e);
    const R g = g_r * Tg[t];
    R fc = 0, ff = 0, fl = 0;
    if (cub) fc = thg ? re * re * re : (R)0.75 * I * re;
    if (fre) ff = rho[rk_index<R>(t)] * re;
    if (los) {
      R fe = losses_ereal ? re * re : I;
      R inv = (fe >= (R)1e-30) ? recip<R>(fe) : (R)0;
      fl = kd[rk_index<R>(t)] * inv * re;
    }
    Fl[u] = (g * sc[2]) * fl;
    if constexpr (RAMAN) {
      Re[u] = re;
      Ir[u] = ram_thg ? re * re : (R)0.75 * I;    // calc_raman / calc_raman_nothg (raman.jl:46-65)
    }
    zin[u] = mk((g * sc[0]) * fc, (g * sc[1]) * ff);
  }
  // without register passes the caller's reads of ev came from Z
  if constexpr (!REGS) __syncthreads();
#pragma unroll
  for (int m = 0; m < SP; m++) k[m] = mk((R)0, (R)0);
  if (cub || fre) {
    if constexpr (REGS) {
      fft_regs_to_lds<R, N, NTH, +1, false, true>(zin, Z, tw, tid);
    } else {
#pragma unroll
      for (int u = 0; u < TP; u++) Z[lds_index(tid + u * NTH)] = zin[u];
      __syncthreads();
      fft_lds<R, N, NTH, +1, true>(Z, tw);
    }
#pragma unroll
    for (int m = 0; m < SP; m++) {
      const int w = tid + m * NTH;
      cx<R> zw = Z[w], zm = conj(Z[(N - w) & (N - 1)]);   // natural order
      cx<R> a2 = zw + zm, b2 = mul_mi(zw - zm);
      k[m] = coef[w] * a2 + coef[NH + w] * b2;
    }
    __syncthreads();
  }
  if constexpr (RAMAN) {
    // Raman: F = conv(I, H) Re(E) with conv = FFT-(hr .* FFT+(I)), hr = Hraman / Nt
    // (convolution!, raman.jl:68-111).  H comes from a real response, so the
    // convolution of the real I is real; its real part is kept.
    if constexpr (REGS) {
#pragma unroll
      for (int u = 0; u < TP; u++) zin[u] = mk(Ir[u], (R)0);
      fft_regs_to_regs<R, N, NTH, +1, false>(zin, zin, Z, tw, tid);
#pragma unroll
      for (int u = 0; u < TP; u++) zin[u] = hr[tid + u * NTH] * zin[u];
      fft_regs_to_regs<R, N, NTH, -1, false>(zin, zin, Z, tw, tid);
    } else {
#pragma unroll
      for (int u = 0; u < TP; u++) Z[lds_index(tid + u * NTH)] = mk(Ir[u], (R)0);
      __syncthreads();
      fft_lds<R, N, NTH, +1>(Z, tw);
      for (int j = tid; j < N; j += NTH) Z[lds_index(j)] = hr[j] * Z[lds_index(j)];
      __syncthreads();
      fft_lds<R, N, NTH, -1>(Z, tw);
#pragma unroll
      for (int u = 0; u < TP; u++) zin[u] = Z[lds_index(tid + u * NTH)];
      __syncthreads();
    }
    // losses and Raman share one complex FFT
#pragma unroll
    for (int u = 0; u < TP; u++) {
      const int t = tid + u * NTH;
      const R g = g_r * Tg[t];
      Ir[u] = (g * sc[3]) * (zin[u].x * Re[u]);
      zin[u] = mk(Fl[u], Ir[u]);
    }
    if constexpr (REGS) {
      fft_regs_to_lds<R, N, NTH, +1, false, true>(zin, Z, tw, tid);
    } else {
#pragma unroll
      for (int u = 0; u < TP; u++) Z[lds_index(tid + u * NTH)] = zin[u];
      __syncthreads();
      fft_lds<R, N, NTH, +1, true>(Z, tw);
    }
#pragma unroll
    for (int m = 0; m < SP; m++) {
      const int w = tid + m * NTH;
      cx<R> zw = Z[w], zm = conj(Z[(N - w) & (N - 1)]);   // natural order
      k[m] = k[m] + coef[2 * NH + w] * (zw + zm) + coef[3 * NH + w] * mul_mi(zw - zm);
    }
    __syncthreads();
  } else if (los) {
    // the losses field is real: transform it as N/2 complex points
    // y[n] = fl[2n] + i fl[2n+1], Y = FFT_{N/2}(y); then with
    // A = Y[w] + conj(Y[-w]), B = Y[w] - conj(Y[-w]) (indices mod N/2),
    // 2 FL[w] = A - i exp(+2 pi i w / N) B for w < N/2.
    R* Zr = reinterpret_cast<R*>(Z);
#pragma unroll
    for (int u = 0; u < TP; u++) {
      const int t = tid + u * NTH;
      Zr[2 * lds_index(t >> 1) + (t & 1)] = Fl[u];
    }
    __syncthreads();
    fft_lds<R, N / 2, NTH, +1, true>(Z, tw + fft_tw_half_offset(N));
    const cx<R>* post = tw + fft_tw_post_offset(N);
#pragma unroll
    for (int m = 0; m < SP; m++) {
      const int w = tid + m * NTH;
      const cx<R> yw = Z[w], ym = conj(Z[(NH - w) & (NH - 1)]);   // natural order
      const cx<R> A = yw + ym, Bm = yw - ym;
      k[m] = k[m] + coef[2 * NH + w] * (A + mul_mi(post[w] * Bm));
    }
    __syncthreads();
  }
}

template <typename R, int N, int NTH>
__device__ __forceinline__ void store_padded(cx<R>* Z, const cx<R>* v) {
  constexpr int SP = (N / 2) / NTH;
  const int tid = threadIdx.x;
#pragma unroll
  for (int m = 0; m < SP; m++) { Z[lds_index(tid + m * NTH)] = v[m]; Z[lds_index(N / 2 + tid + m * NTH)] = mk((R)0, (R)0); }
  __syncthreads();
}

// e[u] <- time-domain field at t = tid + u NTH of the half spectrum U
// (bins tid + m NTH, m < N/2/NTH; the negative frequencies are zero).  Z must
// be free on entry and is free on return.
template <typename R, int N, int NTH>
__device__ __forceinline__ void stage_field(const cx<R>* U, cx<R>* e, cx<R>* Z, const cx<R>* __restrict__ tw) {
  if constexpr (fft_regs_ok<N, NTH>()) {
    fft_regs_to_regs<R, N, NTH, -1, true>(U, e, Z, tw, (int)threadIdx.x);
  } else {
    store_padded<R, N, NTH>(Z, U);
    fft_lds<R, N, NTH, -1>(Z, tw);
#pragma unroll
    for (int u = 0; u < N / NTH; u++) e[u] = Z[lds_index(threadIdx.x + u * NTH)];
  }
}

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

  if (a.plasma) {
    plasma_row<R, A, N, NTH>(Z, rho, kd, a.pp, a.dt, red);
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
  // stage 1 starts from the field the plasma scan read (Z)
#pragma unroll
  for (int u = 0; u < TP; u++) e[u] = Z[lds_index(tid + u * NTH)];
  if constexpr (REGS) __syncthreads();
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
  for (int t = tid; t < N; t += NTH) Z[lds_index(t)] = Et[(size_t)i * N + t];
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
  for (int w = tid; w < N / 2; w += NTH) {
    Z[lds_index(w)] = S[((size_t)i * rs + ro) * (N / 2) + w];
    Z[lds_index(N / 2 + w)] = mk((R)0, (R)0);
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
  for (int w = tid; w < NH; w += NTH) { Z[lds_index(w)] = Srow[w]; Z[lds_index(NH + w)] = mk((R)0, (R)0); }
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
  for (int i = i0; i < i1; i++) s += (double)in[(size_t)i * ncol + j] * wt[i];
  part[(size_t)blockIdx.y * ncol + j] = s;
}

__global__ void colsum_final_kernel(const double* __restrict__ part, int nchunk, int ncol, double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ncol) return;
  double s = 0;
  for (int c = 0; c < nchunk; c++) s += part[(size_t)c * ncol + j];
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
