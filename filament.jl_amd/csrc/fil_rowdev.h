// fil_rowdev.h -- device functions of the row phase shared by the row kernel
// (fil_row.hip) and the non-paraxial stage kernel (fil_nonpar.hip): the rate
// tables, the plasma scan, the RK stage right-hand side and the stage field.
#pragma once

#include "fil_internal.h"
#include "fil_fft.h"
#include "fil_kernels.h"

namespace fil {
// ---------------------------------------------------------------------------
// Rate of one plasma component: TabulatedFunctions.TFunction evaluated as
// log-log linear interpolation, clamped to the end nodes
// (photoionization_avalanche.jl:158,184,196).  `power` receives the local
// log-log slope (tfpower, :196).  It is a reference that is always written: a
// pointer chosen at run time between a local and null would keep that local
// in scratch memory.
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ R tf_eval(const PlasmaParams<R>& pp, int q, R I, R& power) {
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
  power = s;
  const R y = (R)pp.ly[q][i] + s * (lc - ((R)pp.lx0[q] + (R)i * (R)pp.dlx[q]));
  return (sizeof(R) == 8) ? (R)exp2((double)y) : (R)exp2f((float)y);
}
template <typename R>
__device__ __forceinline__ R tf_eval(const PlasmaParams<R>& pp, int q, R I) {
  R unused;
  return tf_eval<R>(pp, q, I, unused);
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
// Writes rho[rk_index(t)], kd[rk_index(t)] (LDS).  CLOBBER_Z: Z may serve as
// the scan's cache and holds no field on return (the caller has read what it
// needs of it beforehand).  Without avalanche every component obeys the
// linear ODE d rho/dt = W(t) (f - rho); one RK3 step of it (the reference's
// ODEIntegrators.RK3, stages at t_j, t_j + dt/2, t_j + dt with the field
// linearly interpolated) is exactly the affine map
//   rho_{j+1} = rho_j + h_j (f - rho_j),
//   h_j = dt [W1/6 + 2/3 W2 (1 - dt W1/2) + 1/6 W3 (1 + dt W1 - 2 dt W2 + dt^2 W1 W2)],
// so the sequential t loop becomes a block-wide prefix scan of affine maps,
// carried in A (FP64 for the fp64-scan form).  With avalanche the RHS is
// quadratic and one lane walks the row sequentially.
// ---------------------------------------------------------------------------
template <typename R, typename A, int N, int NTH, bool CLOBBER_Z = false>
__device__ void plasma_row(cx<R>* Z, R* rho, R* kd, const PlasmaParams<R>& pp, R dt, A* red) {
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
          for (int q = 0; q < nc; q++) W[st][q] = tf_eval<R>(pp, q, Ist[st]);
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
          R w1 = tf_eval<R>(pp, q, Ist[0], pw);
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
  // Pass 1 evaluates the rates of every sample and leaves what pass 2 needs,
  // so that pass 2 evaluates none:
  //   one component:  the gain h_j (rounded to R, and used rounded by both
  //                   passes, so the scan stays consistent) in rho[j+1] and
  //                   K W1(t_j) in kd[j];
  //   two components (CLOBBER_Z): the gains of both (rounded to R) in Z[j],
  //                   K W1(t_j) of component 0 in rho[j+1], of component 1 in
  //                   kd[j];
  // the slots pass 2 overwrites with rho and kdrho.  Each thread owns the
  // slots of its chunk in both passes; Z[j0], which the previous thread reads
  // as its last E1, is written only after the scan's barriers.  Otherwise
  // pass 2 evaluates the rates again.
  const bool cache = (nc == 1), cache2 = CLOBBER_Z && (nc == 2);
  // pass 1: this thread's composed affine map for each component
  Aff<A> m[FIL_MAXCOMP];
#pragma unroll
  for (int q = 0; q < FIL_MAXCOMP; q++) { m[q].a = (A)1; m[q].b = (A)0; }
  R Wprev[FIL_MAXCOMP], Pprev[FIL_MAXCOMP];
  bool have_prev = false;
  cx<R> E1c = mk((R)0, (R)0), hz = mk((R)0, (R)0);
  for (int c = 0; c < C; c++) {
    const int j = j0 + c;
    if (j >= N - 1) break;
    cx<R> E0 = have_prev ? E1c : Z[lds_index(j)], E1 = Z[lds_index(j + 1)];
    E1c = E1;
    cx<R> Em = (R)0.5 * (E0 + E1);
    R I0 = fiarg(E0, pp.ereal), Im = fiarg(Em, pp.ereal), I1 = fiarg(E1, pp.ereal);
    R hq0 = 0, hq1 = 0;   // scalars: the q loop is not unrolled (nc is a run-time count)
#pragma unroll
    for (int q = 0; q < FIL_MAXCOMP; q++) {
      if (q >= nc) break;
      R pw1 = have_prev ? Pprev[q] : (R)0, pw3 = 0;
      R W1 = have_prev ? Wprev[q] : tf_eval<R>(pp, q, I0, pw1);
      R W2 = tf_eval<R>(pp, q, Im);
      R W3 = tf_eval<R>(pp, q, I1, pw3);
      Wprev[q] = W3;
      Pprev[q] = pw3;
      A w1 = W1, w2 = W2, w3 = W3;
      A h = rk3_gain<A>(dtA, w1, w2, w3);
      if (cache || cache2) {
        const R hr = (R)h;
        h = (A)hr;
        const R kw1 = (pp.kdep ? pw1 : (R)pp.K[q]) * W1;
        if (q == 0) rho[rk_index<R>(j + 1)] = cache ? hr : kw1;
        if (q == 0 && cache) kd[rk_index<R>(j)] = kw1;
        if (q == 1) kd[rk_index<R>(j)] = kw1;
        if (q == 0) hq0 = hr;
        if (q == 1) hq1 = hr;
      }
      Aff<A> st; st.a = (A)1 - h; st.b = h * (A)pp.frhont[q];
      m[q] = compose(st, m[q]);
    }
    if (cache2) {
      if (c == 0) hz = mk(hq0, hq1);
      else Z[lds_index(j)] = mk(hq0, hq1);
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
  if (cache2) {
    // the scan's barriers have ordered every pass-1 read of Z[j0]
    if (j0 < N - 1) Z[lds_index(j0)] = hz;
    const A f0 = (A)pp.frhont[0], f1 = (A)pp.frhont[1];
    A r0 = r[0], r1 = r[1];
    for (int c = 0; c < C; c++) {
      const int j = j0 + c;
      if (j >= N - 1) break;
      const cx<R> h = Z[lds_index(j)];
      const R kw0 = rho[rk_index<R>(j + 1)], kw1 = kd[rk_index<R>(j)];
      r0 = ((A)1 - (A)h.x) * r0 + (A)h.x * f0;
      r1 = ((A)1 - (A)h.y) * r1 + (A)h.y * f1;
      rho[rk_index<R>(j + 1)] = (R)(r0 + r1);
      kd[rk_index<R>(j)] = kw0 * (R)(f0 - r0) + kw1 * (R)(f1 - r1);
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
      R W1 = tf_eval<R>(pp, q, I0, pw);
      R W2 = tf_eval<R>(pp, q, Im);
      R W3 = tf_eval<R>(pp, q, I1);
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

}  // namespace fil
