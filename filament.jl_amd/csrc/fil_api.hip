// fil_api.hip -- the C ABI of libfilament_hip.so (declared in include/filament_hip.h).
//
// Device state of one RT problem:
//   S   [Nrp][Nt/2]  positive-frequency spectrum of the field (the z-step state)
//   X   [Nrp][Nt/2]  Hankel-space intermediate of the linear step
//   M1, M2 [Nrp][Nrp] folded QDHT matrices; P [Nrp][Nt/2] propagator for the
//   current dz.  Nrp = Nr rounded up to 64 (padding rows/columns are zero).
// Rows are t/w-contiguous (the transpose of the reference's column-major
// arrays), so the row phase reads whole rows and the Hankel phase reads the
// spectrum as the B operand of a GEMM without a transpose.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "fil_fft.h"
#include "fil_internal.h"
#include "fil_kernels.h"
#include "fil_host.h"

using namespace fil;

namespace fil {

static thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// Per-pass twiddles of an N-point transform (fil_fft.h layout) appended to
// tw starting at entry o; returns the next entry.
size_t append_twiddles(HostArr& tw, size_t o, int N) {
  for (int s = 1; N / s > 1; s *= fft_radix(N, s)) {
    const int rad = fft_radix(N, s);
    if (s < N / rad) {
      const int P = N / (s * rad);
      for (int p = 0; p < P; p++)
        for (int k = 0; k < FIL_TW_STRIDE; k++) {
          const double a = -2.0 * kPi * (double)(((long)p * s * k) % N) / N;
          if (k < rad) tw.setc(o, std::cos(a), std::sin(a));
          o++;
        }
    }
  }
  return o;
}

// ext: also the half-length table and the real-signal post-processing factors
// the row kernel uses (fft_tw_ext_size layout).
int build_twiddles(DevBuf& buf, int N, bool f64, bool ext) {
  const int n = ext ? fft_tw_ext_size(N) : fft_tw_size(N);
  HostArr tw(std::max(1, n), f64, true);
  size_t o = append_twiddles(tw, 0, N);
  if ((int)o != fft_tw_size(N)) return fail(FIL_EINVAL, "internal: twiddle table size");
  if (ext) {
    o = append_twiddles(tw, o, N / 2);
    if ((int)o != fft_tw_post_offset(N)) return fail(FIL_EINVAL, "internal: twiddle table size");
    for (int w = 0; w < N / 2; w++) {
      const double a = 2.0 * kPi * (double)w / N;
      tw.setc(o++, std::cos(a), std::sin(a));
    }
  }
  HIPCHK(upload(buf, tw.b.data(), tw.b.size()));
  return 0;
}

}  // namespace fil

namespace {


double radius(const double* x, const double* y, int n) {
  double ymax = -INFINITY;
  for (int i = 0; i < n; i++) ymax = std::max(ymax, y[i]);
  double lev = ymax * std::exp(-1.0), radl = 0, radr = 0;
  for (int i = 0; i < n; i++) if (y[i] >= lev) { radl = x[i]; break; }
  for (int i = n - 1; i >= 0; i--) if (y[i] >= lev) { radr = x[i]; break; }
  return (std::fabs(radl) + std::fabs(radr)) / 2;
}

}  // namespace

struct fil_handle {
  fil_desc d{};
  int Nr = 0, Nrp = 0, Nt = 0, Nh = 0;
  int B = 1;                 // batch: members side by side, rows (r, member) with r slowest
  bool fp64 = false, bf16x3 = false, scan64 = false;
  hipStream_t st = nullptr;
  bool tables = false, field = false;
  std::vector<char> uploaded;   // members of a batch with a field
  bool pending_filter = false, plasma_valid = false;
  // a corrector output (corr.out, times corr_scale) not yet added to S: the
  // next z-step's row kernel adds it; any other reader of S adds it first
  bool corr_pending = false;
  float corr_scale = 0;
  int tiling = FIL_TILING_RULE;  // MFMA tile shape of the Hankel products (fil_set_tiling)
  double P_dz = NAN;
  double dt = 0;
  double scale[4] = {1, 1, 1, 1};
  std::vector<double> r, rdr, t;
  // device
  DevBuf S, X, Et, M1, M2, M1h, M1l, M2h, M2l, KZ64, KG64, P, tw, Tg, Rg, coef, hr;
  DevBuf tab_ly[FIL_MAXCOMP], tab_sl[FIL_MAXCOMP];
  DevBuf rho_end, rho_max, rho_full, kd_full;
  DevBuf aFr, aIt, aSw, aImax, aFt, aS, ardr, acol;
  // host copy of the analyzer rows of every member, valid until the state changes
  struct {
    bool valid = false;
    std::vector<double> fr, ft, im, re, rm, sp;
    std::vector<char> served;         // members whose analysis was returned
  } acache;
  // XYT geometry (rows are the Nx*Ny points, x fastest)
  bool xyt = false;
  bool tgeo = false;                 // T geometry: one row per pulse, diagonal linear step
  int Nx = 0, Ny = 0;
  double dx = 0, dy = 0;
  std::vector<double> x, y;
  DevBuf twx, twy;                                      // FFT twiddles along x and y
  DevBuf xkt2x, xkt2y, xKX, xKY, xbeta, xwvf, xWg;      // FP64 propagator inputs
  XytProp xprop{};
  // saved states (fil_state_save / fil_state_restore)
  struct Slot {
    DevBuf S, rho_end, rho_max;
    bool used = false, pending_filter = false, plasma_valid = false;
  } slot[FIL_NSLOTS];
  // learned corrector (fil_corrector_*)
  struct Corr {
    bool on = false;
    CorrGeom g{};
    int cap = 0, count = 0, next = 0, bmax = 0;
    double y_scale = 0;
    int64_t adam_t = 0;
    DevBuf w, wbf, m, v, grad;        // flat parameters, BF16 copy, Adam moments, gradient
    DevBuf wfr;                       // the BF16 weights in MFMA operand order
    DevBuf X, Y;                      // stored pairs [cap][4][HW], [cap][2][HW]
    DevBuf Frow, feat, carrier;       // feature scratch, features [4][HW], exp(i w0 t_j) [Nt]
    DevBuf act[3], out, gA, gB, gout; // training activations and gradients [bmax][..][HW]
    DevBuf idx, part, wpart;          // batch indices, loss partials, weight-gradient partials
  } corr;
  PlasmaParams<float> pp32{};
  PlasmaParams<double> pp64{};
  // non-paraxial nonlinear term (fil_desc.nonparaxial): Q(k, w) = QZ(k, w) / QZ(w)
  // Kguard Wguard [Nrp][Nh], the stage right-hand side, the stage derivatives
  // K1, K2, the next stage's input, and the first stage's plasma rows
  bool nonparax = false;
  DevBuf npQ, npR, npK1, npK2, npU, np_rho, np_kd;
  // device z loop (fil_run_adaptive): loop state, log, the state at the last host check
  DevBuf ad_state, ad_log, ad_S, ad_re, ad_rm;
  // timing
  bool timing = false;
  int timing_period = 1;        // time every n-th z-step
  int64_t timing_tick = 0;
  bool timing_now = false;      // the current z-step is timed
  std::vector<hipEvent_t> ev;   // 6 per recorded step: (start, stop) of the row phase and of the two linear phases
  size_t ev_used = 0;
  double ms[3] = {0, 0, 0};
  int64_t count = 0;

  size_t rsz() const { return fp64 ? sizeof(double) : sizeof(float); }
  int rows() const { return Nr * B; }                              // rows of the row-wise kernels
  size_t ns() const { return (size_t)Nrp * B * Nh; }               // elements of the state S
  size_t csz() const { return 2 * rsz(); }
};

// Handles stepped together by one HIP graph per z-step (see fil_group_zstep).
struct fil_group {
  std::vector<fil_handle*> hs;
  int device = 0;
  hipStream_t st = nullptr;            // the graph is launched here
  hipEvent_t fork = nullptr, post = nullptr;
  std::vector<hipEvent_t> join, pre;
  struct Entry {
    double dz;
    uint64_t pattern;                  // bit k: handle k had a pending field filter
    std::vector<int> tiling;           // the handles' tile shapes when recorded
    hipGraphExec_t exec;
  };
  std::vector<Entry> cache;            // at most 4 instantiated graphs
  int64_t captures = 0, replays = 0;
  bool orphaned = false;               // a handle of the group was destroyed
};

// Live groups (their graphs keep the handles' device pointers).
static std::mutex g_groups_mu;
static std::vector<fil_group*> g_groups;

static bool in_live_group(const fil_handle* h) {
  std::lock_guard<std::mutex> lock(g_groups_mu);
  for (auto* g : g_groups)
    for (auto* x : g->hs)
      if (x == h) return true;
  return false;
}

namespace {

// precision-converting copy of a half-spectrum state
template <typename D, typename S>
__global__ void convert_state_kernel(const cx<S>* __restrict__ src, cx<D>* __restrict__ dst, size_t n) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) dst[q] = mk((D)src[q].x, (D)src[q].y);
}

template <typename R>
void put_real(std::vector<unsigned char>& buf, size_t idx, double v) { reinterpret_cast<R*>(buf.data())[idx] = (R)v; }


// round to bf16 (nearest even), returning the bits and the rounded value
uint16_t to_bf16(float f, float* back) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
  uint16_t h = (uint16_t)(r >> 16);
  uint32_t v = (uint32_t)h << 16;
  std::memcpy(back, &v, 4);
  return h;
}

} // namespace

namespace fil {
thread_local LaunchEvents fil_launch_events;
}

namespace {

// The (start, stop) events of the next timed phase; none while timing is off.
int next_pair(fil_handle* h, hipEvent_t* s, hipEvent_t* e) {
  *s = *e = nullptr;
  if (!h->timing || !h->timing_now) return 0;
  while (h->ev.size() < h->ev_used + 2) {
    hipEvent_t x;
    HIPCHK(hipEventCreate(&x));
    h->ev.push_back(x);
  }
  *s = h->ev[h->ev_used];
  *e = h->ev[h->ev_used + 1];
  h->ev_used += 2;
  return 0;
}
// A phase of one kernel: the next fil_launch carries the pair on its dispatch.
int arm(fil_handle* h) {
  hipEvent_t s, e;
  if (int rc = next_pair(h, &s, &e)) return rc;
  fil_launch_events.start = s;
  fil_launch_events.stop = e;
  return 0;
}
// A phase of several kernels (or none): event records around it.
int phase_begin(fil_handle* h, hipEvent_t* stop) {
  hipEvent_t s;
  if (int rc = next_pair(h, &s, stop)) return rc;
  if (s) HIPCHK(hipEventRecord(s, h->st));
  return 0;
}
int phase_end(fil_handle* h, hipEvent_t stop) {
  if (stop) HIPCHK(hipEventRecord(stop, h->st));
  return 0;
}

// Adds a pending corrector output to S now (every reader or writer of the
// state other than the next z-step sees the corrected field).
int corr_materialize(fil_handle* h) {
  if (!h->corr_pending) return 0;
  h->corr_pending = false;
  auto& c = h->corr;
  HIPCHK(launch_corr_apply(h->S.as<cx<float>>(), h->tw.as<cx<float>>(), c.g, h->Nt, c.out.as<float>(),
                           h->corr_scale, h->st));
  h->plasma_valid = false;
  h->acache.valid = false;
  return 0;
}

int flush_timing(fil_handle* h) {
  if (!h->ev_used) return 0;
  HIPCHK(hipStreamSynchronize(h->st));
  for (size_t s = 0; s + 6 <= h->ev_used; s += 6) {
    float a, b, c;
    HIPCHK(hipEventElapsedTime(&a, h->ev[s], h->ev[s + 1]));
    HIPCHK(hipEventElapsedTime(&b, h->ev[s + 2], h->ev[s + 3]));
    HIPCHK(hipEventElapsedTime(&c, h->ev[s + 4], h->ev[s + 5]));
    h->ms[0] += a; h->ms[1] += b; h->ms[2] += c;
    h->count++;
  }
  h->ev_used = 0;
  return 0;
}

// Linear propagator P = exp(i KZ dz) * spectral guard for the step's dz (only
// when dz changes).
template <typename R>
int update_propagator(fil_handle* h, double dz) {
  if (h->P_dz == dz) return 0;
  if (h->xyt)
    HIPCHK(launch_xyt_propagator<R>(h->xprop, dz, h->P.as<cx<R>>(), h->st));
  else
    HIPCHK(launch_propagator<R>(h->KZ64.as<double>(), h->KG64.as<double>(), dz, h->P.as<cx<R>>(), h->Nrp, h->Nh, h->st));
  h->P_dz = dz;
  return 0;
}

// dz_dev: dz is read from device memory by the propagator and the row kernel
// (the device z loop); dz is then unused.  dz_lin (when not NaN): the linear
// step advances by dz_lin instead of dz; with dz = 0 the nonlinear step and
// the plasma solve are skipped (only a pending field filter or corrector
// output is applied before the linear step).
template <typename R>
int zstep_impl(fil_handle* h, double dz, const double* dz_dev = nullptr, double dz_lin = NAN) {
  h->acache.valid = false;
  const bool parts = !dz_dev && !std::isnan(dz_lin);
  const bool linear_only = parts && dz == 0.0;
  if (!parts) dz_lin = dz;
  if (dz_dev) {
    if (h->xyt)
      HIPCHK(launch_xyt_propagator<R>(h->xprop, 0.0, h->P.as<cx<R>>(), h->st, dz_dev));
    else
      HIPCHK(launch_propagator<R>(h->KZ64.as<double>(), h->KG64.as<double>(), 0.0, h->P.as<cx<R>>(), h->Nrp, h->Nh,
                                  h->st, dz_dev));
    h->P_dz = NAN;
  } else if (int rc = update_propagator<R>(h, dz_lin)) {
    return rc;
  }
  RowArgs<R> a{};
  a.S = h->S.as<cx<R>>();
  a.tw = h->tw.as<cx<R>>();
  a.Tg = h->Tg.as<R>();
  a.Rg = h->Rg.as<R>();
  a.coef = h->coef.as<cx<R>>();
  a.hr = h->hr.p ? h->hr.as<cx<R>>() : nullptr;
  for (int q = 0; q < 4; q++) a.sc[q] = (R)h->scale[q];
  a.dz = (R)dz;
  a.dz_dev = dz_dev;
  a.dt = (R)h->dt;
  a.apply_filter = h->pending_filter ? 1 : 0;
  a.nonlinearity = linear_only ? 0 : h->d.nonlinearity;
  a.plasma = linear_only ? 0 : h->d.plasma;
  a.batch = h->B;
  a.responses = h->d.responses;
  a.losses_ereal = h->d.losses_ereal;
  if constexpr (sizeof(R) == 8) a.pp = h->pp64; else a.pp = h->pp32;
  a.rho_end = h->rho_end.as<double>();
  a.rho_max = h->rho_max.as<double>();
  a.rho_full = h->d.keep_plasma ? h->rho_full.as<R>() : nullptr;
  a.kd_full = h->d.keep_plasma ? h->kd_full.as<R>() : nullptr;
  if (h->nonparax)
    if (int rc = corr_materialize(h)) return rc;
  if constexpr (sizeof(R) == 4) {
    if (h->corr_pending) {
      // the row kernel adds the corrector output before the field filter
      a.corr_D = h->corr.out.as<float>();
      a.corr_g = h->corr.g;
      a.corr_scale = h->corr_scale;
    }
  }
  // Timed phases (row, linear 1, linear 2): a failed launch drops the step's events.
  const size_t ev0 = h->ev_used;
  h->timing_now = h->timing && (h->timing_tick++ % h->timing_period) == 0;
#define FIL_TIMED(launch_)                                                                          \
  do {                                                                                              \
    if (int rc_ = arm(h)) return rc_;                                                               \
    const hipError_t e_ = (launch_);                                                                \
    fil_launch_events = LaunchEvents{};                                                             \
    if (e_ != hipSuccess) {                                                                         \
      h->ev_used = ev0;                                                                             \
      return fail(FIL_EDEVICE, std::string(#launch_) + ": " + hipGetErrorString(e_));               \
    }                                                                                               \
  } while (0)
  if (h->nonparax) {
    // three stages: right-hand side rows, then the space operator -> K_s, then
    // Kutta's combination.  RT: K_s = M2 (Q .* (M1 R)); XYT: the rows go
    // straight to K_s and IFFT_xy[Q .* FFT_xy[K_s]] runs in place.
    const size_t n = h->ns();
    cx<R>* S = h->S.as<cx<R>>();
    cx<R>* K1 = h->npK1.as<cx<R>>();
    cx<R>* K2 = h->npK2.as<cx<R>>();
    cx<R>* U = h->npU.as<cx<R>>();
    cx<R>* Rb = h->npR.as<cx<R>>();
    // RT: the third stage's input U is consumed by M1 before M2 writes K3 into it;
    // XYT: K3 goes to Rb, since the stage kernel reads U while it writes
    cx<R>* Ks[3] = {K1, K2, h->xyt ? Rb : U};
    for (int s = 0; s < 3; s++) {
      cx<R>* rows_out = h->xyt ? Ks[s] : Rb;
      if (s == 0)
        FIL_TIMED(launch_np_stage<R>(a, nullptr, rows_out, h->np_rho.as<R>(), h->np_kd.as<R>(), 1, h->rows(),
                                     h->Nt, h->scan64, h->st));
      else
        HIPCHK(launch_np_stage<R>(a, U, rows_out, h->np_rho.as<R>(), h->np_kd.as<R>(), 0, h->rows(), h->Nt,
                                  h->scan64, h->st));
      if (h->xyt) {
        HIPCHK(launch_xyt_linear<R>(Ks[s], h->npQ.as<cx<R>>(), h->xprop, h->twx.as<cx<R>>(), h->twy.as<cx<R>>(),
                                    h->st));
      } else {
        HIPCHK(launch_dht_gemm<R>(h->M1.as<R>(), Rb, h->X.as<cx<R>>(), h->npQ.as<cx<R>>(), h->Nrp, h->Nh * h->B,
                                  h->Nh, h->st, h->tiling));
        HIPCHK(launch_dht_gemm<R>(h->M2.as<R>(), h->X.as<cx<R>>(), Ks[s], nullptr, h->Nrp, h->Nh * h->B, h->Nh,
                                  h->st, h->tiling));
      }
      HIPCHK(launch_rk3_combine<R>(s == 2 ? S : U, S, K1, K2, Ks[2], (R)dz, dz_dev, s + 1, n, h->st));
    }
  } else if (!linear_only || a.apply_filter || a.corr_D) {
    FIL_TIMED(launch_row<R>(a, h->rows(), h->Nt, h->scan64, h->st));
  }
  h->corr_pending = false;
  if (h->tgeo) {
    // propagate!(E, LP, dz) for T (spectral.jl:16-26): E(w) *= exp(i KZ dz) Wguard
    FIL_TIMED(launch_scale_rows<R>(h->S.as<cx<R>>(), h->P.as<cx<R>>(), h->rows(), h->Nh, h->st));
    hipEvent_t e2;
    if (int rc = phase_begin(h, &e2)) return rc;
    if (int rc = phase_end(h, e2)) return rc;
  } else if (h->xyt) {
    // the three transform passes are timed together as the first linear phase
    hipEvent_t e1, e2;
    if (int rc = phase_begin(h, &e1)) return rc;
    HIPCHK(launch_xyt_linear<R>(h->S.as<cx<R>>(), h->P.as<cx<R>>(), h->xprop, h->twx.as<cx<R>>(),
                                h->twy.as<cx<R>>(), h->st));
    if (int rc = phase_end(h, e1)) return rc;
    if (int rc = phase_begin(h, &e2)) return rc;
    if (int rc = phase_end(h, e2)) return rc;
  } else if (sizeof(R) == 4 && h->bf16x3) {
    // X travels as its four BF16 planes (same bytes as FP32 complex)
    FIL_TIMED(launch_dht_gemm_bf16x3_to_planes(h->M1h.as<uint16_t>(), h->M1l.as<uint16_t>(), h->S.as<cx<float>>(),
                                               h->X.as<uint16_t>(), h->P.as<cx<float>>(), h->Nrp, h->Nh * h->B, h->Nh,
                                               h->st, h->tiling));
    FIL_TIMED(launch_dht_gemm_bf16x3_from_planes(h->M2h.as<uint16_t>(), h->M2l.as<uint16_t>(), h->X.as<uint16_t>(),
                                                 h->S.as<cx<float>>(), h->Nrp, h->Nh * h->B, h->Nh, h->st, h->tiling));
  } else {
    FIL_TIMED(launch_dht_gemm<R>(h->M1.as<R>(), h->S.as<cx<R>>(), h->X.as<cx<R>>(), h->P.as<cx<R>>(), h->Nrp,
                                 h->Nh * h->B, h->Nh, h->st, h->tiling));
    FIL_TIMED(launch_dht_gemm<R>(h->M2.as<R>(), h->X.as<cx<R>>(), h->S.as<cx<R>>(), nullptr, h->Nrp, h->Nh * h->B,
                                 h->Nh, h->st, h->tiling));
  }
#undef FIL_TIMED
  h->pending_filter = true;
  h->plasma_valid = h->d.plasma != 0 && !linear_only;
  return 0;
}

// Row kernel without the nonlinear step: applies a pending field filter and
// solves the plasma equation for the current field (Models.jl:66).
template <typename R>
int plasma_only(fil_handle* h) {
  h->acache.valid = false;
  RowArgs<R> a{};
  a.S = h->S.as<cx<R>>();
  a.tw = h->tw.as<cx<R>>();
  a.Tg = h->Tg.as<R>();
  a.Rg = h->Rg.as<R>();
  a.coef = h->coef.as<cx<R>>();
  a.dz = 0;
  a.dt = (R)h->dt;
  a.apply_filter = h->pending_filter ? 1 : 0;
  a.nonlinearity = 0;
  a.plasma = 1;
  a.responses = 0;
  a.batch = h->B;
  if constexpr (sizeof(R) == 8) a.pp = h->pp64; else a.pp = h->pp32;
  a.rho_end = h->rho_end.as<double>();
  a.rho_max = h->rho_max.as<double>();
  a.rho_full = h->d.keep_plasma ? h->rho_full.as<R>() : nullptr;
  a.kd_full = h->d.keep_plasma ? h->kd_full.as<R>() : nullptr;
  HIPCHK(launch_row<R>(a, h->rows(), h->Nt, h->scan64, h->st));
  h->pending_filter = false;
  h->plasma_valid = true;
  return 0;
}

// Applies a pending field filter to every row (the row kernel without the
// nonlinear step and the plasma), so that per-member operations of a batch
// see a consistent state.
template <typename R>
int apply_pending_filter(fil_handle* h) {
  if (int rc = corr_materialize(h)) return rc;
  if (!h->pending_filter) return 0;
  h->acache.valid = false;
  RowArgs<R> a{};
  a.S = h->S.as<cx<R>>();
  a.tw = h->tw.as<cx<R>>();
  a.Tg = h->Tg.as<R>();
  a.Rg = h->Rg.as<R>();
  a.coef = h->coef.as<cx<R>>();
  a.dt = (R)h->dt;
  a.apply_filter = 1;
  a.batch = h->B;
  if constexpr (sizeof(R) == 8) a.pp = h->pp64; else a.pp = h->pp32;
  a.rho_end = h->rho_end.as<double>();
  a.rho_max = h->rho_max.as<double>();
  HIPCHK(launch_row<R>(a, h->rows(), h->Nt, h->scan64, h->st));
  h->pending_filter = false;
  return 0;
}

// Device part of the analyzers: the per-row pass (pending filter, projection,
// Fr / Imax / spectrum rows) and the weighted column sums (RT: 2 pi r dr,
// XYT: dx dy).  Fills host vectors; blocks.
struct AnalyzerRows {
  std::vector<double> fr, ft, im, re, rm, sp;
};

// The first analysis after a state change runs over every member in one pass
// and keeps the rows on the host; each member's first call is served from
// there.  A repeated call for a member analyzes (and projects) that member
// again, as the reference's analyze! does on a second call.
template <typename R>
int analyze_device(fil_handle* h, AnalyzerRows& v, int member) {
  const int Nr = h->Nr, Nt = h->Nt, Nh = h->Nh, B = h->B;
  auto& c = h->acache;
  if (c.valid && c.served[member]) {
    HIPCHK(launch_analyze_rows<R>(h->S.as<cx<R>>(), h->tw.as<cx<R>>(), h->Rg.as<R>(), h->Tg.as<R>(), 0, (R)h->dt,
                                  h->aFr.as<double>(), h->aIt.as<R>(), h->aSw.as<R>(), h->aImax.as<double>(), Nr, Nt,
                                  B, member, 1, h->st));
    HIPCHK(launch_colsum<R>(h->aIt.as<R>(), Nr, Nt, h->ardr.as<double>(), h->acol.as<double>(), h->aFt.as<double>(),
                            h->st));
    HIPCHK(launch_colsum<R>(h->aSw.as<R>(), Nr, Nh + 1, h->ardr.as<double>(), h->acol.as<double>(),
                            h->aS.as<double>(), h->st));
    v.fr.assign(Nr, 0.0); v.im.assign(Nr, 0.0); v.ft.assign(Nt, 0.0); v.sp.assign(Nh + 1, 0.0);
    HIPCHK(hipMemcpyAsync(v.fr.data(), h->aFr.p, (size_t)Nr * 8, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipMemcpyAsync(v.im.data(), h->aImax.p, (size_t)Nr * 8, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipMemcpyAsync(v.ft.data(), h->aFt.p, (size_t)Nt * 8, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipMemcpyAsync(v.sp.data(), h->aS.p, (size_t)(Nh + 1) * 8, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    v.re.resize(Nr); v.rm.resize(Nr);
    for (int i = 0; i < Nr; i++) {
      v.re[i] = c.re[(size_t)i * B + member];
      v.rm[i] = c.rm[(size_t)i * B + member];
    }
    return 0;
  }
  if (!c.valid) {
    // the per-member column sums are the column sums of [Nr][B x ncol]
    if (h->d.plasma && !h->plasma_valid)
      if (int rc = plasma_only<R>(h)) return rc;
    const int n = h->rows();
    HIPCHK(launch_analyze_rows<R>(h->S.as<cx<R>>(), h->tw.as<cx<R>>(), h->Rg.as<R>(), h->Tg.as<R>(),
                                  h->pending_filter ? 1 : 0, (R)h->dt, h->aFr.as<double>(), h->aIt.as<R>(),
                                  h->aSw.as<R>(), h->aImax.as<double>(), n, Nt, 1, 0, B, h->st));
    h->pending_filter = false;
    HIPCHK(launch_colsum<R>(h->aIt.as<R>(), Nr, B * Nt, h->ardr.as<double>(), h->acol.as<double>(),
                            h->aFt.as<double>(), h->st));
    HIPCHK(launch_colsum<R>(h->aSw.as<R>(), Nr, B * (Nh + 1), h->ardr.as<double>(), h->acol.as<double>(),
                            h->aS.as<double>(), h->st));
    c.fr.assign(n, 0.0); c.im.assign(n, 0.0); c.re.assign(n, 0.0); c.rm.assign(n, 0.0);
    c.ft.assign((size_t)B * Nt, 0.0); c.sp.assign((size_t)B * (Nh + 1), 0.0);
    HIPCHK(hipMemcpyAsync(c.fr.data(), h->aFr.p, (size_t)n * 8, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipMemcpyAsync(c.im.data(), h->aImax.p, (size_t)n * 8, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipMemcpyAsync(c.ft.data(), h->aFt.p, c.ft.size() * 8, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipMemcpyAsync(c.sp.data(), h->aS.p, c.sp.size() * 8, hipMemcpyDeviceToHost, h->st));
    if (h->d.plasma) {
      HIPCHK(hipMemcpyAsync(c.re.data(), h->rho_end.p, (size_t)n * 8, hipMemcpyDeviceToHost, h->st));
      HIPCHK(hipMemcpyAsync(c.rm.data(), h->rho_max.p, (size_t)n * 8, hipMemcpyDeviceToHost, h->st));
    }
    HIPCHK(hipStreamSynchronize(h->st));
    c.valid = true;
    c.served.assign(B, 0);
  }
  v.fr.resize(Nr); v.im.resize(Nr); v.re.resize(Nr); v.rm.resize(Nr);
  for (int i = 0; i < Nr; i++) {
    const size_t q = (size_t)i * B + member;
    v.fr[i] = c.fr[q]; v.im[i] = c.im[q]; v.re[i] = c.re[q]; v.rm[i] = c.rm[q];
  }
  v.ft.assign(c.ft.begin() + (size_t)member * Nt, c.ft.begin() + (size_t)(member + 1) * Nt);
  v.sp.assign(c.sp.begin() + (size_t)member * (Nh + 1), c.sp.begin() + (size_t)(member + 1) * (Nh + 1));
  c.served[member] = 1;
  return 0;
}

// The analyzer's row pass over the whole state as analyze_device's first pass
// runs it (the plasma solve when stale, the pending filter, the projection,
// the Imax rows), without the intensity and spectrum rows, the column sums
// and the host copies.
template <typename R>
int analysis_rows(fil_handle* h) {
  if (h->d.plasma && !h->plasma_valid)
    if (int rc = plasma_only<R>(h)) return rc;
  HIPCHK(launch_analyze_rows<R>(h->S.as<cx<R>>(), h->tw.as<cx<R>>(), h->Rg.as<R>(), h->Tg.as<R>(),
                                h->pending_filter ? 1 : 0, (R)h->dt, h->aFr.as<double>(), nullptr, nullptr,
                                h->aImax.as<double>(), h->rows(), h->Nt, 1, 0, h->B, h->st));
  h->pending_filter = false;
  h->acache.valid = false;
  return 0;
}

// Device memory of the device z loop: the loop state, the log, and the copy
// of the state at the start of a chunk (for a stop inside it).
struct AdaptBuffers {
  AdaptState* state;
  double* log;          // [max_steps + 1][4]
  void* S;              // ns() * csz() bytes
  double* re;           // rows() doubles each
  double* rm;
};

// Filament.run's z loop (Filament.jl:129-164) on the device, checked by the
// host every `chunk` steps.  Steps past the stop within a chunk are undone:
// the state at the chunk's start is restored and the chunk's logged steps are
// run again with their dz from the host, which gives the same bits.  With
// `log` given, the logged dz of the steps to run again are read into it (no
// host buffer); without it into a vector.
template <typename R>
int run_adaptive_loop(fil_handle* h, const AdaptBuffers& b, const fil_adapt_params* p, double* z, int max_steps,
                      double* log, int* nsteps, int* reason) {
  if (int rc = corr_materialize(h)) return rc;
  const size_t nb = h->ns() * h->csz(), nr = (size_t)h->rows() * 8;
  AdaptRule rule{};
  rule.dz0 = p->dz0; rule.phimax = p->phimax; rule.phik = p->phik; rule.phip = p->phip; rule.Istop = p->Istop;
  rule.zmax = (float)p->zmax;
  rule.nonlinearity = p->nonlinearity ? 1 : 0;
  rule.plasma = h->d.plasma ? 1 : 0;
  rule.nrows = h->rows();
  rule.max_steps = max_steps;
  const int chunk = p->chunk > 0 ? p->chunk : 32;
  AdaptState* dst = b.state;
  double* dlog = b.log;
  const double* dz_dev = &dst->dz;
  const double* rm = h->d.plasma ? h->rho_max.as<double>() : nullptr;
  AdaptState s{};
  s.z = (float)*z;
  const double head[2] = {(double)s.z, 0.0};    // log row 0: (z, dz) of the start
  // the analysis before the first step (Filament.jl:112), then the first rule evaluation
  if (int rc = analysis_rows<R>(h)) return rc;
  HIPCHK(hipMemcpyAsync(dst, &s, sizeof s, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync(dlog, head, sizeof head, hipMemcpyHostToDevice, h->st));
  HIPCHK(launch_adapt(rule, dst, h->aImax.as<double>(), rm, dlog, h->st));
  HIPCHK(hipMemcpyAsync(&s, dst, sizeof s, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  std::vector<double> lgv;
  while (!s.done) {
    const int k0 = s.steps - 1;                    // steps completed before this chunk
    const int n = std::min(chunk, max_steps - k0);
    const bool plasma_valid0 = h->plasma_valid;
    HIPCHK(hipMemcpyAsync(b.S, h->S.p, nb, hipMemcpyDeviceToDevice, h->st));
    HIPCHK(hipMemcpyAsync(b.re, h->rho_end.p, nr, hipMemcpyDeviceToDevice, h->st));
    HIPCHK(hipMemcpyAsync(b.rm, h->rho_max.p, nr, hipMemcpyDeviceToDevice, h->st));
    for (int i = 0; i < n; i++) {
      if (int rc = zstep_impl<R>(h, 0.0, dz_dev)) return rc;
      if (int rc = analysis_rows<R>(h)) return rc;
      HIPCHK(launch_adapt(rule, dst, h->aImax.as<double>(), rm, dlog, h->st));
    }
    HIPCHK(hipMemcpyAsync(&s, dst, sizeof s, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    // steps the chunk ran that count: s.steps - k0 while running (the last one
    // proposed but not taken), s.steps - k0 once stopped
    const int taken = s.done ? s.steps - k0 : s.steps - 1 - k0;
    if (taken < n) {
      // the stop came inside the chunk: back to its start, then its counted steps again
      double* lg = nullptr;
      if (log) {
        lg = log + 4 * ((size_t)k0 + 1);
      } else {
        lgv.assign((size_t)taken * 4, 0.0);
        lg = lgv.data();
      }
      if (taken)
        HIPCHK(hipMemcpy(lg, dlog + 4 * ((size_t)k0 + 1), (size_t)taken * 4 * 8, hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpyAsync(h->S.p, b.S, nb, hipMemcpyDeviceToDevice, h->st));
      HIPCHK(hipMemcpyAsync(h->rho_end.p, b.re, nr, hipMemcpyDeviceToDevice, h->st));
      HIPCHK(hipMemcpyAsync(h->rho_max.p, b.rm, nr, hipMemcpyDeviceToDevice, h->st));
      h->pending_filter = false;
      h->plasma_valid = plasma_valid0;
      for (int i = 0; i < taken; i++) {
        if (int rc = zstep_impl<R>(h, lg[4 * (size_t)i + 1])) return rc;
        if (int rc = analysis_rows<R>(h)) return rc;
      }
    }
  }
  if (log) HIPCHK(hipMemcpy(log, dlog, ((size_t)s.steps + 1) * 4 * 8, hipMemcpyDeviceToHost));
  HIPCHK(hipStreamSynchronize(h->st));
  *z = (double)s.z;
  *nsteps = s.steps;
  *reason = s.done;
  return 0;
}

// fil_run_adaptive: the loop's buffers are the handle's, allocated on the
// first call that needs them (each on its own: a failed allocation leaves
// that buffer empty and the next call tries it again).
template <typename R>
int run_adaptive_impl(fil_handle* h, const fil_adapt_params* p, double* z, int max_steps, double* log, int* nsteps,
                      int* reason) {
  const size_t nb = h->ns() * h->csz(), nr = (size_t)h->rows() * 8, nlog = ((size_t)max_steps + 1) * 4 * 8;
  auto need = [](DevBuf& b, size_t bytes) {
    if (b.p && b.n >= bytes) return true;
    if (b.alloc(bytes) == hipSuccess) return true;
    b.p = nullptr;
    b.n = 0;
    return false;
  };
  if (!need(h->ad_state, sizeof(AdaptState)) || !need(h->ad_log, nlog) || !need(h->ad_S, nb) ||
      !need(h->ad_re, nr) || !need(h->ad_rm, nr))
    return fail(FIL_ENOMEM, "device allocation failed");
  const AdaptBuffers b{h->ad_state.as<AdaptState>(), h->ad_log.as<double>(), h->ad_S.p, h->ad_re.as<double>(),
                       h->ad_rm.as<double>()};
  return run_adaptive_loop<R>(h, b, p, z, max_steps, log, nsteps, reason);
}

template <typename R>
int analyze_impl(fil_handle* h, int member, double z, fil_diag* d, double* Fr, double* Ft, double* rho_end, double* S) {
  if (int rc = corr_materialize(h)) return rc;
  const int Nr = h->Nr, Nt = h->Nt, Nh = h->Nh;
  AnalyzerRows v;
  if (int rc = analyze_device<R>(h, v, member)) return rc;
  fil_diag o{};
  o.z = z;
  if (h->tgeo) {
    // FieldAnalyzerT (FieldAnalyzers.jl:99-114): Imax, rhomax, duration = radius(t, I),
    // F = sum(I) dt; reported as tau = duration and W = Fmax = F
    o.Imax = v.im[0];
    o.rhomax = v.rm[0];
    o.tau = radius(h->t.data(), v.ft.data(), Nt);
    o.W = o.Fmax = v.fr[0];
    if (d) *d = o;
    if (Fr) Fr[0] = v.fr[0];
    if (Ft) std::memcpy(Ft, v.ft.data(), Nt * 8);
    if (rho_end) rho_end[0] = v.re[0];
    if (S) std::memcpy(S, v.sp.data(), (Nh + 1) * 8);
    return 0;
  }
  // scalar diagnostics (FieldAnalyzers.jl:201-213)
  double Fmax = 0, Imax = 0, rhomax = 0, De = 0, W = 0;
  // maximum() propagates NaN in the reference (FieldAnalyzers.jl:199-203)
  auto nanmax = [](double a, double b) { return (a != a || b != b) ? a + b : std::max(a, b); };
  for (int i = 0; i < Nr; i++) {
    Fmax = nanmax(Fmax, v.fr[i]);
    Imax = nanmax(Imax, v.im[i]);
    rhomax = nanmax(rhomax, v.rm[i]);
    De += v.re[i] * h->rdr[i];
    W += v.fr[i] * h->rdr[i];
  }
  o.Fmax = Fmax; o.Imax = Imax; o.rhomax = rhomax;
  o.De = 2 * kPi * De;
  o.rfil = 2 * radius(h->r.data(), v.fr.data(), Nr);
  o.rpl = 2 * radius(h->r.data(), v.re.data(), Nr);
  o.tau = radius(h->t.data(), v.ft.data(), Nt);
  o.W = 2 * kPi * W;
  if (d) *d = o;
  if (Fr) std::memcpy(Fr, v.fr.data(), Nr * 8);
  if (Ft) std::memcpy(Ft, v.ft.data(), Nt * 8);
  if (rho_end) std::memcpy(rho_end, v.re.data(), Nr * 8);
  if (S) std::memcpy(S, v.sp.data(), (Nh + 1) * 8);
  return 0;
}

// FieldAnalyzers.analyze! XYT (FieldAnalyzers.jl:308-353)
template <typename R>
int analyze_xyt_impl(fil_handle* h, double z, fil_diag_xyt* d, double* Fxy, double* Ft, double* rho_end, double* S) {
  const int Nr = h->Nr, Nt = h->Nt, Nh = h->Nh, Nx = h->Nx, Ny = h->Ny;
  AnalyzerRows v;
  if (int rc = analyze_device<R>(h, v, 0)) return rc;
  fil_diag_xyt o{};
  o.z = z;
  const int imax = (int)(std::max_element(v.fr.begin(), v.fr.end()) - v.fr.begin());   // findmax, x fastest
  o.Fmax = v.fr[imax];
  auto nanmax = [](double a, double b) { return (a != a || b != b) ? a + b : std::max(a, b); };
  o.Imax = 0;
  o.rhomax = 0;
  for (int i = 0; i < Nr; i++) {
    o.Imax = nanmax(o.Imax, v.im[i]);
    o.rhomax = nanmax(o.rhomax, v.rm[i]);
  }
  double De = 0, W = 0;
  for (int i = 0; i < Nr; i++) De += v.re[i];
  o.De = De * h->dx * h->dy;
  for (int j = 0; j < Nt; j++) W += v.ft[j];
  o.W = W * h->dt;
  const int ixm = imax % Nx, iym = imax / Nx;
  std::vector<double> px(Nx), py(Ny);
  for (int ix = 0; ix < Nx; ix++) px[ix] = v.fr[(size_t)ix + (size_t)Nx * iym];
  for (int iy = 0; iy < Ny; iy++) py[iy] = v.fr[(size_t)ixm + (size_t)Nx * iy];
  o.ax = radius(h->x.data(), px.data(), Nx);
  o.ay = radius(h->y.data(), py.data(), Ny);
  if (d) *d = o;
  if (Fxy) std::memcpy(Fxy, v.fr.data(), (size_t)Nr * 8);
  if (Ft) std::memcpy(Ft, v.ft.data(), Nt * 8);
  if (rho_end) std::memcpy(rho_end, v.re.data(), (size_t)Nr * 8);
  if (S) std::memcpy(S, v.sp.data(), (Nh + 1) * 8);
  return 0;
}

bool valid_handle(fil_handle* h) { return h != nullptr; }

// ---------------------------------------------------------------------------
// Learned corrector helpers
// ---------------------------------------------------------------------------
static_assert(corr_nparam() == FIL_CORR_NPARAM, "FIL_CORR_NPARAM must match the network");

uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double uniform01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }
double normal01(uint64_t& s) {
  double u1 = uniform01(s), u2 = uniform01(s);
  if (u1 < 1e-300) u1 = 1e-300;
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
}

int corr_bias_off(int l) { return corr_woff(l) + corr_cout(l) * corr_cin(l) * 9; }

// network forward over B samples of `in` (gathered by idx when given) into
// corr.out, keeping the activations for a backward pass (train); in: pooled
// features [B][4][H][W] (idx selects stored samples), or with from_rows the
// per-row features Frow of one sample (pooled as the kernel reads them)
int corr_forward(fil_handle* h, const float* in, const int* idx, int B, bool from_rows = false, bool train = false) {
  auto& c = h->corr;
  float* act[CORR_LAYERS - 1];
  for (int l = 0; l < CORR_LAYERS - 1; l++) act[l] = c.act[l].as<float>();
  HIPCHK(launch_conv_net(in, idx, from_rows ? c.g.pr : 0, c.wfr.as<uint16_t>(), c.w.as<float>(),
                         train ? act : nullptr, c.out.as<float>(), B, c.g.H, c.g.W, h->st));
  return 0;
}

// gradient of the loss (corr.gout already filled) into corr.grad
int corr_backward(fil_handle* h, const float* in, const int* idx, int B) {
  auto& c = h->corr;
  const uint16_t* wb = c.wbf.as<uint16_t>();
  float* gr = c.grad.as<float>();
  const int H = c.g.H, W = c.g.W;
  float* g = c.gout.as<float>();
  float* bufs[2] = {c.gA.as<float>(), c.gB.as<float>()};
  for (int l = CORR_LAYERS - 1; l >= 0; l--) {
    const float* a_in = (l == 0) ? in : c.act[l - 1].as<float>();
    HIPCHK(launch_conv_bwd_weight(l, g, a_in, l == 0 ? idx : nullptr, gr + corr_woff(l), gr + corr_bias_off(l),
                                  c.wpart.as<float>(), B, H, W, h->st));
    if (l > 0) {
      float* d = bufs[l & 1];
      HIPCHK(launch_conv_bwd_data(l, g, wb + corr_woff(l), c.act[l - 1].as<float>(), d, B, H, W, h->st));
      g = d;
    }
  }
  return 0;
}

// loss over B samples whose targets are corr.Y[idx[b]]; fills corr.gout; returns the partial count
int corr_loss(fil_handle* h, const int* idx, int B, int* nblk) {
  auto& c = h->corr;
  HIPCHK(launch_corr_loss(c.out.as<float>(), c.Y.as<float>(), idx, (float)(1.0 / c.y_scale), c.gout.as<float>(),
                          c.part.as<double>(), B, c.g.H * c.g.W, nblk, h->st));
  return 0;
}

int corr_read_loss(fil_handle* h, int nblk, int B, double* loss) {
  std::vector<double> part(nblk);
  HIPCHK(hipMemcpyAsync(part.data(), h->corr.part.p, nblk * sizeof(double), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  double s = 0;
  for (double v : part) s += v;
  if (loss) *loss = s / ((double)B * 2 * h->corr.g.H * h->corr.g.W);
  return 0;
}

// Corrector entry points start here: the previous apply's output (corr.out)
// is added to the state before anything can overwrite it.
int corr_ready(fil_handle* h) {
  if (!h) return fail(FIL_EINVAL, "null handle");
  if (!h->corr.on) return fail(FIL_ESTATE, "fil_corrector_init first");
  if (h->corr_pending) {
    HIPCHK(hipSetDevice(h->d.device));
    if (int rc = corr_materialize(h)) return rc;
  }
  return 0;
}

}  // namespace

extern "C" {

const char* fil_last_error(void) { return g_err.c_str(); }

int fil_create(const fil_desc* d, fil_handle** out) {
  if (!d || !out) return fail(FIL_EINVAL, "null argument");
  *out = nullptr;
  if (d->abi_version != FIL_ABI_VERSION) return fail(FIL_EINVAL, "ABI version mismatch");
  if (d->geometry != FIL_GEOM_RT && d->geometry != FIL_GEOM_XYT && d->geometry != FIL_GEOM_T)
    return fail(FIL_EINVAL, "geometry must be RT, XYT or T");
  const bool xyt = d->geometry == FIL_GEOM_XYT;
  const bool tgeo = d->geometry == FIL_GEOM_T;
  auto pow2 = [](int n) { return n > 0 && (n & (n - 1)) == 0; };
  if (xyt && (!pow2(d->Nx) || !pow2(d->Ny) || d->Nx < 16 || d->Ny < 16 || d->Nx > 1024 || d->Ny > 1024))
    return fail(FIL_EINVAL, "XYT: Nx and Ny must be powers of two in [16, 1024]");
  if (xyt && d->guard3d_bug && d->Ny > d->Nt) return fail(FIL_EINVAL, "guard3d_bug needs Ny <= Nt (the reference indexes the t window with y)");
  const int batch = d->batch < 1 ? 1 : d->batch;
  if (batch > 1 && xyt) return fail(FIL_EINVAL, "batches are defined for the RT geometry");
  if (batch > 4096) return fail(FIL_EINVAL, "batch must be <= 4096");
  if (d->precision != FIL_FP32 && d->precision != FIL_FP64) return fail(FIL_EINVAL, "precision must be FIL_FP32 or FIL_FP64");
  const int Nt = d->Nt;
  if (Nt < 256 || Nt > 8192 || (Nt & (Nt - 1))) return fail(FIL_EINVAL, "Nt must be a power of two in [256, 8192]");
  if (d->precision == FIL_FP64 && Nt > 4096) return fail(FIL_EINVAL, "FP64 supports Nt <= 4096");
  if (!xyt && !tgeo && (d->Nr < 2 || d->Nr > 16384)) return fail(FIL_EINVAL, "Nr must be in [2, 16384]");
  if (tgeo && (d->form & FIL_FORM_DHT_BF16X3)) return fail(FIL_EINVAL, "the T geometry has no Hankel transform");
  if (d->plasma && (d->ncomp < 1 || d->ncomp > FIL_MAXCOMP)) return fail(FIL_EINVAL, "ncomp must be in [1, 4]");
  if ((d->form & FIL_FORM_DHT_BF16X3) && d->precision != FIL_FP32)
    return fail(FIL_EINVAL, "the split-BF16 Hankel form needs FIL_FP32");
  if (d->nonparaxial && ((d->geometry != FIL_GEOM_RT && d->geometry != FIL_GEOM_XYT) ||
                         (d->form & FIL_FORM_DHT_BF16X3)))
    return fail(FIL_EINVAL, "the non-paraxial nonlinear term is built for the RT geometry with the FP32/FP64 Hankel "
                            "products and for XYT");
  if ((d->responses & FIL_RESP_RAMAN_THG) && (d->responses & FIL_RESP_RAMAN_NOTHG))
    return fail(FIL_EINVAL, "choose one Raman response");
  if ((d->responses & FIL_RESP_CUBIC_THG) && (d->responses & FIL_RESP_CUBIC_NOTHG))
    return fail(FIL_EINVAL, "choose one cubic response");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (d->device < 0 || d->device >= ndev) return fail(FIL_EINVAL, "no such HIP device");
  HIPCHK(hipSetDevice(d->device));
  fil_handle* h = new fil_handle();
  h->d = *d;
  h->xyt = xyt;
  h->tgeo = tgeo;
  h->Nx = xyt ? d->Nx : 0;
  h->Ny = xyt ? d->Ny : 0;
  h->Nr = xyt ? d->Nx * d->Ny : (tgeo ? 1 : d->Nr);
  h->B = batch;
  h->Nrp = tgeo ? 1 : (h->Nr + 63) / 64 * 64;
  h->Nt = Nt;
  h->Nh = Nt / 2;
  h->fp64 = d->precision == FIL_FP64;
  h->bf16x3 = (d->form & FIL_FORM_DHT_BF16X3) != 0;
  h->scan64 = (d->form & FIL_FORM_SCAN_FP64) != 0 || h->fp64;
  hipError_t e = hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking);
  if (e != hipSuccess) { delete h; return fail(FIL_EDEVICE, hipGetErrorString(e)); }
  const size_t cz = h->csz(), rz = h->rsz();
  // P: the propagator [Nrp][Nh]; XYT sizes its own when the tables are set
  const size_t NS = h->ns(), NP = (size_t)h->Nrp * h->Nh, NR = (size_t)h->rows();
  bool ok = h->S.alloc(NS * cz) == hipSuccess && h->X.alloc(xyt || tgeo ? 0 : NS * cz) == hipSuccess &&
            h->P.alloc(xyt ? 0 : NP * cz) == hipSuccess && h->Et.alloc((size_t)h->Nrp * Nt * cz) == hipSuccess &&
            h->rho_end.alloc(NR * 8) == hipSuccess && h->rho_max.alloc(NR * 8) == hipSuccess &&
            h->aFr.alloc(NR * 8) == hipSuccess && h->aImax.alloc(NR * 8) == hipSuccess &&
            h->aIt.alloc(NR * Nt * rz) == hipSuccess && h->aSw.alloc(NR * (h->Nh + 1) * rz) == hipSuccess &&
            h->aFt.alloc((size_t)batch * Nt * 8) == hipSuccess &&
            h->aS.alloc((size_t)batch * (h->Nh + 1) * 8) == hipSuccess &&
            h->acol.alloc(std::min<size_t>(COLSUM_MAXCHUNK, (h->Nr + 63) / 64) * batch * Nt * 8) == hipSuccess;
  if (ok && d->keep_plasma)
    ok = h->rho_full.alloc(NR * Nt * rz) == hipSuccess && h->kd_full.alloc(NR * Nt * rz) == hipSuccess;
  h->nonparax = d->nonparaxial && d->nonlinearity;
  if (ok && h->nonparax)
    ok = h->npQ.alloc(xyt ? 0 : NP * cz) == hipSuccess && h->npR.alloc(NS * cz) == hipSuccess &&
         h->npK1.alloc(NS * cz) == hipSuccess && h->npK2.alloc(NS * cz) == hipSuccess &&
         h->npU.alloc(NS * cz) == hipSuccess &&
         h->np_rho.alloc(d->plasma ? NR * Nt * rz : 0) == hipSuccess &&
         h->np_kd.alloc(d->plasma ? NR * Nt * rz : 0) == hipSuccess;
  if (!ok) { delete h; return fail(FIL_ENOMEM, "device allocation failed"); }
  // the padding rows of S and X must be zero for the Hankel products
  hipError_t em = hipMemset(h->S.p, 0, NS * cz);
  if (em == hipSuccess && !xyt && !tgeo) em = hipMemset(h->X.p, 0, NS * cz);
  if (em == hipSuccess) em = hipMemset(h->rho_end.p, 0, NR * 8);
  if (em == hipSuccess) em = hipMemset(h->rho_max.p, 0, NR * 8);
  if (em == hipSuccess && d->keep_plasma) em = hipMemset(h->rho_full.p, 0, NR * Nt * rz);
  if (em == hipSuccess && d->keep_plasma) em = hipMemset(h->kd_full.p, 0, NR * Nt * rz);
  // the stage buffers' padding rows enter the Hankel products (times zero)
  for (DevBuf* b : {&h->npQ, &h->npR, &h->npK1, &h->npK2, &h->npU})
    if (em == hipSuccess && b->p) em = hipMemset(b->p, 0, b->n);
  if (em != hipSuccess) {
    delete h;
    return fail(FIL_EDEVICE, std::string("clearing the device state failed: ") + hipGetErrorString(em));
  }
  *out = h;
  return FIL_OK;
}

}  // extern "C"

namespace {


int set_tables_impl(fil_handle* h, const fil_tables* t, const fil_tables_xyt* xt) {
  const int Nr = h->Nr, Nrp = h->Nrp, Nt = h->Nt, Nh = h->Nh;
  const bool f64 = h->fp64;
  const bool bug = h->xyt && h->d.guard3d_bug;
  if (!t->Wguard || !t->Tguard || !t->t) return fail(FIL_EINVAL, "missing t grid or guard table");
  if (!(t->dt > 0)) return fail(FIL_EINVAL, "dt must be positive");
  h->dt = t->dt;
  h->t.assign(t->t, t->t + Nt);
  std::vector<double> wt(Nr);    // analyzer weights of the rows
  if (h->tgeo) {
    if (!t->KZ) return fail(FIL_EINVAL, "missing linear-propagator table");
    wt[0] = 1.0;                  // the analyzer's Ft is then I(t)
  } else if (!h->xyt) {
    if (!t->r || !t->dr || !t->KZ || !t->Kguard || !t->Rguard || !t->dht_T || !t->dht_RJ || !t->dht_JV)
      return fail(FIL_EINVAL, "missing linear-propagator table");
    h->r.assign(t->r, t->r + Nr);
    h->rdr.resize(Nr);
    for (int i = 0; i < Nr; i++) { h->rdr[i] = t->r[i] * t->dr[i]; wt[i] = 2 * kPi * h->rdr[i]; }
  } else {
    if (!xt->x || !xt->y || !xt->kt2x || !xt->kt2y || !xt->Xguard || !xt->Yguard || !xt->KXguard || !xt->KYguard ||
        !xt->beta || !xt->w_vf)
      return fail(FIL_EINVAL, "missing XYT table");
    if (!(xt->dx > 0) || !(xt->dy > 0) || !(xt->zu > 0)) return fail(FIL_EINVAL, "dx, dy and zu must be positive");
    h->x.assign(xt->x, xt->x + h->Nx);
    h->y.assign(xt->y, xt->y + h->Ny);
    h->dx = xt->dx;
    h->dy = xt->dy;
    for (int i = 0; i < Nr; i++) wt[i] = xt->dx * xt->dy;
  }
  HIPCHK(upload(h->ardr, wt.data(), Nr * 8));

  if (h->xyt) {
    // propagator inputs (spectral.jl:131-184, Guards.jl:162-195) and the x / y FFT twiddles
    std::vector<double> beta(2 * (size_t)Nh), wvf(Nh);
    for (int w = 0; w < Nh; w++) { beta[2 * w] = xt->beta[2 * w]; beta[2 * w + 1] = xt->beta[2 * w + 1]; wvf[w] = xt->w_vf[w]; }
    HIPCHK(upload(h->xkt2x, xt->kt2x, h->Nx * 8));
    HIPCHK(upload(h->xkt2y, xt->kt2y, h->Ny * 8));
    HIPCHK(upload(h->xKX, xt->KXguard, h->Nx * 8));
    // KY(iy) Wguard(iy) with the reference's 3-D filter quirk (Guards.jl:314);
    // the stored quarter of P needs kt2 and the guards even in kx and in ky
    std::vector<double> gy(xt->KYguard, xt->KYguard + h->Ny);
    if (bug)
      for (int j = 0; j < h->Ny; j++) gy[j] *= t->Wguard[j];
    int xsym = 1, ysym = 1;
    for (int i = 1; i < h->Nx / 2; i++)
      if (xt->kt2x[i] != xt->kt2x[h->Nx - i] || xt->KXguard[i] != xt->KXguard[h->Nx - i]) xsym = 0;
    for (int j = 1; j < h->Ny / 2; j++)
      if (xt->kt2y[j] != xt->kt2y[h->Ny - j] || gy[j] != gy[h->Ny - j]) ysym = 0;
    HIPCHK(upload(h->xKY, gy.data(), h->Ny * 8));
    HIPCHK(upload(h->xbeta, beta.data(), beta.size() * 8));
    HIPCHK(upload(h->xwvf, wvf.data(), wvf.size() * 8));
    HIPCHK(upload(h->xWg, t->Wguard, (size_t)Nt * 8));
    XytProp& xp = h->xprop;
    xp.Nx = h->Nx; xp.Ny = h->Ny; xp.Nh = Nh;
    xp.kparaxial = xt->kparaxial != 0;
    xp.guard3d_bug = bug ? 1 : 0;
    xp.xsym = xsym;
    xp.ysym = ysym;
    xp.Px = xsym ? h->Nx / 2 + 1 : h->Nx;
    xp.Py = ysym ? h->Ny / 2 + 1 : h->Ny;
    HIPCHK(h->P.alloc((size_t)xp.Px * xp.Py * Nh * h->csz()));
    xp.kt2x = h->xkt2x.as<double>(); xp.kt2y = h->xkt2y.as<double>();
    xp.KX = h->xKX.as<double>(); xp.KY = h->xKY.as<double>();
    xp.beta = h->xbeta.as<double>(); xp.wvf = h->xwvf.as<double>(); xp.Wg = h->xWg.as<double>();
    xp.zu = xt->zu;
    xp.norm = 1.0 / ((double)h->Nx * h->Ny);
    h->P_dz = NAN;
    if (h->nonparax) {
      // the non-paraxial prefactor table, in the propagator's layout (dz-independent)
      HIPCHK(h->npQ.alloc((size_t)xp.Px * xp.Py * Nh * h->csz()));
      HIPCHK(h->fp64 ? launch_xyt_qz<double>(xp, h->npQ.as<cx<double>>(), h->st)
                     : launch_xyt_qz<float>(xp, h->npQ.as<cx<float>>(), h->st));
    }
    if (int rc = build_twiddles(h->twx, h->Nx, f64, false)) return rc;
    if (int rc = build_twiddles(h->twy, h->Ny, f64, false)) return rc;
  }

  if (h->tgeo) {
    // T: KZ(w) and Wguard(w) (spectral.jl:55-75, Guards.jl:78-89); the linear
    // step is the diagonal multiply by exp(i KZ dz) Wguard
    std::vector<double> kz((size_t)Nh * 2), kg(Nh);
    for (int w = 0; w < Nh; w++) {
      kz[2 * w] = t->KZ[2 * w];
      kz[2 * w + 1] = t->KZ[2 * w + 1];
      kg[w] = t->Wguard[w];
    }
    HIPCHK(upload(h->KZ64, kz.data(), kz.size() * 8));
    HIPCHK(upload(h->KG64, kg.data(), kg.size() * 8));
    h->P_dz = NAN;
  }
  if (!h->xyt && !h->tgeo) {
  // folded QDHT matrices M1 = T diag(RJ), M2 = diag(1/RJ) T, zero padded
  std::vector<double> M1((size_t)Nrp * Nrp, 0.0), M2((size_t)Nrp * Nrp, 0.0);
  for (int m = 0; m < Nr; m++)
    for (int n = 0; n < Nr; n++) {
      double T = t->dht_T[m + (size_t)Nr * n];
      M1[(size_t)m * Nrp + n] = T * t->dht_RJ[n];
      M2[(size_t)m * Nrp + n] = T / t->dht_RJ[m];
    }
  {
    // stored transposed (MT[k][m]) for the FP32 / FP64 GEMM kernels
    HostArr a(M1.size(), f64, false), b(M2.size(), f64, false);
    for (int m = 0; m < Nrp; m++)
      for (int n = 0; n < Nrp; n++) {
        a.set((size_t)n * Nrp + m, M1[(size_t)m * Nrp + n]);
        b.set((size_t)n * Nrp + m, M2[(size_t)m * Nrp + n]);
      }
    HIPCHK(upload(h->M1, a.b.data(), a.b.size()));
    HIPCHK(upload(h->M2, b.b.data(), b.b.size()));
  }
  if (h->bf16x3) {
    std::vector<uint16_t> h1(M1.size()), l1(M1.size()), h2(M2.size()), l2(M2.size());
    for (size_t q = 0; q < M1.size(); q++) {
      float back, back2;
      float v = (float)M1[q];
      h1[q] = to_bf16(v, &back);
      l1[q] = to_bf16((float)(M1[q] - (double)back), &back2);
      v = (float)M2[q];
      h2[q] = to_bf16(v, &back);
      l2[q] = to_bf16((float)(M2[q] - (double)back), &back2);
    }
    HIPCHK(upload(h->M1h, h1.data(), h1.size() * 2));
    HIPCHK(upload(h->M1l, l1.data(), l1.size() * 2));
    HIPCHK(upload(h->M2h, h2.data(), h2.size() * 2));
    HIPCHK(upload(h->M2l, l2.data(), l2.size() * 2));
  }
  // propagator inputs in FP64: KZ and Kguard*Wguard on [Nrp][Nh]
  {
    std::vector<double> kz((size_t)Nrp * Nh * 2, 0.0), kg((size_t)Nrp * Nh, 0.0);
    for (int m = 0; m < Nr; m++)
      for (int w = 0; w < Nh; w++) {
        size_t src = m + (size_t)Nr * w;
        kz[2 * ((size_t)m * Nh + w)] = t->KZ[2 * src];
        kz[2 * ((size_t)m * Nh + w) + 1] = t->KZ[2 * src + 1];
        kg[(size_t)m * Nh + w] = t->Kguard[src] * t->Wguard[w];
      }
    HIPCHK(upload(h->KZ64, kz.data(), kz.size() * 8));
    HIPCHK(upload(h->KG64, kg.data(), kg.size() * 8));
    h->P_dz = NAN;
  }
  if (h->nonparax) {
    // Q(k, w) = QZ(k, w) / QZ(w) Kguard Wguard: the non-paraxial prefactor
    // (nonlinear_propagators.jl:139-160, 232-242) over the paraxial one the
    // stage right-hand side carries, with func!'s spectral filter (:61)
    if (!t->QZ_kw || !t->QZ) return fail(FIL_EINVAL, "missing QZ_kw or QZ table");
    HostArr q((size_t)Nrp * Nh, f64, true);
    for (int m = 0; m < Nr; m++)
      for (int w = 0; w < Nh; w++) {
        const size_t src = m + (size_t)Nr * w;
        const std::complex<double> qk(t->QZ_kw[2 * src], t->QZ_kw[2 * src + 1]), q0(t->QZ[2 * w], t->QZ[2 * w + 1]);
        const std::complex<double> v = (q0 == 0.0) ? 0.0 : qk / q0 * (t->Kguard[src] * t->Wguard[w]);
        q.setc((size_t)m * Nh + w, v.real(), v.imag());
      }
    HIPCHK(upload(h->npQ, q.b.data(), q.b.size()));
  }
  }  // RT linear-step tables
  // twiddles, field guards: Rg per row, Tg per t sample (Guards.jl:265-318).  The
  // XYT filter is X(x) Y(y) T(t); with guard3d_bug the reference's kernel
  // multiplies by T at the y index instead (Guards.jl:314), which the tables
  // express as Rg = X Y T(iy) and Tg = 1.
  {
    if (int rc = build_twiddles(h->tw, Nt, f64, true)) return rc;
    HostArr tg(Nt, f64, false), rg(Nr, f64, false);
    for (int j = 0; j < Nt; j++) tg.set(j, bug ? 1.0 : t->Tguard[j]);
    if (h->tgeo) {
      rg.set(0, 1.0);             // Guard1D: the field filter is Tguard only
    } else if (!h->xyt) {
      for (int i = 0; i < Nr; i++) rg.set(i, t->Rguard[i]);
    } else {
      for (int iy = 0; iy < h->Ny; iy++)
        for (int ix = 0; ix < h->Nx; ix++)
          rg.set((size_t)ix + (size_t)h->Nx * iy, xt->Xguard[ix] * xt->Yguard[iy] * (bug ? t->Tguard[iy] : 1.0));
    }
    HIPCHK(upload(h->Tg, tg.b.data(), tg.b.size()));
    HIPCHK(upload(h->Rg, rg.b.data(), rg.b.size()));
  }
  // response coefficients i QZ Rnl d(w) / (2 Nt), d(0) = 1, d(w>0) = 2
  {
    HostArr cf((size_t)4 * Nh, f64, true);
    const double* R[4] = {nullptr, nullptr, nullptr, nullptr};
    const bool raman = h->d.nonlinearity && (h->d.responses & (FIL_RESP_RAMAN_THG | FIL_RESP_RAMAN_NOTHG));
    if (h->d.nonlinearity) {
      if (h->d.responses & (FIL_RESP_CUBIC_THG | FIL_RESP_CUBIC_NOTHG)) R[0] = t->Rnl_cubic;
      if (h->d.responses & FIL_RESP_FREE) R[1] = t->Rnl_free;
      if (h->d.responses & FIL_RESP_LOSSES) R[2] = t->Rnl_losses;
      if (raman) R[3] = t->Rnl_raman;
      if (!t->QZ) return fail(FIL_EINVAL, "missing QZ table");
      for (int q = 0; q < 3; q++)
        if ((q == 0 && (h->d.responses & (FIL_RESP_CUBIC_THG | FIL_RESP_CUBIC_NOTHG)) && !R[0]) ||
            (q == 1 && (h->d.responses & FIL_RESP_FREE) && !R[1]) ||
            (q == 2 && (h->d.responses & FIL_RESP_LOSSES) && !R[2]))
          return fail(FIL_EINVAL, "missing Rnl table");
      if (raman && (!R[3] || !t->Hraman)) return fail(FIL_EINVAL, "missing Raman tables (Rnl_raman, Hraman)");
    }
    // Raman response spectrum (raman.jl:15-29), divided by Nt: the
    // convolution (raman.jl:68-111) is FFT-(Hraman .* (1/Nt) FFT+(x))
    if (raman) {
      HostArr hv((size_t)Nt, f64, true);
      for (int j = 0; j < Nt; j++) hv.setc(j, t->Hraman[2 * j] / Nt, t->Hraman[2 * j + 1] / Nt);
      HIPCHK(upload(h->hr, hv.b.data(), hv.b.size()));
    }
    // Each response's real signal is scaled by s_q = median_w |i QZ Rnl| before
    // it shares a complex FFT with another one; the table holds coef / s_q.
    for (int q = 0; q < 4; q++) {
      std::vector<std::complex<double>> c(Nh, std::complex<double>(0, 0));
      std::vector<double> mag;
      if (R[q]) {
        for (int w = 0; w < Nh; w++) {
          std::complex<double> qz(t->QZ[2 * w], t->QZ[2 * w + 1]);
          std::complex<double> rn(R[q][2 * w], R[q][2 * w + 1]);
          c[w] = std::complex<double>(0, 1) * qz * rn;
          if (w > 0 && std::abs(c[w]) > 0) mag.push_back(std::abs(c[w]));
        }
      }
      double s = 1.0;
      if (!mag.empty()) {
        std::nth_element(mag.begin(), mag.begin() + mag.size() / 2, mag.end());
        s = mag[mag.size() / 2];
      }
      h->scale[q] = s;
      for (int w = 0; w < Nh; w++) {
        std::complex<double> v = c[w] * ((w == 0) ? 1.0 : 2.0) / (2.0 * Nt * s);
        cf.setc((size_t)q * Nh + w, v.real(), v.imag());
      }
    }
    HIPCHK(upload(h->coef, cf.b.data(), cf.b.size()));
  }
  // plasma parameters
  if (h->d.plasma) {
    PlasmaParams<double> p{};
    p.ncomp = h->d.ncomp;
    p.ereal = h->d.plasma_ereal;
    p.kdep = h->d.plasma_kdep;
    p.rho0 = t->rho0;
    p.avalanche = 0;
    const double L2 = std::log2(10.0);
    PlasmaParams<float> p32{};
    for (int q = 0; q < h->d.ncomp; q++) {
      int n = t->ntab[q];
      if (n < 2 || !t->tab_log10x[q] || !t->tab_log10y[q]) return fail(FIL_EINVAL, "rate table needs >= 2 nodes");
      const double* lx = t->tab_log10x[q];
      const double* ly = t->tab_log10y[q];
      double d0 = lx[1] - lx[0];
      if (!(d0 > 0)) return fail(FIL_EINVAL, "rate table x must increase");
      for (int m = 1; m < n; m++)
        if (std::fabs((lx[m] - lx[m - 1]) - d0) > 1e-9 * std::fabs(d0) + 1e-12)
          return fail(FIL_EINVAL, "rate table must be uniformly spaced in log10(x)");
      std::vector<double> ly2(n), sl(n - 1);
      for (int m = 0; m < n; m++) ly2[m] = ly[m] * L2;
      for (int m = 0; m < n - 1; m++) sl[m] = (ly[m + 1] - ly[m]) / (lx[m + 1] - lx[m]);
      p.frhont[q] = t->frhont[q];
      p.Rava[q] = t->Rava[q];
      if (t->Rava[q] != 0) p.avalanche = 1;
      p.K[q] = t->K[q];
      p.ntab[q] = n;
      p.lx0[q] = lx[0] * L2;
      p.dlx[q] = d0 * L2;
      p.inv_dlx[q] = 1.0 / (d0 * L2);
      p.xmin[q] = std::pow(10.0, lx[0]);
      if (f64) {
        HIPCHK(upload(h->tab_ly[q], ly2.data(), n * 8));
        HIPCHK(upload(h->tab_sl[q], sl.data(), (n - 1) * 8));
        p.ly[q] = h->tab_ly[q].as<double>();
        p.slope[q] = h->tab_sl[q].as<double>();
      } else {
        std::vector<float> a(ly2.begin(), ly2.end()), b(sl.begin(), sl.end());
        HIPCHK(upload(h->tab_ly[q], a.data(), n * 4));
        HIPCHK(upload(h->tab_sl[q], b.data(), (n - 1) * 4));
        p32.ly[q] = h->tab_ly[q].as<float>();
        p32.slope[q] = h->tab_sl[q].as<float>();
      }
    }
    h->pp64 = p;
    p32.ncomp = p.ncomp; p32.ereal = p.ereal; p32.kdep = p.kdep; p32.avalanche = p.avalanche;
    p32.rho0 = (float)p.rho0;
    for (int q = 0; q < FIL_MAXCOMP; q++) {
      p32.frhont[q] = (float)p.frhont[q]; p32.Rava[q] = (float)p.Rava[q]; p32.K[q] = (float)p.K[q];
      p32.ntab[q] = p.ntab[q]; p32.lx0[q] = (float)p.lx0[q]; p32.dlx[q] = (float)p.dlx[q];
      p32.inv_dlx[q] = (float)p.inv_dlx[q]; p32.xmin[q] = (float)p.xmin[q];
    }
    h->pp32 = p32;
  }
  h->tables = true;
  h->acache.valid = false;
  return FIL_OK;
}

}  // namespace

extern "C" {

int fil_set_tables(fil_handle* h, const fil_tables* t) {
  if (!valid_handle(h) || !t) return fail(FIL_EINVAL, "null argument");
  if (h->xyt) return fail(FIL_EINVAL, "XYT handles take fil_set_tables_xyt");
  if (in_live_group(h))
    return fail(FIL_ESTATE, "the handle belongs to a group whose graphs hold its tables: destroy the group first");
  HIPCHK(hipSetDevice(h->d.device));
  return set_tables_impl(h, t, nullptr);
}

int fil_set_tables_xyt(fil_handle* h, const fil_tables* t, const fil_tables_xyt* x) {
  if (!valid_handle(h) || !t || !x) return fail(FIL_EINVAL, "null argument");
  if (!h->xyt) return fail(FIL_EINVAL, "fil_set_tables_xyt needs an XYT handle");
  if (in_live_group(h))
    return fail(FIL_ESTATE, "the handle belongs to a group whose graphs hold its tables: destroy the group first");
  HIPCHK(hipSetDevice(h->d.device));
  return set_tables_impl(h, t, x);
}

int fil_upload_member(fil_handle* h, int member, const double* E, int64_t ld) {
  if (!valid_handle(h) || !E) return fail(FIL_EINVAL, "null argument");
  if (!h->tables) return fail(FIL_ESTATE, "fil_set_tables first");
  if (member < 0 || member >= h->B) return fail(FIL_EINVAL, "member out of range");
  if (ld < h->Nr) return fail(FIL_EINVAL, "ld < Nr");
  HIPCHK(hipSetDevice(h->d.device));
  // the other members' pending filter is applied before this one is replaced
  if (h->B > 1 && h->field) {
    int rc = h->fp64 ? apply_pending_filter<double>(h) : apply_pending_filter<float>(h);
    if (rc) return rc;
  }
  const int Nr = h->Nr, Nt = h->Nt;
  HostArr a((size_t)h->Nrp * Nt, h->fp64, true);
  for (int j = 0; j < Nt; j++)
    for (int i = 0; i < Nr; i++) {
      size_t s = (size_t)i + (size_t)ld * j;
      a.setc((size_t)i * Nt + j, E[2 * s], E[2 * s + 1]);
    }
  HIPCHK(hipMemcpyAsync(h->Et.p, a.b.data(), a.b.size(), hipMemcpyHostToDevice, h->st));
  if (h->fp64)
    HIPCHK(launch_to_spectrum<double>(h->Et.as<cx<double>>(), h->S.as<cx<double>>(), h->tw.as<cx<double>>(), Nr, Nt,
                                      h->B, member, h->st));
  else
    HIPCHK(launch_to_spectrum<float>(h->Et.as<cx<float>>(), h->S.as<cx<float>>(), h->tw.as<cx<float>>(), Nr, Nt,
                                     h->B, member, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  if (h->uploaded.size() != (size_t)h->B) h->uploaded.assign(h->B, 0);
  h->uploaded[member] = 1;
  h->acache.valid = false;
  h->field = std::all_of(h->uploaded.begin(), h->uploaded.end(), [](char c) { return c != 0; });
  h->pending_filter = false;
  h->plasma_valid = false;
  h->corr_pending = false;
  return FIL_OK;
}

int fil_upload_field(fil_handle* h, const double* E, int64_t ld) {
  if (!valid_handle(h)) return fail(FIL_EINVAL, "null handle");
  if (h->B > 1) return fail(FIL_EINVAL, "batched handles take fil_upload_member");
  return fil_upload_member(h, 0, E, ld);
}

int fil_download_member(fil_handle* h, int member, double* E, int64_t ld) {
  if (!valid_handle(h) || !E) return fail(FIL_EINVAL, "null argument");
  if (!h->field) return fail(FIL_ESTATE, "no field uploaded");
  if (member < 0 || member >= h->B) return fail(FIL_EINVAL, "member out of range");
  if (ld < h->Nr) return fail(FIL_EINVAL, "ld < Nr");
  HIPCHK(hipSetDevice(h->d.device));
  if (int rc = corr_materialize(h)) return rc;
  const int Nr = h->Nr, Nt = h->Nt;
  if (h->fp64)
    HIPCHK(launch_to_time<double>(h->S.as<cx<double>>(), h->Et.as<cx<double>>(), h->tw.as<cx<double>>(),
                                  h->Rg.as<double>(), h->Tg.as<double>(), h->pending_filter, Nr, Nt, h->B, member,
                                  h->st));
  else
    HIPCHK(launch_to_time<float>(h->S.as<cx<float>>(), h->Et.as<cx<float>>(), h->tw.as<cx<float>>(),
                                 h->Rg.as<float>(), h->Tg.as<float>(), h->pending_filter, Nr, Nt, h->B, member,
                                 h->st));
  std::vector<unsigned char> buf((size_t)Nr * Nt * h->csz());
  HIPCHK(hipMemcpyAsync(buf.data(), h->Et.p, buf.size(), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  for (int i = 0; i < Nr; i++)
    for (int j = 0; j < Nt; j++) {
      size_t s = (size_t)i * Nt + j, o = (size_t)i + (size_t)ld * j;
      if (h->fp64) {
        E[2 * o] = reinterpret_cast<double*>(buf.data())[2 * s];
        E[2 * o + 1] = reinterpret_cast<double*>(buf.data())[2 * s + 1];
      } else {
        E[2 * o] = reinterpret_cast<float*>(buf.data())[2 * s];
        E[2 * o + 1] = reinterpret_cast<float*>(buf.data())[2 * s + 1];
      }
    }
  return FIL_OK;
}

int fil_download_field(fil_handle* h, double* E, int64_t ld) {
  if (!valid_handle(h)) return fail(FIL_EINVAL, "null handle");
  if (h->B > 1) return fail(FIL_EINVAL, "batched handles take fil_download_member");
  return fil_download_member(h, 0, E, ld);
}

int fil_zstep(fil_handle* h, double z, double dz) {
  (void)z;
  if (!valid_handle(h)) return fail(FIL_EINVAL, "null handle");
  if (!h->tables || !h->field) return fail(FIL_ESTATE, "tables and field must be set before fil_zstep");
  if (!std::isfinite(dz)) return fail(FIL_EINVAL, "dz must be finite");
  HIPCHK(hipSetDevice(h->d.device));
  return h->fp64 ? zstep_impl<double>(h, dz) : zstep_impl<float>(h, dz);
}

int fil_zstep_parts(fil_handle* h, double z, double dz_nonlinear, double dz_linear) {
  (void)z;
  if (!valid_handle(h)) return fail(FIL_EINVAL, "null handle");
  if (!h->tables || !h->field) return fail(FIL_ESTATE, "tables and field must be set before fil_zstep_parts");
  if (!std::isfinite(dz_nonlinear) || !std::isfinite(dz_linear)) return fail(FIL_EINVAL, "dz must be finite");
  if (dz_nonlinear < 0) return fail(FIL_EINVAL, "dz_nonlinear must be >= 0");
  if (h->nonparax && dz_nonlinear != dz_linear)
    return fail(FIL_EINVAL, "the non-paraxial nonlinear term takes equal distances (fil_zstep)");
  HIPCHK(hipSetDevice(h->d.device));
  return h->fp64 ? zstep_impl<double>(h, dz_nonlinear, nullptr, dz_linear)
                 : zstep_impl<float>(h, dz_nonlinear, nullptr, dz_linear);
}

int fil_run_adaptive(fil_handle* h, const fil_adapt_params* p, double* z, int max_steps, double* log, int* nsteps,
                     int* reason) {
  if (!valid_handle(h) || !p || !z || !nsteps || !reason) return fail(FIL_EINVAL, "null argument");
  if (!h->tables || !h->field) return fail(FIL_ESTATE, "tables and field must be set before fil_run_adaptive");
  if (h->B != 1) return fail(FIL_EINVAL, "fil_run_adaptive takes single-pulse handles");
  if (max_steps < 0 || p->chunk < 0) return fail(FIL_EINVAL, "max_steps and chunk must be >= 0");
  if (!std::isfinite(p->dz0) || !std::isfinite(*z)) return fail(FIL_EINVAL, "dz0 and z must be finite");
  HIPCHK(hipSetDevice(h->d.device));
  return h->fp64 ? run_adaptive_impl<double>(h, p, z, max_steps, log, nsteps, reason)
                 : run_adaptive_impl<float>(h, p, z, max_steps, log, nsteps, reason);
}

// ---------------------------------------------------------------------------
// fil_run_adaptive_noalloc: the device z loop with its buffers in `scratch`, a
// fixed array of device memory that every GPU the library runs on holds from
// the library's load to the end of the process.  In host code `scratch` names
// a host-side stand-in; its device address on the current device comes from
// hipGetSymbolAddress, once per device, and only that address is used.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[FIL_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on `device` (which must be current).
int scratch_address(int device, unsigned char** out) {
  if (device < 0 || device >= kMaxDevices) return fail(FIL_EINVAL, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(FIL_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[device] = p;
  }
  *out = static_cast<unsigned char*>(g_scratch_dev[device]);
  return 0;
}

size_t round256(size_t n) { return (n + 255) & ~(size_t)255; }

int fil_run_adaptive_noalloc(fil_handle* h, const fil_adapt_params* p, double* z, int max_steps, double* log,
                             int* nsteps, int* reason) {
  if (!valid_handle(h) || !p || !z || !log || !nsteps || !reason)
    return fail(FIL_EINVAL, "null argument (fil_run_adaptive_noalloc needs the log array)");
  if (!h->tables || !h->field) return fail(FIL_ESTATE, "tables and field must be set before fil_run_adaptive_noalloc");
  if (h->B != 1) return fail(FIL_EINVAL, "fil_run_adaptive_noalloc takes single-pulse handles");
  if (max_steps < 0 || p->chunk < 0) return fail(FIL_EINVAL, "max_steps and chunk must be >= 0");
  if (!std::isfinite(p->dz0) || !std::isfinite(*z)) return fail(FIL_EINVAL, "dz0 and z must be finite");
  // scratch layout: loop state | log | state copy | rho_end copy | rho_max copy
  const size_t nb = h->ns() * h->csz(), nr = (size_t)h->rows() * 8;
  const size_t o_log = round256(sizeof(AdaptState));
  const size_t o_S = o_log + round256(((size_t)max_steps + 1) * 4 * 8);
  const size_t o_re = o_S + round256(nb);
  const size_t o_rm = o_re + round256(nr);
  const size_t need = o_rm + round256(nr);
  if (need > (size_t)FIL_NOALLOC_SCRATCH_BYTES)
    return fail(FIL_EINVAL, "fil_run_adaptive_noalloc needs " + std::to_string(need) + " bytes of scratch for this " +
                                "grid and max_steps; its maximum is " + std::to_string((size_t)FIL_NOALLOC_SCRATCH_BYTES));
  HIPCHK(hipSetDevice(h->d.device));
  unsigned char* base = nullptr;
  if (int rc = scratch_address(h->d.device, &base)) return rc;
  const AdaptBuffers b{reinterpret_cast<AdaptState*>(base), reinterpret_cast<double*>(base + o_log), base + o_S,
                       reinterpret_cast<double*>(base + o_re), reinterpret_cast<double*>(base + o_rm)};
  // per-phase timing creates its events on demand: off for this call
  const bool timing = h->timing;
  h->timing = false;
  const int rc = h->fp64 ? run_adaptive_loop<double>(h, b, p, z, max_steps, log, nsteps, reason)
                         : run_adaptive_loop<float>(h, b, p, z, max_steps, log, nsteps, reason);
  h->timing = timing;
  return rc;
}

int fil_group_create(fil_handle* const* hs, int n, fil_group** out) {
  if (!hs || !out || n < 1) return fail(FIL_EINVAL, "null argument or empty group");
  *out = nullptr;
  if (n > FIL_GROUP_MAX) return fail(FIL_EINVAL, "too many handles in a group");
  for (int k = 0; k < n; k++) {
    if (!valid_handle(hs[k])) return fail(FIL_EINVAL, "null handle in group");
    if (hs[k]->d.device != hs[0]->d.device) return fail(FIL_EINVAL, "group handles on different devices");
    for (int q = 0; q < k; q++)
      if (hs[q] == hs[k]) return fail(FIL_EINVAL, "a handle appears twice in the group");
  }
  HIPCHK(hipSetDevice(hs[0]->d.device));
  auto* g = new fil_group();
  g->hs.assign(hs, hs + n);
  g->device = hs[0]->d.device;
  bool ok = hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&g->fork, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&g->post, hipEventDisableTiming) == hipSuccess;
  g->join.assign(n, nullptr);
  g->pre.assign(n, nullptr);
  for (int k = 0; ok && k < n; k++)
    ok = hipEventCreateWithFlags(&g->join[k], hipEventDisableTiming) == hipSuccess &&
         hipEventCreateWithFlags(&g->pre[k], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    fil_group_destroy(g);
    return fail(FIL_EDEVICE, "stream or event creation failed");
  }
  {
    std::lock_guard<std::mutex> lock(g_groups_mu);
    g_groups.push_back(g);
  }
  *out = g;
  return FIL_OK;
}

namespace {

// Records one zstep of every handle of the group as a graph: a fork event on
// the group's stream, each handle's kernels on its own stream (parallel
// branches), joined back.  The handles' host flags are restored afterwards:
// capturing runs nothing.
int group_capture(fil_group* g, double dz, hipGraphExec_t* exec) {
  const int n = (int)g->hs.size();
  std::vector<char> pf(n), pv(n), tm(n);
  for (int k = 0; k < n; k++) {
    pf[k] = g->hs[k]->pending_filter;
    pv[k] = g->hs[k]->plasma_valid;
    tm[k] = g->hs[k]->timing;
    g->hs[k]->timing = false;
  }
  int rc = 0;
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamBeginCapture(g->st, hipStreamCaptureModeRelaxed);
  if (e != hipSuccess) rc = fail(FIL_EDEVICE, hipGetErrorString(e));
  if (!rc && hipEventRecord(g->fork, g->st) != hipSuccess) rc = fail(FIL_EDEVICE, "event record in capture failed");
  for (int k = 0; !rc && k < n; k++) {
    fil_handle* h = g->hs[k];
    if (hipStreamWaitEvent(h->st, g->fork, 0) != hipSuccess) { rc = fail(FIL_EDEVICE, "stream join failed"); break; }
    rc = h->fp64 ? zstep_impl<double>(h, dz) : zstep_impl<float>(h, dz);
    if (!rc && (hipEventRecord(g->join[k], h->st) != hipSuccess || hipStreamWaitEvent(g->st, g->join[k], 0) != hipSuccess))
      rc = fail(FIL_EDEVICE, "branch join failed");
  }
  if (e == hipSuccess) {
    hipError_t e2 = hipStreamEndCapture(g->st, &graph);
    if (!rc && e2 != hipSuccess) rc = fail(FIL_EDEVICE, hipGetErrorString(e2));
  }
  for (int k = 0; k < n; k++) {
    g->hs[k]->pending_filter = pf[k];
    g->hs[k]->plasma_valid = pv[k];
    g->hs[k]->timing = tm[k];
  }
  if (!rc) {
    hipError_t e3 = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    if (e3 != hipSuccess) rc = fail(FIL_EDEVICE, hipGetErrorString(e3));
  }
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipGetLastError();
  return rc;
}

}  // namespace

int fil_group_zstep(fil_group* g, double z, double dz) {
  (void)z;
  if (!g) return fail(FIL_EINVAL, "null group");
  if (g->orphaned) return fail(FIL_ESTATE, "a handle of the group was destroyed: destroy the group");
  if (!std::isfinite(dz)) return fail(FIL_EINVAL, "dz must be finite");
  const int n = (int)g->hs.size();
  for (auto* h : g->hs)
    if (!h->tables || !h->field) return fail(FIL_ESTATE, "tables and field must be set before fil_group_zstep");
  HIPCHK(hipSetDevice(g->device));
  // pending corrector outputs are added outside the graph
  for (auto* h : g->hs)
    if (int rc = corr_materialize(h)) return rc;
  // the propagators for this dz, outside the graph
  uint64_t pattern = 0;
  std::vector<int> tiling(n);
  for (int k = 0; k < n; k++) {
    fil_handle* h = g->hs[k];
    if (int rc = h->fp64 ? update_propagator<double>(h, dz) : update_propagator<float>(h, dz)) return rc;
    if (h->pending_filter) pattern |= 1ull << k;
    tiling[k] = h->tiling;
  }
  hipGraphExec_t exec = nullptr;
  for (auto& c : g->cache)
    if (c.dz == dz && c.pattern == pattern && c.tiling == tiling) exec = c.exec;
  if (!exec) {
    if (g->cache.size() >= 4) {
      HIPCHK(hipStreamSynchronize(g->st));
      (void)hipGraphExecDestroy(g->cache.front().exec);
      g->cache.erase(g->cache.begin());
    }
    if (int rc = group_capture(g, dz, &exec)) return rc;
    g->cache.push_back({dz, pattern, tiling, exec});
    g->captures++;
  }
  // the graph runs after each handle's earlier work and before its later work
  for (int k = 0; k < n; k++) {
    HIPCHK(hipEventRecord(g->pre[k], g->hs[k]->st));
    HIPCHK(hipStreamWaitEvent(g->st, g->pre[k], 0));
  }
  HIPCHK(hipGraphLaunch(exec, g->st));
  HIPCHK(hipEventRecord(g->post, g->st));
  for (auto* h : g->hs) {
    HIPCHK(hipStreamWaitEvent(h->st, g->post, 0));
    // the state after a zstep (as zstep_impl leaves it)
    h->pending_filter = true;
    h->plasma_valid = h->d.plasma != 0;
    h->acache.valid = false;
  }
  g->replays++;
  return FIL_OK;
}

int fil_group_stats(fil_group* g, int64_t* captures, int64_t* replays) {
  if (!g) return fail(FIL_EINVAL, "null group");
  if (captures) *captures = g->captures;
  if (replays) *replays = g->replays;
  return FIL_OK;
}

int fil_group_destroy(fil_group* g) {
  if (!g) return FIL_OK;
  (void)hipSetDevice(g->device);
  if (g->st) (void)hipStreamSynchronize(g->st);
  for (auto& c : g->cache) (void)hipGraphExecDestroy(c.exec);
  for (auto e : g->join) if (e) (void)hipEventDestroy(e);
  for (auto e : g->pre) if (e) (void)hipEventDestroy(e);
  if (g->fork) (void)hipEventDestroy(g->fork);
  if (g->post) (void)hipEventDestroy(g->post);
  if (g->st) (void)hipStreamDestroy(g->st);
  {
    std::lock_guard<std::mutex> lock(g_groups_mu);
    g_groups.erase(std::remove(g_groups.begin(), g_groups.end(), g), g_groups.end());
  }
  delete g;
  return FIL_OK;
}

int fil_analyze_member(fil_handle* h, int member, double z, fil_diag* d, double* Fr, double* Ft, double* rho_end,
                       double* S) {
  if (!valid_handle(h)) return fail(FIL_EINVAL, "null handle");
  if (h->xyt) return fail(FIL_EINVAL, "XYT handles take fil_analyze_xyt");
  if (member < 0 || member >= h->B) return fail(FIL_EINVAL, "member out of range");
  if (!h->tables || !h->field) return fail(FIL_ESTATE, "tables and field must be set before fil_analyze");
  HIPCHK(hipSetDevice(h->d.device));
  return h->fp64 ? analyze_impl<double>(h, member, z, d, Fr, Ft, rho_end, S)
                 : analyze_impl<float>(h, member, z, d, Fr, Ft, rho_end, S);
}

int fil_analyze(fil_handle* h, double z, fil_diag* d, double* Fr, double* Ft, double* rho_end, double* S) {
  if (!valid_handle(h)) return fail(FIL_EINVAL, "null handle");
  if (h->B > 1) return fail(FIL_EINVAL, "batched handles take fil_analyze_member");
  return fil_analyze_member(h, 0, z, d, Fr, Ft, rho_end, S);
}

int fil_analyze_xyt(fil_handle* h, double z, fil_diag_xyt* d, double* Fxy, double* Ft, double* rho_end, double* S) {
  if (!valid_handle(h)) return fail(FIL_EINVAL, "null handle");
  if (!h->xyt) return fail(FIL_EINVAL, "fil_analyze_xyt needs an XYT handle");
  if (!h->tables || !h->field) return fail(FIL_ESTATE, "tables and field must be set before fil_analyze_xyt");
  HIPCHK(hipSetDevice(h->d.device));
  return h->fp64 ? analyze_xyt_impl<double>(h, z, d, Fxy, Ft, rho_end, S)
                 : analyze_xyt_impl<float>(h, z, d, Fxy, Ft, rho_end, S);
}

int fil_download_plasma(fil_handle* h, double* rho, double* kdrho, int64_t ld) {
  if (!valid_handle(h)) return fail(FIL_EINVAL, "null handle");
  if (!h->d.plasma || !h->d.keep_plasma) return fail(FIL_ESTATE, "create the handle with plasma and keep_plasma");
  if (h->B > 1) return fail(FIL_EINVAL, "fil_download_plasma takes single-pulse handles");
  if (ld < h->Nr) return fail(FIL_EINVAL, "ld < Nr");
  HIPCHK(hipSetDevice(h->d.device));
  if (int rc = corr_materialize(h)) return rc;
  if (!h->plasma_valid) {
    int rc = h->fp64 ? plasma_only<double>(h) : plasma_only<float>(h);
    if (rc) return rc;
  }
  const int Nr = h->Nr, Nt = h->Nt;
  std::vector<unsigned char> a((size_t)Nr * Nt * h->rsz()), b((size_t)Nr * Nt * h->rsz());
  HIPCHK(hipMemcpyAsync(a.data(), h->rho_full.p, a.size(), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipMemcpyAsync(b.data(), h->kd_full.p, b.size(), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  for (int i = 0; i < Nr; i++)
    for (int j = 0; j < Nt; j++) {
      size_t s = (size_t)i * Nt + j, o = (size_t)i + (size_t)ld * j;
      double rv = h->fp64 ? reinterpret_cast<double*>(a.data())[s] : reinterpret_cast<float*>(a.data())[s];
      double kv = h->fp64 ? reinterpret_cast<double*>(b.data())[s] : reinterpret_cast<float*>(b.data())[s];
      if (rho) rho[o] = rv;
      if (kdrho) kdrho[o] = kv;
    }
  return FIL_OK;
}

int fil_state_save(fil_handle* h, int slot) {
  if (!valid_handle(h)) return fail(FIL_EINVAL, "null handle");
  if (slot < 0 || slot >= FIL_NSLOTS) return fail(FIL_EINVAL, "slot out of range");
  if (!h->field) return fail(FIL_ESTATE, "no field uploaded");
  HIPCHK(hipSetDevice(h->d.device));
  if (int rc = corr_materialize(h)) return rc;
  auto& s = h->slot[slot];
  const size_t nb = h->ns() * h->csz(), nr = (size_t)h->rows() * 8;
  if (!s.S.p) {
    if (s.S.alloc(nb) != hipSuccess || s.rho_end.alloc(nr) != hipSuccess || s.rho_max.alloc(nr) != hipSuccess)
      return fail(FIL_ENOMEM, "device allocation failed");
  }
  HIPCHK(hipMemcpyAsync(s.S.p, h->S.p, nb, hipMemcpyDeviceToDevice, h->st));
  HIPCHK(hipMemcpyAsync(s.rho_end.p, h->rho_end.p, nr, hipMemcpyDeviceToDevice, h->st));
  HIPCHK(hipMemcpyAsync(s.rho_max.p, h->rho_max.p, nr, hipMemcpyDeviceToDevice, h->st));
  s.used = true;
  s.pending_filter = h->pending_filter;
  s.plasma_valid = h->plasma_valid;
  return FIL_OK;
}

int fil_state_copy(fil_handle* dst, const fil_handle* src) {
  if (!valid_handle(dst) || !src) return fail(FIL_EINVAL, "null handle");
  if (dst == src) return FIL_OK;
  if (dst->d.device != src->d.device) return fail(FIL_EINVAL, "handles on different devices");
  if (dst->Nr != src->Nr || dst->Nt != src->Nt || dst->B != src->B) return fail(FIL_EINVAL, "handles of different grids");
  if (!src->field || !dst->tables) return fail(FIL_ESTATE, "source needs a field and destination its tables");
  HIPCHK(hipSetDevice(dst->d.device));
  // the source's pending corrector output belongs to its state (adding it
  // leaves that state unchanged); the destination's is replaced
  if (int rc = corr_materialize(const_cast<fil_handle*>(src))) return rc;
  dst->corr_pending = false;
  // the source's pending work completes before the destination's stream reads it
  HIPCHK(hipStreamSynchronize(src->st));
  const size_t n = dst->ns();
  const dim3 grid((unsigned)((n + 255) / 256));
  if (dst->fp64 == src->fp64) {
    HIPCHK(hipMemcpyAsync(dst->S.p, src->S.p, n * dst->csz(), hipMemcpyDeviceToDevice, dst->st));
  } else if (dst->fp64) {
    hipLaunchKernelGGL((convert_state_kernel<double, float>), grid, dim3(256), 0, dst->st, src->S.as<cx<float>>(),
                       dst->S.as<cx<double>>(), n);
    HIPCHK(hipGetLastError());
  } else {
    hipLaunchKernelGGL((convert_state_kernel<float, double>), grid, dim3(256), 0, dst->st, src->S.as<cx<double>>(),
                       dst->S.as<cx<float>>(), n);
    HIPCHK(hipGetLastError());
  }
  if (src->d.plasma && dst->d.plasma) {
    HIPCHK(hipMemcpyAsync(dst->rho_end.p, src->rho_end.p, (size_t)dst->rows() * 8, hipMemcpyDeviceToDevice, dst->st));
    HIPCHK(hipMemcpyAsync(dst->rho_max.p, src->rho_max.p, (size_t)dst->rows() * 8, hipMemcpyDeviceToDevice, dst->st));
  }
  dst->field = true;
  dst->acache.valid = false;
  dst->pending_filter = src->pending_filter;
  dst->plasma_valid = src->plasma_valid && src->d.plasma && dst->d.plasma;
  return FIL_OK;
}

int fil_state_restore(fil_handle* h, int slot) {
  if (!valid_handle(h)) return fail(FIL_EINVAL, "null handle");
  if (slot < 0 || slot >= FIL_NSLOTS) return fail(FIL_EINVAL, "slot out of range");
  auto& s = h->slot[slot];
  if (!s.used) return fail(FIL_ESTATE, "slot is empty");
  HIPCHK(hipSetDevice(h->d.device));
  const size_t nb = h->ns() * h->csz(), nr = (size_t)h->rows() * 8;
  HIPCHK(hipMemcpyAsync(h->S.p, s.S.p, nb, hipMemcpyDeviceToDevice, h->st));
  HIPCHK(hipMemcpyAsync(h->rho_end.p, s.rho_end.p, nr, hipMemcpyDeviceToDevice, h->st));
  HIPCHK(hipMemcpyAsync(h->rho_max.p, s.rho_max.p, nr, hipMemcpyDeviceToDevice, h->st));
  h->pending_filter = s.pending_filter;
  h->plasma_valid = s.plasma_valid;
  h->corr_pending = false;
  h->acache.valid = false;
  return FIL_OK;
}

int fil_fluence(fil_handle* h, int slot, double* Fr) {
  if (!valid_handle(h) || !Fr) return fail(FIL_EINVAL, "null argument");
  if (!h->field) return fail(FIL_ESTATE, "no field uploaded");
  if (slot >= FIL_NSLOTS || (slot >= 0 && !h->slot[slot].used)) return fail(FIL_EINVAL, "bad slot");
  HIPCHK(hipSetDevice(h->d.device));
  if (slot < 0)
    if (int rc = corr_materialize(h)) return rc;
  const void* S = slot < 0 ? h->S.p : h->slot[slot].S.p;
  const double scale = (double)h->Nt * h->dt;   // dt sum_t |E|^2 = dt Nt sum_w |S_w|^2
  const int n = h->rows(), B = h->B, Nr = h->Nr;
  if (h->fp64)
    HIPCHK(launch_rowsum_abs2<double>(static_cast<const cx<double>*>(S), n, h->Nh, scale, h->aFr.as<double>(), h->st));
  else
    HIPCHK(launch_rowsum_abs2<float>(static_cast<const cx<float>*>(S), n, h->Nh, scale, h->aFr.as<double>(), h->st));
  if (B == 1) {
    HIPCHK(hipMemcpyAsync(Fr, h->aFr.p, (size_t)n * 8, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    return FIL_OK;
  }
  // device rows are (r, member) with r slowest; the caller gets Fr[member][r]
  std::vector<double> t(n);
  HIPCHK(hipMemcpyAsync(t.data(), h->aFr.p, (size_t)n * 8, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  for (int r = 0; r < Nr; r++)
    for (int b = 0; b < B; b++) Fr[(size_t)b * Nr + r] = t[(size_t)r * B + b];
  return FIL_OK;
}

int fil_synchronize(fil_handle* h) {
  if (!valid_handle(h)) return fail(FIL_EINVAL, "null handle");
  HIPCHK(hipSetDevice(h->d.device));
  if (int rc = corr_materialize(h)) return rc;
  HIPCHK(hipStreamSynchronize(h->st));
  return FIL_OK;
}

int fil_destroy(fil_handle* h) {
  if (!h) return FIL_OK;
  {
    // a group that still holds the handle can no longer be stepped
    std::lock_guard<std::mutex> lock(g_groups_mu);
    for (auto* g : g_groups)
      if (std::find(g->hs.begin(), g->hs.end(), h) != g->hs.end()) g->orphaned = true;
  }
  (void)hipSetDevice(h->d.device);
  (void)hipStreamSynchronize(h->st);
  for (auto e : h->ev) (void)hipEventDestroy(e);
  if (h->st) (void)hipStreamDestroy(h->st);
  delete h;
  return FIL_OK;
}

int fil_set_tiling(fil_handle* h, int tiling) {
  if (!valid_handle(h)) return fail(FIL_EINVAL, "null handle");
  if (tiling != FIL_TILING_RULE && tiling != FIL_TILING_128 && tiling != FIL_TILING_64)
    return fail(FIL_EINVAL, "tiling must be FIL_TILING_RULE, FIL_TILING_128 or FIL_TILING_64");
  if (tiling == FIL_TILING_128 && h->Nrp % 128) return fail(FIL_EINVAL, "128-row tiles need Nr (padded to 64) % 128 == 0");
  h->tiling = tiling;
  return FIL_OK;
}

int fil_get_tiling(fil_handle* h, int* tiling) {
  if (!valid_handle(h) || !tiling) return fail(FIL_EINVAL, "null argument");
  *tiling = h->tiling;
  return FIL_OK;
}

int fil_timing_enable(fil_handle* h, int enable) {
  if (!valid_handle(h)) return fail(FIL_EINVAL, "null handle");
  if (int rc = flush_timing(h)) return rc;
  if (enable < 0) return fail(FIL_EINVAL, "timing period must be >= 0");
  h->timing = enable != 0;
  h->timing_period = enable > 1 ? enable : 1;
  h->timing_tick = 0;
  h->ms[0] = h->ms[1] = h->ms[2] = 0;
  h->count = 0;
  return FIL_OK;
}

int fil_timing_read(fil_handle* h, double* ms_row, double* ms_dht1, double* ms_dht2, int64_t* count) {
  if (!valid_handle(h)) return fail(FIL_EINVAL, "null handle");
  if (int rc = flush_timing(h)) return rc;
  if (ms_row) *ms_row = h->ms[0];
  if (ms_dht1) *ms_dht1 = h->ms[1];
  if (ms_dht2) *ms_dht2 = h->ms[2];
  if (count) *count = h->count;
  return FIL_OK;
}

int fil_corrector_init(fil_handle* h, const fil_corr_desc* cd, uint64_t seed) {
  if (!valid_handle(h) || !cd) return fail(FIL_EINVAL, "null argument");
  if (h->fp64) return fail(FIL_EINVAL, "the corrector needs an FP32 handle");
  if (h->xyt || h->tgeo) return fail(FIL_EINVAL, "the corrector is defined for the RT geometry");
  if (h->B > 1) return fail(FIL_EINVAL, "the corrector takes single-pulse handles");
  if (!h->tables) return fail(FIL_ESTATE, "fil_set_tables first");
  if (h->corr_pending) {
    // the output of the previous network is added before its buffers are replaced
    HIPCHK(hipSetDevice(h->d.device));
    if (int rc = corr_materialize(h)) return rc;
  }
  if (cd->pool_r < 1 || cd->pool_t < 1) return fail(FIL_EINVAL, "pool sizes must be >= 1");
  if (cd->nr < 1 || cd->nr > h->Nr || cd->nr % cd->pool_r) return fail(FIL_EINVAL, "nr must be in [1, Nr] and a multiple of pool_r");
  if (cd->t0 < 0 || cd->nt < 1 || cd->t0 + cd->nt > h->Nt || cd->nt % cd->pool_t)
    return fail(FIL_EINVAL, "t window must lie in [0, Nt) with nt a multiple of pool_t");
  if (cd->capacity < 0 || cd->batch_max < 1) return fail(FIL_EINVAL, "capacity >= 0 and batch_max >= 1 required");
  if (!(cd->a_scale > 0) || !(cd->rho_scale > 0) || !(cd->dt > 0) || !std::isfinite(cd->w0) || !std::isfinite(cd->tmin))
    return fail(FIL_EINVAL, "a_scale, rho_scale and dt must be positive, w0 and tmin finite");
  const int H = cd->nr / cd->pool_r, W = cd->nt / cd->pool_t;
  if ((int64_t)H * W > (1 << 22)) return fail(FIL_EINVAL, "pooled grid too large");
  HIPCHK(hipSetDevice(h->d.device));
  HIPCHK(hipStreamSynchronize(h->st));
  auto& c = h->corr;
  c = fil_handle::Corr();
  c.g.nr = cd->nr; c.g.t0 = cd->t0; c.g.nt = cd->nt; c.g.pr = cd->pool_r; c.g.pt = cd->pool_t;
  c.g.H = H; c.g.W = W;
  c.g.dt = (float)cd->dt;
  c.g.inv_as = (float)(1.0 / cd->a_scale); c.g.inv_rs = (float)(1.0 / cd->rho_scale);
  c.cap = cd->capacity;
  c.bmax = cd->batch_max;
  const size_t HW = (size_t)H * W, NP = FIL_CORR_NPARAM, B = c.bmax;
  bool ok = c.w.alloc(NP * 4) == hipSuccess && c.wbf.alloc(NP * 2) == hipSuccess && c.m.alloc(NP * 4) == hipSuccess &&
            c.wfr.alloc((size_t)corr_frag_size() * 2) == hipSuccess &&
            c.v.alloc(NP * 4) == hipSuccess && c.grad.alloc(NP * 4) == hipSuccess &&
            c.X.alloc(std::max<size_t>(1, (size_t)c.cap) * 4 * HW * 4) == hipSuccess &&
            c.Y.alloc(std::max<size_t>(1, (size_t)c.cap) * 2 * HW * 4) == hipSuccess &&
            c.Frow.alloc((size_t)cd->nr * 4 * W * 4) == hipSuccess && c.feat.alloc(4 * HW * 4) == hipSuccess &&
            c.out.alloc(B * 2 * HW * 4) == hipSuccess && c.gout.alloc(B * 2 * HW * 4) == hipSuccess &&
            c.gA.alloc(B * CORR_CH * HW * 4) == hipSuccess && c.gB.alloc(B * CORR_CH * HW * 4) == hipSuccess &&
            c.idx.alloc(B * 4) == hipSuccess && c.part.alloc((B * 2 * HW / 256 + 1024) * 8) == hipSuccess;
  for (int l = 0; l < 3 && ok; l++) ok = c.act[l].alloc(B * CORR_CH * HW * 4) == hipSuccess;
  if (ok) ok = c.carrier.alloc((size_t)h->Nt * 8) == hipSuccess &&
              c.wpart.alloc((size_t)CORR_WCHUNK_MAX * CORR_CH * (CORR_CH + 1) * 9 * 4) == hipSuccess;
  if (!ok) { c = fil_handle::Corr(); return fail(FIL_ENOMEM, "device allocation failed"); }
  {
    std::vector<float> car(2 * (size_t)h->Nt);
    for (int j = 0; j < h->Nt; j++) {
      const double ph = cd->w0 * (cd->tmin + j * cd->dt);
      car[2 * j] = (float)std::cos(ph);
      car[2 * j + 1] = (float)std::sin(ph);
    }
    HIPCHK(hipMemcpyAsync(c.carrier.p, car.data(), car.size() * 4, hipMemcpyHostToDevice, h->st));
    c.g.carrier = c.carrier.as<cx<float>>();
  }
  // He-normal weights for the ReLU layers, zero last layer and biases
  std::vector<float> w(NP, 0.0f);
  uint64_t s = seed;
  for (int l = 0; l < CORR_LAYERS - 1; l++) {
    const double sd = std::sqrt(2.0 / (corr_cin(l) * 9.0));
    for (int q = 0; q < corr_cout(l) * corr_cin(l) * 9; q++) w[corr_woff(l) + q] = (float)(sd * normal01(s));
  }
  HIPCHK(hipMemcpyAsync(c.w.p, w.data(), NP * 4, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemsetAsync(c.m.p, 0, NP * 4, h->st));
  HIPCHK(hipMemsetAsync(c.v.p, 0, NP * 4, h->st));
  HIPCHK(launch_round_bf16(c.w.as<float>(), c.wbf.as<uint16_t>(), (int)NP, h->st));
  HIPCHK(launch_conv_weight_layout(c.wbf.as<uint16_t>(), c.wfr.as<uint16_t>(), h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  c.on = true;
  return FIL_OK;
}

int fil_corrector_params(fil_handle* h, float* w, int set) {
  if (int rc = corr_ready(h)) return rc;
  if (!w) return fail(FIL_EINVAL, "null argument");
  HIPCHK(hipSetDevice(h->d.device));
  auto& c = h->corr;
  const size_t NP = FIL_CORR_NPARAM;
  if (set) {
    for (size_t q = 0; q < NP; q++)
      if (!std::isfinite(w[q])) return fail(FIL_EINVAL, "parameters must be finite");
    HIPCHK(hipMemcpyAsync(c.w.p, w, NP * 4, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipMemsetAsync(c.m.p, 0, NP * 4, h->st));
    HIPCHK(hipMemsetAsync(c.v.p, 0, NP * 4, h->st));
    HIPCHK(launch_round_bf16(c.w.as<float>(), c.wbf.as<uint16_t>(), (int)NP, h->st));
    HIPCHK(launch_conv_weight_layout(c.wbf.as<uint16_t>(), c.wfr.as<uint16_t>(), h->st));
    c.adam_t = 0;
  } else {
    HIPCHK(hipMemcpyAsync(w, c.w.p, NP * 4, hipMemcpyDeviceToHost, h->st));
  }
  HIPCHK(hipStreamSynchronize(h->st));
  return FIL_OK;
}

int fil_corrector_scale(fil_handle* h, double* y_scale, int set) {
  if (int rc = corr_ready(h)) return rc;
  if (!y_scale) return fail(FIL_EINVAL, "null argument");
  if (set) {
    if (!(*y_scale >= 0) || !std::isfinite(*y_scale)) return fail(FIL_EINVAL, "y_scale must be finite and >= 0");
    h->corr.y_scale = *y_scale;
  } else {
    *y_scale = h->corr.y_scale;
  }
  return FIL_OK;
}

int fil_corrector_features(fil_handle* h, float* X) {
  if (int rc = corr_ready(h)) return rc;
  if (!h->field) return fail(FIL_ESTATE, "no field uploaded");
  HIPCHK(hipSetDevice(h->d.device));
  auto& c = h->corr;
  HIPCHK(launch_corr_features(h->S.as<cx<float>>(), h->tw.as<cx<float>>(), h->pp32, h->d.plasma, c.g, h->Nt,
                              c.Frow.as<float>(), c.feat.as<float>(), h->st));
  if (X) HIPCHK(hipMemcpyAsync(X, c.feat.p, (size_t)4 * c.g.H * c.g.W * 4, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return FIL_OK;
}

int fil_corrector_store_pair(fil_handle* h, const float* X, const float* T) {
  if (int rc = corr_ready(h)) return rc;
  if (!X || !T) return fail(FIL_EINVAL, "null argument");
  auto& c = h->corr;
  if (c.cap < 1) return fail(FIL_EINVAL, "the corrector was created with capacity 0");
  HIPCHK(hipSetDevice(h->d.device));
  const size_t HW = (size_t)c.g.H * c.g.W;
  HIPCHK(hipMemcpyAsync(c.X.as<float>() + (size_t)c.next * 4 * HW, X, 4 * HW * 4, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync(c.Y.as<float>() + (size_t)c.next * 2 * HW, T, 2 * HW * 4, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  c.next = (c.next + 1) % c.cap;
  c.count = std::min(c.count + 1, c.cap);
  return FIL_OK;
}

int fil_corrector_train(fil_handle* h, int iters, int batch, double lr, uint64_t seed, double* loss) {
  if (int rc = corr_ready(h)) return rc;
  auto& c = h->corr;
  if (iters < 1 || batch < 1 || batch > c.bmax) return fail(FIL_EINVAL, "iters >= 1 and 1 <= batch <= batch_max required");
  if (batch > c.count) return fail(FIL_ESTATE, "fewer stored pairs than the batch size");
  if (!(lr > 0) || !std::isfinite(lr)) return fail(FIL_EINVAL, "lr must be positive");
  HIPCHK(hipSetDevice(h->d.device));
  const int HW = c.g.H * c.g.W;
  if (c.y_scale == 0) {
    int nb = 0;
    HIPCHK(launch_sum_squares(c.Y.as<float>(), c.count * 2 * HW, c.part.as<double>(), &nb, h->st));
    std::vector<double> part(nb);
    HIPCHK(hipMemcpyAsync(part.data(), c.part.p, nb * 8, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    double s = 0;
    for (double v : part) s += v;
    const double rms = std::sqrt(s / ((double)c.count * 2 * HW));
    if (!std::isfinite(rms)) return fail(FIL_ESTATE, "stored targets hold non-finite values");
    if (!(rms > 0)) return fail(FIL_ESTATE, "stored targets are all zero");
    c.y_scale = rms;
  }
  // batch indices of every iteration, drawn without replacement within a batch
  std::vector<int> ids((size_t)iters * batch);
  std::vector<int> perm(c.count);
  uint64_t s = seed;
  for (int it = 0; it < iters; it++) {
    for (int q = 0; q < c.count; q++) perm[q] = q;
    for (int q = 0; q < batch; q++) {
      const int r = q + (int)(splitmix64(s) % (uint64_t)(c.count - q));
      std::swap(perm[q], perm[r]);
      ids[(size_t)it * batch + q] = perm[q];
    }
  }
  DevBuf dids;
  HIPCHK(dids.alloc(ids.size() * 4));
  HIPCHK(hipMemcpyAsync(dids.p, ids.data(), ids.size() * 4, hipMemcpyHostToDevice, h->st));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  int nblk = 0;
  for (int it = 0; it < iters; it++) {
    const int* idx = dids.as<int>() + (size_t)it * batch;
    if (int rc = corr_forward(h, c.X.as<float>(), idx, batch, false, true)) return rc;
    if (int rc = corr_loss(h, idx, batch, &nblk)) return rc;
    if (int rc = corr_backward(h, c.X.as<float>(), idx, batch)) return rc;
    c.adam_t++;
    const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)c.adam_t)));
    const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)c.adam_t)));
    HIPCHK(launch_adam(c.w.as<float>(), c.grad.as<float>(), c.m.as<float>(), c.v.as<float>(), c.wbf.as<uint16_t>(),
                       FIL_CORR_NPARAM, (float)lr, b1, b2, eps, c1, c2, h->st));
    HIPCHK(launch_conv_weight_layout(c.wbf.as<uint16_t>(), c.wfr.as<uint16_t>(), h->st));
  }
  if (int rc = corr_read_loss(h, nblk, batch, loss)) return rc;
  return FIL_OK;
}

int fil_corrector_eval(fil_handle* h, int first, int count, double* loss, float* grad) {
  if (int rc = corr_ready(h)) return rc;
  auto& c = h->corr;
  if (first < 0 || count < 1 || count > c.bmax || first + count > c.count)
    return fail(FIL_EINVAL, "pairs [first, first+count) must be stored and count <= batch_max");
  if (!(c.y_scale > 0)) return fail(FIL_ESTATE, "y_scale is not set (train or set it first)");
  HIPCHK(hipSetDevice(h->d.device));
  std::vector<int> ids(count);
  for (int q = 0; q < count; q++) ids[q] = first + q;
  HIPCHK(hipMemcpyAsync(c.idx.p, ids.data(), count * 4, hipMemcpyHostToDevice, h->st));
  int nblk = 0;
  if (int rc = corr_forward(h, c.X.as<float>(), c.idx.as<int>(), count, false, grad != nullptr)) return rc;
  if (int rc = corr_loss(h, c.idx.as<int>(), count, &nblk)) return rc;
  if (grad) {
    if (int rc = corr_backward(h, c.X.as<float>(), c.idx.as<int>(), count)) return rc;
    HIPCHK(hipMemcpyAsync(grad, c.grad.p, (size_t)FIL_CORR_NPARAM * 4, hipMemcpyDeviceToHost, h->st));
  }
  return corr_read_loss(h, nblk, count, loss);
}

int fil_corrector_forward(fil_handle* h, const float* X, float* Y) {
  if (int rc = corr_ready(h)) return rc;
  if (!X || !Y) return fail(FIL_EINVAL, "null argument");
  HIPCHK(hipSetDevice(h->d.device));
  auto& c = h->corr;
  const size_t HW = (size_t)c.g.H * c.g.W;
  HIPCHK(hipMemcpyAsync(c.feat.p, X, 4 * HW * 4, hipMemcpyHostToDevice, h->st));
  if (int rc = corr_forward(h, c.feat.as<float>(), nullptr, 1)) return rc;
  HIPCHK(hipMemcpyAsync(Y, c.out.p, 2 * HW * 4, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return FIL_OK;
}

int fil_corrector_apply(fil_handle* h) {
  if (int rc = corr_ready(h)) return rc;
  if (!h->field) return fail(FIL_ESTATE, "no field uploaded");
  auto& c = h->corr;
  if (!(c.y_scale > 0)) return fail(FIL_ESTATE, "the corrector is untrained (y_scale is 0)");
  HIPCHK(hipSetDevice(h->d.device));
  HIPCHK(launch_corr_features(h->S.as<cx<float>>(), h->tw.as<cx<float>>(), h->pp32, h->d.plasma, c.g, h->Nt,
                              c.Frow.as<float>(), nullptr, h->st));
  if (int rc = corr_forward(h, c.Frow.as<float>(), nullptr, 1, true)) return rc;
  // the output is added to the field by the next z-step's row kernel (before
  // its field filter), or by corr_materialize when anything else comes first
  h->corr_scale = (float)(c.y_scale / (double)c.g.inv_as);
  h->corr_pending = true;
  h->plasma_valid = false;
  h->acache.valid = false;
  return FIL_OK;
}

}  // extern "C"
