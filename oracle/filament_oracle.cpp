// filament_oracle.cpp -- CPU restatement of Filament.jl's per-z-step path.
//
// TEST INFRASTRUCTURE ONLY: the checker for libfilament_hip.so.  Only tests/,
// __graft_entry__.smoke() and bench.py's CPU-baseline leg load it; the product
// never does.
//
// The reference is Julia; its RT path runs only on CUDA in FP32.  This file
// restates the same algorithm in plain C++ (templated on the real type) so the
// FP64 instance is the oracle and the FP32 instance is the "port" CPU baseline.
// Citations are file:line in the Filament.jl tree (src/...).  Behaviour that
// lives in un-vendored packages (HankelTransforms, ODEIntegrators,
// TabulatedFunctions, AnalyticSignals) is restated from their published
// algorithms; each such place says which assumption it makes.

#include "filament_oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// CODATA2018 (src/modules/Constants.jl:9-14 via PhysicalConstants.CODATA2018)
constexpr double C0 = 299792458.0;
constexpr double EPS0 = 8.8541878128e-12;
constexpr double MU0 = 1.25663706212e-6;
constexpr double QE = 1.602176634e-19;
constexpr double ME = 9.1093837015e-31;
constexpr double HBAR = 1.054571817e-34;
constexpr double PI = 3.14159265358979323846;

thread_local std::string g_err;

using cd = std::complex<double>;

// ---------------------------------------------------------------------------
// Medium (src/modules/Media.jl:22-45,121-134,168-191)
// ---------------------------------------------------------------------------
struct Medium {
  int kind;
  double n_const, n2;
  // Peck & Reeder air (test/diffraction/RT/air.jl:4-17)
  double permittivity(double w) const {
    if (kind == ORC_MEDIUM_CONSTANT) return n_const * n_const;
    double sig = 0.0;
    if (w != 0.0) {
      double lam = 2.0 * PI * C0 / w;
      sig = 1.0 / lam * 1e-6;
    }
    double dum = 8060.51 + 2480990.0 / (132.247 - sig * sig) + 17455.7 / (39.32957 - sig * sig);
    double n = 1.0 + dum * 1e-8;
    return n * n;
  }
  double permeability(double) const { return 1.0; }
  // Media.jl:22-26: n = sqrt(eps(|w|) * mu(|w|) + 0im)
  cd refractive_index(double w) const {
    return std::sqrt(cd(permittivity(std::fabs(w)) * permeability(std::fabs(w)), 0.0));
  }
  cd beta(double w) const { return refractive_index(w) * w / C0; }   // Media.jl:29-32
  double k(double w) const { return beta(w).real(); }                // Media.jl:35-38
  // five-point stencil, Media.jl:168-191
  double derivative_k(double x, int n) const {
    double h = (x == 0.0) ? 0.01 : 0.001 * x;
    auto f = [&](double v) { return k(v); };
    if (n == 1) return (f(x - 2 * h) - 8 * f(x - h) + 8 * f(x + h) - f(x + 2 * h)) / (12 * h);
    if (n == 2) return (-f(x - 2 * h) + 16 * f(x - h) - 30 * f(x) + 16 * f(x + h) - f(x + 2 * h)) / (12 * h * h);
    if (n == 3) return (-f(x - 2 * h) + 2 * f(x - h) - 2 * f(x + h) + f(x + 2 * h)) / (2 * h * h * h);
    throw std::runtime_error("Wrong derivative order.");
  }
  double group_velocity(double w) const { return 1.0 / derivative_k(w, 1); }   // Media.jl:76-79
  double chi3(double w) const {                                               // Media.jl:121-125
    double nr = refractive_index(w).real();
    return 4.0 / 3.0 * nr * nr * EPS0 * C0 * n2;
  }
};

// ---------------------------------------------------------------------------
// Bessel J0 zeros: McMahon's asymptotic start + Newton (J0' = -J1).
// HankelTransforms.dhtcoord/dhtfreq (called at src/modules/Grids.jl:205-206)
// place r_n = j_n R / j_{N+1} and nu_n = j_n / (2 pi R).
// ---------------------------------------------------------------------------
void bessel_j0_zeros(int n, double* out) {
  for (int s = 1; s <= n; s++) {
    double b = (s - 0.25) * PI;
    double x = b + 1.0 / (8 * b) - 124.0 / (3.0 * std::pow(8 * b, 3)) + 120928.0 / (15.0 * std::pow(8 * b, 5));
    for (int it = 0; it < 8; it++) {
      double j0 = std::cyl_bessel_j(0.0, x), j1 = std::cyl_bessel_j(1.0, x);
      double dx = j0 / j1;
      x += dx;
      if (std::fabs(dx) < 1e-15 * x) break;
    }
    out[s - 1] = x;
  }
}

// ---------------------------------------------------------------------------
// Radix-2 FFT along one contiguous row.  FFTW conventions used by the
// reference through AbstractFFTs (src/modules/Models.jl:162-171):
//   P * E  : forward, unnormalised, exp(-2 pi i jk/N)   (frequency -> time)
//   P \ E  : inverse, normalised 1/N, exp(+2 pi i jk/N) (time -> frequency)
// ---------------------------------------------------------------------------
template <typename T>
struct FFTPlan {
  int n = 0, logn = 0;
  std::vector<std::complex<T>> tw;   // exp(-2 pi i k / n), k < n/2
  std::vector<int> rev;
  explicit FFTPlan(int n_) : n(n_) {
    if (n < 1 || (n & (n - 1))) throw std::runtime_error("FFT size must be a power of two");
    while ((1 << logn) < n) logn++;
    tw.resize(std::max(1, n / 2));
    for (int k = 0; k < n / 2; k++) {
      double a = -2.0 * PI * k / n;
      tw[k] = std::complex<T>((T)std::cos(a), (T)std::sin(a));
    }
    rev.resize(n);
    for (int i = 0; i < n; i++) {
      int r = 0;
      for (int b = 0; b < logn; b++) if (i & (1 << b)) r |= 1 << (logn - 1 - b);
      rev[i] = r;
    }
  }
  // sign = -1: P*E (forward); sign = +1: P\E (inverse, normalised)
  void run(std::complex<T>* x, int sign) const {
    for (int i = 0; i < n; i++) if (i < rev[i]) std::swap(x[i], x[rev[i]]);
    for (int len = 2; len <= n; len <<= 1) {
      int half = len >> 1, step = n / len;
      for (int i = 0; i < n; i += len) {
        for (int k = 0; k < half; k++) {
          std::complex<T> w = tw[k * step];
          if (sign > 0) w = std::conj(w);
          std::complex<T> u = x[i + k], v = x[i + k + half] * w;
          x[i + k] = u + v;
          x[i + k + half] = u - v;
        }
      }
    }
    if (sign > 0) {
      T s = (T)(1.0 / n);
      for (int i = 0; i < n; i++) x[i] *= s;
    }
  }
};

// AnalyticSignals (called at src/modules/Fields.jl:115, Models.jl:181,
// FieldAnalyzers.jl:188,196).  Assumption: half(N) = N/2 for even N; the
// analytic spectrum keeps bin 1, doubles bins 2..N/2, zeroes N/2+1..N.
template <typename T>
void rspec2aspec(std::complex<T>* S, int n) {
  int nh = n / 2;
  for (int i = 1; i < nh; i++) S[i] *= (T)2;
  for (int i = nh; i < n; i++) S[i] = 0;
}
template <typename T>
void aspec2rspec(std::complex<T>* S, int n) {
  int nh = n / 2;
  for (int i = 1; i < nh; i++) S[i] /= (T)2;
  for (int i = nh; i < n; i++) S[i] = std::conj(S[n - i]);
}

// ---------------------------------------------------------------------------
// Guard windows (src/modules/Guards.jl:322-390)
// ---------------------------------------------------------------------------
std::vector<double> guard_window_left(const std::vector<double>& x, double width, int p = 6) {
  if (width >= x.back() - x.front()) throw std::runtime_error("Guard width is larger or equal than the grid size.");
  int N = (int)x.size();
  std::vector<double> g(N, 1.0);
  if (width == 0) return g;
  double xloc1 = x.front(), xloc2 = x.front() + width;
  for (int i = 0; i < N; i++) {
    double g1 = 0.0, g2 = 1.0;
    if (x[i] >= xloc1) g1 = 1 - std::exp(-std::pow((x[i] - xloc1) / (width / 2), p));
    if (x[i] <= xloc2) g2 = std::exp(-std::pow((x[i] - xloc2) / (width / 2), p));
    g[i] = (g1 + g2) / 2;
  }
  return g;
}
std::vector<double> guard_window_right(const std::vector<double>& x, double width, int p = 6) {
  if (width >= x.back() - x.front()) throw std::runtime_error("Guard width is larger or equal than the grid size.");
  int N = (int)x.size();
  std::vector<double> g(N, 1.0);
  if (width == 0) return g;
  double xloc1 = x.back() - width, xloc2 = x.back();
  for (int i = 0; i < N; i++) {
    double g1 = 1.0, g2 = 0.0;
    if (x[i] >= xloc1) g1 = std::exp(-std::pow((x[i] - xloc1) / (width / 2), p));
    if (x[i] <= xloc2) g2 = 1 - std::exp(-std::pow((x[i] - xloc2) / (width / 2), p));
    g[i] = (g1 + g2) / 2;
  }
  return g;
}
std::vector<double> guard_window_both(const std::vector<double>& x, double width, int p = 6) {
  if (width >= (x.back() - x.front()) / 2) throw std::runtime_error("Guard width is larger or equal than the grid size.");
  auto l = guard_window_left(x, width, p), r = guard_window_right(x, width, p);
  for (size_t i = 0; i < x.size(); i++) l[i] = l[i] + r[i] - 1;
  return l;
}

// radius at level 1/e (src/modules/FieldAnalyzers.jl:357-380)
double radius(const double* x, const double* y, int n, double level = std::exp(-1.0)) {
  double ymax = -INFINITY;
  for (int i = 0; i < n; i++) ymax = std::max(ymax, y[i]);
  double ylevel = ymax * level, radl = 0, radr = 0;
  for (int i = 0; i < n; i++) if (y[i] >= ylevel) { radl = x[i]; break; }
  for (int i = n - 1; i >= 0; i--) if (y[i] >= ylevel) { radr = x[i]; break; }
  return (std::fabs(radl) + std::fabs(radr)) / 2;
}

// ---------------------------------------------------------------------------
// Tabulated ionisation rate (TabulatedFunctions.TFunction, constructed at
// src/modules/plasma_equations/photoionization_avalanche.jl:34 with x scaled
// by 1/Iu and y by tu).  Assumption (package un-vendored): linear
// interpolation of log10(y) against log10(x), clamped to the end values
// outside the table; tfpower is the log-log slope of the bracketing interval.
// ---------------------------------------------------------------------------
struct TabFunc {
  std::vector<double> lx, ly;
  double eval(double x, bool power) const {
    int n = (int)lx.size();
    double l = (x > 0) ? std::log10(x) : -INFINITY;
    int i;
    if (!(l > lx[0])) i = 0;
    else if (l >= lx[n - 1]) i = n - 2;
    else { i = (int)(std::upper_bound(lx.begin(), lx.end(), l) - lx.begin()) - 1; }
    double slope = (ly[i + 1] - ly[i]) / (lx[i + 1] - lx[i]);
    if (power) return slope;
    double lc = std::min(std::max(l, lx[0]), lx[n - 1]);
    return std::pow(10.0, ly[i] + slope * (lc - lx[i]));
  }
};

// ---------------------------------------------------------------------------
// The state: grids, operator tables, field and plasma arrays.
// ---------------------------------------------------------------------------
struct Base {
  virtual ~Base() {}
  orc_config cfg;
  int Nr = 1, Nt = 0, Nh = 0;       // Nr: number of rows (r nodes, or Nx * Ny points)
  bool rt = true, xyt = false;
  int Nx = 0, Ny = 0;
  std::vector<double> x, y, kx, ky;
  double dx = 0, dy = 0;
  Medium med{};
  // FP64 grids and scalars
  std::vector<double> r, dr, k, t, w, rdr;
  double dt = 0, w0 = 0, n0 = 0, Eu = 0, vf = 0, k0 = 0, phik = 0, phip = 0;
  std::vector<TabFunc> tabs;
  std::vector<double> frhont, Rava, Ks;
  bool qpar = true;
  virtual int precision() const = 0;
};

template <typename T>
struct State : Base {
  using C = std::complex<T>;
  int precision() const override { return sizeof(T); }
  std::vector<C> E, KZ, QZ, Ftmp;
  std::vector<T> rho, kdrho;
  std::vector<T> Rg, Tg, Kg, Wg;
  std::vector<C> Rnl[4];
  bool Ron[4] = {false, false, false, false};
  int Rkind[4] = {0, 0, 0, 0};
  std::vector<C> Hraman;
  std::vector<T> Tm, RJ, JV;   // DHT
  std::unique_ptr<FFTPlan<T>> plan, planx, plany;
  std::vector<T> Rguard_dummy;

  void build(const orc_config& c, const double* E0);
  // hot path pieces
  void time_fwd(C* row) const { plan->run(row, +1); }    // P \ E
  void time_inv(C* row) const { plan->run(row, -1); }    // P * E
  void dht_columns(C* A, int inverse);
  void fft2_columns(C* A, int inverse);
  void plasma_row(int i, const C* Erow, T* rho_row, T* kd_row) const;
  void plasma_solve();
  void rhs_row(int i, const C* Es, C* dE, C* Et, C* F) const;
  void nl_step_row(int i, C* Es, T dz, std::vector<C>& scratch) const;
  void qz_space(C* dE);
  void nl_step_all(C* u, T dz);
  void linear_step(T dz);
  void zstep(T z, T dz);
  void analyze(double z, orc_diag* d, double* Fr, double* Ft, double* rho_end, double* S);
};

template <typename T>
void State<T>::build(const orc_config& c, const double* E0) {
  cfg = c;
  rt = (c.geometry == ORC_GEOM_RT);
  xyt = (c.geometry == ORC_GEOM_XYT);
  Nt = c.Nt;
  Nr = rt ? c.Nr : (xyt ? c.Nx * c.Ny : 1);
  if (Nt < 4 || (Nt & (Nt - 1))) throw std::runtime_error("Nt must be a power of two >= 4");
  if (rt && Nr < 2) throw std::runtime_error("Nr must be >= 2");
  if (xyt) {
    Nx = c.Nx; Ny = c.Ny;
    if (Nx < 2 || Ny < 2 || (Nx & (Nx - 1)) || (Ny & (Ny - 1))) throw std::runtime_error("Nx, Ny must be powers of two");
    if (c.guard3d_bug && Ny > Nt) throw std::runtime_error("guard3d_bug needs Ny <= Nt");
    planx.reset(new FFTPlan<T>(Nx));
    plany.reset(new FFTPlan<T>(Ny));
    // Grids.jl:196-201: range(min, max, length=N), kx = 2 pi fftfreq(N, 1/dx)
    x.resize(Nx); y.resize(Ny); kx.resize(Nx); ky.resize(Ny);
    dx = (c.xmax - c.xmin) / (Nx - 1);
    dy = (c.ymax - c.ymin) / (Ny - 1);
    for (int i = 0; i < Nx; i++) {
      x[i] = c.xmin + i * dx;
      int f = (i < (Nx + 1) / 2) ? i : i - Nx;
      kx[i] = 2 * PI * f / (Nx * dx);
    }
    for (int i = 0; i < Ny; i++) {
      y[i] = c.ymin + i * dy;
      int f = (i < (Ny + 1) / 2) ? i : i - Ny;
      ky[i] = 2 * PI * f / (Ny * dy);
    }
  }
  Nh = Nt / 2;
  med.kind = c.medium; med.n_const = c.n_const; med.n2 = c.n2;
  plan.reset(new FFTPlan<T>(Nt));

  double ku = 1.0 / c.ru, wu = 1.0 / c.tu;   // Units.jl:79-85

  // Grids.jl:204-223
  t.resize(Nt);
  dt = (c.tmax - c.tmin) / (Nt - 1);
  for (int j = 0; j < Nt; j++) t[j] = c.tmin + j * dt;
  w.resize(Nt);
  for (int j = 0; j < Nt; j++) {
    int f = (j < (Nt + 1) / 2) ? j : j - Nt;     // fftfreq ordering
    w[j] = 2 * PI * f / (Nt * dt);
  }
  r.assign(Nr, 0.0); dr.assign(Nr, 1.0); k.assign(Nr, 0.0); rdr.assign(Nr, 1.0);
  std::vector<double> jz;
  if (rt) {
    jz.resize(Nr + 1);
    bessel_j0_zeros(Nr + 1, jz.data());
    double S = jz[Nr];
    for (int i = 0; i < Nr; i++) { r[i] = jz[i] * c.rmax / S; k[i] = jz[i] / c.rmax; }
    for (int i = 0; i < Nr; i++) {   // Grids.jl:227-237
      if (i == 0) dr[i] = r[1] - r[0];
      else if (i == Nr - 1) dr[i] = r[Nr - 1] - r[Nr - 2];
      else dr[i] = (r[i + 1] - r[i - 1]) / 2;
      rdr[i] = r[i] * dr[i];
    }
  }

  // Fields.jl:96-120 (RT) / :310-333 (T)
  w0 = 2 * PI * C0 / c.lam0;
  n0 = med.refractive_index(w0).real();
  Eu = std::sqrt(c.Iu / (n0 * EPS0 * C0 / 2));    // Units.jl:137-141
  vf = med.group_velocity(w0);                      // spectral.jl:87
  k0 = med.k(w0);

  E.assign((size_t)Nr * Nt, C(0));
  for (size_t q = 0; q < E.size(); q++) E[q] = C((T)E0[q], 0);
  for (int i = 0; i < Nr; i++) {   // rsig2asig!: P\E, rspec2aspec, P*E
    C* row = &E[(size_t)i * Nt];
    time_fwd(row); rspec2aspec(row, Nt); time_inv(row);
  }
  rho.assign((size_t)Nr * Nt, 0); kdrho.assign((size_t)Nr * Nt, 0);
  Ftmp.assign((size_t)Nr * Nt, C(0));

  // Guards.jl:95-126 (RT), :78-89 (T), :162-195 (XYT)
  auto Tgd = guard_window_both(t, c.tguard);
  Tg.assign(Tgd.begin(), Tgd.end());
  Rg.assign(Nr, 1);
  if (rt) { auto Rgd = guard_window_right(r, c.rguard); for (int i = 0; i < Nr; i++) Rg[i] = (T)Rgd[i]; }
  Wg.resize(Nt);
  for (int j = 0; j < Nt; j++) Wg[j] = (T)std::exp(-std::pow(std::pow(w[j] * wu, 2) / (c.wguard * c.wguard), 20));
  Kg.assign((size_t)Nr * Nt, rt ? 0 : 1);
  if (xyt) {
    // field filter X(x) Y(y) T(t) and spectral filter KX(kx) KY(ky) W(w) with
    // KX = KY = 1 (Guards.jl:176-184); kernel! (:301-318) takes its third factor
    // at i3 = the y index, which guard3d_bug reproduces: X Y T(iy) and W(iky).
    auto Xg = guard_window_both(x, c.xguard), Yg = guard_window_both(y, c.yguard);
    std::vector<T> Wfull = Wg;
    for (int iy = 0; iy < Ny; iy++)
      for (int ix = 0; ix < Nx; ix++) {
        const int i = ix + Nx * iy;
        Rg[i] = (T)(Xg[ix] * Yg[iy] * (c.guard3d_bug ? Tgd[iy] : 1.0));
        for (int j = 0; j < Nt; j++) Kg[(size_t)i * Nt + j] = c.guard3d_bug ? Wfull[iy] : (T)1;
      }
    if (c.guard3d_bug) {
      for (int j = 0; j < Nt; j++) { Tg[j] = 1; Wg[j] = 1; }
    }
  }
  if (rt) {
    for (int j = 0; j < Nt; j++) {
      double kk = med.k(w[j] * wu);
      double kmax = kk * std::sin(c.kguard * PI / 180.0);   // sind
      if (kmax != 0)
        for (int i = 0; i < Nr; i++)
          Kg[(size_t)i * Nt + j] = (T)std::exp(-std::pow(std::pow(k[i] * ku, 2) / (kmax * kmax), 20));
    }
  }

  // Linear propagator KZ (spectral.jl:78-101 RT, :131-156 XYT, :55-75 T; Kfunc :160-184)
  bool kpar = (rt || xyt) ? (c.kparaxial != 0) : true;   // Input.jl:96-103 forces T paraxial
  KZ.assign((size_t)Nr * Nt, C(0));
  for (int j = 0; j < Nt; j++) {
    double ww = w[j] * wu;
    cd beta = med.beta(ww);
    for (int i = 0; i < Nr; i++) {
      double kt = k[i] * ku;
      if (xyt) {
        const double kxs = kx[i % Nx] / c.xu, kys = ky[i / Nx] / c.yu;
        kt = std::sqrt(kxs * kxs + kys * kys);
      }
      cd K;
      if (kpar) K = (beta != cd(0)) ? beta - kt * kt / (2.0 * beta) : cd(0);
      else K = std::sqrt(beta * beta - kt * kt + cd(0));
      cd kz = K * c.zu - ww / vf * c.zu;
      KZ[(size_t)i * Nt + j] = C((T)kz.real(), (T)kz.imag());
    }
  }
  // Nonlinear prefactor QZ (nonlinear_propagators.jl:139-208; Qfunc :210-242):
  // paraxial mu0 mu w^2 / (2 beta), or with kz = sqrt(beta^2 - kt^2 + 0im)
  // mu0 mu w^2 / (2 kz) (0 where beta, resp. kz, is 0)
  qpar = (rt || xyt) ? (c.qparaxial != 0) : true;
  QZ.assign((size_t)Nr * Nt, C(0));
  for (int j = 0; j < Nt; j++) {
    double ww = w[j] * wu;
    cd beta = med.beta(ww);
    for (int i = 0; i < Nr; i++) {
      cd Q;
      if (qpar) {
        Q = (beta == cd(0)) ? cd(0) : MU0 * med.permeability(ww) * ww * ww / (2.0 * beta);
      } else {
        double kt = k[i] * ku;
        if (xyt) {
          const double kxs = kx[i % Nx] / c.xu, kys = ky[i / Nx] / c.yu;
          kt = std::sqrt(kxs * kxs + kys * kys);
        }
        cd kz = std::sqrt(beta * beta - kt * kt + cd(0));
        Q = (kz == cd(0)) ? cd(0) : MU0 * med.permeability(ww) * ww * ww / (2.0 * kz);
      }
      Q = Q * c.zu / Eu;
      QZ[(size_t)i * Nt + j] = C((T)Q.real(), (T)Q.imag());
    }
  }

  // Responses (medium_responses/*.jl)
  double chi3 = med.chi3(w0);
  double Rk = EPS0 * chi3 * Eu * Eu * Eu;   // cubic.jl:11-12, raman.jl:9-12
  for (int q = 0; q < 4; q++) Rnl[q].assign(Nt, C(0));
  if (c.nonlinearity) {
    if (c.cubic) { Ron[0] = true; Rkind[0] = c.cubic; for (int j = 0; j < Nt; j++) Rnl[0][j] = C((T)Rk, 0); }
    if (c.raman) {
      Ron[1] = true; Rkind[1] = c.raman;
      for (int j = 0; j < Nt; j++) Rnl[1][j] = C((T)Rk, 0);
      // raman.jl:15-29 with the Mlejnek response of test/diffraction/RT/air.jl:48-58
      double tshift = c.tmin + 0.5 * (c.tmax - c.tmin);
      std::vector<cd> H(Nt);
      for (int j = 0; j < Nt; j++) {
        double tt = (t[j] - tshift) * c.tu, h = 0;
        if (tt >= 0) {
          double Om = 20.6e12, Ga = 26e12, La = std::sqrt(Om * Om - Ga * Ga / 4.0);
          h = Om * Om * std::exp(-0.5 * Ga * tt) * std::sin(La * tt) / La;
        }
        H[j] = cd(h * c.tu, 0);
      }
      std::vector<cd> Hs(Nt);
      for (int j = 0; j < Nt; j++) Hs[j] = H[(j + Nt / 2) % Nt];    // ifftshift (even Nt)
      FFTPlan<double> p(Nt);
      p.run(Hs.data(), +1);
      Hraman.resize(Nt);
      for (int j = 0; j < Nt; j++) { cd v = Hs[j] * (double)Nt * dt; Hraman[j] = C((T)v.real(), (T)v.imag()); }
    }
    if (c.current_free) {   // current_free.jl:4-28
      Ron[2] = true;
      double MR = c.mr * ME;
      for (int j = 0; j < Nt; j++) if (w[j] != 0) {
        double ww = w[j] * wu;
        cd v = cd(0, 1) / ww * QE * QE / MR / (cd(c.nuc, 0) - cd(0, 1) * ww) * c.rhou * Eu;
        Rnl[2][j] = C((T)v.real(), (T)v.imag());
      }
    }
    if (c.current_losses) {   // current_losses.jl:4-33
      Ron[3] = true;
      for (int j = 0; j < Nt; j++) if (w[j] != 0) {
        double ww = w[j] * wu;
        cd v = cd(0, 1) / ww * HBAR * w0 * c.rhou / (c.tu * Eu);
        Rnl[3][j] = C((T)v.real(), (T)v.imag());
      }
    }
  }

  // Plasma equation (photoionization_avalanche.jl:1-116)
  tabs.clear(); frhont.clear(); Rava.clear(); Ks.clear();
  if (c.plasma) {
    if (c.ncomp < 1 || c.ncomp > ORC_MAXCOMP) throw std::runtime_error("bad ncomp");
    for (int q = 0; q < c.ncomp; q++) {
      TabFunc tf;
      int n = c.comp_ntab[q];
      if (n < 2) throw std::runtime_error("rate table needs >= 2 rows");
      for (int m = 0; m < n; m++) {
        tf.lx.push_back(std::log10(c.comp_tab_x[q][m] / c.Iu));
        tf.ly.push_back(std::log10(c.comp_tab_y[q][m] * c.tu));
      }
      tabs.push_back(tf);
      frhont.push_back(c.comp_frac[q] * c.pe_rho_nt / c.rhou);
      double Ui = c.comp_Ui[q] * QE;
      double MR = c.pe_mr * ME;
      double sigmaB = QE * QE / MR * c.pe_nuc / (c.pe_nuc * c.pe_nuc + w0 * w0);
      Rava.push_back(sigmaB / Ui * c.tu * Eu * Eu);
      Ks.push_back(std::ceil(Ui / (HBAR * w0)));
    }
  }

  // AdaptiveSteps.jl:46-72
  {
    double mu = med.permeability(w0);
    double QZ0 = MU0 * mu * w0 * w0 / (2 * k0) * c.zu / Eu;
    double Rnl0 = EPS0 * chi3 * 3.0 / 4.0 * Eu * Eu * Eu;
    phik = QZ0 * std::fabs(Rnl0);
    double MR = c.mr * ME;
    cd Rp = cd(0, 1) / w0 * QE * QE / MR / (cd(c.nuc, 0) - cd(0, 1) * w0) * c.rhou * Eu;
    phip = QZ0 * std::fabs(Rp.real());
  }

  // DHT plan (HankelTransforms.plan_dht, used at Fields.jl:108-110).
  // Guizar-Sicairos & Gutierrez-Vega, JOSA A 21, 53 (2004):
  //   T_mn = 2 J0(j_m j_n / S) / (S |J1(j_m)| |J1(j_n)|), S = j_{N+1}
  //   forward f -> JV .* (T * (RJ .* f)), RJ = R/|J1|, JV = |J1|/V, V = S/(2 pi R)
  //   inverse F -> (T * (F ./ JV)) ./ RJ
  if (rt) {
    double S = jz[Nr], R = c.rmax, V = S / (2 * PI * R);
    std::vector<double> J1(Nr);
    for (int i = 0; i < Nr; i++) J1[i] = std::fabs(std::cyl_bessel_j(1.0, jz[i]));
    Tm.resize((size_t)Nr * Nr); RJ.resize(Nr); JV.resize(Nr);
    for (int i = 0; i < Nr; i++) { RJ[i] = (T)(R / J1[i]); JV[i] = (T)(J1[i] / V); }
#pragma omp parallel for schedule(static)
    for (int m = 0; m < Nr; m++)
      for (int n = 0; n < Nr; n++)
        Tm[(size_t)m * Nr + n] = (T)(2.0 * std::cyl_bessel_j(0.0, jz[m] * jz[n] / S) / (S * J1[m] * J1[n]));
  }

  if (c.plasma) plasma_solve();   // Models.jl:57-67 solves once at construction
}

// QDHT of columns 0..Nh-1 of A [Nr][Nt] in place (HankelTransforms `P * E` /
// `P \ E` with region (Nr, Nthalf), Fields.jl:108-110; Models.jl:135-148).
template <typename T>
void State<T>::dht_columns(C* A, int inverse) {
  std::vector<C> X((size_t)Nr * Nh);
  // pre-scale
#pragma omp parallel for schedule(static)
  for (int n = 0; n < Nr; n++) {
    T s = inverse ? (T)1 / JV[n] : RJ[n];
    C* row = A + (size_t)n * Nt;
    for (int j = 0; j < Nh; j++) row[j] *= s;
  }
#pragma omp parallel for schedule(static)
  for (int m = 0; m < Nr; m++) {
    C* x = &X[(size_t)m * Nh];
    const T* Trow = &Tm[(size_t)m * Nr];
    for (int n = 0; n < Nr; n++) {
      T a = Trow[n];
      const C* src = A + (size_t)n * Nt;
      for (int j = 0; j < Nh; j++) x[j] += a * src[j];
    }
  }
#pragma omp parallel for schedule(static)
  for (int m = 0; m < Nr; m++) {
    T s = inverse ? (T)1 / RJ[m] : JV[m];
    C* dst = A + (size_t)m * Nt;
    const C* x = &X[(size_t)m * Nh];
    for (int j = 0; j < Nh; j++) dst[j] = x[j] * s;
  }
}

// One row of the plasma ODE (plasma_equations.jl:46-72 on the GPU, :1-24 on
// the CPU; right-hand sides photoionization_avalanche.jl:119-205).
// ODEIntegrators.RK3 assumption (package un-vendored): Kutta's third-order
// tableau c=(0,1/2,1), a21=1/2, a31=-1, a32=2, b=(1/6,2/3,1/6).
// TabulatedFunctions.linterp of the complex field at the stage time.
template <typename T>
void State<T>::plasma_row(int i, const C* EE, T* rho_row, T* kd_row) const {
  (void)i;
  const int nc = cfg.ncomp;
  const T dtT = (T)dt;
  const T t0 = (T)t[0];
  auto linterp = [&](T tt) -> C {
    T x = (tt - t0) / dtT;
    int j = (int)std::floor(x);
    if (j < 0) j = 0;
    if (j > Nt - 2) j = Nt - 2;
    T a = (tt - ((T)t[j])) / dtT;
    return EE[j] + a * (EE[j + 1] - EE[j]);
  };
  auto fiarg = [&](C e) -> T { return cfg.pe_ereal ? e.real() * e.real() : std::norm(e); };
  T R1[ORC_MAXCOMP];
  auto rates = [&](T tt, T& E2) {
    C e = linterp(tt);
    T I = fiarg(e);
    E2 = e.real() * e.real();
    for (int q = 0; q < nc; q++) R1[q] = (T)tabs[q].eval((double)I, false);
  };
  // func_photoionization_avalanche / func_photoionization
  auto f_ava = [&](const T* u, T tt, T* du) {
    T E2; rates(tt, E2);
    T tot = 0; for (int q = 0; q < nc; q++) tot += u[q];
    for (int q = 0; q < nc; q++) {
      T f = (T)frhont[q], R2 = (T)Rava[q] * E2;
      du[q] = R1[q] * (f - u[q]) + R2 * (f - u[q]) / f * tot;
    }
  };
  auto f_pi = [&](const T* u, T tt, T* du) {
    T E2; rates(tt, E2);
    for (int q = 0; q < nc; q++) du[q] = R1[q] * ((T)frhont[q] - u[q]);
  };
  auto rk3 = [&](auto&& f, T* u, T tt) {
    T k1[ORC_MAXCOMP], k2[ORC_MAXCOMP], k3[ORC_MAXCOMP], tmp[ORC_MAXCOMP];
    f(u, tt, k1);
    for (int q = 0; q < nc; q++) tmp[q] = u[q] + dtT * ((T)0.5 * k1[q]);
    f(tmp, tt + (T)0.5 * dtT, k2);
    for (int q = 0; q < nc; q++) tmp[q] = u[q] + dtT * (-k1[q] + (T)2 * k2[q]);
    f(tmp, tt + dtT, k3);
    for (int q = 0; q < nc; q++) u[q] = u[q] + dtT * (k1[q] / (T)6 + (T)2 * k2[q] / (T)3 + k3[q] / (T)6);
  };
  T u[ORC_MAXCOMP], ut[ORC_MAXCOMP];
  for (int q = 0; q < nc; q++) { u[q] = (T)(cfg.pe_rho0 / cfg.rhou); ut[q] = u[q]; }
  T s = 0; for (int q = 0; q < nc; q++) s += u[q];
  rho_row[0] = s;
  for (int j = 0; j < Nt - 1; j++) {
    T tj = (T)t[j];
    rk3(f_ava, u, tj);
    s = 0; for (int q = 0; q < nc; q++) s += u[q];
    rho_row[j + 1] = s;
    rk3(f_pi, ut, tj);
    // kdrho_func (photoionization_avalanche.jl:176-205) at t_j with the state after step j
    C e = linterp(tj);
    T I = fiarg(e), kd = 0;
    for (int q = 0; q < nc; q++) {
      T r1 = (T)tabs[q].eval((double)I, false);
      T K = cfg.pe_kdep ? (T)tabs[q].eval((double)I, true) : (T)Ks[q];
      kd += K * (r1 * ((T)frhont[q] - ut[q]));
    }
    kd_row[j] = kd;
  }
}

template <typename T>
void State<T>::plasma_solve() {
#pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < Nr; i++)
    plasma_row(i, &E[(size_t)i * Nt], &rho[(size_t)i * Nt], &kdrho[(size_t)i * Nt]);
}

// func! for one row (nonlinear_propagators.jl:38-67; QPARAXIAL branch).
// Es: stage input spectrum (row), dE: output, Et/F: scratch rows.
template <typename T>
void State<T>::rhs_row(int i, const C* Es, C* dE, C* Et, C* F) const {
  const size_t o = (size_t)i * Nt;
  for (int j = 0; j < Nt; j++) Et[j] = Es[j];
  time_inv(Et);                                    // inverse_transform_time!
  for (int j = 0; j < Nt; j++) dE[j] = 0;
  for (int q = 0; q < 4; q++) {
    if (!Ron[q]) continue;
    // resp.calculate (cubic.jl:33-54, raman.jl:46-65, current_free.jl:31-37, current_losses.jl:36-83)
    if (q == 0) {
      for (int j = 0; j < Nt; j++) {
        T re = Et[j].real();
        F[j] = (Rkind[0] == ORC_THG) ? C(re * re * re, 0) : C((T)0.75 * std::norm(Et[j]) * re, 0);
      }
    } else if (q == 1) {
      for (int j = 0; j < Nt; j++) {
        T re = Et[j].real();
        F[j] = (Rkind[1] == ORC_THG) ? C(re * re, 0) : C((T)0.75 * std::norm(Et[j]), 0);
      }
      time_fwd(F);                                 // convolution! (raman.jl:68-111)
      for (int j = 0; j < Nt; j++) F[j] = Hraman[j] * F[j];
      time_inv(F);
      for (int j = 0; j < Nt; j++) F[j] = F[j] * Et[j].real();
    } else if (q == 2) {
      for (int j = 0; j < Nt; j++) F[j] = C(rho[o + j] * Et[j].real(), 0);
    } else {
      for (int j = 0; j < Nt; j++) {
        T fe = cfg.losses_ereal ? Et[j].real() * Et[j].real() : std::norm(Et[j]);
        T inv = (fe >= (T)1e-30) ? (T)1 / fe : (T)0;   // inverse_kernel
        F[j] = C(kdrho[o + j] * inv * Et[j].real(), 0);
      }
    }
    for (int j = 0; j < Nt; j++) F[j] *= Rg[i] * Tg[j];   // apply_field_filter!
    time_fwd(F);                                          // rsig2aspec!
    rspec2aspec(F, Nt);
    for (int j = 0; j < Nt; j++) dE[j] += Rnl[q][j] * F[j];   // update_dE!
  }
  if (qpar) for (int j = 0; j < Nt; j++) dE[j] = C(0, 1) * QZ[o + j] * dE[j];
}

// func!'s non-paraxial end (nonlinear_propagators.jl:58-63) on all rows of dE:
// forward space transform, times i QZ(k, w), spectral filter, inverse.
template <typename T>
void State<T>::qz_space(C* dE) {
  if (rt) dht_columns(dE, 0);
  if (xyt) fft2_columns(dE, 0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < Nr; i++)
    for (int j = 0; j < Nt; j++) {
      const size_t q = (size_t)i * Nt + j;
      dE[q] = C(0, 1) * QZ[q] * dE[q];
      dE[q] = dE[q] * (Kg[q] * Wg[j]);      // apply_spectral_filter! (Guards.jl:232-261)
    }
  if (rt) dht_columns(dE, 1);
  if (xyt) fft2_columns(dE, 1);
}

// ODEIntegrators.step! with RK3 over all rows at once, for the non-paraxial
// right-hand side, which couples the rows through the space transform.
template <typename T>
void State<T>::nl_step_all(C* u, T dz) {
  const size_t n = (size_t)Nr * Nt;
  std::vector<C> k1(n), k2(n), k3(n), ut(n);
  auto rhs = [&](const C* in, C* out) {
#pragma omp parallel
    {
      std::vector<C> Et(Nt), F(Nt);
#pragma omp for schedule(dynamic, 2)
      for (int i = 0; i < Nr; i++) rhs_row(i, in + (size_t)i * Nt, out + (size_t)i * Nt, Et.data(), F.data());
    }
    qz_space(out);
  };
  rhs(u, k1.data());
  for (size_t q = 0; q < n; q++) ut[q] = u[q] + dz * ((T)0.5 * k1[q]);
  rhs(ut.data(), k2.data());
  for (size_t q = 0; q < n; q++) ut[q] = u[q] + dz * (-k1[q] + (T)2 * k2[q]);
  rhs(ut.data(), k3.data());
  for (size_t q = 0; q < n; q++) u[q] = u[q] + dz * (k1[q] / (T)6 + (T)2 * k2[q] / (T)3 + k3[q] / (T)6);
}

// ODEIntegrators.step! with RK3 on one row (nonlinear_propagators.jl:70-74)
template <typename T>
void State<T>::nl_step_row(int i, C* u, T dz, std::vector<C>& s) const {
  C* k1 = &s[0]; C* k2 = &s[Nt]; C* k3 = &s[2 * Nt]; C* ut = &s[3 * Nt];
  C* Et = &s[4 * Nt]; C* F = &s[5 * Nt];
  rhs_row(i, u, k1, Et, F);
  for (int j = 0; j < Nt; j++) ut[j] = u[j] + dz * ((T)0.5 * k1[j]);
  rhs_row(i, ut, k2, Et, F);
  for (int j = 0; j < Nt; j++) ut[j] = u[j] + dz * (-k1[j] + (T)2 * k2[j]);
  rhs_row(i, ut, k3, Et, F);
  for (int j = 0; j < Nt; j++) u[j] = u[j] + dz * (k1[j] / (T)6 + (T)2 * k2[j] / (T)3 + k3[j] / (T)6);
}

// propagate!(E, LP::LinearPropagatorSpectral, dz) (spectral.jl:16-26)
// The 2-D FFT over (x, y) of every t/w column of A [rows][Nt] (Field XYT's
// space plan plan_fft!(E, [1, 2]), Fields.jl:152; forward P * E unnormalised
// exp(-i k x), inverse P \ E normalised, Models.jl:135-148).
template <typename T>
void State<T>::fft2_columns(C* A, int inverse) {
  const int sgn = inverse ? +1 : -1;   // FFTPlan::run: +1 = normalised exp(+), -1 = exp(-)
#pragma omp parallel
  {
    std::vector<C> v(std::max(Nx, Ny));
#pragma omp for schedule(static)
    for (int j = 0; j < Nt; j++) {
      for (int iy = 0; iy < Ny; iy++) {
        for (int ix = 0; ix < Nx; ix++) v[ix] = A[(size_t)(ix + Nx * iy) * Nt + j];
        planx->run(v.data(), sgn);
        for (int ix = 0; ix < Nx; ix++) A[(size_t)(ix + Nx * iy) * Nt + j] = v[ix];
      }
      for (int ix = 0; ix < Nx; ix++) {
        for (int iy = 0; iy < Ny; iy++) v[iy] = A[(size_t)(ix + Nx * iy) * Nt + j];
        plany->run(v.data(), sgn);
        for (int iy = 0; iy < Ny; iy++) A[(size_t)(ix + Nx * iy) * Nt + j] = v[iy];
      }
    }
  }
}

template <typename T>
void State<T>::linear_step(T dz) {
  if (rt) dht_columns(E.data(), 0);
  if (xyt) fft2_columns(E.data(), 0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < Nr; i++)
    for (int j = 0; j < Nt; j++) {
      size_t q = (size_t)i * Nt + j;
      C ph = std::exp(C(0, 1) * KZ[q] * dz);
      E[q] = E[q] * ph;
      E[q] = E[q] * (Kg[q] * Wg[j]);      // apply_spectral_filter! (Guards.jl:283-298)
    }
  if (rt) dht_columns(E.data(), 1);
  if (xyt) fft2_columns(E.data(), 1);
}

// Models.zstep (Models.jl:75-121)
template <typename T>
void State<T>::zstep(T z, T dz) {
  (void)z;
  if (cfg.plasma) plasma_solve();
#pragma omp parallel for schedule(static)
  for (int i = 0; i < Nr; i++) time_fwd(&E[(size_t)i * Nt]);
  if (cfg.nonlinearity && !qpar) {
    nl_step_all(E.data(), dz);
  } else if (cfg.nonlinearity) {
#pragma omp parallel
    {
      std::vector<C> scratch((size_t)6 * Nt);
#pragma omp for schedule(dynamic, 2)
      for (int i = 0; i < Nr; i++) nl_step_row(i, &E[(size_t)i * Nt], dz, scratch);
    }
  }
  linear_step(dz);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < Nr; i++) {
    C* row = &E[(size_t)i * Nt];
    time_inv(row);
    for (int j = 0; j < Nt; j++) row[j] *= Rg[i] * Tg[j];   // apply_field_filter! (Guards.jl:265-280)
  }
}

// FieldAnalyzers.analyze! RT (FieldAnalyzers.jl:166-215) and T (:99-114).
// Note the side effect the reference has: E is projected onto its analytic
// part (aspec2rspec! then rspec2aspec!, :188-197).
template <typename T>
void State<T>::analyze(double z, orc_diag* d, double* Fr, double* Ft, double* rho_end, double* S) {
  const double twopi = 2 * PI;
  int Nw = Nt / 2 + 1;
  std::vector<double> fr(Nr, 0.0), ft(Nt, 0.0), re(Nr, 0.0), sp(Nw, 0.0);
  double Imax = 0, rhomax = 0;
  for (int i = 0; i < Nr; i++) {
    double s = 0;
    for (int j = 0; j < Nt; j++) {
      double I = std::norm(std::complex<double>(E[(size_t)i * Nt + j]));
      s += I;
      ft[j] += I * rdr[i];
      rhomax = std::max(rhomax, (double)rho[(size_t)i * Nt + j]);
    }
    fr[i] = s * dt;
    re[i] = rho[(size_t)i * Nt + Nt - 1];
  }
  // row weights: 2 pi r dr (RT, rdr holds r dr), dx dy (XYT, rdr = 1)
  const double wscale = xyt ? dx * dy : twopi;
  for (int j = 0; j < Nt; j++) ft[j] *= wscale;
  std::vector<C> row(Nt);
  for (int i = 0; i < Nr; i++) {
    C* Er = &E[(size_t)i * Nt];
    time_fwd(Er);
    aspec2rspec(Er, Nt);
    for (int j = 0; j < Nw; j++) {
      std::complex<double> v = std::complex<double>(Er[j]) * (2.0 * Nt * dt);
      sp[j] += std::norm(v) * rdr[i];
    }
    rspec2aspec(Er, Nt);
    time_inv(Er);
    // Imax is taken after the projection (FieldAnalyzers.jl:203)
    for (int j = 0; j < Nt; j++) Imax = std::max(Imax, (double)std::norm(std::complex<double>(Er[j])));
  }
  for (int j = 0; j < Nw; j++) sp[j] *= wscale;
  d->z = z;
  if (xyt) {
    // FieldAnalyzers.jl:330-352
    const int imax = (int)(std::max_element(fr.begin(), fr.end()) - fr.begin());
    d->Fmax = fr[imax];
    d->Imax = Imax;
    d->rhomax = rhomax;
    double De = 0, W = 0;
    for (int i = 0; i < Nr; i++) De += re[i];
    for (int j = 0; j < Nt; j++) W += ft[j];
    d->De = De * dx * dy;
    d->W = W * dt;
    const int ixm = imax % Nx, iym = imax / Nx;
    std::vector<double> px(Nx), py(Ny);
    for (int ix = 0; ix < Nx; ix++) px[ix] = fr[ix + Nx * iym];
    for (int iy = 0; iy < Ny; iy++) py[iy] = fr[ixm + Nx * iy];
    d->rfil = radius(x.data(), px.data(), Nx);   // ax
    d->rpl = radius(y.data(), py.data(), Ny);    // ay
    d->tau = 0;
  } else if (rt) {
    d->Fmax = *std::max_element(fr.begin(), fr.end());
    d->Imax = Imax;
    d->rhomax = rhomax;
    double De = 0, W = 0;
    for (int i = 0; i < Nr; i++) { De += re[i] * rdr[i]; W += fr[i] * rdr[i]; }
    d->De = twopi * De;
    d->rfil = 2 * radius(r.data(), fr.data(), Nr);
    d->rpl = 2 * radius(r.data(), re.data(), Nr);
    d->tau = radius(t.data(), ft.data(), Nt);
    d->W = twopi * W;
  } else {
    std::vector<double> I(Nt);
    double F = 0;
    for (int j = 0; j < Nt; j++) { I[j] = std::norm(std::complex<double>(E[j])); F += I[j]; }
    d->Fmax = 0; d->Imax = Imax; d->rhomax = rhomax; d->De = 0; d->rfil = 0; d->rpl = 0;
    d->tau = radius(t.data(), I.data(), Nt);   // "duration"
    d->W = F * dt;                              // "F"
  }
  if (Fr) std::copy(fr.begin(), fr.end(), Fr);
  if (Ft) std::copy(ft.begin(), ft.end(), Ft);
  if (rho_end) std::copy(re.begin(), re.end(), rho_end);
  if (S) std::copy(sp.begin(), sp.end(), S);
}

// ---------------------------------------------------------------------------
// Spatial geometries R and XY
// ---------------------------------------------------------------------------
struct SpatialBase {
  virtual ~SpatialBase() {}
  virtual int precision() const = 0;
};

template <typename T>
struct Spatial : SpatialBase {
  using C = std::complex<T>;
  int precision() const override { return sizeof(T); }
  orc_spatial_config cfg{};
  bool xy = false;
  int n = 0, Nr = 0, Nx = 0, Ny = 0;
  Medium med{};
  std::vector<double> r, dr, k, x, y, kx, ky, rdr;
  double dx = 0, dy = 0, w0 = 0, n0 = 0, Eu = 0, vf = 0, k0 = 0, phik = 0;
  T Rnl = 0;
  std::vector<C> E, KZ, QZ;
  std::vector<T> Kg, Fg, Xg, Yg, KXg, KYg;
  std::vector<T> Tm, RJ, JV;
  std::unique_ptr<FFTPlan<T>> planx, plany;

  void build(const orc_spatial_config& c, const double* E0) {
    cfg = c;
    xy = c.geometry == ORC_GEOM_XY;
    if (!xy && c.geometry != ORC_GEOM_R) throw std::runtime_error("geometry must be R or XY");
    if (c.alg != 3 && c.alg != 4) throw std::runtime_error("alg must be 3 or 4");
    med.kind = c.medium; med.n_const = c.n_const; med.n2 = c.n2;
    std::vector<double> jz;
    if (!xy) {
      // Grids.jl:42-45 with _grid_spatial_axial (:204-216)
      Nr = n = c.Nr;
      if (Nr < 2) throw std::runtime_error("Nr must be >= 2");
      jz.resize(Nr + 1);
      bessel_j0_zeros(Nr + 1, jz.data());
      const double S = jz[Nr];
      r.resize(Nr); k.resize(Nr); dr.resize(Nr); rdr.resize(Nr);
      for (int i = 0; i < Nr; i++) { r[i] = jz[i] * c.rmax / S; k[i] = jz[i] / c.rmax; }
      for (int i = 0; i < Nr; i++) {
        dr[i] = (i == 0) ? r[1] - r[0] : (i == Nr - 1 ? r[Nr - 1] - r[Nr - 2] : (r[i + 1] - r[i - 1]) / 2);
        rdr[i] = r[i] * dr[i];
      }
    } else {
      // Grids.jl:144-155 with _grid_spatial_rectangular: range(min, max, N), 2 pi fftfreq(N, 1/dx)
      Nx = c.Nx; Ny = c.Ny; n = Nx * Ny;
      if (Nx < 2 || Ny < 2 || (Nx & (Nx - 1)) || (Ny & (Ny - 1))) throw std::runtime_error("Nx, Ny must be powers of two");
      planx.reset(new FFTPlan<T>(Nx));
      plany.reset(new FFTPlan<T>(Ny));
      dx = (c.xmax - c.xmin) / (Nx - 1);
      dy = (c.ymax - c.ymin) / (Ny - 1);
      x.resize(Nx); y.resize(Ny); kx.resize(Nx); ky.resize(Ny);
      for (int i = 0; i < Nx; i++) { x[i] = c.xmin + i * dx; kx[i] = 2 * PI * ((i < (Nx + 1) / 2) ? i : i - Nx) / (Nx * dx); }
      for (int i = 0; i < Ny; i++) { y[i] = c.ymin + i * dy; ky[i] = 2 * PI * ((i < (Ny + 1) / 2) ? i : i - Ny) / (Ny * dy); }
    }
    // Fields.jl:31-51, 123-138; Units.jl:137-141
    w0 = 2 * PI * C0 / c.lam0;
    n0 = med.refractive_index(w0).real();
    Eu = std::sqrt(c.Iu / (n0 * EPS0 * C0 / 2));
    vf = med.group_velocity(w0);
    k0 = med.k(w0);
    E.resize(n);
    for (int q = 0; q < n; q++) E[q] = C((T)E0[2 * q], (T)E0[2 * q + 1]);
    // transverse wavenumber of every spectral point [1/m]
    auto kt_of = [&](int q) {
      if (!xy) return k[q] / c.ru;
      const double a = kx[q % Nx] / c.xu, b = ky[q / Nx] / c.yu;
      return std::sqrt(a * a + b * b);
    };
    // Guards.jl:41-57 (R), :132-157 (XY)
    Kg.assign(n, 1); Fg.assign(n, 1);
    if (!xy) {
      auto Rgd = guard_window_right(r, c.rguard);
      const double kmax = k0 * std::sin(c.kguard * PI / 180.0);
      for (int i = 0; i < n; i++) {
        Fg[i] = (T)Rgd[i];
        Kg[i] = (T)std::exp(-std::pow(std::pow(k[i] / c.ru, 2) / (kmax * kmax), 20));
      }
    } else {
      auto Xd = guard_window_both(x, c.xguard), Yd = guard_window_both(y, c.yguard);
      const double kxmax = k0 * std::sin(c.kxguard * PI / 180.0), kymax = k0 * std::sin(c.kyguard * PI / 180.0);
      Xg.resize(Nx); Yg.resize(Ny); KXg.resize(Nx); KYg.resize(Ny);
      for (int i = 0; i < Nx; i++) {
        Xg[i] = (T)Xd[i];
        KXg[i] = (T)std::exp(-std::pow(std::pow(kx[i] / c.xu, 2) / (kxmax * kxmax), 20));
      }
      for (int i = 0; i < Ny; i++) {
        Yg[i] = (T)Yd[i];
        KYg[i] = (T)std::exp(-std::pow(std::pow(ky[i] / c.yu, 2) / (kymax * kymax), 20));
      }
      for (int q = 0; q < n; q++) { Fg[q] = Xg[q % Nx] * Yg[q / Nx]; Kg[q] = KXg[q % Nx] * KYg[q / Nx]; }
    }
    // KZ (spectral.jl:30-50 R, :104-128 XY; Kfunc :160-184)
    const cd beta = med.beta(w0);
    KZ.resize(n);
    for (int q = 0; q < n; q++) {
      const double kt = kt_of(q);
      cd K;
      if (c.kparaxial) K = (beta != cd(0)) ? beta - kt * kt / (2.0 * beta) : cd(0);
      else K = std::sqrt(beta * beta - kt * kt + cd(0));
      const cd kz = K * c.zu - w0 / vf * c.zu;
      KZ[q] = C((T)kz.real(), (T)kz.imag());
    }
    // QZ (nonlinear_propagators.jl:78-96 R, :162-180 XY; Qfunc :210-242)
    QZ.resize(n);
    const double mu = med.permeability(w0);
    for (int q = 0; q < n; q++) {
      cd Q;
      if (c.qparaxial) {
        Q = (beta == cd(0)) ? cd(0) : MU0 * mu * w0 * w0 / (2.0 * beta);
      } else {
        const double kt = kt_of(q);
        const cd kz = std::sqrt(beta * beta - kt * kt + cd(0));
        Q = (kz == cd(0)) ? cd(0) : MU0 * mu * w0 * w0 / (2.0 * kz);
      }
      Q = Q * c.zu / Eu;
      QZ[q] = C((T)Q.real(), (T)Q.imag());
    }
    // cubic.jl:4-30: Rnl = EPS0 chi3 Eu^3; AdaptiveSteps.jl:46-57: phik
    const double chi3 = med.chi3(w0);
    Rnl = (T)(EPS0 * chi3 * Eu * Eu * Eu);
    const double QZ0 = MU0 * mu * w0 * w0 / (2 * k0) * c.zu / Eu;
    phik = (c.n2 != 0) ? QZ0 * std::fabs(EPS0 * chi3 * 3.0 / 4.0 * Eu * Eu * Eu) : 0.0;
    // Hankel plan (Fields.jl:38-42): the same quasi-discrete transform as RT
    if (!xy) {
      const double S = jz[Nr], R = c.rmax, V = S / (2 * PI * R);
      std::vector<double> J1(Nr);
      for (int i = 0; i < Nr; i++) J1[i] = std::fabs(std::cyl_bessel_j(1.0, jz[i]));
      Tm.resize((size_t)Nr * Nr); RJ.resize(Nr); JV.resize(Nr);
      for (int i = 0; i < Nr; i++) { RJ[i] = (T)(R / J1[i]); JV[i] = (T)(J1[i] / V); }
#pragma omp parallel for schedule(static)
      for (int m = 0; m < Nr; m++)
        for (int j = 0; j < Nr; j++)
          Tm[(size_t)m * Nr + j] = (T)(2.0 * std::cyl_bessel_j(0.0, jz[m] * jz[j] / S) / (S * J1[m] * J1[j]));
    }
  }

  // forward_transform_space! / inverse_transform_space! (Models.jl:135-148)
  void transform(std::vector<C>& A, bool inverse) {
    if (!xy) {
      std::vector<C> X(n, C(0)), B(n);
      for (int j = 0; j < n; j++) B[j] = inverse ? A[j] / JV[j] : RJ[j] * A[j];
#pragma omp parallel for schedule(static)
      for (int m = 0; m < n; m++) {
        C acc(0);
        const T* row = &Tm[(size_t)m * n];
        for (int j = 0; j < n; j++) acc += row[j] * B[j];
        X[m] = acc;
      }
      for (int m = 0; m < n; m++) A[m] = inverse ? X[m] / RJ[m] : JV[m] * X[m];
      return;
    }
    const int sign = inverse ? +1 : -1;   // P * E forward (exp(-i)), P \ E inverse normalised
    for (int iy = 0; iy < Ny; iy++) planx->run(&A[(size_t)iy * Nx], sign);
    std::vector<C> col(Ny);
    for (int ix = 0; ix < Nx; ix++) {
      for (int iy = 0; iy < Ny; iy++) col[iy] = A[(size_t)iy * Nx + ix];
      plany->run(col.data(), sign);
      for (int iy = 0; iy < Ny; iy++) A[(size_t)iy * Nx + ix] = col[iy];
    }
  }

  // func! (nonlinear_propagators.jl:38-67) for the spatial cubic response
  void rhs(const std::vector<C>& U, std::vector<C>& dE) {
    dE.resize(n);
    for (int q = 0; q < n; q++) {
      C F = (T)0.75 * std::norm(U[q]) * U[q];   // calc_cubic_nothg_spatial
      F = F * Fg[q];                             // apply_field_filter!
      dE[q] = Rnl * F;                           // update_dE!
    }
    if (cfg.qparaxial) {
      for (int q = 0; q < n; q++) dE[q] = C(0, 1) * QZ[q] * dE[q];
    } else {
      transform(dE, false);
      for (int q = 0; q < n; q++) dE[q] = C(0, 1) * QZ[q] * dE[q];
      for (int q = 0; q < n; q++) dE[q] = dE[q] * Kg[q];   // apply_spectral_filter!
      transform(dE, true);
    }
  }

  // ODEIntegrators step! (nonlinear_propagators.jl:70-74): RK3 (Kutta) or RK4
  void nl_step(T h) {
    std::vector<C> k1, k2, k3, k4, U(n);
    rhs(E, k1);
    for (int q = 0; q < n; q++) U[q] = E[q] + (T)0.5 * h * k1[q];
    rhs(U, k2);
    if (cfg.alg == 3) {
      for (int q = 0; q < n; q++) U[q] = E[q] + h * ((T)2 * k2[q] - k1[q]);
      rhs(U, k3);
      for (int q = 0; q < n; q++) E[q] += h * (k1[q] / (T)6 + (T)2 * k2[q] / (T)3 + k3[q] / (T)6);
    } else {
      for (int q = 0; q < n; q++) U[q] = E[q] + (T)0.5 * h * k2[q];
      rhs(U, k3);
      for (int q = 0; q < n; q++) U[q] = E[q] + h * k3[q];
      rhs(U, k4);
      for (int q = 0; q < n; q++) E[q] += h / (T)6 * (k1[q] + (T)2 * k2[q] + (T)2 * k3[q] + k4[q]);
    }
  }

  // propagate!(E, LP, dz) (spectral.jl:16-26)
  void linear_step(double dz) {
    transform(E, false);
    for (int q = 0; q < n; q++) {
      const std::complex<double> kz((double)KZ[q].real(), (double)KZ[q].imag());
      const std::complex<double> p = std::exp(std::complex<double>(0, 1) * kz * dz);
      E[q] = E[q] * C((T)p.real(), (T)p.imag());
      E[q] = E[q] * Kg[q];                      // apply_spectral_filter!
    }
    transform(E, true);
  }

  // Models.zstep (Models.jl:75-121) without plasma and t transforms
  void zstep(double dz) {
    if (cfg.nonlinearity) nl_step((T)dz);
    linear_step(dz);
    for (int q = 0; q < n; q++) E[q] = E[q] * Fg[q];   // apply_field_filter!
  }

  // FieldAnalyzerR (FieldAnalyzers.jl:52-70) / XY (:251-266)
  void analyze(double* o) {
    std::vector<double> I(n);
    double Imax = -INFINITY, P = 0;
    long idx = 0;
    for (int q = 0; q < n; q++) {
      I[q] = (double)std::norm(E[q]);
      if (I[q] > Imax) { Imax = I[q]; idx = q; }
    }
    for (int q = 0; q < n; q++) P += xy ? I[q] : I[q] * rdr[q];
    o[0] = Imax; o[1] = o[2] = o[3] = 0;
    if (!xy) {
      o[1] = 2 * radius(r.data(), I.data(), n);
      o[4] = 2 * PI * P;
    } else {
      const int im = (int)(idx % Nx), jm = (int)(idx / Nx);
      std::vector<double> line(Nx), colv(Ny);
      for (int i = 0; i < Nx; i++) line[i] = I[(size_t)jm * Nx + i];
      for (int j = 0; j < Ny; j++) colv[j] = I[(size_t)j * Nx + im];
      o[2] = radius(x.data(), line.data(), Nx);
      o[3] = radius(y.data(), colv.data(), Ny);
      o[4] = P * dx * dy;
    }
    o[5] = (double)idx;
  }
};

SpatialBase* SB(void* h) { return static_cast<SpatialBase*>(h); }

template <typename Fn>
int sdispatch(void* h, Fn&& fn);

template <typename F>
int guarded(F&& f) {
  try { f(); return 0; }
  catch (const std::exception& e) { g_err = e.what(); return 1; }
}

Base* B(void* h) { return static_cast<Base*>(h); }

template <typename Fn>
int dispatch(void* h, Fn&& fn) {
  if (!h) { g_err = "null handle"; return 1; }
  return guarded([&] {
    if (B(h)->precision() == 8) fn(*static_cast<State<double>*>(B(h)));
    else fn(*static_cast<State<float>*>(B(h)));
  });
}

template <typename C>
void put_complex(const std::vector<C>& v, double* out) {
  for (size_t q = 0; q < v.size(); q++) { out[2 * q] = v[q].real(); out[2 * q + 1] = v[q].imag(); }
}

template <typename Fn>
int sdispatch(void* h, Fn&& fn) {
  if (!h) { g_err = "null handle"; return 1; }
  return guarded([&] {
    if (SB(h)->precision() == 8) fn(*static_cast<Spatial<double>*>(SB(h)));
    else fn(*static_cast<Spatial<float>*>(SB(h)));
  });
}

}  // namespace

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n < 1) { g_err = "thread count must be >= 1"; return 1; }
  omp_set_num_threads(n);
#else
  (void)n;
#endif
  return 0;
}

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void* orc_create(const orc_config* cfg, const double* E0, int precision) {
  Base* b = nullptr;
  int rc = guarded([&] {
    if (precision == 8) { auto* s = new State<double>(); b = s; s->build(*cfg, E0); }
    else if (precision == 4) { auto* s = new State<float>(); b = s; s->build(*cfg, E0); }
    else throw std::runtime_error("precision must be 4 or 8");
  });
  if (rc) { delete b; return nullptr; }
  return b;
}

void orc_destroy(void* h) { delete B(h); }

int orc_get_grid(void* h, double* r, double* dr, double* k, double* t, double* w) {
  return dispatch(h, [&](auto& s) {
    std::copy(s.r.begin(), s.r.end(), r); std::copy(s.dr.begin(), s.dr.end(), dr);
    std::copy(s.k.begin(), s.k.end(), k); std::copy(s.t.begin(), s.t.end(), t);
    std::copy(s.w.begin(), s.w.end(), w);
  });
}

int orc_get_grid_xyt(void* h, double* xo, double* yo, double* kxo, double* kyo) {
  return dispatch(h, [&](auto& s) {
    if (!s.xyt) throw std::runtime_error("not an XYT oracle");
    std::copy(s.x.begin(), s.x.end(), xo); std::copy(s.y.begin(), s.y.end(), yo);
    std::copy(s.kx.begin(), s.kx.end(), kxo); std::copy(s.ky.begin(), s.ky.end(), kyo);
  });
}

int orc_get_scalars(void* h, double* o) {
  return dispatch(h, [&](auto& s) {
    o[0] = s.w0; o[1] = s.n0; o[2] = s.Eu; o[3] = s.vf; o[4] = s.k0; o[5] = s.phik; o[6] = s.phip; o[7] = s.dt;
    o[8] = s.med.derivative_k(s.w0, 2);   // k2 (Media.jl:48-53)
  });
}

int orc_get_KZ(void* h, double* KZ) { return dispatch(h, [&](auto& s) { put_complex(s.KZ, KZ); }); }
int orc_get_QZ(void* h, double* QZ) { return dispatch(h, [&](auto& s) { put_complex(s.QZ, QZ); }); }

int orc_get_guards(void* h, double* Rg, double* Tg, double* Kg, double* Wg) {
  return dispatch(h, [&](auto& s) {
    std::copy(s.Rg.begin(), s.Rg.end(), Rg); std::copy(s.Tg.begin(), s.Tg.end(), Tg);
    std::copy(s.Kg.begin(), s.Kg.end(), Kg); std::copy(s.Wg.begin(), s.Wg.end(), Wg);
  });
}

int orc_get_Rnl(void* h, int which, double* Rnl) {
  return dispatch(h, [&](auto& s) {
    if (which < 0 || which > 3) throw std::runtime_error("bad response index");
    put_complex(s.Rnl[which], Rnl);
  });
}

int orc_get_dht(void* h, double* Tm, double* RJ, double* JV) {
  return dispatch(h, [&](auto& s) {
    if (!s.rt) throw std::runtime_error("no DHT in T geometry");
    std::copy(s.Tm.begin(), s.Tm.end(), Tm); std::copy(s.RJ.begin(), s.RJ.end(), RJ);
    std::copy(s.JV.begin(), s.JV.end(), JV);
  });
}

int orc_set_field(void* h, const double* E) {
  return dispatch(h, [&](auto& s) {
    using Cx = typename std::decay_t<decltype(s)>::C;
    using R = typename Cx::value_type;
    for (size_t q = 0; q < s.E.size(); q++) s.E[q] = Cx((R)E[2 * q], (R)E[2 * q + 1]);
  });
}

int orc_get_field(void* h, double* E) { return dispatch(h, [&](auto& s) { put_complex(s.E, E); }); }

int orc_get_plasma(void* h, double* rho, double* kdrho) {
  return dispatch(h, [&](auto& s) {
    if (rho) std::copy(s.rho.begin(), s.rho.end(), rho);
    if (kdrho) std::copy(s.kdrho.begin(), s.kdrho.end(), kdrho);
  });
}

int orc_zstep(void* h, double z, double dz) {
  return dispatch(h, [&](auto& s) {
    using R = typename std::decay_t<decltype(s)>::C::value_type;
    s.zstep((R)z, (R)dz);
  });
}

int orc_plasma_solve(void* h) {
  return dispatch(h, [&](auto& s) { if (s.cfg.plasma) s.plasma_solve(); });
}

int orc_analyze(void* h, double z, orc_diag* d, double* Fr, double* Ft, double* rho_end, double* S) {
  return dispatch(h, [&](auto& s) { s.analyze(z, d, Fr, Ft, rho_end, S); });
}

// AStepKerrPlasma (AdaptiveSteps.jl:96-107) + Filament.jl:132-143
int orc_step_and_analyze(void* h, double* z, double dz0, double phimax, orc_diag* d, double* dz_out) {
  return dispatch(h, [&](auto& s) {
    using R = typename std::decay_t<decltype(s)>::C::value_type;
    double dz = dz0;
    if (s.cfg.nonlinearity) {
      double dzk = phimax / (s.phik * d->Imax);
      double dzp = phimax / (s.phip * d->rhomax);
      dz = std::min(dz0, std::min(dzk, dzp));
    }
    dz = (double)(R)dz;
    *z = (double)((R)*z + (R)dz);
    s.zstep((R)*z, (R)dz);
    s.analyze(*z, d, nullptr, nullptr, nullptr, nullptr);
    *dz_out = dz;
  });
}

int orc_bessel_j0_zeros(int n, double* out) { return guarded([&] { bessel_j0_zeros(n, out); }); }

int orc_fft(int n, double* x, int inverse_normalized) {
  return guarded([&] {
    FFTPlan<double> p(n);
    p.run(reinterpret_cast<std::complex<double>*>(x), inverse_normalized ? +1 : -1);
  });
}

int orc_dht_apply(void* h, double* f, int inverse) {
  return dispatch(h, [&](auto& s) {
    using Cx = typename std::decay_t<decltype(s)>::C;
    using R = typename Cx::value_type;
    if (!s.rt) throw std::runtime_error("no DHT in T geometry");
    int N = s.Nr;
    std::vector<Cx> a(N), b(N, Cx(0));
    for (int n = 0; n < N; n++) a[n] = Cx((R)f[2 * n], (R)f[2 * n + 1]) * (inverse ? (R)1 / s.JV[n] : s.RJ[n]);
    for (int m = 0; m < N; m++) for (int n = 0; n < N; n++) b[m] += s.Tm[(size_t)m * N + n] * a[n];
    for (int m = 0; m < N; m++) {
      Cx v = b[m] * (inverse ? (R)1 / s.RJ[m] : s.JV[m]);
      f[2 * m] = v.real(); f[2 * m + 1] = v.imag();
    }
  });
}

double orc_tf_eval(void* h, int comp, double I, int power) {
  double out = NAN;
  dispatch(h, [&](auto& s) {
    if (comp < 0 || comp >= (int)s.tabs.size()) throw std::runtime_error("bad component");
    out = s.tabs[comp].eval(I, power != 0);
  });
  return out;
}

int orc_linear_step(void* h, double dz) {
  return dispatch(h, [&](auto& s) {
    using R = typename std::decay_t<decltype(s)>::C::value_type;
    s.linear_step((R)dz);
  });
}

int orc_set_plasma(void* h, const double* rho, const double* kdrho) {
  return dispatch(h, [&](auto& s) {
    using R = typename std::decay_t<decltype(s)>::C::value_type;
    for (size_t q = 0; q < s.rho.size(); q++) {
      if (rho) s.rho[q] = (R)rho[q];
      if (kdrho) s.kdrho[q] = (R)kdrho[q];
    }
  });
}

int orc_nl_step(void* h, double* Es, double dz) {
  return dispatch(h, [&](auto& s) {
    using Cx = typename std::decay_t<decltype(s)>::C;
    using R = typename Cx::value_type;
    const int Nt = s.Nt;
    if (!s.qpar) {
      std::vector<Cx> all((size_t)s.Nr * Nt);
      for (size_t q = 0; q < all.size(); q++) all[q] = Cx((R)Es[2 * q], (R)Es[2 * q + 1]);
      s.nl_step_all(all.data(), (R)dz);
      for (size_t q = 0; q < all.size(); q++) { Es[2 * q] = all[q].real(); Es[2 * q + 1] = all[q].imag(); }
      return;
    }
    std::vector<Cx> row(Nt), scratch((size_t)6 * Nt);
    for (int i = 0; i < s.Nr; i++) {
      double* p = Es + 2 * (size_t)i * Nt;
      for (int j = 0; j < Nt; j++) row[j] = Cx((R)p[2 * j], (R)p[2 * j + 1]);
      s.nl_step_row(i, row.data(), (R)dz, scratch);
      for (int j = 0; j < Nt; j++) { p[2 * j] = row[j].real(); p[2 * j + 1] = row[j].imag(); }
    }
  });
}

int orc_nl_rhs(void* h, const double* Es, double* dE) {
  return dispatch(h, [&](auto& s) {
    using Cx = typename std::decay_t<decltype(s)>::C;
    using R = typename Cx::value_type;
    int Nt = s.Nt;
    std::vector<Cx> in(Nt), Et(Nt), F(Nt), out((size_t)s.Nr * Nt);
    for (int i = 0; i < s.Nr; i++) {
      for (int j = 0; j < Nt; j++) in[j] = Cx((R)Es[2 * ((size_t)i * Nt + j)], (R)Es[2 * ((size_t)i * Nt + j) + 1]);
      s.rhs_row(i, in.data(), &out[(size_t)i * Nt], Et.data(), F.data());
    }
    if (!s.qpar) s.qz_space(out.data());
    for (size_t q = 0; q < out.size(); q++) { dE[2 * q] = out[q].real(); dE[2 * q + 1] = out[q].imag(); }
  });
}


void* orc_spatial_create(const orc_spatial_config* cfg, const double* E0, int precision) {
  SpatialBase* b = nullptr;
  int rc = guarded([&] {
    if (precision == 8) { auto* s = new Spatial<double>(); b = s; s->build(*cfg, E0); }
    else if (precision == 4) { auto* s = new Spatial<float>(); b = s; s->build(*cfg, E0); }
    else throw std::runtime_error("precision must be 4 or 8");
  });
  if (rc) { delete b; return nullptr; }
  return b;
}

void orc_spatial_destroy(void* h) { delete SB(h); }

int orc_spatial_set_field(void* h, const double* E) {
  return sdispatch(h, [&](auto& s) {
    using C = typename std::decay_t<decltype(s)>::C;
    using T = typename C::value_type;
    for (int q = 0; q < s.n; q++) s.E[q] = C((T)E[2 * q], (T)E[2 * q + 1]);
  });
}

int orc_spatial_get_field(void* h, double* E) { return sdispatch(h, [&](auto& s) { put_complex(s.E, E); }); }

int orc_spatial_get_grid(void* h, double* a, double* b, double* c, double* d) {
  return sdispatch(h, [&](auto& s) {
    auto put = [](const std::vector<double>& v, double* o) { if (o) std::copy(v.begin(), v.end(), o); };
    if (!s.xy) { put(s.r, a); put(s.dr, b); put(s.k, c); }
    else { put(s.x, a); put(s.y, b); put(s.kx, c); put(s.ky, d); }
  });
}

int orc_spatial_get_tables(void* h, double* KZ, double* QZ, double* Kg, double* Fg, double* s7) {
  return sdispatch(h, [&](auto& s) {
    if (KZ) put_complex(s.KZ, KZ);
    if (QZ) put_complex(s.QZ, QZ);
    for (int q = 0; q < s.n; q++) {
      if (Kg) Kg[q] = s.Kg[q];
      if (Fg) Fg[q] = s.Fg[q];
    }
    if (s7) {
      s7[0] = s.w0; s7[1] = s.n0; s7[2] = s.Eu; s7[3] = s.vf; s7[4] = s.k0; s7[5] = s.phik; s7[6] = s.Rnl;
    }
  });
}

int orc_spatial_zstep(void* h, double dz) { return sdispatch(h, [&](auto& s) { s.zstep(dz); }); }

int orc_spatial_nl_step(void* h, double dz) {
  return sdispatch(h, [&](auto& s) {
    using T = typename std::decay_t<decltype(s)>::C::value_type;
    s.nl_step((T)dz);
  });
}

int orc_spatial_linear_step(void* h, double dz) { return sdispatch(h, [&](auto& s) { s.linear_step(dz); }); }

int orc_spatial_analyze(void* h, double* out6) { return sdispatch(h, [&](auto& s) { s.analyze(out6); }); }

}  // extern "C"
