/*
 * filament_oracle.h -- CPU restatement of Filament.jl's per-z-step propagation
 * path (RT, XYT and T geometries), in FP64 and FP32.
 *
 * THIS IS TEST INFRASTRUCTURE.  It is the checker that the HIP library is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.  The product (libfilament_hip.so) never calls it.
 *
 * Each function in filament_oracle.cpp cites the reference file:line it
 * restates (paths relative to the Filament.jl source tree).
 *
 * Layout: fields are row-major [rows][Nt] (t fastest), complex interleaved;
 * the rows are the r nodes (RT), the (x, y) points with x fastest (XYT) or a
 * single row (T).
 */
#ifndef FILAMENT_ORACLE_H
#define FILAMENT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXCOMP 4

enum { ORC_GEOM_RT = 0, ORC_GEOM_T = 1, ORC_GEOM_XYT = 2 };
enum { ORC_MEDIUM_AIR_PECK_REEDER = 0, ORC_MEDIUM_CONSTANT = 1 };
enum { ORC_OFF = 0, ORC_THG = 1, ORC_NOTHG = 2 };

typedef struct orc_config {
  int32_t geometry;       /* ORC_GEOM_RT or ORC_GEOM_T                       */
  int32_t Nr, Nt;
  double rmax, tmin, tmax;          /* grid units (ru, tu)                   */
  double ru, zu, tu, Iu, rhou;      /* unit set (Units.jl UnitRT)            */
  double lam0;                      /* [m]                                   */
  int32_t medium;                   /* ORC_MEDIUM_*                          */
  double n_const;                   /* refractive index for constant medium  */
  double n2;                        /* [m^2/W]                               */
  int32_t nonlinearity, plasma, kparaxial, qparaxial;
  double rguard, tguard, kguard, wguard;
  int32_t cubic;                    /* ORC_OFF / ORC_THG / ORC_NOTHG         */
  int32_t raman;                    /* ORC_OFF / ORC_THG / ORC_NOTHG         */
  int32_t current_free;             /* 0/1 */
  double nuc, mr;                   /* medium collision freq., reduced mass  */
  int32_t current_losses;           /* 0/1 */
  int32_t losses_ereal;             /* fearg = Re(E)^2 (1) or |E|^2 (0)      */
  int32_t pe_ereal, pe_kdep;        /* plasma equation switches              */
  double pe_rho0, pe_rho_nt, pe_nuc, pe_mr;
  int32_t ncomp;
  double comp_frac[ORC_MAXCOMP];
  double comp_Ui[ORC_MAXCOMP];      /* [eV] */
  int32_t comp_ntab[ORC_MAXCOMP];
  const double* comp_tab_x[ORC_MAXCOMP];   /* raw .tf columns: x [W/m^2]    */
  const double* comp_tab_y[ORC_MAXCOMP];   /*                  y [1/s]      */
  /* XYT (Input.jl:151-178): grid in xu, yu units; guards; the reference's
   * 3-D filter indexes its third factor with the y index (Guards.jl:314),
   * reproduced when guard3d_bug != 0 */
  int32_t Nx, Ny;
  double xmin, xmax, ymin, ymax, xu, yu;
  double xguard, yguard, kxguard, kyguard;
  int32_t guard3d_bug;
} orc_config;

typedef struct orc_diag {
  double z, Fmax, Imax, rhomax, De, rfil, rpl, tau, W;
} orc_diag;

/* precision: 8 = FP64, 4 = FP32.  E0 is the real initial field [Nr][Nt]
 * (the user initial_condition evaluated on the grid, Fields.jl:102). */
void* orc_create(const orc_config* cfg, const double* E0, int precision);
void  orc_destroy(void* h);
const char* orc_last_error(void);
/* OpenMP threads of later calls (the CPU baseline's 1-core figure) */
int orc_set_threads(int n);
int orc_max_threads(void);

/* tables (always returned in FP64, after rounding to the working precision) */
int orc_get_grid(void* h, double* r, double* dr, double* k, double* t, double* w);
int orc_get_grid_xyt(void* h, double* x, double* y, double* kx, double* ky);
int orc_get_scalars(void* h, double* out9);   /* w0, n0, Eu, vf, k0, phik, phip, dt, k2 */
int orc_get_KZ(void* h, double* KZ);          /* [Nr][Nt] complex */
int orc_get_QZ(void* h, double* QZ);          /* [Nr][Nt] complex */
int orc_get_guards(void* h, double* Rg, double* Tg, double* Kg, double* Wg);
int orc_get_Rnl(void* h, int which, double* Rnl);   /* which: 0 cubic,1 raman,2 free,3 losses; [Nt] complex */
int orc_get_dht(void* h, double* T, double* RJ, double* JV);

/* state */
int orc_set_field(void* h, const double* E);  /* complex [Nr][Nt], time domain */
int orc_get_field(void* h, double* E);
int orc_get_plasma(void* h, double* rho, double* kdrho);

/* the hot path */
int orc_zstep(void* h, double z, double dz);            /* Models.jl:75-121 */
int orc_plasma_solve(void* h);                          /* plasma_equations.jl */
/* RT (FieldAnalyzers.jl:166-215): Fr [Nr]; XYT (:308-353): Fr = Fxy [rows],
 * rfil = ax, rpl = ay, tau = 0. */
int orc_analyze(void* h, double z, orc_diag* d, double* Fr, double* Ft,
                double* rho_end, double* S);
/* One reference main-loop iteration without I/O: dz from AdaptiveSteps,
 * zstep, analyze (Filament.jl:132-143).  Returns dz used via *dz_out. */
int orc_step_and_analyze(void* h, double* z, double dz0, double phimax,
                         orc_diag* d, double* dz_out);

/* component checks */
int orc_bessel_j0_zeros(int n, double* out);
int orc_fft(int n, double* x, int inverse_normalized);   /* complex, in place */
int orc_dht_apply(void* h, double* f, int inverse);      /* one column [Nr] complex */
double orc_tf_eval(void* h, int comp, double I, int power);
int orc_linear_step(void* h, double dz);                 /* spectral.jl:16-26 on spectral E */
int orc_nl_rhs(void* h, const double* Es, double* dE);   /* nonlinear_propagators.jl:38-67 */
/* one RK3 step of the nonlinear propagator on spectral rows Es (in place),
 * with the plasma arrays as they stand (nonlinear_propagators.jl:70-74) */
int orc_nl_step(void* h, double* Es, double dz);
/* replaces rho / kdrho [rows][Nt] (either may be NULL) */
int orc_set_plasma(void* h, const double* rho, const double* kdrho);

/* ---------------------------------------------------------------------------
 * The spatial geometries R and XY (Input.jl:37-56, 131-150): a monochromatic
 * field E(r) [Nr] or E(x, y) [Nx x Ny] (x fastest), complex, no plasma; one
 * cubic response (calc_cubic_nothg_spatial, cubic.jl:49-54) with the medium's
 * n2.  The transforms follow the reference's plans literally: the Hankel
 * pair JV .* (T * (RJ .* f)) and (T * (F ./ JV)) ./ RJ, the 2-D FFT as 1-D
 * transforms along x, then y.
 * ------------------------------------------------------------------------- */
enum { ORC_GEOM_R = 3, ORC_GEOM_XY = 4 };

typedef struct orc_spatial_config {
  int32_t geometry;                 /* ORC_GEOM_R or ORC_GEOM_XY */
  int32_t Nr;                       /* R */
  double rmax, ru;                  /* R: grid units */
  int32_t Nx, Ny;                   /* XY: powers of two */
  double xmin, xmax, ymin, ymax, xu, yu;
  double zu, Iu, lam0;
  int32_t medium;                   /* ORC_MEDIUM_* */
  double n_const, n2;
  int32_t nonlinearity, kparaxial, qparaxial;
  int32_t alg;                      /* 3: RK3 (Kutta), 4: RK4 (classic) */
  double rguard, kguard;            /* R */
  double xguard, yguard, kxguard, kyguard;   /* XY */
} orc_spatial_config;

/* E0: complex [n], the initial_condition on the grid (Fields.jl:31-51, 123-138) */
void* orc_spatial_create(const orc_spatial_config* cfg, const double* E0, int precision);
void  orc_spatial_destroy(void* h);
int orc_spatial_set_field(void* h, const double* E);
int orc_spatial_get_field(void* h, double* E);
/* R: a = r, b = dr, c = k (grid units); XY: a = x, b = y, c = kx, d = ky */
int orc_spatial_get_grid(void* h, double* a, double* b, double* c, double* d);
/* KZ, QZ complex [n]; Kg, Fg [n] (spectral and field guards, XY products of
 * the x and y factors); s7 = w0, n0, Eu, vf, k0, phik, Rnl */
int orc_spatial_get_tables(void* h, double* KZ, double* QZ, double* Kg, double* Fg, double* s7);
int orc_spatial_zstep(void* h, double dz);          /* Models.jl:75-121 */
int orc_spatial_nl_step(void* h, double dz);        /* nonlinear_propagators.jl:70-74 alone */
int orc_spatial_linear_step(void* h, double dz);    /* spectral.jl:16-26 alone */
/* FieldAnalyzerR / XY: out6 = Imax, rfil, ax, ay, P, index of the maximum */
int orc_spatial_analyze(void* h, double* out6);

#ifdef __cplusplus
}
#endif
#endif
