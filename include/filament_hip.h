/*
 * filament_hip.h -- C ABI of libfilament_hip.so, the MI355X (gfx950)
 * implementation of Filament.jl's per-z-step propagation path for the
 * axisymmetric RT and the three-dimensional XYT geometries, the T geometry,
 * and the spatial R and XY geometries (the last section).
 *
 * The library replaces the device side of one reference call,
 *   Models.zstep(z, dz, grid, field, guard, model)   (src/modules/Models.jl:75-121)
 * together with the pieces it reaches: the plasma solve (plasma_equations.jl:27-72),
 * the t-transforms (Models.jl:152-171), the nonlinear RK step
 * (nonlinear_propagators.jl:38-74,246-283), the spectral linear step with the
 * Hankel transform (RT) or the 2-D FFT (XYT) (linear_propagators/spectral.jl:16-26,
 * Models.jl:135-148), the guard filters (Guards.jl:199-318) and the RT and XYT
 * field analyzers (FieldAnalyzers.jl:166-215, 308-353).  The host keeps Input.prepare, Units, Grids,
 * Media and the table construction; it hands the tables over once.
 *
 * Conventions
 *  - Every function returns 0 on success and a nonzero FIL_E* code on
 *    failure; fil_last_error() then gives a thread-local message.
 *  - 2-D host arrays use the reference's (Julia) column-major layout: element
 *    (i, j) of an Nr x Nt array is at [i + ld*j], r (or k) index fastest.
 *  - Complex host arrays are interleaved (re, im) doubles.
 *  - The library owns all device memory; callers never see device pointers.
 *  - A handle is not thread-safe; calls on one handle are issued in order on
 *    the handle's own HIP stream.  fil_zstep is asynchronous; fil_analyze,
 *    fil_download_* and fil_synchronize block.
 */
#ifndef FILAMENT_HIP_H
#define FILAMENT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FIL_ABI_VERSION 5
#define FIL_MAXCOMP 4

typedef struct fil_handle fil_handle;

enum {
  FIL_OK = 0,
  FIL_EINVAL = 1,    /* bad argument or unsupported configuration */
  FIL_EDEVICE = 2,   /* HIP runtime failure */
  FIL_ESTATE = 3,    /* call out of order (e.g. zstep before tables) */
  FIL_ENOMEM = 4
};

enum { FIL_GEOM_RT = 0, FIL_GEOM_XYT = 1, FIL_GEOM_T = 2 };

enum { FIL_FP32 = 4, FIL_FP64 = 8 };

/* Forms (bit flags; 0 = the plain port) */
enum {
  FIL_FORM_DHT_BF16X3 = 1,  /* Hankel transform as split-BF16 3-term MFMA products, FP32 accumulate */
  FIL_FORM_SCAN_FP64 = 2    /* plasma-density t-integration carried in FP64 */
};

/* Nonlinear responses (src/modules/medium_responses), bit flags */
enum {
  FIL_RESP_CUBIC_THG = 1,     /* calc_cubic          cubic.jl:33-38  */
  FIL_RESP_CUBIC_NOTHG = 2,   /* calc_cubic_nothg    cubic.jl:41-46  */
  FIL_RESP_FREE = 4,          /* calc_current_free   current_free.jl:31-37 */
  FIL_RESP_LOSSES = 8,        /* calc_current_losses current_losses.jl:36-83 */
  FIL_RESP_RAMAN_THG = 16,    /* calc_raman          raman.jl:46-54  */
  FIL_RESP_RAMAN_NOTHG = 32   /* calc_raman_nothg    raman.jl:57-65  */
};

/* Problem description (replaces the type parameters of Field/Model built by
 * Filament.run, src/Filament.jl:69-75). */
typedef struct fil_desc {
  int32_t abi_version;     /* FIL_ABI_VERSION */
  int32_t geometry;        /* FIL_GEOM_RT, FIL_GEOM_XYT or FIL_GEOM_T (Nr ignored: one row) */
  int32_t Nr, Nt;          /* Nt: power of two, 256..8192 (FP64: ..4096) */
  int32_t precision;       /* FIL_FP32 (the reference's FloatGPU) or FIL_FP64 */
  int32_t form;            /* FIL_FORM_* flags */
  int32_t nonlinearity;    /* NONLINEARITY */
  int32_t plasma;          /* PLASMA */
  int32_t responses;       /* FIL_RESP_* flags (ignored unless nonlinearity) */
  int32_t losses_ereal;    /* current_losses "EREAL" */
  int32_t plasma_ereal;    /* plasma_equation "EREAL" */
  int32_t plasma_kdep;     /* plasma_equation "KDEP" */
  int32_t ncomp;           /* plasma components, 1..FIL_MAXCOMP */
  int32_t keep_plasma;     /* store full rho/kdrho each step (for fil_download_plasma) */
  int32_t device;          /* HIP device ordinal */
  /* XYT only (ignored for RT): grid sizes, powers of two in [16, 1024]; the
   * rows of the field are the Nx*Ny points (x fastest), Nr is ignored. */
  int32_t Nx, Ny;
  /* XYT only: reproduce the reference's 3-D filter kernel, which indexes the
   * t (and w) factor with the y index (Guards.jl:314); 0 = the intended
   * filter X(x) Y(y) T(t) and KX(kx) KY(ky) W(w).  Needs Ny <= Nt. */
  int32_t guard3d_bug;
  /* RT only: number of pulses B (0 or 1 = one) propagated together on one
   * grid with one dz.  Replaces B independent runs of Filament.run
   * (Filament.jl:69-106, main_loop :111-166) on the same Grid/Medium; each pulse has its own
   * field and plasma, the tables are shared.  Batched handles take the
   * *_member entry points and no corrector. */
  int32_t batch;
  /* RT and XYT: QPARAXIAL = false, the non-paraxial prefactor of the
   * nonlinear term (nonlinear_propagators.jl:58-63, 232-242).  RT needs
   * fil_tables.QZ_kw and the FP32/FP64 Hankel products (not
   * FIL_FORM_DHT_BF16X3); XYT evaluates QZ(kx, ky, w) / QZ(w) on the device
   * from fil_tables_xyt.  0: paraxial. */
  int32_t nonparaxial;
} fil_desc;

/* Host-precomputed tables, FP64 (the library rounds them to the working
 * precision).  The library copies everything; the caller keeps ownership. */
typedef struct fil_tables {
  /* Grids (Grids.jl:97-121,204-223) */
  const double* r;         /* [Nr] r nodes (grid units) */
  const double* dr;        /* [Nr] */
  const double* t;         /* [Nt] t nodes (grid units) */
  double dt;               /* t step (grid units) */
  /* Linear propagator (spectral.jl:78-101) and spectral guard (Guards.jl:108-119) */
  const double* KZ;        /* [Nr x Nt] complex, column-major */
  const double* Kguard;    /* [Nr x Nt] */
  const double* Wguard;    /* [Nt] */
  /* Field guard (Guards.jl:105-106) */
  const double* Rguard;    /* [Nr] */
  const double* Tguard;    /* [Nt] */
  /* Hankel plan (HankelTransforms.plan_dht, Fields.jl:108-110):
   * forward f -> JV .* (T * (RJ .* f)), inverse F -> (T * (F ./ JV)) ./ RJ */
  const double* dht_T;     /* [Nr x Nr] */
  const double* dht_RJ;    /* [Nr] */
  const double* dht_JV;    /* [Nr] */
  /* Nonlinear propagator (nonlinear_propagators.jl:139-160, 210-230) and responses */
  const double* QZ;        /* [Nt] complex: QPARAXIAL prefactor (independent of r) */
  const double* Rnl_cubic; /* [Nt] complex (cubic.jl:11-12, broadcast) */
  const double* Rnl_free;  /* [Nt] complex (current_free.jl:13-23) */
  const double* Rnl_losses;/* [Nt] complex (current_losses.jl:11-20) */
  const double* Rnl_raman; /* [Nt] complex (raman.jl:9-12, broadcast) */
  const double* Hraman;    /* [Nt] complex (raman.jl:15-29) */
  /* Plasma equation (photoionization_avalanche.jl:1-116), per component:
   * dimensionless frhont, Rava, K; rate table as log10 pairs of the scaled
   * table (x/Iu, y*tu), which must be uniformly spaced in log10(x). */
  double rho0;             /* initial density per component [rhou] */
  double frhont[FIL_MAXCOMP];
  double Rava[FIL_MAXCOMP];
  double K[FIL_MAXCOMP];
  int32_t ntab[FIL_MAXCOMP];
  const double* tab_log10x[FIL_MAXCOMP];
  const double* tab_log10y[FIL_MAXCOMP];
  /* RT with fil_desc.nonparaxial: QZ(k, w) = mu0 mu w^2 / (2 kz) zu / Eu,
   * kz = sqrt(beta^2 - k^2 + 0im) (nonlinear_propagators.jl:139-160, 232-242),
   * [Nr x Nt] complex, column-major; NULL otherwise. */
  const double* QZ_kw;
} fil_tables;

/* XYT tables (Grids.jl:158-201, Guards.jl:162-195, spectral.jl:131-184), FP64.
 * fil_set_tables_xyt takes these together with the fil_tables fields that the
 * two geometries share: t, dt, Wguard, Tguard, QZ, Rnl_*, the plasma tables
 * (r, dr, KZ, Kguard, Rguard and dht_* are not used).  KZ is evaluated on the
 * device from its separable factors:
 *   KZ = (Kfunc(beta(w), kt) - w/vf) zu,  kt^2 = kt2x[ix] + kt2y[iy],
 *   Kfunc = beta - kt^2/(2 beta) (kparaxial) or sqrt(beta^2 - kt^2 + 0im).
 * When kt2x and KXguard (kt2y and KYguard) are even, i.e. equal at ix and
 * Nx - ix bit for bit as on the FFT frequency grid, the device stores the
 * propagator for ix <= Nx/2 (iy <= Ny/2) only; any tables are accepted. */
typedef struct fil_tables_xyt {
  const double* x;         /* [Nx] grid units */
  const double* y;         /* [Ny] */
  double dx, dy;           /* grid units */
  const double* kt2x;      /* [Nx] (kx * kxu)^2 [1/m^2] */
  const double* kt2y;      /* [Ny] (ky * kyu)^2 */
  const double* Xguard;    /* [Nx] field guard windows */
  const double* Yguard;    /* [Ny] */
  const double* KXguard;   /* [Nx] spectral guards (all ones in the reference, Guards.jl:183-184) */
  const double* KYguard;   /* [Ny] */
  const double* beta;      /* [Nt] complex beta(w) [1/m] at the grid frequencies */
  const double* w_vf;      /* [Nt] w / vf [1/m] (moving frame, spectral.jl:140-150) */
  double zu;               /* [m] */
  int32_t kparaxial;       /* KPARAXIAL */
} fil_tables_xyt;

/* Scalar diagnostics of FieldAnalyzerRT (FieldAnalyzers.jl:125-138) */
typedef struct fil_diag {
  double z, Fmax, Imax, rhomax, De, rfil, rpl, tau, W;
} fil_diag;

const char* fil_last_error(void);

/* Replaces Fields.Field / Models.Model construction (Filament.jl:71-74). */
int fil_create(const fil_desc* d, fil_handle** h);
int fil_set_tables(fil_handle* h, const fil_tables* t);
/* XYT handles (Field/Guard/LinearPropagator XYT, Fields.jl:141-160). */
int fil_set_tables_xyt(fil_handle* h, const fil_tables* t, const fil_tables_xyt* x);

/* Field in the time domain, column-major E[i + ld*j] (Field.E, Fields.jl:102-115
 * after rsig2asig!).  Upload replaces the state; download returns what the
 * reference holds in field.E after the last zstep (guarded, time domain). */
int fil_upload_field(fil_handle* h, const double* E, int64_t ld);
int fil_download_field(fil_handle* h, double* E, int64_t ld);
/* The same for pulse `member` of a batched handle (0 <= member < batch).  Every
 * member is uploaded before the first zstep. */
int fil_upload_member(fil_handle* h, int member, const double* E, int64_t ld);
int fil_download_member(fil_handle* h, int member, double* E, int64_t ld);

/* Models.zstep (Models.jl:75-121).  Asynchronous. */
int fil_zstep(fil_handle* h, double z, double dz);

/* No counterpart in the reference, whose zstep advances the nonlinear and the
 * linear part by the same dz (Models.jl:97-107): the same z-step with the
 * nonlinear step over dz_nonlinear (>= 0) and the linear step (propagator,
 * spectral filter) over dz_linear; the field filter stays pending as after
 * fil_zstep.  dz_nonlinear = 0 skips the nonlinear step and the plasma solve
 * (a pending field filter or corrector output is still applied first).  For
 * callers that compose symmetric (Strang) sequences of the two parts, e.g.
 * L(h/2) [N(h) L(h)]^(n-1) N(h) L(h/2).  fil_zstep_parts(h, z, dz, dz) is
 * fil_zstep(h, z, dz).  Handles with the non-paraxial term take equal
 * distances only.  Asynchronous. */
int fil_zstep_parts(fil_handle* h, double z, double dz_nonlinear, double dz_linear);

/* Several handles (e.g. the per-grid batches of an ensemble of pulses of
 * uneven size) advanced by one z-step each with ONE graph launch: the first
 * call for a (dz, pending-filter pattern) records the handles' zstep kernels
 * as a HIP graph with one branch per handle; later calls replay it.  Replaces
 * the loop over independent runs' Models.zstep calls (Models.jl:75-121,
 * Filament.jl:136).  Ordered after each handle's earlier work and before its
 * later work; asynchronous.  Destroy the group before its handles: while a
 * group lives, fil_set_tables on its handles is refused (its graphs hold their
 * device pointers), and once one of its handles is destroyed the group refuses
 * to step. */
#define FIL_GROUP_MAX 64
typedef struct fil_group fil_group;
int fil_group_create(fil_handle* const* hs, int n, fil_group** g);
int fil_group_zstep(fil_group* g, double z, double dz);
int fil_group_stats(fil_group* g, int64_t* captures, int64_t* replays);
int fil_group_destroy(fil_group* g);

/* MFMA tile shape of a handle's Hankel-transform products (rows of the QDHT
 * matrix per workgroup).  No counterpart in the reference (CUBLAS chooses its
 * own); it is what the ensemble's autotuner (filament_amd.tuning) sets from its
 * tile table.  FIL_TILING_RULE (the default): 128-row tiles when they still give
 * >= 256 workgroups for one pulse, else 64; FIL_TILING_128 needs Nr (padded to
 * 64) divisible by 128.  Every shape sums each output over k in the same order,
 * so the choice changes the speed, never the result. */
#define FIL_TILING_RULE 0
#define FIL_TILING_128 1
#define FIL_TILING_64 2
int fil_set_tiling(fil_handle* h, int tiling);
int fil_get_tiling(fil_handle* h, int* tiling);

/* FieldAnalyzers.analyze! RT (FieldAnalyzers.jl:166-215), including its side
 * effect of projecting the field onto its analytic part.  Fr [Nr], Ft [Nt],
 * rho_end [Nr] and S [Nt/2+1] may be NULL.  Blocks.
 * T handles: FieldAnalyzerT (FieldAnalyzers.jl:99-114) in the same struct:
 * Imax, rhomax, tau = duration = radius(t, I), W = Fmax = F = sum(I) dt, the
 * other fields 0; Ft = I(t), Fr [1] = F. */
int fil_analyze(fil_handle* h, double z, fil_diag* d, double* Fr, double* Ft,
                double* rho_end, double* S);
/* fil_analyze for pulse `member` of a batched handle (the projection applies
 * to every member). */
int fil_analyze_member(fil_handle* h, int member, double z, fil_diag* d, double* Fr, double* Ft,
                       double* rho_end, double* S);

/* FieldAnalyzerXYT (FieldAnalyzers.jl:272-353): Fxy [Nx x Ny] (x fastest),
 * Ft [Nt], rho_end [Nx x Ny], S [Nt/2+1] may be NULL.  ax, ay: radii of Fxy
 * along x and y through its maximum.  Projects the field onto its analytic
 * part like the RT analyzer.  Blocks. */
typedef struct fil_diag_xyt {
  double z, Fmax, Imax, rhomax, De, ax, ay, W;
} fil_diag_xyt;
int fil_analyze_xyt(fil_handle* h, double z, fil_diag_xyt* d, double* Fxy, double* Ft,
                    double* rho_end, double* S);

/* The z loop of Filament.run (Filament.jl:129-164) with the adaptive step of
 * AStepKerrPlasma (AdaptiveSteps.jl:96-107; the Kerr-only rule is phip = 0)
 * kept on the device: per step, dz = min(dz0, phimax/(phik Imax),
 * phimax/(phip rhomax)) from the previous analysis (FP64, then rounded to
 * single precision), z += dz in single precision, zstep, then the analysis
 * (its projection included); the loop ends when z >= zmax, or after a step
 * whose Imax > Istop, or when an analysis gives a NaN Imax or rhomax, or
 * after max_steps steps.  The host checks the loop
 * every `chunk` steps (0: 32) instead of after every step.  Steps and their
 * analyses are the kernels fil_zstep and fil_analyze run, so the state after
 * fil_run_adaptive equals the host loop's (fil_analyze, then zstep(z + dz, dz)
 * with the same dz, ...) bit for bit.  *z: the start on entry, the end on
 * return.  log (may be NULL): [max_steps + 1][4] of (z, dz, Imax, rhomax),
 * row 0 the analysis before the first step.  *nsteps: steps taken; *reason:
 * FIL_STOP_*.  Single-pulse handles (any geometry and form); a pending
 * corrector output is added first.  Blocks. */
#define FIL_STOP_ZMAX 1
#define FIL_STOP_ISTOP 2
#define FIL_STOP_MAXSTEPS 3
#define FIL_STOP_NONFINITE 4   /* an analysis gave a NaN Imax or rhomax (the reference's z becomes NaN) */
typedef struct fil_adapt_params {
  double zmax, dz0, phimax, phik, phip, Istop;
  int nonlinearity;            /* 0: dz = dz0 (AStepKerrPlasma.NONLINEARITY) */
  int chunk;
} fil_adapt_params;
int fil_run_adaptive(fil_handle* h, const fil_adapt_params* p, double* z, int max_steps, double* log,
                     int* nsteps, int* reason);

/* fil_run_adaptive without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  The
 * loop's buffers (loop state, the log, and the copy of the state, rho_end and
 * rho_max at each chunk's start) live in one fixed device array of
 * FIL_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until the
 * process ends; a grid and max_steps whose buffers exceed it are refused
 * with FIL_EINVAL.  `log` is required ([max_steps + 1][4]).  Per-phase timing
 * is off during the call.  The array keeps the last call's contents and is
 * shared by every call on that GPU: calls that overlap on one GPU (from
 * several host threads) race.  Steps, log and state equal fil_run_adaptive's
 * bit for bit. */
#define FIL_NOALLOC_SCRATCH_BYTES (48u << 20)
int fil_run_adaptive_noalloc(fil_handle* h, const fil_adapt_params* p, double* z, int max_steps, double* log,
                             int* nsteps, int* reason);

/* field.rho and field.kdrho (Fields.jl:117-118) as the last plasma solve left
 * them, column-major [i + ld*j].  Blocks. */
int fil_download_plasma(fil_handle* h, double* rho, double* kdrho, int64_t ld);

int fil_synchronize(fil_handle* h);

/* Device-side copies of the z-step state (slots 0..FIL_NSLOTS-1), used to
 * run a coarse step and the fine steps it replaces from the same state.  No
 * counterpart in the reference (its driver never branches the state); the
 * host-side equivalent would be deepcopy(field). */
#define FIL_NSLOTS 4
int fil_state_save(fil_handle* h, int slot);
int fil_state_restore(fil_handle* h, int slot);
/* Copies the z-step state of src into dst (same device and grid), converting
 * between FP32 and FP64 handles; dst's previous state is replaced.  Waits for
 * src's stream, then runs on dst's stream. */
int fil_state_copy(fil_handle* dst, const fil_handle* src);

/* Fluence profile Fr(r) = dt * sum_t |E(r,t)|^2 (FieldAnalyzers.jl:171) of the
 * current state (slot < 0) or of a saved slot, computed on the device from
 * the spectrum by Parseval's identity, before the pending field filter.
 * Fr has Nr entries (batch x Nr for a batched handle, member-major).  Blocks. */
int fil_fluence(fil_handle* h, int slot, double* Fr);
int fil_destroy(fil_handle* h);

/* ---------------------------------------------------------------------------
 * Learned coarse-step corrector.  No counterpart in the reference: it serves
 * a driver that advances by coarse steps (a multiple of the reference's dz)
 * and corrects each coarse result towards what the fine steps would give.
 *
 * Features of a state (FP32 handles only): the field of rows [0, nr) in the
 * time window [t0, t0 + nt), as the envelope A = E exp(i w0 t), averaged over
 * pool_r x pool_t blocks into an H x W grid (H = nr/pool_r, W = nt/pool_t):
 *   X[0] = Re A / a_scale, X[1] = Im A / a_scale, X[2] = |E|^2 / a_scale^2,
 *   X[3] = rho / rho_scale  (plasma density of that field, zero without plasma)
 * X is [4][H][W] floats, W fastest.  The network (4 layers of 3x3
 * convolutions, 4 -> 32 -> 32 -> 32 -> 2 channels, ReLU between, BF16 weights
 * with FP32 accumulation) maps X to Y [2][H][W]; the correction added to the
 * field is a_scale * y_scale * (Y[0] + i Y[1]) exp(-i w0 t), bilinearly
 * interpolated from the block centres.  Training targets are
 * (X_fine[0:2] - X_coarse[0:2]) for the same step, in feature units.
 * Parameters: a flat FP32 vector of FIL_CORR_NPARAM values, layer by layer
 * w[cout][cin][3][3] then b[cout].
 * ------------------------------------------------------------------------- */
#define FIL_CORR_NPARAM 20258

typedef struct fil_corr_desc {
  int32_t nr;              /* window rows [0, nr) */
  int32_t t0, nt;          /* window t samples [t0, t0 + nt) */
  int32_t pool_r, pool_t;  /* block size; nr % pool_r == 0, nt % pool_t == 0 */
  int32_t capacity;        /* training pairs held on the device (ring) */
  int32_t batch_max;       /* largest training batch */
  double w0;               /* carrier frequency, radians per t unit */
  double tmin, dt;         /* t_j = tmin + j dt (t units) */
  double a_scale;          /* field scale of the features */
  double rho_scale;        /* density scale of the features */
} fil_corr_desc;

/* Creates (or re-creates) the handle's corrector with seeded He-initialised
 * weights and a zero last layer (the initial correction is zero). */
int fil_corrector_init(fil_handle* h, const fil_corr_desc* c, uint64_t seed);
/* Get (set = 0) or set (set != 0) the FP32 parameters; setting resets Adam. */
int fil_corrector_params(fil_handle* h, float* w, int set);
/* Get or set the target scale y_scale (0 = not yet trained). */
int fil_corrector_scale(fil_handle* h, double* y_scale, int set);
/* Features of the current state; X (may be NULL) receives [4][H][W].  Blocks. */
int fil_corrector_features(fil_handle* h, float* X);
/* Stores one training pair (X [4][H][W], target T [2][H][W]). */
int fil_corrector_store_pair(fil_handle* h, const float* X, const float* T);
/* Adam on the stored pairs: iters steps of random batches of `batch` pairs.
 * y_scale is set to the RMS of the stored targets if it is 0.  *loss: the
 * mean-square error (in y_scale units) of the last batch.  Blocks. */
int fil_corrector_train(fil_handle* h, int iters, int batch, double lr, uint64_t seed, double* loss);
/* Loss and (grad != NULL) its gradient over stored pairs [first, first+count). Blocks. */
int fil_corrector_eval(fil_handle* h, int first, int count, double* loss, float* grad);
/* Network output Y [2][H][W] for features X [4][H][W] from the host.  Blocks. */
int fil_corrector_forward(fil_handle* h, const float* X, float* Y);
/* Adds the network's correction to the current state.  Asynchronous.  The
 * network runs now; its output is added in time by the next fil_zstep's row
 * kernel (before the pending field filter), or first by any other call that
 * reads or replaces the state (download, analyze, fluence, state save/copy,
 * synchronize, corrector calls). */
int fil_corrector_apply(fil_handle* h);

/* Per-phase device timing with HIP events on the handle's stream (the
 * reference's TimerOutputs sections "plasma/nonlinearity" and "linear",
 * Models.jl:84-119).  names: "row" (plasma + t-FFTs + RK), "dht1", "dht2".
 * The events ride on the kernel dispatches.  enable: 0 off, 1 every z-step,
 * n > 1 every n-th z-step (timed dispatches are separated by gaps of several
 * microseconds, so sampling keeps the timing from slowing the stream).
 * fil_timing_read returns accumulated milliseconds and the number of timed
 * z-steps. */
int fil_timing_enable(fil_handle* h, int enable);
int fil_timing_read(fil_handle* h, double* ms_row, double* ms_dht1, double* ms_dht2, int64_t* count);

/* ---------------------------------------------------------------------------
 * The spatial geometries R and XY: a monochromatic beam without a time axis
 * and without plasma (Input.jl:37-56 R, :131-150 XY force PLASMA = false).
 * A spatial handle replaces Field/Guard/LinearPropagator/NonlinearPropagator R
 * (Fields.jl:31-51, Guards.jl:41-57, spectral.jl:30-50,
 * nonlinear_propagators.jl:78-96) and XY (Fields.jl:123-138, Guards.jl:132-157,
 * spectral.jl:104-128, nonlinear_propagators.jl:162-180) and runs
 * Models.zstep for them (Models.jl:75-121: no plasma, no t transforms):
 *   nonlinear RK step of func! (nonlinear_propagators.jl:38-74) with the
 *     spatial cubic response 3/4 |E|^2 E (cubic.jl:49-54), the field filter
 *     on the response and i QZ (paraxial) or the space transform, i QZ,
 *     spectral filter and inverse transform (non-paraxial);
 *   linear step: forward transform (R: Hankel, XY: 2-D FFT), exp(i KZ dz),
 *     spectral filter, inverse transform (spectral.jl:16-26);
 *   field filter (Guards.jl:199-215, kernel! :263-278).
 * Arrays are complex interleaved (re, im) doubles in the reference's layout:
 * R [Nr]; XY [Nx x Ny], x fastest (E[ix + Nx iy]).
 * ------------------------------------------------------------------------- */
enum { FIL_GEOM_R = 3, FIL_GEOM_XY = 4 };
/* ODEIntegrators algorithms of the nonlinear step (ALG in the input file) */
enum { FIL_ALG_RK3 = 3, FIL_ALG_RK4 = 4 };

typedef struct fil_spatial fil_spatial;

typedef struct fil_spatial_desc {
  int32_t abi_version;     /* FIL_ABI_VERSION */
  int32_t geometry;        /* FIL_GEOM_R or FIL_GEOM_XY */
  int32_t Nr;              /* R: points, 2..16384 */
  int32_t Nx, Ny;          /* XY: powers of two in [16, 4096] */
  int32_t precision;       /* FIL_FP32 (FloatGPU) or FIL_FP64 */
  int32_t nonlinearity;    /* NONLINEARITY: the cubic response */
  int32_t nonparaxial;     /* QPARAXIAL = false */
  int32_t alg;             /* FIL_ALG_RK3 (Kutta) or FIL_ALG_RK4 (classic) */
  int32_t device;          /* HIP device ordinal */
} fil_spatial_desc;

typedef struct fil_spatial_tables {
  /* KZ (spectral.jl:30-50 R, :104-128 XY; Kfunc :160-184): complex [n] */
  const double* KZ;
  /* spectral guards (Guards.jl:41-57, 132-157): R Kguard [Nr]; XY KXguard [Nx], KYguard [Ny] */
  const double* Kguard;
  const double* KXguard;
  const double* KYguard;
  /* field guards: R Rguard [Nr]; XY Xguard [Nx], Yguard [Ny] */
  const double* Rguard;
  const double* Xguard;
  const double* Yguard;
  /* R: Hankel plan (as fil_tables.dht_*) and r, dr [Nr] (grid units) */
  const double* dht_T;
  const double* dht_RJ;
  const double* dht_JV;
  const double* r;
  const double* dr;
  /* XY: grid x [Nx], y [Ny] and steps (grid units) */
  const double* x;
  const double* y;
  double dx, dy;
  /* nonlinear term: Rnl of the cubic response (cubic.jl:4-30, a real scalar)
   * and QZ (QZfunc R nonlinear_propagators.jl:78-96, XY :162-180), complex [n]
   * indexed like the field (paraxial: applied in real space, as the
   * reference's 1im * QZ .* dE; non-paraxial: in transform space) */
  double Rnl;
  const double* QZ;
} fil_spatial_tables;

/* FieldAnalyzerR (FieldAnalyzers.jl:17-70) and FieldAnalyzerXY (:218-266):
 * Imax, R: rfil = 2 radius(r, I), P = 2 pi sum(I r dr); XY: ax = radius(x,
 * I[:, jmax]), ay = radius(y, I[imax, :]), P = sum(I) dx dy.  imax, jmax: the
 * (first, x-fastest) position of the maximum (R: jmax = 0). */
typedef struct fil_diag_spatial {
  double z, Imax, rfil, ax, ay, P;
  int64_t imax, jmax;
} fil_diag_spatial;

int fil_spatial_create(const fil_spatial_desc* d, fil_spatial** h);
int fil_spatial_set_tables(fil_spatial* h, const fil_spatial_tables* t);
/* field.E (Fields.jl:31-51, 123-138): complex [n]; upload replaces the state. */
int fil_spatial_upload_field(fil_spatial* h, const double* E);
int fil_spatial_download_field(fil_spatial* h, double* E);
/* Models.zstep (Models.jl:75-121) for R and XY.  Asynchronous. */
int fil_spatial_zstep(fil_spatial* h, double z, double dz);
/* analyze! R / XY.  I [n] (|E|^2), and for the XY zdata of plot.h5
 * (WritePlots.jl:422-459) Izx = I[:, Ny/2] and Izy = I[Nx/2, :] (1-based
 * indices of the reference), may be NULL.  Blocks. */
int fil_spatial_analyze(fil_spatial* h, double z, fil_diag_spatial* d, double* I, double* Izx, double* Izy);
int fil_spatial_synchronize(fil_spatial* h);
int fil_spatial_destroy(fil_spatial* h);

#ifdef __cplusplus
}
#endif
#endif
