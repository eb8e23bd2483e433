"""Host-side setup of the RT and XYT problems in FP64 (the part the reference keeps in
Julia on the host): units, grids, medium, guards and the operator tables that
are handed to libfilament_hip.so once.

Each function cites the reference file:line it follows (src/...).
"""
import math

import numpy as np
import scipy.special as sps

# CODATA2018 (src/modules/Constants.jl:9-14)
C0 = 299792458.0
EPS0 = 8.8541878128e-12
MU0 = 1.25663706212e-6
QE = 1.602176634e-19
ME = 9.1093837015e-31
HBAR = 1.054571817e-34


# ---------------------------------------------------------------------------
# Units (src/modules/Units.jl:67-85, 137-141)
# ---------------------------------------------------------------------------
class UnitRT:
    def __init__(self, ru, zu, tu, Iu, rhou):
        self.r, self.k, self.z, self.t = ru, 1.0 / ru, zu, tu
        self.w, self.lam, self.I, self.rho = 1.0 / tu, tu, Iu, rhou


class UnitXYT:
    """Units.jl:110-130 (XYT): the same time, intensity and density units as
    RT, with x and y in place of r."""
    def __init__(self, xu, yu, zu, tu, Iu, rhou):
        self.x, self.y, self.kx, self.ky = xu, yu, 1.0 / xu, 1.0 / yu
        self.z, self.t, self.w, self.lam, self.I, self.rho = zu, tu, 1.0 / tu, tu, Iu, rhou


class UnitT:
    """Units.jl:47-61 (T): time, intensity and density units only."""
    def __init__(self, zu, tu, Iu, rhou):
        self.z, self.t, self.w, self.lam, self.I, self.rho = zu, tu, 1.0 / tu, tu, Iu, rhou


class UnitR:
    """Units.jl:28-41 (R): r, k, z and intensity units."""
    def __init__(self, ru, zu, Iu):
        self.r, self.k, self.z, self.I = ru, 1.0 / ru, zu, Iu


class UnitXY:
    """Units.jl:90-104 (XY): x, y, kx, ky, z and intensity units."""
    def __init__(self, xu, yu, zu, Iu):
        self.x, self.y, self.kx, self.ky, self.z, self.I = xu, yu, 1.0 / xu, 1.0 / yu, zu, Iu


def field_unit(unit, n0):
    """Units.E (Units.jl:137-141)."""
    return math.sqrt(unit.I / (n0 * EPS0 * C0 / 2))


# ---------------------------------------------------------------------------
# Grids (src/modules/Grids.jl:97-121, 204-237)
# ---------------------------------------------------------------------------
def _step(x):
    n = len(x)
    d = np.empty(n)
    d[0] = x[1] - x[0]
    d[-1] = x[-1] - x[-2]
    d[1:-1] = (x[2:] - x[:-2]) / 2
    return d


class GridRT:
    def __init__(self, rmax, Nr, tmin, tmax, Nt):
        # HankelTransforms.dhtcoord / dhtfreq: r_n = j_n R / j_{N+1}, k_n = j_n / R
        j = sps.jn_zeros(0, Nr + 1)
        self.jzeros = j
        self.rmax, self.Nr = rmax, Nr
        self.r = j[:Nr] * rmax / j[Nr]
        self.k = j[:Nr] / rmax
        self.dr = _step(self.r)
        self.tmin, self.tmax, self.Nt = tmin, tmax, Nt
        self.t = np.linspace(tmin, tmax, Nt)
        self.dt = (tmax - tmin) / (Nt - 1)
        self.w = 2 * np.pi * np.fft.fftfreq(Nt, self.dt)


class GridR:
    """Grids.jl:33-45 (R): the axial grid of _grid_spatial_axial (:204-216),
    the same r, dr and k as the RT grid's."""
    def __init__(self, rmax, Nr):
        j = sps.jn_zeros(0, Nr + 1)
        self.jzeros = j
        self.rmax, self.Nr = rmax, Nr
        self.r = j[:Nr] * rmax / j[Nr]
        self.k = j[:Nr] / rmax
        self.dr = _step(self.r)


class GridXY:
    """Grids.jl:127-155 (XY): uniform x and y (_grid_spatial_rectangular),
    kx = 2 pi fftfreq(Nx, 1/dx)."""
    def __init__(self, xmin, xmax, Nx, ymin, ymax, Ny):
        self.Nx, self.Ny = Nx, Ny
        self.x = np.linspace(xmin, xmax, Nx)
        self.dx = (xmax - xmin) / (Nx - 1)
        self.kx = 2 * np.pi * np.fft.fftfreq(Nx, self.dx)
        self.y = np.linspace(ymin, ymax, Ny)
        self.dy = (ymax - ymin) / (Ny - 1)
        self.ky = 2 * np.pi * np.fft.fftfreq(Ny, self.dy)


class GridT:
    """Grids.jl:78-91: t = range(tmin, tmax, Nt), w = 2 pi fftfreq(Nt, 1/dt)."""
    def __init__(self, tmin, tmax, Nt):
        self.tmin, self.tmax, self.Nt = tmin, tmax, Nt
        self.t = np.linspace(tmin, tmax, Nt)
        self.dt = (tmax - tmin) / (Nt - 1)
        self.w = 2 * np.pi * np.fft.fftfreq(Nt, self.dt)


class GridXYT:
    """Grids.jl:158-201: uniform x, y (range(min, max, length=N)),
    kx = 2 pi fftfreq(Nx, 1/dx); t as for RT."""
    def __init__(self, xmin, xmax, Nx, ymin, ymax, Ny, tmin, tmax, Nt):
        self.Nx, self.Ny, self.Nt = Nx, Ny, Nt
        self.x = np.linspace(xmin, xmax, Nx)
        self.dx = (xmax - xmin) / (Nx - 1)
        self.kx = 2 * np.pi * np.fft.fftfreq(Nx, self.dx)
        self.y = np.linspace(ymin, ymax, Ny)
        self.dy = (ymax - ymin) / (Ny - 1)
        self.ky = 2 * np.pi * np.fft.fftfreq(Ny, self.dy)
        self.tmin, self.tmax = tmin, tmax
        self.t = np.linspace(tmin, tmax, Nt)
        self.dt = (tmax - tmin) / (Nt - 1)
        self.w = 2 * np.pi * np.fft.fftfreq(Nt, self.dt)


# ---------------------------------------------------------------------------
# Media (src/modules/Media.jl:8-191)
# ---------------------------------------------------------------------------
class Medium:
    def __init__(self, permittivity, permeability, n2):
        self.permittivity, self.permeability, self.n2 = permittivity, permeability, n2

    def refractive_index(self, w):
        w = np.abs(np.asarray(w, dtype=float))
        eps = np.vectorize(self.permittivity, otypes=[float])(w)
        mu = np.vectorize(self.permeability, otypes=[float])(w)
        return np.sqrt(eps * mu + 0j)

    def beta(self, w):
        return self.refractive_index(w) * np.asarray(w, dtype=float) / C0

    def k(self, w):
        return np.real(self.beta(w))

    def derivative_k(self, x, n):
        h = 0.01 if x == 0 else 0.001 * x
        f = lambda v: float(self.k(v))
        if n == 1:
            return (f(x - 2 * h) - 8 * f(x - h) + 8 * f(x + h) - f(x + 2 * h)) / (12 * h)
        if n == 2:
            return (-f(x - 2 * h) + 16 * f(x - h) - 30 * f(x) + 16 * f(x + h) - f(x + 2 * h)) / (12 * h * h)
        if n == 3:
            return (-f(x - 2 * h) + 2 * f(x - h) - 2 * f(x + h) + f(x + 2 * h)) / (2 * h ** 3)
        raise ValueError("Wrong derivative order.")

    def group_velocity(self, w):
        return 1.0 / self.derivative_k(w, 1)

    def chi3(self, w):
        n = float(np.real(self.refractive_index(w)))
        return 4.0 / 3.0 * n * n * EPS0 * C0 * self.n2

    def critical_power(self, w):
        lam = 2 * np.pi * C0 / w
        n = float(np.real(self.refractive_index(w)))
        return 3.79 * lam ** 2 / (8.0 * np.pi * abs(n) * abs(self.n2))


def peck_reeder_air(w):
    """E.R. Peck and K. Reeder, JOSA 62, 958 (1972) -- the air permittivity
    of the reference's inputs (test/diffraction/RT/air.jl:4-17)."""
    if w == 0.0:
        sig = 0.0
    else:
        lam = 2.0 * np.pi * C0 / w
        sig = 1.0 / lam * 1e-6
    dum = 8060.51 + 2480990.0 / (132.247 - sig ** 2) + 17455.7 / (39.32957 - sig ** 2)
    n = 1.0 + dum * 1e-8
    return n * n


def unit_permeability(w):
    return 1.0


# ---------------------------------------------------------------------------
# Guards (src/modules/Guards.jl:78-195, windows :322-390)
# ---------------------------------------------------------------------------
def guard_window_left(x, width, p=6):
    if width >= x[-1] - x[0]:
        raise ValueError("Guard width is larger or equal than the grid size.")
    if width == 0:
        return np.ones(len(x))
    xloc1, xloc2 = x[0], x[0] + width
    g1 = np.where(x >= xloc1, 1 - np.exp(-((x - xloc1) / (width / 2)) ** p), 0.0)
    g2 = np.where(x <= xloc2, np.exp(-((x - xloc2) / (width / 2)) ** p), 1.0)
    return (g1 + g2) / 2


def guard_window_right(x, width, p=6):
    if width >= x[-1] - x[0]:
        raise ValueError("Guard width is larger or equal than the grid size.")
    if width == 0:
        return np.ones(len(x))
    xloc1, xloc2 = x[-1] - width, x[-1]
    g1 = np.where(x >= xloc1, np.exp(-((x - xloc1) / (width / 2)) ** p), 1.0)
    g2 = np.where(x <= xloc2, 1 - np.exp(-((x - xloc2) / (width / 2)) ** p), 0.0)
    return (g1 + g2) / 2


def guard_window_both(x, width, p=6):
    if width >= (x[-1] - x[0]) / 2:
        raise ValueError("Guard width is larger or equal than the grid size.")
    return guard_window_left(x, width, p) + guard_window_right(x, width, p) - 1


class GuardRT:
    def __init__(self, unit, grid, medium, rguard, tguard, kguard, wguard):
        self.Rguard = guard_window_right(grid.r, rguard)
        self.Tguard = guard_window_both(grid.t, tguard)
        kk = medium.k(grid.w * unit.w)
        kmax = kk * math.sin(math.radians(kguard))
        with np.errstate(divide="ignore", invalid="ignore"):
            Kg = np.exp(-(((grid.k[:, None] * unit.k) ** 2 / kmax[None, :] ** 2) ** 20))
        Kg[:, kmax == 0] = 0.0
        self.Kguard = Kg                                             # [Nr, Nt]
        self.Wguard = np.exp(-(((grid.w * unit.w) ** 2 / wguard ** 2) ** 20))


class GuardT:
    """Guards.jl:78-89 (Guard1D): the field window T(t) and the spectral W(w)."""
    def __init__(self, unit, grid, tguard, wguard):
        self.Tguard = guard_window_both(grid.t, tguard)
        self.Wguard = np.exp(-(((grid.w * unit.w) ** 2 / wguard ** 2) ** 20))


class GuardXYT:
    """Guards.jl:162-195: field windows X(x), Y(y), T(t); the spectral KX, KY
    guards are ones in the reference (its kmax lines are commented out,
    :176-184); W(w) = exp(-((w wu)^2 / wguard^2)^20)."""
    def __init__(self, unit, grid, xguard, yguard, tguard, wguard):
        self.Xguard = guard_window_both(grid.x, xguard)
        self.Yguard = guard_window_both(grid.y, yguard)
        self.Tguard = guard_window_both(grid.t, tguard)
        self.KXguard = np.ones(grid.Nx)
        self.KYguard = np.ones(grid.Ny)
        self.Wguard = np.exp(-(((grid.w * unit.w) ** 2 / wguard ** 2) ** 20))


class GuardR:
    """Guards.jl:41-57 (R): Rguard = guard_window_right(r, rguard) and
    Kguard = exp(-((k ku)^2 / kmax^2)^20), kmax = k0 sind(kguard)."""
    def __init__(self, unit, grid, medium, w0, rguard, kguard):
        self.Rguard = guard_window_right(grid.r, rguard)
        kmax = float(medium.k(w0)) * math.sin(math.radians(kguard))
        self.Kguard = np.exp(-(((grid.k * unit.k) ** 2 / kmax ** 2) ** 20))


class GuardXY:
    """Guards.jl:132-157 (XY): field windows X(x), Y(y) (guard_window_both)
    and the spectral KX(kx), KY(ky) with kmax = k0 sind(kxguard), sind(kyguard)."""
    def __init__(self, unit, grid, medium, w0, xguard, yguard, kxguard, kyguard):
        self.Xguard = guard_window_both(grid.x, xguard)
        self.Yguard = guard_window_both(grid.y, yguard)
        k0 = float(medium.k(w0))
        kxmax, kymax = k0 * math.sin(math.radians(kxguard)), k0 * math.sin(math.radians(kyguard))
        self.KXguard = np.exp(-(((grid.kx * unit.kx) ** 2 / kxmax ** 2) ** 20))
        self.KYguard = np.exp(-(((grid.ky * unit.ky) ** 2 / kymax ** 2) ** 20))


def kt_spatial(unit, grid):
    """Transverse wavenumber [1/m] of every spectral point: R k ku [Nr];
    XY sqrt((kx kxu)^2 + (ky kyu)^2) [Nx, Ny]."""
    if hasattr(grid, "k"):
        return np.asarray(grid.k) * unit.k
    return np.sqrt((grid.kx[:, None] * unit.kx) ** 2 + (grid.ky[None, :] * unit.ky) ** 2)


def kz_spatial(unit, grid, medium, w0, paraxial):
    """KZ R (spectral.jl:30-50) / XY (:104-128): (Kfunc(w0, kt) - w0 / vf) zu,
    Kfunc = beta - kt^2 / (2 beta) (paraxial) or sqrt(beta^2 - kt^2 + 0im)
    (:160-184)."""
    vf = medium.group_velocity(w0)
    beta = complex(medium.beta(w0))
    kt = kt_spatial(unit, grid)
    if paraxial:
        K = (beta - kt ** 2 / (2 * beta)) if beta != 0 else np.zeros(kt.shape, dtype=complex)
    else:
        K = np.sqrt(beta ** 2 - kt ** 2 + 0j)
    return K * unit.z - w0 / vf * unit.z


def qz_spatial(unit, grid, medium, w0, Eu, paraxial):
    """QZ R (nonlinear_propagators.jl:78-96) / XY (:162-180): Qfunc(w0, kt)
    zu / Eu; paraxial mu0 mu w0^2 / (2 beta) (the same everywhere), else
    mu0 mu w0^2 / (2 kz), kz = sqrt(beta^2 - kt^2 + 0im), 0 where kz = 0
    (:210-242)."""
    beta = complex(medium.beta(w0))
    mu = medium.permeability(w0)
    kt = kt_spatial(unit, grid)
    if paraxial:
        Q = np.full(kt.shape, 0.0 if beta == 0 else MU0 * mu * w0 ** 2 / (2 * beta), dtype=complex)
    else:
        kz = np.sqrt(beta ** 2 - kt ** 2 + 0j)
        with np.errstate(divide="ignore", invalid="ignore"):
            Q = np.where(kz == 0, 0.0, MU0 * mu * w0 ** 2 / (2 * kz))
    return Q * unit.z / Eu


def kz_factors_xyt(unit, grid, medium, w0, vf=None):
    """The separable factors of KZ XYT (spectral.jl:131-156): beta(w) at the grid
    frequencies, w/vf, and kt^2 = (kx kxu)^2 + (ky kyu)^2 split over x and y.
    ``vf`` overrides the frame velocity (Media.group_velocity, a 5-point
    derivative that two FP64 implementations round differently)."""
    vf = medium.group_velocity(w0) if vf is None else vf
    w = grid.w * unit.w
    beta = np.asarray(medium.beta(w), dtype=complex)
    return dict(beta=beta, w_vf=w / vf, kt2x=(grid.kx * unit.kx) ** 2, kt2y=(grid.ky * unit.ky) ** 2)


def kz_table_xyt(unit, grid, medium, w0, paraxial):
    """Full KZ[ix, iy, iw] (spectral.jl:131-156, Kfunc :160-184), for checks."""
    f = kz_factors_xyt(unit, grid, medium, w0)
    kt2 = f["kt2x"][:, None, None] + f["kt2y"][None, :, None]
    beta = f["beta"][None, None, :]
    if paraxial:
        with np.errstate(divide="ignore", invalid="ignore"):
            K = np.where(beta != 0, beta - kt2 / (2 * beta), 0.0)
    else:
        K = np.sqrt(beta ** 2 - kt2 + 0j)
    return (K - f["w_vf"][None, None, :]) * unit.z


# ---------------------------------------------------------------------------
# Linear propagator KZ (linear_propagators/spectral.jl:78-101,160-184)
# ---------------------------------------------------------------------------
def kz_table(unit, grid, medium, w0, paraxial):
    vf = medium.group_velocity(w0)
    w = grid.w * unit.w
    beta = medium.beta(w)[None, :]
    kt = (grid.k * unit.k)[:, None]
    if paraxial:
        with np.errstate(divide="ignore", invalid="ignore"):
            K = np.where(beta != 0, beta - kt ** 2 / (2 * beta), 0.0)
    else:
        K = np.sqrt(beta ** 2 - kt ** 2 + 0j)
    return K * unit.z - (w / vf * unit.z)[None, :]


def kz_table_t(unit, grid, medium, w0):
    """KZ T (spectral.jl:55-75): kt = 0 and paraxial (Input.jl:96-103 forces
    it for T), i.e. (beta(w) - w / vf) zu."""
    vf = medium.group_velocity(w0)
    w = grid.w * unit.w
    beta = np.asarray(medium.beta(w), dtype=complex)
    return beta * unit.z - w / vf * unit.z


# ---------------------------------------------------------------------------
# Nonlinear prefactor (nonlinear_propagators.jl:139-160, 210-230), paraxial
# ---------------------------------------------------------------------------
def qz_vector(unit, grid, medium, w0, Eu):
    w = grid.w * unit.w
    beta = medium.beta(w)
    mu = np.vectorize(medium.permeability, otypes=[float])(w)
    with np.errstate(divide="ignore", invalid="ignore"):
        Q = np.where(beta == 0, 0.0, MU0 * mu * w ** 2 / (2 * beta))
    return Q * unit.z / Eu


def qz_table_rt(unit, grid, medium, w0, Eu):
    """Non-paraxial QZ(k, w) of QZfunc RT with Qfunc_nonparaxial
    (nonlinear_propagators.jl:139-160, 232-242): mu0 mu w^2 / (2 kz) zu / Eu,
    kz = sqrt(beta^2 - kt^2 + 0im), 0 where kz = 0; [Nr, Nt] complex."""
    w = grid.w * unit.w
    beta = np.asarray(medium.beta(w), dtype=complex)
    mu = np.vectorize(medium.permeability, otypes=[float])(w)
    kt = np.asarray(grid.k) * unit.k
    kz = np.sqrt(beta[None, :] ** 2 - kt[:, None] ** 2 + 0j)
    with np.errstate(divide="ignore", invalid="ignore"):
        Q = np.where(kz == 0, 0.0, MU0 * mu[None, :] * w[None, :] ** 2 / (2 * kz))
    return Q * unit.z / Eu


# ---------------------------------------------------------------------------
# Responses (medium_responses/*.jl)
# ---------------------------------------------------------------------------
def rnl_cubic(unit, medium, w0, Eu, Nt, n2=None):
    """cubic.jl:4-30 (scalar, broadcast over w); also the Raman Rnl
    (raman.jl:4-14).  n2: the response's own nonlinear index (default: the
    medium's)."""
    n = float(np.real(medium.refractive_index(w0)))
    chi3 = medium.chi3(w0) if n2 is None else 4.0 / 3.0 * n * n * EPS0 * C0 * n2
    return np.full(Nt, EPS0 * chi3 * Eu ** 3, dtype=complex)


def hraman_table(unit, grid, raman_response):
    """Hraman(w) of init_raman (raman.jl:15-30): the response sampled at
    (t - tshift) tu, tshift the window centre, times tu; ifftshift; ifft
    (normalised, exp(+i w t)); times Nt dt."""
    tshift = grid.tmin + 0.5 * (grid.tmax - grid.tmin)
    h = np.array([raman_response((tt - tshift) * unit.t) for tt in grid.t], dtype=float) * unit.t
    H = np.fft.ifft(np.fft.ifftshift(h.astype(complex)))
    return H * grid.Nt * grid.dt


def mlejnek_raman_response(t):
    """Raman response of air (examples/RT/air.jl:47-58, Mlejnek, Wright and
    Moloney, Opt. Lett. 23, 382 (1998)); t in seconds."""
    if t < 0.0:
        return 0.0
    Om, Ga = 20.6e12, 26e12
    La = np.sqrt(Om ** 2 - Ga ** 2 / 4.0)
    return Om ** 2 * np.exp(-0.5 * Ga * t) * np.sin(La * t) / La


def rnl_current_free(unit, grid, Eu, nuc, mr):
    """current_free.jl:4-28."""
    MR = mr * ME
    w = grid.w * unit.w
    out = np.zeros(grid.Nt, dtype=complex)
    nz = w != 0
    out[nz] = 1j / w[nz] * QE ** 2 / MR / (nuc - 1j * w[nz]) * unit.rho * Eu
    return out


def rnl_current_losses(unit, grid, Eu, w0):
    """current_losses.jl:4-33."""
    w = grid.w * unit.w
    out = np.zeros(grid.Nt, dtype=complex)
    nz = w != 0
    out[nz] = 1j / w[nz] * HBAR * w0 * unit.rho / (unit.t * Eu)
    return out


# ---------------------------------------------------------------------------
# Plasma equation (photoionization_avalanche.jl:1-116)
# ---------------------------------------------------------------------------
def read_tf(fname):
    """Two-column rate table (x [W/m^2], y [1/s])."""
    d = np.loadtxt(fname, ndmin=2)
    return d[:, 0].copy(), d[:, 1].copy()


def plasma_tables(unit, w0, Eu, params, basedir="."):
    comps = []
    for comp in params["components"]:
        src = comp["ionization_rate"]
        if isinstance(src, str):
            x, y = read_tf(_resolve(src, basedir))
        else:
            x, y = (np.asarray(v, dtype=float) for v in src)
        Ui = comp["ionization_energy"] * QE
        MR = params["mr"] * ME
        nuc = params["nuc"]
        sigmaB = QE ** 2 / MR * nuc / (nuc ** 2 + w0 ** 2)
        comps.append(dict(
            log10x=np.log10(x / unit.I), log10y=np.log10(y * unit.t),
            frhont=comp["fraction"] * params["rho_nt"] / unit.rho,
            Rava=sigmaB / Ui * unit.t * Eu ** 2,
            K=math.ceil(Ui / (HBAR * w0)),
            x=x, y=y, fraction=comp["fraction"], Ui_eV=comp["ionization_energy"]))
    return dict(components=comps, rho0=params["rho0"] / unit.rho,
                EREAL=bool(params["EREAL"]), KDEP=bool(params["KDEP"]))


def _resolve(fname, basedir):
    import os
    return fname if os.path.isabs(fname) else os.path.join(basedir, fname)


# ---------------------------------------------------------------------------
# QDHT plan (HankelTransforms.plan_dht, used at Fields.jl:108-110):
#   T_mn = 2 J0(j_m j_n / S) / (S |J1(j_m)| |J1(j_n)|),  S = j_{N+1}
#   forward  f -> JV .* (T * (RJ .* f)),  RJ = R/|J1|, JV = |J1|/V, V = S/(2 pi R)
#   inverse  F -> (T * (F ./ JV)) ./ RJ
# Guizar-Sicairos & Gutierrez-Vega, JOSA A 21, 53 (2004).
# ---------------------------------------------------------------------------
def dht_plan(grid):
    j = grid.jzeros
    N = grid.Nr
    S = j[N]
    J1 = np.abs(sps.j1(j[:N]))
    T = 2.0 * sps.j0(np.outer(j[:N], j[:N]) / S) / (S * np.outer(J1, J1))
    V = S / (2 * np.pi * grid.rmax)
    return T, grid.rmax / J1, J1 / V


# ---------------------------------------------------------------------------
# Adaptive step (AdaptiveSteps.jl:46-72, 96-107)
# ---------------------------------------------------------------------------
def phi_kerr(unit, medium, w0, Eu):
    k0 = float(medium.k(w0))
    QZ0 = MU0 * medium.permeability(w0) * w0 ** 2 / (2 * k0) * unit.z / Eu
    return QZ0 * abs(EPS0 * medium.chi3(w0) * 3 / 4 * Eu ** 3)


def phi_plasma(unit, medium, w0, Eu, mr, nuc):
    k0 = float(medium.k(w0))
    MR = mr * ME
    QZ0 = MU0 * medium.permeability(w0) * w0 ** 2 / (2 * k0) * unit.z / Eu
    Rnl0 = 1j / w0 * QE ** 2 / MR / (nuc - 1j * w0) * unit.rho * Eu
    return QZ0 * abs(Rnl0.real)


def rsig2asig(E):
    """AnalyticSignals.rsig2asig! along the last axis (Fields.jl:115): the
    spectrum keeps bin 0, doubles bins 1..N/2-1 and zeroes N/2..N-1."""
    N = E.shape[-1]
    S = np.fft.ifft(E, axis=-1)
    S[..., 1:N // 2] *= 2
    S[..., N // 2:] = 0
    return np.fft.fft(S, axis=-1)
