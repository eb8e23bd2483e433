"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY.  The oracle is the checker for libfilament_hip.so;
only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module.  The product path never does.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

MAXCOMP = 4
GEOM_RT, GEOM_T, GEOM_XYT = 0, 1, 2
MEDIUM_AIR, MEDIUM_CONST = 0, 1
OFF, THG, NOTHG = 0, 1, 2

_dp = ctypes.POINTER(ctypes.c_double)


class OrcConfig(ctypes.Structure):
    _fields_ = [
        ("geometry", ctypes.c_int32), ("Nr", ctypes.c_int32), ("Nt", ctypes.c_int32),
        ("rmax", ctypes.c_double), ("tmin", ctypes.c_double), ("tmax", ctypes.c_double),
        ("ru", ctypes.c_double), ("zu", ctypes.c_double), ("tu", ctypes.c_double),
        ("Iu", ctypes.c_double), ("rhou", ctypes.c_double), ("lam0", ctypes.c_double),
        ("medium", ctypes.c_int32), ("n_const", ctypes.c_double), ("n2", ctypes.c_double),
        ("nonlinearity", ctypes.c_int32), ("plasma", ctypes.c_int32),
        ("kparaxial", ctypes.c_int32), ("qparaxial", ctypes.c_int32),
        ("rguard", ctypes.c_double), ("tguard", ctypes.c_double),
        ("kguard", ctypes.c_double), ("wguard", ctypes.c_double),
        ("cubic", ctypes.c_int32), ("raman", ctypes.c_int32),
        ("current_free", ctypes.c_int32), ("nuc", ctypes.c_double), ("mr", ctypes.c_double),
        ("current_losses", ctypes.c_int32), ("losses_ereal", ctypes.c_int32),
        ("pe_ereal", ctypes.c_int32), ("pe_kdep", ctypes.c_int32),
        ("pe_rho0", ctypes.c_double), ("pe_rho_nt", ctypes.c_double),
        ("pe_nuc", ctypes.c_double), ("pe_mr", ctypes.c_double),
        ("ncomp", ctypes.c_int32),
        ("comp_frac", ctypes.c_double * MAXCOMP), ("comp_Ui", ctypes.c_double * MAXCOMP),
        ("comp_ntab", ctypes.c_int32 * MAXCOMP),
        ("comp_tab_x", _dp * MAXCOMP), ("comp_tab_y", _dp * MAXCOMP),
        ("Nx", ctypes.c_int32), ("Ny", ctypes.c_int32),
        ("xmin", ctypes.c_double), ("xmax", ctypes.c_double), ("ymin", ctypes.c_double), ("ymax", ctypes.c_double),
        ("xu", ctypes.c_double), ("yu", ctypes.c_double),
        ("xguard", ctypes.c_double), ("yguard", ctypes.c_double),
        ("kxguard", ctypes.c_double), ("kyguard", ctypes.c_double),
        ("guard3d_bug", ctypes.c_int32),
    ]


GEOM_R, GEOM_XY = 3, 4


class OrcSpatialConfig(ctypes.Structure):
    _fields_ = [
        ("geometry", ctypes.c_int32), ("Nr", ctypes.c_int32), ("rmax", ctypes.c_double), ("ru", ctypes.c_double),
        ("Nx", ctypes.c_int32), ("Ny", ctypes.c_int32),
        ("xmin", ctypes.c_double), ("xmax", ctypes.c_double), ("ymin", ctypes.c_double), ("ymax", ctypes.c_double),
        ("xu", ctypes.c_double), ("yu", ctypes.c_double),
        ("zu", ctypes.c_double), ("Iu", ctypes.c_double), ("lam0", ctypes.c_double),
        ("medium", ctypes.c_int32), ("n_const", ctypes.c_double), ("n2", ctypes.c_double),
        ("nonlinearity", ctypes.c_int32), ("kparaxial", ctypes.c_int32), ("qparaxial", ctypes.c_int32),
        ("alg", ctypes.c_int32),
        ("rguard", ctypes.c_double), ("kguard", ctypes.c_double),
        ("xguard", ctypes.c_double), ("yguard", ctypes.c_double), ("kxguard", ctypes.c_double),
        ("kyguard", ctypes.c_double),
    ]


class OrcDiag(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("z", "Fmax", "Imax", "rhomax", "De", "rfil", "rpl", "tau", "W")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


def build():
    """Compile liboracle.so from oracle/filament_oracle.cpp (make)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.POINTER(OrcConfig), _dp, ctypes.c_int]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_last_error.restype = ctypes.c_char_p
        for name in ("orc_get_grid", "orc_get_grid_xyt", "orc_get_scalars", "orc_get_KZ", "orc_get_QZ", "orc_get_guards",
                     "orc_get_dht", "orc_set_field", "orc_get_field", "orc_get_plasma",
                     "orc_plasma_solve", "orc_nl_rhs"):
            getattr(L, name).argtypes = None
            getattr(L, name).restype = ctypes.c_int
        L.orc_get_Rnl.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_zstep.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.orc_analyze.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(OrcDiag), _dp, _dp, _dp, _dp]
        L.orc_step_and_analyze.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, ctypes.c_double,
                                           ctypes.POINTER(OrcDiag), _dp]
        L.orc_bessel_j0_zeros.argtypes = [ctypes.c_int, _dp]
        L.orc_fft.argtypes = [ctypes.c_int, _dp, ctypes.c_int]
        L.orc_dht_apply.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int]
        L.orc_tf_eval.restype = ctypes.c_double
        L.orc_tf_eval.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int]
        L.orc_linear_step.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_nl_step.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double]
        L.orc_nl_step.restype = ctypes.c_int
        L.orc_set_plasma.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_set_plasma.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_threads.restype = ctypes.c_int
        L.orc_max_threads.argtypes = []
        L.orc_max_threads.restype = ctypes.c_int
        L.orc_spatial_create.restype = ctypes.c_void_p
        L.orc_spatial_create.argtypes = [ctypes.POINTER(OrcSpatialConfig), _dp, ctypes.c_int]
        L.orc_spatial_destroy.argtypes = [ctypes.c_void_p]
        for name in ("orc_spatial_set_field", "orc_spatial_get_field", "orc_spatial_get_grid",
                     "orc_spatial_get_tables", "orc_spatial_analyze"):
            getattr(L, name).argtypes = None
            getattr(L, name).restype = ctypes.c_int
        for name in ("orc_spatial_zstep", "orc_spatial_nl_step", "orc_spatial_linear_step"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_double]
            getattr(L, name).restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


def set_threads(n):
    """OpenMP threads of the oracle's later calls."""
    _check(lib().orc_set_threads(int(n)))


def max_threads():
    return int(lib().orc_max_threads())


def bessel_j0_zeros(n):
    out = np.zeros(n)
    _check(lib().orc_bessel_j0_zeros(n, _p(out)))
    return out


def fft(x, inverse_normalized=False):
    """Oracle radix-2 FFT: forward (P*E) or normalised inverse (P\\E)."""
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    _check(lib().orc_fft(a.size, _p(a.view(np.float64)), 1 if inverse_normalized else 0))
    return a


class Oracle:
    """One oracle instance for a physical configuration.

    ``cfg`` is a dict of the orc_config fields (physical units as in the
    reference input files); ``E0`` the real initial field [Nr][Nt] (time
    domain, before the analytic-signal conversion)."""

    def __init__(self, cfg, E0, precision=8):
        self._keep = []
        c = OrcConfig()
        tabs = cfg.get("tables", [])
        for k, v in cfg.items():
            if k in ("tables", "comp_frac", "comp_Ui"):
                continue
            setattr(c, k, v)
        for q, frac in enumerate(cfg.get("comp_frac", [])):
            c.comp_frac[q] = frac
        for q, ui in enumerate(cfg.get("comp_Ui", [])):
            c.comp_Ui[q] = ui
        for q, (x, y) in enumerate(tabs):
            x = np.ascontiguousarray(x, dtype=np.float64)
            y = np.ascontiguousarray(y, dtype=np.float64)
            self._keep += [x, y]
            c.comp_ntab[q] = x.size
            c.comp_tab_x[q] = _p(x)
            c.comp_tab_y[q] = _p(y)
        c.ncomp = len(tabs)
        self.cfg = c
        self.geometry = c.geometry
        self.Nr = c.Nr if c.geometry == GEOM_RT else (c.Nx * c.Ny if c.geometry == GEOM_XYT else 1)
        self.Nt = c.Nt
        E0 = np.ascontiguousarray(E0, dtype=np.float64).reshape(self.Nr, self.Nt)
        self.h = lib().orc_create(ctypes.byref(c), _p(E0), precision)
        if not self.h:
            raise RuntimeError("oracle: " + lib().orc_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(ctypes.c_void_p(self.h))
            self.h = None

    @property
    def hp(self):
        return ctypes.c_void_p(self.h)

    def grid_xyt(self):
        x, y = np.zeros(self.cfg.Nx), np.zeros(self.cfg.Ny)
        kx, ky = np.zeros(self.cfg.Nx), np.zeros(self.cfg.Ny)
        _check(lib().orc_get_grid_xyt(self.hp, _p(x), _p(y), _p(kx), _p(ky)))
        return dict(x=x, y=y, kx=kx, ky=ky)

    def grid(self):
        Nr, Nt = self.Nr, self.Nt
        r, dr, k = np.zeros(Nr), np.zeros(Nr), np.zeros(Nr)
        t, w = np.zeros(Nt), np.zeros(Nt)
        _check(lib().orc_get_grid(self.hp, _p(r), _p(dr), _p(k), _p(t), _p(w)))
        return dict(r=r, dr=dr, k=k, t=t, w=w)

    def scalars(self):
        o = np.zeros(9)
        _check(lib().orc_get_scalars(self.hp, _p(o)))
        return dict(zip(("w0", "n0", "Eu", "vf", "k0", "phik", "phip", "dt", "k2"), o))

    def _cplx(self, fn, shape):
        a = np.zeros(shape, dtype=np.complex128)
        _check(fn(self.hp, _p(a.view(np.float64))))
        return a

    def KZ(self):
        return self._cplx(lib().orc_get_KZ, (self.Nr, self.Nt))

    def QZ(self):
        return self._cplx(lib().orc_get_QZ, (self.Nr, self.Nt))

    def Rnl(self, which):
        a = np.zeros(self.Nt, dtype=np.complex128)
        _check(lib().orc_get_Rnl(self.hp, which, _p(a.view(np.float64))))
        return a

    def guards(self):
        Rg, Tg = np.zeros(self.Nr), np.zeros(self.Nt)
        Kg, Wg = np.zeros((self.Nr, self.Nt)), np.zeros(self.Nt)
        _check(lib().orc_get_guards(self.hp, _p(Rg), _p(Tg), _p(Kg), _p(Wg)))
        return dict(Rguard=Rg, Tguard=Tg, Kguard=Kg, Wguard=Wg)

    def dht(self):
        N = self.Nr
        T, RJ, JV = np.zeros((N, N)), np.zeros(N), np.zeros(N)
        _check(lib().orc_get_dht(self.hp, _p(T), _p(RJ), _p(JV)))
        return T, RJ, JV

    def dht_apply(self, f, inverse=False):
        a = np.ascontiguousarray(f, dtype=np.complex128).copy()
        _check(lib().orc_dht_apply(self.hp, _p(a.view(np.float64)), 1 if inverse else 0))
        return a

    def tf(self, comp, I, power=False):
        return lib().orc_tf_eval(self.hp, comp, float(I), 1 if power else 0)

    def field(self):
        return self._cplx(lib().orc_get_field, (self.Nr, self.Nt))

    def set_field(self, E):
        a = np.ascontiguousarray(E, dtype=np.complex128)
        _check(lib().orc_set_field(self.hp, _p(a.view(np.float64))))

    def plasma(self):
        rho = np.zeros((self.Nr, self.Nt))
        kd = np.zeros((self.Nr, self.Nt))
        _check(lib().orc_get_plasma(self.hp, _p(rho), _p(kd)))
        return rho, kd

    def plasma_solve(self):
        _check(lib().orc_plasma_solve(self.hp))

    def zstep(self, z, dz):
        _check(lib().orc_zstep(self.hp, z, dz))

    def linear_step(self, dz):
        _check(lib().orc_linear_step(self.hp, dz))

    def nl_rhs(self, Es):
        a = np.ascontiguousarray(Es, dtype=np.complex128)
        out = np.zeros_like(a)
        _check(lib().orc_nl_rhs(self.hp, _p(a.view(np.float64)), _p(out.view(np.float64))))
        return out

    def nl_step(self, Es, dz):
        """One RK3 step of the nonlinear propagator on spectral rows (plasma as it stands)."""
        a = np.ascontiguousarray(Es, dtype=np.complex128).copy()
        _check(lib().orc_nl_step(self.hp, _p(a.view(np.float64)), float(dz)))
        return a

    def set_plasma(self, rho=None, kdrho=None):
        r = None if rho is None else np.ascontiguousarray(np.broadcast_to(rho, (self.Nr, self.Nt)), dtype=np.float64)
        k = None if kdrho is None else np.ascontiguousarray(np.broadcast_to(kdrho, (self.Nr, self.Nt)),
                                                             dtype=np.float64)
        _check(lib().orc_set_plasma(self.hp, None if r is None else _p(r), None if k is None else _p(k)))

    def analyze(self, z=0.0):
        d = OrcDiag()
        Fr, Ft = np.zeros(self.Nr), np.zeros(self.Nt)
        re, S = np.zeros(self.Nr), np.zeros(self.Nt // 2 + 1)
        _check(lib().orc_analyze(self.hp, z, ctypes.byref(d), _p(Fr), _p(Ft), _p(re), _p(S)))
        out = d.as_dict()
        out.update(Fr=Fr, Ft=Ft, rho_end=re, S=S)
        return out

    def run(self, nsteps, dz0, phimax=float("inf"), z0=0.0, diag_every=0):
        """Reference main loop without I/O (Filament.jl:123-164): per step
        the adaptive dz, zstep and analyze!.  Returns the list of diagnostics
        (initial one first)."""
        d = OrcDiag()
        _check(lib().orc_analyze(self.hp, z0, ctypes.byref(d), None, None, None, None))
        rows = [d.as_dict()]
        z = ctypes.c_double(z0)
        dz = ctypes.c_double(0.0)
        for n in range(nsteps):
            _check(lib().orc_step_and_analyze(self.hp, ctypes.byref(z), dz0, phimax,
                                              ctypes.byref(d), ctypes.byref(dz)))
            rows.append(d.as_dict())
        return rows


class SpatialOracle:
    """The oracle of the spatial geometries R and XY (filament_oracle.h,
    orc_spatial_*).  ``cfg`` is an OrcSpatialConfig, ``E0`` the complex
    initial field ([Nr], or [Nx, Ny])."""

    def __init__(self, cfg, E0, precision=8):
        self.cfg = cfg
        self.xy = cfg.geometry == GEOM_XY
        self.shape = (cfg.Nx, cfg.Ny) if self.xy else (cfg.Nr,)
        self.n = int(np.prod(self.shape))
        e = np.ascontiguousarray(np.asarray(E0, dtype=np.complex128).ravel(order="F"))
        self.h = lib().orc_spatial_create(ctypes.byref(cfg), _p(e.view(np.float64)), int(precision))
        if not self.h:
            raise RuntimeError("oracle: " + lib().orc_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_spatial_destroy(ctypes.c_void_p(self.h))
            self.h = None

    def _h(self):
        return ctypes.c_void_p(self.h)

    def field(self):
        out = np.zeros(self.n, dtype=np.complex128)
        _check(lib().orc_spatial_get_field(self._h(), _p(out.view(np.float64))))
        return out.reshape(self.shape, order="F")

    def set_field(self, E):
        e = np.ascontiguousarray(np.asarray(E, dtype=np.complex128).ravel(order="F"))
        _check(lib().orc_spatial_set_field(self._h(), _p(e.view(np.float64))))

    def grid(self):
        if self.xy:
            a, b, c, d = np.zeros(self.cfg.Nx), np.zeros(self.cfg.Ny), np.zeros(self.cfg.Nx), np.zeros(self.cfg.Ny)
            _check(lib().orc_spatial_get_grid(self._h(), _p(a), _p(b), _p(c), _p(d)))
            return dict(x=a, y=b, kx=c, ky=d)
        a, b, c = np.zeros(self.n), np.zeros(self.n), np.zeros(self.n)
        _check(lib().orc_spatial_get_grid(self._h(), _p(a), _p(b), _p(c), None))
        return dict(r=a, dr=b, k=c)

    def tables(self):
        KZ, QZ = np.zeros(self.n, dtype=np.complex128), np.zeros(self.n, dtype=np.complex128)
        Kg, Fg, s7 = np.zeros(self.n), np.zeros(self.n), np.zeros(7)
        _check(lib().orc_spatial_get_tables(self._h(), _p(KZ.view(np.float64)), _p(QZ.view(np.float64)), _p(Kg),
                                            _p(Fg), _p(s7)))
        sh = lambda a: a.reshape(self.shape, order="F")
        return dict(KZ=sh(KZ), QZ=sh(QZ), Kg=sh(Kg), Fg=sh(Fg), w0=s7[0], n0=s7[1], Eu=s7[2], vf=s7[3], k0=s7[4],
                    phik=s7[5], Rnl=s7[6])

    def zstep(self, dz):
        _check(lib().orc_spatial_zstep(self._h(), float(dz)))

    def nl_step(self, dz):
        _check(lib().orc_spatial_nl_step(self._h(), float(dz)))

    def linear_step(self, dz):
        _check(lib().orc_spatial_linear_step(self._h(), float(dz)))

    def analyze(self):
        o = np.zeros(6)
        _check(lib().orc_spatial_analyze(self._h(), _p(o)))
        return dict(Imax=o[0], rfil=o[1], ax=o[2], ay=o[3], P=o[4], index=int(o[5]))

    def run(self, nsteps, dz):
        for _ in range(nsteps):
            self.zstep(dz)
        return self.field()
