"""RT and XYT models on the host + hot path on libfilament_hip.so.

``RTModel(p)`` / ``XYTModel(p)`` play the role of Filament.run's setup for
geometry "RT" / "XYT" (src/Filament.jl:69-76): Unit, Grid, Field, Medium,
Guard, Model and AStep are built on the host in FP64 and the tables are handed
to the library once.  ``zstep`` is Models.zstep (src/modules/Models.jl:75-121)
on the GPU and ``analyze`` is FieldAnalyzers.analyze! (RT
FieldAnalyzers.jl:166-215, XYT :308-353).

``p`` carries the reference's input-file names (src/Input.jl:106-178 and the
medium / initial-condition files), e.g. for RT ``ru, zu, tu, Iu, rhou, rmax,
Nr, tmin, tmax, Nt, lam0, permittivity, permeability, n2, NONLINEARITY,
PLASMA, KPARAXIAL, QPARAXIAL, rguard, tguard, kguard, wguard, responses,
plasma_equation, initial_condition``; for XYT ``xu, yu`` and ``xmin, xmax, Nx,
ymin, ymax, Ny, xguard, yguard, kxguard, kyguard`` in place of the r entries.
"""
import ctypes
import math

import numpy as np

from . import lib as L
from . import physics as ph
from .responses import alg_order, response_kind


class _Model:
    """Host setup and boundary calls shared by the geometries.  Rows of the
    device field are the spatial points (r, or (x, y) with x fastest); the t
    axis is last."""

    nrows = 0

    def _physics(self, p, basedir):
        u, g = self.unit, self.grid
        self.medium = m = ph.Medium(p["permittivity"], p.get("permeability", ph.unit_permeability), p["n2"])
        self.w0 = 2 * math.pi * ph.C0 / p["lam0"]
        self.n0 = float(np.real(m.refractive_index(self.w0)))
        self.Eu = ph.field_unit(u, self.n0)
        self.nonlinearity = bool(p["NONLINEARITY"])
        self.plasma = bool(p["PLASMA"])
        self.nonparaxial = bool(self.nonlinearity and not p.get("QPARAXIAL", True))
        if self.nonparaxial and not isinstance(self, (RTModel, XYTModel)):
            raise NotImplementedError("QPARAXIAL=false is built for the RT and XYT geometries")
        # responses (nonlinear_propagators.jl:20-27)
        resp_flags = 0
        self.Rnl = {}
        self.losses_ereal = 0
        nuc_dz, mr_dz = p.get("nuc", 0.0), p.get("mr", 1.0)
        if self.nonlinearity:
            for item in p["responses"]:
                kind = response_kind(item["init"])
                if kind == "cubic":
                    resp_flags |= L.RESP_CUBIC_THG if item["THG"] else L.RESP_CUBIC_NOTHG
                    self.Rnl["cubic"] = ph.rnl_cubic(u, m, self.w0, self.Eu, g.Nt, item.get("n2"))
                elif kind == "current_free":
                    resp_flags |= L.RESP_FREE
                    self.Rnl["free"] = ph.rnl_current_free(u, g, self.Eu, item["nuc"], item["mr"])
                elif kind == "current_losses":
                    resp_flags |= L.RESP_LOSSES
                    self.losses_ereal = int(bool(item["EREAL"]))
                    self.Rnl["losses"] = ph.rnl_current_losses(u, g, self.Eu, self.w0)
                elif kind == "raman":
                    resp_flags |= L.RESP_RAMAN_THG if item["THG"] else L.RESP_RAMAN_NOTHG
                    self.Rnl["raman"] = ph.rnl_cubic(u, m, self.w0, self.Eu, g.Nt, item.get("n2"))
                    self.Hraman = ph.hraman_table(u, g, item["raman_response"])
                else:
                    raise ValueError(f"unknown response {kind}")
            # QPARAXIAL: QZ depends on w only, also in XYT (nonlinear_propagators.jl:184-242)
            self.QZ = ph.qz_vector(u, g, m, self.w0, self.Eu)
        else:
            self.QZ = np.zeros(g.Nt, dtype=complex)
        self.responses = resp_flags
        # adaptive step constants (AdaptiveSteps.jl:37-43)
        self.phik = ph.phi_kerr(u, m, self.w0, self.Eu) if p["n2"] != 0 else 0.0
        self.phip = ph.phi_plasma(u, m, self.w0, self.Eu, mr_dz, nuc_dz)
        # plasma (photoionization_avalanche.jl:1-116)
        self.pt = None
        if self.plasma:
            self.pt = ph.plasma_tables(u, self.w0, self.Eu, p["plasma_equation"], basedir)

    def _desc(self, geometry, precision, form, keep_plasma, device):
        d = L.FilDesc()
        d.abi_version = L.ABI_VERSION
        d.geometry = geometry
        d.Nt = self.grid.Nt
        d.precision = precision
        d.form = form
        d.nonlinearity = int(self.nonlinearity)
        d.plasma = int(self.plasma)
        d.responses = self.responses
        d.losses_ereal = self.losses_ereal
        if self.plasma:
            d.plasma_ereal = int(self.pt["EREAL"])
            d.plasma_kdep = int(self.pt["KDEP"])
            d.ncomp = len(self.pt["components"])
        d.keep_plasma = int(keep_plasma)
        d.device = device
        d.nonparaxial = int(getattr(self, "nonparaxial", False))
        return d

    def _create(self, d):
        self._lib = L.load()
        h = ctypes.c_void_p()
        L.check(self._lib.fil_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h

    def _common_tables(self, keep):
        """fil_tables fields the geometries share; ``keep`` holds the arrays alive."""
        g = self.grid

        def arr(a, cplx=False):
            a = np.ascontiguousarray(a, dtype=np.complex128 if cplx else np.float64)
            keep.append(a)
            return L.ptr(a.view(np.float64) if cplx else a)

        t = L.FilTables()
        t.t, t.dt = arr(g.t), g.dt
        t.Wguard = arr(self.guard.Wguard)
        t.Tguard = arr(self.guard.Tguard)
        t.QZ = arr(self.QZ, True)
        zero = np.zeros(g.Nt, dtype=complex)
        t.Rnl_cubic = arr(self.Rnl.get("cubic", zero), True)
        t.Rnl_free = arr(self.Rnl.get("free", zero), True)
        t.Rnl_losses = arr(self.Rnl.get("losses", zero), True)
        t.Rnl_raman = arr(self.Rnl.get("raman", zero), True)
        t.Hraman = arr(getattr(self, "Hraman", zero), True)
        if self.pt is not None:
            t.rho0 = self.pt["rho0"]
            for q, c in enumerate(self.pt["components"]):
                t.frhont[q], t.Rava[q], t.K[q] = c["frhont"], c["Rava"], c["K"]
                t.ntab[q] = len(c["log10x"])
                t.tab_log10x[q] = arr(c["log10x"])
                t.tab_log10y[q] = arr(c["log10y"])
        return t, arr

    # -- layout: rows x t on the device, the reference's array shape on the host
    def _to_rows(self, E):
        return np.asarray(E).reshape(self.nrows, self.grid.Nt, order="F")

    def _from_rows(self, A):
        return A

    # -- boundary ------------------------------------------------------------
    def upload(self, E):
        E = np.asfortranarray(self._to_rows(np.asarray(E, dtype=np.complex128)))
        L.check(self._lib.fil_upload_field(self.h, L.ptr(E.ravel(order="F").view(np.float64)), self.nrows))

    def field(self):
        """field.E after the last step (time domain, guarded), the reference's array shape."""
        n, Nt = self.nrows, self.grid.Nt
        out = np.zeros(n * Nt, dtype=np.complex128)
        L.check(self._lib.fil_download_field(self.h, L.ptr(out.view(np.float64)), n))
        return self._from_rows(out.reshape(Nt, n).T.copy())

    def plasma_arrays(self):
        n, Nt = self.nrows, self.grid.Nt
        rho = np.zeros(n * Nt)
        kd = np.zeros(n * Nt)
        L.check(self._lib.fil_download_plasma(self.h, L.ptr(rho), L.ptr(kd), n))
        return self._from_rows(rho.reshape(Nt, n).T.copy()), self._from_rows(kd.reshape(Nt, n).T.copy())

    def zstep(self, z, dz):
        L.check(self._lib.fil_zstep(self.h, float(z), float(dz)))

    def zstep_parts(self, z, dz_nonlinear, dz_linear):
        """The z-step with the nonlinear step over dz_nonlinear (0: none, and no
        plasma solve) and the linear step over dz_linear (fil_zstep_parts); the
        building block of symmetric (Strang) sequences of the two parts."""
        L.check(self._lib.fil_zstep_parts(self.h, float(z), float(dz_nonlinear), float(dz_linear)))

    def synchronize(self):
        L.check(self._lib.fil_synchronize(self.h))

    def dz_adaptive(self, dz0, phimax, Imax, rhomax):
        """AStepKerrPlasma (AdaptiveSteps.jl:96-107).  Julia's min propagates
        NaN: a NaN Imax or rhomax gives a NaN dz (and z), which ends the run's
        `while z < zmax`."""
        if not self.nonlinearity:
            return dz0
        if math.isnan(Imax) or math.isnan(rhomax):
            return math.nan
        with np.errstate(divide="ignore"):
            dzk = phimax / (self.phik * Imax) if self.phik * Imax != 0 else math.inf
            dzp = phimax / (self.phip * rhomax) if self.phip * rhomax != 0 else math.inf
        return min(dz0, dzk, dzp)

    def run_adaptive(self, z, zmax, dz0, phimax, Istop, max_steps, chunk=0):
        """Filament.run's z loop (Filament.jl:129-164) with the adaptive step
        (AdaptiveSteps.jl:96-107) kept on the device (fil_run_adaptive): z and
        dz in single precision, stop when z >= zmax, after a step whose Imax >
        Istop, or after ``max_steps`` steps.  Returns (z, nsteps, reason, log)
        with log[s] = (z, dz, Imax, rhomax) after step s (row 0: the start)."""
        p = L.FilAdaptParams(zmax=float(zmax), dz0=float(dz0), phimax=float(phimax), phik=float(self.phik),
                             phip=float(self.phip), Istop=float(Istop), nonlinearity=int(bool(self.nonlinearity)),
                             chunk=int(chunk))
        zz = np.array([float(z)])
        log = np.zeros((int(max_steps) + 1, 4))
        n, reason = ctypes.c_int(), ctypes.c_int()
        L.check(self._lib.fil_run_adaptive(self.h, ctypes.byref(p), L.ptr(zz), int(max_steps), L.ptr(log),
                                           ctypes.byref(n), ctypes.byref(reason)))
        return float(zz[0]), n.value, reason.value, log[:n.value + 1]

    def run_adaptive_noalloc(self, z, zmax, dz0, phimax, Istop, max_steps, chunk=0):
        """run_adaptive through fil_run_adaptive_noalloc: the same loop with its
        device buffers in the library's fixed scratch array (no device
        allocation; grids and step counts beyond it are refused)."""
        p = L.FilAdaptParams(zmax=float(zmax), dz0=float(dz0), phimax=float(phimax), phik=float(self.phik),
                             phip=float(self.phip), Istop=float(Istop), nonlinearity=int(bool(self.nonlinearity)),
                             chunk=int(chunk))
        zz = np.array([float(z)])
        log = np.zeros((max(int(max_steps), 0) + 1, 4))
        n, reason = ctypes.c_int(), ctypes.c_int()
        L.check(self._lib.fil_run_adaptive_noalloc(self.h, ctypes.byref(p), L.ptr(zz), int(max_steps), L.ptr(log),
                                                   ctypes.byref(n), ctypes.byref(reason)))
        return float(zz[0]), n.value, reason.value, log[:n.value + 1]

    def state_save(self, slot):
        L.check(self._lib.fil_state_save(self.h, int(slot)))

    def state_restore(self, slot):
        L.check(self._lib.fil_state_restore(self.h, int(slot)))

    def copy_state_from(self, other):
        """Take over the z-step state of another model on the same grid (device copy)."""
        L.check(self._lib.fil_state_copy(self.h, other.h))

    def fluence(self, slot=-1):
        """dt sum_t |E|^2 per row of the current state (slot < 0) or of a saved
        slot, by Parseval on the device, before the pending field filter."""
        Fr = np.zeros(self.nrows)
        L.check(self._lib.fil_fluence(self.h, int(slot), L.ptr(Fr)))
        return Fr

    def set_tiling(self, tiling):
        """MFMA tile shape of the Hankel products (lib.TILING_*, see
        fil_set_tiling); the result does not depend on it."""
        L.check(self._lib.fil_set_tiling(self.h, int(tiling)))

    @property
    def tiling(self):
        t = ctypes.c_int()
        L.check(self._lib.fil_get_tiling(self.h, ctypes.byref(t)))
        return t.value

    def timing(self, enable=True):
        """Per-kernel timing: False off, True every z-step, an integer n > 1
        every n-th z-step."""
        L.check(self._lib.fil_timing_enable(self.h, int(enable)))

    def timing_read(self):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_int64()
        L.check(self._lib.fil_timing_read(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                          ctypes.byref(n)))
        return dict(row_ms=a.value, dht1_ms=b.value, dht2_ms=c.value, count=n.value)

    def close(self):
        if getattr(self, "h", None):
            self._lib.fil_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RTModel(_Model):
    def __init__(self, p, precision=L.FP32, form=0, device=0, keep_plasma=False, basedir=".", KZ=None):
        """``KZ`` overrides the host-computed linear-propagator table (an
        [Nr, Nt] complex array), e.g. to drive the device with another
        implementation's table."""
        self.p = p
        self.unit = u = ph.UnitRT(p["ru"], p["zu"], p["tu"], p["Iu"], p["rhou"])
        self.grid = g = ph.GridRT(p["rmax"], p["Nr"], p["tmin"], p["tmax"], p["Nt"])
        self.nrows = g.Nr
        self.precision, self.form = precision, form
        self._physics(p, basedir)
        m = self.medium
        self.guard = ph.GuardRT(u, g, m, p["rguard"], p["tguard"], p["kguard"], p["wguard"])
        self.KZ = ph.kz_table(u, g, m, self.w0, bool(p["KPARAXIAL"])) if KZ is None else np.asarray(KZ)

        # initial field (Fields.jl:96-120): user IC on the host in FP64, analytic signal
        E0 = np.asarray(p["initial_condition"](g.r, g.t, u.r, u.t, u.I), dtype=float)
        self.E0 = ph.rsig2asig(E0)
        self.T, self.RJ, self.JV = ph.dht_plan(g)

        self._start(self._desc(L.GEOM_RT, precision, form, keep_plasma, device))

    def _start(self, d):
        d.Nr = self.grid.Nr
        self._create(d)
        self._set_tables()
        self.upload(self.E0)

    def reset_field(self, initial_condition):
        """A new pulse on the same grid, medium and tables: the field of
        ``initial_condition(r, t, ru, tu, Iu)`` (Fields.jl:96-120) replaces the
        state (the pending filter and correction are dropped, the plasma is
        solved again from it)."""
        g, u = self.grid, self.unit
        self.E0 = ph.rsig2asig(np.asarray(initial_condition(g.r, g.t, u.r, u.t, u.I), dtype=float))
        self.upload(self.E0)

    def _set_tables(self):
        g = self.grid
        keep = []
        t, arr = self._common_tables(keep)
        t.r, t.dr = arr(g.r), arr(g.dr)
        t.KZ = arr(self.KZ.T, True)              # column-major [Nr x Nt]
        t.Kguard = arr(self.guard.Kguard.T)
        t.Rguard = arr(self.guard.Rguard)
        t.dht_T, t.dht_RJ, t.dht_JV = arr(self.T.T), arr(self.RJ), arr(self.JV)
        if self.nonparaxial:
            QZk = ph.qz_table_rt(self.unit, g, self.medium, self.w0, self.Eu)
            t.QZ_kw = arr(QZk.T, True)         # column-major [Nr x Nt]
        L.check(self._lib.fil_set_tables(self.h, ctypes.byref(t)))

    def analyze(self, z=0.0, arrays=False):
        g = self.grid
        d = L.FilDiag()
        if arrays:
            Fr, Ft, re, S = np.zeros(g.Nr), np.zeros(g.Nt), np.zeros(g.Nr), np.zeros(g.Nt // 2 + 1)
            L.check(self._lib.fil_analyze(self.h, float(z), ctypes.byref(d), L.ptr(Fr), L.ptr(Ft),
                                          L.ptr(re), L.ptr(S)))
            out = d.as_dict()
            out.update(Fr=Fr, Ft=Ft, rho_end=re, S=S)
            return out
        L.check(self._lib.fil_analyze(self.h, float(z), ctypes.byref(d), None, None, None, None))
        return d.as_dict()


class RTEnsemble(RTModel):
    """B pulses on one RT grid advanced together by one handle (fil_desc.batch
    = B): the tables (grid, medium, guards, responses, plasma) are shared, each
    member has its own field and plasma.  Equivalent to B runs of Filament.run
    (src/Filament.jl:69-106) with the same input apart from the initial
    condition, stepped with a common dz.

    ``members``: parameter dicts that agree on everything but
    ``initial_condition`` (e.g. from workloads.rt_ensemble); the first one
    provides the shared tables."""

    def __init__(self, members, precision=L.FP32, form=0, device=0, basedir="."):
        self.B = len(members)
        if self.B < 1:
            raise ValueError("an ensemble needs at least one member")
        for q in members[1:]:
            for k in ("Nr", "Nt", "rmax", "tmin", "tmax", "lam0", "n2", "ru", "zu", "tu", "Iu", "rhou"):
                if q[k] != members[0][k]:
                    raise ValueError(f"ensemble members differ in {k}; group them by grid")
        self.members = members
        super().__init__(members[0], precision=precision, form=form, device=device, keep_plasma=False,
                         basedir=basedir)

    def _start(self, d):
        d.Nr = self.grid.Nr
        d.batch = self.B
        self._create(d)
        self._set_tables()
        g, u = self.grid, self.unit
        self.E0s = [ph.rsig2asig(np.asarray(q["initial_condition"](g.r, g.t, u.r, u.t, u.I), dtype=float))
                    for q in self.members]
        for b, E in enumerate(self.E0s):
            self.upload(E, b)

    def upload(self, E, member=0):
        E = np.asfortranarray(self._to_rows(np.asarray(E, dtype=np.complex128)))
        L.check(self._lib.fil_upload_member(self.h, int(member), L.ptr(E.ravel(order="F").view(np.float64)),
                                            self.nrows))

    def field(self, member=0):
        n, Nt = self.nrows, self.grid.Nt
        out = np.zeros(n * Nt, dtype=np.complex128)
        L.check(self._lib.fil_download_member(self.h, int(member), L.ptr(out.view(np.float64)), n))
        return out.reshape(Nt, n).T.copy()

    def plasma_arrays(self):
        raise NotImplementedError("plasma arrays are kept for single-pulse models")

    def fluence(self, slot=-1):
        """[B, Nr] fluence of every member (see _Model.fluence)."""
        Fr = np.zeros(self.B * self.nrows)
        L.check(self._lib.fil_fluence(self.h, int(slot), L.ptr(Fr)))
        return Fr.reshape(self.B, self.nrows)

    def analyze(self, z=0.0, arrays=False, member=0):
        g = self.grid
        d = L.FilDiag()
        if arrays:
            Fr, Ft, re, S = np.zeros(g.Nr), np.zeros(g.Nt), np.zeros(g.Nr), np.zeros(g.Nt // 2 + 1)
            L.check(self._lib.fil_analyze_member(self.h, int(member), float(z), ctypes.byref(d), L.ptr(Fr),
                                                 L.ptr(Ft), L.ptr(re), L.ptr(S)))
            out = d.as_dict()
            out.update(Fr=Fr, Ft=Ft, rho_end=re, S=S)
            return out
        L.check(self._lib.fil_analyze_member(self.h, int(member), float(z), ctypes.byref(d), None, None, None,
                                             None))
        return d.as_dict()

    def analyze_all(self, z=0.0):
        return [self.analyze(z, member=b) for b in range(self.B)]


class TModel(_Model):
    """The T geometry (Fields.jl:70-93, Guards.jl:78-89, spectral.jl:55-75,
    FieldAnalyzers.jl:74-114): one time row, the linear step a diagonal
    multiply by exp(i KZ dz) Wguard.  ``p`` carries ``zu, tu, Iu, rhou, tmin,
    tmax, Nt, tguard, wguard`` and ``initial_condition(t, tu, Iu)``."""

    def __init__(self, p, precision=L.FP64, form=0, device=0, keep_plasma=False, basedir=".", KZ=None):
        """``KZ`` overrides the host-computed KZ(w) table (Nt complex)."""
        if form & L.FORM_DHT_BF16X3:
            raise ValueError("the T geometry has no Hankel transform")
        self.p = p
        self.unit = u = ph.UnitT(p["zu"], p["tu"], p["Iu"], p["rhou"])
        self.grid = g = ph.GridT(p["tmin"], p["tmax"], p["Nt"])
        self.nrows = 1
        self.precision, self.form = precision, form
        self._physics(p, basedir)
        self.guard = ph.GuardT(u, g, p["tguard"], p["wguard"])
        self.KZ = ph.kz_table_t(u, g, self.medium, self.w0) if KZ is None else np.asarray(KZ).ravel()
        E0 = np.asarray(p["initial_condition"](g.t, u.t, u.I), dtype=float)
        self.E0 = ph.rsig2asig(E0[None, :])
        d = self._desc(L.GEOM_T, precision, form, keep_plasma, device)
        d.Nr = 1
        self._create(d)
        self._set_tables()
        self.upload(self.E0)

    def _from_rows(self, A):
        return A[0]

    def _to_rows(self, E):
        return np.asarray(E).reshape(1, self.grid.Nt)

    def _set_tables(self):
        keep = []
        t, arr = self._common_tables(keep)
        t.KZ = arr(self.KZ, True)
        L.check(self._lib.fil_set_tables(self.h, ctypes.byref(t)))

    def analyze(self, z=0.0, arrays=False):
        """FieldAnalyzerT: Imax, rhomax, duration = radius(t, I), F = sum(I) dt."""
        d = L.FilDiag()
        It = np.zeros(self.grid.Nt)
        L.check(self._lib.fil_analyze(self.h, float(z), ctypes.byref(d), None, L.ptr(It), None, None))
        out = dict(z=d.z, Imax=d.Imax, rhomax=d.rhomax, duration=d.tau, F=d.W)
        if arrays:
            out["I"] = It
        return out


class XYTModel(_Model):
    def __init__(self, p, precision=L.FP32, form=0, device=0, keep_plasma=False, basedir=".", guard3d_bug=False,
                 vf=None):
        """``guard3d_bug``: reproduce the reference's 3-D filter kernel, which
        takes the t (and w) factor at the y index (Guards.jl:314).  ``vf``
        overrides the moving-frame velocity of KZ (e.g. another
        implementation's value)."""
        if form & L.FORM_DHT_BF16X3:
            raise ValueError("the split-BF16 form applies to the RT Hankel transform only")
        self.p = p
        self.unit = u = ph.UnitXYT(p["xu"], p["yu"], p["zu"], p["tu"], p["Iu"], p["rhou"])
        self.grid = g = ph.GridXYT(p["xmin"], p["xmax"], p["Nx"], p["ymin"], p["ymax"], p["Ny"],
                                   p["tmin"], p["tmax"], p["Nt"])
        self.nrows = g.Nx * g.Ny
        self.precision, self.form = precision, form
        self.guard3d_bug = bool(guard3d_bug)
        self._physics(p, basedir)
        self.guard = ph.GuardXYT(u, g, p["xguard"], p["yguard"], p["tguard"], p["wguard"])
        self.kz = ph.kz_factors_xyt(u, g, self.medium, self.w0, vf)
        self.kparaxial = bool(p["KPARAXIAL"])

        # initial field (Fields.jl:141-160)
        E0 = np.asarray(p["initial_condition"](g.x, g.y, g.t, u.x, u.y, u.t, u.I), dtype=float)
        self.E0 = ph.rsig2asig(E0)

        d = self._desc(L.GEOM_XYT, precision, form, keep_plasma, device)
        d.Nx, d.Ny = g.Nx, g.Ny
        d.guard3d_bug = int(self.guard3d_bug)
        self._create(d)
        self._set_tables()
        self.upload(self.E0)

    def _from_rows(self, A):
        g = self.grid
        return A.reshape(g.Nx, g.Ny, g.Nt, order="F")

    def _set_tables(self):
        g = self.grid
        keep = []
        t, arr = self._common_tables(keep)
        x = L.FilTablesXYT()
        x.x, x.y, x.dx, x.dy = arr(g.x), arr(g.y), g.dx, g.dy
        x.kt2x, x.kt2y = arr(self.kz["kt2x"]), arr(self.kz["kt2y"])
        x.Xguard, x.Yguard = arr(self.guard.Xguard), arr(self.guard.Yguard)
        x.KXguard, x.KYguard = arr(self.guard.KXguard), arr(self.guard.KYguard)
        x.beta = arr(self.kz["beta"], True)
        x.w_vf = arr(self.kz["w_vf"])
        x.zu = self.unit.z
        x.kparaxial = int(self.kparaxial)
        L.check(self._lib.fil_set_tables_xyt(self.h, ctypes.byref(t), ctypes.byref(x)))

    def analyze(self, z=0.0, arrays=False):
        g = self.grid
        d = L.FilDiagXYT()
        if arrays:
            Fxy, Ft, re, S = np.zeros(self.nrows), np.zeros(g.Nt), np.zeros(self.nrows), np.zeros(g.Nt // 2 + 1)
            L.check(self._lib.fil_analyze_xyt(self.h, float(z), ctypes.byref(d), L.ptr(Fxy), L.ptr(Ft),
                                              L.ptr(re), L.ptr(S)))
            out = d.as_dict()
            out.update(Fxy=Fxy.reshape(g.Nx, g.Ny, order="F"), Ft=Ft, rho_end=re.reshape(g.Nx, g.Ny, order="F"),
                       S=S)
            return out
        L.check(self._lib.fil_analyze_xyt(self.h, float(z), ctypes.byref(d), None, None, None, None))
        return d.as_dict()


class _SpatialModel:
    """The spatial geometries R and XY (Input.jl:37-56, 131-150): a
    monochromatic beam E(r) or E(x, y), no time axis and no plasma.  Field,
    Guard, LinearPropagator and NonlinearPropagator are built on the host in
    FP64 (Fields.jl:31-51, 123-138; Guards.jl:41-57, 132-157; spectral.jl:30-50,
    104-128; nonlinear_propagators.jl:78-96, 162-180) and handed to a
    fil_spatial handle; ``zstep`` is Models.zstep for these geometries and
    ``analyze`` FieldAnalyzerR / FieldAnalyzerXY (FieldAnalyzers.jl:17-70,
    218-266)."""

    def _setup(self, p, geometry, precision, device, KZ=None):
        u, g = self.unit, self.grid
        self.p = p
        self.precision = precision
        self.medium = m = ph.Medium(p["permittivity"], p.get("permeability", ph.unit_permeability), p["n2"])
        self.w0 = 2 * math.pi * ph.C0 / p["lam0"]
        self.n0 = float(np.real(m.refractive_index(self.w0)))
        self.Eu = ph.field_unit(u, self.n0)
        self.nonlinearity = bool(p["NONLINEARITY"])
        self.nonparaxial = bool(self.nonlinearity and not p.get("QPARAXIAL", True))
        self.alg = alg_order(p.get("ALG"))
        # the responses with a spatial form: the cubic one (cubic.jl:18-19,
        # calc_cubic_nothg_spatial); several add their Rnl F terms
        self.Rnl = 0.0
        if self.nonlinearity:
            for item in p["responses"]:
                kind = response_kind(item["init"])
                if kind != "cubic":
                    raise NotImplementedError(f"response {kind} has no spatial form (geometry {geometry})")
                self.Rnl += float(np.real(ph.rnl_cubic(u, m, self.w0, self.Eu, 1, item.get("n2"))[0]))
        self.QZ = ph.qz_spatial(u, g, m, self.w0, self.Eu, paraxial=not self.nonparaxial)
        self.KZ = ph.kz_spatial(u, g, m, self.w0, bool(p["KPARAXIAL"])) if KZ is None else np.asarray(KZ)
        # AStepKerr (AdaptiveSteps.jl:29-34, 46-57)
        self.phik = ph.phi_kerr(u, m, self.w0, self.Eu) if p["n2"] != 0 else 0.0
        d = L.FilSpatialDesc()
        d.abi_version = L.ABI_VERSION
        d.geometry = geometry
        d.precision = precision
        d.nonlinearity = int(self.nonlinearity)
        d.nonparaxial = int(self.nonparaxial)
        d.alg = self.alg
        d.device = device
        return d

    def _create(self, d, t):
        self._lib = L.load()
        h = ctypes.c_void_p()
        L.check(self._lib.fil_spatial_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h
        L.check(self._lib.fil_spatial_set_tables(self.h, ctypes.byref(t)))

    @staticmethod
    def _arr(keep, a, cplx=False):
        a = np.ascontiguousarray(a, dtype=np.complex128 if cplx else np.float64)
        keep.append(a)
        return L.ptr(a.view(np.float64) if cplx else a)

    def upload(self, E):
        E = np.ascontiguousarray(np.asarray(E, dtype=np.complex128).ravel(order="F"))
        if E.size != self.n:
            raise ValueError(f"field has {E.size} points, the grid {self.n}")
        L.check(self._lib.fil_spatial_upload_field(self.h, L.ptr(E.view(np.float64))))

    def field(self):
        out = np.zeros(self.n, dtype=np.complex128)
        L.check(self._lib.fil_spatial_download_field(self.h, L.ptr(out.view(np.float64))))
        return out.reshape(self.shape, order="F")

    def zstep(self, z, dz):
        L.check(self._lib.fil_spatial_zstep(self.h, float(z), float(dz)))

    def synchronize(self):
        L.check(self._lib.fil_spatial_synchronize(self.h))

    def dz_adaptive(self, dz0, phimax, Imax, rhomax=0.0):
        """AStepKerr (AdaptiveSteps.jl:75-88): min(dz0, phimax / (phik Imax))
        with the nonlinearity, dz0 without; a NaN Imax gives a NaN dz."""
        if not self.nonlinearity:
            return dz0
        if math.isnan(Imax):
            return math.nan
        dzk = phimax / (self.phik * Imax) if self.phik * Imax != 0 else math.inf
        return min(dz0, dzk)

    def close(self):
        if getattr(self, "h", None):
            self._lib.fil_spatial_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RModel(_SpatialModel):
    """Geometry R (Input.jl:37-56): ``p`` with ru, zu, Iu, rmax, Nr, lam0,
    permittivity, permeability, n2, NONLINEARITY, KPARAXIAL, QPARAXIAL,
    rguard, kguard, responses, ALG and initial_condition(r, ru, Iu)."""

    def __init__(self, p, precision=L.FP32, device=0, basedir=".", KZ=None):
        """``KZ`` overrides the host-computed linear-propagator table ([Nr]
        complex), e.g. to drive the device with another implementation's."""
        self.unit = u = ph.UnitR(p["ru"], p["zu"], p["Iu"])
        self.grid = g = ph.GridR(p["rmax"], p["Nr"])
        self.n, self.shape = g.Nr, (g.Nr,)
        d = self._setup(p, L.GEOM_R, precision, device, KZ)
        d.Nr = g.Nr
        self.guard = ph.GuardR(u, g, self.medium, self.w0, p["rguard"], p["kguard"])
        self.T, self.RJ, self.JV = ph.dht_plan(g)
        self.E0 = np.asarray(p["initial_condition"](g.r, u.r, u.I), dtype=complex)
        keep = []
        a = lambda v, c=False: self._arr(keep, v, c)
        t = L.FilSpatialTables()
        t.KZ, t.QZ, t.Rnl = a(self.KZ, True), a(self.QZ, True), self.Rnl
        t.Kguard, t.Rguard = a(self.guard.Kguard), a(self.guard.Rguard)
        t.dht_T, t.dht_RJ, t.dht_JV = a(self.T.T), a(self.RJ), a(self.JV)
        t.r, t.dr = a(g.r), a(g.dr)
        self._create(d, t)
        self.upload(self.E0)

    def analyze(self, z=0.0, arrays=False, full=False):
        """arrays or full: also I = |E|^2 [Nr]."""
        d = L.FilDiagSpatial()
        arrays = arrays or full
        I = np.zeros(self.n) if arrays else None
        L.check(self._lib.fil_spatial_analyze(self.h, float(z), ctypes.byref(d), L.ptr(I), None, None))
        out = d.as_dict()
        if arrays:
            out["I"] = I
        return out


class XYModel(_SpatialModel):
    """Geometry XY (Input.jl:131-150): ``p`` with xu, yu, zu, Iu, xmin, xmax,
    Nx, ymin, ymax, Ny, lam0, medium entries, NONLINEARITY, KPARAXIAL,
    QPARAXIAL, xguard, yguard, kxguard, kyguard, responses, ALG and
    initial_condition(x, y, xu, yu, Iu) -> [Nx, Ny]."""

    def __init__(self, p, precision=L.FP32, device=0, basedir=".", KZ=None):
        """``KZ``: an [Nx, Ny] complex table in place of the host's."""
        self.unit = u = ph.UnitXY(p["xu"], p["yu"], p["zu"], p["Iu"])
        self.grid = g = ph.GridXY(p["xmin"], p["xmax"], p["Nx"], p["ymin"], p["ymax"], p["Ny"])
        self.n, self.shape = g.Nx * g.Ny, (g.Nx, g.Ny)
        d = self._setup(p, L.GEOM_XY, precision, device, KZ)
        d.Nx, d.Ny = g.Nx, g.Ny
        self.guard = ph.GuardXY(u, g, self.medium, self.w0, p["xguard"], p["yguard"], p["kxguard"], p["kyguard"])
        self.E0 = np.asarray(p["initial_condition"](g.x, g.y, u.x, u.y, u.I), dtype=complex)
        keep = []
        a = lambda v, c=False: self._arr(keep, v, c)
        t = L.FilSpatialTables()
        t.KZ, t.QZ, t.Rnl = a(self.KZ.ravel(order="F"), True), a(self.QZ.ravel(order="F"), True), self.Rnl
        t.KXguard, t.KYguard = a(self.guard.KXguard), a(self.guard.KYguard)
        t.Xguard, t.Yguard = a(self.guard.Xguard), a(self.guard.Yguard)
        t.x, t.y, t.dx, t.dy = a(g.x), a(g.y), g.dx, g.dy
        self._create(d, t)
        self.upload(self.E0)

    def analyze(self, z=0.0, arrays=False, full=False):
        """arrays: also the zdata lines Izx = I[:, Ny/2], Izy = I[Nx/2, :]
        (1-based, WritePlots.jl:448-456); full: also I [Nx, Ny]."""
        g = self.grid
        d = L.FilDiagSpatial()
        if arrays or full:
            I = np.zeros(self.n) if full else None
            Izx, Izy = np.zeros(g.Nx), np.zeros(g.Ny)
            L.check(self._lib.fil_spatial_analyze(self.h, float(z), ctypes.byref(d), L.ptr(I), L.ptr(Izx),
                                                  L.ptr(Izy)))
            out = d.as_dict()
            out.update(Izx=Izx, Izy=Izy)
            if full:
                out["I"] = I.reshape(g.Nx, g.Ny, order="F")
            return out
        L.check(self._lib.fil_spatial_analyze(self.h, float(z), ctypes.byref(d), None, None, None))
        return d.as_dict()


class StepGroup:
    """Models advanced together, one graph launch per z-step (fil_group_*):
    e.g. the per-grid RTEnsembles of an ensemble of pulses of uneven size.
    Close the group before its models."""

    def __init__(self, models):
        self._lib = L.load()
        self.models = list(models)
        arr = (ctypes.c_void_p * len(self.models))(*[m.h.value for m in self.models])
        g = ctypes.c_void_p()
        L.check(self._lib.fil_group_create(arr, len(self.models), ctypes.byref(g)))
        self.g = g

    def zstep(self, z, dz):
        L.check(self._lib.fil_group_zstep(self.g, ctypes.c_double(z), ctypes.c_double(dz)))

    def synchronize(self):
        for m in self.models:
            m.synchronize()

    def stats(self):
        c, r = ctypes.c_int64(), ctypes.c_int64()
        L.check(self._lib.fil_group_stats(self.g, ctypes.byref(c), ctypes.byref(r)))
        return dict(captures=c.value, replays=r.value)

    def close(self):
        if getattr(self, "g", None):
            self._lib.fil_group_destroy(self.g)
            self.g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
