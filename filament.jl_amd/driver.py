"""Filament.prepare / Filament.run for the RT, XYT, T, R and XY geometries, host side.

prepare(fname) follows src/Input.jl:20-200: the input file is executed (it
is Python here, with the reference's variable names), the process changes
directory to the input file's directory, then the initial-condition and
medium files are executed into the same namespace; the geometry's branch
(R :37-56, T :81-105, RT :106-130, XY :131-150, XYT :151-178) picks the keys
the run needs.
run(input) follows src/Filament.jl:51-184: build the model, analyze, write
plot.dat (WritePlots.jl:65-87 R, :92-116 T, :121-152 RT, :158-182 XY,
:187-215 XYT) and loop
z-steps with the adaptive step and the Istop check.  The z-step and the
analyzer run in libfilament_hip.so.
"""
import math
import os
import time

import numpy as np

from . import lib as L
from . import physics as ph
from . import responses as R
from .responses import alg_order
from .plotfield import PlotField
from .solver import RModel, RTModel, TModel, XYModel, XYTModel


def _namespace():
    ns = dict(C0=ph.C0, EPS0=ph.EPS0, MU0=ph.MU0, QE=ph.QE, ME=ph.ME, HBAR=ph.HBAR, pi=math.pi,
              math=math, np=np, RK3=R.RK3, RK4=R.RK4, Inf=math.inf,
              init_cubic=R.init_cubic, init_current_free=R.init_current_free,
              init_current_losses=R.init_current_losses, init_raman=R.init_raman,
              init_photoionization_avalanche=R.init_photoionization_avalanche,
              DEFPATHNR="", DEFPATHPE="")
    ns["include"] = lambda *a, **k: None   # response plugins are built in
    ns["FFTW"] = _FFTWShim()
    return ns


class _FFTWShim:
    """The few FFTW calls that initial-condition files make (fftfreq, fft, ifft)."""
    fftfreq = staticmethod(lambda n, fs=1.0: np.fft.fftfreq(n, 1.0 / fs))
    fft = staticmethod(np.fft.fft)
    ifft = staticmethod(np.fft.ifft)


_COMMON_KEYS = ("prefix", "zu", "tu", "Iu", "rhou", "tmin", "tmax", "Nt", "zmax", "dz_initial", "NONLINEARITY",
                "PLASMA", "tguard", "wguard", "dzphimax", "Istop", "z", "lam0", "permittivity", "permeability", "n2",
                "initial_condition")
# R and XY (Input.jl:37-56, 131-150): no t grid, no plasma
_SPATIAL_COMMON_KEYS = ("prefix", "zu", "Iu", "zmax", "dz_initial", "NONLINEARITY", "dzphimax", "Istop", "z", "lam0",
                        "permittivity", "permeability", "n2", "initial_condition")
_SPATIAL = ("R", "XY")
_GEOMETRY_KEYS = {
    "R": ("ru", "rmax", "Nr", "KPARAXIAL", "QPARAXIAL", "rguard", "kguard"),
    "XY": ("xu", "yu", "xmin", "xmax", "Nx", "ymin", "ymax", "Ny", "KPARAXIAL", "QPARAXIAL", "xguard", "yguard",
           "kxguard", "kyguard"),
    "RT": ("ru", "rmax", "Nr", "KPARAXIAL", "QPARAXIAL", "rguard", "kguard"),
    "XYT": ("xu", "yu", "xmin", "xmax", "Nx", "ymin", "ymax", "Ny", "KPARAXIAL", "QPARAXIAL", "xguard", "yguard",
            "kxguard", "kyguard"),
    "T": (),
}


def prepare(fname):
    """Input.prepare (src/Input.jl:20-200), RT, XYT and T branches."""
    file_input = os.path.abspath(fname)
    ns = _namespace()
    exec(compile(open(file_input).read(), file_input, "exec"), ns)
    os.chdir(os.path.dirname(file_input))            # Input.jl:23, a process-wide side effect
    for key in ("file_initial_condition", "file_medium"):
        path = os.path.abspath(ns[key])
        exec(compile(open(path).read(), path, "exec"), ns)
    geometry = ns["geometry"]
    if geometry not in _GEOMETRY_KEYS:
        raise ValueError(f"geometry {geometry!r} is not implemented (R, T, RT, XY, XYT)")
    keys = (_SPATIAL_COMMON_KEYS if geometry in _SPATIAL else _COMMON_KEYS) + _GEOMETRY_KEYS[geometry]
    missing = [k for k in keys if k not in ns]
    if missing:
        raise ValueError(f"input for geometry {geometry} lacks {', '.join(missing)}")
    p = {k: ns[k] for k in keys}
    p["geometry"] = geometry
    p["ALG"] = ns.get("ALG")
    if geometry in _SPATIAL:
        p["PLASMA"] = False          # Input.jl:51, :145
        if ns.get("HTLOAD", False):
            raise NotImplementedError("HTLOAD: the Hankel plan is computed, not loaded from a JLD2 file")
    if geometry == "T":
        # Input.jl:96-103: the T model is paraxial in both terms
        p["KPARAXIAL"] = p["QPARAXIAL"] = True
    if "dz_plothdf" in ns:
        p["dz_plothdf"] = ns["dz_plothdf"] if geometry == "T" else float(np.float32(ns["dz_plothdf"]))
    p["nuc"] = ns.get("nuc", 0.0)
    p["mr"] = ns.get("mr", 1.0)
    p["responses"] = ns.get("responses", []) if p["NONLINEARITY"] else []
    p["plasma_equation"] = ns.get("plasma_equation", {}) if p["PLASMA"] else {}
    p["basedir"] = os.getcwd()
    p["z"] = p["z"] / p["zu"]
    return p


# plot.dat columns: (name, SI unit, unit factor, analyzer key) per geometry
# (WritePlots.jl:92-116 T, :121-152 RT, :187-215 XYT)
def _plotvars(p):
    g = p.get("geometry", "RT")
    zu, Iu = p["zu"], p["Iu"]
    if g == "R":
        ru = p["ru"]
        return [("z", "m", zu, "z"), ("Imax", "W/m^2", Iu, "Imax"), ("rfil", "m", ru, "rfil"),
                ("P", "W", ru ** 2 * Iu, "P")]
    if g == "XY":
        xu, yu = p["xu"], p["yu"]
        return [("z", "m", zu, "z"), ("Imax", "W/m^2", Iu, "Imax"), ("ax", "m", xu, "ax"), ("ay", "m", yu, "ay"),
                ("P", "W", xu * yu * Iu, "P")]
    tu, rhou = p["tu"], p["rhou"]
    if g == "T":
        return [("z", "m", zu, "z"), ("Imax", "W/m^2", Iu, "Imax"), ("Nemax", "1/m^3", rhou, "rhomax"),
                ("duration", "s", tu, "duration"), ("F", "J/m^2", tu * Iu, "F")]
    if g == "XYT":
        xu, yu = p["xu"], p["yu"]
        return [("z", "m", zu, "z"), ("Fmax", "J/m^2", tu * Iu, "Fmax"), ("Imax", "W/m^2", Iu, "Imax"),
                ("Nemax", "1/m^3", rhou, "rhomax"), ("De", "1/m", xu * yu * rhou, "De"), ("ax", "m", xu, "ax"),
                ("ay", "m", yu, "ay"), ("W", "J", xu * yu * tu * Iu, "W")]
    ru = p["ru"]
    return [("z", "m", zu, "z"), ("Fmax", "J/m^2", tu * Iu, "Fmax"), ("Imax", "W/m^2", Iu, "Imax"),
            ("Nemax", "1/m^3", rhou, "rhomax"), ("De", "1/m", ru ** 2 * rhou, "De"), ("rfil", "m", ru, "rfil"),
            ("rpl", "m", ru, "rpl"), ("tau", "s", tu, "tau"), ("W", "J", ru ** 2 * tu * Iu, "W")]


def _fmt(x):
    return "%18.12e" % float(x)


def write_plotdat_header(fname, p):
    pv = _plotvars(p)
    with open(fname, "w") as f:
        f.write("#" + "".join(" %-18s" % v[0] for v in pv) + "\n")
        f.write("#" + "".join(" %-18s" % v[1] for v in pv) + "\n")
        f.write("#" + "".join(" %-18s" % repr(float(v[2])) for v in pv) + "\n")


def write_plotdat_row(fname, d, p=None):
    keys = [v[3] for v in _plotvars(p or {"zu": 1, "tu": 1, "Iu": 1, "rhou": 1, "ru": 1})]
    with open(fname, "a") as f:
        f.write("  " + "".join(_fmt(d[k]) + " " for k in keys) + "\n")


def make_model(p, precision=L.FP32, form=0, device=0):
    """The geometry's model (Filament.jl:69-76)."""
    g = p.get("geometry", "RT")
    if g in _SPATIAL:
        if form:
            raise ValueError("the chosen forms apply to the RT geometry")
        cls = RModel if g == "R" else XYModel
        return cls(p, precision=precision, device=device, basedir=p.get("basedir", "."))
    if alg_order(p.get("ALG")) != 3:
        raise NotImplementedError(f"geometry {g} integrates RK3 only")
    kw = dict(precision=precision, form=form, device=device, basedir=p.get("basedir", "."))
    if g == "RT":
        return RTModel(p, **kw)
    if g == "XYT":
        return XYTModel(p, **kw)
    if g == "T":
        return TModel(p, **kw)
    raise ValueError(f"geometry {g!r} is not implemented")


def run(p, precision=L.FP32, form=0, device=0, verbose=True, plotfield=True):
    """Filament.run (src/Filament.jl:51-184) without info.txt: plot.dat, and
    with ``plotfield`` the field snapshots every dz_plothdf and the per-step
    analyzer profiles of plot.h5 in the ``plot.fdata`` container
    (filament_amd.plotfield).  Returns the list of diagnostics rows."""
    prefix = p["prefix"]
    prefix_dir, prefix_name = os.path.dirname(prefix), os.path.basename(prefix)
    if prefix_dir:
        os.makedirs(prefix_dir, exist_ok=True)
    geometry = p.get("geometry", "RT")
    model = make_model(p, precision=precision, form=form, device=device)
    # the reference keeps z, zmax and dz of the RT and XYT geometries in FloatGPU = Float32
    # (Input.jl:107-108, :152-153; AdaptiveSteps.jl:106), the T geometry's in Float64 (Input.jl:82-83),
    # which fixes the number of steps
    ft = np.float64 if geometry == "T" else np.float32
    z = ft(p["z"])
    arrays = plotfield and geometry in ("RT", "XYT", "XY")
    d = model.analyze(float(z), arrays=arrays)
    fplot = os.path.join(prefix_dir, prefix_name + "plot.dat")
    write_plotdat_header(fplot, p)
    write_plotdat_row(fplot, d, p)
    pf = None
    if plotfield:
        # plot.h5's content and schedule (Filament.jl:93-98; dz_zdata = lam0 / 2 / zu)
        pf = PlotField(os.path.join(prefix_dir, prefix_name + "plot.fdata"), p, model.grid, z,
                       p.get("dz_plothdf", math.inf), p["lam0"] / 2 / p["zu"], ftype=ft)
        pf.write(model, d, z)
    rows = [d]
    zmax = ft(p["zmax"])
    t0 = time.time()
    try:
        while z < zmax:
            if verbose and geometry in _SPATIAL:
                print(f"z={_fmt(z)}[zu] I={_fmt(d['Imax'])}[Iu]")
            elif verbose:
                print(f"z={_fmt(z)}[zu] I={_fmt(d['Imax'])}[Iu] rho={_fmt(d['rhomax'])}[rhou]")
            dz = ft(model.dz_adaptive(p["dz_initial"], p["dzphimax"], d["Imax"], d.get("rhomax", 0.0)))
            z = ft(z + dz)
            model.zstep(float(z), float(dz))
            d = model.analyze(float(z), arrays=arrays)
            write_plotdat_row(fplot, d, p)
            if pf is not None:
                pf.write(model, d, z)
            rows.append(d)
            if d["Imax"] > p["Istop"]:
                if verbose:
                    print(f"Stop (Imax >= Istop): z={z}[zu], z={z * p['zu']}[m]")
                break
    finally:
        model.close()
    if verbose:
        print(f"Run time: {time.time() - t0:.2f} s")
    return rows
