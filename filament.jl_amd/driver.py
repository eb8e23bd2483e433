"""Filament.prepare / Filament.run for the RT geometry, host side.

prepare(fname) follows src/Input.jl:20-200: the input file is executed (it
is Python here, with the reference's variable names), the process changes
directory to the input file's directory, then the initial-condition and
medium files are executed into the same namespace.  run(input) follows
src/Filament.jl:51-184: build the model, analyze, write plot.dat
(WritePlots.jl:121-152 format) and loop z-steps with the adaptive step and
the Istop check.  The z-step and the analyzer run in libfilament_hip.so.
"""
import math
import os
import time

import numpy as np

from . import lib as L
from . import physics as ph
from . import responses as R
from .solver import RTModel


def _namespace():
    ns = dict(C0=ph.C0, EPS0=ph.EPS0, MU0=ph.MU0, QE=ph.QE, ME=ph.ME, HBAR=ph.HBAR, pi=math.pi,
              math=math, np=np, RK3=R.RK3, Inf=math.inf,
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


def prepare(fname):
    """Input.prepare (src/Input.jl:20-200), RT branch."""
    file_input = os.path.abspath(fname)
    ns = _namespace()
    exec(compile(open(file_input).read(), file_input, "exec"), ns)
    os.chdir(os.path.dirname(file_input))            # Input.jl:23, a process-wide side effect
    for key in ("file_initial_condition", "file_medium"):
        path = os.path.abspath(ns[key])
        exec(compile(open(path).read(), path, "exec"), ns)
    geometry = ns["geometry"]
    if geometry != "RT":
        raise ValueError(f"geometry {geometry!r} is not implemented (RT only)")
    p = {k: ns[k] for k in (
        "prefix", "ru", "zu", "tu", "Iu", "rhou", "rmax", "Nr", "tmin", "tmax", "Nt", "zmax", "dz_initial",
        "NONLINEARITY", "PLASMA", "KPARAXIAL", "QPARAXIAL", "rguard", "tguard", "kguard", "wguard",
        "dzphimax", "Istop", "z", "lam0", "permittivity", "permeability", "n2", "initial_condition")}
    p["geometry"] = geometry
    p["nuc"] = ns.get("nuc", 0.0)
    p["mr"] = ns.get("mr", 1.0)
    p["responses"] = ns.get("responses", []) if p["NONLINEARITY"] else []
    p["plasma_equation"] = ns.get("plasma_equation", {}) if p["PLASMA"] else {}
    p["basedir"] = os.getcwd()
    p["z"] = p["z"] / p["zu"]
    return p


# plot.dat columns for RT (WritePlots.jl:121-152)
def _plotvars(p):
    ru, zu, tu, Iu, rhou = p["ru"], p["zu"], p["tu"], p["Iu"], p["rhou"]
    return [("z", "m", zu), ("Fmax", "J/m^2", tu * Iu), ("Imax", "W/m^2", Iu), ("Nemax", "1/m^3", rhou),
            ("De", "1/m", ru ** 2 * rhou), ("rfil", "m", ru), ("rpl", "m", ru), ("tau", "s", tu),
            ("W", "J", ru ** 2 * tu * Iu)]


def _fmt(x):
    return "%18.12e" % float(x)


def write_plotdat_header(fname, p):
    pv = _plotvars(p)
    with open(fname, "w") as f:
        f.write("#" + "".join(" %-18s" % n for n, _, _ in pv) + "\n")
        f.write("#" + "".join(" %-18s" % s for _, s, _ in pv) + "\n")
        f.write("#" + "".join(" %-18s" % repr(float(u)) for _, _, u in pv) + "\n")


def write_plotdat_row(fname, d):
    with open(fname, "a") as f:
        f.write("  " + "".join(_fmt(d[k]) + " " for k in ("z", "Fmax", "Imax", "rhomax", "De", "rfil", "rpl",
                                                          "tau", "W")) + "\n")


def run(p, precision=L.FP32, form=0, device=0, verbose=True):
    """Filament.run (src/Filament.jl:51-184) without the HDF5 and info.txt
    outputs.  Returns the list of diagnostics rows."""
    prefix = p["prefix"]
    prefix_dir, prefix_name = os.path.dirname(prefix), os.path.basename(prefix)
    if prefix_dir:
        os.makedirs(prefix_dir, exist_ok=True)
    model = RTModel(p, precision=precision, form=form, device=device, basedir=p.get("basedir", "."))
    # the reference keeps z, zmax and dz of the RT geometry in FloatGPU = Float32
    # (Input.jl:107-108, AdaptiveSteps.jl:106), which fixes the number of steps
    f32 = np.float32
    z = f32(p["z"])
    d = model.analyze(float(z))
    fplot = os.path.join(prefix_dir, prefix_name + "plot.dat")
    write_plotdat_header(fplot, p)
    write_plotdat_row(fplot, d)
    rows = [d]
    zmax = f32(p["zmax"])
    t0 = time.time()
    try:
        while z < zmax:
            if verbose:
                print(f"z={_fmt(z)}[zu] I={_fmt(d['Imax'])}[Iu] rho={_fmt(d['rhomax'])}[rhou]")
            dz = f32(model.dz_adaptive(p["dz_initial"], p["dzphimax"], d["Imax"], d["rhomax"]))
            z = f32(z + dz)
            model.zstep(float(z), float(dz))
            d = model.analyze(float(z))
            write_plotdat_row(fplot, d)
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
