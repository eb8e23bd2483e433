"""Host-side driver pieces that need no GPU: input-file loading (Input.prepare
semantics, src/Input.jl:20-200) for the RT, T and XYT geometries, and the
plot.dat format of each (WritePlots.jl:92-116 T, :121-152 RT, :187-215 XYT)."""
import os

import numpy as np
import pytest

import _inputs


@pytest.mark.parametrize("case", ["diffraction_rt", "dispersion_rt", "dispersion_t", "diffraction_xyt"])
def test_prepare_reads_generated_input_sets(pkg, case, tmp_path):
    c = _inputs.cases()[case]
    fname = _inputs.write_input_set(case, str(tmp_path / case))
    here = os.getcwd()
    try:
        p = pkg.driver.prepare(fname)
        assert os.getcwd() == str(tmp_path / case)       # Input.jl:23 cd side effect
    finally:
        os.chdir(here)
    assert p["geometry"] == c["geometry"]
    for k, v in {**c["grid"], **c["units"], **c["guards"]}.items():
        assert p[k] == v, k
    assert p["dz_initial"] == c["model"]["zmax"] / c["model"]["steps"]
    assert p["NONLINEARITY"] is False and p["responses"] == []
    assert abs(p["permittivity"](2.35e15) - 1.000275 ** 2) < 1e-6
    q = c["pulse"]
    if c["geometry"] == "RT":
        E = p["initial_condition"](np.array([0.0, 1.0]), np.array([0.0]), 1e-3, 1e-15, 1e16)
        assert E.shape == (2, 1) and abs(E[0, 0] - 1.0) < 1e-12
        assert abs(E[1, 0] - np.exp(-0.5 * (1e-3 / q["a0"]) ** 2)) < 1e-12
    elif c["geometry"] == "T":
        assert p["KPARAXIAL"] and p["QPARAXIAL"]             # Input.jl:96-103
        E = p["initial_condition"](np.array([0.0, 20.0]), 1e-15, 1e16)
        w0 = 2 * np.pi * 299792458.0 / q["lam0"]
        assert abs(E[0] - 1.0) < 1e-12
        assert abs(E[1] - np.exp(-0.5) * np.cos(w0 * 20e-15)) < 1e-12
    else:
        E = p["initial_condition"](np.array([0.0, 1.0]), np.array([0.0]), np.array([0.0]), 1e-3, 1e-3, 1e-15, 1e16)
        assert E.shape == (2, 1, 1) and abs(E[1, 0, 0] - np.exp(-0.5)) < 1e-12


def test_prepare_refuses_unknown_geometry_and_missing_keys(pkg, tmp_path):
    fname = _inputs.write_input_set("dispersion_t", str(tmp_path / "t"))
    src = open(fname).read()
    here = os.getcwd()
    try:
        open(fname, "w").write(src.replace("geometry = 'T'", "geometry = 'Rn'"))
        with pytest.raises(ValueError, match="not implemented"):
            pkg.driver.prepare(fname)
        open(fname, "w").write(src.replace("tguard = 10.0\n", ""))
        with pytest.raises(ValueError, match="tguard"):
            pkg.driver.prepare(fname)
    finally:
        os.chdir(here)


def test_plotdat_format(pkg, tmp_path):
    p = dict(ru=1e-3, zu=1.0, tu=1e-15, Iu=1e16, rhou=2.5e25)
    f = str(tmp_path / "plot.dat")
    pkg.driver.write_plotdat_header(f, p)
    row = dict(z=0.5, Fmax=1.0, Imax=2.0, rhomax=3.0, De=4.0, rfil=5.0, rpl=6.0, tau=7.0, W=8.0)
    pkg.driver.write_plotdat_row(f, row, p)
    lines = open(f).read().splitlines()
    assert lines[0].split()[1:] == ["z", "Fmax", "Imax", "Nemax", "De", "rfil", "rpl", "tau", "W"]
    assert lines[1].split()[1:] == ["m", "J/m^2", "W/m^2", "1/m^3", "1/m", "m", "m", "s", "J"]
    data = np.loadtxt(f, comments="#")
    assert np.allclose(data, [0.5, 1, 2, 3, 4, 5, 6, 7, 8])
    assert lines[3].startswith("  5.000000000000e-01 ")


def test_plotdat_columns_of_t_and_xyt(pkg, tmp_path):
    pt = dict(geometry="T", zu=1.0, tu=1e-15, Iu=1e16, rhou=2.5e25)
    f = str(tmp_path / "t.dat")
    pkg.driver.write_plotdat_header(f, pt)
    pkg.driver.write_plotdat_row(f, dict(z=1.0, Imax=2.0, rhomax=3.0, duration=4.0, F=5.0), pt)
    lines = open(f).read().splitlines()
    assert lines[0].split()[1:] == ["z", "Imax", "Nemax", "duration", "F"]
    assert [float(v) for v in lines[2].split()[1:]] == [1.0, 1e16, 2.5e25, 1e-15, 1e-15 * 1e16]
    assert np.allclose(np.loadtxt(f, comments="#"), [1, 2, 3, 4, 5])
    px = dict(geometry="XYT", xu=1e-3, yu=2e-3, zu=1.0, tu=1e-15, Iu=1e16, rhou=2.5e25)
    f = str(tmp_path / "xyt.dat")
    pkg.driver.write_plotdat_header(f, px)
    pkg.driver.write_plotdat_row(f, dict(z=1, Fmax=2, Imax=3, rhomax=4, De=5, ax=6, ay=7, W=8), px)
    lines = open(f).read().splitlines()
    assert lines[0].split()[1:] == ["z", "Fmax", "Imax", "Nemax", "De", "ax", "ay", "W"]
    assert lines[1].split()[1:] == ["m", "J/m^2", "W/m^2", "1/m^3", "1/m", "m", "m", "J"]
    assert float(lines[2].split()[5]) == 1e-3 * 2e-3 * 2.5e25 and float(lines[2].split()[7]) == 2e-3
    assert np.allclose(np.loadtxt(f, comments="#"), np.arange(1, 9))


def test_adaptive_rule_propagates_nan(pkg):
    """AStepKerrPlasma with a NaN analysis (AdaptiveSteps.jl:100-106): Julia's
    min returns NaN, so dz and z become NaN and `while z < zmax` ends the run
    (Filament.jl:129); a finite analysis gives the usual minimum."""
    import math

    class Rule:
        nonlinearity, phik, phip = True, 2.0, 3.0
    rule = pkg.solver._Model.dz_adaptive
    assert math.isnan(rule(Rule(), 0.1, 0.5, math.nan, 0.0))
    assert math.isnan(rule(Rule(), 0.1, 0.5, 1.0, math.nan))
    assert rule(Rule(), 0.1, 0.5, 1.0, 0.0) == 0.1
    assert rule(Rule(), 0.1, 0.5, 10.0, 0.0) == 0.5 / 20.0
    z = np.float32(0.0) + np.float32(math.nan)
    assert not (z < np.float32(1.0))


def test_plotfield_container_and_schedule(pkg, tmp_path):
    """plot.h5's write schedule (WritePlots.jl:263-283) and content in the
    plot.fdata container: a field dataset when z >= the next field distance
    (then + dz_plothdf) and z differs from the last zdata z; a zdata row when
    z >= the next zdata distance (then z + dz_zdata); Float32 distances."""
    class Stub:
        precision = 4

        def __init__(self):
            self.E = np.zeros((3, 8), dtype=complex)

        def field(self):
            return self.E

    class Grid:
        r = np.arange(3.0)
        t = np.arange(8.0)

    p = dict(geometry="RT", ru=1e-3, zu=1.0, tu=1e-15, Iu=1e16, rhou=2.5e25)
    path = str(tmp_path / "plot.fdata")
    pf = pkg.plotfield.PlotField(path, p, Grid(), 0.0, 0.25, 0.1, ftype=np.float32)
    m = Stub()
    zs = [0.0, 0.05, 0.1, 0.2, 0.26, 0.3, 0.55]
    for k, z in enumerate(zs):
        m.E = (k + 1) * (1.0 + 1j) * np.ones((3, 8))
        d = dict(Fr=np.full(3, k), rho_end=np.zeros(3), Ft=np.full(8, 2 * k), S=np.full(5, 3 * k))
        pf.write(m, d, z)
    meta = pkg.plotfield.meta(path)
    assert meta["version"] == "3.0" and meta["units"]["zu"] == 1.0 and meta["grid"] == {"r": 3, "t": 8}
    assert [(e["dset"], e["z"]) for e in meta["field"]] == [("000", 0.0), ("001", float(np.float32(0.26))),
                                                           ("002", float(np.float32(0.55)))]
    E, z = pkg.plotfield.read_field(path, "001")
    assert E.dtype == np.float32 and np.all(E == 5.0) and z == float(np.float32(0.26))
    # zdata at 0 (next 0.1), 0.1 (next 0.2), 0.2 (0.3), 0.3 (0.4), 0.55
    assert np.allclose(pkg.plotfield.read_zdata(path, "z")[:, 0], np.float32([0.0, 0.1, 0.2, 0.3, 0.55]))
    assert pkg.plotfield.read_zdata(path, "Fzt").shape == (5, 8)
    assert np.allclose(pkg.plotfield.read_zdata(path, "iSzf")[:, 0], [0, 6, 9, 15, 18])
    Er, zm = pkg.plotfield.restart(path, "002")
    assert np.all(Er == 7.0) and zm == float(np.float32(0.55))
