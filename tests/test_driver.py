"""Host-side driver pieces that need no GPU: input-file loading (Input.prepare
semantics) and the plot.dat format (WritePlots.jl:30-59,121-152)."""
import os

import numpy as np

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_prepare_reads_reference_input_set(pkg):
    here = os.getcwd()
    try:
        p = pkg.driver.prepare(os.path.join(GOLD, "diffraction_rt", "input.py"))
        assert os.getcwd() == os.path.join(GOLD, "diffraction_rt")     # Input.jl:23 cd side effect
    finally:
        os.chdir(here)
    assert p["geometry"] == "RT" and p["Nr"] == 500 and p["Nt"] == 4096
    assert p["NONLINEARITY"] is False and p["responses"] == []
    assert abs(p["permittivity"](2.35e15) - 1.000275 ** 2) < 1e-6
    E = p["initial_condition"](np.array([0.0, 1.0]), np.array([0.0]), 1e-3, 1e-15, 1e16)
    assert E.shape == (2, 1) and abs(E[0, 0] - 1.0) < 1e-12 and abs(E[1, 0] - np.exp(-0.5)) < 1e-12


def test_plotdat_format(pkg, tmp_path):
    p = dict(ru=1e-3, zu=1.0, tu=1e-15, Iu=1e16, rhou=2.5e25)
    f = str(tmp_path / "plot.dat")
    pkg.driver.write_plotdat_header(f, p)
    row = dict(z=0.5, Fmax=1.0, Imax=2.0, rhomax=3.0, De=4.0, rfil=5.0, rpl=6.0, tau=7.0, W=8.0)
    pkg.driver.write_plotdat_row(f, row)
    lines = open(f).read().splitlines()
    assert lines[0].split()[1:] == ["z", "Fmax", "Imax", "Nemax", "De", "rfil", "rpl", "tau", "W"]
    assert lines[1].split()[1:] == ["m", "J/m^2", "W/m^2", "1/m^3", "1/m", "m", "m", "s", "J"]
    data = np.loadtxt(f, comments="#")
    assert np.allclose(data, [0.5, 1, 2, 3, 4, 5, 6, 7, 8])
    assert lines[3].startswith("  5.000000000000e-01 ")
