"""The reference's analytic physics tests restated through prepare/run on the
GPU: test/test_diffraction_rt.jl, test/test_dispersion_rt.jl and
test/test_dispersion_t.jl, test/test_diffraction_r.jl (its R part) and
test/test_diffraction_xy.jl with their input sets (written from
tests/golden/reference_inputs.json), plot.dat read back and compared with the
Gaussian-beam / Gaussian-pulse closed forms at the reference's tolerances;
and an XYT Gaussian beam (the closed form of test/test_diffraction_xy.jl) run
the same way."""
import json
import math
import os
import shutil

import numpy as np
import pytest

import _inputs

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _run(pkg, case, precision, tmp_path):
    here = os.getcwd()
    try:
        p = pkg.driver.prepare(_inputs.write_input_set(case, str(tmp_path / case)))
        pkg.driver.run(p, precision=precision, verbose=False)
        data = np.loadtxt(os.path.join("results", "plot.dat"), comments="#").T
        shutil.rmtree("results")
    finally:
        os.chdir(here)
    return data


def _gold():
    return json.load(open(os.path.join(GOLD, "reference", "constants.json")))


@pytest.mark.parametrize("precision", [4, 8])
def test_diffraction_rt(pkg, precision, tmp_path):
    gold = _gold()
    data = _run(pkg, "diffraction_rt", precision, tmp_path)
    z, I, a = data[0], data[2], data[5] * 1e-3
    lam0, a0 = 800e-9, 1e-3
    zd = 2 * math.pi / lam0 * a0 ** 2
    ath = a0 * np.sqrt(1 + (z / zd) ** 2)
    Ith = (a0 / ath) ** 2
    print(f"diffraction_rt precision={precision}: max|a-ath|={np.max(np.abs(a - ath)):.3e} m, "
          f"max|I-Ith|={np.max(np.abs(I - Ith)):.3e}")
    assert len(z) == 101
    assert np.max(np.abs(a - ath)) <= gold["diffraction_rt_radius_tol_m"]
    assert np.max(np.abs(I - Ith)) <= gold["diffraction_rt_intensity_tol"]


@pytest.mark.parametrize("precision", [4, 8])
def test_dispersion_rt(pkg, precision, tmp_path):
    gold = _gold()
    data = _run(pkg, "dispersion_rt", precision, tmp_path)
    z, I = data[0], data[2]
    tau0 = 20e-15
    zd = tau0 ** 2 / gold["k2_s2_per_m"]
    Ith = tau0 / (tau0 * np.sqrt(1 + (z / zd) ** 2))
    print(f"dispersion_rt precision={precision}: max|I-Ith|={np.max(np.abs(I - Ith)):.3e}")
    assert len(z) == 101
    assert np.max(np.abs(I - Ith)) <= gold["dispersion_rt_intensity_tol"]


@pytest.mark.parametrize("precision", [4, 8])
def test_dispersion_t(pkg, precision, tmp_path):
    """test/test_dispersion_t.jl: plot.dat columns z, Imax, Nemax, duration, F."""
    gold = _gold()
    data = _run(pkg, "dispersion_t", precision, tmp_path)
    z, I, tau = data[0], data[1], data[3] * 1e-15     # plot.dat columns in the unit set (tu = 1 fs)
    tau0 = 20e-15
    zd = tau0 ** 2 / gold["k2_s2_per_m"]
    tauth = tau0 * np.sqrt(1 + (z / zd) ** 2)
    Ith = tau0 / tauth
    print(f"dispersion_t precision={precision}: max|tau-tauth|={np.max(np.abs(tau - tauth)):.3e} s, "
          f"max|I-Ith|={np.max(np.abs(I - Ith)):.3e}")
    # z and dz are Float64 in the T geometry (Input.jl:82): 100 steps of 0.2 end a rounding short of zmax
    assert len(z) in (101, 102)
    assert np.max(np.abs(tau - tauth)) <= gold["dispersion_t_duration_tol_s"]
    assert np.max(np.abs(I - Ith)) <= gold["dispersion_rt_intensity_tol"]


@pytest.mark.parametrize("precision", [4, 8])
def test_diffraction_xyt(pkg, precision, tmp_path):
    """An XYT Gaussian beam through prepare/run: the 1/e radii ax, ay of the
    fluence follow a0 sqrt(1 + (z/zd)^2) to one grid step (the level crossing
    lands on grid nodes, test/test_diffraction_xy.jl's criterion at its grid
    step) and the peak intensity (a0/a)^2 tau0/tau(z) to 1e-3 (the 50 fs
    pulse also disperses in air: tau(z) = tau0 sqrt(1 + (z k2/tau0^2)^2)),
    sampled where the grid has its nodes nearest the axis, half a step off it
    in x and in y (an even number of nodes on a symmetric window)."""
    c = _inputs.cases()["diffraction_xyt"]
    data = _run(pkg, "diffraction_xyt", precision, tmp_path)
    xu, yu = c["units"]["xu"], c["units"]["yu"]
    z, I, ax, ay = data[0], data[2], data[5] * xu, data[6] * yu
    g = c["grid"]
    dx = (g["xmax"] - g["xmin"]) / (g["Nx"] - 1) * c["units"]["xu"]
    lam0, a0 = c["pulse"]["lam0"], c["pulse"]["a0"]
    zd = 2 * math.pi / lam0 * a0 ** 2
    ath = a0 * np.sqrt(1 + (z / zd) ** 2)
    tau0 = c["pulse"]["tau0"]
    dy = (g["ymax"] - g["ymin"]) / (g["Ny"] - 1) * c["units"]["yu"]
    Ith = ((a0 / ath) ** 2 * np.exp(-((dx / 2) ** 2 + (dy / 2) ** 2) / ath ** 2)
           / np.sqrt(1 + (z * _gold()["k2_s2_per_m"] / tau0 ** 2) ** 2))
    print(f"diffraction_xyt precision={precision}: max|ax-ath|={np.max(np.abs(ax - ath)):.3e} m, "
          f"max|ay-ath|={np.max(np.abs(ay - ath)):.3e} m (dx {dx:.3e}), max|I-Ith|={np.max(np.abs(I - Ith)):.3e}")
    assert len(z) == 101
    assert np.max(np.abs(ax - ath)) <= dx and np.max(np.abs(ay - ath)) <= dx
    assert np.max(np.abs(I - Ith)) <= 1e-3


@pytest.mark.parametrize("precision", [4, 8])
def test_diffraction_r(pkg, precision, tmp_path):
    """test/test_diffraction_r.jl, geometry R (Nr = 2000, 100 steps to 10 m):
    plot.dat columns z, Imax, rfil, P; the 1/e radius within 0.02 mm of
    a0 sqrt(1 + (z/zd)^2) and isapprox(I, (a0/a)^2, rtol=1e-3)."""
    data = _run(pkg, "diffraction_r", precision, tmp_path)
    z, I, a = data[0], data[1], data[2] * 1e-3
    lam0, a0 = 800e-9, 1e-3
    zd = 2 * math.pi / lam0 * a0 ** 2
    ath = a0 * np.sqrt(1 + (z / zd) ** 2)
    Ith = (a0 / ath) ** 2
    dI = np.linalg.norm(I - Ith) / max(np.linalg.norm(I), np.linalg.norm(Ith))
    print(f"diffraction_r precision={precision}: max|a-ath|={np.max(np.abs(a - ath)):.3e} m, "
          f"|I-Ith|/|I|={dI:.3e}")
    assert len(z) == 101
    assert np.max(np.abs(a - ath)) <= 0.02e-3
    assert dI <= 1e-3


@pytest.mark.parametrize("precision", [4, 8])
def test_diffraction_xy(pkg, precision, tmp_path):
    """test/test_diffraction_xy.jl (Nx = Ny = 2048, 100 steps to 10 m):
    plot.dat columns z, Imax, ax, ay, P; both radii within 0.01 mm of
    a0 sqrt(1 + (z/zd)^2), the intensity within 1e-3 of (a0/a)^2."""
    data = _run(pkg, "diffraction_xy", precision, tmp_path)
    z, I, ax, ay = data[0], data[1], data[2] * 1e-3, data[3] * 1e-3
    lam0, a0 = 800e-9, 1e-3
    zd = 2 * math.pi / lam0 * a0 ** 2
    ath = a0 * np.sqrt(1 + (z / zd) ** 2)
    Ith = (a0 / ath) ** 2
    print(f"diffraction_xy precision={precision}: max|ax-ath|={np.max(np.abs(ax - ath)):.3e} m, "
          f"max|ay-ath|={np.max(np.abs(ay - ath)):.3e} m, max|I-Ith|={np.max(np.abs(I - Ith)):.3e}")
    assert len(z) == 101
    assert np.max(np.abs(ax - ath)) <= 0.01e-3 and np.max(np.abs(ay - ath)) <= 0.01e-3
    assert np.max(np.abs(I - Ith)) <= 1e-3
