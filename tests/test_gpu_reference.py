"""The reference's analytic physics tests (test/test_diffraction_rt.jl,
test/test_dispersion_rt.jl) restated: same input sets (tests/golden/*_rt/),
run through prepare/run on the GPU, plot.dat read back and compared with the
Gaussian-beam / Gaussian-pulse closed forms at the reference's tolerances."""
import json
import math
import os
import shutil

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _run(pkg, case, precision):
    here = os.getcwd()
    try:
        p = pkg.driver.prepare(os.path.join(GOLD, case, "input.py"))
        pkg.driver.run(p, precision=precision, verbose=False)
        data = np.loadtxt(os.path.join("results", "plot.dat"), comments="#").T
        shutil.rmtree("results")
    finally:
        os.chdir(here)
    return data


@pytest.mark.parametrize("precision", [4, 8])
def test_diffraction_rt(pkg, precision):
    gold = json.load(open(os.path.join(GOLD, "reference", "constants.json")))
    data = _run(pkg, "diffraction_rt", precision)
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
def test_dispersion_rt(pkg, precision):
    gold = json.load(open(os.path.join(GOLD, "reference", "constants.json")))
    data = _run(pkg, "dispersion_rt", precision)
    z, I = data[0], data[2]
    tau0 = 20e-15
    zd = tau0 ** 2 / gold["k2_s2_per_m"]
    Ith = tau0 / (tau0 * np.sqrt(1 + (z / zd) ** 2))
    print(f"dispersion_rt precision={precision}: max|I-Ith|={np.max(np.abs(I - Ith)):.3e}")
    assert np.max(np.abs(I - Ith)) <= gold["dispersion_rt_intensity_tol"]
