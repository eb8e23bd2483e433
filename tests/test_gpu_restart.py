"""Field snapshots and restart (WritePlots.jl:222-331, examples/RT/gauss_restart.jl:1-22)
through prepare-style parameters and driver.run on the GPU.

Run A goes from z = 0 to zmax writing a snapshot every dz_plothdf; run B
starts from A's middle snapshot (its real field as the initial condition, its
z as the start) and goes to the same zmax.  B's steps after the restart point
follow A's: the real field of an analytic signal gives the same analytic
signal back (A's field was projected by its analysis) except for the
imaginary part of its w = 0 bin, which the real field cannot carry and the
reference's restart loses as well (here ~5e-8 of the peak intensity), and
every z-step solves the plasma of the field it starts from.  The restart row's plasma columns differ
by design: the reference's analysis after a step reports the plasma of the
field at the step's start, a new model's that of its initial field
(Models.jl:66)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def params(pkg, tmp, name):
    p = pkg.workloads.rt_air(64, 512, 6.0, tmin=-125.0, tmax=125.0, tau0=20e-15, dz=5e-3)
    dz = p["dz_initial"]
    p.update(prefix=os.path.join(str(tmp), name) + "/", zmax=40 * dz, dz_plothdf=float(np.float32(20 * dz)),
             dzphimax=0.9 * dz * 1.0)
    return p


@pytest.mark.parametrize("precision,tol", [(4, 2e-6), (8, 1e-6)])
def test_restart_from_snapshot_continues_the_run(pkg, precision, tol, tmp_path):
    pa = params(pkg, tmp_path, "a")
    m = pkg.RTModel(pa, precision=precision)
    pa["dzphimax"] = 0.9 * pa["dz_initial"] * m.phik * m.analyze(0.0)["Imax"]    # Kerr limit active
    m.close()
    rows_a = pkg.driver.run(pa, precision=precision, verbose=False)
    fa = os.path.join(str(tmp_path), "a", "plot.fdata")
    meta = pkg.plotfield.meta(fa)
    assert [e["dset"] for e in meta["field"]] == ["000", "001", "002"]
    E1, z1 = pkg.plotfield.read_field(fa, "001")
    assert E1.shape == (64, 512) and E1.dtype == (np.float32 if precision == 4 else np.float64)
    E1m, z1m = pkg.plotfield.restart(fa, "001")
    assert z1m == z1 * pa["zu"] and np.array_equal(E1m, E1)
    # zdata: one row per step (dz_zdata = lam0/2 is below every dz)
    Fzr = pkg.plotfield.read_zdata(fa, "Fzr")
    zz = pkg.plotfield.read_zdata(fa, "z")[:, 0]
    assert Fzr.shape == (len(rows_a), 64) and np.allclose(zz, [r["z"] for r in rows_a])
    assert np.allclose(Fzr[-1], pkg.plotfield.read_zdata(fa, "Fzr")[-1])

    pb = params(pkg, tmp_path, "b")
    pb["dzphimax"] = pa["dzphimax"]
    pb["initial_condition"] = lambda r, t, ru, tu, Iu: E1m       # gauss_restart.jl: the stored field
    pb["z"] = z1m / pb["zu"]
    rows_b = pkg.driver.run(pb, precision=precision, verbose=False)
    k = [i for i, r in enumerate(rows_a) if r["z"] == z1][0]
    assert rows_b[0]["z"] == z1 and len(rows_b) == len(rows_a) - k
    for key in ("Fmax", "Imax", "W", "rfil", "tau"):          # field columns of the restart row
        assert abs(rows_b[0][key] - rows_a[k][key]) <= tol * abs(rows_a[k][key]) + 1e-300, key
    for ra, rb in zip(rows_a[k + 1:], rows_b[1:]):
        # dz follows Imax, so z may move by the last bit of the run's single-precision z
        assert abs(ra["z"] - rb["z"]) <= 4e-7 * ra["z"]
        for key in ("Fmax", "Imax", "rhomax", "De", "W"):
            assert abs(rb[key] - ra[key]) <= 10 * tol * abs(ra[key]) + 1e-300, (ra["z"], key)
    # the per-step profiles after the restart (zdata rows) equal run A's
    fb = os.path.join(str(tmp_path), "b", "plot.fdata")
    zb = pkg.plotfield.read_zdata(fb, "z")[:, 0]
    worst = 0.0
    for name in ("Fzr", "Fzt", "iSzf"):
        a, b = pkg.plotfield.read_zdata(fa, name)[k + 1:], pkg.plotfield.read_zdata(fb, name)[1:]
        assert a.shape == b.shape and np.allclose(zz[k + 1:], zb[1:], rtol=4e-7, atol=0)
        worst = max(worst, float(np.max(np.abs(b - a)) / np.max(np.abs(a))))
    print(f"restart fp{8 * precision}: {len(rows_b) - 1} steps after z = {z1:.5g}, zdata profiles max rel. diff "
          f"{worst:.3e}")
    assert worst <= max(10 * tol, 1e-6)                  # the zdata rows are float32
