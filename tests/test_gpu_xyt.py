"""Parity of the XYT geometry on the device against the XYT oracle
(tests/test_oracle_xyt.py pins it), through the C ABI.

Gates (relative L2 field error after the stated steps): FP32 <= 1e-4 and
FP64 <= 1e-10 when the device is given the oracle's frame velocity (its
5-point derivative rounds differently in two FP64 implementations, which is
a ~1e-6 rad/m phase and nothing else), FP64 with the host's own tables
<= 1e-5.  Also: the reference's 3-D filter quirk, the analyzer, and the
512^3 case (energy bookkeeping and timing)."""
import time

import numpy as np
import pytest

import _helpers as H

pytestmark = pytest.mark.gpu


def small(pkg, nonlinearity=True, plasma=True, power=3.0, N=32, Nt=256, raman=False):
    return pkg.workloads.xyt_air(N, N, Nt, power, xmax=4.0, ymax=4.0, tmin=-120.0, tmax=120.0, ax=0.8e-3,
                                 ay=0.6e-3, tau0=20e-15, noise=0.05, nonlinearity=nonlinearity, plasma=plasma,
                                 raman=raman)


def oracle_for(pkg, orc, p, bug=False):
    g = pkg.physics.GridXYT(p["xmin"], p["xmax"], p["Nx"], p["ymin"], p["ymax"], p["Ny"], p["tmin"], p["tmax"], p["Nt"])
    E0 = p["initial_condition"](g.x, g.y, g.t, p["xu"], p["yu"], p["tu"], p["Iu"])
    return orc.Oracle(H.oracle_config_xyt(p, bug), H.xyt_rows(E0), 8)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def run(pkg, p, nsteps, dz, precision, vf=None, bug=False):
    m = pkg.XYTModel(p, precision=precision, guard3d_bug=bug, vf=vf, keep_plasma=True)
    z = 0.0
    for _ in range(nsteps):
        z += dz
        m.zstep(z, dz)
        m.analyze(z)          # the reference's main loop analyzes (and projects) after every step
    return m


@pytest.mark.parametrize("precision,tol,oracle_vf", [(4, 1e-4, False), (8, 1e-5, False), (8, 1e-10, True)])
@pytest.mark.parametrize("nonlinear", [False, True])
def test_xyt_steps_match_oracle(pkg, orc, precision, tol, oracle_vf, nonlinear):
    p = small(pkg, nonlinearity=nonlinear, plasma=nonlinear)
    o = oracle_for(pkg, orc, p)
    dz = 2e-4
    o.run(6, dz)
    vf = o.scalars()["vf"] if oracle_vf else None
    m = run(pkg, p, 6, dz, precision, vf=vf)
    err = rel(H.xyt_rows(m.field()), o.field())
    print(f"xyt nonlinear={nonlinear} precision={precision} oracle_vf={oracle_vf}: rel {err:.3e}")
    assert err < tol


def test_xyt_reference_filter_quirk_matches_oracle(pkg, orc):
    """guard3d_bug: the reference's kernel! takes the t and w factors at the y
    index (Guards.jl:314); the device reproduces the oracle's restatement."""
    p = small(pkg)
    o = oracle_for(pkg, orc, p, bug=True)
    dz = 2e-4
    o.run(4, dz)
    m = run(pkg, p, 4, dz, 8, vf=o.scalars()["vf"], bug=True)
    err = rel(H.xyt_rows(m.field()), o.field())
    print(f"xyt guard3d_bug: rel {err:.3e}")
    assert err < 1e-10
    # and it is a different field from the intended filter's
    m2 = run(pkg, p, 4, dz, 8, vf=o.scalars()["vf"], bug=False)
    assert rel(m2.field(), m.field()) > 1e-6


def test_xyt_analyzer_matches_oracle(pkg, orc):
    p = small(pkg, power=6.0)
    o = oracle_for(pkg, orc, p)
    dz = 2e-4
    o.run(3, dz)
    m = run(pkg, p, 3, dz, 8, vf=o.scalars()["vf"])
    d = m.analyze(3 * dz, arrays=True)
    od = o.analyze(3 * dz)
    for k, ko in (("Fmax", "Fmax"), ("Imax", "Imax"), ("rhomax", "rhomax"), ("De", "De"), ("W", "W"),
                  ("ax", "rfil"), ("ay", "rpl")):
        assert abs(d[k] - od[ko]) <= 1e-9 * max(1e-300, abs(od[ko])), (k, d[k], od[ko])
    assert rel(H.xyt_rows(d["Fxy"][:, :, None])[:, 0], od["Fr"]) < 1e-10
    assert rel(d["Ft"], od["Ft"]) < 1e-10 and rel(d["S"], od["S"]) < 1e-9


def test_xyt_512_cube_energy_and_timing(pkg):
    """The 512^3 case (30 Pcr elliptical beam, 5% smoothed noise): energy is
    conserved by the linear step up to the guards and lost only slowly to the
    plasma; prints the z-step time."""
    p = pkg.workloads.xyt_512()
    lin = dict(p, NONLINEARITY=False, PLASMA=False)
    m = pkg.XYTModel(lin, precision=pkg.lib.FP32)
    W0 = m.analyze(0.0)["W"]
    dz = p["dz_initial"]
    for k in range(3):
        m.zstep((k + 1) * dz, dz)
    W1 = m.analyze(3 * dz)["W"]
    m.close()
    print(f"512^3 linear: W {W0:.6e} -> {W1:.6e}")
    assert abs(W1 / W0 - 1) < 1e-4
    m = pkg.XYTModel(p, precision=pkg.lib.FP32)
    d0 = m.analyze(0.0)
    z = 0.0
    for _ in range(2):
        z += dz
        m.zstep(z, dz)
    m.synchronize()
    m.timing(True)
    t0 = time.perf_counter()
    n = 10
    for _ in range(n):
        z += dz
        m.zstep(z, dz)
    m.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / n
    kt = m.timing_read()
    d = m.analyze(z)
    print(f"512^3 full physics: {ms:.2f} ms/step (row kernel {kt['row_ms'] / kt['count']:.2f} ms, "
          f"xy transforms {kt['dht1_ms'] / kt['count']:.2f} ms); W {d0['W']:.6e} -> {d['W']:.6e}, "
          f"Imax {d0['Imax']:.4f} -> {d['Imax']:.4f}")
    assert np.isfinite(d["W"]) and 0.99 < d["W"] / d0["W"] <= 1.0 + 1e-6


@pytest.mark.parametrize("precision,tol,oracle_vf", [(4, 1e-4, False), (8, 1e-10, True)])
def test_xyt_raman_steps_match_oracle(pkg, orc, precision, tol, oracle_vf):
    """The XYT row kernel with the stimulated Raman response."""
    p = small(pkg, power=6.0, raman=True)
    o = oracle_for(pkg, orc, p)
    dz = 2e-4
    o.run(6, dz)
    m = run(pkg, p, 6, dz, precision, vf=o.scalars()["vf"] if oracle_vf else None)
    err = rel(H.xyt_rows(m.field()), o.field())
    print(f"xyt raman precision={precision}: rel {err:.3e}")
    assert err < tol


def white_noise(pkg, half_width_mm, kparaxial=False, seed=5):
    """Linear-only XYT case on x, y in +-half_width_mm (32 points) whose initial
    field is real white noise: every (kx, ky, w) of the propagator carries
    weight.  At 0.5 mm kt reaches past beta(w) at the lowest frequencies, so
    evanescent and near-cutoff components are part of the field."""
    p = small(pkg, nonlinearity=False, plasma=False)
    E = np.random.default_rng(seed).standard_normal((p["Nx"], p["Ny"], p["Nt"]))
    return dict(p, xmin=-half_width_mm, xmax=half_width_mm, ymin=-half_width_mm, ymax=half_width_mm,
                xguard=0.1 * half_width_mm, yguard=0.1 * half_width_mm, KPARAXIAL=kparaxial,
                initial_condition=lambda *a: E)


@pytest.mark.parametrize("kparaxial", [False, True])
@pytest.mark.parametrize("precision,tol", [(4, 1e-5), (8, 1e-10)])
def test_xyt_propagator_on_white_noise_matches_oracle(pkg, orc, kparaxial, precision, tol):
    """P = exp(i KZ dz) KX KY W / (Nx Ny) at every (kx, ky, w), paraxial and
    not (Kfunc spectral.jl:160-184), from the quarter of it the device stores."""
    p = white_noise(pkg, 0.5, kparaxial)
    o = oracle_for(pkg, orc, p)
    dz = 2e-4
    o.run(3, dz)
    m = run(pkg, p, 3, dz, precision, vf=o.scalars()["vf"])
    kz = m.kz
    kt2 = kz["kt2x"][:, None, None] + kz["kt2y"][None, :, None]
    evanescent = int(np.count_nonzero(kt2 > np.abs(kz["beta"][None, None, : p["Nt"] // 2]) ** 2))
    err = rel(H.xyt_rows(m.field()), o.field())
    print(f"white noise kparaxial={kparaxial} fp{8 * precision}: rel {err:.3e}, {evanescent} components with kt > |beta|")
    assert evanescent > 0
    assert err < tol


@pytest.mark.parametrize("axis", ["x", "y"])
@pytest.mark.parametrize("precision,tol", [(4, 2e-6), (8, 1e-13)])
def test_xyt_propagator_without_symmetric_tables_agrees(pkg, axis, precision, tol):
    """With kt^2 and the guards even in kx and ky (the FFT grid's) the device
    stores P for ix <= Nx/2, iy <= Ny/2 only and reads it for ix, Nx - ix and
    iy, Ny - iy; moving kt2x[1] (kt2y[1]) by one ulp breaks that symmetry and
    the whole x (y) extent is stored and read instead: the fields agree to the
    size of that change (FP64) or to the roundings of the multiply, which the
    compiler may contract differently in the two loops (FP32; a wrong row or
    column of P would differ at O(1))."""
    p = white_noise(pkg, 0.5)
    dz = 2e-4
    ms = run(pkg, p, 3, dz, precision)
    ma = pkg.XYTModel(p, precision=precision, keep_plasma=True)
    key = "kt2x" if axis == "x" else "kt2y"
    ma.kz[key][1] = np.nextafter(ma.kz[key][1], np.inf)
    ma._set_tables()
    ma.upload(ma.E0)
    z = 0.0
    for _ in range(3):
        z += dz
        ma.zstep(z, dz)
        ma.analyze(z)
    err = rel(ma.field(), ms.field())
    print(f"{axis}-symmetric vs whole stored propagator fp{8 * precision}: rel {err:.3e}")
    assert err < tol


def narrow_beam(pkg, qparaxial):
    """An elliptical 6 x 4 um beam (x, y in +-0.1 mm, 64 points; Kerr, free
    current and losses, no plasma), narrow enough that QZ(kx, ky, w) differs
    from its paraxial value by ~1e-3 over the beam's spectrum."""
    p = pkg.workloads.xyt_air(64, 64, 256, 0.5, xmax=0.1, ymax=0.1, tmin=-120.0, tmax=120.0, ax=6e-6, ay=4e-6,
                              tau0=20e-15, noise=0.0, plasma=False)
    return dict(p, xguard=0.01, yguard=0.01, QPARAXIAL=qparaxial)


@pytest.mark.parametrize("precision,tol", [(4, 1e-4), (8, 1e-10)])
def test_xyt_nonparaxial_steps_match_oracle(pkg, orc, precision, tol):
    """QPARAXIAL = false for XYT: func!'s space branch (nonlinear_propagators.jl:58-63,
    QZfunc XYT :184-204) as the three RK stages with IFFT_xy[Q .* FFT_xy[.]]
    between the rows and the combination.  Against the oracle, and distinct
    from the paraxial step by far more than the gate."""
    dz, n = 5e-5, 6
    p = narrow_beam(pkg, False)
    o = oracle_for(pkg, orc, p)
    o.run(n, dz)
    vf = o.scalars()["vf"]
    m = run(pkg, p, n, dz, precision, vf=vf)
    err = rel(H.xyt_rows(m.field()), o.field())
    mp = run(pkg, narrow_beam(pkg, True), n, dz, precision, vf=vf)
    dpar = rel(mp.field(), m.field())
    print(f"xyt non-paraxial fp{8 * precision}: rel {err:.3e} vs oracle; paraxial step differs by {dpar:.3e}")
    assert err < tol
    assert dpar > 1e-4 and dpar > 100 * err
