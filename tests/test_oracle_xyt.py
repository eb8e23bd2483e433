"""Pins the oracle's XYT geometry (Grids.jl:158-201, Guards.jl:162-195,318,
spectral.jl:131-184, FieldAnalyzers.jl:308-353) before it checks the device:
the 2-D FFT conventions against numpy, Gaussian-beam diffraction against its
closed form, the guard tables (with and without the reference's y-index
quirk) against the window formulas, and the full nonlinear step against the
independently pinned RT geometry on a cylindrically symmetric beam."""
import json
import math
import os

import numpy as np

import _helpers as H
from test_oracle import base_cfg

C0 = 299792458.0

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference")


def xyt_params(pkg, N=64, Nt=256, **kw):
    a = dict(xmax=8.0, ymax=8.0, tmin=-250.0, tmax=250.0, ax=1e-3, ay=1e-3, tau0=50e-15, noise=0.0)
    a.update(kw)
    return pkg.workloads.xyt_air(N, N, Nt, 1.0, **a)


def oracle_for(pkg, orc, p, bug=False, precision=8):
    g = pkg.physics.GridXYT(p["xmin"], p["xmax"], p["Nx"], p["ymin"], p["ymax"], p["Ny"], p["tmin"], p["tmax"], p["Nt"])
    E0 = p["initial_condition"](g.x, g.y, g.t, p["xu"], p["yu"], p["tu"], p["Iu"])
    return orc.Oracle(H.oracle_config_xyt(p, bug), H.xyt_rows(E0), precision), g


def test_xyt_grid_matches_host(pkg, orc):
    p = xyt_params(pkg, N=32, Nt=64)
    o, g = oracle_for(pkg, orc, p)
    gx = o.grid_xyt()
    for k in ("x", "y", "kx", "ky"):
        assert np.allclose(gx[k], getattr(g, k), rtol=1e-13, atol=1e-13)


def test_xyt_linear_step_zero_distance_is_identity(pkg, orc):
    """FFT2 forward then normalised inverse: dz = 0 leaves the spectral field
    unchanged (the spectral guards are 1 at these frequencies)."""
    p = xyt_params(pkg, N=32, Nt=64, nonlinearity=False, plasma=False)
    o, _ = oracle_for(pkg, orc, p)
    rng = np.random.default_rng(0)
    E = rng.standard_normal((32 * 32, 64)) + 1j * rng.standard_normal((32 * 32, 64))
    o.set_field(E)
    o.linear_step(0.0)
    W = np.exp(-(((2 * np.pi * np.fft.fftfreq(64, o.scalars()["dt"]) * 1e15) ** 2 / 1e16 ** 2) ** 20))
    assert np.allclose(o.field(), E * W[None, :], rtol=1e-12, atol=1e-12)


def test_xyt_fft2_convention_matches_numpy(pkg, orc):
    """One linear step with the propagator reduced to a known phase: with no
    dispersion term, exp(i KZ dz) multiplies each (kx, ky) mode; the step must
    equal numpy's ifft2(P * fft2(E)) along (x, y) for every t column."""
    p = xyt_params(pkg, N=32, Nt=64, nonlinearity=False, plasma=False)
    o, g = oracle_for(pkg, orc, p)
    rng = np.random.default_rng(1)
    E = rng.standard_normal((32, 32, 64)) + 1j * rng.standard_normal((32, 32, 64))
    o.set_field(H.xyt_rows(E))
    dz = 0.01
    o.linear_step(dz)
    KZ = o.KZ().reshape(32, 32, 64, order="F")
    gd = o.guards()
    Wg = gd["Wguard"]
    ref = np.fft.ifft2(np.fft.fft2(E, axes=(0, 1)) * np.exp(1j * KZ * dz) * Wg[None, None, :], axes=(0, 1))
    got = o.field().reshape(32, 32, 64, order="F")
    assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref))


def test_xyt_guard_tables_and_reference_quirk(pkg, orc):
    p = xyt_params(pkg, N=32, Nt=64)
    ph = pkg.physics
    u = ph.UnitXYT(p["xu"], p["yu"], p["zu"], p["tu"], p["Iu"], p["rhou"])
    for bug in (False, True):
        o, g = oracle_for(pkg, orc, p, bug=bug)
        gd = ph.GuardXYT(u, g, p["xguard"], p["yguard"], p["tguard"], p["wguard"])
        og = o.guards()
        X, Y, T, W = gd.Xguard, gd.Yguard, gd.Tguard, gd.Wguard
        rows = (X[:, None] * Y[None, :]).reshape(-1, order="F")
        if not bug:
            assert np.allclose(og["Rguard"], rows, atol=1e-15)
            assert np.allclose(og["Tguard"], T, atol=1e-15) and np.allclose(og["Wguard"], W, atol=1e-15)
        else:
            # kernel! (Guards.jl:314) takes C[i3] with i3 = the y index
            Ty = (X[:, None] * Y[None, :] * T[None, :32]).reshape(-1, order="F")
            assert np.allclose(og["Rguard"], Ty, atol=1e-15)
            assert np.all(og["Tguard"] == 1) and np.all(og["Wguard"] == 1)
            Wy = np.repeat(np.tile(W[:32][None, :], (32, 1)).reshape(-1, order="F")[:, None], 64, axis=1)
            assert np.allclose(og["Kguard"], Wy, atol=1e-15)


def test_xyt_gaussian_beam_diffraction(pkg, orc):
    """a(z) = a0 sqrt(1 + (z/zd)^2), on-axis I = (a0/a)^2 (the closed form of
    test/test_diffraction_rt.jl, here for the Cartesian beam).  The radius the
    analyzer reports is read off the grid at the 1/e level, so the expected
    value is the same grid readout of the exact profile."""
    p = xyt_params(pkg, N=128, Nt=512, xmax=4.0, ymax=4.0, tmin=-200.0, tmax=200.0, ax=0.5e-3, ay=0.5e-3,
                   tau0=50e-15, nonlinearity=False, plasma=False)
    o, g = oracle_for(pkg, orc, p)
    gold = json.load(open(os.path.join(GOLD, "constants.json")))
    zd = 2 * np.pi / 800e-9 * 0.5e-3 ** 2
    I0 = o.analyze(0.0)["Imax"]
    z = 0.0
    dz = 0.1
    for _ in range(5):
        z += dz
        o.zstep(z, dz)
        d = o.analyze(z)
        a = 0.5e-3 * math.sqrt(1 + (z / zd) ** 2)
        prof = np.exp(-(g.x * 1e-3) ** 2 / a ** 2)
        lev = prof.max() * math.exp(-1)
        idx = np.nonzero(prof >= lev)[0]
        ath = (abs(g.x[idx[0]]) + abs(g.x[idx[-1]])) / 2
        assert abs(d["rfil"] - ath) <= 1e-9            # ax: the same grid points
        assert abs(d["Imax"] / I0 - (0.5e-3 / a) ** 2) <= gold["diffraction_rt_intensity_tol"]


def test_xyt_nonlinear_step_agrees_with_rt(pkg, orc):
    """A cylindrically symmetric beam with the full response set (Kerr, free
    current, losses, multiphoton plasma) propagated by the XYT and the RT
    oracles over 5 cm: the growth of the peak intensity and of the peak
    density by self-focusing agree to the grids' resolution (the Cartesian grid
    has no node on the axis, so absolute peaks differ by ~1%; their growth is
    compared).  A wrong sign, factor or index in the XYT path changes the
    growth far beyond the 10% allowed here."""
    kw = dict(tmin=-120.0, tmax=120.0, tau0=20e-15, noise=0.0)
    px = pkg.workloads.xyt_air(128, 128, 256, 3.0, xmax=4.0, ymax=4.0, ax=0.4e-3, ay=0.4e-3, **kw)
    ox, gx = oracle_for(pkg, orc, px)
    pr = pkg.workloads.rt_air(256, 256, 3.0, rmax=4.0, a0=0.4e-3, **kw)
    gr = pkg.physics.GridRT(pr["rmax"], pr["Nr"], pr["tmin"], pr["tmax"], pr["Nt"])
    orr = orc.Oracle(H.oracle_config(pr), pr["initial_condition"](gr.r, gr.t, pr["ru"], pr["tu"], pr["Iu"]), 8)
    a0, b0 = ox.analyze(0.0), orr.analyze(0.0)
    dz = 5e-3
    for k in range(10):
        ox.zstep((k + 1) * dz, dz)
        orr.zstep((k + 1) * dz, dz)
    a, b = ox.analyze(10 * dz), orr.analyze(10 * dz)
    gi_x, gi_r = a["Imax"] / a0["Imax"] - 1, b["Imax"] / b0["Imax"] - 1
    gr_x, gr_r = a["rhomax"] / a0["rhomax"] - 1, b["rhomax"] / b0["rhomax"] - 1
    print(f"Imax growth XYT {gi_x:.4e} RT {gi_r:.4e}; rhomax growth XYT {gr_x:.4e} RT {gr_r:.4e}")
    assert gi_r > 0.01 and abs(gi_x - gi_r) < 0.1 * gi_r
    assert gr_r > 0.05 and abs(gr_x - gr_r) < 0.1 * gr_r
    assert abs(a["W"] / a0["W"] - 1) < 1e-6 and abs(b["W"] / b0["W"] - 1) < 1e-6


def test_t_mode_reproduces_reference_dispersion_test(pkg, orc):
    """The oracle's T geometry (spectral.jl:55-75, Guards.jl:78-89,
    FieldAnalyzers.jl:99-114) on the reference's test/test_dispersion_t.jl
    case: tau(z) = tau0 sqrt(1 + (z/zd)^2), zd = tau0^2/k2 with the golden k2,
    within the reference's tolerances (0.1 fs, 1e-3); the host's KZ(w) agrees
    with the oracle's to the frame-velocity rounding."""
    p = pkg.workloads.t_dispersion()
    g = pkg.physics.GridT(p["tmin"], p["tmax"], p["Nt"])
    E0 = p["initial_condition"](g.t, p["tu"], p["Iu"])
    o = orc.Oracle(H.oracle_config_t(p), E0[None, :], 8)
    rows = o.run(p["nsteps"], p["dz_initial"])
    k2, tau0 = 2.131584978375e-29, 20e-15
    zd = tau0 ** 2 / k2
    for d in rows[1:]:
        tauth = tau0 * np.sqrt(1 + (d["z"] / zd) ** 2)
        assert abs(d["tau"] * 1e-15 - tauth) <= 0.1e-15
        assert abs(d["Imax"] / rows[0]["Imax"] - tau0 / tauth) <= 1e-3
    ph = pkg.physics
    kz = ph.kz_table_t(ph.UnitT(p["zu"], p["tu"], p["Iu"], p["rhou"]), g,
                       ph.Medium(p["permittivity"], p["permeability"], p["n2"]), 2 * np.pi * ph.C0 / p["lam0"])
    assert np.max(np.abs(kz - o.KZ().ravel())) <= 1e-8 * np.max(np.abs(kz))


def _narrow_beam_xyt(orc, qparaxial, N=64, xmax=0.05, ymax=0.08, ax=0.004, ay=0.006, Nt=256, k=30, A=0.5):
    """A monochromatic elliptical Gaussian beam exp(-x^2/(2ax^2) - y^2/(2ay^2))
    at t-bin k on an XYT grid (mm, fs) with different x and y windows (so kx
    and ky differ) in a constant medium (n = 1), Kerr only, no guards: kt/beta
    reaches ~0.1 over its spectrum."""
    c = base_cfg(geometry=2, Nr=0, Nx=N, Ny=N, xmin=-xmax, xmax=xmax, ymin=-ymax, ymax=ymax, xu=1e-3, yu=1e-3,
                 xguard=0.0, yguard=0.0, kxguard=90.0, kyguard=90.0, guard3d_bug=0, Nt=Nt, tmin=-100.0,
                 tmax=100.0, nonlinearity=1, cubic=2, medium=1, n_const=1.0, rguard=0.0, tguard=0.0,
                 qparaxial=qparaxial)
    o = orc.Oracle(c, np.zeros((N * N, Nt)))
    gx = o.grid_xyt()
    X, Y = np.meshgrid(gx["x"], gx["y"], indexing="ij")
    Es = np.zeros((N * N, Nt), dtype=complex)
    Es[:, k] = H.xyt_rows((A * np.exp(-X ** 2 / (2 * ax * ax) - Y ** 2 / (2 * ay * ay)))[:, :, None])[:, 0]
    return o, gx, Es, k


def test_xyt_nonparaxial_rhs_is_qz_ratio_in_kx_ky_space(orc):
    """func! with QPARAXIAL = false for XYT (nonlinear_propagators.jl:58-63,
    QZfunc XYT :184-204, Qfunc_nonparaxial :232-242): after the 2-D FFT over
    x and y the right-hand side is the paraxial one times beta / kz(kx, ky),
    kz = sqrt(beta^2 - kx^2 - ky^2) (the guards are 1 to double precision
    here).  beta/kz - 1 reaches ~0.1 over the beam's spectrum; a transform
    taken the wrong way round, kx and ky exchanged (the beam is elliptical)
    or the paraxial factor kept fails by far more than the 1e-9 gate."""
    op, gx, Es, k = _narrow_beam_xyt(orc, 1)
    onp, _, _, _ = _narrow_beam_xyt(orc, 0)
    N = 64
    rp, rn = op.nl_rhs(Es)[:, k], onp.nl_rhs(Es)[:, k]
    # rows are ix + N iy: [iy][ix] in C order
    Fp = np.fft.fft2(rp.reshape(N, N))
    Fn = np.fft.fft2(rn.reshape(N, N))
    w = op.grid()["w"][k] * 1e15
    beta = w / C0
    KX, KY = np.meshgrid(gx["kx"] * 1e3, gx["ky"] * 1e3, indexing="xy")   # [iy][ix]
    kz = np.sqrt(beta ** 2 - KX ** 2 - KY ** 2 + 0j)
    expect = Fp * beta / kz
    sup = np.abs(Fp) > 1e-3 * np.abs(Fp).max()
    print(f"max beta/kz - 1 over the spectrum: {np.max(np.abs(beta / kz[sup] - 1)):.2e}")
    assert np.max(np.abs(beta / kz[sup] - 1)) > 1e-3
    assert np.max(np.abs(Fn - expect)) < 1e-9 * np.max(np.abs(Fp))
    assert np.max(np.abs(onp.nl_rhs(Es)[:, np.arange(256) != k])) < 1e-12 * np.max(np.abs(rn))
