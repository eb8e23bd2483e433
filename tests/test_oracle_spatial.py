"""The oracle of the spatial geometries R and XY pinned against what the
mathematics fixes, and the host tables of the product against it:

- Gaussian-beam diffraction (the closed form of test/test_diffraction_r.jl and
  test/test_diffraction_xy.jl) pins the transforms, KZ and the linear step;
- one nonlinear RK step on a field against the exact solution of the
  pointwise ODE dE/dz = i a |E|^2 E, E(h) = E0 exp(i a |E0|^2 h), whose local
  error falls as h^4 (RK3) or h^5 (RK4): a wrong tableau entry changes the
  order;
- a wide beam makes the non-paraxial step the paraxial one;
- a round beam evolves the same in R and in XY (two independent transforms,
  with the Kerr term);
- the host tables of filament_amd (KZ, QZ, guards, Rnl, phik) equal the
  oracle's.
"""
import math

import numpy as np
import pytest

import _helpers as H


def _p(geometry, N, **kw):
    pk = H.load_pkg()
    return pk.workloads.spatial_air(geometry, N, **kw)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


@pytest.mark.parametrize("geometry", ["R", "XY"])
@pytest.mark.parametrize("qparaxial", [True, False])
def test_host_tables_match_oracle(geometry, qparaxial):
    pk = H.load_pkg()
    p = _p(geometry, 64 if geometry == "R" else 32, qparaxial=qparaxial, extent=5.0, guard=1.0, kguard=0.05)
    o = H.spatial_oracle(p)
    t = o.tables()
    u = (pk.physics.UnitR(p["ru"], p["zu"], p["Iu"]) if geometry == "R"
         else pk.physics.UnitXY(p["xu"], p["yu"], p["zu"], p["Iu"]))
    g = (pk.physics.GridR(p["rmax"], p["Nr"]) if geometry == "R"
         else pk.physics.GridXY(p["xmin"], p["xmax"], p["Nx"], p["ymin"], p["ymax"], p["Ny"]))
    m = pk.physics.Medium(p["permittivity"], p["permeability"], p["n2"])
    w0 = 2 * math.pi * pk.physics.C0 / p["lam0"]
    Eu = pk.physics.field_unit(u, float(np.real(m.refractive_index(w0))))
    assert abs(Eu / t["Eu"] - 1) < 1e-13
    KZ = pk.physics.kz_spatial(u, g, m, w0, bool(p["KPARAXIAL"]))
    QZ = pk.physics.qz_spatial(u, g, m, w0, Eu, paraxial=qparaxial)
    # KZ = (Kfunc - w0/vf) zu cancels ~1e7 1/m to ~1e2: the two FP64 five-point
    # derivatives of k(w) behind vf differ in their last bits, i.e. ~1e-13 of beta zu
    beta = abs(complex(m.beta(w0)))
    assert np.max(np.abs(KZ - t["KZ"])) < 1e-12 * beta * p["zu"]
    assert _rel(QZ, t["QZ"]) < 1e-13
    if geometry == "R":
        gd = pk.physics.GuardR(u, g, m, w0, p["rguard"], p["kguard"])
        Kg, Fg = gd.Kguard, gd.Rguard
        assert np.allclose(o.grid()["r"], g.r, rtol=1e-14, atol=0)
    else:
        gd = pk.physics.GuardXY(u, g, m, w0, p["xguard"], p["yguard"], p["kxguard"], p["kyguard"])
        Kg, Fg = np.outer(gd.KXguard, gd.KYguard), np.outer(gd.Xguard, gd.Yguard)
    assert np.max(np.abs(Kg - t["Kg"])) < 1e-13 and np.max(np.abs(Fg - t["Fg"])) < 1e-13
    assert Fg.min() < 1 - 1e-6 and Kg.min() < 1e-6          # the guards are active
    Rnl = float(np.real(pk.physics.rnl_cubic(u, m, w0, Eu, 1, p["n2"])[0]))
    assert abs(Rnl / t["Rnl"] - 1) < 1e-13
    assert abs(pk.physics.phi_kerr(u, m, w0, Eu) / t["phik"] - 1) < 1e-13


def test_r_diffraction_closed_form():
    """test/test_diffraction_r.jl's Gaussian beam on a coarser grid: 1/e
    radius a0 sqrt(1 + (z/zd)^2) to one grid step, on-axis intensity at the
    first node (a0/a)^2 exp(-r0^2/a^2) to 1e-5 with the diffraction length
    of air, n0 k0 a0^2 (the reference's test allows 1e-3 with the vacuum one)."""
    p = _p("R", 400, nonlinearity=False, guard=0.0, kguard=90.0)
    o = H.spatial_oracle(p)
    r = o.grid()["r"] * p["ru"]
    a0, lam0, I0 = 1e-3, p["lam0"], p["I0"] / p["Iu"]
    zd = o.tables()["n0"] * 2 * math.pi / lam0 * a0 ** 2      # the diffraction length in air
    for z in range(1, 11):
        o.zstep(1.0)
        d = o.analyze()
        ath = a0 * math.sqrt(1 + (z / zd) ** 2)
        Ith = I0 * (a0 / ath) ** 2 * math.exp(-r[0] ** 2 / ath ** 2)
        assert abs(d["rfil"] * p["ru"] - ath) <= (r[1] - r[0])
        assert abs(d["Imax"] / Ith - 1) < 1e-5
        assert abs(d["P"] / (math.pi * I0) - 1) < 1e-5    # power pi a0^2 I0 (the r dr quadrature: ~4e-6)


def test_xy_diffraction_closed_form():
    """test/test_diffraction_xy.jl's Gaussian beam on a coarser grid, the
    intensity sampled half a step off the axis in x and y."""
    p = _p("XY", 256, nonlinearity=False, guard=0.0, kguard=90.0)
    o = H.spatial_oracle(p)
    dx = (p["xmax"] - p["xmin"]) / (p["Nx"] - 1) * p["xu"]
    a0, lam0, I0 = 1e-3, p["lam0"], p["I0"] / p["Iu"]
    zd = o.tables()["n0"] * 2 * math.pi / lam0 * a0 ** 2      # the diffraction length in air
    for z in range(1, 11):
        o.zstep(1.0)
        d = o.analyze()
        ath = a0 * math.sqrt(1 + (z / zd) ** 2)
        Ith = I0 * (a0 / ath) ** 2 * math.exp(-2 * (dx / 2) ** 2 / ath ** 2)
        assert abs(d["ax"] * p["xu"] - ath) <= dx and abs(d["ay"] * p["yu"] - ath) <= dx
        assert abs(d["Imax"] / Ith - 1) < 1e-5


@pytest.mark.parametrize("geometry", ["R", "XY"])
@pytest.mark.parametrize("alg,order", [(3, 3), (4, 4)])
def test_nl_step_local_error_order(geometry, alg, order):
    p = _p(geometry, 64 if geometry == "R" else 32, power_pcr=3.0, alg=alg, guard=0.0, extent=5.0)
    o = H.spatial_oracle(p)
    t = o.tables()
    E0 = o.field()
    a = complex(t["QZ"].flat[0]) * t["Rnl"] * 0.75
    assert abs(a.imag) < 1e-12 * abs(a.real)                 # lossless air: a real
    I0 = float(np.max(np.abs(E0) ** 2))
    errs = []
    for h in (0.05, 0.025, 0.0125):
        h = h / (a.real * I0)                                  # peak phase 0.05 .. 0.0125 rad (asymptotic range)
        o.set_field(E0)
        o.nl_step(h)
        exact = E0 * np.exp(1j * a * np.abs(E0) ** 2 * h)
        errs.append(_rel(o.field(), exact))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    print(f"{geometry} RK{alg}: errors {errs}, ratios {r1:.2f} {r2:.2f}")
    expect = 2.0 ** (order + 1)
    assert 0.85 * expect < r1 < 1.15 * expect and 0.9 * expect < r2 < 1.1 * expect
    assert errs[0] < 1e-5


@pytest.mark.parametrize("geometry", ["R", "XY"])
def test_nonparaxial_step_tends_to_paraxial_for_wide_beams(geometry):
    """QZ(kt) -> QZ(0) as kt -> 0: for a beam much wider than the wavelength
    the non-paraxial nonlinear step equals the paraxial one to (kt/k0)^2,
    for a beam of a few wavelengths it does not."""
    def diff(a0, extent):
        out = []
        for qp in (True, False):
            p = _p(geometry, 128 if geometry == "R" else 64, power_pcr=3.0, a0=a0, extent=extent, qparaxial=qp,
                   guard=0.0, kguard=90.0)
            o = H.spatial_oracle(p)
            o.nl_step(1e-3 * (a0 / 1e-3) ** 2)
            out.append(o.field())
        return _rel(out[1], out[0])
    wide, narrow = diff(1e-3, 5.0), diff(4e-6, 0.02)
    print(f"{geometry}: non-paraxial vs paraxial step, wide beam {wide:.2e}, narrow beam {narrow:.2e}")
    assert wide < 1e-7 and narrow > 1e3 * wide


def test_round_beam_same_in_r_and_xy():
    """Kerr self-focusing of a round Gaussian beam, 3 critical powers, 20
    steps: the on-axis intensity of the R and XY oracles agrees to 2e-4 (the
    R profile read at the XY node nearest the axis, half a step off it in x
    and y)."""
    pr = _p("R", 600, power_pcr=3.0, extent=8.0, guard=0.0, kguard=90.0, dz=0.01)
    px = _p("XY", 512, power_pcr=3.0, extent=8.0, guard=0.0, kguard=90.0, dz=0.01)
    orr, oxy = H.spatial_oracle(pr), H.spatial_oracle(px)
    dx = 16.0 / 511 * 1e-3
    for _ in range(20):
        orr.zstep(0.01)
        oxy.zstep(0.01)
    dr, dxy = orr.analyze(), oxy.analyze()
    r = orr.grid()["r"] * 1e-3
    # the on-axis profile I(r) of R at the XY node nearest the axis: r = dx/sqrt(2)
    I_r = np.abs(orr.field()) ** 2
    I_at = np.interp(dx / math.sqrt(2), r, I_r)
    print(f"R Imax {dr['Imax']:.6f} at r0, {I_at:.6f} at dx/sqrt2; XY Imax {dxy['Imax']:.6f}")
    assert dr["Imax"] > 1.005 * pr["I0"] / pr["Iu"]           # self-focusing has raised the peak
    assert abs(dxy["Imax"] / I_at - 1) < 2e-4
