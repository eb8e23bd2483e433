"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Every check is against something the oracle does not compute itself: library
routines (scipy Bessel functions, numpy FFT), closed forms (Gaussian Hankel
pair, self-phase modulation, the RK3 amplification polynomial of a linear
ODE, Gaussian-beam diffraction, Gaussian-pulse dispersion) or the reference's
own golden numbers and fixture files (tests/golden/reference/).
"""
import json
import math
import os

import numpy as np
import pytest
import scipy.special as sps

import _helpers as H

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference")


def base_cfg(**kw):
    c = dict(geometry=0, Nr=64, Nt=256, rmax=10.0, tmin=-200.0, tmax=200.0, ru=1e-3, zu=1.0, tu=1e-15,
             Iu=1e16, rhou=2.5e25, lam0=800e-9, medium=0, n_const=1.0, n2=1e-23, nonlinearity=0, plasma=0,
             kparaxial=0, qparaxial=1, rguard=1.0, tguard=20.0, kguard=45.0, wguard=1e16, cubic=0, raman=0,
             current_free=0, nuc=5e12, mr=1.0, current_losses=0, losses_ereal=0, pe_ereal=0, pe_kdep=0,
             pe_rho0=0.0, pe_rho_nt=5.4e24, pe_nuc=0.0, pe_mr=1.0)
    c.update(kw)
    return c


@pytest.mark.parametrize("n", [8, 257, 501, 1025])
def test_bessel_zeros_match_scipy(orc, n):
    z = orc.bessel_j0_zeros(n)
    ref = sps.jn_zeros(0, n)
    assert np.max(np.abs(z - ref) / ref) < 1e-12


@pytest.mark.parametrize("n", [4, 8, 16, 64, 2048])
def test_fft_conventions(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    # P * E: unnormalised exp(-2 pi i jk/N); P \ E: normalised exp(+...)
    assert np.max(np.abs(orc.fft(x) - np.fft.fft(x))) < 1e-12 * n
    assert np.max(np.abs(orc.fft(x, True) - np.fft.ifft(x))) < 1e-14 * n
    if n <= 64:
        j = np.arange(n)
        brute = np.exp(-2j * np.pi * np.outer(j, j) / n) @ x
        assert np.max(np.abs(orc.fft(x) - brute)) < 1e-12


@pytest.mark.parametrize("N", [8, 16])
def test_dht_matrix_closed_form(orc, N):
    o = orc.Oracle(base_cfg(Nr=N, Nt=256), np.zeros((N, 256)))
    T, RJ, JV = o.dht()
    j = sps.jn_zeros(0, N + 1)
    S = j[N]
    J1 = np.abs(sps.j1(j[:N]))
    Tref = 2 * sps.j0(np.outer(j[:N], j[:N]) / S) / (S * np.outer(J1, J1))
    assert np.max(np.abs(T - Tref)) < 1e-13
    assert np.max(np.abs(RJ - 10.0 / J1)) < 1e-12
    assert np.max(np.abs(JV - J1 / (S / (2 * np.pi * 10.0)))) < 1e-14


def test_dht_gaussian_pair(orc):
    """H[exp(-r^2/a^2)](nu) = pi a^2 exp(-pi^2 a^2 nu^2), nu = k/(2 pi)."""
    N = 256
    o = orc.Oracle(base_cfg(Nr=N, Nt=256), np.zeros((N, 256)))
    g = o.grid()
    a = 1.3
    f = np.exp(-g["r"] ** 2 / a ** 2)
    F = o.dht_apply(f)
    nu = g["k"] / (2 * np.pi)
    assert np.max(np.abs(F - np.pi * a ** 2 * np.exp(-np.pi ** 2 * a ** 2 * nu ** 2))) < 1e-12
    assert np.max(np.abs(o.dht_apply(F, inverse=True) - f)) < 1e-12
    # the QDHT matrix is close to an involution (T T ~ I)
    T, _, _ = o.dht()
    assert np.max(np.abs(T @ T - np.eye(N))) < 1e-2


def test_grid_and_reference_k2(orc):
    o = orc.Oracle(base_cfg(), np.zeros((64, 256)))
    g, s = o.grid(), o.scalars()
    j = sps.jn_zeros(0, 65)
    assert np.allclose(g["r"], j[:64] * 10.0 / j[64], rtol=1e-13)
    assert np.allclose(g["k"], j[:64] / 10.0, rtol=1e-13)
    assert np.allclose(g["w"], 2 * np.pi * np.fft.fftfreq(256, 400.0 / 255), rtol=1e-13)
    gold = json.load(open(os.path.join(GOLD, "constants.json")))
    assert abs(s["k2"] / gold["k2_s2_per_m"] - 1) < 1e-9
    # Kerr phase per unit z and intensity: (w0/c) n2 Iu (AdaptiveSteps.jl:46-57)
    assert abs(s["phik"] / (s["w0"] / 299792458.0 * 1e-23 * 1e16) - 1) < 1e-12


def test_analytic_signal_of_cosine(orc):
    """rsig2asig!: cos(w t) -> exp(-i w t) for w on the frequency grid."""
    Nt, N = 256, 4
    t = np.linspace(-200.0, 200.0, Nt)
    dt = t[1] - t[0]
    w = 2 * np.pi * 17 / (Nt * dt)
    E0 = np.tile(np.cos(w * (t - t[0])), (N, 1))
    o = orc.Oracle(base_cfg(Nr=N, Nt=Nt), E0)
    assert np.max(np.abs(o.field() - np.exp(-1j * w * (t - t[0])))) < 1e-12


def test_rate_table_reference_fixtures(orc):
    """Log-log interpolation reproduces the reference's tables at the nodes
    and the exact power law between them (log-log slope 6.5 for O2, 7.5 for N2)."""
    tabs = [np.loadtxt(os.path.join(GOLD, f"multiphoton_{g}.tf")).T for g in ("O2", "N2")]
    c = base_cfg(plasma=1)
    o = orc.Oracle(dict(c, comp_frac=[0.21, 0.79], comp_Ui=[12.063, 15.576], tables=[tuple(t) for t in tabs]),
                   np.zeros((64, 256)))
    for q, (x, y) in enumerate(tabs):
        xs = x / 1e16
        for i in range(len(x)):
            assert abs(o.tf(q, xs[i]) / (y[i] * 1e-15) - 1) < 1e-9
        slope = 6.5 if q == 0 else 7.5
        I = math.sqrt(xs[3] * xs[4])
        assert abs(o.tf(q, I) / (y[3] * 1e-15 * (I / xs[3]) ** slope) - 1) < 1e-6
        assert abs(o.tf(q, I, power=True) - slope) < 1e-6
        assert o.tf(q, xs[0] / 10) == pytest.approx(y[0] * 1e-15, rel=1e-9)    # clamped below


def _power_law_cfg(**kw):
    x = np.logspace(9, 20, 12)
    y = 2.9e-131 * x ** 8
    return dict(base_cfg(plasma=1, **kw), comp_frac=[1.0], comp_Ui=[12.063], tables=[(x, y)])


def test_plasma_rk3_closed_form_constant_intensity(orc):
    """Constant W: one Kutta RK3 step multiplies (f - rho) by the cubic Taylor
    polynomial of exp(-W dt); kdrho[j] = K W (f - rho[j+1])."""
    Nr, Nt = 2, 256
    A = 11.0
    o = orc.Oracle(_power_law_cfg(Nr=Nr, Nt=Nt), np.zeros((Nr, Nt)))
    o.set_field(np.full((Nr, Nt), A + 0j))
    o.plasma_solve()
    rho, kd = o.plasma()
    s = o.scalars()
    W = 2.9e-131 * (A * A * 1e16) ** 8 * 1e-15
    h = s["dt"] * W
    g = 1 - h + h ** 2 / 2 - h ** 3 / 6
    f = 5.4e24 / 2.5e25
    n = np.arange(Nt)
    ref = f * (1 - g ** n)
    assert h > 0.05
    assert np.max(np.abs(rho[0] - ref)) < 1e-12
    assert np.max(np.abs(kd[0, :-1] - 8 * W * (f - ref[1:]))) < 1e-10
    assert kd[0, -1] == 0.0


def test_plasma_avalanche_closed_form_constant_field(orc):
    """With avalanche (photoionization_avalanche.jl:120-160) and a constant
    real field E = A: d rho/dt = (f - rho)(W + a rho), a = Rava A^2 / f, a
    Riccati equation whose solution is rho = f - 1/v,
    v(t) = a/(W + a f) + (1/f - a/(W + a f)) exp((W + a f) t).  Rava is the
    inverse-bremsstrahlung rate sigma_B I / Ui written through the cycle
    average Re(E)^2 = |E|^2 / 2 and I = (n0 eps0 c / 2) Eu^2 |E|^2:
    sigma_B = e^2 nu / (n0 c eps0 m (nu^2 + w0^2)) gives Rava = e^2 nu /
    (m (nu^2 + w0^2)) Eu^2 / Ui in 1/s, times tu.  The RK3 solution converges
    at third order."""
    A, nu = 10.0, 5e12
    e, me, eps0, c0 = 1.602176634e-19, 9.1093837015e-31, 8.8541878128e-12, 299792458.0
    errs = []
    for Nt in (256, 512, 1024):
        o = orc.Oracle(_power_law_cfg(Nr=2, Nt=Nt, tmin=-100.0, tmax=100.0, pe_nuc=nu), np.zeros((2, Nt)))
        o.set_field(np.full((2, Nt), A + 0j))
        o.plasma_solve()
        rho, _ = o.plasma()
        s = o.scalars()
        n0, Eu, w0 = s["n0"], s["Eu"], s["w0"]
        W = 2.9e-131 * (A * A * 1e16) ** 8 * 1e-15
        Rava = e ** 2 * nu / (me * (nu ** 2 + w0 ** 2)) * Eu ** 2 / (12.063 * e) * 1e-15
        f = 5.4e24 / 2.5e25
        a = Rava * A * A / f
        t = s["dt"] * np.arange(Nt)
        lam = W + a * f
        v = a / lam + (1 / f - a / lam) * np.exp(lam * t)
        ref = f - 1 / v
        assert ref[-1] > 0.5 * f and a * f > 0.2 * W      # both terms matter
        errs.append(np.max(np.abs(rho[0] - ref)) / f)
        assert n0 > 1.0 and abs(Eu ** 2 * n0 * eps0 * c0 / 2 - 1e16) < 1e-6 * 1e16    # Eu from Iu = 1e16 W/m^2
    assert errs[0] < 1e-4 and errs[0] / errs[1] > 6 and errs[1] / errs[2] > 6


def test_plasma_converges_to_exact_solution(orc):
    """rho = f (1 - exp(-int W dt)) for a smooth carrier-free Gaussian field:
    second-order convergence (linear interpolation of E at the RK3 midpoints)."""
    errs = []
    for Nt in (256, 512, 1024):
        t = np.linspace(-200.0, 200.0, Nt)
        A = 6.0 * np.exp(-t ** 2 / (2 * 40.0 ** 2))
        o = orc.Oracle(_power_law_cfg(Nr=2, Nt=Nt), np.zeros((2, Nt)))
        o.set_field(np.tile(A + 0j, (2, 1)))
        o.plasma_solve()
        rho, _ = o.plasma()
        W = 2.9e-131 * (A ** 2 * 1e16) ** 8 * 1e-15
        tt = np.linspace(-200, 200, 200001)
        Wf = 2.9e-131 * ((6.0 * np.exp(-tt ** 2 / (2 * 40.0 ** 2))) ** 2 * 1e16) ** 8 * 1e-15
        total = np.trapezoid(Wf, tt)
        f = 0.216
        errs.append(abs(rho[0, -1] - f * (1 - math.exp(-total))) / (f * (1 - math.exp(-total))))
    assert errs[0] / errs[1] > 3.5 and errs[1] / errs[2] > 3.5


def test_kerr_rhs_is_self_phase_modulation(orc):
    """For a monochromatic field at grid frequency w the cubic (no THG)
    right-hand side is dE/dz = i (w/c) n2 I E (constant-index medium)."""
    Nr, Nt = 4, 512
    t = np.linspace(-100.0, 100.0, Nt)
    dt = t[1] - t[0]
    kidx = 120
    w = 2 * np.pi * kidx / (Nt * dt)                  # grid units (1/fs)
    A = 0.7
    c = base_cfg(Nr=Nr, Nt=Nt, tmin=-100.0, tmax=100.0, nonlinearity=1, cubic=2, medium=1, n_const=1.0,
                 rguard=0.0, tguard=0.0)
    o = orc.Oracle(c, np.zeros((Nr, Nt)))
    Es = np.zeros((Nr, Nt), dtype=complex)
    Es[:, kidx] = A
    dE = o.nl_rhs(Es)
    expect = 1j * (w * 1e15 / 299792458.0) * 1e-23 * (A * A * 1e16) * A
    assert abs(dE[0, kidx] / expect - 1) < 1e-12
    dE[:, kidx] = 0
    assert np.max(np.abs(dE)) < 1e-12 * abs(expect)


def _mlejnek(t_s):
    """examples/RT/air.jl:47-58 (t in seconds), vectorised."""
    Om, Ga = 20.6e12, 26e12
    La = np.sqrt(Om ** 2 - Ga ** 2 / 4)
    return np.where(t_s >= 0, Om ** 2 * np.exp(-0.5 * Ga * np.maximum(t_s, 0)) * np.sin(La * t_s) / La, 0.0)


def test_raman_rhs_constant_intensity_is_scaled_kerr(orc):
    """With a constant |E| the Raman convolution of the intensity is the
    intensity times H0 = dt sum_j h(t_j - tshift) tu (the discrete integral of
    the response), so the Raman right-hand side is H0 times the cubic one
    (same Rnl, raman.jl:4-14 and cubic.jl:4-12)."""
    Nr, Nt = 2, 1024
    tmin, tmax = -1000.0, 1000.0
    c0 = base_cfg(Nr=Nr, Nt=Nt, tmin=tmin, tmax=tmax, nonlinearity=1, medium=1, n_const=1.0, rguard=0.0, tguard=0.0)
    oc = orc.Oracle(dict(c0, cubic=2), np.zeros((Nr, Nt)))
    orm = orc.Oracle(dict(c0, raman=2), np.zeros((Nr, Nt)))
    t = np.linspace(tmin, tmax, Nt)
    dt = t[1] - t[0]
    H0 = dt * np.sum(_mlejnek((t - 0.5 * (tmin + tmax)) * 1e-15) * 1e-15)
    assert abs(H0 - 1) < 1e-2        # the response is normalised: int h dt = Omega^2 / (Gamma^2/4 + Lambda^2) = 1
    Es = np.zeros((Nr, Nt), dtype=complex)
    Es[:, 300] = 0.6
    dc, dr = oc.nl_rhs(Es), orm.nl_rhs(Es)
    assert np.max(np.abs(dr - H0 * dc)) < 1e-12 * np.max(np.abs(dc))


def test_raman_rhs_matches_brute_force_convolution(orc):
    """Raman right-hand side of a short pulse plus a weak continuous probe
    against a direct evaluation of its definition (raman.jl:57-65,68-111,
    nonlinear_propagators.jl:38-67) in a constant-index medium: the circular
    convolution summed directly, dt sum_k hs[j-k] 3/4 |E_k|^2 with hs the
    ifftshift of h((t - tshift) tu) tu, times Re E, then the analytic
    spectrum (1/N) sum_j F_j exp(+i w t_j) with positive bins doubled and
    dE = i (4/3)(w/c) n2 Iu x that (the factor pinned by the Kerr test above).
    A reversed (anti-causal) convolution or a response shifted by half the
    window fails by O(1)."""
    Nr, Nt = 2, 512
    tmin, tmax = -600.0, 600.0
    c0 = base_cfg(Nr=Nr, Nt=Nt, tmin=tmin, tmax=tmax, nonlinearity=1, medium=1, n_const=1.0, rguard=0.0,
                  tguard=0.0, raman=2)
    o = orc.Oracle(c0, np.zeros((Nr, Nt)))
    t = np.linspace(tmin, tmax, Nt)
    dt = t[1] - t[0]
    w = 2 * np.pi * np.fft.fftfreq(Nt, dt)              # 1/fs, the oracle's frequency grid
    # field: Gaussian pulse (10 fs) at -200 fs on a 1e-2 probe, carriers on grid bins
    k1, k2 = 60, 75
    E = (1.0 * np.exp(-((t + 200.0) / 10.0) ** 2) * np.exp(-1j * w[k1] * (t - tmin))
         + 1e-2 * np.exp(-1j * w[k2] * (t - tmin)))
    Es = np.tile(np.fft.ifft(E), (Nr, 1))              # E_j = sum_k Es_k exp(-2 pi i jk/N)
    I = 0.75 * np.abs(E) ** 2
    hs = np.fft.ifftshift(_mlejnek((t - 0.5 * (tmin + tmax)) * 1e-15) * 1e-15)
    conv = np.array([dt * np.sum(hs[(j - np.arange(Nt)) % Nt] * I) for j in range(Nt)])
    F = conv * E.real
    A = np.fft.ifft(F)
    A[1:Nt // 2] *= 2
    A[Nt // 2:] = 0
    expect = 1j * (4.0 / 3.0) * (w * 1e15 / 299792458.0) * 1e-23 * 1e16 * A
    got = o.nl_rhs(Es)[0]
    assert np.max(np.abs(got - expect)) < 1e-9 * np.max(np.abs(expect))
    # causality, read off the direct convolution: the pulse's response is behind it
    j0 = np.argmin(np.abs(t + 200.0))
    assert conv[j0 + 60] > 10 * abs(conv[j0 - 60])


def test_linear_step_gaussian_beam_diffraction(orc):
    """Restates test/test_diffraction_rt.jl on the oracle at reduced Nt and
    step count: a(z) = a0 sqrt(1 + (z/zd)^2), I = (a0/a)^2 on axis."""
    Nr, Nt = 500, 1024
    t = np.linspace(-400.0, 400.0, Nt)
    c = base_cfg(Nr=Nr, Nt=Nt, tmin=-400.0, tmax=400.0, rguard=1.0, tguard=20.0)
    o = orc.Oracle(c, np.zeros((Nr, Nt)))
    g = o.grid()
    w0 = 2 * np.pi * 299792458.0 / 800e-9
    E = (np.exp(-0.5 * (g["r"][:, None] * 1e-3) ** 2 / 1e-3 ** 2) *
         np.exp(-0.5 * (t[None, :] * 1e-15) ** 2 / 100e-15 ** 2) * np.cos(w0 * t[None, :] * 1e-15))
    o = orc.Oracle(c, E)
    rows = o.run(10, 1.0)
    gold = json.load(open(os.path.join(GOLD, "constants.json")))
    zd = 2 * np.pi / 800e-9 * 1e-3 ** 2
    for r in rows:
        ath = 1e-3 * math.sqrt(1 + (r["z"] / zd) ** 2)
        assert abs(r["rfil"] * 1e-3 - ath) <= gold["diffraction_rt_radius_tol_m"]
        assert abs(r["Imax"] - (1e-3 / ath) ** 2) <= gold["diffraction_rt_intensity_tol"]


def test_linear_step_gaussian_pulse_dispersion(orc):
    """Restates test/test_dispersion_rt.jl on the oracle with fewer z steps
    (ru = 1 cm makes diffraction negligible): I = tau0 / tau(z)."""
    Nr, Nt = 500, 2048
    t = np.linspace(-100.0, 100.0, Nt)
    c = base_cfg(Nr=Nr, Nt=Nt, ru=1e-2, tmin=-100.0, tmax=100.0, tguard=10.0)
    o = orc.Oracle(c, np.zeros((Nr, Nt)))
    g = o.grid()
    w0 = 2 * np.pi * 299792458.0 / 800e-9
    E = (np.exp(-0.5 * (g["r"][:, None] * 1e-2) ** 2 / 1e-2 ** 2) *
         np.exp(-0.5 * (t[None, :] * 1e-15) ** 2 / 20e-15 ** 2) * np.cos(w0 * t[None, :] * 1e-15))
    o = orc.Oracle(c, E)
    rows = o.run(20, 1.0)
    gold = json.load(open(os.path.join(GOLD, "constants.json")))
    zd = 20e-15 ** 2 / gold["k2_s2_per_m"]
    for r in rows:
        tauth = 20e-15 * math.sqrt(1 + (r["z"] / zd) ** 2)
        assert abs(r["Imax"] - 20e-15 / tauth) <= gold["dispersion_rt_intensity_tol"]


def test_host_tables_match_oracle(pkg, orc):
    """Two independent restatements of the setup (numpy host code of the
    product, C++ oracle) agree."""
    p = pkg.workloads.rt_air(64, 256, 3.0)
    ph = pkg.physics
    u = ph.UnitRT(p["ru"], p["zu"], p["tu"], p["Iu"], p["rhou"])
    g = ph.GridRT(p["rmax"], p["Nr"], p["tmin"], p["tmax"], p["Nt"])
    m = ph.Medium(p["permittivity"], p["permeability"], p["n2"])
    w0 = 2 * np.pi * ph.C0 / p["lam0"]
    Eu = ph.field_unit(u, float(np.real(m.refractive_index(w0))))
    E0 = p["initial_condition"](g.r, g.t, u.r, u.t, u.I)
    o = orc.Oracle(H.oracle_config(p), E0)
    rel = lambda a, b: np.max(np.abs(a - b)) / np.max(np.abs(b))
    assert rel(ph.kz_table(u, g, m, w0, False), o.KZ()) < 1e-12
    assert rel(ph.qz_vector(u, g, m, w0, Eu), o.QZ()[0]) < 1e-12
    gd = ph.GuardRT(u, g, m, 1.0, 20.0, 45.0, 1e16)
    og = o.guards()
    for k in ("Rguard", "Tguard", "Kguard", "Wguard"):
        assert rel(getattr(gd, k), og[k]) < 1e-12
    T, RJ, JV = ph.dht_plan(g)
    oT, oRJ, oJV = o.dht()
    assert rel(T, oT) < 1e-10 and rel(RJ, oRJ) < 1e-12 and rel(JV, oJV) < 1e-12
    assert rel(ph.rsig2asig(E0), o.field()) < 1e-12
