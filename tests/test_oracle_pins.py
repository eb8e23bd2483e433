"""Pins the oracle's z-direction RK3 step, the THG responses, the free-current
response, the plasma phase constant of the adaptive step and the RT field
analyzer against closed forms, before the device is held to them.

Each closed form is derived from the physics, not from the oracle's code:
  - free current (current_free.jl:4-37): the Drude plasma response,
    dE/dz = -i w_p^2 / (2 c^2 beta) * w / (w + i nu) * E for a monochromatic
    field in a constant density;
  - one RK3 step (nonlinear_propagators.jl:70-74) of that linear equation
    multiplies E by 1 + x + x^2/2 + x^3/6 (x = dz times the rate), and of the
    Kerr equation dE/dz = i k |E|^2 E converges at third order to
    E exp(i k |E|^2 z);
  - cubic with THG (cubic.jl:33-38): Re(E)^3 = A^3 (3 cos wt + cos 3wt) / 4,
    so the lines at w and 3w have equal right-hand sides i (w/c) n2 Iu A^3;
  - phi_plasma (AdaptiveSteps.jl:60-72) is the Drude phase rate at w0 per
    unit density;
  - FieldAnalyzerRT (FieldAnalyzers.jl:166-215) of a Gaussian beam and pulse:
    fluence, power, duration, radius and spectrum in closed form, and the
    discrete Parseval identity between the energy W and the spectrum S.
"""
import math

import numpy as np
import pytest

from test_oracle import base_cfg

C0, EPS0, QE, ME = 299792458.0, 8.8541878128e-12, 1.602176634e-19, 9.1093837015e-31


def mono(Nr, Nt, k, A):
    Es = np.zeros((Nr, Nt), dtype=complex)
    Es[:, k] = A
    return Es


def drude_rate(w, rho_m3, nuc, n):
    """dE/dz / E of the free-current response [1/m] (Drude plasma)."""
    wp2 = QE ** 2 * rho_m3 / (EPS0 * ME)
    beta = n * w / C0
    return -1j * wp2 / (2 * C0 ** 2 * beta) * w / (w + 1j * nuc)


def free_current_oracle(orc, nuc, Nt=512, n_const=1.0003):
    c = base_cfg(Nr=2, Nt=Nt, tmin=-100.0, tmax=100.0, nonlinearity=1, current_free=1, nuc=nuc, mr=1.0,
                 medium=1, n_const=n_const, rguard=0.0, tguard=0.0)
    return orc.Oracle(c, np.zeros((2, Nt)))


@pytest.mark.parametrize("nuc", [0.0, 5e12])
def test_free_current_rhs_is_drude(orc, nuc):
    o = free_current_oracle(orc, nuc)
    g = o.grid()
    k, A, rho = 37, 0.8, 1e-3                         # rho in units of rhou = 2.5e25 m^-3
    o.set_plasma(rho=rho)
    dE = o.nl_rhs(mono(2, 512, k, A))
    w = g["w"][k] * 1e15
    expect = drude_rate(w, rho * 2.5e25, nuc, 1.0003) * A      # zu = 1 m
    assert abs(dE[0, k] / expect - 1) < 1e-12
    dE[:, k] = 0
    assert np.max(np.abs(dE)) < 1e-12 * abs(expect)


def test_nl_step_of_linear_response_is_rk3_polynomial(orc):
    """For dE/dz = a E every third-order three-stage Runge-Kutta step is the
    cubic Taylor polynomial of exp(a dz); a wrong weight or stage
    coefficient breaks it at O(x^2) or O(x^3)."""
    o = free_current_oracle(orc, 5e12)
    g = o.grid()
    k, A, rho = 37, 0.8, 1e-3
    o.set_plasma(rho=rho)
    a = drude_rate(g["w"][k] * 1e15, rho * 2.5e25, 5e12, 1.0003)
    dz = 0.4 / abs(a)
    x = a * dz
    out = o.nl_step(mono(2, 512, k, A), dz)
    expect = A * (1 + x + x ** 2 / 2 + x ** 3 / 6)
    assert abs(out[0, k] / expect - 1) < 1e-12
    assert abs(expect / (A * np.exp(x)) - 1) > 1e-4      # the test tells RK3 from the exact exponential
    out[:, k] = 0
    assert np.max(np.abs(out)) < 1e-12 * A


def test_nl_step_kerr_converges_at_third_order(orc):
    """Self-phase modulation of a monochromatic field (cubic, no THG):
    E(z) = E exp(i (w/c) n2 I z); n RK3 steps over a phase of 0.5 rad have an
    error that falls as n^-3 (ratio 8 per halving; a second-order
    combination gives 4)."""
    Nr, Nt, k, A = 2, 512, 120, 0.7
    c = base_cfg(Nr=Nr, Nt=Nt, tmin=-100.0, tmax=100.0, nonlinearity=1, cubic=2, medium=1, n_const=1.0,
                 rguard=0.0, tguard=0.0)
    o = orc.Oracle(c, np.zeros((Nr, Nt)))
    w = o.grid()["w"][k] * 1e15
    kappa = (w / C0) * 1e-23 * A * A * 1e16            # rad per m
    z = 0.5 / kappa
    errs = []
    for n in (16, 32, 64):
        Es = mono(Nr, Nt, k, A)
        for _ in range(n):
            Es = o.nl_step(Es, z / n)
        errs.append(abs(Es[0, k] - A * np.exp(1j * kappa * z)) / A)
        assert np.max(np.abs(np.delete(Es[0], k))) < 1e-12
    assert errs[0] < 1e-5
    for e1, e2 in zip(errs, errs[1:]):
        assert 7.0 < e1 / e2 < 9.0, errs


def test_cubic_thg_rhs_lines(orc):
    Nr, Nt, k, A = 2, 512, 40, 0.9
    c = base_cfg(Nr=Nr, Nt=Nt, tmin=-100.0, tmax=100.0, nonlinearity=1, cubic=1, medium=1, n_const=1.0,
                 rguard=0.0, tguard=0.0)
    o = orc.Oracle(c, np.zeros((Nr, Nt)))
    w = o.grid()["w"][k] * 1e15
    dE = o.nl_rhs(mono(Nr, Nt, k, A))
    expect = 1j * (w / C0) * 1e-23 * 1e16 * A ** 3
    assert abs(dE[0, k] / expect - 1) < 1e-12
    assert abs(dE[0, 3 * k] / expect - 1) < 1e-12
    dE[:, [k, 3 * k]] = 0
    assert np.max(np.abs(dE)) < 1e-12 * abs(expect)
    # without THG the 3w line is absent and the w line is unchanged
    o2 = orc.Oracle(dict(c, cubic=2), np.zeros((Nr, Nt)))
    dE2 = o2.nl_rhs(mono(Nr, Nt, k, A))
    assert abs(dE2[0, k] / expect - 1) < 1e-12 and abs(dE2[0, 3 * k]) < 1e-12 * abs(expect)


def _mlejnek(t_s):
    """Raman response of air (examples/RT/air.jl:47-58), t in seconds."""
    Om, Ga = 20.6e12, 26e12
    La = np.sqrt(Om ** 2 - Ga ** 2 / 4)
    return np.where(t_s >= 0, Om ** 2 * np.exp(-0.5 * Ga * np.maximum(t_s, 0)) * np.sin(La * t_s) / La, 0.0)


def test_raman_thg_rhs_matches_brute_force_convolution(orc):
    """calc_raman (raman.jl:46-54): F = (h * Re(E)^2) Re(E), the convolution
    summed directly; then the analytic spectrum and the Kerr factor
    i (4/3)(w/c) n2 Iu."""
    Nr, Nt = 2, 512
    tmin, tmax = -600.0, 600.0
    c = base_cfg(Nr=Nr, Nt=Nt, tmin=tmin, tmax=tmax, nonlinearity=1, medium=1, n_const=1.0, rguard=0.0,
                 tguard=0.0, raman=1)
    o = orc.Oracle(c, np.zeros((Nr, Nt)))
    t = np.linspace(tmin, tmax, Nt)
    dt = t[1] - t[0]
    w = 2 * np.pi * np.fft.fftfreq(Nt, dt)
    E = (1.0 * np.exp(-((t + 200.0) / 10.0) ** 2) * np.exp(-1j * w[60] * (t - tmin))
         + 1e-2 * np.exp(-1j * w[75] * (t - tmin)))
    Es = np.tile(np.fft.ifft(E), (Nr, 1))
    I = E.real ** 2
    hs = np.fft.ifftshift(_mlejnek((t - 0.5 * (tmin + tmax)) * 1e-15) * 1e-15)
    conv = np.array([dt * np.sum(hs[(j - np.arange(Nt)) % Nt] * I) for j in range(Nt)])
    Fs = np.fft.ifft(conv * E.real)
    Fs[1:Nt // 2] *= 2
    Fs[Nt // 2:] = 0
    expect = 1j * (4.0 / 3.0) * (w * 1e15 / C0) * 1e-23 * 1e16 * Fs
    got = o.nl_rhs(Es)[0]
    assert np.max(np.abs(got - expect)) < 1e-9 * np.max(np.abs(expect))
    # the THG variant differs from the no-THG one (3/4 |E|^2) by O(1) at the 2w and 3w products
    o2 = orc.Oracle(dict(c, raman=2), np.zeros((Nr, Nt)))
    assert np.max(np.abs(o2.nl_rhs(Es)[0] - got)) > 0.1 * np.max(np.abs(expect))


@pytest.mark.parametrize("nuc", [0.0, 5e12])
def test_phi_plasma_is_drude_phase_rate(orc, nuc):
    """phip rhomax is the phase rate (the imaginary part of the free-current
    rate dE/dz / E) at w0 for the density rhomax in rhou units: rad per zu of
    the Drude response in the medium of index n0."""
    c = base_cfg(Nr=4, Nt=256, nonlinearity=1, nuc=nuc, mr=1.0)
    o = orc.Oracle(c, np.zeros((4, 256)))
    s = o.scalars()
    rate = drude_rate(s["w0"], 2.5e25, nuc, s["n0"])
    assert abs(s["phip"] / abs(rate.imag) - 1) < 1e-12


def test_adaptive_rule_follows_kerr_and_plasma_limits(orc):
    """AStepKerrPlasma (AdaptiveSteps.jl:96-107) in the oracle's main loop:
    each dz is min(dz0, phimax/(phik Imax), phimax/(phip rhomax)) of the
    analysis before the step, in the working precision; z accumulates in it."""
    pk = __import__("_helpers").load_pkg()
    p = pk.workloads.rt_air(32, 256, 8.0, tmin=-125.0, tmax=125.0, tau0=20e-15)
    g = pk.physics.GridRT(p["rmax"], p["Nr"], p["tmin"], p["tmax"], p["Nt"])
    E0 = p["initial_condition"](g.r, g.t, p["ru"], p["tu"], p["Iu"])
    import _helpers as H
    for prec, ftype in ((8, np.float64), (4, np.float32)):
        o = orc.Oracle(H.oracle_config(p), E0, prec)
        s = o.scalars()
        dz0 = 5e-3
        phimax = 0.4 * dz0 * s["phik"] * o.analyze(0.0)["Imax"]
        rows = o.run(6, dz0, phimax=phimax)
        z = ftype(0)
        limits = set()
        for a, b in zip(rows, rows[1:]):
            cand = [dz0, phimax / (s["phik"] * a["Imax"]), phimax / (s["phip"] * a["rhomax"])]
            dz = ftype(min(cand))
            limits.add(int(np.argmin(cand)))
            z = ftype(z + dz)
            assert b["z"] == float(z)
        assert 1 in limits             # the Kerr limit is active


def test_analyzer_gaussian_closed_forms_and_parseval(orc):
    """FieldAnalyzerRT (FieldAnalyzers.jl:166-215) of the analytic signal of
    exp(-r^2/(2a^2)) exp(-t^2/(2 tau0^2)) cos(w0 t), grid units:
      Fr(r) = sqrt(pi) tau0 exp(-r^2/a^2)             (t-sum, spectrally exact)
      Ft(t) = pi a^2 exp(-t^2/tau0^2),  W = pi^(3/2) a^2 tau0
      S(w)  = 2 pi^2 a^2 tau0^2 exp(-(w - w0)^2 tau0^2)  (w > 0)
      tau = radius(t, Ft) = tau0, rfil = 2 radius(r, Fr) = r[1] + a,
    the r-sums to the accuracy of the r dr quadrature on the Hankel nodes
    (1e-5 here); and the discrete Parseval identity sum_w S dw / (2 pi) = W,
    exact for an analytic field without DC content (a lost factor 2 or N dt
    in S fails it by O(1))."""
    from test_oracle import base_cfg as bc
    Nr, Nt, a, tau0 = 256, 2048, 1.0, 20.0
    c = bc(Nr=Nr, Nt=Nt, tmin=-200.0, tmax=200.0, rguard=0.0, tguard=0.0)
    g = orc.Oracle(c, np.zeros((Nr, Nt))).grid()
    r, t = g["r"], g["t"]
    dt = t[1] - t[0]
    w0 = 2 * np.pi * C0 / 800e-9 * 1e-15
    E = np.exp(-r[:, None] ** 2 / (2 * a * a)) * np.exp(-t[None, :] ** 2 / (2 * tau0 ** 2)) * np.cos(w0 * t[None, :])
    d = orc.Oracle(c, E).analyze(0.0)
    Fr_th = math.sqrt(math.pi) * tau0 * np.exp(-r ** 2 / a ** 2)
    assert np.max(np.abs(d["Fr"] - Fr_th)) < 1e-12 * Fr_th.max()
    assert abs(d["Fmax"] / Fr_th[0] - 1) < 1e-12
    Ft_th = math.pi * a * a * np.exp(-t ** 2 / tau0 ** 2)
    assert np.max(np.abs(d["Ft"] - Ft_th)) < 3e-5 * Ft_th.max()
    assert abs(d["W"] / (math.pi ** 1.5 * a * a * tau0) - 1) < 3e-5
    w = np.abs(g["w"][:Nt // 2 + 1])
    S_th = 2 * math.pi ** 2 * a * a * tau0 ** 2 * np.exp(-(w - w0) ** 2 * tau0 ** 2)
    assert np.max(np.abs(d["S"] - S_th)) < 3e-5 * S_th.max()
    dw = 2 * math.pi / (Nt * dt)
    assert abs(np.sum(d["S"]) * dw / (2 * math.pi) / d["W"] - 1) < 1e-12
    assert abs(d["tau"] - tau0) <= dt
    assert abs(d["rfil"] - (r[0] + a)) <= np.max(g["dr"])
    assert abs(d["Imax"] - 1) < 2e-3 and d["rhomax"] == 0 and d["De"] == 0


def _narrow_beam(orc, qparaxial, Nr=128, Nt=256, rmax=0.05, a=0.004, k=30, A=0.5):
    """A monochromatic Gaussian beam exp(-r^2/(2a^2)) at t-bin k (grid units:
    mm, fs) in a constant medium, narrow enough that kt^2/beta^2 ~ 1e-3."""
    c = base_cfg(Nr=Nr, Nt=Nt, rmax=rmax, tmin=-100.0, tmax=100.0, nonlinearity=1, cubic=2, medium=1, n_const=1.0,
                 rguard=0.0, tguard=0.0, kguard=90.0, qparaxial=qparaxial)
    o = orc.Oracle(c, np.zeros((Nr, Nt)))
    g = o.grid()
    Es = np.zeros((Nr, Nt), dtype=complex)
    Es[:, k] = A * np.exp(-g["r"] ** 2 / (2 * a * a))
    return o, g, Es, k


def test_nonparaxial_rhs_is_qz_ratio_in_hankel_space(orc):
    """func! with QPARAXIAL = false (nonlinear_propagators.jl:58-63, Qfunc_nonparaxial
    :232-242): in Hankel space the right-hand side is the paraxial one times
    beta / kz(k), kz = sqrt(beta^2 - k^2), times the spectral filter Kguard
    (kguard = 90 deg: exp(-((k/beta)^2)^20)) and Wguard.  A narrow beam makes
    beta/kz - 1 reach ~1e-3 over its spectrum, and a transform taken the wrong
    way round or the paraxial factor kept fails by far more than the 1e-9 gate."""
    op, g, Es, k = _narrow_beam(orc, 1)
    onp, _, _, _ = _narrow_beam(orc, 0)
    rp, rn = op.nl_rhs(Es)[:, k], onp.nl_rhs(Es)[:, k]
    Hp, Hn = op.dht_apply(rp), op.dht_apply(rn)
    w = g["w"][k] * 1e15
    beta = w / C0
    kt = g["k"] * 1e3
    kz = np.sqrt(beta ** 2 - kt ** 2 + 0j)
    Kg = np.exp(-((kt / beta) ** 2) ** 20)
    Wg = np.exp(-((w ** 2 / 1e16 ** 2) ** 20))
    expect = Hp * beta / kz * Kg * Wg
    sup = np.abs(Hp) > 1e-3 * np.abs(Hp).max()
    assert np.max(np.abs(beta / kz[sup] - 1)) > 5e-4           # the non-paraxial factor matters here
    assert np.max(np.abs(Hn - expect)) < 1e-9 * np.max(np.abs(Hp))
    assert np.max(np.abs(onp.nl_rhs(Es)[:, np.arange(256) != k])) < 1e-12 * np.max(np.abs(rn))


def test_nonparaxial_step_tends_to_paraxial_step_for_wide_beams(orc):
    """The all-rows RK3 of the non-paraxial term (the same Kutta combination)
    on a wide beam, where beta / kz - 1 < 1e-8: it reproduces the paraxial row-wise
    step to that order."""
    op, g, Es, k = _narrow_beam(orc, 1, rmax=10.0, a=1.5)
    onp, _, _, _ = _narrow_beam(orc, 0, rmax=10.0, a=1.5)
    w = g["w"][k] * 1e15
    dz = 0.3 / ((w / C0) * 1e-23 * 0.25 * 1e16)
    sp, sn = op.nl_step(Es, dz), onp.nl_step(Es, dz)
    d = np.max(np.abs(sn - sp)) / np.max(np.abs(sp - Es))
    assert d < 1e-6 and np.max(np.abs(sp - Es)) > 1e-3 * np.max(np.abs(Es))
