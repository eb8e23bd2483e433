"""The RT workloads the throughput figures are quoted on, built from the
reference's input conventions (examples/RT/input.jl, test/diffraction/RT/air.jl,
examples/RT/gauss.jl) with synthetic, seeded initial fields.

RT 256x2048  : air, 1 atm, 800 nm; Nr=256 on [0,10 mm] (Hankel nodes),
               Nt=2048 on [-0.5,0.5] ps; Gaussian a0=1 mm, tau0=35 fs, 3 Pcr;
               1e-3 uniform multiplicative amplitude noise; Kerr n2=1e-23 (no THG),
               free current, MPI losses; O2 multiphoton ionisation K=8,
               sigma_8 = 2.9e-99 s^-1 (cm^2/W)^8, neutrals 5.4e24 m^-3; dz=0.5 mm.
RT 1024x4096 : same physics at Nr=1024, Nt=4096, 5 Pcr.
RT ensemble  : 64 pulses of the same physics, Nr in {256, 512, 1024}, Nt in
               {1024, 2048, 4096}, P ~ U[1, 10] Pcr, a0 ~ U[0.5, 2] mm,
               tau0 ~ U[25, 100] fs (seeded).
XYT 512^3    : air, 800 nm; x, y on [-10, 10] mm (512 points each), t on
               [-250, 250] fs (512); elliptical Gaussian 3 x 2 mm, 50 fs, 30 Pcr,
               5% amplitude noise (white noise smoothed by a Gaussian of 0.1 mm
               correlation length, seeded); dz = 0.2 mm.
"""
import math

import numpy as np

from . import physics as ph
from .responses import (RK3, RK4, init_cubic, init_current_free, init_current_losses, init_photoionization_avalanche,
                        init_raman)


def o2_rate_table(sigma8_cm=2.9e-99, K=8, xmin=1e9, xmax=1e20, n=12):
    """Power-law multiphoton rate W = sigma_K I^K as a two-column table
    (x in W/m^2, y in 1/s), log-uniform in x like the reference's .tf files."""
    x = np.logspace(math.log10(xmin), math.log10(xmax), n)
    y = sigma8_cm * (x * 1e-4) ** K
    return x, y


def gaussian_ic(a0, tau0, I0, lam0, noise=0.0, seed=0):
    """initial_condition(r, t, ru, tu, Iu) of examples/RT/gauss.jl form, with
    multiplicative amplitude noise (1 + noise (2u - 1)), u ~ U[0,1) seeded."""
    def initial_condition(r, t, ru, tu, Iu):
        w0 = 2 * math.pi * ph.C0 / lam0
        rr = np.asarray(r)[:, None] * ru
        tt = np.asarray(t)[None, :] * tu
        E = math.sqrt(I0 / Iu) * np.exp(-0.5 * rr ** 2 / a0 ** 2) * np.exp(-0.5 * tt ** 2 / tau0 ** 2) * np.cos(w0 * tt)
        if noise:
            u = np.random.default_rng(seed).random(E.shape)
            E = E * (1 + noise * (2 * u - 1))
        return E
    return initial_condition


def rt_air(Nr, Nt, power_pcr, rmax=10.0, tmin=-500.0, tmax=500.0, a0=1e-3, tau0=35e-15, noise=1e-3,
           seed=0, n2=1e-23, nonlinearity=True, plasma=True, dz=5e-4, raman=False):
    """Parameter dict in the reference's input-file names (examples/RT/input.jl:1-49, read by src/Input.jl:20-130).
    raman: add the stimulated Raman response of air with the same n2 and no THG
    (the response list of examples/RT/air.jl:88)."""
    lam0 = 800e-9
    medium = ph.Medium(ph.peck_reeder_air, ph.unit_permeability, n2)
    w0 = 2 * math.pi * ph.C0 / lam0
    Pcr = medium.critical_power(w0)
    I0 = power_pcr * Pcr / (math.pi * a0 ** 2)
    nuc, mr = 5e12, 1.0
    p = dict(
        prefix="results/", geometry="RT",
        ru=1e-3, zu=1.0, tu=1e-15, Iu=1e16, rhou=2.5e25,
        rmax=rmax, Nr=Nr, tmin=tmin, tmax=tmax, Nt=Nt, lam0=lam0,
        permittivity=ph.peck_reeder_air, permeability=ph.unit_permeability, n2=n2, nuc=nuc, mr=mr,
        NONLINEARITY=nonlinearity, PLASMA=plasma, KPARAXIAL=False, QPARAXIAL=True,
        rguard=1.0, tguard=20.0, kguard=45.0, wguard=1e16,
        dz_initial=dz, dzphimax=math.inf, Istop=1e3, ALG=RK3(), z=0.0,
        responses=[
            {"init": init_cubic, "THG": False, "n2": n2},
            {"init": init_current_free, "nuc": nuc, "mr": mr},
            {"init": init_current_losses, "EREAL": False},
        ] + ([{"init": init_raman, "THG": False, "n2": n2, "raman_response": ph.mlejnek_raman_response}]
             if raman else []),
        plasma_equation={
            "init": init_photoionization_avalanche, "ALG": RK3(), "EREAL": False, "KDEP": False,
            "rho0": 0.0, "rho_nt": 5.4e24, "nuc": 0.0, "mr": mr,
            "components": [{"fraction": 1.0, "ionization_energy": 12.063, "ionization_rate": o2_rate_table()}],
        },
        initial_condition=gaussian_ic(a0, tau0, I0, lam0, noise, seed),
        I0=I0, Pcr=Pcr, power_pcr=power_pcr,
    )
    return p


def gaussian_ic_r(a0, I0, focus=None, lam0=800e-9, n0=1.0):
    """initial_condition(r, ru, Iu) of test/diffraction/R/gauss.jl form,
    sqrt(I0/Iu) exp(-r^2/(2 a0^2)), with the focusing phase
    exp(-i k0 r^2 / (2 f)) of examples/R/gauss.jl when ``focus`` is given."""
    def initial_condition(r, ru, Iu):
        rr = np.asarray(r) * ru
        E = math.sqrt(I0 / Iu) * np.exp(-0.5 * rr ** 2 / a0 ** 2) + 0j
        if focus:
            k0 = n0 * 2 * math.pi / lam0
            E = E * np.exp(-1j * k0 * rr ** 2 / (2 * focus))
        return E
    return initial_condition


def gaussian_ic_xy(ax, ay, I0, noise=0.0, seed=0):
    """initial_condition(x, y, xu, yu, Iu) of test/diffraction/XY/gauss.jl
    form, sqrt(I0/Iu) exp(-x^2/(2 ax^2) - y^2/(2 ay^2)), times (1 + noise u),
    u ~ U[-1, 1) seeded, when ``noise`` is given."""
    def initial_condition(x, y, xu, yu, Iu):
        xx = np.asarray(x)[:, None] * xu
        yy = np.asarray(y)[None, :] * yu
        E = math.sqrt(I0 / Iu) * np.exp(-0.5 * xx ** 2 / ax ** 2 - 0.5 * yy ** 2 / ay ** 2) + 0j
        if noise:
            E = E * (1 + noise * (2 * np.random.default_rng(seed).random(E.shape) - 1))
        return E
    return initial_condition


def spatial_air(geometry, N, power_pcr=1.0, extent=10.0, a0=1e-3, n2=1e-23, nonlinearity=True, qparaxial=True,
                alg=4, dz=1e-2, guard=1.0, kguard=45.0, noise=0.0):
    """Parameter dict of the R or XY geometry (examples/R/input.jl,
    examples/XY/input.jl): a Gaussian beam of power_pcr critical powers in
    air at 800 nm, Nr = N points on [0, extent] (R) or N x N on
    [-extent, extent]^2 (XY), guards of width ``guard`` and cut-off angle
    ``kguard``, the cubic response."""
    lam0 = 800e-9
    medium = ph.Medium(ph.peck_reeder_air, ph.unit_permeability, n2)
    w0 = 2 * math.pi * ph.C0 / lam0
    Pcr = medium.critical_power(w0)
    I0 = power_pcr * Pcr / (math.pi * a0 ** 2)
    p = dict(prefix="results/", geometry=geometry, zu=1.0, Iu=1e16, lam0=lam0,
             permittivity=ph.peck_reeder_air, permeability=ph.unit_permeability, n2=n2,
             NONLINEARITY=nonlinearity, PLASMA=False, KPARAXIAL=False, QPARAXIAL=qparaxial,
             dz_initial=dz, dzphimax=math.inf, Istop=1e3, ALG=RK4() if alg == 4 else RK3(), z=0.0,
             responses=[{"init": init_cubic, "THG": False, "n2": n2}], I0=I0, Pcr=Pcr, power_pcr=power_pcr)
    if geometry == "R":
        p.update(ru=1e-3, rmax=extent, Nr=N, rguard=guard, kguard=kguard,
                 initial_condition=gaussian_ic_r(a0, I0))
    else:
        p.update(xu=1e-3, yu=1e-3, xmin=-extent, xmax=extent, Nx=N, ymin=-extent, ymax=extent, Ny=N,
                 xguard=guard, yguard=guard, kxguard=kguard, kyguard=kguard,
                 initial_condition=gaussian_ic_xy(a0, a0, I0, noise))
    return p


def gaussian_ic_t(tau0, I0, lam0):
    """initial_condition(t, tu, Iu) of a Gaussian pulse sqrt(I0/Iu)
    exp(-t^2/(2 tau0^2)) cos(w0 t) (the T geometry's form)."""
    def initial_condition(t, tu, Iu):
        w0 = 2 * math.pi * ph.C0 / lam0
        tt = np.asarray(t) * tu
        return math.sqrt(I0 / Iu) * np.exp(-0.5 * tt ** 2 / tau0 ** 2) * np.cos(w0 * tt)
    return initial_condition


def gaussian_ic_xyt(ax, ay, tau0, I0, lam0, noise=0.0, corr=1e-4, seed=0):
    """initial_condition(x, y, t, xu, yu, tu, Iu) of examples/XYT/gauss.jl form:
    an elliptical Gaussian beam and pulse, times (1 + noise N(x, y)) with N
    seeded white noise smoothed by a Gaussian kernel of correlation length
    ``corr`` [m] and scaled to unit standard deviation."""
    def initial_condition(x, y, t, xu, yu, tu, Iu):
        w0 = 2 * math.pi * ph.C0 / lam0
        xx = np.asarray(x)[:, None, None] * xu
        yy = np.asarray(y)[None, :, None] * yu
        tt = np.asarray(t)[None, None, :] * tu
        E = (math.sqrt(I0 / Iu) * np.exp(-0.5 * xx ** 2 / ax ** 2 - 0.5 * yy ** 2 / ay ** 2)
             * np.exp(-0.5 * tt ** 2 / tau0 ** 2) * np.cos(w0 * tt))
        if noise:
            rng = np.random.default_rng(seed)
            nx, ny = len(x), len(y)
            wn = rng.standard_normal((nx, ny))
            kx = 2 * np.pi * np.fft.fftfreq(nx, (x[1] - x[0]) * xu)
            ky = 2 * np.pi * np.fft.fftfreq(ny, (y[1] - y[0]) * yu)
            G = np.exp(-0.5 * corr ** 2 * (kx[:, None] ** 2 + ky[None, :] ** 2))
            N = np.real(np.fft.ifft2(np.fft.fft2(wn) * G))
            N /= N.std()
            E = E * (1 + noise * N)[:, :, None]
        return E
    return initial_condition


def xyt_air(Nx, Ny, Nt, power_pcr, xmax=10.0, ymax=10.0, tmin=-250.0, tmax=250.0, ax=3e-3, ay=2e-3, tau0=50e-15,
            noise=0.05, seed=0, n2=1e-23, nonlinearity=True, plasma=True, dz=2e-4, raman=False):
    """XYT parameter dict in the reference's input-file names (src/Input.jl:151-178,
    examples/XYT/input.jl); raman as in rt_air."""
    lam0 = 800e-9
    medium = ph.Medium(ph.peck_reeder_air, ph.unit_permeability, n2)
    w0 = 2 * math.pi * ph.C0 / lam0
    Pcr = medium.critical_power(w0)
    I0 = power_pcr * Pcr / (math.pi * ax * ay)
    nuc, mr = 5e12, 1.0
    return dict(
        prefix="results/", geometry="XYT",
        xu=1e-3, yu=1e-3, zu=1.0, tu=1e-15, Iu=1e16, rhou=2.5e25,
        xmin=-xmax, xmax=xmax, Nx=Nx, ymin=-ymax, ymax=ymax, Ny=Ny, tmin=tmin, tmax=tmax, Nt=Nt, lam0=lam0,
        permittivity=ph.peck_reeder_air, permeability=ph.unit_permeability, n2=n2, nuc=nuc, mr=mr,
        NONLINEARITY=nonlinearity, PLASMA=plasma, KPARAXIAL=False, QPARAXIAL=True,
        xguard=1.0, yguard=1.0, tguard=20.0, kxguard=45.0, kyguard=45.0, wguard=1e16,
        dz_initial=dz, dzphimax=math.inf, Istop=1e3, ALG=RK3(), z=0.0,
        responses=[
            {"init": init_cubic, "THG": False, "n2": n2},
            {"init": init_current_free, "nuc": nuc, "mr": mr},
            {"init": init_current_losses, "EREAL": False},
        ] + ([{"init": init_raman, "THG": False, "n2": n2, "raman_response": ph.mlejnek_raman_response}]
             if raman else []),
        plasma_equation={
            "init": init_photoionization_avalanche, "ALG": RK3(), "EREAL": False, "KDEP": False,
            "rho0": 0.0, "rho_nt": 5.4e24, "nuc": 0.0, "mr": mr,
            "components": [{"fraction": 1.0, "ionization_energy": 12.063, "ionization_rate": o2_rate_table()}],
        },
        initial_condition=gaussian_ic_xyt(ax, ay, tau0, I0, lam0, noise, 1e-4, seed),
        I0=I0, Pcr=Pcr, power_pcr=power_pcr,
    )


def xyt_512():
    """The XYT case: 512^3, 30 Pcr, 500 steps of 0.2 mm."""
    p = xyt_air(512, 512, 512, 30.0)
    p["nsteps"] = 500
    return p


def rt_256x2048():
    """The CPU-runnable plain-method case: 3 Pcr, 1000 steps of 0.5 mm."""
    p = rt_air(256, 2048, 3.0)
    p["nsteps"] = 1000
    return p


def rt_1024x4096():
    """The precision-form case the z-steps/s metric is quoted on: 5 Pcr, 2000 steps."""
    p = rt_air(1024, 4096, 5.0)
    p["nsteps"] = 2000
    return p


def rt_ensemble(n=64, seed=0, shapes_r=(256, 512, 1024), shapes_t=(1024, 2048, 4096)):
    """The batched ensemble: n independent pulses of uneven size.  Nr is drawn
    from shapes_r and Nt from shapes_t, power uniform on [1, 10] Pcr, a0 on
    [0.5, 2] mm, tau0 on [25, 100] fs (seeded).  All pulses share the air
    physics of rt_air, r on [0, 10] mm and t on [-0.5, 0.5] ps, so pulses of
    equal (Nr, Nt) share a grid and batch together.  Returns the parameter
    dicts in draw order; each carries its draw under 'draw'."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        Nr = int(rng.choice(shapes_r))
        Nt = int(rng.choice(shapes_t))
        P = float(rng.uniform(1.0, 10.0))
        a0 = float(rng.uniform(0.5e-3, 2e-3))
        tau0 = float(rng.uniform(25e-15, 100e-15))
        p = rt_air(Nr, Nt, P, a0=a0, tau0=tau0, seed=seed * 1000 + k)
        p["draw"] = dict(index=k, Nr=Nr, Nt=Nt, power_pcr=P, a0=a0, tau0=tau0)
        out.append(p)
    return out


def group_by_grid(members):
    """Members grouped by (Nr, Nt), in first-appearance order:
    [((Nr, Nt), [indices])]."""
    groups = {}
    for k, p in enumerate(members):
        groups.setdefault((p["Nr"], p["Nt"]), []).append(k)
    return list(groups.items())


def t_air(Nt, I0=1e16, tmin=-100.0, tmax=100.0, tau0=20e-15, nonlinearity=True, plasma=True, n2=1e-23,
          tguard=10.0, dz=1e-2, raman=False):
    """T-geometry parameter dict (test/dispersion/T/input.jl, gauss.jl): a
    Gaussian pulse sqrt(I0/Iu) exp(-t^2 / (2 tau0^2)) cos(w0 t) in air with the
    responses and plasma of rt_air."""
    lam0 = 800e-9
    nuc, mr = 5e12, 1.0
    initial_condition = gaussian_ic_t(tau0, I0, lam0)
    q = rt_air(2, Nt, 1.0, tmin=tmin, tmax=tmax, n2=n2, nonlinearity=nonlinearity, plasma=plasma, dz=dz, raman=raman)
    p = {k: v for k, v in q.items() if k not in ("ru", "rmax", "Nr", "rguard", "kguard", "I0", "Pcr", "power_pcr")}
    p.update(geometry="T", tguard=tguard, initial_condition=initial_condition, I0=I0, nuc=nuc, mr=mr)
    return p


def t_dispersion():
    """test/dispersion/T/input.jl: linear, Nt = 2048 on [-100, 100] fs, a 20 fs
    pulse at 1e12 W/cm^2, 100 steps of 0.2 m."""
    p = t_air(2048, I0=1e16, tmin=-100.0, tmax=100.0, tau0=20e-15, nonlinearity=False, plasma=False,
              tguard=10.0, dz=0.2)
    p["nsteps"] = 100
    return p
