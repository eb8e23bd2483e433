# Initial condition of the reference's RT diffraction test (test/diffraction/RT/gauss.jl).
lam0 = 800e-9   # [m] central wavelength


def initial_condition(r, t, ru, tu, Iu):
    a0 = 1e-3   # [m] initial beam radius
    t0 = 100e-15   # [s] initial pulse duration
    I0 = 1e12 * 1e4   # [W/m^2] initial intensity
    w0 = 2 * pi * C0 / lam0
    rr = np.asarray(r)[:, None] * ru
    tt = np.asarray(t)[None, :] * tu
    return np.sqrt(I0 / Iu) * np.exp(-0.5 * rr ** 2 / a0 ** 2) * np.exp(-0.5 * tt ** 2 / t0 ** 2) * np.cos(w0 * tt)
