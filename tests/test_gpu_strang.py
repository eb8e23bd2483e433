"""fil_zstep_parts (the z-step with separate nonlinear and linear distances)
and the symmetric (Strang) coarse step on the device.

  - fil_zstep_parts(dz, dz) is fil_zstep(dz), to the bit, in every form;
  - a chained Strang sequence L(h/2) [N(h) L(h)]^2 N(h) L(h/2) matches the
    same sequence composed from the FP64 oracle's nonlinear and linear steps
    (with the field filter and the analytic projection between the calls,
    as the device's half-spectrum state takes them): <= 1e-4 FP32, <= 1e-10
    FP64 (the oracle's KZ, as the full-step parity tests use it);
  - over 0.3 m of a self-focusing pulse, Strang coarse steps of 8 dz stay
    closer to the fine (dz) steps than Lie coarse steps of 8 dz do, by the
    fluence profile."""
import numpy as np
import pytest

import _helpers as H

pytestmark = pytest.mark.gpu


def small(pkg, power=6.0):
    return pkg.workloads.rt_air(64, 512, power, tmin=-125.0, tmax=125.0, tau0=20e-15, noise=1e-3)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("precision,form", [(4, 0), (8, 0), (4, 3)])
def test_parts_with_equal_distances_is_zstep(pkg, precision, form):
    p = small(pkg)
    dz = 5e-3
    out = []
    for parts in (False, True):
        m = pkg.RTModel(p, precision=precision, form=form)
        for k in range(4):
            if parts:
                m.zstep_parts((k + 1) * dz, dz, dz)
            else:
                m.zstep((k + 1) * dz, dz)
        out.append(m.field())
        m.close()
    assert np.array_equal(out[0], out[1])


def oracle_parts(o, filt, dz_nl, dz_lin):
    """fil_zstep_parts on the oracle: the projection the device state takes
    with the pending field filter, the plasma solve and RK3 step over dz_nl
    (none for 0), the linear step over dz_lin, the field filter."""
    o.analyze(0.0)
    E = o.field()
    Es = np.fft.ifft(E, axis=1)                 # forward t-transform P \\ E
    if dz_nl > 0:
        o.plasma_solve()
        Es = o.nl_step(Es, dz_nl)
    o.set_field(Es)
    o.linear_step(dz_lin)
    o.set_field(np.fft.fft(o.field(), axis=1) * filt)


@pytest.mark.parametrize("precision,tol", [(4, 1e-4), (8, 1e-10)])
def test_strang_sequence_matches_oracle(pkg, orc, precision, tol):
    p = small(pkg)
    g = pkg.physics.GridRT(p["rmax"], p["Nr"], p["tmin"], p["tmax"], p["Nt"])
    E0 = p["initial_condition"](g.r, g.t, p["ru"], p["tu"], p["Iu"])
    o = orc.Oracle(H.oracle_config(p), E0, 8)
    gd = o.guards()
    filt = gd["Rguard"][:, None] * gd["Tguard"][None, :]
    h = 4 * 5e-3
    seq = [(0.0, h / 2), (h, h), (h, h), (h, h / 2)]
    for dn, dl in seq:
        oracle_parts(o, filt, dn, dl)
    o.analyze(0.0)
    m = pkg.RTModel(p, precision=precision, KZ=o.KZ() if precision == 8 else None)
    Ea = m.field()                        # the analytic initial field
    z = 0.0
    for dn, dl in seq:
        z += dl
        m.zstep_parts(z, dn, dl)
    m.analyze(z)                          # the pending filter and the projection, as on the oracle side
    E, Eo = m.field(), o.field()
    m.close()
    e = rel(E, Eo)
    print(f"Strang sequence fp{8 * precision}: rel L2 {e:.2e}, field change {rel(Eo, Ea):.2e}")
    assert rel(Eo, Ea) > 1e-3             # the sequence moved the field
    assert e <= tol


def test_strang_coarse_steps_track_fine_steps_closer_than_lie(pkg):
    p = small(pkg, power=6.0)
    dz, n, F = 5e-3, 64, 8
    L = pkg.lib

    def fluence_after(run):
        m = pkg.RTModel(p, precision=L.FP64)
        try:
            run(m)
            return m.analyze(n * dz, arrays=True)["Fr"]
        finally:
            m.close()

    def fine(m):
        for k in range(n):
            m.zstep((k + 1) * dz, dz)

    def coarse(scheme):
        def go(m):
            st = pkg.CoarseStepper(m, factor=F, check_every=10 ** 6, tol=1.0, scheme=scheme)
            st.advance(0.0, dz, n)
        return go

    Ff = fluence_after(fine)
    eL = rel(fluence_after(coarse("lie")), Ff)
    eS = rel(fluence_after(coarse("strang")), Ff)
    print(f"fluence vs fine steps after {n * dz:.2f} m: Lie x{F} {eL:.2e}, Strang x{F} {eS:.2e}")
    assert eL > 1e-6
    assert eS < 0.5 * eL
