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


@pytest.mark.parametrize("power,a0,tau0", [(3.0, 1.5e-3, 35e-15), (8.0, 0.8e-3, 50e-15)])
def test_symmetric_coarse_steps_hold_the_gate_on_other_pulses(pkg, power, a0, tau0):
    """The chosen form's coarse stepping (split-BF16 Hankel, FP64 scan, 25 dz in
    the symmetric order, a fine-step check every 800 dz with fallback at 1e-4)
    on pulses other than the bench's, RT 512 x 2048 over 0.5 m, against the
    FP64 fine steps: within the 1e-3 fluence gate, and closer than the same
    coarse steps in the zstep order."""
    L = pkg.lib
    p = pkg.workloads.rt_air(512, 2048, power, a0=a0, tau0=tau0, seed=7)
    dz, n, F = p["dz_initial"], 1000, 25
    form = L.FORM_DHT_BF16X3 | L.FORM_SCAN_FP64

    def fluence(m):
        return m.analyze(n * dz, arrays=True)["Fr"]

    m64 = pkg.RTModel(p, precision=L.FP64)
    try:
        for k in range(n):
            m64.zstep((k + 1) * dz, dz)
        F64 = fluence(m64)
    finally:
        m64.close()
    err, stats = {}, {}
    for scheme in ("strang", "lie"):
        m = pkg.RTModel(p, precision=L.FP32, form=form)
        try:
            st = pkg.CoarseStepper(m, factor=F, check_every=800 // F, tol=1e-4, scheme=scheme)
            st.advance(0.0, dz, n)
            err[scheme], stats[scheme] = rel(fluence(m), F64), st.stats()
        finally:
            m.close()
    print(f"{power} Pcr, a0 {a0 * 1e3:.1f} mm, tau0 {tau0 * 1e15:.0f} fs, 0.5 m, 25 dz: fluence error symmetric "
          f"{err['strang']:.2e}, zstep order {err['lie']:.2e}; checks {stats['strang']['checks']}, failed "
          f"{stats['strang']['failed_checks']}")
    assert err["strang"] <= 1e-3
    assert err["strang"] < err["lie"]
