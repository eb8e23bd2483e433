"""Filament.run's z loop kept on the device (fil_run_adaptive) against the same
loop driven from the host (Filament.jl:129-164 as filament_amd.driver.run runs
it: analyze, dz = AStepKerrPlasma(Imax, rhomax) (AdaptiveSteps.jl:96-107), z
and dz in single precision, zstep, analyze, Istop test).  The device loop runs
the same step and analysis kernels, so the gate is equality: the same number
of steps and stop reason, the same z, dz, Imax and rhomax after every step and
the same final field, bit for bit.  The host checks the device loop every
`chunk` steps; chunks that the stop ends early (steps taken past it are undone
and the counted ones run again) are covered by picking a chunk that does not
divide the step count."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FORMS = [("fp32", 4, 0), ("fp32-bf16x3-scan64", 4, 3), ("fp64", 8, 0)]


def host_loop(pkg, m, z, zmax, dz0, phimax, Istop, max_steps):
    L = pkg.lib
    f32 = np.float32
    z, zmax = f32(z), f32(zmax)
    d = m.analyze(float(z))
    log = [(float(z), 0.0, d["Imax"], d["rhomax"])]
    reason = L.STOP_ZMAX
    while z < zmax:
        if len(log) - 1 >= max_steps:
            reason = L.STOP_MAXSTEPS
            break
        dz = f32(m.dz_adaptive(dz0, phimax, d["Imax"], d["rhomax"]))
        z = f32(z + dz)
        m.zstep(float(z), float(dz))
        d = m.analyze(float(z))
        log.append((float(z), float(dz), d["Imax"], d["rhomax"]))
        if d["Imax"] > Istop:
            reason = L.STOP_ISTOP
            break
    return float(z), len(log) - 1, reason, np.array(log)


def workload(pkg):
    return pkg.workloads.rt_air(128, 512, 8.0, a0=1e-3, dz=4e-3, seed=3)


def kerr_phimax(pkg, p, precision=4, form=0, model=None):
    """phimax for which the Kerr rule sets dz = 0.9 dz0 at the start, from a
    model of its own (every analysis projects the field again, so the models
    compared must see the same calls)."""
    m = (model or pkg.RTModel)(p, precision=precision, form=form)
    try:
        return 0.9 * p["dz_initial"] * m.phik * m.analyze(0.0)["Imax"]
    finally:
        m.close()


@pytest.mark.parametrize("name,precision,form", FORMS)
@pytest.mark.parametrize("stop", ["zmax", "istop", "maxsteps"])
def test_device_loop_equals_host_loop(pkg, name, precision, form, stop):
    p = workload(pkg)
    dz0 = p["dz_initial"]
    # the Kerr rule sets dz from the first step on and shrinks it as the beam focuses
    phimax = kerr_phimax(pkg, p, precision, form)
    zmax, Istop, max_steps = 40 * dz0, np.inf, 1000
    if stop == "istop":
        # a threshold the run crosses halfway
        m = pkg.RTModel(p, precision=precision, form=form)
        try:
            _, _, _, lg0 = host_loop(pkg, m, 0.0, zmax, dz0, phimax, np.inf, max_steps)
        finally:
            m.close()
        assert lg0[-1, 2] > lg0[0, 2], "the beam does not focus: no Istop stop to test"
        Istop = 0.5 * (lg0[0, 2] + lg0[-1, 2])
    a = pkg.RTModel(p, precision=precision, form=form)
    b = pkg.RTModel(p, precision=precision, form=form)
    try:
        if stop == "maxsteps":
            max_steps = 17
        zh, nh, rh, lh = host_loop(pkg, a, 0.0, zmax, dz0, phimax, Istop, max_steps)
        chunk = next(c for c in (5, 7, 6) if nh % c)
        zd, nd, rd, ld = b.run_adaptive(0.0, zmax, dz0, phimax, Istop, max_steps, chunk=chunk)
        print(f"{name} {stop}: {nd} steps (host {nh}), chunk {chunk}, z {zd:.6g}, reason {rd}, "
              f"dz {ld[1, 1]:.4g}..{ld[-1, 1]:.4g}, Imax {ld[0, 2]:.4g}..{ld[-1, 2]:.4g}", flush=True)
        want = {"zmax": pkg.lib.STOP_ZMAX, "istop": pkg.lib.STOP_ISTOP, "maxsteps": pkg.lib.STOP_MAXSTEPS}[stop]
        assert rh == want and rd == want
        assert nd == nh and zd == zh
        assert len(set(ld[1:, 1])) > 1 and ld[1:, 1].max() < dz0, "dz did not adapt"
        np.testing.assert_array_equal(ld, lh)
        np.testing.assert_array_equal(b.field(), a.field())
        assert b.analyze(zd) == a.analyze(zh)
        # the handle keeps stepping from where the loop left it
        a.zstep(zh + dz0, dz0)
        b.zstep(zd + dz0, dz0)
        np.testing.assert_array_equal(b.field(), a.field())
    finally:
        a.close()
        b.close()


def test_device_loop_chunks_and_start_at_zmax(pkg):
    """The chunk length changes when the host looks, never the steps; a start
    at zmax takes no step."""
    p = workload(pkg)
    dz0 = p["dz_initial"]
    phimax = kerr_phimax(pkg, p)
    ms = [pkg.RTModel(p) for _ in range(3)]
    try:
        out = [m.run_adaptive(0.0, 23 * dz0, dz0, phimax, np.inf, 1000, chunk=c) for m, c in zip(ms, (1, 4, 0))]
        for (z, n, r, lg), m in zip(out[1:], ms[1:]):
            assert (z, n, r) == out[0][:3]
            np.testing.assert_array_equal(lg, out[0][3])
            np.testing.assert_array_equal(m.field(), ms[0].field())
        z, n, r, lg = ms[0].run_adaptive(out[0][0], out[0][0], dz0, phimax, np.inf, 10)
        assert (n, r, lg.shape) == (0, pkg.lib.STOP_ZMAX, (1, 4))
    finally:
        for m in ms:
            m.close()


def geometry_case(pkg, geometry):
    if geometry == "T":
        return pkg.TModel, pkg.workloads.t_air(1024, I0=6e17, tmin=-150.0, tmax=150.0, tau0=15e-15, dz=2e-3)
    return pkg.XYTModel, pkg.workloads.xyt_air(32, 32, 256, 3.0, xmax=4.0, ymax=4.0, tmin=-120.0, tmax=120.0,
                                                ax=0.8e-3, ay=0.6e-3, tau0=20e-15, noise=0.05)


@pytest.mark.parametrize("precision", [4, 8])
@pytest.mark.parametrize("geometry", ["T", "XYT"])
def test_device_loop_other_geometries(pkg, geometry, precision):
    """The T (one row, diagonal linear step) and XYT (2-D transform linear
    step) handles run the same device loop; stop inside a chunk."""
    model, p = geometry_case(pkg, geometry)
    dz0 = p["dz_initial"]
    phimax = kerr_phimax(pkg, p, precision, 0, model)
    a = model(p, precision=precision)
    b = model(p, precision=precision)
    try:
        zh, nh, rh, lh = host_loop(pkg, a, 0.0, 30 * dz0, dz0, phimax, np.inf, 1000)
        chunk = next(c for c in (7, 5, 6) if nh % c)
        zd, nd, rd, ld = b.run_adaptive(0.0, 30 * dz0, dz0, phimax, np.inf, 1000, chunk=chunk)
        print(f"{geometry} fp{8 * precision}: {nd} steps (host {nh}), chunk {chunk}, "
              f"dz {ld[1, 1]:.4g}..{ld[-1, 1]:.4g}, Imax {ld[0, 2]:.4g}..{ld[-1, 2]:.4g}", flush=True)
        assert (nd, zd, rd) == (nh, zh, rh) == (nh, zh, pkg.lib.STOP_ZMAX)
        assert ld[1, 1] < np.float32(dz0), "the first step did not adapt"
        np.testing.assert_array_equal(ld, lh)
        np.testing.assert_array_equal(b.field(), a.field())
    finally:
        a.close()
        b.close()


def test_device_loop_rejects_batches_and_bad_arguments(pkg):
    """Batched handles (one dz for pulses whose rules differ) and negative
    step limits or chunks are refused with an error, the state untouched."""
    p = workload(pkg)
    dz0 = p["dz_initial"]
    ens = pkg.RTEnsemble([p, pkg.workloads.rt_air(128, 512, 6.0, a0=1e-3, dz=4e-3, seed=4)])
    m = pkg.RTModel(p)
    try:
        with pytest.raises(pkg.lib.FilamentError, match="single-pulse"):
            ens.run_adaptive(0.0, 10 * dz0, dz0, np.inf, np.inf, 10)
        E0 = m.field()
        with pytest.raises(pkg.lib.FilamentError, match=">= 0"):
            m.run_adaptive(0.0, 10 * dz0, dz0, np.inf, np.inf, 10, chunk=-1)
        with pytest.raises(pkg.lib.FilamentError, match=">= 0"):
            m.run_adaptive(0.0, 10 * dz0, dz0, np.inf, np.inf, -1)
        np.testing.assert_array_equal(m.field(), E0)
    finally:
        ens.close()
        m.close()


@pytest.mark.parametrize("precision,tol", [(4, 1e-4), (8, 1e-6)])
def test_device_loop_matches_oracle(pkg, orc, precision, tol):
    """fil_run_adaptive against the oracle's main loop (Filament.jl:129-164
    with AStepKerrPlasma, AdaptiveSteps.jl:96-107) on the same input with the
    Kerr limit active: the same number of steps and stop, and z, dz, Imax,
    rhomax after every step and the final field within the precision's gate.
    The reference's rule is FP64 arithmetic (dz0, phimax, phik and phip are
    the input file's Float64 values, Input.jl:116, AdaptiveSteps.jl:30-72) on
    the analyzer's Imax/rhomax, with dz converted once to the field's type;
    the device loop keeps z and dz in single precision as the RT run does
    (Input.jl:107), the FP64 oracle keeps them in FP64, so the z values agree
    to single precision.  The FP64 handle runs on the oracle's KZ table."""
    import _helpers as H
    p = workload(pkg)
    dz0 = p["dz_initial"]
    g = pkg.physics.GridRT(p["rmax"], p["Nr"], p["tmin"], p["tmax"], p["Nt"])
    E0 = p["initial_condition"](g.r, g.t, p["ru"], p["tu"], p["Iu"])
    o = orc.Oracle(H.oracle_config(p), E0, 8)
    phimax = kerr_phimax(pkg, p, 8)
    n = 24
    rows = o.run(n, dz0, phimax=phimax)
    zo = np.array([r["z"] for r in rows])
    # stop by zmax halfway between the oracle's 20th and 21st z
    zmax = 0.5 * (zo[20] + zo[21])
    m = pkg.RTModel(p, precision=precision, KZ=o.KZ() if precision == 8 else None)
    try:
        z, nd, reason, lg = m.run_adaptive(0.0, zmax, dz0, phimax, np.inf, 1000, chunk=6)
        assert (nd, reason) == (21, pkg.lib.STOP_ZMAX)
        ref = np.array([(r["z"], 0.0, r["Imax"], r["rhomax"]) for r in rows[:nd + 1]])
        ref[1:, 1] = np.diff(ref[:, 0])
        rel = lambda a, b: np.max(np.abs(a - b) / np.abs(b))
        e = dict(z=rel(lg[1:, 0], ref[1:, 0]), dz=rel(lg[1:, 1], ref[1:, 1]), Imax=rel(lg[:, 2], ref[:, 2]),
                 rhomax=rel(lg[:, 3], ref[:, 3]))
        # the oracle's field after the same 21 steps
        o2 = orc.Oracle(H.oracle_config(p), E0, 8)
        o2.run(nd, dz0, phimax=phimax)
        ef = float(np.linalg.norm(m.field() - o2.field()) / np.linalg.norm(o2.field()))
        print(f"device loop vs oracle fp{8 * precision}: {nd} steps, dz {lg[1, 1]:.5g}..{lg[-1, 1]:.5g} (dz0 {dz0}), "
              f"rel errors {e}, field {ef:.3e}", flush=True)
        assert lg[1:, 1].max() < 0.95 * dz0, "the Kerr limit is not active"
        assert e["z"] < max(tol, 2e-7) and e["dz"] < max(tol, 2e-7)
        assert e["Imax"] < 10 * tol and e["rhomax"] < 100 * tol
        assert ef < tol
    finally:
        m.close()


@pytest.mark.parametrize("precision", [4, 8])
def test_device_loop_stops_on_nan_analysis(pkg, precision):
    """A field with a NaN sample: the analysis's Imax is NaN (maximum(abs2, E)
    propagates it, FieldAnalyzers.jl:201), which makes the reference's dz and z
    NaN and ends its loop before a step.  The device loop stops with
    FIL_STOP_NONFINITE and takes no step; the host's rule returns NaN.  (No
    plasma, so no rate-table lookup sees the NaN intensity.)"""
    p = pkg.workloads.rt_air(64, 256, 3.0, tmin=-125.0, tmax=125.0, tau0=20e-15, plasma=False)
    m = pkg.RTModel(p, precision=precision)
    try:
        E = m.field()
        E[3, 100] = np.nan
        m.upload(E)
        d = m.analyze(0.0)
        assert np.isnan(d["Imax"])
        assert np.isnan(m.dz_adaptive(1e-3, 0.1, d["Imax"], d["rhomax"]))
        z, n, reason, lg = m.run_adaptive(0.0, 1.0, 1e-3, 0.1, np.inf, 50, chunk=4)
        print(f"fp{8 * precision}: {n} steps, reason {reason}, log {lg}")
        assert (n, reason) == (0, pkg.lib.STOP_NONFINITE) and z == 0.0 and np.isnan(lg[0, 2])
    finally:
        m.close()
