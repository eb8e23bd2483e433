"""Batched RT handles (fil_desc.batch): B pulses on one grid advanced by one
handle must give each pulse's own single-handle result.  The batch shares the
row kernel and the two DHT products of every step (the products run over the
B x Nt/2 columns with the shared propagator), so the arithmetic per pulse is
the same; the gates are equality to rounding of one FP32 operation order
(<= 1e-6 relative after the steps) and <= 1e-13 in FP64.

The ensemble test draws the 64-pulse ensemble (Nr in {256, 512, 1024}, Nt in
{1024, 2048, 4096}, P ~ U[1, 10] Pcr, a0 ~ U[0.5, 2] mm, tau0 ~ U[25, 100] fs),
groups it by grid, and checks one pulse of every group against its own handle;
it prints the ensemble's step time."""
import ctypes
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def members(pkg, Nr, Nt, B, seed=1):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(B):
        P, a0, tau0 = rng.uniform(1, 10), rng.uniform(0.5e-3, 2e-3), rng.uniform(25e-15, 100e-15)
        out.append(pkg.workloads.rt_air(Nr, Nt, P, a0=a0, tau0=tau0, seed=k))
    return out


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


FORMS = [("fp32", 4, 0, 1e-6), ("fp32-bf16x3-scan64", 4, 3, 1e-6), ("fp64", 8, 0, 1e-13)]


@pytest.mark.parametrize("name,precision,form,tol", FORMS)
def test_batch_equals_single_pulses(pkg, name, precision, form, tol):
    ps = members(pkg, 96, 1024, 3)
    ens = pkg.RTEnsemble(ps, precision=precision, form=form)
    singles = [pkg.RTModel(p, precision=precision, form=form) for p in ps]
    dz = 5e-4
    z = 0.0
    for k in range(6):
        z += dz
        ens.zstep(z, dz)
        for m in singles:
            m.zstep(z, dz)
        if k % 2 == 1:   # the analyzer projects the state: run it on some steps, as a driver would
            # at k = 3 twice in a row: a repeated call analyzes the projected field again
            for rep in range(2 if k == 3 else 1):
                for b, m in enumerate(singles):
                    d1, d0 = ens.analyze(z, arrays=True, member=b), m.analyze(z, arrays=True)
                    for key in ("Fmax", "Imax", "rhomax", "De", "W", "rfil", "rpl", "tau"):
                        assert abs(d1[key] - d0[key]) <= tol * max(abs(d0[key]), 1e-300), (name, b, key)
                    for key in ("Fr", "Ft", "rho_end", "S"):
                        assert rel(d1[key], d0[key]) <= tol, (name, b, key, rep)
    F = ens.fluence()
    for b, m in enumerate(singles):
        e = rel(ens.field(b), m.field())
        print(f"{name} member {b}: field rel {e:.3e}")
        assert e <= tol
        assert rel(F[b], m.fluence()) <= tol
    # a saved and restored batch state continues identically
    ens.state_save(0)
    E_before = [ens.field(b) for b in range(3)]
    ens.zstep(z + dz, dz)
    ens.state_restore(0)
    for b in range(3):
        assert np.array_equal(ens.field(b), E_before[b])
    ens.close()
    for m in singles:
        m.close()


def test_batch_member_errors(pkg):
    ps = members(pkg, 64, 1024, 2)
    ens = pkg.RTEnsemble(ps)
    L = pkg.lib
    lib = L.load()
    E = np.zeros(64 * 1024, dtype=np.complex128)
    assert lib.fil_upload_member(ens.h, 2, L.ptr(E.view(np.float64)), 64) != 0
    assert lib.fil_download_field(ens.h, L.ptr(E.view(np.float64)), 64) != 0   # batched: member entry points
    assert lib.fil_analyze(ens.h, ctypes.c_double(0.0), None, None, None, None, None) != 0
    ens.close()


@pytest.mark.parametrize("form", [0, 3])
def test_tile_shapes_give_identical_steps(pkg, form):
    """128-row and 64-row tiles of the Hankel products sum every output over
    k in the same order: the z-steps agree to the bit (FP32 and split-BF16)."""
    p = pkg.workloads.rt_air(256, 1024, 3.0, seed=2)
    dz = p["dz_initial"]
    out = {}
    for t in (pkg.lib.TILING_128, pkg.lib.TILING_64):
        m = pkg.RTModel(p, form=form)
        m.set_tiling(t)
        assert m.tiling == t
        for k in range(3):
            m.zstep((k + 1) * dz, dz)
        out[t] = m.field()
        m.close()
    assert np.array_equal(out[pkg.lib.TILING_128], out[pkg.lib.TILING_64])


def test_tiling_rejects_bad_shapes(pkg):
    m = pkg.RTModel(pkg.workloads.rt_air(64, 512, 1.0))     # 64 rows: no 128-row tiles
    with pytest.raises(pkg.lib.FilamentError):
        m.set_tiling(pkg.lib.TILING_128)
    with pytest.raises(pkg.lib.FilamentError):
        m.set_tiling(7)
    m.set_tiling(pkg.lib.TILING_64)
    m.close()


def test_ensemble_64_groups_match_single_pulses(pkg):
    W = pkg.workloads
    ps = W.rt_ensemble(64, seed=0)
    groups = W.group_by_grid(ps)
    assert sum(len(ix) for _, ix in groups) == 64
    dz = 5e-4
    nsteps = 4
    ens = [(shape, ix, pkg.RTEnsemble([ps[k] for k in ix], form=3)) for shape, ix in groups]
    # tile table of this ensemble, timed on its own handles (state restored)
    table = pkg.tuning.autotune([e for _, _, e in ens], dz, steps=3, log=print)
    assert len(table) == len(ens)
    t_setup = time.perf_counter()
    for k in range(nsteps):
        for _, _, e in ens:
            e.zstep((k + 1) * dz, dz)
    for shape, ix, e in ens:
        single = pkg.RTModel(ps[ix[-1]], form=3)    # the rule's tiles: same result as the table's
        for k in range(nsteps):
            single.zstep((k + 1) * dz, dz)
        err = rel(e.field(len(ix) - 1), single.field())
        d1, d0 = e.analyze(nsteps * dz, member=len(ix) - 1), single.analyze(nsteps * dz)
        print(f"group {shape} x{len(ix)}: member {ix[-1]} field rel {err:.3e}, Imax {d1['Imax']:.4f} vs {d0['Imax']:.4f}")
        assert err <= 1e-6
        assert abs(d1["Imax"] - d0["Imax"]) <= 1e-6 * d0["Imax"]
        single.close()
    # step time of the whole ensemble (all groups, back to back on one stream each)
    for _, _, e in ens:
        e.synchronize()
    t0 = time.perf_counter()
    n = 10
    for k in range(n):
        for _, _, e in ens:
            e.zstep((nsteps + k + 1) * dz, dz)
    for _, _, e in ens:
        e.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / n
    print(f"64-pulse ensemble ({len(groups)} grids), eager: {ms:.2f} ms per ensemble z-step, "
          f"{64 * 1e3 / ms:.0f} pulse z-steps/s; setup-to-now {time.perf_counter() - t_setup:.1f} s")
    # the same steps as one graph launch per z-step
    grp = pkg.StepGroup([e for _, _, e in ens])
    z = (nsteps + n) * dz
    for k in range(2):
        z += dz
        grp.zstep(z, dz)
    grp.synchronize()
    t0 = time.perf_counter()
    for k in range(n):
        z += dz
        grp.zstep(z, dz)
    grp.synchronize()
    msg = (time.perf_counter() - t0) * 1e3 / n
    print(f"64-pulse ensemble, one graph launch per z-step: {msg:.2f} ms per ensemble z-step, "
          f"{64 * 1e3 / msg:.0f} pulse z-steps/s ({grp.stats()})")
    grp.close()
    for _, _, e in ens:
        d = e.analyze_all(z)
        assert all(np.isfinite(x["W"]) and x["W"] > 0 for x in d)
        e.close()


def test_group_graph_replays_equal_eager_steps(pkg):
    """fil_group_zstep (one graph launch per z-step over several handles)
    against the same handles stepped eagerly: every replay is compared with
    the eager result, with downloads and analyses (eager work on the handles'
    own streams, which also flips the pending-filter pattern the graphs are
    keyed on) between replays."""
    def make():
        a = pkg.RTEnsemble(members(pkg, 96, 1024, 3, seed=2), form=3)
        b = pkg.RTEnsemble(members(pkg, 64, 2048, 2, seed=3), form=3)
        x = pkg.XYTModel(pkg.workloads.xyt_air(32, 32, 256, 3.0, xmax=4.0, ymax=4.0, tmin=-120.0, tmax=120.0,
                                               ax=0.8e-3, ay=0.6e-3, tau0=20e-15, noise=0.05))
        return [a, b, x]

    eager, graphed = make(), make()
    grp = pkg.StepGroup(graphed)
    dz = 5e-4
    z = 0.0
    for k in range(7):
        z += dz
        for m in eager:
            m.zstep(z, dz)
        grp.zstep(z, dz)
        if k in (2, 5):
            for m in eager + graphed:
                if isinstance(m, pkg.RTEnsemble):
                    m.analyze_all(z)
                else:
                    m.analyze(z)
        for me, mg in zip(eager, graphed):
            if isinstance(me, pkg.RTEnsemble):
                for b in range(me.B):
                    assert np.array_equal(me.field(b), mg.field(b)), (k, b)
            else:
                assert np.array_equal(me.field(), mg.field()), k
    st = grp.stats()
    print(f"group: {st}")
    assert st["replays"] == 7 and st["captures"] == 2   # pending filter off (after an analysis) and on
    grp.close()
    for m in eager + graphed:
        m.close()


def test_group_guards_its_handles_tables(pkg):
    """A group's graphs hold its handles' device pointers: fil_set_tables on a
    grouped handle is refused until the group is destroyed, and a group one of
    whose handles was destroyed refuses to step (no replay of a graph that
    points at freed memory)."""
    L = pkg.lib
    p = pkg.workloads.rt_air(64, 256, 3.0, tmin=-125.0, tmax=125.0, tau0=20e-15)
    a, b = pkg.RTModel(p), pkg.RTModel(p)
    grp = pkg.StepGroup([a, b])
    try:
        grp.zstep(1e-3, 1e-3)
        grp.synchronize()
        with pytest.raises(L.FilamentError, match="destroy the group first"):
            a._set_tables()
        grp.close()
        a._set_tables()                                  # allowed once the group is gone
        grp = pkg.StepGroup([a, b])
        b.close()
        with pytest.raises(L.FilamentError, match="was destroyed"):
            grp.zstep(2e-3, 1e-3)
    finally:
        grp.close()
        a.close()
        b.close()
