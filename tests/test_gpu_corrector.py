"""The learned corrector on the device (fil_corrector_*) against its CPU
restatement (oracle/corrector.py): features, forward, loss gradient, the
applied correction, and that device training reduces the loss."""
import importlib.util
import os

import numpy as np
import pytest

import _helpers as H

pytestmark = pytest.mark.gpu

ROOT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..")


@pytest.fixture(scope="module")
def oc():
    spec = importlib.util.spec_from_file_location("oracle_corrector", os.path.join(ROOT, "oracle", "corrector.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def small(pkg, power=6.0):
    return pkg.workloads.rt_air(64, 512, power, tmin=-125.0, tmax=125.0, tau0=20e-15, noise=1e-3)


# pooled maps: one whose sizes are whole network tiles (2 or 4 x 16 pixels) and
# one that leaves partial tiles at both edges
POOLED = [(8, 32), (7, 21)]


def make(pkg, p, rho_scale=1.0, capacity=16, batch_max=8, seed=0, pooled=(8, 32)):
    m = pkg.RTModel(p, precision=pkg.lib.FP32)
    geom = pkg.corrector.window_geometry(m.grid, m.unit, m.w0, r_window=5e-3, t_window=(-62e-15, 62e-15),
                                         pooled=pooled)
    c = pkg.corrector.Corrector(m, geom=geom, capacity=capacity, batch_max=batch_max, rho_scale=rho_scale, seed=seed)
    return m, c


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_features_match_oracle(pkg, orc, oc):
    p = small(pkg)
    o = orc.Oracle(H.oracle_config(p), p["initial_condition"](*_grid_args(pkg, p)), 8)
    m, c = make(pkg, p)
    E = m.field()                      # time-domain field the features are taken from (no filter pending)
    o.set_field(E)
    o.plasma_solve()
    rho = o.plasma()[0]
    rs = float(rho.max())
    assert rs > 0
    m2, c2 = make(pkg, p, rho_scale=rs)
    X = c2.features()
    Xo = oc.features(E, rho, c2.geom)
    # the envelope is nearly real, so Re A and Im A are compared as one complex field
    for chans, tol in (((0, 1), 1e-5), ((2,), 1e-5), ((3,), 1e-4)):
        e = rel(X[list(chans)], Xo[list(chans)])
        print(f"feature channels {chans}: rel {e:.2e}")
        assert e < tol


def _grid_args(pkg, p):
    g = pkg.physics.GridRT(p["rmax"], p["Nr"], p["tmin"], p["tmax"], p["Nt"])
    return g.r, g.t, p["ru"], p["tu"], p["Iu"]


@pytest.mark.parametrize("pooled", POOLED)
def test_forward_matches_oracle(pkg, oc, pooled):
    m, c = make(pkg, small(pkg), pooled=pooled)
    rng = np.random.default_rng(3)
    w = (rng.standard_normal(pkg.lib.CORR_NPARAM) * 0.1).astype(np.float32)
    c.set_params(w)
    assert np.array_equal(c.params(), w)
    X = rng.standard_normal((4, c.H, c.W)).astype(np.float32)
    Y = c.forward(X)
    Yo, _ = oc.forward(w, X[None])
    e = rel(Y, Yo[0])
    print(f"forward rel {e:.2e}")
    assert e < 1e-5


@pytest.mark.parametrize("pooled", POOLED)
def test_gradient_matches_oracle(pkg, oc, pooled):
    m, c = make(pkg, small(pkg), pooled=pooled)
    rng = np.random.default_rng(4)
    w = (rng.standard_normal(pkg.lib.CORR_NPARAM) * 0.1).astype(np.float32)
    c.set_params(w)
    B = 4
    Xs = rng.standard_normal((B, 4, c.H, c.W)).astype(np.float32)
    Ts = (rng.standard_normal((B, 2, c.H, c.W)) * 0.3).astype(np.float32)
    for b in range(B):
        c.store_pair(Xs[b], Ts[b])
    c.y_scale = 0.7
    loss, g = c.eval(0, B, grad=True)
    lo, go = oc.loss_and_grad(w, Xs, Ts, 0.7)
    print(f"loss {loss:.6e} vs {lo:.6e}; grad rel {rel(g, go):.2e}")
    assert abs(loss - lo) <= 1e-5 * lo
    assert rel(g, go) < 1e-4
    # per layer too, so a small layer cannot hide behind a large one
    o = 0
    for ci, co in oc.LAYERS:
        n = co * ci * 9 + co
        assert rel(g[o:o + n], go[o:o + n]) < 1e-4
        o += n


@pytest.mark.parametrize("pooled", POOLED)
def test_apply_matches_oracle(pkg, oc, pooled):
    p = small(pkg)
    m, c = make(pkg, p, rho_scale=1e-17, pooled=pooled)
    rng = np.random.default_rng(5)
    w = (rng.standard_normal(pkg.lib.CORR_NPARAM) * 0.1).astype(np.float32)
    c.set_params(w)
    c.y_scale = 1e-3
    E0 = m.field()
    X = c.features()
    D = c.forward(X)
    c.apply()
    E1 = m.field()
    # the same correction on the CPU, then the projection onto the stored half
    # spectrum: bins 0..Nt/2-1 of the reference's forward transform P \ E
    # (exp(+i ...), numpy's ifft), where the carrier exp(-i w0 t) lives
    Ec = oc.apply(E0, D, c.geom, c.y_scale * c.geom["a_scale"])
    Nt = p["Nt"]
    F = np.fft.ifft(Ec, axis=1)
    F[:, Nt // 2:] = 0
    Ec = np.fft.fft(F, axis=1)
    dg, do = E1 - E0, Ec - E0
    e = rel(dg, do)
    print(f"apply: |dE| {np.abs(do).max():.3e}, rel {e:.2e}")
    assert np.abs(do).max() > 1e-6
    assert e < 1e-3


def test_deferred_apply_in_zstep(pkg):
    """fil_corrector_apply leaves its output to the next z-step's row kernel,
    which adds it before the pending field filter.  That step must give the
    state that adding the correction first (fil_synchronize does) and then
    stepping gives; a step without the correction differs by far more."""
    p = small(pkg)
    dz = p["dz_initial"]
    out = {}
    for mode in ("fused", "materialized", "none"):
        m, c = make(pkg, p, rho_scale=1e-17)
        rng = np.random.default_rng(5)
        c.set_params((rng.standard_normal(pkg.lib.CORR_NPARAM) * 0.1).astype(np.float32))
        c.y_scale = 1e-3
        m.zstep(dz, dz)                  # leaves its field filter pending
        if mode != "none":
            c.apply()
        if mode == "materialized":
            m.synchronize()              # adds the correction now
        m.zstep(2 * dz, dz)
        out[mode] = m.field()
        m.close()
    d_corr = rel(out["fused"], out["none"])
    d_path = rel(out["fused"], out["materialized"])
    print(f"deferred apply: correction changes the step by {d_corr:.2e}, fused vs materialized {d_path:.2e}")
    assert d_corr > 1e-6
    assert d_path < 1e-3 * d_corr


def test_training_reduces_loss(pkg):
    p = small(pkg)
    m, c = make(pkg, p, capacity=32, batch_max=8)
    dz = p["dz_initial"]
    # a learnable target: a fixed local function of the real features
    z = 0.0
    for k in range(24):
        m.zstep(z + dz, dz)
        z += dz
        X = c.features()
        T = np.stack([0.5 * X[0] * X[2], -0.3 * X[1] + 0.2 * X[0]]).astype(np.float32)
        c.store_pair(X, T)
    c.y_scale = float(np.sqrt(np.mean(T ** 2)))
    l0 = c.eval(0, 8)
    l1 = c.train(300, batch=8, lr=2e-3, seed=1)
    l2 = c.eval(0, 8)
    print(f"loss {l0:.3e} -> {l1:.3e} (last batch) -> {l2:.3e}")
    assert l2 < 0.1 * l0
