"""Pins the CPU restatement of the learned corrector (oracle/corrector.py)
against brute force, finite differences and closed forms.  CPU only."""
import importlib.util
import os
import struct

import numpy as np
import pytest

ROOT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..")


@pytest.fixture(scope="module")
def oc():
    spec = importlib.util.spec_from_file_location("oracle_corrector", os.path.join(ROOT, "oracle", "corrector.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_parameter_count(oc):
    assert oc.nparam() == 20258


def test_bf16_round_is_nearest_even(oc):
    def bits(f):
        return struct.unpack("<I", struct.pack("<f", f))[0]

    def from_bits(u):
        return struct.unpack("<f", struct.pack("<I", u))[0]

    one = bits(1.0)
    cases = {
        from_bits(one + 0x8000): 1.0,                      # tie, even below -> down
        from_bits(one + 0x18000): from_bits(one + 0x20000),  # tie, odd below -> up
        from_bits(one + 0x7FFF): 1.0,                      # below half -> down
        from_bits(one + 0x8001): from_bits(one + 0x10000),  # above half -> up
        -3.0: -3.0,
    }
    got = oc.bf16_round(np.array(list(cases.keys()), dtype=np.float32))
    assert np.array_equal(got, np.array(list(cases.values())))


def test_conv3x3_matches_brute_force(oc):
    rng = np.random.default_rng(1)
    B, ci, co, H, W = 2, 3, 4, 5, 6
    x = rng.standard_normal((B, ci, H, W))
    Wt = rng.standard_normal((co, ci, 3, 3))
    b = rng.standard_normal(co)
    ref = np.zeros((B, co, H, W))
    for n in range(B):
        for o in range(co):
            for y in range(H):
                for xx in range(W):
                    s = b[o]
                    for c in range(ci):
                        for dy in (-1, 0, 1):
                            for dx in (-1, 0, 1):
                                if 0 <= y + dy < H and 0 <= xx + dx < W:
                                    s += Wt[o, c, dy + 1, dx + 1] * x[n, c, y + dy, xx + dx]
                    ref[n, o, y, xx] = s
    assert np.allclose(oc.conv3x3(x, Wt, b), ref, rtol=1e-12, atol=1e-12)
    # the transpose is the adjoint: <conv(x), g> = <x, conv^T(g)>
    g = rng.standard_normal((B, co, H, W))
    lhs = np.sum(oc.conv3x3(x, Wt, np.zeros(co)) * g)
    rhs = np.sum(x * oc.conv3x3_transpose(g, Wt))
    assert abs(lhs - rhs) < 1e-9 * abs(lhs)


def test_loss_gradient_matches_finite_differences(oc):
    rng = np.random.default_rng(2)
    B, H, W = 2, 4, 5
    p = rng.standard_normal(oc.nparam()) * 0.2
    X = rng.standard_normal((B, 4, H, W))
    T = rng.standard_normal((B, 2, H, W)) * 3.0
    ys = 2.5
    loss, g = oc.loss_and_grad(p, X, T, ys, round_weights=False)
    # probe weights and biases of every layer
    offs, o = [], 0
    for ci, co in oc.LAYERS:
        n = co * ci * 9
        offs += list(rng.integers(o, o + n, 6)) + list(rng.integers(o + n, o + n + co, 2))
        o += n + co
    eps = 1e-6
    for q in offs:
        pp, pm = p.copy(), p.copy()
        pp[q] += eps
        pm[q] -= eps
        fd = (oc.loss_and_grad(pp, X, T, ys, False)[0] - oc.loss_and_grad(pm, X, T, ys, False)[0]) / (2 * eps)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(fd)), (q, fd, g[q])
    assert loss > 0


def test_features_closed_forms(oc):
    Nr, Nt = 16, 64
    geom = dict(nr=8, t0=16, nt=32, pool_r=2, pool_t=4, w0=0.7, tmin=-10.0, dt=0.25, a_scale=2.0, rho_scale=1e-3)
    t = geom["tmin"] + np.arange(Nt) * geom["dt"]
    c = 1.5 - 0.5j
    # a plane carrier: the envelope is the constant c
    E = np.tile(c * np.exp(-1j * geom["w0"] * t), (Nr, 1))
    rho = np.tile(np.arange(Nt, dtype=float), (Nr, 1)) * 1e-3
    X = oc.features(E, rho, geom)
    assert X.shape == (4, 4, 8)
    assert np.allclose(X[0], c.real / 2.0) and np.allclose(X[1], c.imag / 2.0)
    assert np.allclose(X[2], abs(c) ** 2 / 4.0)
    # block average of a ramp: mean of samples t0 + 4b .. t0 + 4b + 3
    expect = (16 + 4 * np.arange(8) + 1.5)
    assert np.allclose(X[3], np.tile(expect, (4, 1)))


def test_upsample_reproduces_linear_fields(oc):
    geom = dict(nr=8, t0=0, nt=32, pool_r=2, pool_t=4)
    H, W = 4, 8
    hh, ww = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    D = np.stack([1.0 + 0.5 * hh - 0.25 * ww, np.full((H, W), 3.0)])
    up = oc.upsample(D, geom)
    # interior samples: the pooled coordinate of row i is (i + 1/2)/pr - 1/2
    y = (np.arange(8) + 0.5) / 2 - 0.5
    x = (np.arange(32) + 0.5) / 4 - 0.5
    exact = 1.0 + 0.5 * y[:, None] - 0.25 * x[None, :]
    inner = (slice(1, 7), slice(2, 30))
    assert np.allclose(up[0][inner], exact[inner])
    assert np.allclose(up[1], 3.0)
