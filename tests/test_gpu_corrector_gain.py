"""The learned corrector earns its place: trained on the device on FP64
fine-step targets of an ensemble of pulses (RT 512 x 2048, powers 1-10 Pcr,
radii 0.5-2 mm, durations 25-100 fs), it brings coarse steps of 8 dz closer
to the FP64 fine-step solution of a pulse it has not seen, in both the
fluence profile and the field."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FACTOR = 8
NFINE = 400


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_corrector_reduces_coarse_step_error(pkg):
    L, C = pkg.lib, pkg.corrector
    form = L.FORM_DHT_BF16X3 | L.FORM_SCAN_FP64

    def make(d):
        return pkg.workloads.rt_air(512, 2048, d["power_pcr"], a0=d["a0"], tau0=d["tau0"], seed=d["seed"])

    draws = C.ensemble_draws(8, seed=0)
    train, held = draws[:-1], draws[-1]
    p = make(held)
    dz = p["dz_initial"]
    m0 = pkg.RTModel(p, precision=L.FP32, form=form)
    geom = C.window_geometry(m0.grid, m0.unit, m0.w0)
    m0.close()
    st, info = C.fit(pkg.RTModel, make, train, geom, dz, FACTOR, form, npairs=32, iters=2000, seed=0)
    print(f"fit: {info}")
    assert info["last_batch_loss"] < 0.1          # fraction of the target power left unexplained

    def run(prec, fac, state=None):
        m = pkg.RTModel(p, precision=prec, form=form if prec == L.FP32 else 0)
        c = None
        if state is not None:
            c = C.Corrector(m, geom=geom, capacity=1, batch_max=1)
            c.load_state(state)
        z = 0.0
        for _ in range(NFINE // fac):
            m.zstep(z + fac * dz, fac * dz)
            z += fac * dz
            if c is not None:
                c.apply()
        F, E = m.analyze(z, arrays=True)["Fr"], m.field()
        m.close()
        return F, E

    F64, E64 = run(L.FP64, 1)
    Fc, Ec = run(L.FP32, FACTOR)
    Fk, Ek = run(L.FP32, FACTOR, st)
    ec, ek = rel(Fc, F64), rel(Fk, F64)
    gc, gk = rel(Ec, E64), rel(Ek, E64)
    print(f"held-out pulse {held}: fluence err coarse {ec:.3e} corrected {ek:.3e}; field err coarse {gc:.3e} "
          f"corrected {gk:.3e}")
    assert ek < ec and gk < gc
