"""The T geometry on the device (Fields.jl:70-93, spectral.jl:55-75,
Guards.jl:78-89, FieldAnalyzers.jl:74-114): the reference's own dispersion
test (test/test_dispersion_t.jl) at its size and tolerances, and parity with
the oracle's T mode with the full response set and plasma."""
import numpy as np
import pytest

import _helpers as H

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("precision", [4, 8])
def test_t_dispersion_reference(pkg, precision):
    """tau(z) = tau0 sqrt(1 + (z/zd)^2), zd = tau0^2 / k2 with the reference's
    golden k2; |tau - tau_th| <= 0.1 fs and |I - tau0/tau| <= 1e-3."""
    p = pkg.workloads.t_dispersion()
    m = pkg.TModel(p, precision=precision)
    k2, tau0 = 2.131584978375e-29, 20e-15
    zd = tau0 ** 2 / k2
    d = m.analyze(0.0)
    I0 = d["Imax"]
    z = 0.0
    dz = p["dz_initial"]
    worst_tau = worst_I = 0.0
    for _ in range(p["nsteps"]):
        z += dz
        m.zstep(z, dz)
        d = m.analyze(z)
        tauth = tau0 * np.sqrt(1 + (z / zd) ** 2)
        worst_tau = max(worst_tau, abs(d["duration"] * 1e-15 - tauth))
        worst_I = max(worst_I, abs(d["Imax"] / I0 - tau0 / tauth))
    print(f"T dispersion precision={precision}: max |tau - th| {worst_tau * 1e15:.4f} fs, max |I - th| {worst_I:.2e}")
    assert worst_tau <= 0.1e-15 and worst_I <= 1e-3
    # a repeated analysis sees the field without the negative frequencies its
    # first one's field filter left (the half-spectrum state drops them): a
    # 1e-11-level change of F here, below the FP32 rounding
    assert abs(d["F"] / m.analyze(z)["F"] - 1) < (1e-6 if precision == 4 else 1e-10)


@pytest.mark.parametrize("precision,tol,oracle_kz", [(4, 1e-4, False), (8, 1e-10, True)])
def test_t_steps_match_oracle(pkg, orc, precision, tol, oracle_kz):
    p = pkg.workloads.t_air(1024, I0=6e17, tmin=-150.0, tmax=150.0, tau0=15e-15, raman=True)
    g = pkg.physics.GridT(p["tmin"], p["tmax"], p["Nt"])
    E0 = p["initial_condition"](g.t, p["tu"], p["Iu"])
    o = orc.Oracle(H.oracle_config_t(p), E0[None, :], 8)
    dz = 2e-3
    rows = o.run(20, dz)
    m = pkg.TModel(p, precision=precision, keep_plasma=True, KZ=o.KZ() if oracle_kz else None)
    z = 0.0
    m.analyze(0.0)
    for _ in range(20):
        z += dz
        m.zstep(z, dz)
        d = m.analyze(z)
    err = rel(m.field(), o.field()[0])
    print(f"T precision={precision}: rel {err:.3e}, Imax {d['Imax']:.6e} vs {rows[-1]['Imax']:.6e}, "
          f"rhomax {d['rhomax']:.4e} vs {rows[-1]['rhomax']:.4e}")
    assert rows[-1]["rhomax"] > 0 and abs(rows[-1]["Imax"] / rows[0]["Imax"] - 1) > 1e-3   # nonlinear regime
    assert err < tol
    assert abs(d["Imax"] - rows[-1]["Imax"]) <= 10 * tol * rows[-1]["Imax"]
    assert abs(d["F"] - rows[-1]["W"]) <= 10 * tol * rows[-1]["W"]
