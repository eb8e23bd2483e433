"""The RT workloads the throughput figure is quoted on, on the device against
the FP64 oracle at their full sizes (Filament.run's loop: zstep then the
analysis, Filament.jl:129-164, with the analyzer's projection every step).

  RT 256 x 2048, 3 Pcr, 1000 steps of 0.5 mm (0.5 m): the plain port in FP64
      (relative L2 field error <= 1e-5) and FP32 (<= 1e-4) over the whole run.
  RT 1024 x 4096, 5 Pcr: the plain FP32 port and the precision form (FP32
      field, split-BF16 Hankel products, FP64 plasma scan) over 24 steps.
  RT 64 x 512, 6 Pcr: the precision form over 30 steps.

Beside the field error each case states the errors of the fluence profile,
the peak intensity, the peak electron density and the on-axis spectrum."""
import numpy as np
import pytest

import _helpers as H

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def diagnostics_gpu(m, z):
    d = m.analyze(z, arrays=True)
    E = m.field()
    return E, dict(Fr=d["Fr"], Imax=d["Imax"], rhomax=d["rhomax"], onaxis=np.abs(np.fft.ifft(E[0])) ** 2)


def diagnostics_oracle(o, z):
    d = o.analyze(z)
    E = o.field()
    return E, dict(Fr=d["Fr"], Imax=d["Imax"], rhomax=d["rhomax"], onaxis=np.abs(np.fft.ifft(E[0])) ** 2)


def errors(E, d, Eo, do):
    nh = len(d["onaxis"]) // 2
    return dict(field=rel_l2(E, Eo), fluence=rel_l2(d["Fr"], do["Fr"]),
                peak_intensity=abs(d["Imax"] / do["Imax"] - 1), max_electron_density=abs(d["rhomax"] / do["rhomax"] - 1),
                onaxis_spectrum=rel_l2(d["onaxis"][:nh], do["onaxis"][:nh]))


def run_oracle(orc, p, nsteps, dz):
    g = H.load_pkg().physics.GridRT(p["rmax"], p["Nr"], p["tmin"], p["tmax"], p["Nt"])
    E0 = p["initial_condition"](g.r, g.t, p["ru"], p["tu"], p["Iu"])
    o = orc.Oracle(H.oracle_config(p), E0, 8)
    o.run(nsteps, dz)
    return o


def run_device(pkg, p, nsteps, dz, precision, form=0):
    m = pkg.RTModel(p, precision=precision, form=form)
    m.analyze(0.0)
    z = 0.0
    for _ in range(nsteps):
        z += dz
        m.zstep(z, dz)
        m.analyze(z)
    return m, z


@pytest.fixture(scope="module")
def rt256_oracle(pkg, orc):
    p = pkg.workloads.rt_256x2048()
    n, dz = p["nsteps"], p["dz_initial"]
    print(f"oracle RT 256x2048: {n} steps", flush=True)
    o = run_oracle(orc, p, n, dz)
    return p, diagnostics_oracle(o, n * dz)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("precision,tol", [(8, 1e-5), (4, 1e-4)])
def test_rt256x2048_full_run_matches_oracle(pkg, rt256_oracle, precision, tol):
    p, (Eo, do) = rt256_oracle
    n, dz = p["nsteps"], p["dz_initial"]
    m, z = run_device(pkg, p, n, dz, precision)
    try:
        E, d = diagnostics_gpu(m, z)
    finally:
        m.close()
    e = errors(E, d, Eo, do)
    print(f"RT 256x2048 fp{8 * precision}, {n} steps ({z:.3f} m): " + ", ".join(f"{k} {v:.2e}" for k, v in e.items()),
          flush=True)
    assert e["field"] <= tol
    # |E|^2 quantities carry about twice the field's relative error
    assert e["fluence"] <= 2 * tol and e["peak_intensity"] <= 2 * tol and e["onaxis_spectrum"] <= 2 * tol
    assert e["max_electron_density"] <= 20 * tol          # Ne ~ I^8: eight times the intensity error and more


@pytest.fixture(scope="module")
def rt1024_oracle(pkg, orc):
    p = pkg.workloads.rt_1024x4096()
    n, dz = 24, p["dz_initial"]
    print(f"oracle RT 1024x4096: {n} steps", flush=True)
    o = run_oracle(orc, p, n, dz)
    return p, n, diagnostics_oracle(o, n * dz)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("form_name", ["plain", "precision"])
def test_rt1024x4096_forms_match_oracle(pkg, rt1024_oracle, form_name):
    L = pkg.lib
    p, n, (Eo, do) = rt1024_oracle
    form = 0 if form_name == "plain" else L.FORM_DHT_BF16X3 | L.FORM_SCAN_FP64
    m, z = run_device(pkg, p, n, p["dz_initial"], L.FP32, form)
    try:
        E, d = diagnostics_gpu(m, z)
    finally:
        m.close()
    e = errors(E, d, Eo, do)
    print(f"RT 1024x4096 {form_name}, {n} steps: " + ", ".join(f"{k} {v:.2e}" for k, v in e.items()), flush=True)
    assert e["field"] <= 1e-4 and e["fluence"] <= 1e-4
    assert e["peak_intensity"] <= 2e-4 and e["onaxis_spectrum"] <= 2e-4
    assert e["max_electron_density"] <= 2e-3


def test_rt64x512_precision_form_matches_oracle(pkg, orc):
    L = pkg.lib
    p = pkg.workloads.rt_air(64, 512, 6.0, tmin=-125.0, tmax=125.0, tau0=20e-15)
    n, dz = 30, 5e-3
    o = run_oracle(orc, p, n, dz)
    Eo, do = diagnostics_oracle(o, n * dz)
    res = {}
    for name, form in (("plain", 0), ("precision", L.FORM_DHT_BF16X3 | L.FORM_SCAN_FP64)):
        m, z = run_device(pkg, p, n, dz, L.FP32, form)
        try:
            res[name] = errors(*diagnostics_gpu(m, z), Eo, do)
        finally:
            m.close()
        print(f"RT 64x512 {name}, {n} steps: " + ", ".join(f"{k} {v:.2e}" for k, v in res[name].items()))
    e = res["precision"]
    assert e["field"] <= 1e-4 and e["fluence"] <= 1e-4 and e["peak_intensity"] <= 2e-4
    assert e["max_electron_density"] <= 2e-3 and e["onaxis_spectrum"] <= 2e-4
