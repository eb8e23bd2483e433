"""fil_run_adaptive_noalloc (the device z loop with its buffers in the
library's fixed scratch array) against fil_run_adaptive through the C ABI:
the same steps, stop, log and field bit for bit, on two calls in a row, the
second on a smaller grid (and the other precision) than the first; requests
beyond the scratch array and a missing log are refused with the state
untouched."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def kerr_phimax(pkg, p, precision):
    m = pkg.RTModel(p, precision=precision)
    try:
        return 0.9 * p["dz_initial"] * m.phik * m.analyze(0.0)["Imax"]
    finally:
        m.close()


CASES = [
    # (params, precision, zmax in dz0, max_steps, chunk): stop by zmax inside a chunk, then by the step limit
    (dict(Nr=128, Nt=512, power=8.0, extra=dict(a0=1e-3, seed=3)), 4, 30, 1000, 7),
    (dict(Nr=64, Nt=256, power=6.0, extra=dict(tmin=-125.0, tmax=125.0, tau0=20e-15, seed=5)), 8, 100, 9, 4),
]


def test_noalloc_loop_equals_allocating_loop_on_two_calls(pkg):
    L = pkg.lib
    for k, (c, precision, nz, max_steps, chunk) in enumerate(CASES):
        p = pkg.workloads.rt_air(c["Nr"], c["Nt"], c["power"], dz=4e-3, **c["extra"])
        dz0 = p["dz_initial"]
        phimax = kerr_phimax(pkg, p, precision)
        a = pkg.RTModel(p, precision=precision)
        b = pkg.RTModel(p, precision=precision)
        try:
            ref = a.run_adaptive(0.0, nz * dz0, dz0, phimax, np.inf, max_steps, chunk=chunk)
            got = b.run_adaptive_noalloc(0.0, nz * dz0, dz0, phimax, np.inf, max_steps, chunk=chunk)
            print(f"call {k + 1}: RT {c['Nr']}x{c['Nt']} fp{8 * precision}: {got[1]} steps, reason {got[2]}, "
                  f"z {got[0]:.6g} (allocating loop: {ref[1]} steps, reason {ref[2]})", flush=True)
            assert got[:3] == ref[:3]
            assert got[2] == (L.STOP_ZMAX if k == 0 else L.STOP_MAXSTEPS)
            assert got[3][1:, 1].max() < dz0
            np.testing.assert_array_equal(got[3], ref[3])
            np.testing.assert_array_equal(b.field(), a.field())
        finally:
            a.close()
            b.close()


def test_noalloc_refuses_beyond_scratch_and_without_log(pkg):
    L = pkg.lib
    p = pkg.workloads.rt_air(64, 256, 3.0, tmin=-125.0, tmax=125.0, tau0=20e-15)
    m = pkg.RTModel(p)
    try:
        E0 = m.field()
        big = L.NOALLOC_SCRATCH_BYTES // 32           # log rows alone exceed the array
        with pytest.raises(L.FilamentError, match="maximum"):
            m.run_adaptive_noalloc(0.0, 1.0, 1e-3, np.inf, np.inf, big)
        prm = L.FilAdaptParams(zmax=1.0, dz0=1e-3, phimax=np.inf, phik=m.phik, phip=m.phip, Istop=np.inf,
                               nonlinearity=1, chunk=0)
        zz = np.array([0.0])
        n, r = ctypes.c_int(), ctypes.c_int()
        rc = m._lib.fil_run_adaptive_noalloc(m.h, ctypes.byref(prm), L.ptr(zz), 5, None, ctypes.byref(n),
                                             ctypes.byref(r))
        assert rc == 1 and b"log" in m._lib.fil_last_error()
        np.testing.assert_array_equal(m.field(), E0)
    finally:
        m.close()
