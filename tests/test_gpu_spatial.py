"""The spatial geometries R and XY on the GPU (fil_spatial_*) against the
oracle (oracle/filament_oracle.cpp, Spatial): z-steps with the Kerr term and
diffraction in both precisions, RK3 and RK4, paraxial and non-paraxial
nonlinear prefactor, with active guards; the analyzer against the oracle's on
the same field; and the input guards of the C ABI."""
import math

import numpy as np
import pytest

import _helpers as H

pytestmark = pytest.mark.gpu

TOL = {4: 1e-4, 8: 1e-10}


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


def case(pkg, geometry, beam, qparaxial=True, alg=4, nonlinearity=True):
    """wide: a 1 mm beam (non-paraxial corrections ~1e-8); narrow: a 2 um
    beam, where QZ(kt) differs from QZ(0) by ~1e-2.  dz: 1/40 of the
    diffraction length, or a 0.05 rad Kerr phase when that is smaller (20
    steps stay short of the collapse)."""
    if geometry == "R":
        N = 250                      # not a multiple of 64
    else:
        N = 128
    if beam == "wide":
        kw = dict(a0=1e-3, extent=5.0, guard=1.0, kguard=2.0)
    else:
        kw = dict(a0=2e-6, extent=0.012, guard=0.002, kguard=60.0)
    p = pkg.workloads.spatial_air(geometry, N, power_pcr=3.0, qparaxial=qparaxial, alg=alg,
                                  nonlinearity=nonlinearity, **kw)
    if geometry == "XY":
        p["Ny"] = 64                 # non-square, and y spans half the x window
        p["ymin"], p["ymax"] = p["xmin"] / 2, p["xmax"] / 2
    m = pkg.physics.Medium(p["permittivity"], p["permeability"], p["n2"])
    w0 = 2 * math.pi * pkg.physics.C0 / p["lam0"]
    zd = float(m.k(w0)) * kw["a0"] ** 2
    u = pkg.physics.UnitR(1e-3, 1.0, p["Iu"])
    Eu = pkg.physics.field_unit(u, float(np.real(m.refractive_index(w0))))
    phik = pkg.physics.phi_kerr(u, m, w0, Eu)
    p["dz_initial"] = min(zd / 40, 0.05 / (phik * p["I0"] / p["Iu"]))
    return p


def model(pkg, p, precision, KZ=None):
    cls = pkg.RModel if p["geometry"] == "R" else pkg.XYModel
    return cls(p, precision=precision, KZ=KZ)


def run_both(pkg, orc, p, precision, nsteps):
    """The device driven with the oracle's KZ: KZ = (Kfunc - w0/vf) zu cancels
    ~1e7 1/m, and the two FP64 five-point derivatives behind vf differ by
    ~1e-13 of it, a phase of ~1e-6 over these runs."""
    o = H.spatial_oracle(p)
    m = model(pkg, p, precision, KZ=o.tables()["KZ"])
    np.testing.assert_allclose(m.field(), o.field(), rtol=0, atol=1e-6 * np.abs(o.field()).max())
    dz = p["dz_initial"]
    for k in range(nsteps):
        o.zstep(dz)
        m.zstep((k + 1) * dz, dz)
    return m, o


@pytest.mark.parametrize("geometry", ["R", "XY"])
@pytest.mark.parametrize("precision", [4, 8])
@pytest.mark.parametrize("alg", [3, 4])
@pytest.mark.parametrize("beam,qparaxial", [("wide", True), ("narrow", True), ("narrow", False)])
def test_spatial_steps_match_oracle(pkg, orc, geometry, precision, alg, beam, qparaxial):
    p = case(pkg, geometry, beam, qparaxial, alg)
    m, o = run_both(pkg, orc, p, precision, 20)
    E, Eo = m.field(), o.field()
    err = rel(E, Eo)
    # the Kerr term matters: the linear run differs far beyond the gate
    q = case(pkg, geometry, beam, qparaxial, alg, nonlinearity=False)
    q["dz_initial"] = p["dz_initial"]
    m0 = model(pkg, q, precision, KZ=o.tables()["KZ"])
    for k in range(20):
        m0.zstep((k + 1) * q["dz_initial"], q["dz_initial"])
    effect = rel(m0.field(), Eo)
    m.close()
    m0.close()
    print(f"{geometry} {beam} qparaxial={qparaxial} RK{alg} precision={precision}: rel_l2 {err:.3e} "
          f"(linear run {effect:.3e})")
    assert err < TOL[precision]
    assert effect > 100 * TOL[4]


@pytest.mark.parametrize("geometry", ["R", "XY"])
def test_nonparaxial_term_matters_for_narrow_beams(pkg, orc, geometry):
    """The non-paraxial and paraxial runs of the 2 um beam differ beyond the
    FP32 gate (so the parity above checks the non-paraxial branch)."""
    fields = []
    for qp in (True, False):
        p = case(pkg, geometry, "narrow", qp, 4)
        m = model(pkg, p, 8)
        for k in range(20):
            m.zstep((k + 1) * p["dz_initial"], p["dz_initial"])
        fields.append(m.field())
        m.close()
    d = rel(fields[1], fields[0])
    print(f"{geometry}: non-paraxial vs paraxial {d:.3e}")
    assert d > 30 * TOL[4]


@pytest.mark.parametrize("geometry", ["R", "XY"])
@pytest.mark.parametrize("precision", [4, 8])
def test_spatial_analyzer_matches_oracle(pkg, orc, geometry, precision):
    """FieldAnalyzerR / XY (FieldAnalyzers.jl:52-70, 251-266) of the device
    against the oracle's on the device's own field: the peak and its index,
    the radii (grid nodes) and the power; and the XY zdata lines."""
    p = case(pkg, geometry, "wide")
    m, o = run_both(pkg, orc, p, precision, 10)
    E = m.field()
    o.set_field(E)
    d, do = m.analyze(full=True), o.analyze()
    itol = 1e-6 if precision == 4 else 1e-13
    assert abs(d["Imax"] / do["Imax"] - 1) < itol
    assert abs(d["P"] / do["P"] - 1) < itol
    I = np.abs(E) ** 2
    np.testing.assert_allclose(d["I"], I, rtol=itol, atol=itol * I.max())   # FP32 flushes ~1e-71 to 0
    if geometry == "R":
        assert d["imax"] == do["index"] and d["jmax"] == 0
        assert d["rfil"] == pytest.approx(do["rfil"], rel=1e-12)
    else:
        Nx = p["Nx"]
        assert d["imax"] + Nx * d["jmax"] == do["index"]
        assert d["ax"] == pytest.approx(do["ax"], rel=1e-12) and d["ay"] == pytest.approx(do["ay"], rel=1e-12)
        Ny = p["Ny"]
        np.testing.assert_allclose(d["Izx"], I[:, Ny // 2 - 1], rtol=itol, atol=itol * I.max())
        np.testing.assert_allclose(d["Izy"], I[Nx // 2 - 1, :], rtol=itol, atol=itol * I.max())
    m.close()


def test_spatial_abi_refuses_bad_input(pkg):
    """fil_spatial_create refuses what the kernels do not cover; zstep before
    the field is uploaded is an order error."""
    import ctypes
    L = pkg.lib
    lib = L.load()

    def create(**kw):
        d = L.FilSpatialDesc()
        d.abi_version, d.geometry, d.Nr, d.Nx, d.Ny, d.precision, d.alg = L.ABI_VERSION, L.GEOM_XY, 0, 64, 64, 4, 4
        for k, v in kw.items():
            setattr(d, k, v)
        h = ctypes.c_void_p()
        return lib.fil_spatial_create(ctypes.byref(d), ctypes.byref(h)), h

    assert create(Nx=96)[0] == 1                      # not a power of two
    assert create(Nx=8192)[0] == 1                    # beyond 4096
    assert create(alg=5)[0] == 1
    assert create(geometry=L.GEOM_RT)[0] == 1
    assert create(abi_version=L.ABI_VERSION - 1)[0] == 1
    assert create(geometry=L.GEOM_R, Nr=1)[0] == 1
    rc, h = create()
    assert rc == 0
    assert lib.fil_spatial_zstep(h, 0.0, 1e-3) == 3    # tables not set: FIL_ESTATE
    assert lib.fil_spatial_destroy(h) == 0


@pytest.mark.parametrize("precision", [4, 8])
def test_xy_lines_shorter_than_a_workgroup(pkg, orc, precision):
    """Nx = 32, Ny = 16: the x-line kernel runs its 64 threads over 32 points
    (the loads and writes of the threads past the line are skipped), the y
    columns are 16 points; against the oracle."""
    p = case(pkg, "XY", "wide")
    p["Nx"], p["Ny"] = 32, 16
    m, o = run_both(pkg, orc, p, precision, 20)
    err = rel(m.field(), o.field())
    m.close()
    print(f"XY 32 x 16 precision={precision}: rel_l2 {err:.3e}")
    assert err < TOL[precision]
