"""The C-ABI library loads and exports every entry point declared in
include/filament_hip.h; its structs agree with the ctypes mirror.  No
compute calls (no GPU needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..")
HEADER = os.path.join(ROOT, "include", "filament_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fil_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for n in ("fil_create", "fil_set_tables", "fil_upload_field", "fil_download_field", "fil_zstep",
              "fil_analyze", "fil_download_plasma", "fil_destroy", "fil_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol(pkg):
    L = pkg.lib.load()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert set(pkg.lib.EXPORTS) <= set(declared_functions())


def test_struct_layouts_match_header(pkg):
    # fil_desc: 20 int32 fields; fil_diag: 9 doubles; fil_diag_xyt: 8 doubles
    assert ctypes.sizeof(pkg.lib.FilDesc) == 20 * 4
    assert ctypes.sizeof(pkg.lib.FilDiag) == 9 * 8
    assert ctypes.sizeof(pkg.lib.FilDiagXYT) == 8 * 8
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = text[text.index("typedef struct fil_desc"):text.index("} fil_desc;")]
    fields = re.findall(r"int32_t\s+([a-zA-Z0-9_, ]+);", body)
    names = [f.strip() for group in fields for f in group.split(",")]
    assert names == [n for n, _ in pkg.lib.FilDesc._fields_]
    body = text[text.index("typedef struct fil_tables"):text.index("} fil_tables;")]
    names = re.findall(r"\b([a-zA-Z0-9_]+)(?:\[FIL_MAXCOMP\])?;", body)
    assert names == [n for n, _ in pkg.lib.FilTables._fields_]
    body = text[text.index("typedef struct fil_tables_xyt"):text.index("} fil_tables_xyt;")]
    names = re.findall(r"\b([a-zA-Z0-9_]+)(?:\[FIL_MAXCOMP\])?;", body)
    names = [n for group in re.findall(r"(?:double\*?|int32_t)\s+([a-zA-Z0-9_, ]+);", body) for n in
             (x.strip() for x in group.split(","))]
    assert names == [n for n, _ in pkg.lib.FilTablesXYT._fields_]
    body = text[text.index("typedef struct fil_diag_xyt"):text.index("} fil_diag_xyt;")]
    names = [n.strip() for n in re.search(r"double ([^;]+);", body).group(1).split(",")]
    assert names == [n for n, _ in pkg.lib.FilDiagXYT._fields_]
    body = text[text.index("typedef struct fil_corr_desc"):text.index("} fil_corr_desc;")]
    names = [f.strip() for group in re.findall(r"(?:int32_t|double)\s+([a-zA-Z0-9_, ]+);", body)
             for f in group.split(",")]
    assert names == [n for n, _ in pkg.lib.FilCorrDesc._fields_]
    assert ctypes.sizeof(pkg.lib.FilCorrDesc) == 7 * 4 + 4 + 5 * 8   # 4 bytes of padding before w0
    body = text[text.index("typedef struct fil_adapt_params"):text.index("} fil_adapt_params;")]
    names = [f.strip() for group in re.findall(r"(?:int|double)\s+([a-zA-Z0-9_, ]+);", body)
             for f in group.split(",")]
    assert names == [n for n, _ in pkg.lib.FilAdaptParams._fields_]
    assert ctypes.sizeof(pkg.lib.FilAdaptParams) == 6 * 8 + 2 * 4
    for name in ("ZMAX", "ISTOP", "MAXSTEPS"):
        assert int(re.search(rf"#define FIL_STOP_{name} (\d+)", text).group(1)) == getattr(pkg.lib, f"STOP_{name}")
    nparam = int(re.search(r"#define FIL_CORR_NPARAM (\d+)", text).group(1))
    # 4 -> 32 -> 32 -> 32 -> 2 channels of 3x3 weights plus biases
    assert nparam == pkg.lib.CORR_NPARAM == (32 * 4 * 9 + 32) + 2 * (32 * 32 * 9 + 32) + (2 * 32 * 9 + 2)


def test_errors_are_reported_without_a_device(pkg):
    L = pkg.lib.load()
    h = ctypes.c_void_p()
    rc = L.fil_create(None, ctypes.byref(h))
    assert rc == 1 and b"null" in L.fil_last_error()
    d = pkg.lib.FilDesc()
    d.abi_version = 99
    assert L.fil_create(ctypes.byref(d), ctypes.byref(h)) == 1
    assert b"ABI" in L.fil_last_error()
    c = pkg.lib.FilCorrDesc()
    assert L.fil_corrector_init(None, ctypes.byref(c), 0) == 1
    assert L.fil_corrector_apply(None) == 1 and b"null" in L.fil_last_error()
    assert L.fil_state_save(None, 0) == 1
    assert L.fil_zstep_parts(None, 0.0, 1e-3, 5e-4) == 1 and b"null" in L.fil_last_error()
    n, r = ctypes.c_int(), ctypes.c_int()
    assert L.fil_run_adaptive(None, None, None, 0, None, ctypes.byref(n), ctypes.byref(r)) == 1
    assert L.fil_upload_member(None, 0, None, 0) == 1
    assert L.fil_analyze_member(None, 0, 0.0, None, None, None, None, None) == 1
    g = ctypes.c_void_p()
    assert L.fil_group_create(None, 0, ctypes.byref(g)) == 1 and g.value is None
    assert L.fil_group_zstep(None, 0.0, 1e-3) == 1
    assert L.fil_group_destroy(None) == 0
    # a batch is refused for the XYT geometry before any device call
    d = pkg.lib.FilDesc()
    d.abi_version, d.geometry, d.Nx, d.Ny, d.Nt, d.precision, d.ncomp, d.batch = pkg.lib.ABI_VERSION, 1, 16, 16, 256, 4, 1, 2
    assert L.fil_create(ctypes.byref(d), ctypes.byref(h)) == 1 and b"batch" in L.fil_last_error()


def _struct_fields(text, name):
    body = text[text.index(f"typedef struct {name} {{"):text.index(f"}} {name};")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    out = []
    for decl in re.findall(r"(?:int32_t|int64_t|double|const double\*)\s+([a-zA-Z0-9_, ]+);", body):
        out += [f.strip() for f in decl.split(",")]
    return out


def test_spatial_structs_match_header(pkg):
    """fil_spatial_desc / fil_spatial_tables / fil_diag_spatial field by field."""
    text = open(HEADER).read()
    L = pkg.lib
    for cname, pyname in (("fil_spatial_desc", "FilSpatialDesc"), ("fil_spatial_tables", "FilSpatialTables"),
                          ("fil_diag_spatial", "FilDiagSpatial")):
        assert _struct_fields(text, cname) == [n for n, _ in getattr(L, pyname)._fields_], cname
    assert ctypes.sizeof(L.FilSpatialDesc) == 10 * 4
    assert ctypes.sizeof(L.FilDiagSpatial) == 6 * 8 + 2 * 8
    assert ctypes.sizeof(L.FilSpatialTables) == 14 * 8 + 2 * 8 + 8 + 8
    for name, val in (("FIL_GEOM_R", L.GEOM_R), ("FIL_GEOM_XY", L.GEOM_XY), ("FIL_ALG_RK3", L.ALG_RK3),
                      ("FIL_ALG_RK4", L.ALG_RK4)):
        assert int(re.search(rf"\b{name} = (\d+)", text).group(1)) == val


def test_spatial_errors_are_reported_without_a_device(pkg):
    L = pkg.lib.load()
    h = ctypes.c_void_p()
    assert L.fil_spatial_create(None, ctypes.byref(h)) == 1 and b"null" in L.fil_last_error()
    d = pkg.lib.FilSpatialDesc()
    d.abi_version, d.geometry, d.precision, d.alg, d.Nx, d.Ny = 99, pkg.lib.GEOM_XY, 4, 4, 64, 64
    assert L.fil_spatial_create(ctypes.byref(d), ctypes.byref(h)) == 1 and b"ABI" in L.fil_last_error()
    d.abi_version = pkg.lib.ABI_VERSION
    d.geometry = pkg.lib.GEOM_RT
    assert L.fil_spatial_create(ctypes.byref(d), ctypes.byref(h)) == 1 and b"geometry" in L.fil_last_error()
    d.geometry, d.Nx = pkg.lib.GEOM_XY, 100
    assert L.fil_spatial_create(ctypes.byref(d), ctypes.byref(h)) == 1 and b"power" in L.fil_last_error()
    d.Nx, d.alg = 64, 2
    assert L.fil_spatial_create(ctypes.byref(d), ctypes.byref(h)) == 1 and b"alg" in L.fil_last_error()
    assert L.fil_spatial_zstep(None, 0.0, 1e-3) == 1
    assert L.fil_spatial_upload_field(None, None) == 1
    assert L.fil_spatial_destroy(None) == 0
