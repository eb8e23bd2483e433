"""ctypes binding of libfilament_hip.so (include/filament_hip.h).

This is the reference-side FFI a host would write: the Julia host would hold
the same structs behind ``ccall``.  Loading fails loudly when the library has
not been built -- there is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libfilament_hip.so")

ABI_VERSION = 5
MAXCOMP = 4
GEOM_RT, GEOM_XYT, GEOM_T, GEOM_R, GEOM_XY = 0, 1, 2, 3, 4
ALG_RK3, ALG_RK4 = 3, 4
FP32, FP64 = 4, 8
FORM_DHT_BF16X3, FORM_SCAN_FP64 = 1, 2
RESP_CUBIC_THG, RESP_CUBIC_NOTHG, RESP_FREE, RESP_LOSSES = 1, 2, 4, 8
RESP_RAMAN_THG, RESP_RAMAN_NOTHG = 16, 32
TILING_RULE, TILING_128, TILING_64 = 0, 1, 2

_dp = ctypes.POINTER(ctypes.c_double)


class FilDesc(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "abi_version", "geometry", "Nr", "Nt", "precision", "form", "nonlinearity", "plasma",
        "responses", "losses_ereal", "plasma_ereal", "plasma_kdep", "ncomp", "keep_plasma", "device",
        "Nx", "Ny", "guard3d_bug", "batch", "nonparaxial")]


class FilTables(ctypes.Structure):
    _fields_ = [
        ("r", _dp), ("dr", _dp), ("t", _dp), ("dt", ctypes.c_double),
        ("KZ", _dp), ("Kguard", _dp), ("Wguard", _dp), ("Rguard", _dp), ("Tguard", _dp),
        ("dht_T", _dp), ("dht_RJ", _dp), ("dht_JV", _dp),
        ("QZ", _dp), ("Rnl_cubic", _dp), ("Rnl_free", _dp), ("Rnl_losses", _dp),
        ("Rnl_raman", _dp), ("Hraman", _dp),
        ("rho0", ctypes.c_double),
        ("frhont", ctypes.c_double * MAXCOMP), ("Rava", ctypes.c_double * MAXCOMP),
        ("K", ctypes.c_double * MAXCOMP), ("ntab", ctypes.c_int32 * MAXCOMP),
        ("tab_log10x", _dp * MAXCOMP), ("tab_log10y", _dp * MAXCOMP),
        ("QZ_kw", _dp),
    ]


class FilTablesXYT(ctypes.Structure):
    _fields_ = [("x", _dp), ("y", _dp), ("dx", ctypes.c_double), ("dy", ctypes.c_double),
                ("kt2x", _dp), ("kt2y", _dp), ("Xguard", _dp), ("Yguard", _dp), ("KXguard", _dp), ("KYguard", _dp),
                ("beta", _dp), ("w_vf", _dp), ("zu", ctypes.c_double), ("kparaxial", ctypes.c_int32)]


class FilDiagXYT(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("z", "Fmax", "Imax", "rhomax", "De", "ax", "ay", "W")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class FilDiag(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("z", "Fmax", "Imax", "rhomax", "De", "rfil", "rpl", "tau", "W")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class FilCorrDesc(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("nr", "t0", "nt", "pool_r", "pool_t", "capacity", "batch_max")] + \
               [(n, ctypes.c_double) for n in ("w0", "tmin", "dt", "a_scale", "rho_scale")]


class FilSpatialDesc(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "abi_version", "geometry", "Nr", "Nx", "Ny", "precision", "nonlinearity", "nonparaxial", "alg", "device")]


class FilSpatialTables(ctypes.Structure):
    _fields_ = [("KZ", _dp), ("Kguard", _dp), ("KXguard", _dp), ("KYguard", _dp),
                ("Rguard", _dp), ("Xguard", _dp), ("Yguard", _dp),
                ("dht_T", _dp), ("dht_RJ", _dp), ("dht_JV", _dp), ("r", _dp), ("dr", _dp),
                ("x", _dp), ("y", _dp), ("dx", ctypes.c_double), ("dy", ctypes.c_double),
                ("Rnl", ctypes.c_double), ("QZ", _dp)]


class FilDiagSpatial(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("z", "Imax", "rfil", "ax", "ay", "P")] + \
               [("imax", ctypes.c_int64), ("jmax", ctypes.c_int64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


CORR_NPARAM = 20258

STOP_ZMAX, STOP_ISTOP, STOP_MAXSTEPS, STOP_NONFINITE = 1, 2, 3, 4
NOALLOC_SCRATCH_BYTES = 48 << 20


class FilAdaptParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("zmax", "dz0", "phimax", "phik", "phip", "Istop")] + \
               [("nonlinearity", ctypes.c_int32), ("chunk", ctypes.c_int32)]

EXPORTS = ("fil_last_error", "fil_create", "fil_set_tables", "fil_set_tables_xyt", "fil_upload_field",
           "fil_download_field", "fil_upload_member", "fil_download_member", "fil_zstep", "fil_zstep_parts",
           "fil_analyze",
           "fil_analyze_member", "fil_analyze_xyt", "fil_download_plasma", "fil_synchronize", "fil_destroy",
           "fil_timing_enable", "fil_timing_read", "fil_state_save", "fil_state_restore", "fil_state_copy",
           "fil_fluence", "fil_group_create", "fil_group_zstep", "fil_group_stats", "fil_group_destroy",
           "fil_set_tiling", "fil_get_tiling", "fil_run_adaptive", "fil_run_adaptive_noalloc",
           "fil_corrector_init", "fil_corrector_params", "fil_corrector_scale", "fil_corrector_features",
           "fil_corrector_store_pair", "fil_corrector_train", "fil_corrector_eval", "fil_corrector_forward",
           "fil_corrector_apply", "fil_spatial_create", "fil_spatial_set_tables", "fil_spatial_upload_field",
           "fil_spatial_download_field", "fil_spatial_zstep", "fil_spatial_analyze", "fil_spatial_synchronize",
           "fil_spatial_destroy")

_lib = None


class FilamentError(RuntimeError):
    pass


def load():
    """Load libfilament_hip.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise FilamentError(f"{LIBPATH} is missing: build it first (make -C filament.jl_amd/csrc)")
    L = ctypes.CDLL(LIBPATH)
    L.fil_last_error.restype = ctypes.c_char_p
    L.fil_create.argtypes = [ctypes.POINTER(FilDesc), ctypes.POINTER(ctypes.c_void_p)]
    L.fil_set_tables.argtypes = [ctypes.c_void_p, ctypes.POINTER(FilTables)]
    L.fil_set_tables_xyt.argtypes = [ctypes.c_void_p, ctypes.POINTER(FilTables), ctypes.POINTER(FilTablesXYT)]
    L.fil_analyze_xyt.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(FilDiagXYT), _dp, _dp, _dp, _dp]
    L.fil_upload_field.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int64]
    L.fil_download_field.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int64]
    L.fil_upload_member.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_int64]
    L.fil_download_member.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_int64]
    L.fil_analyze_member.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.POINTER(FilDiag), _dp,
                                     _dp, _dp, _dp]
    L.fil_group_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    L.fil_group_zstep.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
    L.fil_group_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
    L.fil_group_destroy.argtypes = [ctypes.c_void_p]
    L.fil_set_tiling.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.fil_get_tiling.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    L.fil_zstep.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
    L.fil_zstep_parts.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    L.fil_run_adaptive.argtypes = [ctypes.c_void_p, ctypes.POINTER(FilAdaptParams), _dp, ctypes.c_int, _dp,
                                   ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.fil_run_adaptive_noalloc.argtypes = L.fil_run_adaptive.argtypes
    L.fil_analyze.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(FilDiag), _dp, _dp, _dp, _dp]
    L.fil_download_plasma.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int64]
    L.fil_synchronize.argtypes = [ctypes.c_void_p]
    L.fil_destroy.argtypes = [ctypes.c_void_p]
    L.fil_timing_enable.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.fil_timing_read.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_int64)]
    L.fil_state_save.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.fil_state_restore.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.fil_fluence.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
    L.fil_state_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    _fp = ctypes.POINTER(ctypes.c_float)
    L.fil_corrector_init.argtypes = [ctypes.c_void_p, ctypes.POINTER(FilCorrDesc), ctypes.c_uint64]
    L.fil_corrector_params.argtypes = [ctypes.c_void_p, _fp, ctypes.c_int]
    L.fil_corrector_scale.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int]
    L.fil_corrector_features.argtypes = [ctypes.c_void_p, _fp]
    L.fil_corrector_store_pair.argtypes = [ctypes.c_void_p, _fp, _fp]
    L.fil_corrector_train.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_uint64,
                                      _dp]
    L.fil_corrector_eval.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _fp]
    L.fil_corrector_forward.argtypes = [ctypes.c_void_p, _fp, _fp]
    L.fil_corrector_apply.argtypes = [ctypes.c_void_p]
    L.fil_spatial_create.argtypes = [ctypes.POINTER(FilSpatialDesc), ctypes.POINTER(ctypes.c_void_p)]
    L.fil_spatial_set_tables.argtypes = [ctypes.c_void_p, ctypes.POINTER(FilSpatialTables)]
    L.fil_spatial_upload_field.argtypes = [ctypes.c_void_p, _dp]
    L.fil_spatial_download_field.argtypes = [ctypes.c_void_p, _dp]
    L.fil_spatial_zstep.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
    L.fil_spatial_analyze.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(FilDiagSpatial), _dp, _dp, _dp]
    L.fil_spatial_synchronize.argtypes = [ctypes.c_void_p]
    L.fil_spatial_destroy.argtypes = [ctypes.c_void_p]
    for name in EXPORTS:
        if name != "fil_last_error":
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise FilamentError(f"libfilament_hip error {rc}: {load().fil_last_error().decode()}")


def ptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def fptr(a):
    """float32 array -> float*"""
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if a is not None else None
