"""Loader of the in-tree HIP library ``libpgalf_density.so`` (ctypes).

The library is the product: there is no CPU fallback.  If it is missing, or a
device call is attempted without a GPU, the error is raised loudly.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# PGALF_LIB selects another build of the library (e.g. one with tuned
# compile-time capacities); the default is the in-tree build.
LIB_PATH = os.environ.get("PGALF_LIB") or os.path.join(_HERE, "libpgalf_density.so")

_lib = None

_c_p = ctypes.c_void_p
_c_ll = ctypes.c_longlong
_c_i = ctypes.c_int
_c_f = ctypes.c_float
_c_d = ctypes.c_double


class pgalf_grid_t(ctypes.Structure):
    """pgalf_grid_t of include/pgalf_density.h."""
    _fields_ = [("nx", _c_i), ("ny", _c_i), ("nz", _c_i),
                ("xmin", _c_d), ("ymin", _c_d), ("zmin", _c_d), ("cellsize", _c_d)]


class pgalf_schedule_t(ctypes.Structure):
    """pgalf_schedule_t of include/pgalf_density.h."""
    _fields_ = [("nmax", _c_ll), ("knn_packed", _c_i), ("knn_xcd_chunk", _c_i),
                ("tsc_unit_entries", _c_i)]


class pgalf_shells_t(ctypes.Structure):
    """pgalf_shells_t of include/pgalf_density.h (CSR shells)."""
    _fields_ = [("part_ptr", _c_p), ("part", _c_p), ("core_ptr", _c_p), ("cores", _c_p),
                ("maxshell", _c_ll), ("maxcores", _c_ll), ("nshell", _c_ll)]


# name -> (restype, argtypes); every symbol of include/pgalf_density.h and
# include/pgalf_catalog.h
SIGNATURES = {
    "pgalf_last_error": (_c_i, []),
    "pgalf_last_error_string": (ctypes.c_char_p, []),
    "pgalf_version": (ctypes.c_char_p, []),
    "assign_density_TSC": (None, [_c_p, _c_i, _c_p, _c_i, _c_i, _c_i, _c_d, _c_d, _c_d,
                                  _c_d, _c_i]),
    "gaussian_Smoothing": (None, [_c_p, _c_i, _c_i, _c_i, _c_d, _c_f]),
    "findDen": (None, [_c_p, _c_i, _c_p, _c_d, _c_d, _c_d, _c_d, _c_p, _c_i, _c_i, _c_i]),
    "findsphdensity": (None, [_c_p, _c_i, _c_p, _c_i, _c_p]),
    "starfindsphdensity": (None, [_c_p, _c_i, _c_p, _c_i, _c_p]),
    "pgalf_near_lists": (_c_i, [_c_p, _c_i, _c_i, _c_p]),
    "pgalf_lag_density_stage": (_c_i, [_c_p, _c_i, _c_i, _c_i, _c_p, _c_p, _c_p, _c_p, _c_p]),
    "pgalf_dev_tsc": (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_p, _c_ll, _c_i,
                             ctypes.POINTER(pgalf_grid_t), _c_p, _c_p]),
    "pgalf_dev_smooth": (_c_i, [_c_p, ctypes.POINTER(pgalf_grid_t), _c_f, _c_p]),
    "pgalf_dev_findden": (_c_i, [_c_p, _c_p, _c_p, _c_ll, ctypes.POINTER(pgalf_grid_t), _c_p,
                                 _c_p, _c_p]),
    "pgalf_dev_knn": (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_p, _c_ll, _c_i, _c_i, _c_p, _c_p,
                             _c_i, _c_p]),
    "pgalf_dev_knn_noalloc": (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_p, _c_ll, _c_i, _c_i, _c_p,
                                     _c_p, _c_i, _c_p]),
    "pgalf_dev_knn_noalloc_max": (_c_ll, []),
    "pgalf_dev_density_stage": (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_ll,
                                       ctypes.POINTER(pgalf_grid_t), _c_f, _c_i, _c_p, _c_p,
                                       _c_p, _c_p, _c_p]),
    "pgalf_dev_density_stage_backward": (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_ll,
                                                ctypes.POINTER(pgalf_grid_t), _c_f, _c_i,
                                                _c_p, _c_p, _c_i, _c_p, _c_p, _c_p, _c_p,
                                                _c_p, _c_p, _c_p]),
    "pgalf_dev_status": (_c_i, [_c_p]),
    "pgalf_grid_cores": (_c_i, [_c_p, _c_i, _c_p, _c_p, _c_i, _c_i, _c_i, _c_d, _c_d, _c_d,
                                _c_d, _c_i, _c_f, _c_i, _c_p, _c_i, _c_p]),
    "pgalf_finddenpeak": (_c_i, [_c_p, _c_i, _c_p, _c_i, _c_i, _c_p, _c_p, _c_i, _c_p]),
    "pgalf_dev_grid_cores": (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_ll, _c_i, _c_p,
                                    ctypes.POINTER(pgalf_grid_t), _c_p, _c_f, _c_i, _c_p, _c_ll,
                                    _c_p, _c_p]),
    "pgalf_dev_finddenpeak": (_c_i, [_c_p, _c_i, _c_p, _c_ll, _c_p, _c_i, _c_f, _c_p, _c_ll,
                                     _c_p, _c_p]),
    "pgalf_merging_peak": (_c_i, [_c_p, _c_i, _c_p, _c_i, _c_i]),
    "pgalf_dev_merging_peak": (_c_i, [_c_p, _c_p, _c_p, _c_ll, _c_p, _c_p, _c_i, _c_f, _c_p,
                                      _c_p, _c_p, _c_p]),
    "pgalf_find_core_density": (_c_i, [_c_p, _c_i, _c_p, _c_i, _c_p, _c_i, _c_p, _c_p, _c_p]),
    "pgalf_shell_thresholds": (_c_i, [_c_p, _c_i, _c_p, _c_i, _c_p]),
    "pgalf_shell_particles_peaks": (_c_i, [_c_i, _c_p, _c_i, _c_p, _c_i, _c_p, _c_p, _c_p, _c_p,
                                           _c_i, _c_p]),
    "pgalf_dev_find_core_density": (_c_i, [_c_p, _c_ll, _c_p, _c_i, _c_p, _c_p, _c_p, _c_i,
                                           _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p]),
    "pgalf_dev_shell_thresholds": (_c_i, [_c_p, _c_ll, _c_p, _c_i, _c_p, _c_p, _c_p]),
    "pgalf_dev_shell_particles_peaks": (_c_i, [_c_p, _c_ll, _c_p, _c_i, _c_p, _c_p, _c_i, _c_p,
                                               _c_i, _c_p, _c_p, _c_p, _c_p]),
    "pgalf_catalog_shell_layout": (_c_ll, [_c_i, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p]),
    "pgalf_dev_catalog_watershed": (_c_i, [_c_i, _c_p, _c_p, _c_p, _c_p, _c_i, _c_p, _c_p, _c_p,
                                           _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p,
                                           _c_p, _c_p, _c_ll, _c_p, _c_p, _c_p]),
    "pgalf_batch_create": (_c_p, [_c_i, _c_p, _c_p, _c_p, _c_i]),
    "pgalf_batch_grid_offset": (_c_ll, [_c_p, _c_i]),
    "pgalf_batch_run": (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_i,
                               _c_p]),
    "pgalf_batch_status": (_c_i, [_c_p, _c_p]),
    "pgalf_batch_run_grad": (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_i,
                                    _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p,
                                    _c_i, _c_f, _c_p, _c_p, _c_p, _c_p, _c_i, _c_p]),
    "pgalf_dev_peak_climb": (_c_i, [_c_p, _c_p, _c_p, _c_ll, ctypes.POINTER(pgalf_grid_t),
                                    _c_p, _c_i, _c_f, _c_p, _c_p, _c_p, _c_p, _c_p]),
    "pgalf_batch_workspace_bytes": (_c_ll, [_c_p, _c_i]),
    "pgalf_batch_destroy": (None, [_c_p]),
    "pgalf_set_schedule": (_c_i, [ctypes.POINTER(pgalf_schedule_t), _c_i, _c_i]),
    "pgalf_get_schedule": (_c_i, [_c_ll, ctypes.POINTER(pgalf_schedule_t),
                                  ctypes.POINTER(_c_i)]),
    "pgalf_catalog_open": (_c_p, [ctypes.c_char_p, ctypes.c_char_p, _c_i]),
    "pgalf_catalog_header": (_c_i, [_c_p, _c_p]),
    "pgalf_catalog_next": (_c_ll, [_c_p, _c_ll, _c_ll, _c_i, ctypes.POINTER(_c_ll)]),
    "pgalf_catalog_fetch": (_c_i, [_c_p] + [_c_p] * 14),
    "pgalf_catalog_records": (_c_p, [_c_p, ctypes.POINTER(_c_ll)]),
    "pgalf_catalog_close": (None, [_c_p]),
    "pgalf_writer_open": (_c_p, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]),
    "pgalf_writer_halo": (_c_i, [_c_p, _c_i, _c_p, _c_p, _c_p, _c_i]),
    "pgalf_writer_close": (_c_i, [_c_p]),
    "pgalf_timing_enable": (None, [_c_i]),
    "pgalf_timing_reset": (None, []),
    "pgalf_timing_get": (_c_i, [ctypes.c_char_p, ctypes.POINTER(_c_d), ctypes.POINTER(_c_i)]),
}


class PgalfError(RuntimeError):
    """Error reported by the library (status code and message)."""

    def __init__(self, code, msg):
        super().__init__("pgalf error %d: %s" % (code, msg))
        self.code = code


def load():
    """Load the library (once).  Raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("%s is missing: build it with galaxyfinder_amd.build.build() "
                               "(hipcc for gfx950)" % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(code=None):
    """Raise PgalfError if the last call (or `code`) failed."""
    lib = load()
    c = lib.pgalf_last_error() if code is None else code
    if c != 0:
        raise PgalfError(c, lib.pgalf_last_error_string().decode())
