"""ctypes loader of libmom6hip.so.  There is no CPU fallback: if the library
is missing, or a call fails, an exception says so."""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmom6hip.so")


class Mom6hError(RuntimeError):
    pass


class MConfig(C.Structure):
    """mom6h_config (include/mom6hip.h)."""
    _fields_ = [("ni", C.c_int), ("nj", C.c_int), ("nk", C.c_int), ("bt_nsteps", C.c_int),
                ("dt", C.c_double), ("dx_km", C.c_double),
                ("periodic_x", C.c_int), ("periodic_y", C.c_int), ("walls_y", C.c_int),
                ("Kh", C.c_double), ("Ah", C.c_double),
                ("laplacian", C.c_int), ("biharmonic", C.c_int), ("no_slip", C.c_int),
                ("Kv", C.c_double), ("Kv_ml", C.c_double), ("Kv_extra_bbl", C.c_double),
                ("Hmix", C.c_double), ("Hbbl", C.c_double),
                ("coriolis_scheme", C.c_int), ("monotonic", C.c_int), ("meridional", C.c_int),
                ("precision", C.c_int), ("smagorinsky", C.c_int), ("leith", C.c_int), ("better_bound", C.c_int),
                ("coupled", C.c_int), ("ray3d", C.c_int)]


_lib = None
# models and ensembles not yet closed: closed by an exit hook registered when
# the library loads, so that no device call is made while the interpreter
# tears its modules down (in no defined order) or after the HIP runtime and a
# profiler's tool library have run their own exit handlers
_LIVE: "weakref.WeakSet" = weakref.WeakSet()
P = C.c_void_p
D = C.POINTER(C.c_double)

_SIGS = {
    "mom6h_last_error": ([], C.c_char_p),
    "mom6h_version": ([], C.c_char_p),
    "mom6h_set_device": ([C.c_int], C.c_int),
    "mom6h_device_synchronize": ([], C.c_int),
    "mom6h_malloc": ([C.POINTER(P), C.c_size_t], C.c_int),
    "mom6h_free": ([P], C.c_int),
    "mom6h_memcpy_h2d": ([P, P, C.c_size_t], C.c_int),
    "mom6h_memcpy_d2h": ([P, P, C.c_size_t], C.c_int),
    "mom6h_memcpy_d2d": ([P, P, C.c_size_t], C.c_int),
    "mom6h_memset": ([P, C.c_int, C.c_size_t], C.c_int),
    "mom6h_stream_create": ([C.POINTER(P)], C.c_int),
    "mom6h_stream_destroy": ([P], C.c_int),
    "mom6h_stream_synchronize": ([P], C.c_int),
    "mom6h_seeded_normal": ([P, C.c_long, C.c_ulonglong, C.c_int, C.c_int, C.c_double, C.c_double,
                             P, C.c_long, P], C.c_int),
    "mom6h_bt_cfl": ([C.c_double, C.c_int, C.c_double, C.c_double, D, D, C.POINTER(C.c_int),
                      C.POINTER(C.c_int)], C.c_int),
    "mom6h_barotropic_init": ([C.POINTER(P)] + [C.c_int] * 8 + [C.c_int, C.c_double, C.c_double, C.c_int]
                              + [D] * 13 + [C.c_int, C.c_int], C.c_int),
    "mom6h_btstep": ([P] + [P] * 6 + [P], C.c_int),
    "mom6h_btstep_noalloc": ([P] + [P] * 6 + [P], C.c_int),
    "mom6h_barotropic_end": ([P], C.c_int),
    "mom6h_model_diagnostics": ([P, D], C.c_int),
    "mom6h_diag_stats_work_doubles": ([], C.c_long),
    "mom6h_diag_product_sum": ([C.c_int] * 9 + [P, P, C.c_int, P, P], C.c_int),
    "mom6h_diag_ke": ([C.c_int] * 9 + [P, P, P, P], C.c_int),
    "mom6h_diag_mass": ([C.c_int] * 9 + [P, P, P], C.c_int),
    "mom6h_diag_stats": ([C.c_int] * 9 + [P, P, P, P], C.c_int),
    "mom6h_model_create": ([C.POINTER(P), C.POINTER(MConfig)], C.c_int),
    "mom6h_model_field": ([P, C.c_char_p, C.POINTER(P), C.POINTER(C.c_long)], C.c_int),
    "mom6h_model_field_bytes": ([P, C.c_char_p, C.POINTER(C.c_int)], C.c_int),
    "mom6h_model_upload": ([P, C.c_char_p, P, C.c_long], C.c_int),
    "mom6h_model_fill": ([P, C.c_char_p, P, C.c_long], C.c_int),
    "mom6h_model_download": ([P, C.c_char_p, P, C.c_long], C.c_int),
    "mom6h_model_prepare": ([P], C.c_int),
    "mom6h_model_step": ([P, C.c_int], C.c_int),
    "mom6h_model_synchronize": ([P], C.c_int),
    "mom6h_model_set_option": ([P, C.c_char_p, C.c_double], C.c_int),
    "mom6h_model_stage_times": ([P, D, C.c_int, C.c_char_p, C.c_int], C.c_int),
    "mom6h_model_main_stream": ([P, C.POINTER(P)], C.c_int),
    "mom6h_model_destroy": ([P], C.c_int),
    "mom6h_model_autotune": ([P, C.c_int, C.c_char_p, C.c_int], C.c_int),
    "mom6h_ensemble_create": ([C.POINTER(P), C.POINTER(MConfig), C.c_int], C.c_int),
    "mom6h_ensemble_member": ([P, C.c_int, C.POINTER(P)], C.c_int),
    "mom6h_ensemble_prepare": ([P], C.c_int),
    "mom6h_ensemble_step": ([P, C.c_int], C.c_int),
    "mom6h_ensemble_synchronize": ([P], C.c_int),
    "mom6h_ensemble_set_option": ([P, C.c_char_p, C.c_double], C.c_int),
    "mom6h_ensemble_autotune": ([P, C.c_int, C.c_char_p, C.c_int], C.c_int),
    "mom6h_ensemble_destroy": ([P], C.c_int),
}


def lib() -> C.CDLL:
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise Mom6hError(f"{LIB_PATH} is missing: build it with "
                             "`python -m hackathon_mom6_miniapp_amd.build` (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
        atexit.register(close_all)
    return _lib


def track(obj) -> None:
    """Register a live model or ensemble (it has a close() method)."""
    _LIVE.add(obj)


def close_all() -> None:
    """Close every model and ensemble still open (ensembles first: they own
    their members' device state).  Runs at interpreter exit."""
    live = list(_LIVE)
    for obj in sorted(live, key=lambda o: 0 if hasattr(o, "members") else 1):
        try:
            obj.close()
        except Exception:
            pass
    _LIVE.clear()


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().mom6h_last_error().decode(errors="replace")
        raise Mom6hError(f"{what or 'mom6h call'} failed (code {rc}): {msg}")
