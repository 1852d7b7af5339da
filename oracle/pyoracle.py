"""ctypes binding of the C oracle (oracle/mom6_oracle.c) -- test infrastructure."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build(force: bool = False) -> str:
    """Compile the oracle with gcc (no FMA contraction, no fast-math)."""
    src = os.path.join(HERE, "mom6_oracle.c")
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(src):
        subprocess.check_call(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off",
                               "-fno-fast-math", "-o", LIB, src, "-lm"])
    return LIB


class OmConfig(C.Structure):
    _fields_ = [("ni", C.c_int), ("nj", C.c_int), ("nk", C.c_int), ("bt_nsteps", C.c_int),
                ("dt", C.c_double), ("dx_km", C.c_double), ("bc_x", C.c_int), ("bc_y", C.c_int),
                ("coriolis_mode", C.c_int), ("f0", C.c_double), ("beta", C.c_double),
                ("Kh", C.c_double), ("Ah", C.c_double),
                ("laplacian", C.c_int), ("biharmonic", C.c_int), ("smagorinsky", C.c_int),
                ("no_slip", C.c_int), ("bound_Kh", C.c_int), ("better_bound", C.c_int),
                ("Kv", C.c_double), ("Kv_ml", C.c_double), ("Kv_extra_bbl", C.c_double),
                ("Hmix", C.c_double), ("Hbbl", C.c_double), ("ray_bottom", C.c_double),
                ("coriolis_scheme", C.c_int), ("monotonic", C.c_int),
                ("simple_2nd", C.c_int), ("upwind_1st", C.c_int), ("meridional", C.c_int),
                ("diag", C.c_int), ("leith", C.c_int), ("coupled", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = C.CDLL(build())
        _lib.om_create.restype = C.c_void_p
        _lib.om_create.argtypes = [C.POINTER(OmConfig)]
        _lib.om_destroy.argtypes = [C.c_void_p]
        _lib.om_field.restype = C.POINTER(C.c_double)
        _lib.om_field.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_long)]
        _lib.om_prepare.argtypes = [C.c_void_p]
        _lib.om_step.argtypes = [C.c_void_p, C.c_int]
        _lib.om_verify.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        _lib.om_diag.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        _lib.om_config_default.argtypes = [C.POINTER(OmConfig)]
        d = C.POINTER(C.c_double)
        _lib.om_compute_transports.argtypes = [C.c_void_p, d, d, d, d, d]
        _lib.om_hor_visc.argtypes = [C.c_void_p, d, d, d, d, d]
        _lib.om_coradcalc.argtypes = [C.c_void_p, d, d, d, d, d, d, d]
        _lib.om_vert_visc.argtypes = [C.c_void_p, d, d, d, C.c_double]
        _lib.om_btstep.argtypes = [C.c_void_p, d, d, d, d, d, d]
        _lib.om_continuity_ppm.argtypes = [C.c_void_p, d, d, d, d, C.c_double]
        _lib.om_continuity_ppm_y.argtypes = [C.c_void_p, d, d, d, d, C.c_double]
        _lib.om_halo_fill.argtypes = [C.c_void_p, d, C.c_int]
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.f_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """One oracle model: grid + control structures + driver arrays."""

    def __init__(self, cfg, case: dict | None = None, diag: bool = False):
        L = lib()
        c = OmConfig()
        L.om_config_default(C.byref(c))
        for name in ("ni", "nj", "nk", "bt_nsteps", "dt", "dx_km", "bc_x", "bc_y", "Kh", "Ah",
                     "Kv", "Kv_ml", "Kv_extra_bbl", "Hmix", "Hbbl", "ray_bottom", "coriolis_scheme"):
            setattr(c, name, getattr(cfg, name))
        c.laplacian = int(cfg.laplacian)
        c.biharmonic = int(cfg.biharmonic)
        c.smagorinsky = int(cfg.smagorinsky)
        c.leith = int(cfg.leith)
        c.no_slip = int(cfg.no_slip)
        c.better_bound = int(cfg.better_bound)
        c.coupled = int(cfg.coupled)
        c.monotonic = int(cfg.monotonic)
        c.meridional = int(getattr(cfg, "meridional", False))
        c.diag = int(diag)
        c.coriolis_mode = 0 if cfg.coriolis == "reference" else 1
        c.f0 = cfg.f0
        c.beta = cfg.beta if cfg.coriolis == "betaplane" else 0.0
        self.cfg = cfg
        self.m = L.om_create(C.byref(c))
        if case is not None:
            self.load(case)

    def __del__(self):
        if getattr(self, "m", None):
            lib().om_destroy(self.m)
            self.m = None

    def field(self, name: str) -> np.ndarray:
        n = C.c_long()
        p = lib().om_field(self.m, name.encode(), C.byref(n))
        if not p:
            raise KeyError(name)
        cfg = self.cfg
        n2 = cfg.n1 * cfg.n2
        if n.value == n2:
            shape = (cfg.n1, cfg.n2)
        elif n.value % n2 == 0:
            shape = (cfg.n1, cfg.n2, n.value // n2)
        else:
            shape = (n.value,)
        return np.ctypeslib.as_array(p, shape=(n.value,)).reshape(shape, order="F")

    def load(self, case: dict):
        # (at nk = 1 a 3D field has the size of a 2D one and field() returns it 2D)
        for name, arr in case.items():
            f = self.field(name)
            f[...] = np.asarray(arr).reshape(f.shape, order="F")
        f = self.field("h")
        f[...] = np.asarray(case["h0"]).reshape(f.shape, order="F")
        lib().om_prepare(self.m)

    def step(self, niter: int = 1):
        lib().om_step(self.m, niter)

    def verify(self):
        out = (C.c_double * 7)()
        lib().om_verify(self.m, out)
        return list(out)

    def diagnostics(self) -> dict:
        """om_diag: KE, diffu_sum, diffv_sum, mass as (min, max, mean, rms)."""
        out = (C.c_double * 16)()
        lib().om_diag(self.m, out)
        names = ("KE", "diffu_sum", "diffv_sum", "mass")
        return {n: tuple(out[4 * q:4 * q + 4]) for q, n in enumerate(names)}

    def halo_fill(self, a: np.ndarray):
        lib().om_halo_fill(self.m, _ptr(a), a.shape[2] if a.ndim == 3 else 1)

    # module-level calls on explicit arrays
    def call(self, fn: str, *arrays, scalar=None):
        args = [self.m] + [_ptr(a) for a in arrays]
        if scalar is not None:
            args.append(C.c_double(scalar))
        getattr(lib(), fn)(*args)
