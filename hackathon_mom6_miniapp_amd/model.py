"""Host-side driver of one split-RK2 run on the GPU (the Python counterpart of
app/rk2_cuda_c_driver.F90): builds the device model from a Config and a case,
advances it with mom6h_model_step and reads fields back."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np

from ._lib import MConfig, check, lib, track
from .state import BC_PERIODIC, BC_WALL, Config, Seeded

# inputs uploaded before mom6h_model_prepare (2D first: masks feed the 3D recipes)
_UPLOAD = ("eta", "ubt", "vbt", "taux", "tauy", "CoriolisBu",
           "mask2dT", "mask2dCu", "mask2dCv", "mask2dBu", "u", "v", "h0")


def to_mconfig(cfg: Config) -> MConfig:
    c = MConfig()
    c.ni, c.nj, c.nk, c.bt_nsteps = cfg.ni, cfg.nj, cfg.nk, cfg.bt_nsteps
    c.dt, c.dx_km = cfg.dt, cfg.dx_km
    c.periodic_x = int(cfg.bc_x == BC_PERIODIC)
    c.periodic_y = int(cfg.bc_y == BC_PERIODIC)
    c.walls_y = int(cfg.bc_y == BC_WALL)
    c.Kh, c.Ah = cfg.Kh, cfg.Ah
    c.laplacian, c.biharmonic, c.no_slip = int(cfg.laplacian), int(cfg.biharmonic), int(cfg.no_slip)
    c.smagorinsky, c.leith, c.better_bound = int(cfg.smagorinsky), int(cfg.leith), int(cfg.better_bound)
    c.coupled = int(cfg.coupled)
    c.ray3d = int(cfg.ray_profile != "bottom")
    c.Kv, c.Kv_ml, c.Kv_extra_bbl, c.Hmix, c.Hbbl = cfg.Kv, cfg.Kv_ml, cfg.Kv_extra_bbl, cfg.Hmix, cfg.Hbbl
    c.coriolis_scheme, c.monotonic = cfg.coriolis_scheme, int(cfg.monotonic)
    c.meridional = int(cfg.meridional)
    if cfg.precision not in ("fp64", "mixed"):
        raise ValueError(f"precision must be 'fp64' or 'mixed', not {cfg.precision!r}")
    c.precision = 1 if cfg.precision == "mixed" else 0
    return c


class Model:
    def __init__(self, cfg: Config, case: dict | None, graph: bool = False, handle=None):
        """A model run on the device.  ``handle``: wrap a member of an
        ensemble (created and prepared by the ensemble) instead of creating one."""
        self.cfg = cfg
        self.L = lib()
        self.owned = handle is None
        self.m = C.c_void_p()
        if handle is not None:
            self.m = handle
            return
        mc = to_mconfig(cfg)
        check(self.L.mom6h_model_create(C.byref(self.m), C.byref(mc)), "mom6h_model_create")
        track(self)
        self.fill_inputs(case)
        check(self.L.mom6h_model_set_option(self.m, b"graph", float(graph)), "set_option")
        check(self.L.mom6h_model_prepare(self.m), "mom6h_model_prepare")

    def fill_inputs(self, case: dict):
        """Upload (or generate on the device) the run's inputs; before prepare."""
        cfg = self.cfg
        for name in _UPLOAD:
            val = case[name]
            if isinstance(val, Seeded):
                self.generate(name, val)
            else:
                self.upload(name, val)
        check(self.L.mom6h_device_synchronize(), "synchronize")
        check(self.L.mom6h_memcpy_d2d(self.device_ptr("h"), self.device_ptr("h0"),
                                      self.field_bytes("h0") * self._count("h0")), "h = h0")
        ray = np.full((cfg.n1, cfg.n2), cfg.ray_bottom, order="F")
        self.upload("ray_u", ray)
        self.upload("ray_v", ray)
        if cfg.ray_profile != "bottom":
            self.upload("ray_u3d", case["Ray_u"])
            self.upload("ray_v3d", case["Ray_v"])

    def close(self):
        if self.m and self.owned:
            self.L.mom6h_model_destroy(self.m)
        self.m = C.c_void_p()

    def __del__(self):
        # at interpreter exit the exit hook has closed it already; make no
        # device call while modules are being torn down
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def _count(self, name: str) -> int:
        n = C.c_long()
        check(self.L.mom6h_model_field(self.m, name.encode(), None, C.byref(n)), f"field {name}")
        return n.value

    def device_ptr(self, name: str) -> int:
        p, n = C.c_void_p(), C.c_long()
        check(self.L.mom6h_model_field(self.m, name.encode(), C.byref(p), C.byref(n)), f"field {name}")
        return p.value

    def upload(self, name: str, arr):
        a = np.asfortranarray(np.asarray(arr, dtype=np.float64))
        check(self.L.mom6h_model_upload(self.m, name.encode(), a.ctypes.data, a.size), f"upload {name}")

    def field_bytes(self, name: str) -> int:
        b = C.c_int()
        check(self.L.mom6h_model_field_bytes(self.m, name.encode(), C.byref(b)), f"field {name}")
        return b.value

    def generate(self, name: str, rec: Seeded):
        """Fill a 3D field from its seeded recipe on the device (FP64 values,
        rounded once for an FP32 field)."""
        n = self._count(name)
        mask = self.device_ptr(rec.mask) if rec.mask else None
        args = (n, rec.seed, rec.field, int(rec.mean is not None), rec.mean or 0.0, rec.sigma,
                mask, self.cfg.n1 * self.cfg.n2, None)
        if self.field_bytes(name) == 8:
            check(self.L.mom6h_seeded_normal(self.device_ptr(name), *args), f"generate {name}")
            return
        tmp = C.c_void_p()
        check(self.L.mom6h_malloc(C.byref(tmp), 8 * n), "malloc")
        try:
            check(self.L.mom6h_seeded_normal(tmp, *args), f"generate {name}")
            check(self.L.mom6h_device_synchronize(), "synchronize")
            check(self.L.mom6h_model_fill(self.m, name.encode(), tmp, n), f"fill {name}")
        finally:
            self.L.mom6h_free(tmp)

    def get(self, name: str) -> np.ndarray:
        n = self._count(name)
        out = np.empty(n, dtype=np.float64)
        check(self.L.mom6h_model_download(self.m, name.encode(), out.ctypes.data, n), f"download {name}")
        n2 = self.cfg.n1 * self.cfg.n2
        shape = (self.cfg.n1, self.cfg.n2) if n == n2 else (self.cfg.n1, self.cfg.n2, n // n2)
        return out.reshape(shape, order="F")

    def step(self, niter: int = 1):
        check(self.L.mom6h_model_step(self.m, niter), "mom6h_model_step")

    def synchronize(self):
        check(self.L.mom6h_model_synchronize(self.m), "mom6h_model_synchronize")

    def set_option(self, key: str, value: float):
        check(self.L.mom6h_model_set_option(self.m, key.encode(), float(value)), f"option {key}")

    def stage_times(self) -> dict:
        ms = (C.c_double * 16)()
        names = C.create_string_buffer(512)
        n = self.L.mom6h_model_stage_times(self.m, ms, 16, names, 512)
        keys = names.value.decode().split(",")
        return {keys[i]: ms[i] for i in range(n)}

    def main_stream(self) -> int:
        s = C.c_void_p()
        check(self.L.mom6h_model_main_stream(self.m, C.byref(s)), "main_stream")
        return s.value or 0

    def autotune(self, reps: int = 5) -> str:
        buf = C.create_string_buffer(512)
        check(self.L.mom6h_model_autotune(self.m, reps, buf, 512), "mom6h_model_autotune")
        return buf.value.decode()

    def diagnostics(self) -> dict:
        """The serial driver's diagnostics (option "diag" must be on): KE and
        mass of the current state, diffu_sum and diffv_sum of the last
        predictor; each (min, max, mean, rms)."""
        out = (C.c_double * 16)()
        check(self.L.mom6h_model_diagnostics(self.m, out), "mom6h_model_diagnostics")
        return {name: tuple(out[4 * q:4 * q + 4]) for q, name in enumerate(DIAG_NAMES)}


# order of mom6h_model_diagnostics' 16 values
DIAG_NAMES = ("KE", "diffu_sum", "diffv_sum", "mass")


class Ensemble:
    """Members of different sizes advanced together: one launch per stage over
    every member (mom6h_ensemble_*).  ``members`` are Model views on the
    members' device state."""

    def __init__(self, cfgs: list, cases: list, graph: bool = False):
        self.L = lib()
        self.e = C.c_void_p()
        arr = (MConfig * len(cfgs))(*[to_mconfig(c) for c in cfgs])
        check(self.L.mom6h_ensemble_create(C.byref(self.e), arr, len(cfgs)), "mom6h_ensemble_create")
        track(self)
        self.members = []
        for q, (cfg, case) in enumerate(zip(cfgs, cases)):
            h = C.c_void_p()
            check(self.L.mom6h_ensemble_member(self.e, q, C.byref(h)), "mom6h_ensemble_member")
            mv = Model(cfg, None, handle=h)
            mv.fill_inputs(case)
            self.members.append(mv)
        check(self.L.mom6h_ensemble_set_option(self.e, b"graph", float(graph)), "set_option")
        check(self.L.mom6h_ensemble_prepare(self.e), "mom6h_ensemble_prepare")

    def step(self, niter: int = 1):
        check(self.L.mom6h_ensemble_step(self.e, niter), "mom6h_ensemble_step")

    def synchronize(self):
        check(self.L.mom6h_ensemble_synchronize(self.e), "mom6h_ensemble_synchronize")

    def set_option(self, key: str, value: float):
        check(self.L.mom6h_ensemble_set_option(self.e, key.encode(), float(value)), f"option {key}")

    def autotune(self, reps: int = 5) -> str:
        buf = C.create_string_buffer(512)
        check(self.L.mom6h_ensemble_autotune(self.e, reps, buf, 512), "mom6h_ensemble_autotune")
        return buf.value.decode()

    def close(self):
        for mv in self.members:
            mv.m = C.c_void_p()
        if self.e:
            self.L.mom6h_ensemble_destroy(self.e)
            self.e = C.c_void_p()

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass
