"""Host-side set-up of a split-RK2 run: configuration, grid inputs and initial state.

Everything here runs once on the host before the first step and produces plain
float64 arrays in the reference's Fortran layout (column-major, data domain
``isd=1..ied=ni+2``, ``jsd=1..jed=nj+2``, compute domain ``isc=4..iec=ni-1``;
see ``src/common/mom6_types.F90:145-148``).  Arrays are numpy arrays of shape
``(n1, n2)`` or ``(n1, n2, nk)`` with ``order='F'``, so ``arr.ravel(order='F')``
is exactly the buffer the reference and the HIP library index.

Two initial states exist:

* ``"reference"`` -- the analytic state of ``app/rk2_serial_driver.F90:466-506``
  (sin/cos/exp with the driver's ``3.14159`` literal);
* ``"seeded"`` -- the counter-based recipe used by the benchmark configurations:
  ``h = 4000/nk + N(0, 0.5 m)``, ``u, v ~ N(0, 0.05 m/s)``, ``eta = ubt = vbt = 0``,
  optionally plus a Gaussian sea-surface bump.  Normals are Irwin-Hall sums of 12
  SplitMix64 uniforms, so the same (seed, field, index) gives the same bits on
  any host.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

BC_STATIC, BC_PERIODIC, BC_WALL = 0, 1, 2

# physical constants, src/common/mom6_types.F90:80-84
G_EARTH = 9.80
RHO_0 = 1035.0
OMEGA = 7.2921e-5
PI = 3.14159265358979323846
EARTH_RADIUS = 6.371e6


@dataclasses.dataclass
class Config:
    """Run configuration.  Defaults are the reference serial driver's values
    (``app/rk2_serial_driver.F90:91-93,135-147``)."""

    ni: int = 180
    nj: int = 180
    nk: int = 75
    niter: int = 10
    bt_nsteps: int = 30
    dt: float = 300.0
    dx_km: float = 10.0
    bc_x: int = BC_STATIC
    bc_y: int = BC_STATIC
    # Coriolis parameter at q points: "reference" (f0 = 2 Omega sin 45,
    # beta = 2 Omega cos 45 / R), "fplane" (f0) or "betaplane" (f0 + beta y)
    coriolis: str = "reference"
    f0: float = 1.0e-4
    beta: float = 0.0
    Kh: float = 100.0
    Ah: float = 1.0e10
    laplacian: bool = True
    biharmonic: bool = False
    # hor_visc_init's other closures (mom6_hor_visc_serial.F90:160-193):
    # Smagorinsky and Leith viscosities, no-slip walls, thickness-aware bounds
    smagorinsky: bool = False
    leith: bool = False
    no_slip: bool = False
    better_bound: bool = False
    Kv: float = 1.0e-4
    Kv_ml: float = 1.0e-2
    Kv_extra_bbl: float = 1.0e-2
    Hmix: float = 50.0
    Hbbl: float = 10.0
    ray_bottom: float = 1.0e-4
    # Rayleigh drag profile: "bottom" (the reference driver's: ray_bottom in
    # the bottom layer only) or "linear" (Ray(k) = ray_bottom*k/nk at every
    # level, the 3D vertvisc_type fields)
    ray_profile: str = "bottom"
    coriolis_scheme: int = 1
    monotonic: bool = False
    # continuity also takes a meridional PPM pass after the zonal one (no
    # counterpart in the reference, whose continuity is zonal only)
    meridional: bool = False
    # continuity's barotropic-transport targets from the barotropic solver
    # (uhbt = Datu*ubt_av, vhbt = Datv*vbt_av of the concurrent btstep call);
    # the reference leaves them zero
    coupled: bool = False
    # "fp64": the reference's arithmetic throughout; "mixed": FP32 layer
    # fields and stencils, FP64 vertical solve, barotropic solver and 2D fields
    precision: str = "fp64"

    # ---- derived bounds (mom6_types.F90:145-148)
    @property
    def n1(self) -> int:
        return self.ni + 2

    @property
    def n2(self) -> int:
        return self.nj + 2

    @property
    def bounds(self):
        """(isd, ied, jsd, jed, isc, iec, jsc, jec), 1-based."""
        return (1, self.ni + 2, 1, self.nj + 2, 4, self.ni - 1, 4, self.nj - 1)

    @property
    def periodic(self) -> bool:
        return self.bc_x == BC_PERIODIC or self.bc_y == BC_PERIODIC

    @property
    def cells(self) -> int:
        """Cell updates per step as the benchmark metric counts them: ni*nj*nk."""
        return self.ni * self.nj * self.nk


# ---------------------------------------------------------------- random ---
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(z: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser (Steele, Lea & Flood 2014) on uint64 arrays."""
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def seeded_normal(seed: int, field: int, count: int, sigma: float) -> np.ndarray:
    """N(0, sigma) samples for element indices 0..count-1 of one field.

    Element ``e`` of field ``f`` draws 12 uniforms ``U_r = (x_r >> 11) * 2**-53``
    from ``x_r = splitmix64(seed * 2**32 + f * 2**52 + 12 * e + r)`` (mod 2**64),
    and returns ``sigma * ((U_0 + ... + U_11) - 6)`` summed left to right."""
    e = np.arange(count, dtype=np.uint64)
    with np.errstate(over="ignore"):
        base = (np.uint64(seed) << np.uint64(32)) + (np.uint64(field) << np.uint64(52)) + e * np.uint64(12)
    acc = np.zeros(count, dtype=np.float64)
    scale = 2.0 ** -53
    for r in range(12):
        with np.errstate(over="ignore"):
            x = _splitmix64(base + np.uint64(r))
        acc = acc + (x >> np.uint64(11)).astype(np.float64) * scale
    return sigma * (acc - 6.0)


# ------------------------------------------------------------------ grid ---
def coriolis_bu(cfg: Config) -> np.ndarray:
    """CoriolisBu(i,j) = f0 + beta*(real(j-1-nj/2) - 0.5)*dx (mom6_types.F90:189-190,226)."""
    dx_m = cfg.dx_km * 1000.0
    if cfg.coriolis == "reference":
        f0 = 2.0 * OMEGA * math.sin(45.0 * PI / 180.0)
        beta = 2.0 * OMEGA * math.cos(45.0 * PI / 180.0) / EARTH_RADIUS
    elif cfg.coriolis == "fplane":
        f0, beta = cfg.f0, 0.0
    elif cfg.coriolis == "betaplane":
        f0, beta = cfg.f0, cfg.beta
    else:
        raise ValueError(f"unknown coriolis mode {cfg.coriolis!r}")
    j = np.arange(1, cfg.n2 + 1)
    row = f0 + beta * ((j - 1 - cfg.nj // 2).astype(np.float64) - 0.5) * dx_m
    return np.asfortranarray(np.broadcast_to(row[None, :], (cfg.n1, cfg.n2)).copy())


def masks(cfg: Config):
    """(mask2dT, mask2dCu, mask2dCv, mask2dBu).  All ocean except with walls in y:
    rows outside jsc..jec are land and the v/q lines J = jsc-1 and J = jec are closed."""
    mT = np.ones((cfg.n1, cfg.n2), order="F")
    mCu, mCv, mBu = mT.copy(order="F"), mT.copy(order="F"), mT.copy(order="F")
    if cfg.bc_y == BC_WALL:
        _, _, _, _, _, _, jsc, jec = cfg.bounds
        j = np.arange(1, cfg.n2 + 1)
        land = (j < jsc) | (j > jec)
        closed = (j < jsc) | (j >= jec)
        mT[:, land] = 0.0
        mCu[:, land] = 0.0
        mCv[:, closed] = 0.0
        mBu[:, closed] = 0.0
    return mT, mCu, mCv, mBu


# ----------------------------------------------------------------- state ---
@dataclasses.dataclass
class Seeded:
    """A 3D field of the seeded state, generated where it is used (host or GPU,
    same bits): ``sigma*(sum of 12 uniforms - 6)``, plus ``mean`` if given,
    times the 2D mask named by ``mask`` if given."""

    seed: int
    field: int
    sigma: float
    mean: float | None = None
    mask: str | None = None

    def host(self, cfg: Config, case: dict) -> np.ndarray:
        n3 = cfg.n1 * cfg.n2 * cfg.nk
        a = seeded_normal(self.seed, self.field, n3, self.sigma)
        if self.mean is not None:
            a = self.mean + a
        a = a.reshape((cfg.n1, cfg.n2, cfg.nk), order="F")
        if self.mask is not None:
            a = a * case[self.mask][:, :, None]
        return np.asfortranarray(a)


def make_case(cfg: Config, init: str = "seeded", seed: int = 1234, bump: tuple | None = None,
              lazy3d: bool = False) -> dict:
    """Inputs of one run: u, v, h0 (3D); eta, ubt, vbt, taux, tauy, CoriolisBu,
    mask2dT, mask2dCu, mask2dCv, mask2dBu (2D).  ``bump = (amplitude_m, radius_m)``
    adds a Gaussian sea-surface bump centred on the compute domain."""
    n1, n2, nk = cfg.n1, cfg.n2, cfg.nk
    shp3, shp2 = (n1, n2, nk), (n1, n2)
    if init == "reference":  # rk2_serial_driver.F90:466-506 (pi written 3.14159)
        p = 3.14159
        i = np.arange(1, n1 + 1, dtype=np.float64)[:, None, None]
        j = np.arange(1, n2 + 1, dtype=np.float64)[None, :, None]
        k = np.arange(1, nk + 1, dtype=np.float64)[None, None, :]
        h0 = 4000.0 / nk + 10.0 * np.sin((i - 1) / cfg.ni * p) * np.cos((j - 1) / cfg.nj * p) * np.exp(-k / 20.0)
        u = 0.1 * np.sin((j - 1) / cfg.nj * p * 2.0) * np.exp(-k / 30.0) + 0.0 * i
        v = 0.1 * np.cos((i - 1) / cfg.ni * p * 2.0) * np.exp(-k / 30.0) + 0.0 * j
        i2, j2 = i[:, :, 0], j[:, :, 0]
        eta = 0.5 * np.sin((i2 - 1) / cfg.ni * p * 2.0) * np.cos((j2 - 1) / cfg.nj * p * 2.0)
        ubt = 0.05 * np.sin((j2 - 1) / cfg.nj * p) + 0.0 * i2
        vbt = 0.05 * np.cos((i2 - 1) / cfg.ni * p) + 0.0 * j2
    elif init == "seeded":
        walls = cfg.bc_y == BC_WALL
        h0 = Seeded(seed, 0, 0.5, mean=4000.0 / nk)
        u = Seeded(seed, 1, 0.05, mask="mask2dCu" if walls else None)
        v = Seeded(seed, 2, 0.05, mask="mask2dCv" if walls else None)
        eta = np.zeros(shp2)
        ubt = np.zeros(shp2)
        vbt = np.zeros(shp2)
    else:
        raise ValueError(f"unknown init {init!r}")
    if bump is not None:
        amp, radius = bump
        dx_m = cfg.dx_km * 1000.0
        _, _, _, _, isc, iec, jsc, jec = cfg.bounds
        xc, yc = 0.5 * (isc + iec), 0.5 * (jsc + jec)
        x = (np.arange(1, n1 + 1) - xc)[:, None] * dx_m
        y = (np.arange(1, n2 + 1) - yc)[None, :] * dx_m
        eta = eta + amp * np.exp(-(x * x + y * y) / (radius * radius))
    jj = np.arange(1, n2 + 1, dtype=np.float64)[None, :]
    taux = 0.1 * np.sin((jj - 1) / cfg.nj * 3.14159) + np.zeros(shp2)   # :516-522
    tauy = np.zeros(shp2)
    mT, mCu, mCv, mBu = masks(cfg)
    if cfg.bc_y == BC_WALL:  # no flow on land rows and through the closed faces
        if init != "seeded":
            u = u * mCu[:, :, None]
            v = v * mCv[:, :, None]
        ubt = ubt * mCu
        vbt = vbt * mCv
    out = dict(u=u, v=v, h0=h0, eta=eta, ubt=ubt, vbt=vbt, taux=taux, tauy=tauy,
               CoriolisBu=coriolis_bu(cfg), mask2dT=mT, mask2dCu=mCu, mask2dCv=mCv, mask2dBu=mBu)
    if cfg.ray_profile == "linear":   # Ray(k) = ray_bottom*k/nk, as (ray_bottom*k)/nk
        kk = np.arange(1, nk + 1, dtype=np.float64)
        ray = np.broadcast_to((cfg.ray_bottom * kk / float(nk))[None, None, :], shp3)
        out["Ray_u"] = ray.copy()
        out["Ray_v"] = ray.copy()
    elif cfg.ray_profile != "bottom":
        raise ValueError(f"unknown ray_profile {cfg.ray_profile!r}")
    case = {key: (val if isinstance(val, Seeded) else np.asfortranarray(np.asarray(val, dtype=np.float64)))
            for key, val in out.items()}
    if not lazy3d:
        for key in ("u", "v", "h0"):
            if isinstance(case[key], Seeded):
                case[key] = case[key].host(cfg, case)
    return case


INPUT_ORDER = ("u", "v", "h0", "eta", "ubt", "vbt", "taux", "tauy", "CoriolisBu",
               "mask2dT", "mask2dCu", "mask2dCv", "mask2dBu")
