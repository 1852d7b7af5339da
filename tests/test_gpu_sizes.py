"""Parity at the workload sizes and across tile seams.

The kernels tile the grid: the fused momentum stencil in tiles of 64 compute
cells in x, continuity in blocks of 62 faces, the blocked barotropic kernel in
66 x 32 tiles.  The small cases of test_gpu_model.py fit one tile in x; the
cases here cross at least three seams in every kernel, and two run the
workloads' own sizes:

* 180x180x75, doubly periodic f-plane, bt_nsteps = 30, 10 steps (fused, with
  the autotuned momentum launch shape);
* 720x720x75, beta plane, periodic in x, walls in y, Kh = 1e3, bt_nsteps = 50,
  one step;
* 200x140x12, periodic-x/walls-y and doubly periodic, module and fused modes,
  with and without the meridional pass;
* the large-grid physics (bt_nsteps = 100, a Gaussian sea-surface bump,
  graph-captured steps) at 200x200x20, and its mixed form against FP64 after
  100 steps;
* an ensemble whose members reach 50 and 75 levels and 140 columns, each FP64
  member against the oracle directly.

Bar: bit-exact (np.array_equal) against the oracle for FP64; max relative
error <= 1e-5 in h, u and v for the mixed form against FP64."""
from dataclasses import replace

import numpy as np
import pytest

from hackathon_mom6_miniapp_amd.state import BC_PERIODIC, BC_STATIC, BC_WALL, Config, make_case
from tests.cases import COMPARE, COMPARE_FUSED, MERID_COMPARE

pytestmark = pytest.mark.gpu


def _diff(g, o, f):
    if g.shape != o.shape:
        return f"{f}: shape {g.shape} vs {o.shape}"
    if np.array_equal(g, o):
        return None
    d = np.abs(g - o)
    return f"{f}: max|diff|={d.max():.3e} at {np.unravel_index(d.argmax(), d.shape)} n={(d > 0).sum()}"


def _compare(gpu, ora, cfg, fused, merid=False):
    isd, ied, jsd, jed, isc, iec, jsc, jec = cfg.bounds
    fields = (COMPARE if not fused else COMPARE_FUSED + ("uh",)) + ((MERID_COMPARE + ("vh",)) if merid else ())
    bad = []
    for f in fields:
        g, o = gpu.get(f), ora.field(f)
        if fused and f == "uh":     # the fused stencil keeps its transports on chip
            g, o = g[isc - 2:iec, jsc - 1:jec], o[isc - 2:iec, jsc - 1:jec]
        if fused and f == "vh":
            g, o = g[isc - 1:iec, jsc - 2:jec], o[isc - 1:iec, jsc - 2:jec]
        msg = _diff(g, o, f)
        if msg:
            bad.append(msg)
    return bad


def _device_seeded(cfg, seed, **kw):
    """A model whose seeded u, v, h0 are generated on the device, and the case
    the oracle takes (those arrays read back: the device recipe equals the
    host recipe bit for bit, test_gpu_model.py)."""
    from hackathon_mom6_miniapp_amd.model import Model

    case = make_case(cfg, init="seeded", seed=seed, lazy3d=True)
    gpu = Model(cfg, case, **kw)
    for f in ("u", "v", "h0"):
        case[f] = gpu.get(f)
    return gpu, case


SEAM_CASES = {
    "seam_periodic_walls": Config(ni=200, nj=140, nk=12, bt_nsteps=20, bc_x=BC_PERIODIC, bc_y=BC_WALL,
                                  coriolis="betaplane", f0=1e-4, beta=2e-11, Kh=1e3),
    "seam_doubly_periodic": Config(ni=200, nj=140, nk=12, bt_nsteps=20, bc_x=BC_PERIODIC, bc_y=BC_PERIODIC,
                                   coriolis="fplane", f0=1e-4),
}


@pytest.mark.parametrize("merid", [False, True])
@pytest.mark.parametrize("mode", ["modules", "fused"])
@pytest.mark.parametrize("name", sorted(SEAM_CASES))
def test_tile_seams_bit_exact(name, mode, merid):
    """200 x 140 columns: 4 momentum tiles (modules: 4 hor_visc and CorAdCalc
    tiles), 4 continuity blocks and 4 barotropic tiles in x (5 or more in y),
    three steps."""
    from hackathon_mom6_miniapp_amd.model import Model
    from oracle.pyoracle import Oracle

    cfg = replace(SEAM_CASES[name], meridional=merid)
    print(f"[gpu seams] case={name} mode={mode} meridional={merid}", flush=True)
    case = make_case(cfg, init="seeded", seed=31)
    gpu = Model(cfg, case)
    gpu.set_option("fused", 1.0 if mode == "fused" else 0.0)
    gpu.step(3)
    gpu.synchronize()
    ora = Oracle(cfg, case)
    ora.step(3)
    bad = _compare(gpu, ora, cfg, mode == "fused", merid)
    gpu.close()
    assert not bad, "\n".join(bad)


def test_fplane_180x180x75_ten_steps_bit_exact():
    """The oracle configuration at full size: 180x180x75, doubly periodic,
    f-plane f = 1e-4, bt_nsteps = 30, 10 steps, fused with the autotuned
    momentum launch shape."""
    from hackathon_mom6_miniapp_amd.model import Model
    from oracle.pyoracle import Oracle

    cfg = Config(ni=180, nj=180, nk=75, bt_nsteps=30, bc_x=BC_PERIODIC, bc_y=BC_PERIODIC,
                 coriolis="fplane", f0=1e-4)
    print("[gpu full] 180x180x75 fplane: case", flush=True)
    case = make_case(cfg, init="seeded", seed=180)
    gpu = Model(cfg, case)
    print(f"[gpu full] autotune: {gpu.autotune(3)}", flush=True)
    gpu.step(10)
    gpu.synchronize()
    print("[gpu full] oracle 10 steps", flush=True)
    ora = Oracle(cfg, case)
    ora.step(10)
    bad = _compare(gpu, ora, cfg, fused=True)
    gpu.close()
    assert not bad, "\n".join(bad)


def test_beta_walls_720x720x75_one_step_bit_exact():
    """The drop-in workload at full size: 720x720x75, beta plane, periodic in
    x, walls in y, Kh = 1e3, bottom drag, bt_nsteps = 50, one step."""
    from oracle.pyoracle import Oracle

    cfg = Config(ni=720, nj=720, nk=75, bt_nsteps=50, bc_x=BC_PERIODIC, bc_y=BC_WALL,
                 coriolis="betaplane", f0=1e-4, beta=2e-11, Kh=1e3, Kv=1e-4, ray_bottom=1e-4)
    print("[gpu full] 720x720x75 beta/walls: device state", flush=True)
    gpu, case = _device_seeded(cfg, seed=720)
    gpu.autotune(3)
    gpu.step(1)
    gpu.synchronize()
    print("[gpu full] oracle one step (about 10-30 s)", flush=True)
    ora = Oracle(cfg, case)
    ora.step(1)
    print("[gpu full] compare", flush=True)
    bad = _compare(gpu, ora, cfg, fused=True)
    gpu.close()
    assert not bad, "\n".join(bad)


def _bump_cfg(precision="fp64"):
    return Config(ni=200, nj=200, nk=20, bt_nsteps=100, bc_x=BC_PERIODIC, bc_y=BC_WALL,
                  coriolis="betaplane", f0=1e-4, beta=2e-11, Kh=1e3, precision=precision)


BUMP = (0.5, 100.0e3)


def test_large_grid_physics_bump_graph_bit_exact():
    """bt_nsteps = 100 with a Gaussian sea-surface bump (0.5 m, 100 km),
    graph-captured steps, FP64, against the oracle."""
    from hackathon_mom6_miniapp_amd.model import Model
    from oracle.pyoracle import Oracle

    cfg = _bump_cfg()
    print("[gpu bump] 200x200x20 bt=100 graph", flush=True)
    case = make_case(cfg, init="seeded", seed=1440, bump=BUMP)
    assert case["eta"].max() > 0.49
    gpu = Model(cfg, case, graph=True)
    gpu.step(3)
    gpu.synchronize()
    ora = Oracle(cfg, case)
    ora.step(3)
    bad = _compare(gpu, ora, cfg, fused=True)
    gpu.close()
    assert not bad, "\n".join(bad)


def test_large_grid_physics_bump_mixed_within_bound():
    """The same workload in the mixed form, graph-captured, against FP64 after
    100 steps: max relative error <= 1e-5 in h, u and v; eta bit-identical."""
    from hackathon_mom6_miniapp_amd.model import Model

    case = make_case(_bump_cfg(), init="seeded", seed=1441, bump=BUMP)
    ref = Model(_bump_cfg(), case, graph=True)
    mix = Model(_bump_cfg("mixed"), case, graph=True)
    print("[gpu bump] mixed vs fp64, 100 steps", flush=True)
    ref.step(100)
    mix.step(100)
    ref.synchronize()
    mix.synchronize()
    errs = {}
    for f in ("h", "u", "v"):
        a, b = mix.get(f), ref.get(f)
        errs[f] = np.abs(a - b).max() / np.abs(b).max()
    print(f"[gpu bump] max rel errors {errs}", flush=True)
    assert np.array_equal(mix.get("eta"), ref.get("eta"))
    ref.close()
    mix.close()
    assert all(e <= 1e-5 for e in errs.values()), errs


DEEP_MEMBERS = [
    Config(ni=140, nj=60, nk=50, bt_nsteps=16, bc_x=BC_PERIODIC, bc_y=BC_WALL, coriolis="betaplane", f0=1e-4,
           beta=2e-11, Kh=1e3),
    Config(ni=48, nj=40, nk=75, bt_nsteps=16, bc_x=BC_PERIODIC, bc_y=BC_PERIODIC, coriolis="fplane", f0=1e-4),
    Config(ni=70, nj=34, nk=25, bt_nsteps=16, bc_x=BC_PERIODIC, bc_y=BC_WALL, coriolis="betaplane", f0=1e-4,
           beta=2e-11, Kh=1e3),
]


def test_deep_ensemble_members_bit_exact_with_oracle():
    """An ensemble reaching 75 levels (the deep batched continuity block) and
    140 columns: every FP64 member against the oracle directly."""
    from hackathon_mom6_miniapp_amd.model import Ensemble
    from oracle.pyoracle import Oracle

    cases = [make_case(c, init="seeded", seed=500 + q) for q, c in enumerate(DEEP_MEMBERS)]
    print("[gpu deep ensemble] fp64, 2 steps", flush=True)
    ens = Ensemble(DEEP_MEMBERS, cases)
    ens.step(2)
    ens.synchronize()
    bad = []
    for q, (cfg, case) in enumerate(zip(DEEP_MEMBERS, cases)):
        ora = Oracle(cfg, case)
        ora.step(2)
        bad += [f"member {q}: {m}" for m in _compare(ens.members[q], ora, cfg, fused=True)]
    ens.close()
    assert not bad, "\n".join(bad)


RAGGED_MEMBERS = [
    Config(ni=70, nj=30, nk=20, bt_nsteps=12, bc_x=BC_PERIODIC, bc_y=BC_WALL, coriolis="betaplane", f0=1e-4,
           beta=2e-11, Kh=1e3),
    Config(ni=40, nj=28, nk=40, bt_nsteps=12, bc_x=BC_PERIODIC, bc_y=BC_PERIODIC, coriolis="fplane", f0=1e-4),
    Config(ni=36, nj=24, nk=60, bt_nsteps=12, bc_x=BC_PERIODIC, bc_y=BC_WALL, coriolis="betaplane", f0=1e-4,
           beta=2e-11, Kh=1e3),
]


def test_ragged_ensemble_members_bit_exact_with_oracle():
    """Members whose depths fall below their block shape's even spread (20 of
    8 x 4 slots, 40 of 8 x 7, 60 of 8 x 10): the batched continuity runs one
    launch per shape with the plain level mapping; every FP64 member against
    the oracle directly."""
    from hackathon_mom6_miniapp_amd.model import Ensemble
    from oracle.pyoracle import Oracle

    cases = [make_case(c, init="seeded", seed=700 + q) for q, c in enumerate(RAGGED_MEMBERS)]
    print("[gpu ragged ensemble] fp64, 2 steps", flush=True)
    ens = Ensemble(RAGGED_MEMBERS, cases)
    ens.step(2)
    ens.synchronize()
    bad = []
    for q, (cfg, case) in enumerate(zip(RAGGED_MEMBERS, cases)):
        ora = Oracle(cfg, case)
        ora.step(2)
        bad += [f"member {q}: {m}" for m in _compare(ens.members[q], ora, cfg, fused=True)]
    ens.close()
    assert not bad, "\n".join(bad)


def test_deep_ensemble_mixed_within_bound():
    """The same members in the mixed form (one launch per stage) against the
    FP64 ensemble after 100 steps."""
    from hackathon_mom6_miniapp_amd.model import Ensemble

    cases = [make_case(c, init="seeded", seed=600 + q) for q, c in enumerate(DEEP_MEMBERS)]
    print("[gpu deep ensemble] mixed vs fp64, 100 steps", flush=True)
    ref = Ensemble(DEEP_MEMBERS, cases)
    mix = Ensemble([replace(c, precision="mixed") for c in DEEP_MEMBERS], cases, graph=True)
    ref.step(100)
    mix.step(100)
    ref.synchronize()
    mix.synchronize()
    worst = {}
    for q in range(len(DEEP_MEMBERS)):
        for f in ("h", "u", "v"):
            a, b = mix.members[q].get(f), ref.members[q].get(f)
            worst[(q, f)] = np.abs(a - b).max() / np.abs(b).max()
    print(f"[gpu deep ensemble] worst {max(worst.values()):.2e}", flush=True)
    ref.close()
    mix.close()
    assert all(e <= 1e-5 for e in worst.values()), worst
