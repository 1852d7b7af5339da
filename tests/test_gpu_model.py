"""End-to-end parity of the HIP split-RK2 step against the oracle.

Bar: bit-exact (np.array_equal) on every array the reference driver holds, for
static, periodic and walled grids, both limiters and all Coriolis schemes."""
import numpy as np
import pytest

from hackathon_mom6_miniapp_amd.state import make_case
from tests.cases import CASES, CLOSURE_CASES, COMPARE, COMPARE_FUSED, MERID_CASES, MERID_COMPARE

pytestmark = pytest.mark.gpu


def _run_pair(name, fused=True, general=False):
    from hackathon_mom6_miniapp_amd.model import Model
    from oracle.pyoracle import Oracle

    cfg, init, niter = {**CASES, **MERID_CASES, **CLOSURE_CASES}[name]
    case = make_case(cfg, init=init, seed=11)
    gpu = Model(cfg, case)
    gpu.set_option("fused", 1.0 if fused else 0.0)
    gpu.set_option("general", 1.0 if general else 0.0)
    gpu.step(niter)
    gpu.synchronize()
    ora = Oracle(cfg, case)
    ora.step(niter)
    return gpu, ora


@pytest.mark.parametrize("mode", ["modules", "general", "fused"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_model_bit_exact(name, mode):
    """modules: the module entry points (on these uniform grids hor_visc and
    CorAdCalc take their tiled kernels); general: the same with the general
    hor_visc and CorAdCalc kernels; fused: the fused momentum stencil."""
    print(f"[gpu parity] case={name} mode={mode}", flush=True)
    gpu, ora = _run_pair(name, fused=(mode == "fused"), general=(mode == "general"))
    bad = []
    fields = COMPARE if mode != "fused" else COMPARE_FUSED + ("uh",)
    cfg = CASES[name][0]
    isd, ied, jsd, jed, isc, iec, jsc, jec = cfg.bounds
    for f in fields:
        g, o = gpu.get(f), ora.field(f)
        if mode == "fused" and f == "uh":
            g, o = g[isc - 2:iec, jsc - 1:jec], o[isc - 2:iec, jsc - 1:jec]
        assert g.shape == o.shape, f
        if not np.array_equal(g, o):
            d = np.abs(g - o)
            bad.append(f"{f}: max|diff|={d.max():.3e} at {np.unravel_index(d.argmax(), d.shape)} n={(d > 0).sum()}")
    gpu.close()
    assert not bad, "\n".join(bad)


@pytest.mark.parametrize("name", sorted(CASES))
def test_vert_visc_coefficients_bit_exact(name):
    """The module path keeps vert_visc's coupling coefficients a and thickness
    h_u (the u columns of the last call); masked columns (the walls) take a = 0
    and h_u = h_neglect without a solve of their own."""
    print(f"[gpu vv coefficients] case={name}", flush=True)
    gpu, ora = _run_pair(name, fused=False)
    for f in ("vv_a_u", "vv_h_u"):
        g, o = gpu.get(f), ora.field(f)
        assert g.shape == o.shape, (f, g.shape, o.shape)
        assert np.array_equal(g, o), f"{f}: max|diff| = {np.abs(g - o).max():.3e}"
    gpu.close()


def test_fused_stencil_staging_checks():
    """The fused stencil stages the masks as bytes and Coriolis as one value
    per row, which the model checks on the host.  On the beta-plane case the
    fused stencil runs (its stage timer counts, the module timers do not);
    with a CoriolisBu that varies along i the model keeps to the module
    kernels, bit-exact with the oracle."""
    from hackathon_mom6_miniapp_amd.model import Model
    from oracle.pyoracle import Oracle

    cfg, init, niter = CASES["zonal_periodic_walls_beta"]
    for vary in (False, True):
        print(f"[gpu fused staging] Coriolis varying along i: {vary}", flush=True)
        case = make_case(cfg, init=init, seed=11)
        if vary:
            f = np.array(case["CoriolisBu"], order="F")
            f += 1.0e-6 * np.arange(f.shape[0], dtype=np.float64)[:, None]
            case["CoriolisBu"] = f
        gpu = Model(cfg, case)
        gpu.set_option("fused", 1.0)
        gpu.set_option("profile", 1.0)
        gpu.step(niter)
        gpu.synchronize()
        st = gpu.stage_times()
        if vary:
            assert st["momentum"] == 0.0 and st["coriolis"] > 0.0, st
        else:
            assert st["momentum"] > 0.0 and st["coriolis"] == 0.0, st
        ora = Oracle(cfg, case)
        ora.step(niter)
        for name in ("u", "v", "h", "eta", "up", "vp"):
            assert np.array_equal(gpu.get(name), ora.field(name)), (vary, name)
        gpu.close()


def test_seeded_state_generated_on_device_matches_host():
    """The device generator (mom6h_seeded_normal) reproduces the host recipe bit for bit."""
    from hackathon_mom6_miniapp_amd.model import Model
    from hackathon_mom6_miniapp_amd.state import BC_PERIODIC, BC_WALL, Config

    cfg = Config(ni=30, nj=26, nk=7, bt_nsteps=10, bc_x=BC_PERIODIC, bc_y=BC_WALL,
                 coriolis="betaplane", f0=1e-4, beta=2e-11)
    lazy = make_case(cfg, init="seeded", seed=99, lazy3d=True)
    host = make_case(cfg, init="seeded", seed=99)
    m = Model(cfg, lazy)
    isd, ied, jsd, jed, isc, iec, jsc, jec = cfg.bounds
    for f in ("u", "v", "h0"):
        # prepare() refilled the periodic x halo: compare the compute columns
        assert np.array_equal(m.get(f)[isc - 1:iec], host[f][isc - 1:iec]), f
    m.close()


def test_graph_replay_matches_eager():
    """A step captured as a hipGraph and replayed gives the eager result after
    every call, with eager work and synchronisations in between."""
    from hackathon_mom6_miniapp_amd.model import Model

    cfg, init, _ = CASES["zonal_periodic_walls_beta"]
    case = make_case(cfg, init=init, seed=5)
    eager = Model(cfg, case)
    graph = Model(cfg, case, graph=True)
    # calls of 1, 3 and 2 steps: a call of n steps replays n barotropic graphs
    # on the second stream beside n 3D graphs on the first, joined once
    for rep, n in enumerate((1, 3, 2)):
        print(f"[gpu graph] replay {rep}: {n} steps", flush=True)
        eager.step(n)
        eager.synchronize()
        graph.step(n)
        graph.synchronize()
        for f in ("u", "v", "h", "eta", "ubt_av", "vbt_av", "uh", "up"):
            assert np.array_equal(eager.get(f), graph.get(f)), (rep, f)
    eager.close()
    graph.close()


def test_serialized_streams_match_overlapped():
    """The barotropic chain run on the 3D stream (the per-stage timing mode)
    gives the same state as the two overlapping streams."""
    from hackathon_mom6_miniapp_amd.model import Model

    cfg, init, _ = CASES["zonal_periodic_walls_beta"]
    case = make_case(cfg, init=init, seed=6)
    conc = Model(cfg, case)
    ser = Model(cfg, case)
    ser.set_option("serialize", 1.0)
    ser.set_option("profile", 1.0)
    print("[gpu serialize] 3 steps", flush=True)
    conc.step(3)
    conc.synchronize()
    ser.step(3)
    ser.synchronize()
    assert set(ser.stage_times()) >= {"momentum", "vert_visc", "continuity", "btstep"}
    for f in ("u", "v", "h", "eta", "ubt", "vbt", "ubt_av", "vbt_av", "uh"):
        assert np.array_equal(conc.get(f), ser.get(f)), f
    conc.close()
    ser.close()


@pytest.mark.parametrize("mode", ["modules", "fused"])
@pytest.mark.parametrize("name", sorted(MERID_CASES))
def test_meridional_pass_bit_exact(name, mode):
    """Continuity with the meridional pass after the zonal one, against the
    oracle (whose meridional pass is its zonal pass on the transposed grid)."""
    print(f"[gpu parity] case={name} mode={mode}", flush=True)
    gpu, ora = _run_pair(name, fused=(mode == "fused"))
    cfg = MERID_CASES[name][0]
    isd, ied, jsd, jed, isc, iec, jsc, jec = cfg.bounds
    fields = (COMPARE if mode == "modules" else COMPARE_FUSED) + MERID_COMPARE + ("vh",)
    bad = []
    for f in fields:
        g, o = gpu.get(f), ora.field(f)
        if mode == "fused" and f == "vh":   # the fused stencil keeps its transports on chip
            g, o = g[isc - 1:iec, jsc - 2:jec], o[isc - 1:iec, jsc - 2:jec]
        assert g.shape == o.shape, f
        if not np.array_equal(g, o):
            d = np.abs(g - o)
            bad.append(f"{f}: max|diff|={d.max():.3e} at {np.unravel_index(d.argmax(), d.shape)} n={(d > 0).sum()}")
    gpu.close()
    assert not bad, "\n".join(bad)


@pytest.mark.parametrize("shape", [(4, 1, 8), (16, 2, 16), (8, 4, 4), (8, 3, 4)])
def test_launch_shapes_give_identical_results(shape):
    """Every momentum tile / layer split / barotropic tile choice computes the
    same values (the autotuner only picks among equals)."""
    from hackathon_mom6_miniapp_amd.model import Model

    ty, nkz, bt = shape
    print(f"[gpu shapes] momentum {ty}x{nkz}, bt tile {bt}", flush=True)
    cfg, init, _ = CASES["zonal_periodic_walls_beta"]
    case = make_case(cfg, init=init, seed=12)
    ref = Model(cfg, case)
    alt = Model(cfg, case)
    alt.set_option("momentum_tile", ty)
    alt.set_option("momentum_splits", nkz)
    alt.set_option("bt_tile", bt)
    ref.step(2)
    alt.step(2)
    ref.synchronize()
    alt.synchronize()
    for f in ("u", "v", "h", "up", "vp", "eta", "ubt_av", "vbt_av", "eta_av"):
        assert np.array_equal(ref.get(f), alt.get(f)), f
    ref.close()
    alt.close()


@pytest.mark.parametrize("name", ["doubly_periodic_fplane", "zonal_periodic_walls_beta", "static_seeded_ragged"])
@pytest.mark.parametrize("per_launch,rows", [(2, 1), (3, 0), (4, 1), (4, 0), (5, 1)])
def test_blocked_barotropic_substeps_match_one_per_launch(name, per_launch, rows):
    """Several barotropic substeps per launch (the blocked kernel, periodic
    halos as images; coefficients staged per row or as 2D fields) give the
    values of one launch per substep."""
    from hackathon_mom6_miniapp_amd.model import Model

    print(f"[gpu bt blocks] case={name} per_launch={per_launch} rows={rows}", flush=True)
    cfg, init, _ = CASES[name]
    case = make_case(cfg, init=init, seed=21)
    one = Model(cfg, case)
    blk = Model(cfg, case)
    one.set_option("bt_substeps", 1)
    blk.set_option("bt_substeps", per_launch)
    blk.set_option("bt_rows", rows)
    one.step(2)
    blk.step(2)
    one.synchronize()
    blk.synchronize()
    for f in ("eta", "ubt", "vbt", "ubt_av", "vbt_av", "eta_av", "u", "v", "h"):
        assert np.array_equal(one.get(f), blk.get(f)), f
    one.close()
    blk.close()


def _sweep_cfg(nk, meridional):
    from hackathon_mom6_miniapp_amd.state import BC_PERIODIC, BC_WALL, Config

    return Config(ni=24, nj=10, nk=nk, bt_nsteps=6, bc_x=BC_PERIODIC, bc_y=BC_WALL, coriolis="betaplane",
                  f0=1e-4, beta=2e-11, Kh=1e3, meridional=meridional)


@pytest.mark.parametrize("direction", ["zonal", "meridional"])
def test_every_level_count_bit_exact(direction):
    """Continuity's block kernels choose their shape from nk (levels per wave,
    waves per block) and fold the columns in batches with a remainder: every
    nk from 1 to 80 (zonal; the meridional pass on a set covering every shape
    and remainder) matches the oracle bit for bit, the barotropic face
    coefficients included."""
    from hackathon_mom6_miniapp_amd.model import Model
    from oracle.pyoracle import Oracle

    merid = direction == "meridional"
    levels = range(1, 81) if not merid else (1, 2, 3, 4, 5, 6, 7, 9, 11, 17, 31, 33, 38, 67, 75, 80)
    fields = COMPARE + (MERID_COMPARE + ("vh",) if merid else ())
    bad = []
    for nk in levels:
        print(f"[gpu levels] {direction} nk={nk}", flush=True)
        cfg = _sweep_cfg(nk, merid)
        case = make_case(cfg, init="seeded", seed=100 + nk)
        gpu = Model(cfg, case)
        gpu.set_option("fused", 0.0)
        gpu.step(1)
        gpu.synchronize()
        ora = Oracle(cfg, case)
        ora.step(1)
        for f in fields:
            g, o = gpu.get(f), ora.field(f)   # (the oracle returns nk = 1 fields 2D)
            if g.size != o.size or not np.array_equal(g.ravel(order="F"), o.ravel(order="F")):
                bad.append(f"nk={nk} {f}")
        gpu.close()
    assert not bad, ", ".join(bad[:40])


@pytest.mark.parametrize("name", sorted(CLOSURE_CASES))
def test_viscosity_closures_bit_exact(name):
    """Smagorinsky and Leith viscosities, the thickness-aware bounds and
    no-slip walls (the general hor_visc kernel), every array the driver
    holds, against the oracle (pinned by golden vectors of the same
    configurations)."""
    print(f"[gpu closures] case={name}", flush=True)
    gpu, ora = _run_pair(name, fused=False)
    cfg = CLOSURE_CASES[name][0]
    bad = []
    for f in COMPARE:
        g, o = gpu.get(f), ora.field(f)
        if not np.array_equal(g, o):
            d = np.abs(g - o)
            bad.append(f"{f}: max|diff|={d.max():.3e} at {np.unravel_index(d.argmax(), d.shape)} n={(d > 0).sum()}")
    gpu.close()
    assert not bad, "\n".join(bad)
