"""The batched ensemble (one launch per stage over members of different sizes)
against the same models stepped one by one: bit-identical per member, in the
FP64 form (itself bit-exact with the oracle) and the mixed-precision form."""
from dataclasses import replace

import numpy as np
import pytest

from hackathon_mom6_miniapp_amd.state import BC_PERIODIC, BC_STATIC, BC_WALL, Config, make_case

pytestmark = pytest.mark.gpu

MEMBERS = [
    Config(ni=40, nj=32, nk=10, bt_nsteps=16, bc_x=BC_PERIODIC, bc_y=BC_WALL, coriolis="betaplane", f0=1e-4,
           beta=2e-11, Kh=1e3),
    Config(ni=52, nj=44, nk=25, bt_nsteps=16, bc_x=BC_PERIODIC, bc_y=BC_PERIODIC, coriolis="fplane", f0=1e-4),
    Config(ni=36, nj=60, nk=12, bt_nsteps=16, bc_x=BC_STATIC, bc_y=BC_STATIC),
]
FIELDS = ("u", "v", "h", "htmp", "up", "vp", "eta", "ubt", "vbt", "ubt_av", "vbt_av", "eta_av", "u_cor", "du_cor",
          "FA_u_EE", "FA_u_W0", "uBT_WW", "BT_h_u")


def _members(precision="fp64", meridional=False, biharmonic=False):
    cfgs = [replace(c, precision=precision, meridional=meridional, biharmonic=biharmonic, Ah=1e11)
            for c in MEMBERS]
    cases = [make_case(c, init="seeded", seed=40 + q) for q, c in enumerate(cfgs)]
    return cfgs, cases


def _check_against_single(cfgs, cases, steps, fields, graph=False, options=None):
    from hackathon_mom6_miniapp_amd.model import Ensemble, Model

    ens = Ensemble(cfgs, cases, graph=graph)
    for k, v in (options or {}).items():
        ens.set_option(k, v)
    ens.step(steps)
    ens.synchronize()
    bad = []
    for q, (cfg, case) in enumerate(zip(cfgs, cases)):
        single = Model(cfg, case)
        single.step(steps)
        single.synchronize()
        for f in fields:
            if not np.array_equal(ens.members[q].get(f), single.get(f)):
                bad.append(f"member {q} ({cfg.ni}x{cfg.nj}x{cfg.nk}): {f}")
        single.close()
    ens.close()
    assert not bad, "\n".join(bad)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ensemble_matches_single_runs(precision):
    print(f"[gpu ensemble] precision={precision}", flush=True)
    cfgs, cases = _members(precision)
    _check_against_single(cfgs, cases, 3, FIELDS)


@pytest.mark.parametrize("per_launch", [1, 2, 3, 5])
def test_ensemble_barotropic_substeps_per_launch(per_launch):
    """The ensemble's barotropic chain with 1, 2 or 4 substeps per launch (the
    single models run the default) gives the same members."""
    print(f"[gpu ensemble] bt_substeps={per_launch}", flush=True)
    cfgs, cases = _members("fp64")
    _check_against_single(cfgs, cases, 2, ("eta", "ubt", "vbt", "ubt_av", "vbt_av", "eta_av", "u", "h"),
                          options={"bt_substeps": per_launch})


def test_ensemble_meridional_biharmonic_mixed():
    print("[gpu ensemble] mixed + meridional + biharmonic", flush=True)
    cfgs, cases = _members("mixed", meridional=True, biharmonic=True)
    _check_against_single(cfgs, cases, 2, FIELDS + ("v_cor", "FA_v_NN", "BT_h_v"))


def test_ensemble_graph_and_autotune():
    """A captured ensemble step replays to the eager result; the autotuner
    leaves the state alone and its choice gives the same results."""
    from hackathon_mom6_miniapp_amd.model import Ensemble

    print("[gpu ensemble] graph + autotune", flush=True)
    cfgs, cases = _members("mixed")
    eager = Ensemble(cfgs, cases)
    graph = Ensemble(cfgs, cases, graph=True)
    eager.step(1)
    graph.step(1)
    eager.synchronize()
    graph.synchronize()
    before = [m.get("u") for m in graph.members]
    report = graph.autotune(3)
    print(f"[gpu ensemble] {report}", flush=True)
    assert "chosen" in report
    for q, m in enumerate(graph.members):
        assert np.array_equal(m.get("u"), before[q])
    for rep, n in enumerate((1, 3)):
        eager.step(n)
        graph.step(n)
        eager.synchronize()
        graph.synchronize()
        for q in range(len(cfgs)):
            for f in ("u", "v", "h", "eta"):
                assert np.array_equal(eager.members[q].get(f), graph.members[q].get(f)), (rep, q, f)
    eager.close()
    graph.close()


def test_model_autotune_keeps_state():
    from hackathon_mom6_miniapp_amd.model import Model

    cfg = MEMBERS[0]
    case = make_case(cfg, init="seeded", seed=3)
    a, b = Model(cfg, case), Model(cfg, case)
    print(f"[gpu autotune] {b.autotune(3)}", flush=True)
    a.step(2)
    b.step(2)
    a.synchronize()
    b.synchronize()
    for f in ("u", "v", "h"):
        assert np.array_equal(a.get(f), b.get(f)), f
    a.close()
    b.close()
