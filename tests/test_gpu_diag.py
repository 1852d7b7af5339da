"""The diagnostics on the device against the oracle's restatement of the
serial driver's (itself checked against the reference driver's printout,
test_cpu_diag.py).

Bar: the per-point fields -- the shared work array of the diffu*h / diffv*h
column sums, the KE field, the mass field -- bit-identical; min and max
identical; mean and rms within 1e-10 of the rms (the serial sum adds the
values one by one, the device sums them as double-double, so the two differ
by the serial sum's rounding, below count * 1.1e-16 relative: at most 4e-11
for the 3.4e5 values here)."""
import os
import re
import subprocess

import numpy as np
import pytest

from hackathon_mom6_miniapp_amd.state import BC_PERIODIC, BC_STATIC, BC_WALL, Config, make_case

pytestmark = pytest.mark.gpu

DIAG_CASES = {
    "static_reference_40": (Config(ni=40, nj=40, nk=10, bt_nsteps=20), "reference", 3),
    "seam_periodic_walls_200x140": (Config(ni=200, nj=140, nk=12, bt_nsteps=20, bc_x=BC_PERIODIC, bc_y=BC_WALL,
                                           coriolis="betaplane", f0=1e-4, beta=2e-11, Kh=1e3), "seeded", 2),
}
TOL = 1e-10


def _check_stats(got, want, name):
    assert got[0] == want[0] and got[1] == want[1], (name, got, want)
    scale = max(abs(want[3]), 1e-300)
    assert abs(got[2] - want[2]) <= TOL * scale, (name, "mean", got[2], want[2])
    assert abs(got[3] - want[3]) <= TOL * scale, (name, "rms", got[3], want[3])


@pytest.mark.parametrize("mode", ["fused", "modules"])
@pytest.mark.parametrize("name", sorted(DIAG_CASES))
def test_model_diagnostics_match_oracle(name, mode):
    from hackathon_mom6_miniapp_amd.model import Model
    from oracle.pyoracle import Oracle

    cfg, init, steps = DIAG_CASES[name]
    print(f"[gpu diag] case={name} mode={mode}", flush=True)
    case = make_case(cfg, init=init, seed=77)
    gpu = Model(cfg, case)
    gpu.set_option("fused", 1.0 if mode == "fused" else 0.0)
    gpu.set_option("diag", 1.0)
    ora = Oracle(cfg, case, diag=True)
    bad = []
    for it in range(steps):
        gpu.step(1)
        ora.step(1)
        g, o = gpu.diagnostics(), ora.diagnostics()
        for key in ("diffu_sum", "diffv_sum", "KE", "mass"):
            _check_stats(g[key], o[key], f"step {it + 1} {key}")
        isd, ied, jsd, jed, isc, iec, jsc, jec = cfg.bounds
        for f in ("diag_work2d", "diag_mass"):
            if not np.array_equal(gpu.get(f), ora.field(f)):
                bad.append(f"step {it + 1} {f}")
        gk, ok = gpu.get("diag_ke"), ora.field("diag_ke")
        if not np.array_equal(gk[isc - 1:iec, jsc - 1:jec], ok[isc - 1:iec, jsc - 1:jec]):
            bad.append(f"step {it + 1} diag_ke")
    gpu.close()
    assert not bad, bad


def test_fortran_driver_diag_lines_match_oracle():
    """`rk2_hip_driver 40 40 10 3 20 1` prints the reference driver's Diag
    lines (diffu_sum, diffv_sum each iteration; KE, mass at the end) with the
    oracle's values."""
    from hackathon_mom6_miniapp_amd import diag
    from hackathon_mom6_miniapp_amd.build import DRIVER
    from oracle.pyoracle import Oracle

    print("[gpu diag] rk2_hip_driver 40 40 10 3 20 1", flush=True)
    assert os.path.exists(DRIVER), f"{DRIVER} is not built"
    res = subprocess.run([DRIVER, "40", "40", "10", "3", "20", "1"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
    assert "Diagnostics: ENABLED" in res.stdout
    got = [line for line in res.stdout.splitlines() if "Diag:" in line]
    cfg = Config(ni=40, nj=40, nk=10, bt_nsteps=20)
    o = Oracle(cfg, make_case(cfg, init="reference"), diag=True)
    want = []
    for _ in range(3):
        o.step(1)
        d = o.diagnostics()
        want += [diag.stats_line(n, d[n]) for n in ("diffu_sum", "diffv_sum")]
    want += [diag.stats_line(n, d[n]) for n in ("KE", "mass")]
    assert got == want


def test_diagnostics_need_the_option_and_the_fp64_form():
    from dataclasses import replace

    from hackathon_mom6_miniapp_amd._lib import Mom6hError
    from hackathon_mom6_miniapp_amd.model import Model

    cfg = DIAG_CASES["seam_periodic_walls_200x140"][0]
    case = make_case(cfg, init="seeded", seed=1)
    m = Model(cfg, case)
    with pytest.raises(Mom6hError, match="diag"):
        m.diagnostics()
    m.close()
    mix = Model(replace(cfg, precision="mixed"), case)
    with pytest.raises(Mom6hError, match="FP64"):
        mix.set_option("diag", 1.0)
    mix.close()
