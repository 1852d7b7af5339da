"""The Fortran host side (iso_c_binding modules + rk2_hip_driver) against the
oracle: the reference driver's own configuration (static halos, analytic
state, Kh = 100) run through the Fortran module API, bit-exact on u, v, h and
eta, and its verify_state printout equal to the oracle's; and the driver's
options that reach the module switches the reference driver leaves off (the
hor_visc closures through hor_visc_set_closures_hip, Rayleigh drag at every
level through vert_visc_cra_ray3d_hip)."""
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from hackathon_mom6_miniapp_amd.build import DRIVER
from hackathon_mom6_miniapp_amd.state import Config, make_case

pytestmark = pytest.mark.gpu


OPTION_CASES = {
    "": {},
    "smagorinsky,bound": dict(smagorinsky=True, better_bound=True),
    "leith,ray3d": dict(leith=True, ray_profile="linear"),
}


@pytest.mark.parametrize("options", sorted(OPTION_CASES))
@pytest.mark.parametrize("ni,nk,niter,bt", [(24, 6, 2, 10), (40, 12, 3, 20)])
def test_fortran_driver_bit_exact(ni, nk, niter, bt, options):
    from oracle.pyoracle import Oracle

    print(f"[gpu fortran] rk2_hip_driver {ni} {ni} {nk} {niter} {bt} options={options!r}", flush=True)
    assert os.path.exists(DRIVER), f"{DRIVER} is not built (hackathon_mom6_miniapp_amd/build.py)"
    cfg = Config(ni=ni, nj=ni, nk=nk, bt_nsteps=bt, **OPTION_CASES[options])
    case = make_case(cfg, init="reference")
    with tempfile.TemporaryDirectory() as tmp:
        fin, fout = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        with open(fin, "wb") as f:
            for name in ("u", "v", "h0", "eta", "ubt", "vbt", "CoriolisBu"):
                f.write(np.asfortranarray(case[name]).tobytes(order="F"))
        args = [DRIVER, str(ni), str(ni), str(nk), str(niter), str(bt), "0", fin, fout] + ([options] if options else [])
        res = subprocess.run(args, capture_output=True, text=True, timeout=120)
        assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
        raw = np.fromfile(fout, dtype=np.float64)
    n3, n2 = cfg.n1 * cfg.n2 * nk, cfg.n1 * cfg.n2
    assert raw.size == 3 * n3 + n2
    got = {"u": raw[:n3], "v": raw[n3:2 * n3], "h": raw[2 * n3:3 * n3], "eta": raw[3 * n3:]}
    ora = Oracle(cfg, case)
    ora.step(niter)
    for name, g in got.items():
        o = ora.field(name).reshape(-1, order="F")
        assert np.array_equal(g, o), f"{name}: max|diff| = {np.abs(g - o).max():.3e}"
    # verify_state printout (ES20.12) equals the oracle's values
    tail = res.stdout.split("Verification:", 1)[-1]
    printed = [float(x) for x in re.findall(r":\s+([-+0-9.E]+)\s*$", tail, flags=re.M)[:7]]
    ref = ora.verify()
    assert len(printed) == 7, res.stdout
    for p, r in zip(printed, ref):
        assert float(f"{r:.12E}") == p
    assert len(re.findall(r"RK2 iteration\s+\d+:", res.stdout)) == niter
