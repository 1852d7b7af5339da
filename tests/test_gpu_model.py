"""End-to-end parity of the HIP split-RK2 step against the oracle.

Bar: bit-exact (np.array_equal) on every array the reference driver holds, for
static, periodic and walled grids, both limiters and all Coriolis schemes."""
import numpy as np
import pytest

from hackathon_mom6_miniapp_amd.state import make_case
from tests.cases import CASES, COMPARE

pytestmark = pytest.mark.gpu


def _run_pair(name):
    from hackathon_mom6_miniapp_amd.model import Model
    from oracle.pyoracle import Oracle

    cfg, init, niter = CASES[name]
    case = make_case(cfg, init=init, seed=11)
    gpu = Model(cfg, case)
    gpu.step(niter)
    gpu.synchronize()
    ora = Oracle(cfg, case)
    ora.step(niter)
    return gpu, ora


@pytest.mark.parametrize("name", sorted(CASES))
def test_model_bit_exact(name):
    print(f"[gpu parity] case={name}", flush=True)
    gpu, ora = _run_pair(name)
    bad = []
    for f in COMPARE:
        g, o = gpu.get(f), ora.field(f)
        assert g.shape == o.shape, f
        if not np.array_equal(g, o):
            d = np.abs(g - o)
            bad.append(f"{f}: max|diff|={d.max():.3e} at {np.unravel_index(d.argmax(), d.shape)} n={(d > 0).sum()}")
    gpu.close()
    assert not bad, "\n".join(bad)
