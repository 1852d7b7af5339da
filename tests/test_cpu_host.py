"""Host-side logic and the C ABI library (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

from hackathon_mom6_miniapp_amd.state import (BC_PERIODIC, BC_WALL, Config, Seeded, make_case, masks,
                                              seeded_normal)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "mom6hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mom6h_\w+)\s*\(", txt)))


def test_library_exports_every_declared_function():
    path = os.path.join(ROOT, "hackathon_mom6_miniapp_amd", "libmom6hip.so")
    if not os.path.exists(path):
        pytest.skip("libmom6hip.so not built")
    lib = ctypes.CDLL(path)
    names = _header_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_library_reports_errors_without_a_gpu_call():
    from hackathon_mom6_miniapp_amd._lib import lib
    L = lib()
    assert L.mom6h_version().startswith(b"mom6hip")
    # NULL control structure: rejected on the host, before any HIP call
    rc = L.mom6h_model_step(None, 1)
    assert rc == 1001 and b"NULL" in L.mom6h_last_error()


def test_bounds_follow_the_reference_grid():
    cfg = Config(ni=720, nj=720, nk=75)
    assert cfg.bounds == (1, 722, 1, 722, 4, 719, 4, 719)
    assert cfg.cells == 720 * 720 * 75


def test_step_bytes_match_the_algorithmic_figure():
    import bench
    cfg = bench.workload_config()
    assert abs(bench.step_bytes(cfg) / 1e9 - 23.28) < 0.01


def test_seeded_normal_statistics_and_determinism():
    a = seeded_normal(5, 1, 200000, 0.05)
    assert np.array_equal(a, seeded_normal(5, 1, 200000, 0.05))
    assert abs(a.mean()) < 1e-3 and abs(a.std() - 0.05) < 1e-3
    assert not np.array_equal(a[:1000], seeded_normal(5, 2, 1000, 0.05))
    assert np.abs(a).max() <= 6 * 0.05


def test_lazy_seeded_fields_equal_the_host_fields():
    cfg = Config(ni=20, nj=16, nk=4, bc_x=BC_PERIODIC, bc_y=BC_WALL, coriolis="betaplane", beta=2e-11)
    lazy = make_case(cfg, seed=9, lazy3d=True)
    host = make_case(cfg, seed=9)
    for k in ("u", "v", "h0"):
        assert isinstance(lazy[k], Seeded)
        assert np.array_equal(lazy[k].host(cfg, lazy), host[k])


def test_wall_masks_close_the_domain():
    cfg = Config(ni=20, nj=16, nk=2, bc_x=BC_PERIODIC, bc_y=BC_WALL)
    mT, mCu, mCv, mBu = masks(cfg)
    isd, ied, jsd, jed, isc, iec, jsc, jec = cfg.bounds
    assert mT[:, jsc - 1:jec].all() and not mT[:, :jsc - 1].any() and not mT[:, jec:].any()
    assert not mCv[:, jec - 1].any() and not mCv[:, jsc - 2].any() and mCv[:, jsc - 1:jec - 1].all()
    case = make_case(cfg, seed=1)
    assert not case["v"][:, jec - 1, :].any() and not case["u"][:, :jsc - 1, :].any()


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


@pytest.mark.parametrize("dt,nsteps,status", [(300.0, 50, 0), (300.0, 4, 2), (300.0, 10, 1)])
def test_bt_cfl_matches_closed_form(dt, nsteps, status):
    """mom6h_bt_cfl against the closed form of check_bt_cfl
    (app/rk2_serial_driver.F90:625-662): c = sqrt(9.8*4000), CFL = c*dt/n/dx,
    CFL_2d = sqrt(2)*CFL, recommended substeps ceil(2.5*dt*c/dx)."""
    import math

    from hackathon_mom6_miniapp_amd._lib import lib
    L = lib()
    cfl, cfl2, nmin, st = ctypes.c_double(), ctypes.c_double(), ctypes.c_int(), ctypes.c_int()
    assert L.mom6h_bt_cfl(ctypes.c_double(dt), nsteps, ctypes.c_double(10.0), ctypes.c_double(4000.0),
                          ctypes.byref(cfl), ctypes.byref(cfl2), ctypes.byref(nmin), ctypes.byref(st)) == 0
    c = math.sqrt(9.8 * 4000.0)
    assert cfl.value == pytest.approx(c * dt / nsteps / 1.0e4, rel=1e-15)
    assert cfl2.value == pytest.approx(cfl.value * math.sqrt(2.0), rel=1e-15)
    assert nmin.value == math.ceil(dt * c / 1.0e4 * 2.5) == 15
    assert st.value == status


def test_traffic_summary_holds_every_roofline_kernel():
    """bench.py takes each roofline kernel's HBM bytes per launch from the
    committed counter summary; every one of them is there and plausible
    (between the kernel's algorithmic bytes / 2 and 4x them)."""
    import bench

    n2 = 723 * 723   # the 720x720 grid with its halos
    for stage, kernel in bench.ROOFLINE_KERNELS.items():
        t = bench.measured_traffic(kernel)
        assert t is not None, kernel
        a3, a2 = bench.STAGE_ARRAYS[stage]
        alg = 8.0 * (a3 * n2 * 75 + a2 * n2)
        assert 0.5 * alg < t < 4.0 * alg, (kernel, t, alg)
