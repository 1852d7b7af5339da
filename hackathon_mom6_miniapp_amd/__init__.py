"""MI355X-native split-RK2 dynamical core of the MOM6 mini-app.

The compute path is libmom6hip.so (hand-written HIP for gfx950, C ABI in
include/mom6hip.h).  This package holds the host side: run configuration and
initial state (state.py), the ctypes binding (_lib.py), the run driver
(model.py) and the build step (build.py).
"""
from .state import Config, make_case, BC_STATIC, BC_PERIODIC, BC_WALL  # noqa: F401

__all__ = ["Config", "make_case", "BC_STATIC", "BC_PERIODIC", "BC_WALL"]
