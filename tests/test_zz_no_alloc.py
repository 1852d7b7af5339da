"""mom6h_btstep_noalloc (the barotropic step with its scratch in a static
device array, no allocation) against mom6h_btstep (the control structure's
scratch): bit-identical outputs over consecutive calls on one GPU, the second
on a smaller grid (the shared array then holds the first call's larger
contents), the third on the first grid again; and the size limit is refused
before any device work."""
import ctypes as C

import numpy as np
import pytest

from hackathon_mom6_miniapp_amd.state import seeded_normal

pytestmark = pytest.mark.gpu


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class _Grid:
    """A barotropic control structure and device inputs/outputs on one grid."""

    def __init__(self, L, ni, nj, nstep, px, py, seed, uniform):
        from hackathon_mom6_miniapp_amd._lib import check
        self.L, self.check = L, check
        n1, n2 = ni + 2, nj + 2
        self.n = n1 * n2
        shp = (n1, n2)
        dx = 1.0e4
        Datu = np.full(shp, 4000.0 * dx, order="F")
        Datv = Datu.copy(order="F")
        if not uniform:     # coefficients varying in i and j: the general kernel
            Datu = Datu * (1.0 + 0.01 * seeded_normal(seed, 7, self.n, 1.0).reshape(shp, order="F"))
        j = np.arange(1, n2 + 1, dtype=np.float64)
        f = 1.0e-4 + 2.0e-11 * ((j - 1 - nj // 2) - 0.5) * dx
        f4 = np.asfortranarray(np.broadcast_to((0.25 * f)[None, None, :], (4, n1, n2)).copy())
        g = np.full(shp, 9.8, order="F")
        rem = np.full(shp, 0.999, order="F")
        Ia = np.full(shp, 1.0 / (dx * dx), order="F")
        Id = np.full(shp, 1.0 / dx, order="F")
        self.keep = [Datu, Datv, g, f4, rem, Ia, Id]
        self.cs = C.c_void_p()
        check(L.mom6h_barotropic_init(C.byref(self.cs), 1, n1, 1, n2, 4, ni - 1, 4, nj - 1, nstep, 300.0, 0.2, 0,
                                      _dptr(Datu), _dptr(Datv), _dptr(g), _dptr(g), _dptr(g), _dptr(g),
                                      _dptr(f4), _dptr(f4), _dptr(rem), _dptr(rem), _dptr(Ia), _dptr(Id),
                                      _dptr(Id), px, py), "mom6h_barotropic_init")
        self.inputs = [np.asfortranarray(seeded_normal(seed, q, self.n, s).reshape(shp, order="F"))
                       for q, s in ((0, 0.1), (1, 0.01), (2, 0.01))]
        self.d_in = [self._dev(a) for a in self.inputs]
        self.d_out = [self._dev(np.zeros(shp)) for _ in range(3)]

    def _dev(self, a):
        p = C.c_void_p()
        self.check(self.L.mom6h_malloc(C.byref(p), a.nbytes), "malloc")
        self.check(self.L.mom6h_memcpy_h2d(p, np.asfortranarray(a).ctypes.data, a.nbytes), "h2d")
        return p

    def run(self, entry):
        for p in self.d_out:     # outputs cleared between the two entry points
            self.check(self.L.mom6h_memset(p, 0, 8 * self.n), "memset")
        self.check(getattr(self.L, entry)(self.cs, *self.d_in, *self.d_out, None), entry)
        self.check(self.L.mom6h_device_synchronize(), "synchronize")
        out = []
        for p in self.d_out:
            a = np.empty(self.n)
            self.check(self.L.mom6h_memcpy_d2h(a.ctypes.data, p, 8 * self.n), "d2h")
            out.append(a)
        return out

    def close(self):
        for p in self.d_in + self.d_out:
            self.L.mom6h_free(p)
        self.L.mom6h_barotropic_end(self.cs)


def test_noalloc_btstep_matches_btstep_over_consecutive_calls():
    from hackathon_mom6_miniapp_amd._lib import lib
    L = lib()
    big = _Grid(L, 200, 150, 20, 1, 0, seed=3, uniform=True)       # blocked substeps, per-row coefficients
    small = _Grid(L, 90, 60, 12, 0, 0, seed=4, uniform=False)      # the general substep kernel
    for name, g in (("first call, 200x150", big), ("second call, 90x60", small), ("third call, 200x150", big)):
        print(f"[gpu noalloc] {name}", flush=True)
        ref = g.run("mom6h_btstep")
        got = g.run("mom6h_btstep_noalloc")
        for q, (a, b) in enumerate(zip(ref, got)):
            assert np.array_equal(a, b), (name, ("u_av", "v_av", "eta_av")[q])
        assert np.abs(ref[2]).max() > 0.0
    big.close()
    small.close()


def test_noalloc_btstep_refuses_grids_beyond_its_scratch():
    from hackathon_mom6_miniapp_amd._lib import lib
    L = lib()
    g = _Grid(L, 1100, 1000, 4, 0, 0, seed=5, uniform=True)   # 1102 x 1002 > 1,048,576 points
    rc = L.mom6h_btstep_noalloc(g.cs, *g.d_in, *g.d_out, None)
    assert rc == 1001 and b"at most" in L.mom6h_last_error()
    g.close()
