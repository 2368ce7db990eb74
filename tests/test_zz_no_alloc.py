"""rc_energy_noalloc (the energy with its scratch in a static device array,
no allocation in the call) against rc_energy, through the C ABI: two calls in
a row, the second on a smaller curve (the shared scratch keeps the first
call's contents), both bit-identical to rc_energy on the same input; a set
beyond the scratch bound is refused; an endpoint outside [0, V) gives a NaN
energy without reading out of bounds. The caller's device buffers come from
the HIP runtime directly (hipMalloc / hipMemcpy), as such a caller holds them."""
import ctypes as C
import sys

import numpy as np
import pytest

from rcurves import _lib as L
from rcurves import core, curves

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class Hip:
    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        self.bufs.append(p)
        return p

    def h2d(self, dst, arr):
        arr = np.ascontiguousarray(arr)
        assert self.h.hipMemcpy(dst, arr.ctypes.data_as(C.c_void_p), C.c_size_t(arr.nbytes), 1) == 0

    def d2h(self, arr, src):
        assert self.h.hipMemcpy(arr.ctypes.data_as(C.c_void_p), src, C.c_size_t(arr.nbytes), 2) == 0

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def noalloc_energy(hip, v, e, params):
    V, E = len(v), len(e)
    soa = np.ascontiguousarray(np.asarray(v, dtype=np.float64).T).reshape(-1)
    d_x = hip.alloc(soa.nbytes)
    d_e = hip.alloc(4 * 2 * E)
    d_o = hip.alloc(8)
    hip.h2d(d_x, soa)
    hip.h2d(d_e, np.ascontiguousarray(np.asarray(e, dtype=np.int32)))
    st = L.lib().rc_energy_noalloc(d_x, V, d_e, E, C.byref(params), d_o, None)
    out = np.zeros(1)
    if st == 0:
        hip.sync()
        hip.d2h(out, d_o)
    return st, float(out[0])


def test_noalloc_energy_matches_rc_energy_on_two_calls():
    hip = Hip()
    try:
        for (a, b) in ((3.0, 6.0), (2.0, 4.5)):
            p = L.Params(a, b, 1e-10)
            v1, e1 = curves.torus_knot(3, 5, 1024, jitter=0.02, seed=1)
            v2, e2 = curves.trefoil_jittered(200, 0.01, seed=3)
            for v, e in ((v1, e1), (v2, e2)):
                st, en = noalloc_energy(hip, v, e, p)
                ref = core.calculate_energy(v, e, None, a, b, 1e-10)
                log(len(v), a, b, st, en, ref)
                assert st == L.RC_OK
                assert en == ref
    finally:
        hip.free()


def test_noalloc_energy_limits():
    hip = Hip()
    try:
        p = L.Params(3.0, 6.0, 1e-10)
        v, e = curves.fourier_curve(16448, seed=1)      # E = 16448 > 16384
        st, _ = noalloc_energy(hip, v, e, p)
        assert st == L.RC_UNSUPPORTED
        v, e = curves.trefoil_jittered(64, 0.01, seed=1)
        e = e.copy()
        e[5, 1] = 64 + 7                                 # endpoint outside [0, V)
        st, en = noalloc_energy(hip, v, e, p)
        assert st == L.RC_OK and np.isnan(en)
        # the next valid call is unaffected by the previous contents
        v, e = curves.trefoil_jittered(64, 0.01, seed=1)
        st, en = noalloc_energy(hip, v, e, p)
        assert st == L.RC_OK and en == core.calculate_energy(v, e, None, 3.0, 6.0, 1e-10)
    finally:
        hip.free()
