"""edmd_pcf_noalloc (g(r) with device inputs and outputs and the pair counts
in the library's fixed per-GPU array, no allocation) against edmd_pcf (the
allocating entry point, host arrays), through the C ABI: bit for bit, on two
calls in a row, the second with fewer positions and fewer bins, so that
counts left in the shared array by the first call would show.  The test's
own device buffers come from the HIP runtime the library is linked with."""
import ctypes as C

import numpy as np
import pytest

import edmd_fixtures as F
from conftest import bits

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip(gpu):
    gpu.lib()  # the library's own runtime is loaded first
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class DevBuf:
    def __init__(self, h, nbytes):
        self.h = h
        self.p = C.c_void_p()
        assert h.hipMalloc(C.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.nbytes <= self.nbytes
        assert self.h.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, n):
        out = np.empty(n)
        assert self.h.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.p:
            self.h.hipFree(self.p)
            self.p = C.c_void_p()


def _noalloc(gpu, h, x, y, Lx, Ly, dr, max_r, cap):
    bufs = [DevBuf(h, 8 * len(x)), DevBuf(h, 8 * len(y)), DevBuf(h, 8 * cap), DevBuf(h, 8 * cap)]
    try:
        bx, by, br, bg = bufs
        bx.put(x)
        by.put(y)
        br.put(np.full(cap, -1.0))
        bg.put(np.full(cap, -1.0))
        nb = C.c_int(cap)
        rc = gpu.lib().edmd_pcf_noalloc(len(x), bx.p, by.p, float(Lx), float(Ly), float(dr),
                                        float(max_r), br.p, bg.p, C.byref(nb))
        gpu.check(rc)
        r, g = br.get(cap), bg.get(cap)
        return r[:nb.value], g[:nb.value], np.concatenate([r[nb.value:], g[nb.value:]])
    finally:
        for b in bufs:
            b.free()


def test_pcf_noalloc_matches_pcf_twice(gpu):
    h = _hip(gpu)
    big = F.lattice_state(6000, 0.70, seed=41, lattice="fit")
    small = F.lattice_state(700, 0.45, seed=42)
    cases = [(big, 0.05, big["Lx"] / 2), (small, 0.1, small["Lx"] / 3)]
    cap = 4096
    for st, dr, max_r in cases:
        r0, g0 = gpu.pcf(st["x"], st["y"], st["Lx"], st["Ly"], dr, max_r)
        r1, g1, rest = _noalloc(gpu, h, st["x"], st["y"], st["Lx"], st["Ly"], dr, max_r, cap)
        assert len(r1) == len(r0) == int(max_r / dr)
        np.testing.assert_array_equal(bits(r1), bits(r0))
        np.testing.assert_array_equal(bits(g1), bits(g0))
        assert np.all(rest == -1.0)  # nothing written past the bins
        assert g0.max() > 0


def test_pcf_noalloc_rejects_bad_sizes(gpu):
    h = _hip(gpu)
    st = F.lattice_state(100, 0.5, seed=1)
    with pytest.raises(gpu.EdmdError) as e:
        _noalloc(gpu, h, st["x"], st["y"], st["Lx"], st["Ly"], 1e-4, 10.0, 4)
    assert e.value.code == 22 and "32768" in str(e.value)
    # output arrays shorter than the bin count: an error, nothing written
    with pytest.raises(gpu.EdmdError) as e:
        _noalloc(gpu, h, st["x"], st["y"], st["Lx"], st["Ly"], 0.1, 5.0, 4)
    assert e.value.code == 22
