"""rad_process_runs_static (the entry point that allocates nothing) against
rad_process_runs on the same item buffers: the triplet stream and its write
index equal bit for bit, on two calls in a row, the second with a smaller
input, with device memory from the HIP runtime (ctypes)."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class Dev:
    """Device buffers through libamdhip64 (hipMalloc / hipMemcpy / hipFree)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(4, nbytes)) == 0
        self.bufs.append(p)
        return p

    def put(self, arr):
        a = np.ascontiguousarray(arr)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p, n, dtype):
        out = np.zeros(n, dtype)
        if n:
            assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def _static(rad, s, dev, d_item, k, cap):
    f = rad.lib().rad_process_runs_static
    dh, di, de = dev.alloc(4 * cap), dev.alloc(4 * cap), dev.alloc(4 * cap)
    dn = dev.alloc(4)
    rc = f(s._ctx, d_item, k, cap, dh, di, de, dn)
    return rc, dh, di, de, dn


def test_static_runs_equal_allocating_entry(rad):
    sc = rad.Scene.boxes(nboxes=30, room=(8.0, 3.0, 8.0), target_tris=6000, nlights=4, seed=3)
    s = rad.Solver(hemicube=64, hemicubes=4)
    s.upload(sc)
    dev = Dev()
    try:
        for k, shooters in ((4, [0, 17, sc.npatch // 2, sc.npatch - 1]), (2, [5, 99])):
            say(f"no-alloc run form: {k} hemicubes")
            item = s.render(shooters)
            h, i, e = s.process_runs(k)
            d_item = dev.put(item)
            cap = len(h) + 8
            rc, dh, di, de, dn = _static(rad, s, dev, d_item, k, cap)
            assert rc == 0, rad.lib().rad_last_error(s._ctx)
            s.finish()
            n = int(dev.get(dn, 1, np.uint32)[0])
            assert n == len(h)
            assert np.array_equal(dev.get(dh, n, np.uint32), h)
            assert np.array_equal(dev.get(di, n, np.uint32), i)
            assert np.array_equal(dev.get(de, n, np.float32).view(np.uint32), e.view(np.uint32))
        # a short output array: the write index still counts every triplet,
        # and only cap entries are written
        rc, dh, di, de, dn = _static(rad, s, dev, d_item, 2, 4)
        assert rc == 0
        s.finish()
        assert int(dev.get(dn, 1, np.uint32)[0]) == len(h)
        assert np.array_equal(dev.get(di, 4, np.uint32), i[:4])
        # beyond the fixed scratch: refused before any launch
        rc, *_ = _static(rad, s, dev, d_item, 2000, 0)
        assert rc == -5
        assert b"scratch" in rad.lib().rad_last_error(s._ctx)
    finally:
        s.finish()
        dev.free()
