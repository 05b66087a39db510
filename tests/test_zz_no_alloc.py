"""m3d_gen_cylinders_noalloc (the allocation-free variant of
m3d_gen_cylinders, fixed per-GPU device buffer) against m3d_gen_cylinders,
bit for bit, through the C-ABI: two calls in a row, the second on a smaller
grid with fewer cylinders, so a stale count or cylinder left in the shared
buffer by the first call would show."""
import ctypes as ct

import numpy as np
import pytest

import mat3d_amd as m
from mat3d_amd import gen

pytestmark = pytest.mark.gpu

i32p, dp = ct.POINTER(ct.c_int32), ct.POINTER(ct.c_double)
H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def hip():
    """The HIP runtime libmat3d.so runs on (its NEEDED libamdhip64.so.7), for
    the caller's own device buffer: a caller of the allocation-free entry
    point allocates its inputs and outputs itself."""
    m.lib()
    H = ct.CDLL("libamdhip64.so.7")
    H.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
    H.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
    H.hipFree.argtypes = [ct.c_void_p]
    return H


class DeviceIds:
    def __init__(self, host):
        self.H = hip()
        self.n = host.size
        self.p = ct.c_void_p()
        assert self.H.hipMalloc(ct.byref(self.p), 4 * self.n) == 0
        assert self.H.hipMemcpy(self.p, host.ctypes.data, 4 * self.n, H2D) == 0

    def get(self):
        out = np.empty(self.n, dtype=np.int32)
        assert self.H.hipMemcpy(out.ctypes.data, self.p, 4 * self.n, D2H) == 0
        return out

    def free(self):
        self.H.hipFree(self.p)


def draw(n, count, seed, radius=3.0, length=16.0):
    rng = np.random.default_rng(seed)
    cyl = np.empty((count, 12))
    for k in range(count):
        v, c, lo, hi = gen._fibre_draw(rng, n, radius, length / 2)
        cyl[k, :3], cyl[k, 3:6], cyl[k, 6:9], cyl[k, 9:12] = v, c, lo, hi
    return cyl


def both(n, cyl, target, batch=32, radius=3.0, length=16.0):
    L = m.lib()
    ids_h = np.full((n, n, n), 1, dtype=np.int32)
    u1, c1 = ct.c_int(-1), ct.c_longlong(-1)
    assert L.m3d_gen_cylinders(n, n, n, cyl.ctypes.data_as(dp), len(cyl), length / 2, radius, 0, batch, target,
                               ids_h.ctypes.data_as(i32p), ct.byref(u1), ct.byref(c1)) == 0
    dev = DeviceIds(np.full(n * n * n, 1, dtype=np.int32))
    u2, c2 = ct.c_int(-1), ct.c_longlong(-1)
    rc = L.m3d_gen_cylinders_noalloc(n, n, n, cyl.ctypes.data_as(dp), len(cyl), length / 2, radius, 0, batch, target,
                                     dev.p, ct.byref(u2), ct.byref(c2))
    assert rc == 0, L.m3d_gen_last_error()
    got = dev.get().reshape(n, n, n)
    dev.free()
    return (ids_h, u1.value, c1.value), (got, u2.value, c2.value)


def test_noalloc_equals_allocating_entry_point_twice():
    for n, count, seed, frac in ((96, 2048, 5, 0.3), (40, 160, 6, 0.2)):
        cyl = draw(n, count, seed)
        (a, ua, ca), (b, ub, cb) = both(n, cyl, int(frac * n**3))
        print(f"[noalloc] n {n} cylinders {count}: used {ua}/{ub} count {ca}/{cb}", flush=True)
        assert ua == ub and ca == cb
        assert ca >= int(frac * n**3) or ua + 32 > count
        assert np.array_equal(a, b)
        assert int((b == 0).sum()) == cb


def test_noalloc_rejects_more_cylinders_than_its_buffer():
    L = m.lib()
    cyl = draw(16, 8193, 1)
    dev = DeviceIds(np.zeros(16**3, dtype=np.int32))
    rc = L.m3d_gen_cylinders_noalloc(16, 16, 16, cyl.ctypes.data_as(dp), len(cyl), 8.0, 3.0, 0, 32, 100, dev.p,
                                     None, None)
    dev.free()
    assert rc == -1
    assert b"8192" in L.m3d_gen_last_error()
