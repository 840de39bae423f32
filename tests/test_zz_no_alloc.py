"""ax_crossings_noalloc (the crossing detector without device allocation)
against ax_crossings: the same points, bit for bit, and the same count, on
two calls in a row -- the second on a smaller lattice -- into caller-owned
device buffers, plus a truncated call and the size limit."""
import ctypes
import sys

import numpy as np
import pytest

import axions_amd as ax
from axions_amd.abi import AX_EINVAL
from axions_amd.lattice import lib

pytestmark = pytest.mark.gpu

hip = ctypes.CDLL("libamdhip64.so")
hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
hip.hipFree.argtypes = [ctypes.c_void_p]
hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
hip.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
D2H = 2


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ok(rc):
    assert rc == 0, f"HIP error {rc}"


class DeviceBuffer:
    def __init__(self, nbytes):
        self.nbytes = nbytes
        self.ptr = ctypes.c_void_p()
        ok(hip.hipMalloc(ctypes.byref(self.ptr), nbytes))
        ok(hip.hipMemset(self.ptr, 0xFF, nbytes))

    def read(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        if count:
            ok(hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H))
        return out

    def free(self):
        ok(hip.hipFree(self.ptr))


def free_bytes():
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    ok(hip.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)))
    return f.value


def noalloc(lat, xyz, cap, cnt):
    rc = lib.ax_crossings_noalloc(lat.handle, xyz.ptr, cap, cnt.ptr)
    assert rc == 0, lib.ax_last_error().decode()
    lat.sync()
    n = int(cnt.read(np.int64, 1)[0])
    return xyz.read(np.float64, 3 * min(n, cap)).reshape(-1, 3), n


def test_noalloc_crossings_equal_allocating_entry_point():
    say("[noalloc] 64^3 fp64 then 40x24x32 fp32 into one pair of device buffers")
    cap = 400000
    xyz, cnt = DeviceBuffer(3 * 8 * cap), DeviceBuffer(8)
    try:
        with ax.Lattice(ax.ah_params(64)) as lat:
            lat.init_random_phase(42)
            lat.step(30)
            ref, n_ref = lat.crossings()
            assert 0 < n_ref <= cap
            got, n = noalloc(lat, xyz, cap, cnt)
            assert n == n_ref
            assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
        with ax.Lattice(ax.ah_params(40, ny=24, nz=32, precision=32)) as lat:
            lat.init_random_phase(7)
            lat.step(12)
            ref, n_ref = lat.crossings()
            assert 0 < n_ref <= cap
            before = free_bytes()
            got, n = noalloc(lat, xyz, cap, cnt)
            assert free_bytes() == before          # nothing allocated on the device
            assert n == n_ref
            assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
            # a capacity below the count: the first cap points, the full count
            small = n_ref // 3
            got, n = noalloc(lat, xyz, small, cnt)
            assert n == n_ref
            assert np.array_equal(got.view(np.uint64), ref[:small].view(np.uint64))
    finally:
        xyz.free()
        cnt.free()


def test_noalloc_crossings_refuse_lattices_beyond_the_scratch():
    say("[noalloc] size limit")
    cnt = DeviceBuffer(8)
    try:
        with ax.Lattice(ax.ah_params(520, ny=512, nz=512, precision=32)) as lat:
            rc = lib.ax_crossings_noalloc(lat.handle, None, 0, cnt.ptr)
            assert rc == AX_EINVAL
            assert b"exceeds" in lib.ax_last_error()
    finally:
        cnt.free()
