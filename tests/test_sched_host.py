"""Host bookkeeping of the ensemble pool (axions_amd/csrc/ax_sched.h), on the
CPU: the HBM arena allocator against a brute-force model, the per-request
byte count and the xi expression."""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

from axions_amd.abi import lattice_bytes

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def h():
    d = tempfile.mkdtemp(prefix="ax_sched_")
    so = os.path.join(d, "libsched.so")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", "-o", so,
                           os.path.join(HERE, "csrc", "sched_harness.cpp")])
    L = ctypes.CDLL(so)
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    L.arena_new.restype = vp
    L.arena_new.argtypes = [sz]
    L.arena_delete.argtypes = [vp]
    L.arena_alloc.restype = ctypes.c_longlong
    L.arena_alloc.argtypes = [vp, sz]
    L.arena_release.argtypes = [vp, sz]
    for f in ("arena_capacity", "arena_in_use", "arena_peak", "arena_largest_free", "arena_free_ranges"):
        getattr(L, f).restype = sz
        getattr(L, f).argtypes = [vp]
    L.lattice_bytes.restype = sz
    L.lattice_bytes.argtypes = [ctypes.c_int, sz, sz]
    L.xi_of.restype = ctypes.c_double
    L.xi_of.argtypes = [ctypes.c_double] * 7
    ip = ctypes.POINTER(ctypes.c_int)
    L.fill_segments.restype = ctypes.c_longlong
    L.fill_segments.argtypes = [ctypes.c_int, ip, ip, ip, ctypes.c_longlong, ctypes.c_int]
    L.xcd_item.restype = ctypes.c_int
    L.xcd_item.argtypes = [ctypes.c_int, ctypes.c_int]
    return L


def test_arena_random_against_model(h):
    cap = 1 << 20
    a = h.arena_new(cap + 100)          # rounded down to the 256 B alignment
    assert h.arena_capacity(a) == cap
    rng = np.random.default_rng(0)
    live = {}                           # offset -> size (rounded)
    peak = 0
    for it in range(4000):
        if live and rng.random() < 0.45:
            off = int(rng.choice(list(live)))
            assert h.arena_release(a, off) == 1
            del live[off]
        else:
            b = int(rng.integers(1, 40000))
            off = h.arena_alloc(a, b)
            need = (b + 255) // 256 * 256
            # first fit: the lowest free gap that holds `need`
            edges = sorted(live.items())
            gaps, pos = [], 0
            for o, s in edges:
                if o > pos:
                    gaps.append((pos, o - pos))
                pos = o + s
            if pos < cap:
                gaps.append((pos, cap - pos))
            fit = [g for g in gaps if g[1] >= need]
            if not fit:
                assert off == -1
                continue
            assert off == fit[0][0] and off % 256 == 0
            live[off] = need
        used = sum(live.values())
        peak = max(peak, used)
        assert h.arena_in_use(a) == used
        assert h.arena_peak(a) == peak
        # free ranges are coalesced: one per maximal gap
        edges = sorted(live.items())
        ngaps, pos = 0, 0
        for o, s in edges:
            ngaps += o > pos
            pos = o + s
        ngaps += pos < cap
        assert h.arena_free_ranges(a) == ngaps
    assert h.arena_release(a, 12345) == 0  # not a block
    for off in list(live):
        h.arena_release(a, off)
    assert h.arena_in_use(a) == 0 and h.arena_largest_free(a) == cap and h.arena_free_ranges(a) == 1
    h.arena_delete(a)


def test_zero_size_and_exhaustion(h):
    a = h.arena_new(1024)
    offs = [h.arena_alloc(a, 0) for _ in range(4)]
    assert offs == [0, 256, 512, 768] and h.arena_alloc(a, 1) == -1
    h.arena_delete(a)


@pytest.mark.parametrize("n,prec", [(64, 32), (96, 32), (256, 32), (64, 64), (17, 64)])
def test_lattice_bytes(h, n, prec):
    es = 8 if prec == 64 else 4
    assert h.lattice_bytes(n, es, 4352) == lattice_bytes(n, prec)
    assert lattice_bytes(n, prec) >= 10 * n ** 3 * es


def test_xi_expression(h):
    # the expression of ax_diagnostics: (2/3)(dz Wxy + dy Wxz + dx Wyz) / (V dx dy dz)
    for w in [(0, 0, 0), (10, 20, 30), (123456, 7, 99)]:
        V, dx, dy, dz = 64.0 ** 3, 0.5, 0.6, 0.7
        ref = (2.0 / 3.0) * (w[0] * dz + w[1] * dy + w[2] * dx) / (V * dx * dy * dz)
        assert h.xi_of(*w, V, dx, dy, dz) == ref


def items_of(n, xseg):
    """Work items of a chunk: z chunks of 64 x rows x x segments per lattice."""
    return sum(((m + 63) // 64) * m * -(-m // x) for m, x in zip(n, xseg))


def fill_model(n, xseg, want, floor=4):
    """The chunk fill restated: while short of want items, halve every
    segment of the longest length (not below floor)."""
    xs = list(xseg)
    while items_of(n, xs) < want:
        longest = max(xs)
        if longest <= floor:
            break
        xs = [max(longest // 2, floor) if x == longest else x for x in xs]
    return xs, items_of(n, xs)


def test_fill_segments(h):
    rng = np.random.default_rng(3)
    arr = lambda v: (ctypes.c_int * len(v))(*v)
    for case in range(300):
        count = int(rng.integers(1, 12))
        n = [int(rng.choice([64, 96, 128, 192, 256])) for _ in range(count)]
        xseg = [int(rng.choice([4, 8, 16, 32])) for _ in range(count)]
        want = int(rng.integers(1, 40000))
        xs = arr(xseg)
        work = h.fill_segments(count, arr(n), arr([(m + 63) // 64 for m in n]), xs, want, 4)
        got = list(xs)
        ref, ref_work = fill_model(n, xseg, want)
        assert got == ref and work == ref_work == items_of(n, got), (n, xseg, want, got, ref)
        # segments only shorten; the chunk is full or every segment is at the floor
        assert all(g <= x for g, x in zip(got, xseg))
        assert work >= want or all(g == 4 for g in got) or work >= items_of(n, xseg)
    # a full chunk is left as it is
    xs = arr([32, 32])
    assert h.fill_segments(2, arr([256, 256]), arr([4, 4]), xs, 100, 4) == 2 * 4 * 256 * 8
    assert list(xs) == [32, 32]


@pytest.mark.parametrize("q", [1, 2, 7, 64, 1000, 12345])
def test_xcd_item_order(h, q):
    """w -> (w mod 8) q + w / 8 is a bijection of [0, 8q) that gives the
    workgroups of XCD k (w = k mod 8) the contiguous items [k q, (k+1) q),
    in order."""
    its = np.array([h.xcd_item(w, q) for w in range(8 * q)])
    assert np.array_equal(np.sort(its), np.arange(8 * q))
    for k in range(8):
        mine = its[k::8]
        assert np.array_equal(mine, np.arange(k * q, (k + 1) * q))
