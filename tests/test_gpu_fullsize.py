"""GPU checks at the full sizes of the benchmark runs: the 512^3 Abelian-Higgs
lattice in FP64 and FP32 (1 GiB FP64 field planes, the 3.2 GB edge-phase
buffer of the diagnostics), the 256^3 Moore lattice, and an ensemble pool
holding N = 128, 192 and 256 realizations.

The marching kernels are checked against the one-thread-per-site kernels
(AX_FORCE_SITE_KERNEL, AX_DIAG_SITE_KERNEL), which share every per-site
expression with the oracle and are checked against it bit for bit at the
smaller sizes of test_gpu_parity.py.  Bar: states and windings bit-exact;
energy and Gauss sums within 1e-13 relative (same per-site terms, another
summation order); the winding field obeys the cube sum rule (the oriented
sum over the six faces of every cube is 0).
"""
import os
import sys

import numpy as np
import pytest

import axions_amd as ax
from axions_amd.abi import AX_REQ_DONE

pytestmark = pytest.mark.gpu
GB = 1 << 30


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits_equal(a, b):
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


class env:
    def __init__(self, name):
        self.name = name

    def __enter__(self):
        os.environ[self.name] = "1"

    def __exit__(self, *a):
        os.environ.pop(self.name, None)


def march_vs_site(p, warm, steps, seed=42):
    """State after `warm` marching steps, then `steps` more on the march and,
    from the same state, on the site kernel: both results."""
    with ax.Lattice(p) as lat:
        assert lat.tile()[1], "the marching kernel is expected at this size"
        lat.init_random_phase(seed)
        lat.step(warm)
        s0 = lat.get_state()
        lat.step(steps)
        a = lat.get_state()
        lat.set_state(s0, timestep=warm)
        del s0
        with env("AX_FORCE_SITE_KERNEL"):
            assert not lat.tile()[1]
            lat.step(steps)
        b = lat.get_state()
        lat.sync()
    return a, b


@pytest.mark.parametrize("precision", [64, 32])
def test_ah_512_march_equals_site_kernel(precision):
    p = ax.ah_params(512, precision=precision)
    say(f"[fullsize] AH 512^3 fp{precision}: 20 steps, then 2 march vs 2 site")
    a, b = march_vs_site(p, 20, 2)
    assert np.all(np.isfinite(a))
    assert bits_equal(a, b)


def cube_sums(w):
    """Oriented winding sum over the 6 faces of every cube (0 when flux is conserved)."""
    wxy, wxz, wyz = (w[i].astype(np.int16) for i in range(3))
    return (np.roll(wxy, -1, 2) - wxy) - (np.roll(wxz, -1, 1) - wxz) + (np.roll(wyz, -1, 0) - wyz)


def test_ah_512_diagnostics_march_equals_site_and_cube_sum():
    p = ax.ah_params(512)
    say("[fullsize] AH 512^3 fp64 diagnostics: march vs site kernels, cube sum")
    with ax.Lattice(p) as lat:
        lat.init_random_phase(42)
        lat.step(40)
        d_m = lat.diagnostics(winding=True)
        w_m = lat.winding()
        with env("AX_DIAG_SITE_KERNEL"):
            d_s = lat.diagnostics(winding=True)
            w_s = lat.winding()
    assert d_m["finite"] == d_s["finite"] == 1
    assert np.array_equal(w_m, w_s)
    assert d_m["wound_plaquettes"] == d_s["wound_plaquettes"] > 0
    assert rel(d_m["energy"], d_s["energy"]) < 1e-13
    assert rel(d_m["gauss_mean"], d_s["gauss_mean"]) < 1e-13
    w = w_m.reshape(3, 512, 512, 512)
    assert int(np.abs(w.astype(np.int64)).sum()) == d_m["wound_plaquettes"]
    assert not cube_sums(w).any()


def test_moore_256_march_equals_site_kernel():
    p = ax.moore_params(256, dt=0.2)
    say("[fullsize] Moore 256^3 fp64: 12 steps, then 3 march vs 3 site")
    a, b = march_vs_site(p, 12, 3)
    assert np.all(np.isfinite(a))
    assert bits_equal(a, b)


def single_run(n, seed, steps, chunk):
    """Final state and xi samples (t = 0, every chunk end, last step) of one
    FP32 realization through the single-lattice API."""
    p = ax.ah_params(n, precision=32)
    ts, wound, xi = [], [], []
    with ax.Lattice(p) as lat:
        lat.init_random_phase(seed)
        t = 0
        while True:
            d = lat.diagnostics(winding=True)
            ts.append(t)
            wound.append(d["wound_plaquettes"])
            xi.append(d["xi"])
            if t == steps:
                break
            k = min(chunk, steps - t)
            lat.step(k)
            t += k
        return lat.get_state(), np.array(ts), np.array(wound), np.array(xi)


def test_pool_128_192_256_alone_equals_mixed():
    """Realizations of the largest ensemble sizes (steps = N dx / (2 dt) =
    2.5 N) inside a mixed pool give the same bits as alone through the
    single-lattice path."""
    say("[fullsize] pool N = 128, 192, 256 mixed vs alone")
    reqs = [(128, 1001, 320), (192, 1002, 480), (256, 1003, 640)]
    others = [(64, 2001, 160), (96, 2002, 240), (64, 2003, 160)]
    chunk = 10
    with ax.Pool(ax.ah_params(8, precision=32), hbm_budget=8 * GB, chunk_steps=chunk, autotune=0,
                 ic_threads=4) as pool:
        ids = []
        for (n, seed, steps), (on, oseed, osteps) in zip(reqs, others):
            ids.append(pool.submit(seed, n, steps, keep_state=True))
            pool.submit(oseed, on, osteps)
        assert pool.wait(timeout=300)
        s = pool.stats()
        assert s["completed"] == 6 and s["failed"] == 0 and s["max_resident"] >= 2
        for rid, (n, seed, steps) in zip(ids, reqs):
            say(f"[fullsize] pool N={n} vs single lattice")
            assert pool.info(rid)["status"] == AX_REQ_DONE
            st, ts, wound, xi = single_run(n, seed, steps, chunk)
            pts, pw, px = pool.xi(rid)
            assert np.array_equal(pts, ts)
            assert np.array_equal(pw, wound)
            assert np.array_equal(px.view(np.uint64), xi.view(np.uint64))
            assert bits_equal(pool.get_state(rid), st)
            pool.release(rid)


def test_big_planes_march_equals_site_kernel():
    """A lattice whose field planes exceed 4 GiB (520 x 1024 x 1024 FP64,
    4.36 GB per plane, 43.6 GB of state): the march with 64-bit plane
    addressing against the site kernel, and the marching diagnostics against
    the site diagnostics; and the reference's production lattice (1001^3
    FP64, 80 GB) takes the marching kernel."""
    p = ax.ah_params(520, ny=1024, nz=1024)
    say("[fullsize] AH 520x1024x1024 fp64 (4.36 GB planes): 4 steps, then 1 march vs 1 site")
    a, b = march_vs_site(p, 4, 1)
    assert np.all(np.isfinite(a))
    assert bits_equal(a, b)
    del a, b
    with ax.Lattice(p) as lat:
        lat.init_random_phase(42)
        lat.step(3)
        d_m = lat.diagnostics(winding=False)
        with env("AX_DIAG_SITE_KERNEL"):
            d_s = lat.diagnostics(winding=False)
    assert d_m["finite"] == d_s["finite"] == 1
    assert rel(d_m["energy"], d_s["energy"]) < 1e-13
    assert rel(d_m["gauss_mean"], d_s["gauss_mean"]) < 1e-13
    say("[fullsize] AH 1001^3 fp64 (80 GB): marching kernel")
    with ax.Lattice(ax.ah_params(1001)) as lat:
        assert lat.tile()[1]
        lat.step(1)
        lat.sync()
    say("[fullsize] Moore 1001^3 fp64 (112 GB, moore_mpi_evolution.cpp's lattice): marching kernel")
    with ax.Lattice(ax.moore_params(1001)) as lat:
        assert lat.tile()[1]
        lat.step(2)
        lat.sync()
