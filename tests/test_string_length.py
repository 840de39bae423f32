"""String length by chaining crossing points (ic_w_evolution.cpp:961-1087).

The library's grid-bucketed chaining (ax_string_length) against the
oracle's all-pairs restatement of the reference's loop, and both against
the compiled reference's own printed length for its seeded random initial
condition (tests/golden/icw_len_*.json)."""
import json
import os

import numpy as np
import pytest

import axions_amd as ax
from axions_amd.abi import ah_params
from oracle import oracle_py as orc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def minstd_uniform(seed, count, a=-1.0, b=1.0):
    """libstdc++ minstd_rand + uniform_real_distribution<double>(a, b):
    generate_canonical<double, 53> takes two draws (log2 r truncates to 30),
    sum = (g1 - 1) + (g2 - 1) * r, divided by r*r rounded to double."""
    m, mult = 2147483647, 48271
    x = seed % m or 1
    r = np.float64(2147483646.0)
    tmp2 = np.float64(float(2147483646 ** 2))
    out = np.empty(count)
    for i in range(count):
        x = (mult * x) % m
        s = np.float64(x - 1) * np.float64(1.0)
        x = (mult * x) % m
        s = s + np.float64(x - 1) * r
        u = s / tmp2
        if u >= 1.0:
            u = np.nextafter(1.0, 0.0)
        out[i] = u * (b - a) + a
    return out


def test_minstd_known_answer():
    """C++ [rand.predef]: the 10000th output of minstd_rand is 399268537."""
    x = 1
    for _ in range(10000):
        x = (48271 * x) % 2147483647
    assert x == 399268537


@pytest.mark.parametrize("name", ["icw_len_split", "icw_len_total"])
def test_length_matches_reference(name):
    fx = json.load(open(os.path.join(GOLDEN, name + ".json")))
    nx, ny, nz = fx["lattice"]
    dx, dy, dz = fx["spacing"]
    V = nx * ny * nz
    u = minstd_uniform(fx["ic"]["seed"], 2 * V)
    p = ah_params(nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=dz, g=0.0)
    st = np.zeros(10 * V)
    s = st.reshape(5, 2, V)
    s[0, :] = u[0::2]          # Re phi, both levels
    s[1, :] = u[1::2]
    pts, cnt = orc.crossings(p, st, 0)
    assert cnt > 1000
    lib = ax.string_length(pts, nx, ny, nz, dx, dy, dz)
    brute = orc.string_length(pts, nx, ny, nz, dx, dy, dz)
    assert lib == brute
    if fx["split"]:
        assert [lib[1], lib[2]] == fx["length"]
    else:
        assert [lib[0]] == fx["length"]


def test_grid_equals_all_pairs_on_random_clouds():
    rng = np.random.default_rng(0)
    for trial in range(30):
        n = int(rng.integers(0, 2500))
        nx, ny, nz = (int(v) for v in rng.integers(6, 40, 3))
        dx, dy, dz = rng.uniform(0.2, 0.8, 3)
        x0, y0, z0 = (int(0.5 * (k - 1)) for k in (nx, ny, nz))
        pts = np.stack([(rng.integers(0, nx, n) - x0 + rng.choice([0.0, 0.5, 1.0], n)) * dx,
                        (rng.integers(0, ny, n) - y0 + rng.random(n)) * dy,
                        (rng.integers(0, nz, n) - z0 + rng.random(n)) * dz], 1)
        for pxy, pz in ((True, True), (False, True), (True, False), (False, False)):
            assert ax.string_length(pts, nx, ny, nz, dx, dy, dz, pxy, pz) == \
                orc.string_length(pts, nx, ny, nz, dx, dy, dz, pxy, pz), (trial, pxy, pz)


def test_straight_string_length():
    """A straight z string through a 12^3 box: one crossing per xy face layer,
    chained into a closed loop of length N dz (the last step back to the start
    adds the cutoff distance, the reference's rule)."""
    n, d = 12, 0.5
    z0 = int(0.5 * (n - 1))
    pts = np.array([[0.1, 0.2, (k - z0) * d + 0.0] for k in range(n)])
    tot, inner, outer = ax.string_length(pts, n, n, n, d, d, d)
    cut = 3 * d * d
    assert tot == pytest.approx((n - 1) * d + cut)
    assert ax.string_length(pts[:0], n, n, n, d, d, d) == (0.0, 0.0, 0.0)
