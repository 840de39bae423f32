"""The oracle's straight-string pipeline against the compiled reference
(singleSemiAnalytic.cpp SOR profile, straight_ic.cpp lattice) and against
what the mathematics fixes.  CPU only."""
import hashlib
import json
import math
import os

import numpy as np
import pytest

from axions_amd.abi import AX_BC_FIXED, ah_params
from oracle import oracle_py as orc

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def fixture(name):
    with open(os.path.join(GOLD, f"{name}.json")) as f:
        return json.load(f)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def sor_args(c):
    return (c["nx"], c["h"], c["w"], c["tol"], c["n"], c["lam"], c["eta"], c["g"], c["max_iter"])


@pytest.mark.parametrize("name", ["sor_gauged", "sor_cap"])
def test_sor_oracle_matches_reference(name):
    fx = fixture(name)
    prof, it, res = orc.sor_profile(*sor_args(fx["consts"]))
    assert it == fx["iterations"]
    assert np.array_equal(prof[fx["sample_idx"]], np.array(fx["sample"]))
    assert sha(prof) == fx["profile_sha256"]
    if it < fx["consts"]["max_iter"]:
        assert res <= fx["consts"]["tol"]


def test_sor_restart_continues_exactly():
    """Starting from a previous output (the reference's ic_type 1) continues
    the same sweeps: 120 + 80 sweeps == 200 sweeps, bit for bit."""
    args = (301, 0.05, 1.5, 1e-12, 1.0, 1.0, 1.0, 0.7)
    a, ia, _ = orc.sor_profile(*args, 120)
    b, ib, _ = orc.sor_profile(*args, 80, init=a)
    c, ic_, _ = orc.sor_profile(*args, 200)
    assert (ia, ib, ic_) == (120, 80, 200)
    assert np.array_equal(b, c)


def test_sor_zero_sweeps_when_tol_not_below_one():
    """tolTest starts at 1 (singleSemiAnalytic.cpp:80): tol >= 1 runs no sweep."""
    p, it, res = orc.sor_profile(51, 0.1, 1.5, 1.0, 1.0, 1.0, 1.0, 0.0, 100)
    assert it == 0 and res == 1.0
    f = [math.tanh(0.5 * 0.1 * i) for i in range(1, 51)]     # libm tanh, as the reference
    assert np.array_equal(p[1:, 0], np.array(f))
    assert p[0, 0] == 0 and not p[:, 1].any()


def test_sor_global_profile_far_field():
    """Global string (g = 0): far from the core f -> eta - n^2 / (lambda eta
    r^2), the large-r limit of f'' + f'/r - n^2 f / r^2 - lambda/2 (f^2 -
    eta^2) f = 0 (the equation the residual relaxes).  A stays 0."""
    nx, h, lam = 1201, 0.05, 1.0
    prof, it, res = orc.sor_profile(nx, h, 1.5, 1e-7, 1.0, lam, 1.0, 0.0, 400000)
    assert res <= 1e-7 and not prof[:, 1].any()
    for j in (300, 400, 500):                      # r = 15, 20, 25
        r = j * h
        assert abs(prof[j, 0] - (1 - 1 / (lam * r * r))) < 2e-4
    f = prof[:, 0]
    assert np.all(np.diff(f[:-1]) > 0)             # monotone from 0 to 1


def test_sor_gauged_profile_limits():
    """Gauged string: f -> eta and A -> n/g away from the core."""
    fx = fixture("sor_gauged")
    prof, _, _ = orc.sor_profile(*sor_args(fx["consts"]))
    assert abs(prof[-2, 0] - 1) < 1e-4 and abs(prof[-2, 1] - 1) < 1e-4


def straight_profile_for(fx):
    src = fixture(fx["profile"])
    prof, _, _ = orc.sor_profile(*sor_args(src["consts"]))
    assert sha(prof) == src["profile_sha256"]
    return prof[:fx["SORnx"]]


def test_straight_oracle_matches_reference_gauged():
    fx = fixture("straight_gauged")
    nx, ny, nz = fx["lattice"]
    dx, dy, dz = fx["spacing"]
    st, off = orc.straight_string(nx, ny, nz, dx, dy, dz, straight_profile_for(fx), fx["SORa"], fx["n"], fx["g"])
    assert off == fx["off_grid"]
    assert np.array_equal(st.T[fx["sample_idx"]], np.array(fx["sample"]))
    assert sha(st) == fx["ic_sha256"]


def test_straight_oracle_matches_reference_shipped():
    """straight_ic.cpp's own constants on the shipped 30001-point profile
    (read from the reference run's output when it is present here)."""
    path = os.path.join(REPO, "oracle", "_ref", "sor_ref", "Data", "SOR_Fields.txt")
    if not os.path.exists(path):
        pytest.skip("reference SOR output not built here")
    fx = fixture("straight_ref")
    prof = np.loadtxt(path, dtype=np.float64)
    assert sha(prof) == fixture("sor_ref")["profile_sha256"]
    nx, ny, nz = fx["lattice"]
    dx, dy, dz = fx["spacing"]
    st, off = orc.straight_string(nx, ny, nz, dx, dy, dz, prof[:fx["SORnx"]], fx["SORa"], fx["n"], fx["g"])
    assert off == fx["off_grid"] == 0
    assert sha(st) == fx["ic_sha256"]


def test_straight_interpolation_exact_on_quadratics():
    """Three-point quadratic interpolation reproduces a quadratic profile at
    every distance (and the linear branch a linear one): |phi| and r^2 |A|
    / r equal the profile evaluated at the site's r up to rounding."""
    dr = 0.1
    r = dr * np.arange(400)
    fq = 0.3 + 0.2 * r - 0.01 * r * r
    aq = -0.1 + 0.05 * r + 0.002 * r * r
    prof = np.stack([fq, aq], axis=1)
    nx, ny, nz, dx, dy = 31, 28, 2, 0.37, 0.41
    g = 0.8
    st, off = orc.straight_string(nx, ny, nz, dx, dy, 0.5, prof, dr, 1.0, g)
    assert off == 0
    x = (np.arange(nx) - (nx - 1) // 2) * dx
    y = (np.arange(ny) - (ny - 1) // 2) * dy
    X, Y = np.meshgrid(x, y, indexing="ij")
    R = np.sqrt(X * X + Y * Y)
    f0 = st[0].reshape(nx, ny, nz)[:, :, 0]
    f1 = st[1].reshape(nx, ny, nz)[:, :, 0]
    t0 = st[2].reshape(nx, ny, nz)[:, :, 0]
    t1 = st[3].reshape(nx, ny, nz)[:, :, 0]
    m = R > 1.5 * dr                                   # quadratic branch
    mag = np.hypot(f0, f1)
    assert np.allclose(mag[m], (0.3 + 0.2 * R - 0.01 * R * R)[m], rtol=1e-12, atol=1e-13)
    # A = A(r) (-y, x) / r^2, stored as dx g A_x, dy g A_y
    amag = np.hypot(t0 / (dx * g), t1 / (dy * g)) * R
    assert np.allclose(amag[m], np.abs(-0.1 + 0.05 * R + 0.002 * R * R)[m], rtol=1e-12, atol=1e-13)
    # direction: phi along (x, y)/r, A along (-y, x)
    assert np.allclose(f0[m] * Y[m], f1[m] * X[m], atol=1e-12)
    assert not st[4].any()
    lin = np.stack([0.5 + 2.0 * r, 1.0 - r], axis=1)
    st2, _ = orc.straight_string(5, 5, 1, 0.01, 0.01, 0.5, lin, dr, 1.0, g)
    R2 = np.sqrt(sum(np.meshgrid(((np.arange(5) - 2) * 0.01) ** 2, ((np.arange(5) - 2) * 0.01) ** 2,
                                  indexing="ij"))).ravel()
    nz_ = R2 > 0
    assert np.allclose(np.hypot(st2[0], st2[1])[nz_], (0.5 + 2.0 * R2)[nz_], rtol=1e-13)


def test_straight_string_phase_and_flux_around_core():
    """Around any rectangle enclosing the core the phase of phi turns by
    exactly +2 pi (phi points along (x, y)); far out, where A = n/g, the link
    variables summed around the rectangle approach g * (circulation of A) =
    2 pi n.  (The string passes through the site (x0, y0), where phi = 0, so
    per-plaquette windings next to it are not defined.)"""
    fx = fixture("straight_gauged")
    nx, ny, nz, dx, dy = 40, 33, 2, 0.45, 0.55
    prof = straight_profile_for(fx)
    cols, off = orc.straight_string(nx, ny, nz, dx, dy, 0.5, prof, fx["SORa"], 1.0, 1.0)
    assert off > 0
    f = (cols[0] + 1j * cols[1]).reshape(nx, ny, nz)[:, :, 0]
    tx = cols[2].reshape(nx, ny, nz)[:, :, 0]
    ty = cols[3].reshape(nx, ny, nz)[:, :, 0]
    x0, y0 = (nx - 1) // 2, (ny - 1) // 2
    for half in (1, 3, 8, 15):
        xs, xe, ys, ye = x0 - half, x0 + half, y0 - half, y0 + half
        loop = ([(i, ys) for i in range(xs, xe)] + [(xe, j) for j in range(ys, ye)] +
                [(i, ye) for i in range(xe, xs, -1)] + [(xs, j) for j in range(ye, ys, -1)])
        ph = sum(np.angle(f[b] * np.conj(f[a])) for a, b in zip(loop, loop[1:] + loop[:1]))
        assert abs(ph - 2 * np.pi) < 1e-12
        if half == 15:
            circ = (sum(tx[i, ys] for i in range(xs, xe)) + sum(ty[xe, j] for j in range(ys, ye)) -
                    sum(tx[i, ye] for i in range(xs, xe)) - sum(ty[xs, j] for j in range(ys, ye)))
            assert abs(circ - 2 * np.pi) < 0.05 * 2 * np.pi
