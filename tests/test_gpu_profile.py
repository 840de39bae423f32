"""GPU parity of the straight-string pipeline: the pipelined SOR profile
(ax_sor_profile) and the lattice interpolation (ax_init_straight_string)
against the compiled reference's outputs (tests/golden) and the oracle.
Bar: bit-exact profile, sweep count and lattice values."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import axions_amd as ax
from axions_amd.abi import AX_BC_FIXED
from oracle import oracle_py as orc

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def fixture(name):
    with open(os.path.join(GOLD, f"{name}.json")) as f:
        return json.load(f)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def gpu_sor(c, init=None):
    return ax.sor_profile(nx=c["nx"], h=c["h"], w=c["w"], tol=c["tol"], n=c["n"], lam=c["lam"], eta=c["eta"],
                          g=c["g"], max_iter=c["max_iter"], init=init)


@pytest.mark.parametrize("name", ["sor_gauged", "sor_cap", "sor_ref"])
def test_sor_profile_matches_reference(name):
    """sor_ref is the reference's own constants: 30001 points, 23278 sweeps."""
    fx = fixture(name)
    c = fx["consts"]
    say(f"[sor] {name} nx={c['nx']}")
    prof, it, res = gpu_sor(c)
    assert it == fx["iterations"]
    assert np.array_equal(prof[fx["sample_idx"]], np.array(fx["sample"]))
    assert sha(prof) == fx["profile_sha256"]
    if it < c["max_iter"]:
        assert res <= c["tol"]


@pytest.mark.parametrize("case", [
    # (nx, h, w, tol, n, lam, eta, g, max_iter)
    (3, 0.1, 1.5, 1e-9, 1.0, 1.0, 1.0, 0.0, 50),         # one interior point
    (4, 0.1, 1.2, 1e-9, 1.0, 2.0, 1.0, 1.0, 257),        # one sweep past a full launch
    (130, 0.07, 1.5, 1e-6, 2.0, 1.5, 0.8, 0.6, 1000),    # n = 2, eta != 1
    (301, 0.05, 1.5, 1e-12, 1.0, 1.0, 1.0, 0.7, 5),      # fewer sweeps than lanes
    (51, 0.1, 1.5, 1.0, 1.0, 1.0, 1.0, 0.0, 100),        # tol >= 1: no sweep
    (90, 0.1, 1.5, 1e-3, 1.0, 1.0, 1.0, 0.0, 0),         # max_iter 0
])
def test_sor_profile_matches_oracle(case):
    nx, h, w, tol, n, lam, eta, g, mi = case
    say(f"[sor-oracle] {case}")
    ref, rit, rres = orc.sor_profile(nx, h, w, tol, n, lam, eta, g, mi)
    prof, it, res = ax.sor_profile(nx=nx, h=h, w=w, tol=tol, n=n, lam=lam, eta=eta, g=g, max_iter=mi)
    assert it == rit
    assert res == rres
    assert np.array_equal(prof.view(np.uint64), ref.view(np.uint64))


def test_sor_restart_equals_continuous():
    """A start from a previous output (the reference's ic_type 1) continues
    the same sweeps: 300 + 213 == 513, bit for bit, and equal to the oracle."""
    kw = dict(nx=400, h=0.05, w=1.5, tol=1e-14, n=1.0, lam=1.0, eta=1.0, g=0.7)
    a, _, _ = ax.sor_profile(max_iter=300, **kw)
    b, ib, _ = ax.sor_profile(max_iter=213, init=a, **kw)
    c, ic_, _ = ax.sor_profile(max_iter=513, **kw)
    assert (ib, ic_) == (213, 513)
    assert np.array_equal(b, c)
    o, _, _ = orc.sor_profile(400, 0.05, 1.5, 1e-14, 1.0, 1.0, 1.0, 0.7, 513)
    assert np.array_equal(c, o)


def test_sor_rejects_bad_arguments():
    with pytest.raises(ax.AxError):
        ax.sor_profile(nx=2)
    with pytest.raises(ax.AxError):
        ax.sor_profile(nx=10, h=0.0)


def lattice_columns(lat, p):
    """Level-0 and level-1 planes as [2, 5, V]."""
    st = lat.get_state().reshape(5, 2, -1)
    return st[:, 0, :], st[:, 1, :]


@pytest.mark.parametrize("name", ["straight_gauged", "straight_ref"])
def test_straight_string_matches_reference(name):
    fx = fixture(name)
    src = fixture(fx["profile"])
    say(f"[straight] {name} from {fx['profile']}")
    prof, _, _ = gpu_sor(src["consts"])
    assert sha(prof) == src["profile_sha256"]
    nx, ny, nz = fx["lattice"]
    dx, dy, dz = fx["spacing"]
    p = ax.ah_params(nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=dz, g=fx["g"], boundary=AX_BC_FIXED)
    with ax.Lattice(p) as lat:
        lat.init_random_phase(3)                  # overwritten completely
        off = lat.init_straight_string(prof[:fx["SORnx"]], fx["SORa"], n=fx["n"])
        assert lat.timestep == 0
        c0, c1 = lattice_columns(lat, p)
    assert off == fx["off_grid"]
    assert np.array_equal(c0.T[fx["sample_idx"]], np.array(fx["sample"]))
    assert sha(c0) == fx["ic_sha256"]
    assert np.array_equal(c0, c1)


@pytest.mark.parametrize("precision", [64, 32])
def test_straight_string_matches_oracle_and_evolves(precision):
    """Odd/even sizes, anisotropic spacings, an explicit g unlike the
    lattice's, sites beyond the profile; then 30 steps with fixed edges
    equal the oracle's (FP64: bit-exact; FP32: the float restatement)."""
    fx = fixture("sor_gauged")
    prof, _, _ = gpu_sor(fx["consts"])
    nx, ny, nz = 37, 30, 64
    p = ax.ah_params(nx, ny=ny, nz=nz, dx=0.6, dy=0.45, dz=0.5, dt=0.1, g=0.9, boundary=AX_BC_FIXED,
                     precision=precision)
    say(f"[straight-evolve] fp{precision}")
    cols, off_o = orc.straight_string(nx, ny, nz, 0.6, 0.45, 0.5, prof[:500], 0.02, 1.0, 1.3)
    want = np.repeat(cols[:, None, :], 2, axis=1).ravel()
    if precision == 32:
        want = want.astype(np.float32)
    with ax.Lattice(p) as lat:
        off = lat.init_straight_string(prof[:500], 0.02, n=1.0, g=1.3)
        assert off == off_o > 0
        assert np.array_equal(lat.get_state(), want.astype(np.float64))
        lat.step(30)
        got = lat.get_state()
    if precision == 64:
        orc.step(p, want, 0, 30, 0, 16)
    else:
        orc.step_r(p, want, 0, 30, 16)
    assert np.array_equal(got, want.astype(np.float64))


def test_straight_string_global_model():
    """Global model: phi in both parity buffers, zero momentum; one
    leapfrog step equals the oracle's."""
    prof, _, _ = ax.sor_profile(nx=3001, h=0.01, w=1.5, tol=1e-2, n=1.0, lam=1.0, eta=1.0, g=0.0,
                                max_iter=2000)
    p = ax.global_params(21, ny=18, nz=7)
    say("[straight-global]")
    cols, _ = orc.straight_string(21, 18, 7, p.dx, p.dy, p.dz, prof, 0.01, 1.0, 0.0)
    V = 21 * 18 * 7
    want = np.concatenate([cols[0], cols[1], np.zeros(2 * V)])
    with ax.Lattice(p) as lat:
        lat.init_straight_string(prof, 0.01, n=1.0, g=0.0)
        assert np.array_equal(lat.get_state(), want)
        lat.step(3)
        got = lat.get_state()
    orc.global_step(p, want, 3)
    assert np.array_equal(got, want)


def test_straight_string_rejects_moore_and_bad_profiles():
    prof = np.zeros((10, 2))
    with ax.Lattice(ax.moore_params(8, boundary=AX_BC_FIXED)) as lat:
        with pytest.raises(ax.AxError):
            lat.init_straight_string(prof, 0.1)
    with ax.Lattice(ax.ah_params(8)) as lat:
        with pytest.raises(ax.AxError):
            lat.init_straight_string(prof[:1], 0.1)
        with pytest.raises(ax.AxError):
            lat.init_straight_string(prof, 0.0)
