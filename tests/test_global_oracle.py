"""The global-string momentum-form leapfrog (leapfrog.cpp) in the oracle,
against the compiled reference's own outputs and the invariants the
mathematics fixes."""
import hashlib
import json
import os

import numpy as np
import pytest

from axions_amd.abi import AX_BC_PERIODIC, AxParams, global_params
from oracle import oracle_py as orc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


@pytest.mark.parametrize("name", ["global_neumann", "global_aniso"])
def test_oracle_matches_reference(name):
    fx = json.load(open(os.path.join(GOLDEN, name + ".json")))
    p = AxParams()
    for k, v in fx["params"].items():
        setattr(p, k, v)
    st = orc.init_random_phase(p, fx["seed"])
    e = orc.global_step(p, st, fx["nt"])
    assert np.array_equal(e, np.array(fx["energy"]))            # one-thread loop order: exact
    V = p.nx * p.ny * p.nz
    fin = st[:2 * V].reshape(2, V)
    assert sha(fin) == fx["final_sha256"]
    assert np.array_equal(fin[:, fx["final_sample_idx"]].T, np.array(fx["final_sample"]))


def test_vacuum_is_static_and_has_zero_energy():
    p = global_params(8)
    st = np.zeros(4 * 8 ** 3)
    st[:8 ** 3] = 1.0
    e = orc.global_step(p, st, 5)
    assert np.all(e == 0) and np.all(st[:8 ** 3] == 1.0) and np.all(st[8 ** 3:] == 0)


def test_periodic_translation_invariance():
    p = global_params(7, ny=6, nz=9, boundary=AX_BC_PERIODIC)
    st = orc.init_random_phase(p, 3)
    s = st.reshape(4, 7, 6, 9)
    s[2:] = np.random.default_rng(1).uniform(-0.3, 0.3, s[2:].shape)
    sh = np.ascontiguousarray(np.roll(s, (2, -1, 4), axis=(1, 2, 3))).ravel()
    orc.global_step(p, st, 6)
    orc.global_step(p, sh, 6)
    assert np.array_equal(np.roll(sh.reshape(4, 7, 6, 9), (-2, 1, -4), axis=(1, 2, 3)).ravel(), st)


def test_small_amplitude_plane_wave():
    """Linearised about phi = 1, Im phi obeys the massless lattice wave
    equation.  Started with zero momentum, the leapfrog's discrete solution
    is A cos(kz) cos(w (n + 1/2) dt) / cos(w dt / 2) with
    w dt = 2 asin((dt/dz) sin(k dz / 2)); a dropped or doubled term in the
    Laplacian, the momentum or the field update changes it at O(1)."""
    n, dz, dt, A = 16, 0.25, 0.05, 1e-6
    p = global_params(4, ny=4, nz=n, dx=dz, dt=dt, boundary=AX_BC_PERIODIC)
    k = 2 * np.pi / (n * dz)
    z = np.arange(n) * dz
    st = np.zeros(4 * 16 * n)
    s = st.reshape(4, 4, 4, n)
    s[0] = 1.0
    s[1] = A * np.cos(k * z)
    w = (2 / dt) * np.arcsin((dt / dz) * np.sin(k * dz / 2))
    steps = 200
    orc.global_step(p, st, steps)
    got = s[1, 0, 0]
    ref = A * np.cos(k * z) * np.cos(w * (steps + 0.5) * dt) / np.cos(w * dt / 2)
    assert np.max(np.abs(got - ref)) < 1e-6 * A
