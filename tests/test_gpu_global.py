"""GPU parity of the global-string momentum-form leapfrog (leapfrog.cpp):
fields bit-exact against the oracle and the compiled reference's outputs,
the per-step energy within 1e-12 (summation order differs)."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import axions_amd as ax
from axions_amd.abi import AX_BC_PERIODIC, AxParams
from oracle import oracle_py as orc

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


CASES = {
    "neumann": (lambda: ax.global_params(12, ny=10, nz=14), 25),
    "periodic_aniso": (lambda: ax.global_params(9, ny=11, nz=13, dx=0.3, dy=0.35, dz=0.28, dt=0.07,
                                                boundary=AX_BC_PERIODIC), 20),
    "neumann_f32": (lambda: ax.global_params(10, ny=12, nz=9, precision=32), 20),
    "periodic_f32": (lambda: ax.global_params(8, ny=9, nz=16, boundary=AX_BC_PERIODIC, precision=32), 20),
}


@pytest.mark.parametrize("name", list(CASES))
def test_global_vs_oracle(name):
    build, nsteps = CASES[name]
    p = build()
    say(f"[global] {name} {p.nx}x{p.ny}x{p.nz}")
    st = orc.init_random_phase(p, 17)
    V = p.nx * p.ny * p.nz
    st[2 * V:] = np.random.default_rng(2).uniform(-0.2, 0.2, 2 * V)     # non-zero momentum
    if p.precision == 32:
        st = st.astype(np.float32)
    with ax.Lattice(p) as lat:
        lat.set_state(st.astype(np.float64))
        assert np.array_equal(lat.get_state(), st.astype(np.float64))
        for n in range(nsteps):
            lat.step(1)
            e = orc.global_step(p, st, 1)[0]
            d = lat.diagnostics(winding=False)
            assert d["timestep"] == n + 1 and d["finite"] == 1
            assert rel(d["energy"], e) < 1e-12, (n, d["energy"], e)
        g = lat.get_state()
        assert np.array_equal(g.view(np.uint64), st.astype(np.float64).view(np.uint64))


@pytest.mark.parametrize("name", ["global_neumann", "global_aniso"])
def test_global_matches_reference_fixture(name):
    fx = json.load(open(os.path.join(GOLDEN, name + ".json")))
    p = AxParams()
    for k, v in fx["params"].items():
        setattr(p, k, v)
    say(f"[global golden] {name}")
    with ax.Lattice(p) as lat:
        lat.init_random_phase(fx["seed"])
        d0 = lat.diagnostics(winding=False)
        assert np.isnan(d0["energy"])
        for n in range(fx["nt"]):
            lat.step(1)
            d = lat.diagnostics(winding=False)
            assert rel(d["energy"], fx["energy"][n]) < 1e-12, (n, d["energy"], fx["energy"][n])
        V = p.nx * p.ny * p.nz
        fin = lat.get_state()[:2 * V].reshape(2, V)
        assert sha(fin) == fx["final_sha256"]


def test_global_winding_and_crossings():
    """Winding of the global model = the gauge-free winding of the same
    field (links zero) in the oracle; crossings as for any scalar field."""
    p = ax.global_params(12, ny=10, nz=14, boundary=AX_BC_PERIODIC)
    say("[global winding]")
    with ax.Lattice(p) as lat:
        lat.init_random_phase(5)
        lat.step(7)
        g = lat.get_state()
        w = lat.winding()
        pts, cnt = lat.crossings()
    V = p.nx * p.ny * p.nz
    pa = ax.ah_params(12, ny=10, nz=14, dx=p.dx, dt=p.dt)
    sa = np.zeros(10 * V)
    sa.reshape(5, 2, V)[0, :] = g[:V]
    sa.reshape(5, 2, V)[1, :] = g[V:2 * V]
    assert np.array_equal(w, orc.winding(pa, sa, 0, 0))
    po, co = orc.crossings(pa, sa, 0)
    assert cnt == co and np.array_equal(pts, po)
