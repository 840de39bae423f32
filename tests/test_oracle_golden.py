"""Pins the oracle (oracle/oracle.cpp) to the reference.

tests/golden/*.json were written by tests/golden/make_golden.py from the
reference's own programs (evolution.cpp, mooreEvol.cpp), compiled from
sjcott/Axions by oracle/ref/build_ref.py and fed the seeded IC of
mpi_evolution.cpp:287-312.  Configurations marked "libm" ran the reference
with glibc sin/cos and are replayed by the oracle with glibc; "shared" ran
it with the shared ax_sin/ax_cos routed in by a -include shim.  Bit-exact
equality is required for the fields, every per-step energy and Gauss value
and every string-crossing frame."""
import glob
import hashlib
import json
import os

import numpy as np
import pytest

from axions_amd.abi import AxParams, AX_BC_PERIODIC, AX_MODEL_MOORE
from oracle import oracle_py as orc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = sorted(os.path.basename(p)[:-5] for p in glob.glob(os.path.join(GOLDEN, "*.json"))
                  if not os.path.basename(p).startswith(("global_", "icw_", "sor_", "straight_")))
# global_*: leapfrog.cpp (test_global_oracle.py); icw_*: string length (test_string_length.py);
# sor_*, straight_*: straight-string pipeline (test_profile_oracle.py)
# evo_c1_* (64^3 x 200) and evo_c2_* (64^3 x 1000) take minutes in the oracle:
# the GPU runs them against the fixture directly (test_gpu_parity.py)
FAST = [f for f in FIXTURES if not f.startswith(("evo_c1", "evo_c2"))]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def load(name):
    fx = json.load(open(os.path.join(GOLDEN, name + ".json")))
    p = AxParams()
    for k, v in fx["params"].items():
        setattr(p, k, v)
    return fx, p


def replay(name, nthreads=4):
    """Run the oracle over the fixture's schedule; returns (fx, p, state, E, G, frames_ok)."""
    fx, p = load(name)
    tm = 1 if fx["trig"] == "libm" else 0
    st = orc.init_random_phase(p, fx["seed"], 0)
    frames = {f["step"]: f for f in fx["frames"]}
    E, G, bad = [], [], []
    for n in range(fx["nt"]):
        d = orc.diag(p, st, n, tm)
        E.append(d["energy"])
        G.append(d["gauss"])
        if n in frames:
            pts, c = orc.crossings(p, st, n)
            if c != frames[n]["count"] or sha(pts) != frames[n]["sha256"]:
                bad.append(n)
        orc.step(p, st, n, 1, tm, nthreads)
    return fx, p, st, np.array(E), np.array(G), bad


@pytest.mark.parametrize("name", FAST)
def test_oracle_matches_reference(name):
    fx, p, st, E, G, bad = replay(name)
    cur, _ = orc.levels(p, st, fx["nt"])
    flat = cur.reshape(cur.shape[0], -1)
    assert np.array_equal(flat[:, fx["final_sample_idx"]].T, np.array(fx["final_sample"]))
    assert sha(cur) == fx["final_sha256"]
    assert np.array_equal(E, np.array(fx["energy"]))
    if p.model == AX_MODEL_MOORE and p.boundary != AX_BC_PERIODIC:
        # mooreEvol.cpp:697 divides its running mean by (nx-2)(ny-2)(nz-2) a second time
        G = G / ((p.nx - 2) * (p.ny - 2) * (p.nz - 2))
    assert np.array_equal(G, np.array(fx["gauss"]))
    assert not bad, f"crossing frames differ at steps {bad}"


def test_mt19937_known_answer():
    # C++ [rand.predef]: the 10000th consecutive output of a default-constructed
    # mt19937 (seed 5489) is 4123659995
    assert int(orc.mt19937(5489, 10000)[-1]) == 4123659995


def test_phases_match_libstdcxx():
    import axions_amd
    a = axions_amd.random_phases(42, 200000)
    b = orc.phases(42, 200000)
    assert np.array_equal(a, b)
    assert a.min() >= 0 and a.max() < 2 * np.pi
