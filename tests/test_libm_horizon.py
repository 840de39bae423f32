"""How far the reference as shipped (glibc sin/cos) drifts from the library's
trajectory (the shared sin/cos of include/axions/ax_trig.h) on the seeded
64^3 plain Abelian-Higgs run, through the oracle, which is bit-exact against
the reference built with either trig (test_oracle_golden.py).

Pinned here: the first TimeStep at which the largest field difference,
relative to the largest field magnitude, exceeds 1e-12 (444), and that it
stays below 1e-6 up to there; that about 3 % of the link-angle sines and
cosines and 0.3 % of the plaquette sines of the initial steps differ from
glibc in the last bit.  oracle/libm_horizon.py runs the whole 1000 steps
(energy 4e-12 and xi 0 relative difference at step 1000)."""
import os
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "oracle"))

import libm_horizon  # noqa: E402


def test_glibc_trajectory_stays_within_1e12_until_step_444():
    r = libm_horizon.run(n=64, steps=450, threads=max(1, min(8, os.cpu_count() or 1)), report=(200,),
                         sample_every=100)
    assert r["first_above_1e-12"] == 444
    assert r["first_above_1e-6"] is None
    at = r["at"]["200"]
    assert at["field_rel"] < 1e-12 and at["energy_rel"] < 1e-14 and at["xi_rel"] == 0
    tr = r["trig"]
    assert 0.02 < tr["link_sin_frac"] < 0.05 and 0.02 < tr["link_cos_frac"] < 0.05
    assert 0.001 < tr["plaq_sin_frac"] < 0.01
