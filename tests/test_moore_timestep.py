"""The Moore model's own time step, dt = 0.3 (mooreEvol.cpp:26, with dx = 0.7
and the couplings of :23-36), is unstable from the seeded random-phase
initial condition of the 256^3 benchmark run: the oracle's periodic
trajectory overflows within a few tens of steps at every size tried (32^3:
first non-finite value after TimeStep 22; 64^3: 19; 96^3: 16), so a timed
run at dt = 0.3 would time NaN arithmetic.  The benchmark therefore runs the
Moore model at dt = 0.2, which stays finite."""
import numpy as np

from axions_amd.abi import moore_params
from oracle import oracle_py as orc


def first_nonfinite(p, limit, threads=8):
    st = orc.init_random_phase(p, 42, 0)
    for s in range(limit):
        orc.step(p, st, s, 1, 0, threads)
        if not np.all(np.isfinite(st)):
            return s + 1
    return None


def test_reference_time_step_diverges_from_random_phases():
    assert first_nonfinite(moore_params(32, dt=0.3), 60) == 22


def test_benchmark_time_step_stays_finite():
    assert first_nonfinite(moore_params(24, dt=0.2), 300) is None
