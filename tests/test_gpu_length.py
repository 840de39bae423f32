"""String length by chaining (ic_w_evolution.cpp:961-1087) fed from the GPU:
the face crossings the library finds on an evolved lattice, chained by
ax_string_length's cell grid, against the oracle -- its own trajectory, its
crossings and its all-pairs chaining (pinned on the CPU against the
reference's printed lengths, test_string_length.py).  Bar: crossing lists
and the three lengths (total, inner, outer) bit-exact."""
import sys

import numpy as np
import pytest

import axions_amd as ax
from axions_amd.abi import AX_BC_FIXED, AX_BC_NEUMANN, AX_BC_PERIODIC
from oracle import oracle_py as orc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = {
    # the 512^3 run's couplings, periodic
    "periodic_32": (lambda: ax.ah_params(32), 30),
    "periodic_rect": (lambda: ax.ah_params(24, ny=20, nz=64, dx=0.6, dt=0.1, g=0.8), 24),
    # non-periodic x/y (no minimum image across x/y), z periodic / fixed
    "neumann_zper": (lambda: ax.ah_params(20, ny=18, nz=22, g=1.0, alpha=2.0, damped_nt=4, theta_damp_form=1,
                                          boundary=AX_BC_NEUMANN, boundary_z=AX_BC_PERIODIC), 16),
    "fixed_zfixed": (lambda: ax.ah_params(22, ny=20, nz=18, g=1.0, theta_damp_form=1, boundary=AX_BC_FIXED,
                                          boundary_z=AX_BC_FIXED), 16),
}


@pytest.mark.parametrize("name", list(CASES))
def test_chained_length_of_gpu_crossings(name):
    build, steps = CASES[name]
    p = build()
    say(f"[length] {name} {p.nx}x{p.ny}x{p.nz} after {steps} steps")
    st = orc.init_random_phase(p, 3, 0)
    orc.step(p, st, 0, steps, 0, 16)
    pts_o, n_o = orc.crossings(p, st, steps)
    pxy, pz = p.boundary == AX_BC_PERIODIC, p.boundary_z == AX_BC_PERIODIC
    want = orc.string_length(pts_o, p.nx, p.ny, p.nz, p.dx, p.dy, p.dz, pxy, pz)
    with ax.Lattice(p) as lat:
        lat.init_random_phase(3)
        lat.step(steps)
        pts, n = lat.crossings()
    assert n == n_o > 0
    assert np.array_equal(pts.view(np.uint64), pts_o.view(np.uint64))
    got = ax.string_length(pts, p.nx, p.ny, p.nz, p.dx, p.dy, p.dz, pxy, pz)
    assert got == want, (got, want)
    assert got[0] > 0 and abs(got[0] - (got[1] + got[2])) <= 1e-12 * got[0]
