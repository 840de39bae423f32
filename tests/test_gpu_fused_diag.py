"""ax_step_diag: the diagnostics of every `every`-th TimeStep taken inside the
step kernels (FP64 periodic Abelian-Higgs: the step before stores the
divergence of its links, the step at t computes energy, Gauss residual,
finiteness and edge phases from its own operands and link trigonometry)
against the separate diagnostics pass (ax_diagnostics) between plain steps.

Bar: the fields after the run bit for bit (the fused steps update exactly as
the plain ones), windings equal, energy and Gauss sums within 1e-13 relative
(the same per-site arithmetic; the partial sums may be grouped otherwise).
"""
import sys

import numpy as np
import pytest

import axions_amd as ax

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def separate(p, seed, t0, nsteps, every, xseg=None):
    out = []
    with ax.Lattice(p) as lat:
        if xseg is not None:
            lat.set_tile(xseg)
        lat.init_random_phase(seed)
        lat.step(t0)
        for _ in range(nsteps):
            if lat.timestep % every == 0:
                out.append(lat.diagnostics(winding=True))
            lat.step(1)
        return out, lat.get_state()


def fused(p, seed, t0, calls, every, xseg=None):
    out = []
    with ax.Lattice(p) as lat:
        if xseg is not None:
            lat.set_tile(xseg)
        lat.init_random_phase(seed)
        lat.step(t0)
        for n in calls:
            out += lat.step_diag(n, every, winding=True)
        return out, lat.get_state()


def same(a, b):
    assert len(a) == len(b) > 0
    for x, y in zip(a, b):
        assert x["timestep"] == y["timestep"]
        assert x["finite"] == y["finite"] == 1
        assert x["wound_plaquettes"] == y["wound_plaquettes"]
        assert rel(x["energy"], y["energy"]) < 1e-13, (x, y)
        assert rel(x["gauss_mean"], y["gauss_mean"]) < 1e-13, (x, y)
        assert rel(x["xi"], y["xi"]) < 1e-15


CASES = {
    # dx, dy, dz, g powers of two: the multiplied form of every division
    "pow2_plain": (lambda: ax.ah_params(24, ny=16, nz=64), 3, [10, 10, 7], 5, None),
    # dx a power of two, g not
    "pow2_g075": (lambda: ax.ah_params(20, ny=12, nz=128, g=0.75), 2, [13, 9], 4, None),
    # no powers of two, PRS friction and damping, short z tail, 8-plane segments
    "aniso_friction": (lambda: ax.ah_params(17, ny=10, nz=70, dx=0.7, dy=0.6, dz=0.55, dt=0.13, lam=1.5, eta=0.9,
                                            g=0.8, alpha=2.0, damped_nt=6, dampFac=0.4), 4, [9, 12], 3, 8),
    # the benchmark's couplings, first call starts off-cadence
    "c2_couplings_64": (lambda: ax.ah_params(64), 7, [20, 20, 13], 20, None),
    # whole z chunks (the LDS prefetch) with true divisions, friction and
    # damping, 3-plane segments
    "aniso_friction_128": (lambda: ax.ah_params(9, ny=10, nz=128, dx=0.7, dy=0.6, dz=0.55, dt=0.13, lam=1.5,
                                                eta=0.9, g=0.8, alpha=2.0, damped_nt=6, dampFac=0.4,
                                                theta_damp_form=1), 3, [10, 8], 3, 3),
}


@pytest.mark.parametrize("name", list(CASES))
def test_fused_diagnostics_equal_separate_pass(name):
    build, t0, calls, every, xseg = CASES[name]
    p = build()
    say(f"[fused] {name} {p.nx}x{p.ny}x{p.nz} from t={t0}, calls {calls}, every {every}")
    d_sep, st_sep = separate(p, 5, t0, sum(calls), every, xseg)
    d_fus, st_fus = fused(p, 5, t0, calls, every, xseg)
    assert np.array_equal(st_fus.view(np.uint64), st_sep.view(np.uint64))
    same(d_fus, d_sep)


def test_fused_diagnostics_fp32_and_boundaries_take_the_separate_pass():
    """Lattices outside the fused path (FP32, non-periodic) give exactly the
    separate pass's numbers."""
    for p in (ax.ah_params(16, ny=12, nz=64, precision=32),
              ax.ah_params(14, ny=12, nz=64, alpha=2.0, damped_nt=4, boundary=ax.AX_BC_FIXED)):
        d_sep, st_sep = separate(p, 9, 2, 12, 4)
        d_fus, st_fus = fused(p, 9, 2, [12], 4)
        assert np.array_equal(st_fus.view(np.uint64), st_sep.view(np.uint64))
        assert d_fus == d_sep


def test_fused_diagnostics_512():
    """The 512^3 run itself: 40 steps with diagnostics every 20."""
    p = ax.ah_params(512)
    say("[fused] 512^3 fp64, 40 steps, every 20")
    d_sep, st_sep = separate(p, 42, 20, 40, 20)
    d_fus, st_fus = fused(p, 42, 20, [20, 20], 20)
    assert np.array_equal(st_fus.view(np.uint64), st_sep.view(np.uint64))
    same(d_fus, d_sep)


@pytest.mark.parametrize("precision", [64, 32])
def test_step_diag_waits_for_its_last_steps(precision):
    """ax_step_diag returns only after every step it launched has run: the
    steps after its last output too, and when the range holds no output
    (FP64 takes the fused path, FP32 the separate pass).  Checked by time:
    the call lasts at least as long as the step kernels it launched."""
    import time
    p = ax.ah_params(256, precision=precision)
    say(f"[step_diag wait] 256^3 fp{precision}")
    with ax.Lattice(p) as lat:
        lat.init_random_phase(3)
        lat.step_diag(2, 20)          # buffers and code paths warm
        for n, every in ((30, 20), (12, 1000)):
            lat.kernel_timer(True)
            lat.kernel_time()
            t0 = time.perf_counter()
            lat.step_diag(n, every, winding=True)
            host_ms = (time.perf_counter() - t0) * 1e3
            kms, ksteps = lat.kernel_time()
            lat.kernel_timer(False)
            assert ksteps >= n - 2 and kms > 1.0, (n, every, kms, ksteps)
            assert host_ms >= 0.9 * kms, (n, every, host_ms, kms)
