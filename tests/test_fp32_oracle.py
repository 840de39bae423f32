"""The storage-type oracle step (orc_step_r) that defines the FP32 variant.

Its double instantiation must be the reference-pinned FP64 step bit for bit
(which pins the operand order the float instantiation shares); the float
instantiation is then checked for the invariants the mathematics fixes and
for closeness to FP64 over a short horizon.
"""
import numpy as np
import pytest

from axions_amd.abi import AX_BC_FIXED, ah_params
from oracle import oracle_py as orc

CONFIGS = {
    "plain": dict(n=10),
    "damped_form0": dict(n=9, alpha=3.0, damped_nt=4, dampFac=0.5, theta_damp_form=0),
    "damped_form1": dict(n=9, alpha=3.0, damped_nt=4, dampFac=0.5, theta_damp_form=1),
    "aniso": dict(n=8, ny=7, nz=11, dx=0.7, dy=0.6, dz=0.55, dt=0.13, lam=1.5, eta=0.9, g=0.8),
    "g0": dict(n=8, g=0.0, lam=1.0, dt=0.05),
    "fixed": dict(n=9, boundary=AX_BC_FIXED),
}


def params(cfg, precision):
    c = dict(cfg)
    n = c.pop("n")
    return ah_params(n, precision=precision, **c)


def test_steps_cover_precisions():
    p = ah_params(6, precision=32)
    st = orc.init_random_phase(p, 1).astype(np.float32)
    orc.step_r(p, st, 0, 1)
    with pytest.raises(AssertionError):
        orc.step_r(p, st.astype(np.float64), 0, 1)


@pytest.mark.parametrize("name", list(CONFIGS))
def test_double_instantiation_is_the_fp64_step(name):
    p = params(CONFIGS[name], 64)
    a = orc.init_random_phase(p, 5)
    rng = np.random.default_rng(2)
    s = a.reshape(5, 2, -1)
    s[2:] += rng.uniform(-0.3, 0.3, s[2:].shape)      # non-zero links and momenta
    b = a.copy()
    orc.step(p, a, 0, 7, 0, 2)
    orc.step_r(p, b, 0, 7, 2)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_float_vacuum_is_static():
    p = ah_params(8, precision=32)
    st = np.zeros(5 * 2 * 8 ** 3, dtype=np.float32)
    st.reshape(5, 2, -1)[0] = 1.0                    # phi = eta = 1, links 0
    ref = st.copy()
    orc.step_r(p, st, 0, 5)
    assert np.array_equal(st, ref)


def test_float_translation_invariance():
    p = ah_params(8, ny=6, nz=10, precision=32)
    st = orc.init_random_phase(p, 9).astype(np.float32)
    s = st.reshape(5, 2, 8, 6, 10)
    s[2:] += np.random.default_rng(3).uniform(-0.4, 0.4, s[2:].shape).astype(np.float32)
    sh = np.ascontiguousarray(np.roll(s, (3, -2, 5), axis=(2, 3, 4))).ravel()
    orc.step_r(p, st, 0, 4)
    orc.step_r(p, sh, 0, 4)
    back = np.roll(sh.reshape(5, 2, 8, 6, 10), (-3, 2, -5), axis=(2, 3, 4))
    assert np.array_equal(back.ravel(), st)


def test_float_tracks_fp64_over_a_short_horizon():
    p64 = ah_params(10)
    p32 = ah_params(10, precision=32)
    a = orc.init_random_phase(p64, 4)
    b = a.astype(np.float32)
    orc.step_r(p64, a, 0, 10, 2)
    orc.step_r(p32, b, 0, 10, 2)
    err = np.max(np.abs(b.astype(np.float64) - a)) / np.max(np.abs(a))
    assert 1e-9 < err < 1e-4, err      # really float arithmetic, and close
