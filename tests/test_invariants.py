"""Invariants and closed forms that pin the parts of the oracle no reference
output covers (the per-plaquette winding has no reference counterpart) and
cross-check the rest: vacuum, lattice translation, gauge invariance,
straight-string winding, cube sum rule, linear dispersion, energy
conservation.  CPU only (the oracle).  On the GPU the windings are compared
with the oracle's bit for bit (test_gpu_parity.py) and the cube sum rule is
checked on the 512^3 lattice (test_gpu_fullsize.py)."""
import math

import numpy as np
import pytest

import axions_amd as ax
from oracle import oracle_py as orc


def state_from_levels(p, cur, prv):
    nc = cur.shape[0]
    st = np.empty((nc, 2) + cur.shape[1:])
    st[:, 1] = cur      # TimeStep 0 reads level 1 as current
    st[:, 0] = prv
    return st.ravel().copy()


def test_vacuum_is_static():
    p = ax.ah_params(8, alpha=3.0, damped_nt=5)
    cur = np.zeros((5, 8, 8, 8))
    cur[0] = p.eta
    st = state_from_levels(p, cur, cur)
    st0 = st.copy()
    orc.step(p, st, 0, 12, 0, 1)
    assert np.array_equal(st, st0)
    d = orc.diag(p, st, 12, 0)
    assert d["energy"] == 0.0 and d["gauss"] == 0.0


@pytest.mark.parametrize("model", ["ah", "moore"])
def test_translation_commutes_with_step(model):
    p = ax.ah_params(10, ny=9, nz=12, alpha=1.0, damped_nt=3) if model == "ah" else ax.moore_params(10, ny=9, nz=12)
    st = orc.init_random_phase(p, 5, 0)
    orc.step(p, st, 0, 3, 0, 1)                  # gauge field no longer zero
    shift = (3, -2, 5)
    nc = ax.ncomp(p)
    rolled = np.roll(st.reshape(nc, 2, p.nx, p.ny, p.nz), shift, axis=(2, 3, 4)).ravel().copy()
    orc.step(p, st, 3, 6, 0, 1)
    orc.step(p, rolled, 3, 6, 0, 1)
    expect = np.roll(st.reshape(nc, 2, p.nx, p.ny, p.nz), shift, axis=(2, 3, 4)).ravel()
    assert np.array_equal(rolled, expect)
    w = orc.winding(p, st, 9, 0).reshape(3, p.nx, p.ny, p.nz)
    wr = orc.winding(p, rolled, 9, 0).reshape(3, p.nx, p.ny, p.nz)
    assert np.array_equal(np.roll(w, shift, axis=(1, 2, 3)), wr)


def straight_string(p, xc, yc, width=1.0, eta=1.0):
    x = (np.arange(p.nx) * p.dx)[:, None, None]
    y = (np.arange(p.ny) * p.dy)[None, :, None]
    r = np.sqrt((x - xc) ** 2 + (y - yc) ** 2)
    ang = np.arctan2(y - yc, x - xc)
    f = eta * np.tanh(r / width)
    cur = np.zeros((5, p.nx, p.ny, p.nz))
    cur[0] = np.broadcast_to(f * np.cos(ang), (p.nx, p.ny, p.nz))
    cur[1] = np.broadcast_to(f * np.sin(ang), (p.nx, p.ny, p.nz))
    return cur


def test_straight_string_winding():
    p = ax.ah_params(16, nz=6)
    ic, jc = 7, 8                       # core at the centre of plaquette (7..8, 8..9)
    cur = straight_string(p, (ic + 0.5) * p.dx, (jc + 0.5) * p.dy)
    st = state_from_levels(p, cur, cur)
    w = orc.winding(p, st, 0, 0).reshape(3, p.nx, p.ny, p.nz).astype(int)
    for k in range(p.nz):
        assert w[0, ic, jc, k] == 1
        inner = w[0, 2:14, 2:14, k].copy()
        inner[ic - 2, jc - 2] = 0
        assert not inner.any()
        assert w[0, :, :, k].sum() == 0          # closed periodic layer: net flux zero
    assert not w[1:, 2:14, 2:14, :].any()        # no flux through xz, yz faces near the string
    # the reversed winding gives -1
    cur[1] *= -1
    st = state_from_levels(p, cur, cur)
    w2 = orc.winding(p, st, 0, 0).reshape(3, p.nx, p.ny, p.nz)
    assert (w2[0, ic, jc, :] == -1).all()


def cube_sums(w):
    """Oriented winding sum over the 6 faces of every cube (0 when flux is conserved)."""
    wxy, wxz, wyz = w[0], w[1], w[2]
    return (np.roll(wxy, -1, 2) - wxy) - (np.roll(wxz, -1, 1) - wxz) + (np.roll(wyz, -1, 0) - wyz)


@pytest.mark.parametrize("steps", [0, 10, 40])
def test_cube_sum_rule(steps):
    p = ax.ah_params(12)
    st = orc.init_random_phase(p, 9, 0)
    orc.step(p, st, 0, steps, 0, 4)
    w = orc.winding(p, st, steps, 0).reshape(3, 12, 12, 12).astype(int)
    assert np.abs(w).sum() > 0
    assert not cube_sums(w).any()


def gauge_transform(p, st, n, alpha):
    """phi -> e^{i alpha} phi, theta_a(s) -> theta_a(s) + alpha(s+a) - alpha(s), both levels."""
    s = st.reshape(5, 2, p.nx, p.ny, p.nz).copy()
    c, sn = np.cos(alpha), np.sin(alpha)
    for lev in range(2):
        re, im = s[0, lev].copy(), s[1, lev].copy()
        s[0, lev] = c * re - sn * im
        s[1, lev] = sn * re + c * im
        for a in range(3):
            s[2 + a, lev] += np.roll(alpha, -1, axis=a) - alpha
    return s.ravel().copy()


def test_gauge_invariance():
    p = ax.ah_params(10)
    st = orc.init_random_phase(p, 21, 0)
    orc.step(p, st, 0, 5, 0, 2)
    alpha = np.random.default_rng(0).uniform(-0.5, 0.5, (10, 10, 10))
    gt = gauge_transform(p, st, 5, alpha)
    assert np.array_equal(orc.winding(p, st, 5, 0), orc.winding(p, gt, 5, 0))
    e0, e1 = orc.diag(p, st, 5, 0), orc.diag(p, gt, 5, 0)
    assert abs(e1["energy"] - e0["energy"]) < 1e-11 * abs(e0["energy"])
    assert abs(e1["gauss"] - e0["gauss"]) < 1e-9 * max(abs(e0["gauss"]), 1e-12)
    # covariance of the dynamics over a few steps (no chaos amplification yet)
    orc.step(p, st, 5, 3, 0, 2)
    orc.step(p, gt, 5, 3, 0, 2)
    back = gauge_transform(p, st, 8, alpha)
    assert np.max(np.abs(back - gt)) < 1e-10


def test_linear_dispersion():
    """Small Higgs-mode plane wave around the vacuum: the leapfrog lattice
    dispersion (4/dt^2) sin^2(w dt/2) = lambda eta^2 + (4/dx^2) sin^2(k dx/2)."""
    p = ax.ah_params(32, ny=3, nz=3, dx=0.5, dt=0.05, lam=2.0, eta=1.0, g=1.0)
    k = 2 * math.pi * 3 / (32 * p.dx)
    m2 = p.lambda_ * p.eta ** 2
    lat_k2 = (4 / p.dx ** 2) * math.sin(k * p.dx / 2) ** 2
    w = 2 / p.dt * math.asin(p.dt / 2 * math.sqrt(m2 + lat_k2))
    eps = 1e-7
    x = np.arange(32) * p.dx
    cur = np.zeros((5, 32, 3, 3))
    prv = np.zeros((5, 32, 3, 3))
    cur[0] = p.eta + eps * np.cos(k * x)[:, None, None]
    prv[0] = p.eta + eps * np.cos(k * x + w * p.dt)[:, None, None]   # one step in the past
    st = state_from_levels(p, cur, prv)
    nsteps = 200
    orc.step(p, st, 0, nsteps, 0, 1)
    now, _ = orc.levels(p, st, nsteps)
    expect = p.eta + eps * np.cos(k * x - w * p.dt * nsteps)
    assert np.max(np.abs(now[0, :, 0, 0] - expect)) < 2e-3 * eps


def test_energy_conservation_first_order_estimator():
    """Without friction the energy error shrinks linearly with dt: the
    reference's estimator (evolution.cpp:403-425) pairs the backward
    difference (phi^n - phi^{n-1})/dt with level-n gradients and potential,
    an O(dt) mismatch, while the leapfrog itself is second order."""
    drifts = []
    for dt in (0.04, 0.02):
        p = ax.ah_params(8, dt=dt)
        cur = straight_string(p, 3.7 * p.dx, 4.2 * p.dy)
        cur[0] += 0.05 * np.sin(2 * np.pi * np.arange(8) / 8)[None, None, :]
        st = state_from_levels(p, cur, cur)
        T = 1.0
        n = int(round(T / dt))
        es = []
        for i in range(n):
            es.append(orc.diag(p, st, i, 0)["energy"])
            orc.step(p, st, i, 1, 0, 1)
        es = np.array(es[1:])
        drifts.append(np.max(np.abs(es - es[0])) / es[0])
    assert drifts[0] < 0.1 and 1.6 < drifts[0] / drifts[1] < 2.6, drifts
