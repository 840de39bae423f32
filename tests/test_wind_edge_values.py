"""The marching winding passes store per edge e_a(s) = delta_a(s) + theta_a(s)
rounded to FP32 and form W = rint((e_a(s) + e_b(s+a) - e_a(s+b) - e_b(s)) / 2pi)
(axions_amd/csrc/ax_wind.h), a regrouping of the per-plaquette sum of the
oracle's site winding.  The exact sum is 2 pi times an integer (the phases of
phi telescope around the plaquette), so an evaluation whose rounding stays
below pi gives the same integer.  Checked here with numpy's own
sin/cos/arctan2, so the check shares no code with the oracle or the GPU
(numpy's phases could take the other branch for an edge phase within
rounding of +-pi, which the GPU avoids by computing them with the oracle's
arithmetic; these states have none), against the oracle's windings: on
random-phase states evolved up to 300 steps, on anisotropic and FP32
states, on both Moore scalars (charges 2 and 1: e = delta + q theta), and
with link angles shifted by up to 2^16 pi, where FP32 rounds e to 2^-7 and
the integers still agree.  CPU only; the GPU passes are compared
with the oracle in test_gpu_parity.py, test_gpu_fp32.py and
test_gpu_pool.py."""
import numpy as np
import pytest

import axions_amd as ax
from oracle import oracle_py as orc


def edge_value_windings(p, st, n, field=0):
    """W of every xy, xz, yz plaquette from FP32 edge values, numpy trig;
    Moore: scalar `field` with its charge times the links."""
    nc = ax.ncomp(p)
    lev = st.reshape(nc, 2, p.nx, p.ny, p.nz)[:, (n + 1) % 2].astype(np.float64)
    moore = p.model == ax.AX_MODEL_MOORE
    c0, q = (2 * field, float(p.q1 if field == 0 else p.q2)) if moore else (0, 1.0)
    pr, pi_, th = lev[c0], lev[c0 + 1], q * lev[nc - 3:nc]
    e = []
    for a in range(3):
        nr, ni = np.roll(pr, -1, axis=a), np.roll(pi_, -1, axis=a)
        c, s = np.cos(th[a]), np.sin(th[a])
        wr, wi = c * nr + s * ni, c * ni - s * nr
        re, im = pr * wr + pi_ * wi, pr * wi - pi_ * wr
        e.append((np.arctan2(im, re) + th[a]).astype(np.float32).astype(np.float64))
    w = []
    for a, b in ((0, 1), (0, 2), (1, 2)):
        tot = e[a] + np.roll(e[b], -1, axis=a) - np.roll(e[a], -1, axis=b) - e[b]
        w.append(np.rint(tot * 0.15915494309189535).astype(np.int64))
    return np.stack(w)


def oracle_windings(p, st, n, field=0):
    return orc.winding(p, st, n, field).reshape(3, p.nx, p.ny, p.nz).astype(np.int64)


@pytest.mark.parametrize("field", [0, 1])
def test_edge_value_windings_moore(field):
    """Moore: both scalars (charges 2 and 1 times the one link field)."""
    p = ax.moore_params(12, ny=10, nz=14, dt=0.2)
    st = orc.init_random_phase(p, 13, 0)
    done, wound = 0, 0
    for n in (0, 12, 40):
        if n > done:
            orc.step(p, st, done, n - done, 0, 1)
            done = n
        w_o = oracle_windings(p, st, n, field)
        assert np.array_equal(edge_value_windings(p, st, n, field), w_o), (field, n)
        wound += int(np.abs(w_o).sum())
    assert wound > 0


@pytest.mark.parametrize("name,build,steps", [
    ("plain_16", lambda: ax.ah_params(16), (0, 40, 300)),
    ("aniso_damped", lambda: ax.ah_params(12, ny=10, nz=14, dx=0.7, dy=0.6, dz=0.55, dt=0.13, lam=1.5, eta=0.9,
                                          g=0.8, alpha=2.0, damped_nt=10, dampFac=0.3), (0, 25, 120)),
    ("fp32_16", lambda: ax.ah_params(16, ny=12, nz=10, precision=32), (0, 30, 150)),
])
def test_edge_value_windings_equal_oracle(name, build, steps):
    p = build()
    st = orc.init_random_phase(p, 11, 0)
    done, wound = 0, 0
    for n in steps:
        if n > done:
            orc.step(p, st, done, n - done, 0, 1)
            done = n
        w_o = oracle_windings(p, st, n)
        assert np.array_equal(edge_value_windings(p, st, n), w_o), (name, n)
        wound += int(np.abs(w_o).sum())
    assert wound > 0


def test_edge_value_windings_with_large_link_angles():
    """Link angles shifted by 2 pi k (|k| up to 2^15): the plaquette sums
    move by 2 pi times integers, both forms follow them, and FP32 edge
    values of magnitude ~2^17 still give the oracle's integers (as int8,
    the oracle's and the winding field's type)."""
    p = ax.ah_params(12, ny=10, nz=8)
    st = orc.init_random_phase(p, 4, 0)
    orc.step(p, st, 0, 60, 0, 1)
    n = 60
    v = st.reshape(5, 2, p.nx, p.ny, p.nz)
    cur = (n + 1) % 2
    rng = np.random.default_rng(3)
    for kmax in (1, 100, 1 << 15):
        shifted = v.copy()
        k = rng.integers(-kmax, kmax + 1, size=(3, p.nx, p.ny, p.nz))
        shifted[2:5, cur] += 2 * np.pi * k
        s2 = shifted.ravel().copy()
        w = edge_value_windings(p, s2, n)
        # the oracle returns the windings as int8 (as the GPU's field): compare mod 256
        assert np.array_equal(w.astype(np.int8), oracle_windings(p, s2, n).astype(np.int8)), kmax
        if kmax > 1:
            assert np.abs(w).max() > 127
