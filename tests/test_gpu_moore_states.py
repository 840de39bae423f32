"""Moore FP64 marching step with a non-power-of-two dx^2 on states with an
exact vacuum block (Laplacian numerators +-0, plaquette sines 0) and a block
of tiny fields (numerators near 1e-280, products underflowing, outside the
band that allows the march's short divisions): the marching kernel must
equal the site kernel bit for bit, as it does on random states; and the
march's two division paths must agree."""
import os
import sys

import numpy as np
import pytest

import axions_amd as ax

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(p, st, nsteps, site):
    if site:
        os.environ["AX_FORCE_SITE_KERNEL"] = "1"
    try:
        with ax.Lattice(p) as lat:
            assert lat.tile()[1] == (not site)
            lat.set_state(st, timestep=0)
            lat.step(nsteps)
            return lat.get_state()
    finally:
        os.environ.pop("AX_FORCE_SITE_KERNEL", None)


@pytest.mark.parametrize("case", ["vacuum_block", "tiny_block", "random"])
def test_moore_march_equals_site_kernel_on_special_states(case):
    p = ax.moore_params(12, ny=10, nz=64, dt=0.2)   # dx = 0.7: dx^2 is not a power of two
    say(f"[moore div] {case}")
    with ax.Lattice(p) as lat:
        lat.init_random_phase(5)
        lat.step(3)
        st = lat.get_state()
    f = st.reshape(7, 2, p.nx, p.ny, p.nz)
    if case == "vacuum_block":
        # phi_1 = phi_2 = (1, 0), theta = 0 on both levels over a block that
        # covers whole z chunks' interiors: Laplacian numerators there are 0
        f[:, :, 2:9, 2:8, 5:40] = 0.0
        f[0, :, 2:9, 2:8, 5:40] = 1.0
        f[2, :, 2:9, 2:8, 5:40] = 1.0
    elif case == "tiny_block":
        # fields of order 1e-280 (and theta 0) over a block
        rng = np.random.default_rng(7)
        blk = f[:, :, 3:8, 2:7, 10:30]
        blk[:4] = rng.standard_normal(blk[:4].shape) * 1e-280
        blk[4:] = 0.0
    out_m = run(p, st, 4, site=False)
    out_s = run(p, st, 4, site=True)
    assert np.array_equal(out_m.view(np.uint64), out_s.view(np.uint64))


def test_moore_band_divisions_equal_plain_divisions():
    """The FP64 march divides by dx^2 through the once-per-wave reciprocal
    when the previous step's band flag allows it; with AX_MOORE_PLAIN_DIV it
    always divides.  Both must give the same bits over a run that restarts
    from the host mid-way (the flags restart with it)."""
    p = ax.moore_params(16, ny=12, nz=64, dt=0.2)
    say("[moore div] band vs plain")
    outs = []
    for plain in ("", "1"):
        if plain:
            os.environ["AX_MOORE_PLAIN_DIV"] = "1"
        try:
            with ax.Lattice(p) as lat:
                lat.init_random_phase(9)
                lat.step(17)
                mid = lat.get_state()
                lat.set_state(mid, timestep=17)
                lat.step(14)
                outs.append(lat.get_state())
        finally:
            os.environ.pop("AX_MOORE_PLAIN_DIV", None)
    assert np.array_equal(outs[0].view(np.uint64), outs[1].view(np.uint64))


def same_bits_or_both_nan(a, b):
    """Bit equality, except that two NaNs count as equal (their payloads come
    from whichever operand an instruction propagates)."""
    na, nb = np.isnan(a), np.isnan(b)
    return np.array_equal(na, nb) and np.array_equal(a[~na].view(np.uint64), b[~nb].view(np.uint64))


def edge_state(p, case):
    with ax.Lattice(p) as lat:
        lat.init_random_phase(11)
        st = lat.get_state()
    f = st.reshape(7, 2, p.nx, p.ny, p.nz)
    rng = np.random.default_rng(13)
    blk = (slice(None), slice(None), slice(2, 9), slice(1, 7), slice(3, 50))
    shape = f[blk].shape
    sign = np.where(rng.random(shape) < 0.5, -1.0, 1.0)
    if case == "low_edge":
        # scalars just above 2^-250 (inside the band), links 0: the first
        # step writes values near the lower edge, the second divides them
        b = f[blk]
        b[:4] = sign[:4] * (2.0 ** -250) * rng.uniform(1.0, 1.5, shape[1:])
        b[4:] = 0.0
    elif case == "high_edge_links":
        # link angles near 2^39: after one step still inside the band, so
        # the second step runs the short divisions with trig arguments up to
        # q * 2^40 and plaquette angles built from such values
        b = f[blk]
        base = sign[4:] * (2.0 ** 39) * rng.uniform(1.0, 1.9, shape[1:])
        b[4:, 0] = base[:, 0]
        b[4:, 1] = base[:, 0] + rng.uniform(-0.1, 0.1, shape[2:])   # small theta-dot
    elif case == "mixed_zero_and_edges":
        # signed zeros next to both band edges in one wave
        b = f[blk]
        choice = rng.integers(0, 4, shape)
        vals = np.select([choice == 0, choice == 1, choice == 2],
                         [0.0 * sign, sign * 2.0 ** -249, sign * 2.0 ** 39], sign * rng.uniform(0.5, 1.5, shape))
        b[:] = vals
    elif case == "huge_links":
        # link angles between 2^55 and 2^200: outside the band, every step
        # must take the plain divisions (sines and cosines are no longer
        # bounded there, so results may overflow)
        b = f[blk]
        b[4:] = sign[4:] * 2.0 ** rng.uniform(55, 200, shape[1:])
    return st


@pytest.mark.parametrize("case", ["low_edge", "high_edge_links", "mixed_zero_and_edges", "huge_links"])
def test_moore_band_edges_equal_plain_and_site(case):
    """Periodic FP64 Moore states at the edges of the division band (and far
    beyond it): the band-guarded march, the march with plain divisions and
    the site kernel give the same bits over several steps."""
    p = ax.moore_params(12, ny=10, nz=64, dt=0.2)   # dx = 0.7: dx^2 not a power of two
    say(f"[moore band] {case}")
    st = edge_state(p, case)
    out_band = run(p, st, 5, site=False)
    os.environ["AX_MOORE_PLAIN_DIV"] = "1"
    try:
        out_plain = run(p, st, 5, site=False)
    finally:
        os.environ.pop("AX_MOORE_PLAIN_DIV", None)
    out_site = run(p, st, 5, site=True)
    assert same_bits_or_both_nan(out_band, out_plain)
    assert same_bits_or_both_nan(out_band, out_site)
