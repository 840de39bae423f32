"""The shared sin/cos/atan2 of include/axions/ax_trig.h (used by the HIP
kernels and by the oracle): accuracy against mpmath, IEEE special values,
and internal consistency.  The host build is exercised through the oracle
library; the device build is compared bit for bit in test_gpu_parity."""
import math

import mpmath as mp
import numpy as np
import pytest

from oracle import oracle_py as orc

mp.mp.dps = 40


def _ulp_err(vals, fn, args):
    errs = []
    for v, a in zip(vals, args):
        ex = fn(*[mp.mpf(float(x)) for x in a])
        if ex == 0:
            errs.append(0.0 if v == 0 else np.inf)
            continue
        errs.append(float(abs(mp.mpf(float(v)) - ex) / math.ulp(float(ex))))
    return np.array(errs)


@pytest.mark.parametrize("lo,hi", [(0.0, 1.0), (-10.0, 10.0), (-1e3, 1e3), (-2.0 ** 20, 2.0 ** 20)])
def test_sin_cos_accuracy(lo, hi):
    x = np.random.default_rng(11).uniform(lo, hi, 3000)
    es = _ulp_err(orc.trig(0, x), mp.sin, [(v,) for v in x])
    ec = _ulp_err(orc.trig(1, x), mp.cos, [(v,) for v in x])
    assert es.max() < 0.9 and ec.max() < 0.9, (es.max(), ec.max())


def test_sin_cos_bounded_below_2_pow_50():
    """The premise of the Moore march's division band: for every argument
    below 2^50 in magnitude (q*theta with in-band theta <= 2^40 and |q| <=
    2^8, or a plaquette angle) the reduction leaves |r| <~ pi/4 and sin/cos
    stay within [-1, 1]; far beyond it they do not (which is why the band
    stops at 2^40)."""
    rng = np.random.default_rng(21)
    x = np.concatenate([rng.uniform(-1, 1, 20000) * 2.0 ** rng.integers(0, 51, 20000),
                        np.array([2.0 ** 50, -2.0 ** 50, 2.0 ** 50 - 1, 2.0 ** 48 * 3])])
    s, c = orc.trig(0, x), orc.trig(1, x)
    assert np.max(np.abs(s)) <= 1.0 and np.max(np.abs(c)) <= 1.0
    assert np.max(np.abs(s * s + c * c - 1)) < 1e-13
    far = orc.trig(0, np.array([2.0 ** 60, 2.0 ** 200]))
    assert not np.all(np.abs(far) <= 1.0)


def test_sin_cos_special_values():
    x = np.array([0.0, -0.0, 1e-300, -1e-300, 5e-324, np.inf, -np.inf, np.nan])
    s, c = orc.trig(0, x), orc.trig(1, x)
    assert s[0] == 0 and not np.signbit(s[0]) and s[1] == 0 and np.signbit(s[1])
    assert c[0] == 1.0 and c[1] == 1.0
    assert s[2] == 1e-300 and s[3] == -1e-300 and s[4] == 5e-324
    assert np.all(np.isnan(s[5:])) and np.all(np.isnan(c[5:]))
    # sin(pi/2) rounds to 1, cos(pi) to -1
    assert orc.trig(0, np.array([math.pi / 2]))[0] == 1.0
    assert orc.trig(1, np.array([math.pi]))[0] == -1.0


def test_sin_cos_symmetry_and_identity():
    x = np.random.default_rng(3).uniform(-50, 50, 20000)
    s, c = orc.trig(0, x), orc.trig(1, x)
    assert np.array_equal(orc.trig(0, -x), -s)          # odd
    assert np.array_equal(orc.trig(1, -x), c)           # even
    assert np.max(np.abs(s * s + c * c - 1)) < 4e-16


def test_atan2():
    rng = np.random.default_rng(5)
    y, x = rng.normal(size=3000), rng.normal(size=3000)
    e = _ulp_err(orc.trig(2, y, x), mp.atan2, list(zip(y, x)))
    assert e.max() < 2.5, e.max()
    pairs = [(0.0, 1.0), (-0.0, 1.0), (0.0, -1.0), (-0.0, -1.0), (1.0, 0.0), (-1.0, 0.0),
             (0.0, 0.0), (-0.0, -0.0), (1.0, 1.0), (-1.0, -1.0), (3.0, -4.0)]
    for a, b in pairs:
        got = orc.trig(2, np.array([a]), np.array([b]))[0]
        ref = math.atan2(a, b)
        assert got == ref and np.signbit(got) == np.signbit(ref), (a, b, got, ref)


def test_sincosf_accuracy():
    x = np.random.default_rng(9).uniform(-100, 100, 3000).astype(np.float32).astype(np.float64)
    for fn, f in ((3, mp.sin), (4, mp.cos)):
        got = orc.trig(fn, x)
        ex = np.array([float(f(mp.mpf(float(v)))) for v in x])
        ulp = np.spacing(np.abs(ex).astype(np.float32)).astype(np.float64)
        assert np.max(np.abs(got - ex) / ulp) < 1.6


def test_sin_small_form_is_bit_identical():
    """The k = 0 short form used by the tiled kernel equals ax_sin bit for bit
    on its whole domain |RN(x*2/pi)| <= 0.5, including the boundary and zeros."""
    rng = np.random.default_rng(17)
    lim = 0.5 * math.pi / 2
    x = np.concatenate([rng.uniform(-lim, lim, 400000),
                        rng.standard_normal(20000) * 1e-3,
                        10.0 ** rng.uniform(-320, 0, 20000) * rng.choice([-1, 1], 20000),
                        np.array([0.0, -0.0, 5e-324, -5e-324, 1e-200])])
    edge = [lim]
    lo = hi = lim
    for _ in range(200):     # the doubles on either side of the k = 0 boundary
        lo, hi = np.nextafter(lo, 0), np.nextafter(hi, 2)
        edge += [lo, hi]
    edge = np.array(edge)
    x = np.concatenate([x, edge, -edge])
    small = orc.trig(5, x)
    ok = ~np.isnan(small)
    assert ok.sum() > 400000
    full = orc.trig(0, x[ok])
    assert np.array_equal(small[ok].view(np.uint64), full.view(np.uint64))
    # the predicate is exactly "reduction picks k = 0"
    kpick = np.rint(x * (2 / math.pi)) + 0.0
    assert np.array_equal(ok, kpick == 0)


def test_sinf_small_form_is_bit_identical():
    """Float counterpart: ax_sinf_small equals the sine of ax_sincosf bit for
    bit wherever the float reduction picks k = 0."""
    rng = np.random.default_rng(23)
    lim = np.float32(0.5 * math.pi / 2)
    x = np.concatenate([rng.uniform(-lim, lim, 400000).astype(np.float32),
                        (rng.standard_normal(20000) * 1e-3).astype(np.float32),
                        (10.0 ** rng.uniform(-44, 0, 20000) * rng.choice([-1, 1], 20000)).astype(np.float32),
                        np.array([0.0, -0.0, 1e-45, -1e-45], dtype=np.float32)])
    edge = [lim]
    lo = hi = lim
    for _ in range(200):
        lo, hi = np.nextafter(lo, np.float32(0)), np.nextafter(hi, np.float32(2))
        edge += [lo, hi]
    edge = np.array(edge, dtype=np.float32)
    x = np.concatenate([x, edge, -edge]).astype(np.float64)
    small = orc.trig(6, x)
    ok = ~np.isnan(small)
    assert ok.sum() > 400000
    full = orc.trig(3, x[ok])
    assert np.array_equal(small[ok].view(np.uint64), full.view(np.uint64))
    kpick = np.rint(x.astype(np.float32) * np.float32(2 / math.pi)) + np.float32(0.0)
    assert np.array_equal(ok, kpick == 0)


def _small_domain_points(lim, rng, n=400000):
    x = np.concatenate([rng.uniform(-lim, lim, n), rng.standard_normal(20000) * 1e-3,
                        10.0 ** rng.uniform(-320, 0, 20000) * rng.choice([-1, 1], 20000),
                        np.array([0.0, -0.0, 5e-324, -5e-324, 1e-200])])
    edge = [lim]
    lo = hi = lim
    for _ in range(200):
        lo, hi = np.nextafter(lo, 0), np.nextafter(hi, 2)
        edge += [lo, hi]
    edge = np.array(edge)
    return np.concatenate([x, edge, -edge])


def test_cos_small_form_is_bit_identical():
    """The cos member of the k = 0 sincos (the link fast path of the marching
    kernel) equals ax_cos bit for bit wherever ax_sin_is_small holds."""
    x = _small_domain_points(0.5 * math.pi / 2, np.random.default_rng(23))
    small = orc.trig(7, x)
    ok = ~np.isnan(small)
    assert ok.sum() > 400000
    assert np.array_equal(small[ok].view(np.uint64), orc.trig(1, x[ok]).view(np.uint64))
    assert np.array_equal(ok, (np.rint(x * (2 / math.pi)) + 0.0) == 0)


def test_cosf_small_form_is_bit_identical():
    """Same for binary32: ax_sincosf_small's cos equals ax_sincosf's cos."""
    x = _small_domain_points(0.5 * math.pi / 2, np.random.default_rng(29)).astype(np.float32).astype(np.float64)
    small = orc.trig(8, x)
    ok = ~np.isnan(small)
    assert ok.sum() > 400000
    assert np.array_equal(small[ok], orc.trig(4, x[ok]))
