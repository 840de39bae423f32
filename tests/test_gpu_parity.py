"""GPU parity: the HIP library against the oracle (the CPU restatement pinned to
the reference in test_oracle_golden.py) and against the reference's own
golden fixtures, through the C ABI.

Bar: fields, winding numbers and crossing points bit-exact; energy and Gauss
sums (summed in a different order on the GPU) within 1e-12 relative.
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import axions_amd as ax
from axions_amd.abi import (AX_BC_ABSORBING, AX_BC_FIXED, AX_BC_NEUMANN, AX_BC_PERIODIC, AX_MODEL_MOORE,
                             AxParams)
from oracle import oracle_py as orc

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NT = max(1, min(16, os.cpu_count() or 1))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def bits_equal(a, b):
    return np.array_equal(np.asarray(a, dtype=np.float64).view(np.uint64),
                          np.asarray(b, dtype=np.float64).view(np.uint64))


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


CASES = {
    "ah_plain_16": (lambda: ax.ah_params(16), 40),
    "ah_g0_16": (lambda: ax.ah_params(16, g=0.0, lam=1.0, dt=0.05, alpha=3.0, damped_nt=20), 40),
    "ah_damp_form0_16": (lambda: ax.ah_params(16, alpha=3.0, damped_nt=15, dampFac=0.5, theta_damp_form=0), 30),
    "ah_damp_form1_16": (lambda: ax.ah_params(16, alpha=3.0, damped_nt=15, dampFac=0.5, theta_damp_form=1), 30),
    "ah_aniso": (lambda: ax.ah_params(12, ny=10, nz=14, dx=0.7, dy=0.6, dz=0.55, dt=0.13, lam=1.5, eta=0.9,
                                      g=0.8, alpha=2.0, damped_nt=10, dampFac=0.3), 25),
    "ah_odd_sizes": (lambda: ax.ah_params(17, ny=19, nz=70), 12),
    "ah_plain_64": (lambda: ax.ah_params(64), 20),
    # nz a multiple of 64 routes AH FP64 through the tiled z-row kernel
    "tiled_rect_damped": (lambda: ax.ah_params(40, ny=24, nz=128, alpha=3.0, damped_nt=10, dampFac=0.5), 16),
    "tiled_aniso_damped": (lambda: ax.ah_params(10, ny=12, nz=192, dx=0.7, dy=0.6, dz=0.55, dt=0.13, lam=1.5,
                                                 eta=0.9, g=0.8, alpha=2.0, damped_nt=8, dampFac=0.3,
                                                 theta_damp_form=1), 14),
    "tiled_aniso_plain": (lambda: ax.ah_params(70, ny=5, nz=64, dx=0.7, dy=0.6, dz=0.55, dt=0.13), 12),
    "tiled_512_row": (lambda: ax.ah_params(33, ny=4, nz=512), 6),
    "moore_periodic_12": (lambda: ax.moore_params(12), 25),
    # periodic Moore with nz >= 32 takes the marching kernel
    "moore_tiled_64": (lambda: ax.moore_params(20, ny=12, nz=64, dt=0.2), 12),
    "moore_tiled_96_aniso": (lambda: ax.moore_params(9, ny=8, nz=96, dx=0.6, dt=0.17), 10),
    "moore_fixed_12": (lambda: ax.moore_params(12, boundary=AX_BC_FIXED), 25),
    # non-periodic boundaries on the marching kernel (nz tail 0 or >= 32):
    # interior planes/rows/z sites stepped, faces by the boundary kernels
    "march_fixed_zper": (lambda: ax.ah_params(14, ny=11, nz=64, alpha=2.0, damped_nt=5, boundary=AX_BC_FIXED,
                                              boundary_z=AX_BC_PERIODIC), 10),
    "march_neumann_zneu": (lambda: ax.ah_params(70, ny=9, nz=64, g=0.8, alpha=2.0, damped_nt=5, dampFac=0.4,
                                                theta_damp_form=1, boundary=AX_BC_NEUMANN,
                                                boundary_z=AX_BC_NEUMANN), 10),
    "march_absorbing_zfixed": (lambda: ax.ah_params(13, ny=12, nz=96, dy=0.45, g=1.0, alpha=2.0, damped_nt=5,
                                                    boundary=AX_BC_ABSORBING, boundary_z=AX_BC_FIXED), 10),
    "march_absorbing_zper": (lambda: ax.ah_params(11, ny=10, nz=128, boundary=AX_BC_ABSORBING,
                                                  boundary_z=AX_BC_PERIODIC), 10),
    "moore_march_fixed": (lambda: ax.moore_params(12, ny=10, nz=64, dt=0.2, boundary=AX_BC_FIXED), 10),
    "moore_march_neumann": (lambda: ax.moore_params(11, ny=12, nz=96, dt=0.2, boundary=AX_BC_NEUMANN), 10),
}


# the one-thread-per-site kernel too (every lattice marches by default)
SITE_CASES = ["ah_plain_16", "ah_g0_16", "ah_damp_form0_16", "ah_damp_form1_16", "ah_aniso", "ah_odd_sizes",
              "moore_periodic_12", "moore_fixed_12", "march_neumann_zneu", "march_absorbing_zfixed"]


@pytest.mark.parametrize("name,site", [(n, False) for n in CASES] + [(n, True) for n in SITE_CASES])
def test_step_bitexact_vs_oracle(name, site, monkeypatch):
    build, nsteps = CASES[name]
    p = build()
    say(f"[parity] {name} {p.nx}x{p.ny}x{p.nz} steps={nsteps} {'site' if site else 'march'}")
    if site:
        monkeypatch.setenv("AX_FORCE_SITE_KERNEL", "1")
    st = orc.init_random_phase(p, 42, 0)
    with ax.Lattice(p) as lat:
        assert lat.tile()[1] == (not site)
        lat.init_random_phase(42, 0)
        assert bits_equal(lat.get_state(), st), "initial condition differs"
        done = 0
        for chunk in (1, 2, nsteps - 3):
            lat.step(chunk)
            orc.step(p, st, done, chunk, 0, NT)
            done += chunk
            g = lat.get_state()
            if not bits_equal(g, st):
                d = np.nonzero(g != st)[0]
                pytest.fail(f"{name}: {d.size} values differ after {done} steps, first flat index {d[0]}")


def test_tiled_general_sine_path():
    """Large plaquette angles force the general ax_sin path of the tiled kernel."""
    p = ax.ah_params(36, ny=6, nz=128)
    say("[parity] tiled, large plaquette angles")
    rng = np.random.default_rng(4)
    st = orc.init_random_phase(p, 5, 0).reshape(5, 2, -1)
    st[2:5] = rng.uniform(-3.0, 3.0, st[2:5].shape)
    st = st.ravel().copy()
    with ax.Lattice(p) as lat:
        lat.set_state(st)
        lat.step(5)
        orc.step(p, st, 0, 5, 0, NT)
        assert bits_equal(lat.get_state(), st)


@pytest.mark.parametrize("name", ["ah_damp_form1_16", "ah_aniso", "moore_periodic_12", "moore_fixed_12"])
def test_diagnostics_vs_oracle(name):
    build, nsteps = CASES[name]
    p = build()
    say(f"[diag] {name}")
    st = orc.init_random_phase(p, 7, 0)
    with ax.Lattice(p) as lat:
        lat.init_random_phase(7, 0)
        for n in range(nsteps):
            if n % 5 == 0:
                d = lat.diagnostics(winding=True)
                o = orc.diag(p, st, n, 0)
                assert d["timestep"] == n and d["finite"] == 1
                assert rel(d["energy"], o["energy"]) < 1e-12, (n, d["energy"], o["energy"])
                if o["gauss"] != 0:
                    assert rel(d["gauss_mean"], o["gauss"]) < 1e-12, (n, d["gauss_mean"], o["gauss"])
                else:
                    assert d["gauss_mean"] == 0
                w = orc.winding(p, st, n, 0)
                assert d["wound_plaquettes"] == int(np.abs(w.astype(np.int64)).sum())
            lat.step(1)
            orc.step(p, st, n, 1, 0, NT)


@pytest.mark.parametrize("name,field", [("ah_plain_16", 0), ("ah_aniso", 0), ("moore_periodic_12", 0),
                                        ("moore_periodic_12", 1)])
def test_winding_bitexact(name, field):
    build, nsteps = CASES[name]
    p = build()
    say(f"[winding] {name} field={field}")
    st = orc.init_random_phase(p, 3, 0)
    with ax.Lattice(p) as lat:
        lat.init_random_phase(3, 0)
        done = 0
        for target in (0, 7, 15):
            if target > done:
                lat.step(target - done)
                orc.step(p, st, done, target - done, 0, NT)
                done = target
            wg = lat.winding(field)
            wo = orc.winding(p, st, done, field)
            assert np.array_equal(wg, wo), f"{np.count_nonzero(wg != wo)} plaquettes differ at step {done}"


@pytest.mark.parametrize("name", ["ah_plain_16", "ah_aniso", "ah_odd_sizes"])
def test_crossings_bitexact(name):
    build, nsteps = CASES[name]
    p = build()
    say(f"[crossings] {name}")
    st = orc.init_random_phase(p, 11, 0)
    with ax.Lattice(p) as lat:
        lat.init_random_phase(11, 0)
        lat.step(6)
        orc.step(p, st, 0, 6, 0, NT)
        pg, cg = lat.crossings()
        po, co = orc.crossings(p, st, 6)
        assert cg == co and bits_equal(pg, po)
        # a short output buffer reports the full count and fills only cap points
        pc, cc = lat.crossings(cap=5)
        assert cc == co and bits_equal(pc, po[:5])


def test_nonfinite_is_flagged():
    p = ax.ah_params(12)
    say("[nonfinite]")
    with ax.Lattice(p) as lat:
        lat.init_random_phase(1)
        st = lat.get_state()
        st[12345 % st.size] = np.nan
        lat.set_state(st)
        d = lat.diagnostics(winding=False)
        assert d["finite"] == 0
        lat.init_random_phase(1)
        assert lat.diagnostics(winding=False)["finite"] == 1


GOLDEN_SHARED = [n for n in ("evo_ah_g0", "evo_ah_g1_shared", "evo_ah_aniso", "moore_shared", "evo_c1_shared",
                              "evo_c2_1000_shared",
                              "evo_bc_neumann", "evo_bc_neumann_z", "evo_bc_absorbing", "evo_bc_absorbing_zfixed",
                              "evo_bc_fixed_zper", "evo_bc_fixed_zneu", "moore_neumann")
                 if os.path.exists(os.path.join(GOLDEN, n + ".json"))]


@pytest.mark.parametrize("name", GOLDEN_SHARED)
def test_gpu_matches_reference_fixture(name):
    """The GPU run reproduces the reference program's own outputs: its final
    field bit for bit, its energy and Gauss series to 1e-12, its crossing
    frames bit for bit.  evo_c2_1000_shared is the 512^3 run's couplings and
    cadence (plain, dx = 0.5, dt = 0.1, lambda = 2, g = eta = 1; energy and
    Gauss every step, crossing frames every 20) at 64^3 over 1000 TimeSteps."""
    fx = json.load(open(os.path.join(GOLDEN, name + ".json")))
    p = AxParams()
    for k, v in fx["params"].items():
        setattr(p, k, v)
    say(f"[golden] {name} {p.nx}x{p.ny}x{p.nz} nt={fx['nt']}")
    frames = {f["step"]: f for f in fx["frames"]}
    with ax.Lattice(p) as lat:
        lat.init_random_phase(fx["seed"], 0)
        for n in range(fx["nt"]):
            d = lat.diagnostics(winding=False)
            assert rel(d["energy"], fx["energy"][n]) < 1e-12, (n, d["energy"], fx["energy"][n])
            G = d["gauss_mean"]
            if p.model == AX_MODEL_MOORE and p.boundary != AX_BC_PERIODIC:
                G = G / ((p.nx - 2) * (p.ny - 2) * (p.nz - 2))
            if fx["gauss"][n] != 0:
                assert rel(G, fx["gauss"][n]) < 1e-12, (n, G, fx["gauss"][n])
            if n in frames:
                pts, c = lat.crossings()
                assert c == frames[n]["count"] and sha(pts) == frames[n]["sha256"], f"frame at step {n}"
            lat.step(1)
        cur, _ = lat.levels()
        assert sha(cur) == fx["final_sha256"]


TAILS = {
    # z tails of 1..31 sites: the last chunk marches with its other lanes idle
    "ah_tail1_fp64": (lambda: ax.ah_params(10, ny=9, nz=65), 12),
    "ah_tail2_fp32": (lambda: ax.ah_params(9, ny=10, nz=66, precision=32), 12),
    "ah_tail17_damped": (lambda: ax.ah_params(11, ny=8, nz=81, alpha=2.0, damped_nt=4, dx=0.6, dt=0.11), 12),
    "ah_tail31_fp32_aniso": (lambda: ax.ah_params(8, ny=12, nz=95, dx=0.7, dy=0.6, dz=0.55, dt=0.13,
                                                  precision=32), 12),
    "ah_nz17": (lambda: ax.ah_params(12, ny=10, nz=17), 12),
    "moore_tail5": (lambda: ax.moore_params(9, ny=8, nz=69, dt=0.2), 10),
    "moore_tail20_fp32": (lambda: ax.moore_params(8, ny=9, nz=20, dt=0.2, precision=32), 10),
}


@pytest.mark.parametrize("name", list(TAILS))
def test_short_z_tails_march_bitexact_vs_oracle(name, monkeypatch):
    """The march against the oracle (FP32 Moore, which the oracle does not
    restate, against the site kernel that defines it)."""
    build, nsteps = TAILS[name]
    p = build()
    say(f"[tails] {name} {p.nx}x{p.ny}x{p.nz} fp{p.precision}")
    fp32 = p.precision == 32
    st = orc.init_random_phase(p, 42, 0)
    if fp32:
        st = st.astype(np.float32)
    with ax.Lattice(p) as lat:
        assert lat.tile()[1], "every z tail marches"
        lat.init_random_phase(42, 0)
        lat.step(nsteps)
        if fp32 and p.model == AX_MODEL_MOORE:
            g = lat.get_state()
            monkeypatch.setenv("AX_FORCE_SITE_KERNEL", "1")
            with ax.Lattice(p) as site:
                assert not site.tile()[1]
                site.init_random_phase(42, 0)
                site.step(nsteps)
                assert bits_equal(g, site.get_state())
        elif fp32:
            orc.step_r(p, st, 0, nsteps, NT)
            g = lat.get_state().astype(np.float32)
            assert np.array_equal(g.view(np.uint32), st.view(np.uint32))
        else:
            orc.step(p, st, 0, nsteps, 0, NT)
            assert bits_equal(lat.get_state(), st)


@pytest.mark.parametrize("name", ["ah_plain_64", "tiled_aniso_damped", "ah_odd_sizes", "tiled_rect_damped",
                                  "moore_tiled_64", "moore_tiled_96_aniso"])
def test_big_plane_addressing_bitexact_vs_oracle(name):
    """The 64-bit plane addressing of lattices whose field planes reach 4 GiB
    (ah_march's, and moore_march's global path), forced on small lattices
    (AX_BIG_PLANES): the same bits as the oracle."""
    import os
    build, nsteps = CASES[name]
    p = build()
    say(f"[big planes] {name} {p.nx}x{p.ny}x{p.nz}")
    st = orc.init_random_phase(p, 42, 0)
    os.environ["AX_BIG_PLANES"] = "1"
    try:
        with ax.Lattice(p) as lat:
            lat.init_random_phase(42, 0)
            lat.step(nsteps)
            g = lat.get_state()
            d = lat.diagnostics(winding=True)
    finally:
        os.environ.pop("AX_BIG_PLANES", None)
    orc.step(p, st, 0, nsteps, 0, NT)
    assert bits_equal(g, st)
    o = orc.diag(p, st, nsteps, 0)
    assert rel(d["energy"], o["energy"]) < 1e-12


LONG = {
    # periodic Moore on the march (the band-flag division path after the
    # first step), 800 TimeSteps as the 256^3 run
    "moore_march_800": (lambda: ax.moore_params(24, ny=20, nz=64, dt=0.2), 800, (1, 100, 400, 800)),
    # Abelian-Higgs with non-power-of-two spacings (true divisions), 1000 TimeSteps
    "ah_aniso_march_1000": (lambda: ax.ah_params(24, ny=20, nz=64, dx=0.7, dy=0.6, dz=0.55, dt=0.13, lam=1.5,
                                                 eta=0.9, g=0.8), 1000, (1, 100, 500, 1000)),
}


@pytest.mark.parametrize("name", list(LONG))
def test_long_run_bitexact_vs_oracle(name):
    """Long trajectories, so that late-time branches (the small-angle sine
    vote once plaquette angles grow, the plain-mode overflow guard, the Moore
    band flags) are exercised: the GPU march equals the oracle bit for bit at
    every checkpoint."""
    build, nsteps, checks = LONG[name]
    p = build()
    say(f"[long] {name} {p.nx}x{p.ny}x{p.nz} {nsteps} steps")
    st = orc.init_random_phase(p, 42, 0)
    with ax.Lattice(p) as lat:
        assert lat.tile()[1]
        lat.init_random_phase(42, 0)
        done = 0
        for c in checks:
            lat.step(c - done)
            orc.step(p, st, done, c - done, 0, NT)
            done = c
            g = lat.get_state()
            assert np.all(np.isfinite(g))
            if not bits_equal(g, st):
                d = np.nonzero(g != st)[0]
                pytest.fail(f"{name}: {d.size} values differ after {done} steps, first flat index {d[0]}")


@pytest.mark.parametrize("precision", [32])
def test_fp32_long_run_bitexact_vs_oracle(precision):
    """The FP32 march over 1000 TimeSteps against the oracle's binary32
    instantiation of the same step."""
    p = ax.ah_params(24, ny=20, nz=64, precision=precision)
    say("[long] fp32 24x20x64 1000 steps")
    st = orc.init_random_phase(p, 42, 0).astype(np.float32)
    with ax.Lattice(p) as lat:
        lat.init_random_phase(42, 0)
        lat.step(1000)
        orc.step_r(p, st, 0, 1000, NT)
        g = lat.get_state().astype(np.float32)
        assert np.all(np.isfinite(g))
        assert np.array_equal(g.view(np.uint32), st.view(np.uint32))


@pytest.mark.parametrize("shape,precision,dx,g,energy_form", [
    ((24, 20, 64), 64, 0.5, 1.0, 0), ((13, 10, 96), 64, 0.5, 1.0, 0), ((9, 11, 17), 64, 0.5, 1.0, 0),
    ((16, 12, 100), 32, 0.5, 1.0, 0),
    ((20, 18, 64), 64, 0.7, 1.3, 0),     # spacings and coupling not powers of two: true divisions
    ((16, 16, 64), 64, 0.5, 0.5, 1),     # powers of two, the mpi_evolution.cpp energy form
    ((12, 14, 64), 32, 0.25, 2.0, 1),
    ((12, 10, 64), 64, 0.5, 0.75, 0)])   # dx a power of two, g not
def test_diag_march_equals_site_kernel(shape, precision, dx, g, energy_form):
    """The marching diagnostics (periodic Abelian-Higgs: energy/Gauss with the
    edge phases stored once per edge, windings from them) give the site
    kernels' windings bit for bit and their energy and Gauss sums to 1e-13
    (same per-site terms, another summation order), with and without the
    power-of-two reciprocals."""
    import os
    nx, ny, nz = shape
    p = ax.ah_params(nx, ny=ny, nz=nz, dx=dx, dt=0.1 * dx / 0.5, g=g, alpha=2.0, damped_nt=3, dampFac=0.5,
                     energy_form=energy_form, precision=precision)
    say(f"[diag-march] {shape} fp{precision} dx={dx} g={g} energy_form={energy_form}")
    with ax.Lattice(p) as lat:
        lat.init_random_phase(17)
        lat.step(7)
        d_m = lat.diagnostics(winding=True)
        w_m = lat.winding()
        os.environ["AX_DIAG_SITE_KERNEL"] = "1"
        try:
            d_s = lat.diagnostics(winding=True)
            w_s = lat.winding()
        finally:
            os.environ.pop("AX_DIAG_SITE_KERNEL", None)
    assert np.array_equal(w_m, w_s)
    assert d_m["wound_plaquettes"] == d_s["wound_plaquettes"] > 0
    assert rel(d_m["energy"], d_s["energy"]) < 1e-13
    assert rel(d_m["gauss_mean"], d_s["gauss_mean"]) < 1e-13
    assert d_m["finite"] == d_s["finite"] == 1


@pytest.mark.parametrize("shape,precision,bc,bcz", [((14, 11, 64), 64, AX_BC_FIXED, AX_BC_PERIODIC),
                                                    ((12, 10, 70), 64, AX_BC_ABSORBING, AX_BC_FIXED),
                                                    ((96, 9, 64), 32, AX_BC_NEUMANN, AX_BC_NEUMANN),
                                                    ((10, 12, 64), 64, AX_BC_PERIODIC, AX_BC_FIXED)])
def test_diag_march_bulk_region_equals_site_kernel(shape, precision, bc, bcz):
    """Non-periodic Abelian-Higgs lattices: the diagnostics march sums energy
    and Gauss residual over the bulk region only, as the site kernel does
    (1e-13: the same per-site terms, another summation order), counts
    non-finite values over every site, and the windings (the site kernel on
    these lattices) agree."""
    import os
    nx, ny, nz = shape
    p = ax.ah_params(nx, ny=ny, nz=nz, g=0.8, alpha=2.0, damped_nt=4, dampFac=0.5, boundary=bc, boundary_z=bcz,
                     precision=precision)
    say(f"[diag bulk] {shape} fp{precision} bc={bc} bcz={bcz}")
    with ax.Lattice(p) as lat:
        lat.init_random_phase(29)
        lat.step(8)
        d_m = lat.diagnostics(winding=True)
        os.environ["AX_DIAG_SITE_KERNEL"] = "1"
        try:
            d_s = lat.diagnostics(winding=True)
        finally:
            os.environ.pop("AX_DIAG_SITE_KERNEL", None)
    assert d_m["finite"] == d_s["finite"] == 1
    assert d_m["wound_plaquettes"] == d_s["wound_plaquettes"]
    assert rel(d_m["energy"], d_s["energy"]) < 1e-13, (d_m, d_s)
    assert rel(d_m["gauss_mean"], d_s["gauss_mean"]) < 1e-13, (d_m, d_s)


@pytest.mark.parametrize("shape,precision", [((12, 10, 70), 64), ((16, 16, 64), 32), ((9, 12, 130), 64),
                                             ((10, 8, 33), 32)])
def test_moore_winding_march_equals_site_kernel(shape, precision):
    """Periodic Moore windings of both scalars (charges q1 = 2, q2 = 1) from
    the marching edge-value passes equal the site kernel's bit for bit, and
    so do the diagnostics' wound counts (z tails, rows not a multiple of 8)."""
    import os
    nx, ny, nz = shape
    p = ax.moore_params(nx, ny=ny, nz=nz, dt=0.2, precision=precision)
    say(f"[moore winding march] {shape} fp{precision}")
    with ax.Lattice(p) as lat:
        lat.init_random_phase(23)
        lat.step(9)
        w_m = [lat.winding(f) for f in (0, 1)]
        d_m = lat.diagnostics(winding=True)
        os.environ["AX_DIAG_SITE_KERNEL"] = "1"
        try:
            w_s = [lat.winding(f) for f in (0, 1)]
            d_s = lat.diagnostics(winding=True)
        finally:
            os.environ.pop("AX_DIAG_SITE_KERNEL", None)
    for f in (0, 1):
        assert np.array_equal(w_m[f], w_s[f]), f
    assert np.abs(w_m[0].astype(np.int64)).sum() > 0
    assert d_m["wound_plaquettes"] == d_s["wound_plaquettes"]
    assert rel(d_m["energy"], d_s["energy"]) < 1e-13 and rel(d_m["gauss_mean"], d_s["gauss_mean"]) < 1e-13


@pytest.mark.parametrize("shape,precision,bc", [((12, 10, 70), 64, AX_BC_PERIODIC), ((16, 9, 64), 32, AX_BC_PERIODIC),
                                                ((14, 12, 66), 64, AX_BC_FIXED), ((12, 11, 40), 64, AX_BC_NEUMANN)])
def test_moore_diag_march_equals_site_kernel(shape, precision, bc):
    """Moore energy and Gauss law by the diagnostics march against the site
    kernel: the same per-site terms summed in another order (1e-13), over
    the bulk region where the lattice is not periodic, finiteness over every
    site."""
    import os
    nx, ny, nz = shape
    p = ax.moore_params(nx, ny=ny, nz=nz, dt=0.2, boundary=bc, precision=precision)
    say(f"[moore diag march] {shape} fp{precision} bc={bc}")
    with ax.Lattice(p) as lat:
        lat.init_random_phase(31)
        lat.step(7)
        d_m = lat.diagnostics(winding=False)
        os.environ["AX_DIAG_SITE_KERNEL"] = "1"
        try:
            d_s = lat.diagnostics(winding=False)
        finally:
            os.environ.pop("AX_DIAG_SITE_KERNEL", None)
    assert d_m["finite"] == d_s["finite"] == 1
    assert rel(d_m["energy"], d_s["energy"]) < 1e-13, (d_m, d_s)
    assert rel(d_m["gauss_mean"], d_s["gauss_mean"]) < 1e-13, (d_m, d_s)


@pytest.mark.parametrize("shape,precision", [((24, 16, 192), 64), ((20, 24, 64), 32), ((12, 10, 128), 64),
                                             ((16, 8, 100), 32)])
def test_work_orders_give_the_same_state(shape, precision):
    """The marching kernels' work order (natural, XCD-aware per chunk, XCD-aware
    whole-row bands) only permutes the work items: the state after a few
    steps is the same bits in every order, for the step and for the
    diagnostics and windings that march."""
    import os
    nx, ny, nz = shape
    say(f"[orders] {shape} fp{precision}")
    res = {}
    for order in ("", "natural", "chunks"):
        if order:
            os.environ["AX_TILE_ORDER"] = order
        try:
            p = ax.ah_params(nx, ny=ny, nz=nz, g=1.0, alpha=1.0, damped_nt=2, precision=precision)
            with ax.Lattice(p) as lat:
                lat.init_random_phase(5)
                lat.step(6)
                d = lat.diagnostics(winding=True)
                res[order] = (lat.get_state(), lat.winding(), d)
        finally:
            os.environ.pop("AX_TILE_ORDER", None)
    st0, w0, d0 = res[""]
    for order in ("natural", "chunks"):
        st, w, d = res[order]
        assert np.array_equal(st.view(np.uint64), st0.view(np.uint64)), order
        assert np.array_equal(w, w0), order
        assert d["wound_plaquettes"] == d0["wound_plaquettes"]
        assert rel(d["energy"], d0["energy"]) < 1e-13 and rel(d["gauss_mean"], d0["gauss_mean"]) < 1e-13


@pytest.mark.parametrize("shape,xseg", [((16, 12, 128), 1), ((16, 12, 128), 3), ((40, 8, 64), 32),
                                        ((9, 10, 192), 0)])
def test_prefetch_march_equals_plain_march(shape, xseg, monkeypatch):
    """The FP64 march that prefetches each plane's previous level and next
    plane into LDS (plain runs, power-of-two spacings, whole chunks) gives the
    bits of the march without it, for segments of 1, 3, 32 planes and whole
    rows."""
    nx, ny, nz = shape
    p = ax.ah_params(nx, ny=ny, nz=nz)
    say(f"[prefetch] {shape} xseg={xseg}")
    outs = []
    for off in (False, True):
        if off:
            monkeypatch.setenv("AX_NO_PREFETCH", "1")
        with ax.Lattice(p) as lat:
            lat.set_tile(xseg)
            lat.init_random_phase(8)
            lat.step(11)
            outs.append(lat.get_state())
    assert bits_equal(outs[0], outs[1])


PREFETCH_VARIANTS = {
    # PRS friction and the damping window (theta damping, both forms)
    "friction": lambda: ax.ah_params(16, ny=12, nz=128, alpha=3.0, damped_nt=6, dampFac=0.5),
    "friction_form1": lambda: ax.ah_params(16, ny=12, nz=128, alpha=3.0, damped_nt=6, theta_damp_form=1),
    # anisotropic spacings that are not powers of two: true divisions
    "aniso": lambda: ax.ah_params(16, ny=12, nz=128, dx=0.7, dy=0.6, dz=0.55, dt=0.1),
    # evolution.cpp's own schedule: absorbing x/y faces, g = 0, friction
    "absorbing": lambda: ax.ah_params(18, ny=14, nz=64, g=0.0, lam=1.0, dt=0.05, alpha=3.0, damped_nt=4,
                                      boundary=AX_BC_ABSORBING),
    # Neumann x/y with fixed z: the march skips boundary planes and rows, the
    # z-face lanes do not store
    "neumann_zfixed": lambda: ax.ah_params(14, ny=10, nz=128, boundary=AX_BC_NEUMANN, boundary_z=AX_BC_FIXED),
}


@pytest.mark.parametrize("name", sorted(PREFETCH_VARIANTS))
def test_prefetch_march_equals_plain_march_variants(name, monkeypatch):
    """Every FP64 variant of the march takes the LDS prefetch when its chunks
    are whole: friction products, true divisions and the non-periodic faces
    give the bits of the march without the prefetch (itself bit-exact against
    the oracle in the other tests), over segments of 3 planes and whole rows;
    the prefetched run also matches the oracle directly."""
    p = PREFETCH_VARIANTS[name]()
    outs = []
    for xseg in (3, 0):
        for off in (False, True):
            say(f"[prefetch] {name} xseg={xseg} prefetch={'off' if off else 'on'}")
            if off:
                monkeypatch.setenv("AX_NO_PREFETCH", "1")
            else:
                monkeypatch.delenv("AX_NO_PREFETCH", raising=False)
            with ax.Lattice(p) as lat:
                lat.set_tile(xseg)
                lat.init_random_phase(5)
                lat.step(9)
                outs.append(lat.get_state())
        assert bits_equal(outs[-2], outs[-1])
    st = orc.init_random_phase(p, 5, 0)
    orc.step(p, st, 0, 9, 0, NT)
    assert bits_equal(outs[0], st)
