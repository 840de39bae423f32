"""GPU parity of the FP32 variant and of the marching kernel's shape options.

FP32 has no reference counterpart; its definition is the oracle's
storage-type step orc_step_r (whose double instantiation is the pinned FP64
step, tests/test_fp32_oracle.py).  Bar: bit-exact fields and windings,
energy and Gauss sums (FP64 over the float state) within 1e-12.
"""
import sys

import numpy as np
import pytest

import axions_amd as ax
from axions_amd.abi import AX_BC_ABSORBING, AX_BC_FIXED, AX_BC_NEUMANN, AX_BC_PERIODIC
from oracle import oracle_py as orc

pytestmark = pytest.mark.gpu
NT = 16


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


F32 = {
    # marching kernel: full, partial (96 = 64 + 32, 100 = 64 + 36) and long rows
    "f32_tiled_64": (lambda: ax.ah_params(16, ny=12, nz=64, precision=32), 20, True),
    "f32_tiled_96": (lambda: ax.ah_params(10, ny=9, nz=96, precision=32), 16, True),
    "f32_tiled_100_damped": (lambda: ax.ah_params(7, ny=8, nz=100, alpha=3.0, damped_nt=6, dampFac=0.5,
                                                  precision=32), 14, True),
    "f32_tiled_aniso_form1": (lambda: ax.ah_params(9, ny=6, nz=128, dx=0.7, dy=0.6, dz=0.55, dt=0.13, lam=1.5,
                                                   eta=0.9, g=0.8, alpha=2.0, damped_nt=5, dampFac=0.3,
                                                   theta_damp_form=1, precision=32), 12, True),
    # short rows (one partial chunk) and fixed edges; the last field says
    # the case also runs on the site kernel (AX_FORCE_SITE_KERNEL)
    "f32_site_odd": (lambda: ax.ah_params(11, ny=13, nz=17, precision=32), 16, False),
    "f32_site_fixed": (lambda: ax.ah_params(10, boundary=AX_BC_FIXED, precision=32), 16, False),
    # non-periodic boundaries (boundary kernels + bulk update)
    "f32_neumann_zneu": (lambda: ax.ah_params(11, ny=9, nz=8, dx=0.6, dz=0.55, g=0.8, alpha=2.0, damped_nt=5,
                                              dampFac=0.4, boundary=AX_BC_NEUMANN, boundary_z=AX_BC_NEUMANN,
                                              precision=32), 16, False),
    "f32_absorbing_zper": (lambda: ax.ah_params(10, ny=12, nz=9, g=1.0, alpha=2.0, damped_nt=5,
                                                boundary=AX_BC_ABSORBING, precision=32), 16, False),
    "f32_absorbing_zfixed": (lambda: ax.ah_params(9, ny=8, nz=10, dy=0.45, g=0.7, boundary=AX_BC_ABSORBING,
                                                  boundary_z=AX_BC_FIXED, precision=32), 16, False),
    # the same boundaries on the marching kernel (DPP z shifts, edge masks)
    "f32_march_neumann_zneu": (lambda: ax.ah_params(66, ny=9, nz=64, g=0.8, alpha=2.0, damped_nt=5,
                                                    boundary=AX_BC_NEUMANN, boundary_z=AX_BC_NEUMANN,
                                                    precision=32), 12, True),
    "f32_march_absorbing_zfixed": (lambda: ax.ah_params(10, ny=9, nz=96, g=0.7, boundary=AX_BC_ABSORBING,
                                                        boundary_z=AX_BC_FIXED, precision=32), 12, True),
    "f32_march_fixed_zper": (lambda: ax.ah_params(12, ny=8, nz=100, boundary=AX_BC_FIXED,
                                                  boundary_z=AX_BC_PERIODIC, precision=32), 12, True),
}


CASES32 = [(n, False) for n in F32] + [(n, True) for n, v in F32.items() if not v[2]]


@pytest.mark.parametrize("name,site", CASES32)
def test_fp32_step_bitexact_vs_oracle(name, site, monkeypatch):
    """Every lattice marches (whatever its z tail or boundary); the short and
    non-periodic cases also run on the one-thread-per-site kernel."""
    build, nsteps, _ = F32[name]
    p = build()
    say(f"[fp32] {name} {p.nx}x{p.ny}x{p.nz} steps={nsteps} {'site' if site else 'march'}")
    if site:
        monkeypatch.setenv("AX_FORCE_SITE_KERNEL", "1")
    st = orc.init_random_phase(p, 42).astype(np.float32)
    with ax.Lattice(p) as lat:
        assert lat.tile()[1] == (not site)
        lat.init_random_phase(42)
        assert np.array_equal(lat.get_state(), st.astype(np.float64)), "initial condition differs"
        done = 0
        for chunk in (1, 2, nsteps - 3):
            lat.step(chunk)
            orc.step_r(p, st, done, chunk, NT)
            done += chunk
            g = lat.get_state()
            if not np.array_equal(g, st.astype(np.float64)):
                d = np.nonzero(g != st)[0]
                pytest.fail(f"{name}: {d.size} values differ after {done} steps, first flat index {d[0]}")


F64_SHAPES = {
    "f64_tiled_96": (lambda: ax.ah_params(9, ny=7, nz=96), 10),
    "f64_tiled_100_damped": (lambda: ax.ah_params(6, ny=7, nz=100, alpha=3.0, damped_nt=5, dampFac=0.5), 10),
}


@pytest.mark.parametrize("name", list(F64_SHAPES))
def test_fp64_partial_chunks_bitexact(name):
    build, nsteps = F64_SHAPES[name]
    p = build()
    say(f"[fp64-partial] {name} {p.nx}x{p.ny}x{p.nz}")
    st = orc.init_random_phase(p, 8)
    with ax.Lattice(p) as lat:
        assert lat.tile()[1]
        lat.init_random_phase(8)
        lat.step(nsteps)
        orc.step(p, st, 0, nsteps, 0, NT)
        assert np.array_equal(lat.get_state().view(np.uint64), st.view(np.uint64))


@pytest.mark.parametrize("precision", [64, 32])
def test_tile_shape_does_not_change_results(precision):
    p = ax.ah_params(13, ny=5, nz=64, precision=precision)
    say(f"[tiles] fp{precision}")
    outs = []
    for xseg in (1, 4, 5, 13, 0):
        with ax.Lattice(p) as lat:
            lat.set_tile(xseg)
            lat.init_random_phase(6)
            lat.step(9)
            outs.append(lat.get_state())
    for o in outs[1:]:
        assert np.array_equal(o.view(np.uint64), outs[0].view(np.uint64))


def test_fp32_moore_march_equals_site_kernel():
    """FP32 Moore has no oracle of its own: the marching kernel must equal
    the site kernel (the FP32 definition) bit for bit."""
    import os
    p = ax.moore_params(12, ny=10, nz=64, dt=0.2, precision=32)
    say("[fp32 moore] march vs site")
    outs = []
    for force in ("", "1"):
        if force:
            os.environ["AX_FORCE_SITE_KERNEL"] = "1"
        try:
            with ax.Lattice(p) as lat:
                assert lat.tile()[1] == (not force)
                lat.init_random_phase(21)
                lat.step(12)
                outs.append(lat.get_state())
        finally:
            os.environ.pop("AX_FORCE_SITE_KERNEL", None)
    assert np.array_equal(outs[0].view(np.uint64), outs[1].view(np.uint64))


def test_fp32_diagnostics_and_winding_vs_oracle():
    """FP64 reductions over the FP32 state: the oracle's FP64 diagnostics of
    the same (exactly widened) state."""
    p = ax.ah_params(12, ny=10, nz=64, precision=32)
    p64 = ax.ah_params(12, ny=10, nz=64)
    say("[fp32-diag]")
    st = orc.init_random_phase(p, 13).astype(np.float32)
    with ax.Lattice(p) as lat:
        lat.init_random_phase(13)
        for n in range(0, 16, 5):
            if n:
                lat.step(5)
                orc.step_r(p, st, n - 5, 5, NT)
            d = lat.diagnostics(winding=True)
            wide = st.astype(np.float64)
            o = orc.diag(p64, wide, n, 0)
            assert rel(d["energy"], o["energy"]) < 1e-12
            if o["gauss"] != 0:
                assert rel(d["gauss_mean"], o["gauss"]) < 1e-12
            w = orc.winding(p64, wide, n, 0)
            assert np.array_equal(lat.winding(), w)
            assert d["wound_plaquettes"] == int(np.abs(w.astype(np.int64)).sum())
