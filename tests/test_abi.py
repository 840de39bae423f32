"""The C ABI library loads without a GPU and exports every entry point that
include/axions/*.h declares; host-only entry points behave."""
import ctypes
import glob
import os
import re

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    """ax_* functions declared by the C-ABI headers (ax_trig.h is a
    header-only math library of inline AX_HD functions, not exported)."""
    names = set()
    for h in glob.glob(os.path.join(REPO, "include", "axions", "*.h")):
        text = open(h).read()
        if "AX_HD" in text:
            continue
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:[A-Za-z_][\w \*]*?)\b(ax_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_exports_every_declared_symbol():
    import axions_amd
    lib = ctypes.CDLL(os.path.join(REPO, "axions_amd", "libaxions.so"))
    names = declared_symbols()
    assert len(names) >= 18, names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert axions_amd.lib is not None


def test_param_validation_without_gpu():
    import axions_amd as ax
    p = ax.ah_params(2)
    n = ctypes.c_size_t()
    rc = ax.lib.ax_state_size(ctypes.byref(p), None, ctypes.byref(n))
    assert rc == ax.abi.AX_EINVAL
    assert b"3 sites" in ax.lib.ax_last_error()
    p = ax.ah_params(8, ny=9, nz=10)
    assert ax.lib.ax_state_size(ctypes.byref(p), None, ctypes.byref(n)) == 0
    assert n.value == 5 * 2 * 8 * 9 * 10
    p = ax.moore_params(8)
    assert ax.lib.ax_state_size(ctypes.byref(p), None, ctypes.byref(n)) == 0
    assert n.value == 7 * 2 * 8 ** 3
    p.precision = 16
    assert ax.lib.ax_state_size(ctypes.byref(p), None, ctypes.byref(n)) == ax.abi.AX_EINVAL


def test_default_params_are_the_plain_set():
    import axions_amd as ax
    p = ax.AxParams()
    ax.lib.ax_default_params(ctypes.byref(p))
    assert (p.nx, p.dx, p.dt, p.lambda_, p.g, p.eta, p.alpha) == (64, 0.5, 0.1, 2.0, 1.0, 1.0, 0.0)
    assert p.precision == 64


def test_strerror():
    import axions_amd as ax
    assert ax.lib.ax_strerror(-3) == b"HBM budget exceeded"
    assert np.isfinite(ax.random_phases(1, 10)).all()


def test_sor_params_and_validation_without_gpu():
    """singleSemiAnalytic.cpp:19-32 defaults; bad sizes fail before any GPU work."""
    import axions_amd as ax
    p = ax.sor_params()
    assert (p.nx, p.h, p.w, p.max_iter, p.tol, p.n, p.lambda_, p.eta, p.g) == \
        (30001, 0.01, 1.5, 2000000, 1e-3, 1.0, 1.0, 1.0, 0.0)
    for bad in (dict(nx=2), dict(nx=10, h=0.0), dict(nx=10, h=-1.0)):
        try:
            ax.sor_profile(**bad)
        except ax.AxError as e:
            assert e.code == -1
        else:
            raise AssertionError(bad)


def test_boundary_combinations_validated_without_gpu():
    """Supported (xyBC, zBC) pairs per model; everything else is refused."""
    import axions_amd as ax
    from axions_amd.abi import AX_BC_ABSORBING, AX_BC_FIXED, AX_BC_NEUMANN, AX_BC_PERIODIC
    ok = [ax.ah_params(8, boundary=b, boundary_z=bz) for b in range(4) for bz in range(3)]
    ok += [ax.moore_params(8, boundary=b) for b in (AX_BC_PERIODIC, AX_BC_FIXED, AX_BC_NEUMANN)]
    ok += [ax.global_params(8, boundary=b) for b in (AX_BC_PERIODIC, AX_BC_NEUMANN)]
    n = ctypes.c_size_t()
    for p in ok:
        assert ax.lib.ax_state_size(ctypes.byref(p), None, ctypes.byref(n)) == 0, (p.model, p.boundary, p.boundary_z)
    bad = [ax.ah_params(8, boundary_z=AX_BC_ABSORBING), ax.ah_params(8, boundary=4),
           ax.ah_params(8, nz=9, boundary_z=AX_BC_NEUMANN),          # z Neumann needs nz <= nx
           ax.moore_params(8, boundary=AX_BC_ABSORBING), ax.global_params(8, boundary=AX_BC_FIXED)]
    m = ax.moore_params(8, boundary=AX_BC_NEUMANN)
    m.boundary_z = AX_BC_PERIODIC
    g = ax.global_params(8)
    g.boundary_z = AX_BC_FIXED
    for p in bad + [m, g]:
        assert ax.lib.ax_state_size(ctypes.byref(p), None, ctypes.byref(n)) == -1, (p.model, p.boundary, p.boundary_z)
