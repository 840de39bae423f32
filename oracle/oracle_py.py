"""TEST INFRASTRUCTURE ONLY: ctypes wrapper around oracle/liboracle.so, the
CPU restatement of the reference (oracle/oracle.cpp).  Imported only by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg."""
import ctypes
import os
import subprocess

import numpy as np

from axions_amd.abi import AxParams, ncomp, state_len

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.POINTER(AxParams)
        vp = ctypes.c_void_p
        L.orc_mt19937.argtypes = [ctypes.c_uint32, ctypes.c_size_t, vp]
        L.orc_phases.argtypes = [ctypes.c_uint32, ctypes.c_size_t, vp]
        L.orc_init_random_phase.argtypes = [P, ctypes.c_uint32, ctypes.c_int, vp]
        L.orc_step.argtypes = [P, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_step_r.argtypes = [P, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_global_step.argtypes = [P, vp, ctypes.c_int, vp]
        L.orc_string_length.argtypes = [vp, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_int, vp]
        L.orc_diag.argtypes = [P, vp, ctypes.c_int, ctypes.c_int, vp]
        L.orc_winding.argtypes = [P, vp, ctypes.c_int, ctypes.c_int, vp]
        L.orc_crossings.argtypes = [P, vp, ctypes.c_int, vp, ctypes.c_size_t,
                                    ctypes.POINTER(ctypes.c_size_t)]
        L.orc_trig.argtypes = [ctypes.c_int, vp, vp, vp, ctypes.c_size_t]
        L.orc_sor_profile.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp, vp, vp]
        L.orc_sor_profile.restype = ctypes.c_int
        L.orc_straight_string.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_double, vp, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_double, vp]
        L.orc_straight_string.restype = ctypes.c_longlong
        _LIB = L
    return _LIB


def mt19937(seed, n):
    out = np.empty(n, dtype=np.uint32)
    lib().orc_mt19937(seed, n, out.ctypes.data)
    return out


def phases(seed, n):
    out = np.empty(n, dtype=np.float64)
    lib().orc_phases(seed, n, out.ctypes.data)
    return out


def init_random_phase(p, seed, mode=0):
    st = np.empty(state_len(p), dtype=np.float64)
    lib().orc_init_random_phase(ctypes.byref(p), seed, mode, st.ctypes.data)
    return st


def step(p, st, ts0, nsteps, trig_mode=0, nthreads=1):
    assert st.dtype == np.float64 and st.flags.c_contiguous and st.size == state_len(p)
    assert lib().orc_step(ctypes.byref(p), st.ctypes.data, ts0, nsteps, trig_mode, nthreads) == 0


def step_r(p, st, ts0, nsteps, nthreads=1):
    """AH step in the storage type of p.precision (float32 state for 32)."""
    dt = np.float32 if p.precision == 32 else np.float64
    assert st.dtype == dt and st.flags.c_contiguous and st.size == state_len(p)
    rc = lib().orc_step_r(ctypes.byref(p), st.ctypes.data, ts0, nsteps, nthreads)
    assert rc == 0


def global_step(p, st, nsteps):
    """Global-model steps (leapfrog.cpp) on st = [Re phi, Im phi, Re pi, Im pi];
    returns the energy of each step."""
    dt = np.float32 if p.precision == 32 else np.float64
    assert st.dtype == dt and st.flags.c_contiguous and st.size == 4 * p.nx * p.ny * p.nz
    e = np.empty(max(nsteps, 1))
    assert lib().orc_global_step(ctypes.byref(p), st.ctypes.data, nsteps, e.ctypes.data) == 0
    return e[:nsteps]


def string_length(xyz, nx, ny, nz, dx, dy, dz, pxy=True, pz=True):
    """(total, inner, outer) by the reference's all-pairs chaining."""
    pts = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1, 3))
    out = np.zeros(3)
    lib().orc_string_length(pts.ctypes.data if pts.size else None, pts.shape[0], nx, ny, nz, dx, dy, dz,
                            int(pxy), int(pz), out.ctypes.data)
    return float(out[0]), float(out[1]), float(out[2])


def sor_profile(nx, h, w, tol, n, lam, eta, g, max_iter, init=None):
    """singleSemiAnalytic.cpp's SOR, one thread: (profile [nx, 2], sweeps,
    last residual)."""
    par = np.array([h, w, tol, n, lam, eta, g], dtype=np.float64)
    out = np.zeros((nx, 2))
    res = np.zeros(1)
    ini = None if init is None else np.ascontiguousarray(init, dtype=np.float64)
    it = lib().orc_sor_profile(nx, max_iter, par.ctypes.data, None if ini is None else ini.ctypes.data,
                               out.ctypes.data, res.ctypes.data)
    return out, it, float(res[0])


def straight_string(nx, ny, nz, dx, dy, dz, prof, dr, n, g):
    """straight_ic.cpp's five ic.txt columns as [5, nx*ny*nz] and the number
    of sites beyond the profile."""
    pr = np.ascontiguousarray(prof, dtype=np.float64).reshape(-1, 2)
    st = np.zeros((5, nx * ny * nz))
    off = lib().orc_straight_string(nx, ny, nz, dx, dy, dz, pr.ctypes.data, pr.shape[0], dr, n, g, st.ctypes.data)
    return st, int(off)


def diag(p, st, n, trig_mode=0):
    out = np.zeros(3)
    assert lib().orc_diag(ctypes.byref(p), st.ctypes.data, n, trig_mode, out.ctypes.data) == 0
    return {"energy": out[0], "gauss": out[1], "finite": bool(out[2])}


def winding(p, st, n, field=0):
    out = np.empty(3 * p.nx * p.ny * p.nz, dtype=np.int8)
    lib().orc_winding(ctypes.byref(p), st.ctypes.data, n, field, out.ctypes.data)
    return out


def crossings(p, st, n, cap=None):
    cap = cap if cap is not None else 6 * p.nx * p.ny * p.nz
    xyz = np.empty((max(cap, 1), 3), dtype=np.float64)
    cnt = ctypes.c_size_t(0)
    lib().orc_crossings(ctypes.byref(p), st.ctypes.data, n, xyz.ctypes.data, cap, ctypes.byref(cnt))
    return xyz[:min(cnt.value, cap)].copy(), cnt.value


def trig(fn, x, y=None):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x) if y is None else np.ascontiguousarray(y, dtype=np.float64)
    out = np.empty_like(x)
    lib().orc_trig(fn, x.ctypes.data, y.ctypes.data, out.ctypes.data, x.size)
    return out


def levels(p, st, n):
    """(current, previous) level views [comp][x][y][z] of the state after n steps."""
    s = st.reshape(ncomp(p), 2, p.nx, p.ny, p.nz)
    return s[:, (n + 1) % 2], s[:, n % 2]
