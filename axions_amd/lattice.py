"""Python mirror of the C ABI (include/axions/axions.h) over libaxions.so.

The reference drives its lattice from one main() (mpi_evolution.cpp:64-998);
`Lattice` exposes the same steps as methods: seeded IC, TimeStep loop,
per-step diagnostics, string detection, field read-back.  Every call goes to
the HIP library; there is no Python or CPU fallback -- a missing or
unloadable libaxions.so raises at import.
"""
import ctypes
import os

import numpy as np

from .abi import AxDiag, AxParams, AxSorParams, AX_OK, ncomp, state_len

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libaxions.so")


class AxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"axions error {code}: {msg}")
        self.code = code


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() or "
                          "python axions_amd/build.py")
    L = ctypes.CDLL(LIB_PATH)
    vp, P, sz = ctypes.c_void_p, ctypes.POINTER(AxParams), ctypes.c_size_t
    i32 = ctypes.c_int32
    sig = {
        "ax_default_params": (None, [P]),
        "ax_strerror": (ctypes.c_char_p, [ctypes.c_int]),
        "ax_last_error": (ctypes.c_char_p, []),
        "ax_state_size": (ctypes.c_int, [P, ctypes.POINTER(i32), ctypes.POINTER(sz)]),
        "ax_create": (ctypes.c_int, [P, ctypes.POINTER(vp)]),
        "ax_destroy": (None, [vp]),
        "ax_get_params": (ctypes.c_int, [vp, P]),
        "ax_init_random_phase": (ctypes.c_int, [vp, ctypes.c_uint32, i32]),
        "ax_set_state": (ctypes.c_int, [vp, vp, sz]),
        "ax_get_state": (ctypes.c_int, [vp, vp, sz]),
        "ax_set_timestep": (ctypes.c_int, [vp, i32]),
        "ax_get_timestep": (ctypes.c_int, [vp, ctypes.POINTER(i32)]),
        "ax_step": (ctypes.c_int, [vp, i32]),
        "ax_sync": (ctypes.c_int, [vp]),
        "ax_diagnostics": (ctypes.c_int, [vp, i32, ctypes.POINTER(AxDiag)]),
        "ax_step_diag": (ctypes.c_int, [vp, i32, i32, i32, vp, sz, ctypes.POINTER(sz)]),
        "ax_winding": (ctypes.c_int, [vp, vp, sz]),
        "ax_winding_field": (ctypes.c_int, [vp, i32, vp, sz]),
        "ax_crossings": (ctypes.c_int, [vp, vp, sz, ctypes.POINTER(sz)]),
        "ax_crossings_noalloc": (ctypes.c_int, [vp, vp, sz, vp]),
        "ax_random_phases": (ctypes.c_int, [ctypes.c_uint32, sz, vp]),
        "ax_kernel_timer": (ctypes.c_int, [vp, i32]),
        "ax_set_tile": (ctypes.c_int, [vp, i32]),
        "ax_string_length": (ctypes.c_int, [vp, sz, i32, i32, i32, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double, i32, i32, vp]),
        "ax_get_tile": (ctypes.c_int, [vp, ctypes.POINTER(i32), ctypes.POINTER(i32)]),
        "ax_kernel_time": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)]),
        "ax_sor_default_params": (None, [ctypes.POINTER(AxSorParams)]),
        "ax_sor_profile": (ctypes.c_int, [ctypes.POINTER(AxSorParams), vp, vp, ctypes.POINTER(i32),
                                          ctypes.POINTER(ctypes.c_double)]),
        "ax_init_straight_string": (ctypes.c_int, [vp, vp, i32, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                   ctypes.POINTER(ctypes.c_int64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    return L


lib = _load()


def _check(rc):
    if rc != AX_OK:
        raise AxError(rc, lib.ax_last_error().decode() or lib.ax_strerror(rc).decode())


def string_length(xyz, nx, ny, nz, dx, dy, dz, periodic_xy=True, periodic_z=True):
    """(total, inner, outer) chained string length of crossing points
    (ic_w_evolution.cpp:961-1087)."""
    pts = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1, 3))
    out = np.zeros(3)
    _check(lib.ax_string_length(pts.ctypes.data if pts.size else None, pts.shape[0], nx, ny, nz, dx, dy, dz,
                                1 if periodic_xy else 0, 1 if periodic_z else 0, out.ctypes.data))
    return float(out[0]), float(out[1]), float(out[2])


def sor_params(**kw):
    """ax_sor_params with singleSemiAnalytic.cpp's constants, overridden by
    keyword (nx, max_iter, h, w, tol, n, lam, eta, g)."""
    p = AxSorParams()
    lib.ax_sor_default_params(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, "lambda_" if k in ("lam", "lambda") else k, v)
    return p


def sor_profile(params=None, init=None, **kw):
    """Straight-string radial profile by SOR (singleSemiAnalytic.cpp:54-134)
    on the GPU: (profile [nx, 2] of (f, A), sweeps, last residual).  init:
    a starting profile [nx, 2] (the reference's ic_type 1) or None (tanh)."""
    p = params if params is not None else sor_params(**kw)
    out = np.empty((p.nx, 2), dtype=np.float64)
    ini = None
    if init is not None:
        ini = np.ascontiguousarray(init, dtype=np.float64)
        if ini.shape != (p.nx, 2):
            raise ValueError(f"init profile shape {ini.shape} != ({p.nx}, 2)")
    it = ctypes.c_int32()
    res = ctypes.c_double()
    _check(lib.ax_sor_profile(ctypes.byref(p), None if ini is None else ini.ctypes.data, out.ctypes.data,
                              ctypes.byref(it), ctypes.byref(res)))
    return out, it.value, res.value


def random_phases(seed, n):
    """Phases of the seeded IC: std::mt19937(seed) through
    std::uniform_real_distribution<double>(0, 2pi) (mpi_evolution.cpp:290-301)."""
    out = np.empty(n, dtype=np.float64)
    _check(lib.ax_random_phases(seed, n, out.ctypes.data))
    return out


class Lattice:
    """One lattice resident in HBM.  Parameters use the reference's names."""

    def __init__(self, params: AxParams):
        self.params = params
        h = ctypes.c_void_p()
        _check(lib.ax_create(ctypes.byref(params), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib.ax_destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    @property
    def shape(self):
        p = self.params
        return (p.nx, p.ny, p.nz)

    def init_random_phase(self, seed=42, moore_phase_mode=0):
        _check(lib.ax_init_random_phase(self._h, seed, moore_phase_mode))

    def init_straight_string(self, profile, dr, n=1.0, g=None):
        """Straight string along z from a radial profile [nprof, 2] of
        (f, A) spaced dr (straight_ic.cpp:59-146); g defaults to the
        lattice's coupling.  Returns the number of sites beyond the
        profile."""
        pr = np.ascontiguousarray(profile, dtype=np.float64).reshape(-1, 2)
        off = ctypes.c_int64()
        gg = self.params.g if g is None else g
        _check(lib.ax_init_straight_string(self._h, pr.ctypes.data, pr.shape[0], dr, n, gg, ctypes.byref(off)))
        return off.value

    def set_state(self, st, timestep=0):
        st = np.ascontiguousarray(st, dtype=np.float64)
        _check(lib.ax_set_state(self._h, st.ctypes.data, st.size))
        _check(lib.ax_set_timestep(self._h, timestep))

    def get_state(self):
        st = np.empty(state_len(self.params), dtype=np.float64)
        _check(lib.ax_get_state(self._h, st.ctypes.data, st.size))
        return st

    @property
    def timestep(self):
        t = ctypes.c_int32()
        _check(lib.ax_get_timestep(self._h, ctypes.byref(t)))
        return t.value

    def step(self, n=1):
        _check(lib.ax_step(self._h, n))

    def sync(self):
        _check(lib.ax_sync(self._h))

    def set_tile(self, xseg):
        """x planes one wave of the marching step advances (0 = whole x extent)."""
        _check(lib.ax_set_tile(self._h, xseg))

    def tile(self):
        """(xseg, uses_marching_kernel)."""
        x, t = ctypes.c_int32(), ctypes.c_int32()
        _check(lib.ax_get_tile(self._h, ctypes.byref(x), ctypes.byref(t)))
        return x.value, bool(t.value)

    def kernel_timer(self, enable=True):
        _check(lib.ax_kernel_timer(self._h, 1 if enable else 0))

    def kernel_time(self):
        """(GPU milliseconds, steps) of the step kernels timed since the last call."""
        ms, n = ctypes.c_double(), ctypes.c_int64()
        _check(lib.ax_kernel_time(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def diagnostics(self, winding=True):
        d = AxDiag()
        _check(lib.ax_diagnostics(self._h, 1 if winding else 0, ctypes.byref(d)))
        return {f: getattr(d, f) for f, _ in d._fields_}

    def step_diag(self, n, every, winding=True):
        """n TimeSteps with the diagnostics of every TimeStep t % every == 0
        taken before its step (fused into the step where the library can):
        a list of diagnostics dicts."""
        cap = n // max(every, 1) + 2
        arr = (AxDiag * cap)()
        cnt = ctypes.c_size_t()
        _check(lib.ax_step_diag(self._h, n, every, 1 if winding else 0, ctypes.cast(arr, ctypes.c_void_p), cap,
                                ctypes.byref(cnt)))
        return [{f: getattr(arr[i], f) for f, _ in AxDiag._fields_} for i in range(cnt.value)]

    def winding(self, field=0):
        p = self.params
        out = np.empty(3 * p.nx * p.ny * p.nz, dtype=np.int8)
        _check(lib.ax_winding_field(self._h, field, out.ctypes.data, out.size))
        return out

    def crossings(self, cap=None):
        p = self.params
        cap = 6 * p.nx * p.ny * p.nz if cap is None else cap
        xyz = np.empty((max(cap, 1), 3), dtype=np.float64)
        cnt = ctypes.c_size_t()
        _check(lib.ax_crossings(self._h, xyz.ctypes.data, cap, ctypes.byref(cnt)))
        return xyz[:min(cap, cnt.value)].copy(), cnt.value

    def levels(self, st=None):
        """(current, previous) levels [comp][x][y][z] of a state array."""
        st = self.get_state() if st is None else st
        n = self.timestep
        s = st.reshape(ncomp(self.params), 2, *self.shape)
        return s[:, (n + 1) % 2], s[:, n % 2]
