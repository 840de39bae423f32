"""Python mirror of the ensemble-pool C ABI (ax_pool_* in include/axions/axions.h).

A Pool serves many independent periodic Abelian-Higgs realizations
(seed, N, steps) on one GPU with continuous batching: requests are admitted
in arrival order against the pool's HBM budget, every resident lattice
advances in the same persistent-kernel launches (one captured hipGraph per
chunk of steps), and each request streams back its xi(t) series.  The
reference runs one realization per job (mpi_evolution.cpp:48, :354) and has
no counterpart.
"""
import ctypes

import numpy as np

from .abi import AxParams, AxPoolConfig, AxPoolStats, AxRequestInfo, AX_POOL_KEEP_STATE  # noqa: F401
from .lattice import _check, lib

_vp, _i32, _i64, _sz = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_size_t
for _name, (_res, _args) in {
    "ax_pool_default_config": (None, [ctypes.POINTER(AxPoolConfig)]),
    "ax_pool_create": (ctypes.c_int, [ctypes.POINTER(AxParams), ctypes.POINTER(AxPoolConfig), ctypes.POINTER(_vp)]),
    "ax_pool_destroy": (None, [_vp]),
    "ax_pool_submit": (ctypes.c_int, [_vp, ctypes.c_uint32, _i32, _i32, _i32, ctypes.POINTER(_i64)]),
    "ax_pool_wait": (ctypes.c_int, [_vp, _i64, ctypes.c_double]),
    "ax_pool_info": (ctypes.c_int, [_vp, _i64, ctypes.POINTER(AxRequestInfo)]),
    "ax_pool_xi": (ctypes.c_int, [_vp, _i64, _vp, _vp, _vp, _sz, ctypes.POINTER(_sz)]),
    "ax_pool_get_state": (ctypes.c_int, [_vp, _i64, _vp, _sz]),
    "ax_pool_release": (ctypes.c_int, [_vp, _i64]),
    "ax_pool_stats_get": (ctypes.c_int, [_vp, ctypes.POINTER(AxPoolStats)]),
    "ax_pool_tune": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(_i32), ctypes.POINTER(ctypes.c_double)]),
}.items():
    _f = getattr(lib, _name)
    _f.restype, _f.argtypes = _res, _args


def default_config(**kw):
    c = AxPoolConfig()
    lib.ax_pool_default_config(ctypes.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


class Pool:
    def __init__(self, params: AxParams, **config):
        self.params = params
        self._cfg = default_config(**config)
        self._h = ctypes.c_void_p()
        _check(lib.ax_pool_create(ctypes.byref(params), ctypes.byref(self._cfg), ctypes.byref(self._h)))

    def close(self):
        if self._h:
            lib.ax_pool_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        self.close()

    def submit(self, seed, n, steps, keep_state=False):
        rid = ctypes.c_int64()
        _check(lib.ax_pool_submit(self._h, seed, n, steps, AX_POOL_KEEP_STATE if keep_state else 0,
                                  ctypes.byref(rid)))
        return rid.value

    def wait(self, rid=-1, timeout=-1.0):
        """True when done, False on timeout."""
        rc = lib.ax_pool_wait(self._h, rid, timeout)
        if rc == 1:
            return False
        _check(rc)
        return True

    def info(self, rid):
        r = AxRequestInfo()
        _check(lib.ax_pool_info(self._h, rid, ctypes.byref(r)))
        return {f: getattr(r, f) for f, _ in AxRequestInfo._fields_}

    def xi(self, rid):
        """(timesteps, wound plaquettes, xi) arrays of the samples so far."""
        n = ctypes.c_size_t()
        _check(lib.ax_pool_xi(self._h, rid, None, None, None, 0, ctypes.byref(n)))
        ts = np.empty(n.value, dtype=np.int32)
        w = np.empty(n.value, dtype=np.int64)
        x = np.empty(n.value, dtype=np.float64)
        _check(lib.ax_pool_xi(self._h, rid, ts.ctypes.data, w.ctypes.data, x.ctypes.data, n.value,
                              ctypes.byref(n)))
        return ts, w, x

    def get_state(self, rid):
        n = self.info(rid)["n"]
        out = np.empty(10 * n ** 3, dtype=np.float64)
        _check(lib.ax_pool_get_state(self._h, rid, out.ctypes.data, out.size))
        return out

    def release(self, rid):
        _check(lib.ax_pool_release(self._h, rid))

    def stats(self):
        s = AxPoolStats()
        _check(lib.ax_pool_stats_get(self._h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in AxPoolStats._fields_}

    def tune(self, n):
        """Autotune the tile x extent for requests of edge n: (xseg, ms per step)."""
        x, ms = ctypes.c_int32(), ctypes.c_double()
        _check(lib.ax_pool_tune(self._h, n, ctypes.byref(x), ctypes.byref(ms)))
        return x.value, ms.value
