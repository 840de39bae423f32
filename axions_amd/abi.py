"""ctypes mirror of the C ABI structs in include/axions/axions.h.

Only layouts and constants live here; loading libaxions.so is done by
axions_amd.lattice.  Parameter names are the reference's file-scope consts
(evolution.cpp:23-59, mpi_evolution.cpp:17-60, mooreEvol.cpp:19-42).
"""
import ctypes
import math

AX_OK = 0
AX_EINVAL = -1
AX_ENOMEM = -2
AX_EBUDGET = -3
AX_EHIP = -4
AX_ENONFINITE = -5
AX_ENOTFOUND = -6
AX_ENODEV = -7

AX_MODEL_AH = 0
AX_MODEL_MOORE = 1
AX_MODEL_GLOBAL = 2
AX_BC_PERIODIC = 0
AX_BC_FIXED = 1
AX_BC_NEUMANN = 2
AX_BC_NEUMANN_XY = AX_BC_NEUMANN
AX_BC_ABSORBING = 3


class AxParams(ctypes.Structure):
    _fields_ = [
        ("model", ctypes.c_int32),
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("lambda_", ctypes.c_double), ("eta", ctypes.c_double), ("g", ctypes.c_double),
        ("lambda1", ctypes.c_double), ("lambda2", ctypes.c_double), ("lambda12", ctypes.c_double),
        ("eta1", ctypes.c_double), ("eta2", ctypes.c_double),
        ("q1", ctypes.c_int32), ("q2", ctypes.c_int32),
        ("alpha", ctypes.c_double), ("scaling", ctypes.c_double),
        ("damped_nt", ctypes.c_int32),
        ("dampFac", ctypes.c_double),
        ("theta_damp_form", ctypes.c_int32),
        ("energy_form", ctypes.c_int32),
        ("boundary", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("boundary_z", ctypes.c_int32),
    ]


class AxSorParams(ctypes.Structure):
    """ax_sor_params: constants of singleSemiAnalytic.cpp:19-32."""
    _fields_ = [
        ("nx", ctypes.c_int32),
        ("max_iter", ctypes.c_int32),
        ("h", ctypes.c_double),
        ("w", ctypes.c_double),
        ("tol", ctypes.c_double),
        ("n", ctypes.c_double),
        ("lambda_", ctypes.c_double),
        ("eta", ctypes.c_double),
        ("g", ctypes.c_double),
    ]


class AxDiag(ctypes.Structure):
    _fields_ = [
        ("timestep", ctypes.c_int32),
        ("finite", ctypes.c_int32),
        ("energy", ctypes.c_double),
        ("gauss_mean", ctypes.c_double),
        ("wound_plaquettes", ctypes.c_int64),
        ("string_length", ctypes.c_double),
        ("xi", ctypes.c_double),
    ]


def ah_params(n=64, ny=None, nz=None, dx=0.5, dt=0.1, lam=2.0, eta=1.0, g=1.0, alpha=0.0,
              scaling=1.0, damped_nt=0, dampFac=0.5, theta_damp_form=0, energy_form=0,
              precision=64, boundary=AX_BC_PERIODIC, dy=None, dz=None, boundary_z=None):
    """Abelian-Higgs parameters; defaults are the plain set (dx=0.5, dt=0.1,
    lambda=2, g=eta=1, no friction).  boundary is the reference's xyBC,
    boundary_z its zBC (default: fixed when boundary is fixed, else
    periodic)."""
    p = AxParams()
    p.model = AX_MODEL_AH
    p.nx, p.ny, p.nz = n, (ny or n), (nz or n)
    p.dx = dx
    p.dy = dx if dy is None else dy
    p.dz = dx if dz is None else dz
    p.dt = dt
    p.lambda_, p.eta, p.g = lam, eta, g
    p.lambda1 = p.lambda2 = 2.0
    p.lambda12 = 0.0
    p.eta1 = p.eta2 = 1.0
    p.q1, p.q2 = 2, 1
    p.alpha, p.scaling = alpha, scaling
    p.damped_nt, p.dampFac = damped_nt, dampFac
    p.theta_damp_form, p.energy_form = theta_damp_form, energy_form
    p.boundary, p.precision = boundary, precision
    if boundary_z is None:
        boundary_z = AX_BC_FIXED if boundary == AX_BC_FIXED else AX_BC_PERIODIC
    p.boundary_z = boundary_z
    return p


def moore_params(n=64, ny=None, nz=None, dx=0.7, dt=0.3, lambda1=2.0, lambda2=2.0, lambda12=0.0,
                 eta1=1.0, eta2=1.0, q1=2, q2=1, g=None, boundary=AX_BC_PERIODIC, precision=64):
    """Moore-model parameters at mooreEvol.cpp:23-36 defaults;
    g = sqrt(0.5*lambda1*eta1^2/(q1^2+q2^2)) as mooreEvol.cpp:36.  z is
    periodic with periodic x/y, else fixed (mooreEvol.cpp:187-189)."""
    p = AxParams()
    p.model = AX_MODEL_MOORE
    p.nx, p.ny, p.nz = n, (ny or n), (nz or n)
    p.dx = p.dy = p.dz = dx
    p.dt = dt
    p.lambda_, p.eta = 2.0, 1.0
    p.lambda1, p.lambda2, p.lambda12 = lambda1, lambda2, lambda12
    p.eta1, p.eta2 = eta1, eta2
    p.q1, p.q2 = q1, q2
    p.g = math.sqrt(0.5 * lambda1 * eta1 ** 2 / (q1 * q1 + q2 * q2)) if g is None else g
    p.alpha, p.scaling, p.damped_nt, p.dampFac = 0.0, 1.0, 0, 0.0
    p.theta_damp_form, p.energy_form = 0, 0
    p.boundary, p.precision = boundary, precision
    p.boundary_z = AX_BC_PERIODIC if boundary == AX_BC_PERIODIC else AX_BC_FIXED
    return p


def global_params(n=64, ny=None, nz=None, dx=0.25, dt=0.1, dy=None, dz=None, boundary=AX_BC_NEUMANN_XY,
                  precision=64):
    """Global-string parameters at leapfrog.cpp:74-81 defaults (dx = 0.25,
    dt = 0.1, Neumann x/y faces, periodic z); the potential is the program's
    fixed (|phi|^2 - 1) phi."""
    p = AxParams()
    p.model = AX_MODEL_GLOBAL
    p.nx, p.ny, p.nz = n, (ny or n), (nz or n)
    p.dx, p.dy, p.dz = dx, (dy if dy is not None else dx), (dz if dz is not None else dx)
    p.dt = dt
    p.lambda_, p.eta, p.g = 2.0, 1.0, 0.0
    p.scaling = 1.0
    p.boundary, p.precision = boundary, precision
    p.boundary_z = AX_BC_PERIODIC
    return p


def ncomp(p):
    return {AX_MODEL_AH: 5, AX_MODEL_MOORE: 7, AX_MODEL_GLOBAL: 4}[p.model]


def state_len(p):
    """Host state length: two levels per component, except the global model
    (field and momentum, one level each)."""
    levels = 1 if p.model == AX_MODEL_GLOBAL else 2
    return ncomp(p) * levels * p.nx * p.ny * p.nz


AX_POOL_KEEP_STATE = 1
AX_REQ_QUEUED, AX_REQ_RUNNING, AX_REQ_DONE, AX_REQ_FAILED = 0, 1, 2, 3


class AxPoolConfig(ctypes.Structure):
    _fields_ = [
        ("hbm_budget", ctypes.c_size_t),
        ("max_lattices", ctypes.c_int32),
        ("chunk_steps", ctypes.c_int32),
        ("ic_threads", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
        ("autotune", ctypes.c_int32),
        ("xseg_default", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class AxRequestInfo(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_int64),
        ("seed", ctypes.c_uint32),
        ("n", ctypes.c_int32), ("steps", ctypes.c_int32), ("status", ctypes.c_int32),
        ("nsamples", ctypes.c_int32), ("xseg", ctypes.c_int32),
        ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double),
    ]


class AxPoolStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_int64), ("admitted", ctypes.c_int64),
        ("completed", ctypes.c_int64), ("failed", ctypes.c_int64),
        ("chunks", ctypes.c_int64),
        ("resident", ctypes.c_int32), ("queued", ctypes.c_int32),
        ("max_resident", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("site_updates", ctypes.c_double),
        ("elapsed", ctypes.c_double),
        ("budget", ctypes.c_size_t), ("in_use", ctypes.c_size_t), ("peak_in_use", ctypes.c_size_t),
        ("step_ms", ctypes.c_double),
    ]


def lattice_bytes(n, precision, pad_bytes=4352):
    """Device bytes the pool holds for one n^3 request: 5 components x 2
    levels of planes, each padded by pad_bytes, then (256 B aligned) the
    3 x n^3 FP32 edge values of its winding pass, rounded to 256 B."""
    es = 8 if precision == 64 else 4
    V = n ** 3
    b = 10 * (V + (pad_bytes + es - 1) // es) * es
    b = (b + 255) // 256 * 256 + 3 * V * 4
    return (b + 255) // 256 * 256
