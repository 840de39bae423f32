"""axions_amd - MI355X-native lattice evolution of the Abelian-Higgs and Moore
cosmic-string models (a drop-in for the hot path of sjcott/Axions).

The compute path is libaxions.so (hand-written gfx950 HIP kernels behind the
C ABI in include/axions/axions.h); this package is its Python mirror.
"""
from .abi import (AX_BC_FIXED, AX_BC_NEUMANN_XY, AX_BC_PERIODIC, AX_MODEL_AH, AX_MODEL_GLOBAL,  # noqa: F401
                  AX_MODEL_MOORE, AxDiag, AxParams, ah_params, global_params, moore_params, ncomp, state_len)
from .lattice import AxError, Lattice, lib, random_phases, sor_params, sor_profile, string_length  # noqa: F401
from .pool import Pool  # noqa: F401
