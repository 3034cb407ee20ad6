"""ctypes binding of ``libtpsph.so`` (the C ABI declared in ``include/tpsph.h``).

The library is the only compute path: importing this module on a machine without the built
library raises, and nothing here falls back to a CPU implementation."""
from __future__ import annotations

import ctypes as C
import os

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG_DIR, "libtpsph.so")

TPSPH_OK = 0
TPSPH_ERR_ARGUMENT = 1
TPSPH_ERR_HIP = 2
TPSPH_ERR_CAPACITY = 3
TPSPH_ERR_STATE = 4

TPSPH_MEM_DEVICE = 0
TPSPH_MEM_HOST = 1

PAIR_FLUID_FLUID = 0
PAIR_FLUID_WALL = 1
PAIR_WALL_FLUID = 2

SEARCH_AUTO = 0
SEARCH_LISTS = 1
SEARCH_CELLS = 2
PAIR_SEARCH = {"auto": SEARCH_AUTO, "lists": SEARCH_LISTS, "cells": SEARCH_CELLS}

PRECISION_FP64 = 0
PRECISION_MIXED = 1

DENSITY_CONTINUITY = 0
DENSITY_SUMMATION = 1

TIME_CELL_LIST = 0
TIME_GATHER = 1
TIME_WALL = 2
TIME_FLUID = 3
TIME_STAGE = 4
TIME_ADJOINT = 5
TIMED_KERNELS = {"cell_list": 0, "gather": 1, "wall": 2, "fluid": 3, "stage": 4, "adjoint": 5}

EXPORTED_SYMBOLS = (
    "tpsph_abi_version", "tpsph_create", "tpsph_kick", "tpsph_drift", "tpsph_get_pressures",
    "tpsph_get_density", "tpsph_get_sound_speed",
    "tpsph_neighbors_csr", "tpsph_kick_pairs_csr", "tpsph_rk_stage", "tpsph_solve_ck2n54", "tpsph_synchronize",
    "tpsph_enable_timing", "tpsph_kernel_time", "tpsph_last_error", "tpsph_destroy",
    "tpsph_kick_noalloc", "tpsph_integrals", "tpsph_rebuild_count",
    "tpsph_kick_vjp", "tpsph_adjoint_forward", "tpsph_adjoint_reverse", "tpsph_axpby",
    "tpsph_sort_particles",
)


class TpsphParams(C.Structure):
    """Mirror of ``tpsph_params`` (include/tpsph.h)."""

    _fields_ = [
        ("ndims", C.c_int32),
        ("kernel", C.c_int32),
        ("memory", C.c_int32),
        ("clip_fluid_eos", C.c_int32),
        ("clip_wall_pressure", C.c_int32),
        ("use_viscosity", C.c_int32),
        ("use_density_diffusion", C.c_int32),
        ("cell_divisions", C.c_int32),
        ("x_cell_divisions", C.c_int32),
        ("pair_search", C.c_int32),
        ("smoothing_length", C.c_double),
        ("reference_density", C.c_double),
        ("sound_speed", C.c_double),
        ("exponent", C.c_double),
        ("background_pressure", C.c_double),
        ("alpha", C.c_double),
        ("beta", C.c_double),
        ("epsilon", C.c_double),
        ("delta", C.c_double),
        ("acceleration", C.c_double * 3),
        ("pressure_offset", C.c_double),
        ("domain_min", C.c_double * 3),
        ("domain_max", C.c_double * 3),
        ("skin", C.c_double),
        ("adaptive_sound_speed", C.c_int32),
        ("precision", C.c_int32),
        ("mach_number_target", C.c_double),
        ("min_sound_speed", C.c_double),
        ("max_sound_speed", C.c_double),
        ("wall_viscosity", C.c_int32),
        ("density_calculator", C.c_int32),
        ("wall_alpha", C.c_double),
        ("wall_beta", C.c_double),
        ("wall_epsilon", C.c_double),
        ("wall_nu", C.c_double),
    ]

    def __init__(self, *args, **kw):
        super().__init__(*args, **kw)
        self.skin = -1.0


class TpsphError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"tpsph error {code}: {message}")
        self.code = code


_lib = None


def load(path: str | None = None):
    """Load the library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with trixiparticles.jl_amd/build.py "
                           "(there is no CPU fallback for the right-hand side)")
    # torch ships its own HIP runtime under the same soname as the system's: whichever is
    # loaded first serves the whole process.  Load torch's first (when torch is installed),
    # so that the device buffers torch allocates and this library's calls share one runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    P = C.c_void_p
    dp = C.POINTER(C.c_double)
    lib.tpsph_abi_version.restype = C.c_int
    lib.tpsph_create.argtypes = [C.POINTER(TpsphParams), C.c_int64, C.c_int64, P, P, P, P,
                                 C.POINTER(C.c_void_p)]
    lib.tpsph_kick.argtypes = [P, P, P, P, C.c_double]
    lib.tpsph_drift.argtypes = [P, P, P, P, C.c_double]
    lib.tpsph_get_pressures.argtypes = [P, P, P, P]
    lib.tpsph_get_density.argtypes = [P, P]
    lib.tpsph_get_sound_speed.argtypes = [P, dp]
    lib.tpsph_neighbors_csr.argtypes = [P, P, C.c_int32, P, P, C.c_int64]
    lib.tpsph_kick_pairs_csr.argtypes = [P, C.c_int32, P, P, C.c_int64]
    lib.tpsph_rk_stage.argtypes = [P, P, P, P, C.c_double, C.c_double, C.c_double, C.c_int64,
                                   C.c_int32]
    lib.tpsph_solve_ck2n54.argtypes = [P, P, P, C.c_double, C.c_double, C.c_int64]
    lib.tpsph_axpby.argtypes = [P, C.c_int64, P, C.c_double, P, C.c_double]
    lib.tpsph_kick_noalloc.argtypes = [C.POINTER(TpsphParams), C.c_int64, C.c_int64, P, P, P,
                                       P, P, P, C.c_double, P]
    lib.tpsph_integrals.argtypes = [P, P, P, P]
    lib.tpsph_kick_vjp.argtypes = [P, P, P, P, C.c_double, P, P, P]
    lib.tpsph_adjoint_forward.argtypes = [P, P, P, C.c_double, C.c_double, C.c_int64,
                                          C.c_int64, P, P]
    lib.tpsph_adjoint_reverse.argtypes = [P, P, P, P, P, P]
    lib.tpsph_rebuild_count.argtypes = [P, C.POINTER(C.c_int64)]
    lib.tpsph_sort_particles.argtypes = [P, P, P, P]
    lib.tpsph_synchronize.argtypes = [P]
    lib.tpsph_enable_timing.argtypes = [P, C.c_int32]
    lib.tpsph_kernel_time.argtypes = [P, C.c_int32, C.POINTER(C.c_int64), dp]
    lib.tpsph_last_error.argtypes = [P]
    lib.tpsph_last_error.restype = C.c_char_p
    lib.tpsph_destroy.argtypes = [P]
    lib.tpsph_destroy.restype = None
    for name in ("tpsph_create", "tpsph_kick", "tpsph_drift", "tpsph_get_pressures",
                 "tpsph_get_density", "tpsph_get_sound_speed",
                 "tpsph_neighbors_csr", "tpsph_kick_pairs_csr", "tpsph_rk_stage", "tpsph_solve_ck2n54",
                 "tpsph_synchronize", "tpsph_enable_timing", "tpsph_kernel_time",
                 "tpsph_kick_noalloc", "tpsph_integrals", "tpsph_rebuild_count",
                 "tpsph_kick_vjp", "tpsph_adjoint_forward", "tpsph_adjoint_reverse",
                 "tpsph_axpby", "tpsph_sort_particles"):
        getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib


def check(code: int, handle=None):
    if code != TPSPH_OK:
        msg = load().tpsph_last_error(handle).decode(errors="replace")
        raise TpsphError(code, msg)
