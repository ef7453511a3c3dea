"""ctypes binding of libmpm_hip.so (the C ABI declared in include/mpm_hip.h).

This is the Python equivalent of the ``ccall`` layer a Julia package extension
would use.  Loading fails loudly when the library has not been built: there is
no CPU fallback behind this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmpm_hip.so")

MPM_OK = 0
MPM_ABI_VERSION = 2  # mpm_upload_materials takes phir (version 1: without it)
ERRORS = {-1: "MPM_E_ARG", -2: "MPM_E_UNSUPPORTED", -3: "MPM_E_HIP", -4: "MPM_E_NOMEM",
          -5: "MPM_E_STATE", -6: "MPM_E_RANGE"}
MPM_FP64, MPM_FP32 = 0, 1
MPM_UGIMP, MPM_LINEAR = 0, 1
MPM_MUSL, MPM_USL, MPM_USF = 0, 1, 2
SCHEME = {"MUSL": MPM_MUSL, "USL": MPM_USL, "USF": MPM_USF}
MPM_VA_A, MPM_VA_V = 0, 1
MPM_OPT_SORT_INTERVAL, MPM_OPT_TILES, MPM_OPT_PROFILE_MASK, MPM_OPT_SCATTER = 0, 1, 2, 3
MPM_OPT_FUSED_REMAP, MPM_OPT_TILE_PARTIALS, MPM_OPT_REMAP_SCATTER = 4, 5, 6
MPM_SCATTER_ROUNDS, MPM_SCATTER_OWNED, MPM_SCATTER_OWNED_LAYERS = 0, 1, 2
MPM_SCATTER_ROUNDS_LAYERED, MPM_SCATTER_PAIRS = 3, 4
SCATTER = {"rounds": MPM_SCATTER_ROUNDS, "owned": MPM_SCATTER_OWNED,
           "owned_layers": MPM_SCATTER_OWNED_LAYERS, "rounds_layered": MPM_SCATTER_ROUNDS_LAYERED,
           "pairs": MPM_SCATTER_PAIRS}
CONSTITUTIVE = {"linearelastic": 0, "druckerprager": 1, "mohrcoulomb": 2}


class MPMConfig(C.Structure):
    _fields_ = [("dim", C.c_int32), ("precision", C.c_int32), ("basis", C.c_int32),
                ("sort_interval", C.c_int32), ("np", C.c_int64), ("nmat", C.c_int64),
                ("nnx", C.c_int64), ("nny", C.c_int64), ("nnz", C.c_int64),
                ("x1", C.c_double), ("y1", C.c_double), ("z1", C.c_double),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("lpx", C.c_double), ("lpy", C.c_double), ("lpz", C.c_double)]


class MPMArgs(C.Structure):
    _fields_ = [("scheme", C.c_int32), ("va", C.c_int32), ("constitutive", C.c_int32),
                ("mvl", C.c_int32), ("flip", C.c_double), ("pic", C.c_double),
                ("zeta_s", C.c_double), ("gravity", C.c_double), ("Te", C.c_double),
                ("gravity_x", C.c_double), ("gravity_y", C.c_double)]


class MPMError(RuntimeError):
    """A non-zero status from libmpm_hip.so."""

    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None

_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_I64 = C.POINTER(C.c_int64)
_CTX = C.c_void_p

_PROTOS = {
    "mpm_abi_version": (C.c_int, []),
    "mpm_create": (C.c_int, [C.POINTER(MPMConfig), C.POINTER(_CTX)]),
    "mpm_upload_particles": (C.c_int, [_CTX, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D, _I64]),
    "mpm_upload_materials": (C.c_int, [_CTX, _D, _D, _D, _D, _D, _D, _D, _D, _D]),
    "mpm_upload_bc": (C.c_int, [_CTX, _I64, _D, _I64, _D, _I64, _D]),
    "mpm_step": (C.c_int, [_CTX, C.POINTER(MPMArgs), C.c_double, C.c_double]),
    "mpm_run": (C.c_int, [_CTX, C.POINTER(MPMArgs), C.c_double, C.c_double, C.c_int64, _D]),
    "mpm_run_graph": (C.c_int, [_CTX, C.POINTER(MPMArgs), C.c_double, C.c_double, C.c_int64, _D]),
    "mpm_sync": (C.c_int, [_CTX]),
    "mpm_download_particles": (C.c_int, [_CTX, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D]),
    "mpm_destroy": (None, [_CTX]),
    "mpm_last_error": (C.c_char_p, [_CTX]),
    "mpm_profile_enable": (C.c_int, [_CTX, C.c_int]),
    "mpm_profile_read": (C.c_int, [_CTX, _D, _I64, C.c_int]),
    "mpm_profile_kernel_name": (C.c_char_p, [C.c_int]),
    "mpm_profile_kernel_count": (C.c_int, []),
    "mpm_count_active_nodes": (C.c_int, [_CTX, _I64]),
    "mpm_tpeak": (C.c_int, [C.c_int64, C.c_int, _D]),
    "mpm_mlp_param_count": (C.c_int, []),
    "mpm_mlp_create": (C.c_int, [C.c_int64, C.POINTER(_CTX)]),
    "mpm_mlp_destroy": (None, [_CTX]),
    "mpm_mlp_last_error": (C.c_char_p, [_CTX]),
    "mpm_mlp_set_params": (C.c_int, [_CTX, _F, C.c_int64]),
    "mpm_mlp_get_params": (C.c_int, [_CTX, _F, C.c_int64]),
    "mpm_mlp_forward": (C.c_int, [_CTX, _F, C.c_int64, _F]),
    "mpm_mlp_grad": (C.c_int, [_CTX, _F, _F, C.c_int64, _F, _F]),
    "mpm_mlp_train": (C.c_int, [_CTX, _F, _F, C.c_int64, C.c_int64, C.c_int64, C.c_double,
                                C.c_uint64, _F]),
    "mpm_mlp_bench": (C.c_int, [_CTX, C.c_int64, C.c_int, _D]),
    "mpm_set_surrogate": (C.c_int, [_CTX, _CTX, C.c_double]),
    "mpm_set_option": (C.c_int, [_CTX, C.c_int, C.c_int64]),
    "mpm_surrogate_stats": (C.c_int, [_CTX, _I64, _I64]),
    "mpm_return_log": (C.c_int, [_CTX, C.c_int64, C.c_int64]),
    "mpm_return_log_read": (C.c_int, [_CTX, _F, _F, C.c_int64, _I64]),
    "mpm_count_active_nodes_noalloc": (C.c_int, [_CTX, _I64]),
}


def load(path: str | None = None):
    """Load libmpm_hip.so and declare its prototypes.  Raises if it is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(f"libmpm_hip.so not built at {p}; run __graft_entry__.build() "
                          "(hipcc --offload-arch=gfx950).  There is no CPU fallback.")
    lib = C.CDLL(p)
    for name, (res, args) in _PROTOS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if lib.mpm_abi_version() != MPM_ABI_VERSION:
        raise ImportError("libmpm_hip.so ABI version mismatch")
    if path is None:
        _lib = lib
    return lib


def exported_symbols():
    return list(_PROTOS)


def dptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.flags.c_contiguous)
    return a.ctypes.data_as(_D)


def fptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_F)


def iptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.int64 and (a.flags.f_contiguous or a.flags.c_contiguous)
    return a.ctypes.data_as(_I64)


def check(lib, ctx, rc):
    if rc != MPM_OK:
        msg = lib.mpm_last_error(ctx)
        raise MPMError(rc, msg.decode() if msg else "")
    return rc
