"""ctypes binding of libloopstructural_amd.so (the C ABI in include/).

Device buffers are torch tensors on the current HIP device; their
``data_ptr()`` and torch's current stream are handed to the library.  There is
no fallback: if the library or a GPU is missing, every compute call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import torch

from . import build as _build

LSAMD_OK = 0
LSAMD_ERR_ARG = 1
LSAMD_ERR_HIP = 2
LSAMD_ERR_ALLOC = 3
LSAMD_ERR_INJECTED = 9
FP64 = 0
FP32 = 1
ROWS_VALUE, ROWS_NORMAL, ROWS_GRADIENT, ROWS_TANGENT = 0, 1, 2, 3


class LsamdError(RuntimeError):
    """A non-zero status returned by libloopstructural_amd."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"libloopstructural_amd error {code}: {msg}")
        self.code = code


class Grid(C.Structure):
    _fields_ = [("origin", C.c_double * 3), ("step", C.c_double * 3),
                ("nnodes", C.c_int64 * 3), ("axis_nodes", C.c_void_p * 3)]


class FdReg(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("w2_xx", "w2_yy", "w2_zz", "w2_xy", "w2_xz", "w2_yz",
                                          "w2_bx", "w2_by", "w2_bz", "ridge2")]


class SolveOpts(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("maxiter", C.c_int64),
                ("precision", C.c_int32), ("check_every", C.c_int32),
                ("inner_rtol", C.c_double), ("max_outer", C.c_int64),
                ("jacobi", C.c_int32), ("pad", C.c_int32)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int64), ("outer", C.c_int64), ("converged", C.c_int32),
                ("pad", C.c_int32), ("residual", C.c_double), ("threshold", C.c_double)]


class LsmrOpts(C.Structure):
    _fields_ = [("atol", C.c_double), ("btol", C.c_double), ("conlim", C.c_double),
                ("maxiter", C.c_int64), ("check_every", C.c_int32), ("scaling", C.c_int32)]


class LsmrInfo(C.Structure):
    _fields_ = [("iterations", C.c_int64), ("istop", C.c_int32), ("pad", C.c_int32),
                ("normr", C.c_double), ("normar", C.c_double), ("normA", C.c_double),
                ("condA", C.c_double), ("normx", C.c_double)]


# largest point count of lsamd_fd_evaluate_adjoint_noalloc
NOALLOC_MAX_POINTS = 65536

_P = C.c_void_p
_I64 = C.c_int64
_I32 = C.c_int32
_D = C.c_double

# name -> argtypes (restype is int status unless listed in _RESTYPE)
SIGNATURES = {
    "lsamd_last_error": [],
    "lsamd_version": [],
    "lsamd_device_count": [_P],
    "lsamd_fd_locate": [_P, _P, _I64, _P, _P, _P, _P, _P],
    "lsamd_fd_build_rows": [_P, _I32, _P, _I64, _I64, _I32, _D, _P, _P, _P, _P, _P, _P],
    "lsamd_fd_system_create": [_P, _P, _P, _P],
    "lsamd_fd_system_destroy": [_P],
    "lsamd_fd_system_set_rows": [_P, _I64, _I64, _P, _P, _P, _P, _P, _P],
    "lsamd_fd_system_apply": [_P, _I32, _P, _P, _P],
    "lsamd_fd_system_rhs": [_P, _P, _P],
    "lsamd_fd_system_diag": [_P, _P, _P],
    "lsamd_fd_system_solve": [_P, _P, _P, _P, _P, _P],
    "lsamd_fd_system_lsmr": [_P, _P, _P, _P, _P],
    "lsamd_fd_system_solve_multi": [_P, _P, _I32, _P, _P, _P, _P],
    "lsamd_fd_system_rhs_rows": [_P, _P, _P, _P],
    "lsamd_fd_evaluate": [_P, _P, _P, _I64, _P, _P, _P],
    "lsamd_rows_adjoint": [_P, _P, _P, _P, _P, _P],
    "lsamd_rows_normalise": [_I64, _I32, _P, _P, _P, _P],
    "lsamd_fd_nearest_d2": [_P, _P, _I64, _P, _P],
    "lsamd_fd_evaluate_adjoint": [_P, _P, _I64, _P, _P, _P],
    "lsamd_fd_evaluate_adjoint_noalloc": [_P, _P, _I64, _P, _P, _P],
    "lsamd_fd_system_evaluate_adjoint": [_P, _P, _I64, _P, _P, _P],
    "lsamd_fd_system_time_stencil": [_P, _I32, _I32, _P, _P],
    "lsamd_fd_system_time_multi": [_P, _I32, _I32, _P, _P, _P],
    "lsamd_fd_system_mg_build": [_P, _P],
    "lsamd_fd_systems_solve_mg": [_P, _I32, _P, _P, _P, _P],
    "lsamd_fd_system_stencil_info": [_P, _I32, _P, _P],
    "lsamd_p1_locate": [_P, _P, _I64, _P, _P, _P, _P, _P, _P],
    "lsamd_p1_elements": [_P, _P, _P, _P],
    "lsamd_p1_tet_gradients": [_P, _P, _I64, _P, _P, _P],
    "lsamd_p1_n_pairs": [_P],
    "lsamd_p1_pairs": [_P, _P, _P, _P, _P, _P, _P, _P, _D, _P],
    "lsamd_p1_build_rows": [_P, _I32, _P, _I64, _I64, _D, _D, _P, _P, _P, _P, _P, _P],
    "lsamd_p1_evaluate": [_P, _P, _P, _I64, _P, _P, _P],
    "lsamd_p1_system_create": [_P, _D, _D, _P],
    "lsamd_p1_operator_table": [_P, _D, _P, _P],
    "lsamd_batch_create": [_I32, _P, _P, _P],
    "lsamd_batch_offsets": [_P, _P],
    "lsamd_batch_set_rows": [_P, _P, _I64, _I64, _P, _P, _P, _P, _P, _P],
    "lsamd_batch_solve": [_P, _P, _P, _P, _P, _P],
}
_RESTYPE = {"lsamd_last_error": C.c_char_p, "lsamd_p1_n_pairs": C.c_int64}

_lib = None


def lib_path() -> str:
    return _build.LIB_PATH


def load() -> C.CDLL:
    """Load the in-tree library (build it first if it is absent)."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        _build.build()
    lib = C.CDLL(path)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, C.c_int)
    _lib = lib
    return lib


def check(rc: int):
    if rc != LSAMD_OK:
        msg = load().lsamd_last_error().decode(errors="replace")
        raise LsamdError(rc, msg)


def call(name: str, *args):
    check(getattr(load(), name)(*args))


def device() -> torch.device:
    """The HIP device compute runs on; raises when no GPU is present."""
    if not torch.cuda.is_available():
        raise RuntimeError(
            "loopstructural_amd needs a ROCm GPU (MI355X/gfx950): no HIP device is visible. "
            "There is no CPU fallback.")
    return torch.device("cuda", torch.cuda.current_device())


def stream() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t) -> C.c_void_p:
    if t is None:
        return C.c_void_p(None)
    return C.c_void_p(t.data_ptr())


def make_grid(origin, step, nnodes) -> Grid:
    g = Grid()
    for a in range(3):
        g.origin[a] = float(origin[a])
        g.step[a] = float(step[a])
        g.nnodes[a] = int(nnodes[a])
    return g
