"""ctypes binding of libsfdhip.so (the C ABI declared in include/sfd.h).

The library is the only compute path: if it is missing or fails to load,
every entry point raises ``SfdError`` -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsfdhip.so")
CSRC = os.path.join(os.path.dirname(_HERE), "csrc")

SFD_OK = 0
SFD_ERR_NULL = -1
SFD_ERR_SHAPE = -2
SFD_ERR_ARG = -3
SFD_ERR_OOM = -4
SFD_ERR_HIP = -5

SFD_F32 = 1
SFD_F64 = 2
SFD_F32X = 3
SFD_REAL_TIME = 0
SFD_IMAG_TIME = 1
SFD_FORWARD = -1
SFD_BACKWARD = 1
SFD_BATCH_GRAPH = 1
SFD_BATCH_ESCALATE = 2

#: every symbol include/sfd.h declares
EXPORTS = (
    "sfd_ctx_create", "sfd_ctx_destroy", "sfd_ctx_set_stream", "sfd_split_step",
    "sfd_lap_rot", "sfd_lap_rot_axis", "sfd_solve_nl", "sfd_norm2", "sfd_normalize", "sfd_energy", "sfd_split_step_energy", "sfd_fft_axis", "sfd_alloc", "sfd_free",
    "sfd_upload", "sfd_download", "sfd_synchronize", "sfd_last_error", "sfd_version",
    "sfd_autotune", "sfd_ctx_get_variant", "sfd_ctx_set_variant",
    "sfd_batch_create", "sfd_batch_destroy", "sfd_batch_imag_run", "sfd_batch_autotune", "sfd_adjoint",
    "sfd_project", "sfd_inner", "sfd_ctx_time_sweeps", "sfd_ctx_sweep_times", "sfd_split_step_noalloc",
    "sfd_batch_create_nd", "sfd_batch_imag_solve",
)


class SfdError(RuntimeError):
    """A nonzero return code of the C ABI (the reference raises ArgumentError /
    error() for the same conditions, Fields.jl:65-70, Grids.jl:69-76)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


class ShapeError(SfdError, ValueError):
    pass


class ArgError(SfdError, ValueError):
    pass


_lib = None


def build(verbose: bool = False) -> str:
    """Compile libsfdhip.so for gfx950 with hipcc (Makefile in csrc/)."""
    subprocess.run(["make", "-s", "-B", "-j4", "-C", CSRC], check=True,
                   stdout=None if verbose else subprocess.DEVNULL)
    return LIB_PATH


def load():
    """Load libsfdhip.so and declare argument types; raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SfdError(SFD_ERR_HIP, f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    vp, i64p, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double)
    c_int, c_dbl, c_i64, c_sz = ctypes.c_int, ctypes.c_double, ctypes.c_int64, ctypes.c_size_t
    sig = {
        "sfd_ctx_create": ([ctypes.POINTER(vp), c_int, i64p, dp, dp, c_int, vp], c_int),
        "sfd_ctx_destroy": ([vp], c_int),
        "sfd_ctx_set_stream": ([vp, vp], c_int),
        "sfd_split_step": ([vp, vp, vp, c_dbl, c_dbl, c_dbl, c_dbl, c_i64, c_int, c_int], c_int),
        "sfd_split_step_noalloc": ([vp, vp, vp, c_dbl, c_dbl, c_dbl, c_dbl, c_i64, c_int, c_int], c_int),
        "sfd_lap_rot": ([vp, vp, c_dbl, c_dbl, c_dbl, c_int], c_int),
        "sfd_lap_rot_axis": ([vp, vp, c_int, c_dbl, c_dbl, c_dbl, c_int], c_int),
        "sfd_solve_nl": ([vp, vp, vp, c_dbl, c_dbl, c_int], c_int),
        "sfd_norm2": ([vp, vp, dp], c_int),
        "sfd_normalize": ([vp, vp, dp], c_int),
        "sfd_energy": ([vp, vp, vp, c_dbl, c_dbl, c_dbl, dp], c_int),
        "sfd_split_step_energy": ([vp, vp, vp, c_dbl, c_dbl, c_dbl, c_dbl, c_i64, c_int, c_int, dp], c_int),
        "sfd_fft_axis": ([vp, vp, vp, c_int, c_int], c_int),
        "sfd_alloc": ([vp, c_sz, ctypes.POINTER(vp)], c_int),
        "sfd_free": ([vp, vp], c_int),
        "sfd_upload": ([vp, vp, vp, c_sz], c_int),
        "sfd_download": ([vp, vp, vp, c_sz], c_int),
        "sfd_synchronize": ([vp], c_int),
        "sfd_last_error": ([], ctypes.c_char_p),
        "sfd_version": ([], c_int),
        "sfd_autotune": ([vp, vp, ctypes.POINTER(ctypes.c_float)], c_int),
        "sfd_ctx_get_variant": ([vp, c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int), ctypes.POINTER(c_int)],
                                c_int),
        "sfd_ctx_set_variant": ([vp, c_int, c_int], c_int),
        "sfd_batch_create": ([ctypes.POINTER(vp), c_int, i64p, dp, c_int, vp], c_int),
        "sfd_batch_create_nd": ([ctypes.POINTER(vp), c_int, ctypes.POINTER(c_int), i64p, dp, c_int, vp], c_int),
        "sfd_batch_imag_solve": ([vp, ctypes.POINTER(vp), ctypes.POINTER(vp), dp, dp, c_dbl, c_dbl, c_i64, c_dbl,
                                  c_int, c_int, i64p, dp, ctypes.POINTER(vp), ctypes.POINTER(c_int)], c_int),
        "sfd_batch_destroy": ([vp], c_int),
        "sfd_batch_autotune": ([vp, ctypes.POINTER(ctypes.c_float)], c_int),
        "sfd_adjoint": ([vp, vp, vp, vp, c_dbl, c_dbl, c_dbl, c_dbl, c_i64, c_i64, vp, dp, dp, vp], c_int),
        "sfd_project": ([vp, vp, vp, c_int, dp], c_int),
        "sfd_inner": ([vp, vp, vp, dp], c_int),
        "sfd_ctx_time_sweeps": ([vp, c_int], c_int),
        "sfd_ctx_sweep_times": ([vp, dp, i64p, c_int], c_int),
        "sfd_batch_imag_run": ([vp, ctypes.POINTER(vp), ctypes.POINTER(vp), dp, c_dbl, c_dbl, c_i64, c_dbl, c_int,
                                c_int, i64p, dp], c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def check(rc: int) -> None:
    if rc == SFD_OK:
        return
    msg = load().sfd_last_error().decode(errors="replace")
    if rc == SFD_ERR_SHAPE:
        raise ShapeError(rc, msg)
    if rc in (SFD_ERR_ARG, SFD_ERR_NULL):
        raise ArgError(rc, msg)
    raise SfdError(rc, msg)
