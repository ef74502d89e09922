"""Checkpoint files in the reference's HDF5 layout (IO/Writers.jl:41-134),
through libsfdio.so (include/sfd_io.h).

``WriterSave(f)`` + ``write(w; prefix, istep)`` / ``read(w; prefix, istep)``
mirror the reference's ``WriterSave``, ``write!`` and ``read!``: the file
``{prefix}-{istep}.h5`` holds datasets ``re`` and ``im`` of the field's real
type over the global grid, x fastest.  Files move both ways between this
build and a Julia run of the reference.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib as L

IO_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libsfdio.so")
_io = None


class CheckpointError(L.SfdError):
    """A nonzero return code of libsfdio.so."""


def load_io():
    """Load libsfdio.so; raises if it was not built (no HDF5 headers)."""
    global _io
    if _io is not None:
        return _io
    if not os.path.exists(IO_PATH):
        raise CheckpointError(-2, f"{IO_PATH} is not built (the build found no HDF5 C library headers)")
    lib = ctypes.CDLL(IO_PATH)
    i64p = ctypes.POINTER(ctypes.c_int64)
    lib.sfd_io_last_error.restype = ctypes.c_char_p
    lib.sfd_io_last_error.argtypes = []
    for name, buf in (("sfd_io_write", ctypes.c_void_p), ("sfd_io_read", ctypes.c_void_p)):
        fn = getattr(lib, name)
        fn.restype = ctypes.c_int
        fn.argtypes = [ctypes.c_char_p, ctypes.c_int, i64p, ctypes.c_int, buf]
    _io = lib
    return lib


def _check(lib, rc):
    if rc != 0:
        raise CheckpointError(rc, lib.sfd_io_last_error().decode())


def filename(prefix: str, istep: int) -> str:
    """"{prefix}-{istep}.h5" (Writers.jl:48, 74)."""
    return f"{prefix}-{int(istep)}.h5"


def write_array(path: str, psi: np.ndarray) -> None:
    """Write a host complex array (x-fastest shape, complex64 or complex128)."""
    lib = load_io()
    arr = np.ascontiguousarray(psi)
    if arr.dtype not in (np.complex64, np.complex128):
        raise L.ArgError(L.SFD_ERR_ARG, f"complex64 or complex128 expected, got {arr.dtype}")
    n = (ctypes.c_int64 * arr.ndim)(*reversed(arr.shape))
    rb = 4 if arr.dtype == np.complex64 else 8
    _check(lib, lib.sfd_io_write(path.encode(), arr.ndim, n, rb, arr.ctypes.data_as(ctypes.c_void_p)))


def read_array(path: str, shape, dtype=np.complex128) -> np.ndarray:
    """Read ``re``/``im`` of a checkpoint into a host array of x-fastest
    ``shape``, converting to ``dtype``."""
    lib = load_io()
    arr = np.empty(tuple(shape), dtype=dtype)
    if arr.dtype not in (np.complex64, np.complex128):
        raise L.ArgError(L.SFD_ERR_ARG, f"complex64 or complex128 expected, got {arr.dtype}")
    n = (ctypes.c_int64 * arr.ndim)(*reversed(arr.shape))
    rb = 4 if arr.dtype == np.complex64 else 8
    _check(lib, lib.sfd_io_read(path.encode(), arr.ndim, n, rb, arr.ctypes.data_as(ctypes.c_void_p)))
    return arr


class WriterSave:
    """WriterSave(f) (Writers.jl:10-20): the HDF5 saver of a field."""

    def __init__(self, f):
        self.f = f


def write(w: WriterSave, prefix: str = "res", istep: int = 0, Δt: float = 1.0) -> str:
    """write!(w; prefix, istep, Δt) (Writers.jl:41-93): the field, copied to
    the host, into "{prefix}-{istep}.h5"; returns the path.  Δt is accepted
    and unused, as in the reference."""
    path = filename(prefix, istep)
    write_array(path, w.f.to_numpy())
    return path


def read(w: WriterSave, prefix: str = "res", istep: int = 0, Δt: float = 1.0) -> None:
    """read!(w; prefix, istep, Δt) (Writers.jl:115-134): "{prefix}-{istep}.h5"
    into the field (either stored real type, converted to the field's)."""
    dtype = np.complex64 if w.f.g.FT is np.float32 else np.complex128
    w.f.set(read_array(filename(prefix, istep), w.f.g.shape, dtype))
