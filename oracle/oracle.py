"""ctypes front end of the FP64 CPU oracle (``sfd_oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, by ``__graft_entry__.smoke``
(as the checker) and by the CPU-baseline leg of ``bench.py``.  The HIP product
path never imports this module.

Every function follows the reference file:line listed in ``sfd_oracle.c``.
Arrays are numpy complex128 in the reference's memory order: x fastest, i.e.
a numpy array of shape ``(nz, ny, nx)`` / ``(ny, nx)`` in C order.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")
_lib = None

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_dbl_p = ctypes.POINTER(ctypes.c_double)


def build() -> str:
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_split_step.argtypes = [ctypes.c_int, _c_int_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int64, ctypes.c_int, ctypes.c_int]
        L.orc_split_step.restype = ctypes.c_int
        L.orc_sweep_axis.argtypes = [ctypes.c_int, _c_int_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, ctypes.c_int,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_sweep_axis.restype = ctypes.c_int
        L.orc_fft_axis.argtypes = [ctypes.c_int, _c_int_p, _c_dbl_p, ctypes.c_int, ctypes.c_int]
        L.orc_fft_axis.restype = ctypes.c_int
        L.orc_norm.argtypes = [ctypes.c_int, _c_int_p, _c_dbl_p, _c_dbl_p, _c_dbl_p]
        L.orc_norm.restype = ctypes.c_double
        L.orc_energy.argtypes = [ctypes.c_int, _c_int_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_double, _c_dbl_p]
        L.orc_energy.restype = ctypes.c_int
        L.orc_axis_tables.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _c_dbl_p, _c_dbl_p]
        L.orc_axis_tables.restype = ctypes.c_int
        L.orc_adjoint.argtypes = [ctypes.c_int, _c_int_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int64,
                                  _c_dbl_p, _c_dbl_p, _c_dbl_p]
        L.orc_adjoint.restype = ctypes.c_int
        L.orc_overlap_loss.argtypes = [ctypes.c_int, _c_int_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int64]
        L.orc_overlap_loss.restype = ctypes.c_double
        L.orc_num_threads.argtypes = []
        L.orc_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _geom(shape_xfirst, bounds):
    nd = len(shape_xfirst)
    n = (ctypes.c_int * 3)(*(list(shape_xfirst) + [1] * (3 - nd)))
    lo = (ctypes.c_double * 3)(*([float(b[0]) for b in bounds] + [0.0] * (3 - nd)))
    hi = (ctypes.c_double * 3)(*([float(b[1]) for b in bounds] + [1.0] * (3 - nd)))
    return nd, n, lo, hi


def _xfirst(psi):
    return tuple(reversed(psi.shape))


def _cbuf(psi):
    if psi.dtype != np.complex128 or not psi.flags.c_contiguous:
        raise TypeError("oracle arrays must be C-contiguous complex128")
    return psi.ctypes.data_as(_c_dbl_p)


def _vbuf(V, psi):
    if V is None:
        return None, None
    V = np.ascontiguousarray(V, dtype=np.float64)
    if V.shape != psi.shape:
        raise ValueError("V must have the shape of psi")
    return V, V.ctypes.data_as(_c_dbl_p)


def split_step(psi, bounds, *, V=None, coeffD=-0.5, beta=1.0, Omega=0.0, dt=1e-3,
               nsteps=1, order=2, imag=False):
    """Advance ``psi`` in place by ``nsteps`` Split1/Split2 steps
    (splitting.jl:78-82,149-209); ``imag`` selects the imaginary-time variant."""
    nd, n, lo, hi = _geom(_xfirst(psi), bounds)
    Vk, vp = _vbuf(V, psi)
    rc = lib().orc_split_step(nd, n, lo, hi, _cbuf(psi), vp, coeffD, beta, Omega, dt,
                              int(nsteps), int(order), int(bool(imag)))
    if rc != 0:
        raise ValueError(f"orc_split_step failed ({rc})")
    return psi


def sweep_axis(psi, bounds, axis, *, coeffD=-0.5, Omega=0.0, tau=1e-3, imag=False):
    nd, n, lo, hi = _geom(_xfirst(psi), bounds)
    rc = lib().orc_sweep_axis(nd, n, lo, hi, _cbuf(psi), int(axis), coeffD, Omega, tau, int(bool(imag)))
    if rc != 0:
        raise ValueError(f"orc_sweep_axis failed ({rc})")
    return psi


def fft_axis(psi, axis, sign):
    nd, n, _, _ = _geom(_xfirst(psi), [(0, 1)] * psi.ndim)
    rc = lib().orc_fft_axis(nd, n, _cbuf(psi), int(axis), int(sign))
    if rc != 0:
        raise ValueError(f"orc_fft_axis failed ({rc})")
    return psi


def norm(psi, bounds):
    nd, n, lo, hi = _geom(_xfirst(psi), bounds)
    return lib().orc_norm(nd, n, lo, hi, _cbuf(psi))


def energy(psi, bounds, *, V=None, coeffD=-0.5, beta=1.0, Omega=0.0):
    """(EOm, EDelta, Ebeta, E, mu) as in GrossPitaevskii.jl:61-136."""
    nd, n, lo, hi = _geom(_xfirst(psi), bounds)
    Vk, vp = _vbuf(V, psi)
    out = (ctypes.c_double * 5)()
    rc = lib().orc_energy(nd, n, lo, hi, _cbuf(psi), vp, coeffD, beta, Omega, out)
    if rc != 0:
        raise ValueError(f"orc_energy failed ({rc})")
    return tuple(out)


def axis_tables(n, lo, hi):
    """Coordinates (Grids.jl:78-82) and wavenumbers (Plans.jl:829) of one axis."""
    c = np.empty(n)
    k = np.empty(n)
    lib().orc_axis_tables(int(n), float(lo), float(hi), c.ctypes.data_as(_c_dbl_p), k.ctypes.data_as(_c_dbl_p))
    return c, k


def adjoint(psi0, target, bounds, *, V=None, coeffD=-0.5, beta=1.0, Omega=0.0, dt=1e-3, nsteps=1):
    """(loss, dL/dV field, dL/dbeta) of L = 1 - |<target|psi_N>|^2 after nsteps
    real-time Strang steps (reverse sweep, all states stored)."""
    psi0 = np.ascontiguousarray(psi0, dtype=np.complex128)
    target = np.ascontiguousarray(target, dtype=np.complex128)
    nd, n, lo, hi = _geom(_xfirst(psi0), bounds)
    Vk, vp = _vbuf(V, psi0)
    gV = np.zeros(psi0.shape)
    gb = ctypes.c_double()
    loss = ctypes.c_double()
    rc = lib().orc_adjoint(nd, n, lo, hi, _cbuf(psi0), vp, _cbuf(target), coeffD, beta, Omega, dt, int(nsteps),
                           gV.ctypes.data_as(_c_dbl_p), ctypes.byref(gb), ctypes.byref(loss))
    if rc != 0:
        raise ValueError(f"orc_adjoint failed ({rc})")
    return loss.value, gV, gb.value


def overlap_loss(psi0, target, bounds, *, V=None, coeffD=-0.5, beta=1.0, Omega=0.0, dt=1e-3, nsteps=1):
    psi0 = np.ascontiguousarray(psi0, dtype=np.complex128)
    target = np.ascontiguousarray(target, dtype=np.complex128)
    nd, n, lo, hi = _geom(_xfirst(psi0), bounds)
    Vk, vp = _vbuf(V, psi0)
    return lib().orc_overlap_loss(nd, n, lo, hi, _cbuf(psi0), vp, _cbuf(target), coeffD, beta, Omega, dt,
                                  int(nsteps))


def num_threads() -> int:
    return lib().orc_num_threads()
