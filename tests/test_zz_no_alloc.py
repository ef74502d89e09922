"""sfd_split_step_noalloc through the C ABI against sfd_split_step, the entry
point it mirrors: two calls in a row on one GPU (the second on a smaller grid,
so its scratch holds the first call's leftovers), each compared with
sfd_split_step on a copy of the same input.  Both run the same kernels on
the same tables, so the results are expected bitwise equal; the bar written
here is that of the parity tests (FP64 1e-12, FP32 2e-5 relative L2)."""
import ctypes

import numpy as np
import pytest
import torch

from helpers import random_potential, rel_l2, smooth_state

pytestmark = pytest.mark.gpu

SFD_F32, SFD_F64 = 1, 2


def _ctx(lib, shape_xfirst, bounds, prec):
    nd = len(shape_xfirst)
    n = (ctypes.c_int64 * 3)(*(list(shape_xfirst) + [1] * (3 - nd)))
    lo = (ctypes.c_double * 3)(*([b[0] for b in bounds] + [0.0] * (3 - nd)))
    hi = (ctypes.c_double * 3)(*([b[1] for b in bounds] + [1.0] * (3 - nd)))
    h = ctypes.c_void_p()
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib.sfd_ctx_create(ctypes.byref(h), nd, n, lo, hi, prec, ctypes.c_void_p(stream))
    assert rc == 0, lib.sfd_last_error()
    return h


CASES = [
    # (extents x-first, bounds, precision, Omega, mode (1 = imaginary time), order, steps, with V)
    ((64, 32, 16), ((-4.0, 4.0), (-3.0, 3.0), (-2.0, 2.0)), SFD_F64, 0.3, 1, 2, 7, True),
    ((32, 16), ((-3.0, 3.0), (-2.0, 2.0)), SFD_F32, 0.0, 0, 2, 5, False),
    ((128, 64), ((-5.0, 5.0), (-4.0, 4.0)), SFD_F32, 0.4, 1, 1, 6, True),
    ((16, 8), ((-2.0, 2.0), (-2.0, 2.0)), SFD_F64, 0.0, 0, 1, 3, True),
]


def _run_pair(gpu, lib, case):
    shape, bounds, prec, Om, mode, order, steps, with_v = case
    ctx = _ctx(lib, shape, bounds, prec)
    try:
        cdt = torch.complex128 if prec == SFD_F64 else torch.complex64
        rdt = torch.float64 if prec == SFD_F64 else torch.float32
        psi0 = smooth_state(shape, bounds, seed=len(shape) + steps)
        a = torch.as_tensor(psi0).to(device="cuda", dtype=cdt).contiguous()
        b = a.clone()
        V = torch.as_tensor(random_potential(shape, bounds)).to(device="cuda", dtype=rdt) if with_v else None
        vp = ctypes.c_void_p(V.data_ptr()) if with_v else None
        args = (-0.5, 3.0, Om, 2e-3, steps, order, mode)
        rc = lib.sfd_split_step_noalloc(ctx, ctypes.c_void_p(a.data_ptr()), vp, *args)
        assert rc == 0, lib.sfd_last_error()
        rc = lib.sfd_split_step(ctx, ctypes.c_void_p(b.data_ptr()), vp, *args)
        assert rc == 0, lib.sfd_last_error()
        torch.cuda.synchronize()
        err = rel_l2(a.cpu().numpy(), b.cpu().numpy())
        assert err <= (1e-12 if prec == SFD_F64 else 2e-5), err
        assert rel_l2(a.cpu().numpy(), psi0) > 1e-6  # the call did advance the field
        return err, torch.equal(a, b)
    finally:
        lib.sfd_ctx_destroy(ctx)


@pytest.mark.parametrize("first, second", [(0, 1), (2, 3)], ids=["3d-imag-rot-then-2d", "2d-imag-rot-then-small"])
def test_noalloc_matches_split_step_twice(gpu, first, second):
    lib = gpu.load()
    for i in (first, second):  # the second call's input is smaller than the first's
        err, same = _run_pair(gpu, lib, CASES[i])
        assert same, f"case {i}: rel L2 {err:.3e}"


def test_noalloc_rejects_grids_beyond_its_scratch(gpu):
    """A rotating FP64 1024x1024 grid needs 32 MiB of phase tables: refused with
    SFD_ERR_ARG, the field untouched."""
    lib = gpu.load()
    ctx = _ctx(lib, (1024, 1024), ((-8.0, 8.0), (-8.0, 8.0)), SFD_F64)
    try:
        psi = torch.ones((1024, 1024), dtype=torch.complex128, device="cuda")
        rc = lib.sfd_split_step_noalloc(ctx, ctypes.c_void_p(psi.data_ptr()), None, -0.5, 1.0, 0.5, 1e-3, 1, 2, 0)
        assert rc == -3
        assert b"scratch" in lib.sfd_last_error()
        torch.cuda.synchronize()
        assert bool(torch.all(psi == 1))
        # without rotation the same grid fits
        rc = lib.sfd_split_step_noalloc(ctx, ctypes.c_void_p(psi.data_ptr()), None, -0.5, 1.0, 0.0, 1e-3, 1, 2, 0)
        assert rc == 0, lib.sfd_last_error()
    finally:
        lib.sfd_ctx_destroy(ctx)
