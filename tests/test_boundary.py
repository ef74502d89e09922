"""The C-ABI boundary without a GPU: the library loads, exports every symbol
include/sfd.h declares, and rejects bad arguments before touching HIP (and
the checkpoint library exports what include/sfd_io.h declares)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header="sfd.h"):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sfd_[a-z0-9_]+)\s*\(", txt)))


def test_header_matches_binding_table():
    import sfdhip
    assert sorted(sfdhip.EXPORTS) == _declared()


def test_library_exports_every_declared_symbol():
    import sfdhip
    lib = sfdhip.load()
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.sfd_version() == 101
    assert lib.sfd_last_error() is not None


def test_checkpoint_library_exports_every_declared_symbol():
    """include/sfd_io.h is served by libsfdio.so (built when the HDF5 C
    library headers are present)."""
    from sfdhip import io
    if not os.path.exists(io.IO_PATH):
        pytest.skip("libsfdio.so not built (no HDF5 headers at build time)")
    lib = io.load_io()
    names = _declared("sfd_io.h")
    assert names == ["sfd_io_last_error", "sfd_io_read", "sfd_io_write"]
    for name in names:
        assert hasattr(lib, name), name
    assert lib.sfd_io_write(None, 2, None, 8, None) == -1
    assert b"null" in lib.sfd_io_last_error()


def _create(lib, ndim, n, lo, hi, prec):
    h = ctypes.c_void_p()
    nn = (ctypes.c_int64 * 3)(*n)
    a = (ctypes.c_double * 3)(*lo)
    b = (ctypes.c_double * 3)(*hi)
    return lib.sfd_ctx_create(ctypes.byref(h), ndim, nn, a, b, prec, None), h


@pytest.mark.parametrize("ndim,n,lo,hi,prec,code", [
    (4, (8, 8, 8), (0, 0, 0), (1, 1, 1), 2, -2),      # ndim out of range
    (1, (8, 8, 8), (0, 0, 0), (1, 1, 1), 2, -2),
    (2, (1025, 8, 1), (0, 0, 0), (1, 1, 1), 2, -2),   # not a power of two, above 1024
    (3, (8, 8, 4096), (0, 0, 0), (1, 1, 1), 1, -2),   # power of two above 2048
    (3, (8, 8, 1), (0, 0, 0), (1, 1, 1), 1, -2),      # below 2
    (2, (8, 8, 1), (0, 1, 0), (1, 1, 1), 2, -3),      # empty axis (Grids.jl:75)
    (2, (8, 8, 1), (0, 0, 0), (1, 1, 1), 7, -3),      # bad precision
])
def test_ctx_create_rejects(ndim, n, lo, hi, prec, code):
    import sfdhip
    lib = sfdhip.load()
    rc, h = _create(lib, ndim, n, lo, hi, prec)
    assert rc == code
    assert not h.value
    assert lib.sfd_last_error()


def test_null_arguments():
    import sfdhip
    lib = sfdhip.load()
    assert lib.sfd_ctx_create(None, 2, None, None, None, 2, None) == -1
    assert lib.sfd_split_step(None, None, None, -0.5, 1.0, 0.0, 1e-3, 1, 2, 0) == -1
    assert lib.sfd_norm2(None, None, None) == -1
    assert lib.sfd_energy(None, None, None, -0.5, 1.0, 0.0, None) == -1
    assert lib.sfd_fft_axis(None, None, None, 0, -1) == -1
    assert lib.sfd_autotune(None, None, None) == -1
    assert lib.sfd_adjoint(None, None, None, None, -0.5, 1.0, 0.0, 1e-3, 1, 0, None, None, None, None) == -1
    assert lib.sfd_inner(None, None, None, None) == -1
    assert lib.sfd_project(None, None, None, 1, None) == -1
    assert lib.sfd_batch_create(None, 1, None, None, 2, None) == -1
    assert lib.sfd_batch_imag_run(None, None, None, None, -0.5, 1e-3, 1, 0.0, 1, 0, None, None) == -1
    assert lib.sfd_batch_autotune(None, None) == -1
    assert lib.sfd_batch_destroy(None) in (0, -1)
    assert lib.sfd_ctx_destroy(None) == 0
    assert b"null" in lib.sfd_last_error()


def test_error_mapping_to_exceptions():
    import sfdhip
    from sfdhip import _lib as L
    lib = sfdhip.load()
    rc, _ = _create(lib, 2, (1025, 8, 1), (0, 0, 0), (1, 1, 1), 2)
    with pytest.raises(sfdhip.ShapeError):
        L.check(rc)
    rc, _ = _create(lib, 2, (8, 8, 1), (0, 0, 0), (1, 1, 1), 9)
    with pytest.raises(sfdhip.ArgError):
        L.check(rc)


def test_grid_validation():
    import sfdhip
    with pytest.raises(sfdhip.ArgError):
        sfdhip.Grid((0, 8), ((0, 1), (0, 1)))
    with pytest.raises(sfdhip.ArgError):
        sfdhip.Grid((8, 8), ((1, 1), (0, 1)))
    with pytest.raises(sfdhip.ArgError):
        sfdhip.Grid((8, 8), ((0, 1), (0, 1)), FT=np.int32)
    g = sfdhip.Grid((32, 16), ((-4, 4), (-1, 1)))
    assert g.Δx == 0.25 and g.Δy == 0.125
    assert g.shape == (16, 32)
    assert g.x[0] == -4.0 and g.x[-1] == 3.75


def test_product_path_has_no_cpu_fallback():
    """Without a HIP device the field constructor raises instead of running
    anything on the CPU."""
    import torch
    import sfdhip
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(sfdhip.SfdError):
        sfdhip.Field(sfdhip.Grid((8, 8), ((0, 1), (0, 1))))
    pkg = os.path.dirname(sfdhip.__file__)
    for fn in os.listdir(pkg):
        if fn.endswith(".py"):
            assert "oracle" not in open(os.path.join(pkg, fn)).read(), fn
