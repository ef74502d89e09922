"""Argument checks of the Python API before any raw device pointer reaches a
kernel: host tensors, wrong dtypes and wrong shapes raise ArgError /
ShapeError (a host pointer would page-fault the GPU, a narrower dtype or a
smaller shape would be read past its end)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture
def model(gpu):
    g = gpu.Grid((16, 8), ((-2.0, 2.0), (-2.0, 2.0)))
    f = gpu.Field(g)
    f.set(np.ones(g.shape, dtype=np.complex128))
    p = gpu.GrossPitaevskiiParameters(pot=gpu.PotentialZero(f), coeffΔ=-0.5, β=1.0)
    return f, gpu.NumModelSplit2(f, p, 1e-3, 1, 1)


def _bad_tensors(f, dtype=None):
    dtype = dtype or f.ϕ.dtype
    shape = f.g.shape
    narrow = torch.complex64 if dtype == torch.complex128 else torch.float32
    return {
        "host": torch.zeros(shape, dtype=dtype),
        "dtype": torch.zeros(shape, dtype=narrow, device="cuda"),
        "shape": torch.zeros((shape[0], shape[1] // 2), dtype=dtype, device="cuda"),
        "strided": torch.zeros((shape[0], 2 * shape[1]), dtype=dtype, device="cuda")[:, ::2],
    }


@pytest.mark.parametrize("kind", ["host", "dtype", "shape", "strided"])
def test_adjoint_target_checked(gpu, model, kind):
    from sfdhip import _lib as L
    f, m = model
    with pytest.raises((L.ArgError, L.ShapeError)):
        gpu.adjoint_gradient(m, _bad_tensors(f)[kind], 3)


@pytest.mark.parametrize("kind", ["host", "dtype", "shape", "strided"])
def test_inner_and_fft_operands_checked(gpu, model, kind):
    from sfdhip import _lib as L
    f, _ = model
    bad = _bad_tensors(f)[kind]
    with pytest.raises((L.ArgError, L.ShapeError)):
        gpu.inner(f, bad, f.ϕ)
    with pytest.raises((L.ArgError, L.ShapeError)):
        gpu.inner(f, f.ϕ, bad)
    with pytest.raises((L.ArgError, L.ShapeError)):
        gpu.mul_x(bad, f, f.ϕ)
    with pytest.raises((L.ArgError, L.ShapeError)):
        gpu.ldiv_y(torch.empty_like(f.ϕ), f, bad)


@pytest.mark.parametrize("kind", ["host", "dtype", "shape", "strided"])
def test_project_operands_checked(gpu, model, kind):
    from sfdhip import _lib as L
    f, _ = model
    good_field = torch.zeros(f.g.shape, dtype=torch.float64, device="cuda")
    good_basis = torch.zeros((3,) + f.g.shape, dtype=torch.float64, device="cuda")
    bad_field = _bad_tensors(f, torch.float64)[kind]
    with pytest.raises((L.ArgError, L.ShapeError)):
        gpu.project(f, bad_field, good_basis)
    bad_basis = {
        "host": good_basis.cpu(),
        "dtype": good_basis.float(),
        "shape": good_basis[:, :, :4].contiguous(),
        "strided": torch.zeros((3,) + (f.g.shape[0], 2 * f.g.shape[1]), dtype=torch.float64, device="cuda")[..., ::2],
    }[kind]
    with pytest.raises((L.ArgError, L.ShapeError)):
        gpu.project(f, good_field, bad_basis)
    assert np.all(gpu.project(f, good_field, good_basis) == 0.0)


def test_valid_operands_pass(gpu, model):
    f, m = model
    t = torch.ones_like(f.ϕ)
    loss, gV, gb = gpu.adjoint_gradient(m, t, 2)
    assert np.isfinite(loss)
    out = torch.empty_like(f.ϕ)
    gpu.mul_x(out, f, f.ϕ)
    assert abs(gpu.inner(f, f.ϕ, f.ϕ) - 16.0) < 1e-12
