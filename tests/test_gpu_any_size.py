"""Grids of any extent: line lengths that are not powers of two run
Bluestein's chirp transform on the power of two m >= 2n - 1 (and 2048-point
lines run as two interleaved 1024-point halves).  The sizes are those of the
reference's own suites: the odd 29/31/33 grids of runtests.jl:351-360, the
48² Thomas-Fermi test of runtests.jl:614-622 and the 96³ Bao & Du case of
benchmarks/validate_expected.jl:128-131; the oracle transforms any n by a
direct DFT.  Bars: FP64 1e-12, F32X 1e-6, FP32 2e-5 relative L2 (measured at
most 3.1e-6 after 10 steps: the chirp multiplies are plain FP32 complex
products, not the exactly invertible lifting rotations of the power-of-two
sweeps, so FP32 Bluestein round trips carry more rounding)."""
import numpy as np
import pytest
import torch

from helpers import PRECS, gpu_run, random_potential, rel_l2, smooth_state

pytestmark = pytest.mark.gpu

TOL = {"f64": 1e-12, "f32": 2e-5, "f32x": 1e-6}

CASES = [
    # shape (x-first), bounds, Omega, with V
    ((48, 48), ((-8, 8), (-8, 8)), 0.0, True),
    ((29, 31), ((-3, 3), (-4, 4)), 0.5, True),
    ((33, 29), ((-4, 4), (-3, 3)), 0.0, False),
    ((29, 31, 33), ((-3, 3), (-3, 3), (-3, 3)), 0.3, True),
    ((3, 5, 7), ((-1, 1), (-1, 1), (-1, 1)), 0.0, True),
    ((100, 12, 6), ((-5, 5), (-2, 2), (-1, 1)), 0.2, False),
    ((1000, 3), ((-10, 10), (-1, 1)), 0.0, True),
    ((2048, 8), ((-10, 10), (-1, 1)), 0.0, True),
    ((8, 2048), ((-1, 1), (-10, 10)), 0.2, False),
    ((24, 2048), ((-2, 2), (-10, 10)), 0.0, True),
    # power-of-two x with Bluestein y / z at Omega = 0: chained steps around
    # Bluestein sweeps (real time: merged flows; imaginary time: [x x])
    ((64, 48), ((-6, 6), (-5, 5)), 0.0, True),
    ((32, 24, 20), ((-4, 4), (-3, 3), (-3, 3)), 0.0, True),
]


def _ids(c):
    return "x".join(map(str, c[0])) if isinstance(c, tuple) else None


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("prec", ["f64", "f32", "f32x"])
@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_split_parity_any_size(gpu, orc, case, prec, order):
    shape, bounds, Om, withV = case
    psi = smooth_state(shape, bounds, seed=len(shape) + shape[0])
    V = random_potential(shape, bounds) if withV else None
    kw = dict(V=V, coeffD=-0.5, beta=3.0, Omega=Om, dt=2e-3, nsteps=10, order=order)
    ref = orc.split_step(psi.copy(), bounds, **kw)
    out = gpu_run(gpu, psi, bounds, prec=prec, **kw)
    err = rel_l2(out, ref)
    print(f"{shape} {prec} order {order}: rel L2 = {err:.3e}")
    assert err < TOL[prec]


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_bao_du_3d_size(gpu, orc, prec):
    """96³ (the grid of the Bao & Du 3D case), V = ½(x² + y² + z²), β = 100,
    3 Strang steps against the oracle."""
    shape, bounds = (96, 96, 96), ((-6, 6),) * 3
    psi = smooth_state(shape, bounds, seed=3)
    from helpers import mesh
    x, y, z = mesh(shape, bounds)
    V = np.ascontiguousarray(np.broadcast_to(0.5 * (x ** 2 + y ** 2 + z ** 2), psi.shape))
    kw = dict(V=V, coeffD=-0.5, beta=100.0, dt=1e-3, nsteps=3, order=2)
    ref = orc.split_step(psi.copy(), bounds, **kw)
    out = gpu_run(gpu, psi, bounds, prec=prec, **kw)
    err = rel_l2(out, ref)
    print(f"96³ {prec}: rel L2 = {err:.3e}")
    assert err < TOL[prec]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("shape", [(29, 31, 33), (2048, 8), (48, 36)], ids=lambda s: "x".join(map(str, s)))
def test_fft_axis_any_size(gpu, prec, shape):
    """mul_d! / ldiv_d! along every axis against numpy.fft (forward
    unnormalised, backward 1/n)."""
    FT, compute = PRECS[prec]
    bounds = tuple((-1.0, 1.0) for _ in shape)
    g = gpu.Grid(shape, bounds, FT=FT, compute=compute)
    f = gpu.Field(g)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(g.shape) + 1j * rng.standard_normal(g.shape)
    f.set(x)
    out = torch.empty_like(f.ϕ)
    nd = len(shape)
    for a in range(nd):
        npax = nd - 1 - a
        (gpu.mul_x, gpu.mul_y, gpu.mul_z)[a](out, f, f.ϕ)
        ref = np.fft.fft(x, axis=npax)
        err = rel_l2(out.cpu().numpy(), ref)
        assert err < (1e-13 if prec == "f64" else 1e-5), (a, err)
        (gpu.ldiv_x, gpu.ldiv_y, gpu.ldiv_z)[a](out, f, f.ϕ)
        ref = np.fft.ifft(x, axis=npax)
        err = rel_l2(out.cpu().numpy(), ref)
        assert err < (1e-13 if prec == "f64" else 1e-5), (a, err)


@pytest.mark.parametrize("case", [CASES[0], CASES[1], CASES[3], CASES[5]], ids=_ids)
def test_energy_any_size(gpu, orc, case):
    shape, bounds, Om, withV = case
    psi = smooth_state(shape, bounds, seed=5)
    V = random_potential(shape, bounds) if withV else None
    ref = orc.energy(psi, bounds, V=V, coeffD=-0.5, beta=7.0, Omega=Om)
    g = gpu.Grid(shape, bounds)
    f = gpu.Field(g)
    f.set(psi)
    pot = gpu.PotentialArray(f, V) if V is not None else gpu.PotentialZero(f)
    m = gpu.NumModelSplit2(f, gpu.GrossPitaevskiiParameters(pot=pot, coeffΔ=-0.5, β=7.0, Ω=Om), 1e-3, 1, 1)
    out = gpu.energy_all(m)
    for k in range(5):
        assert out[k] == pytest.approx(ref[k], rel=1e-12, abs=1e-12 * abs(ref[3])), (k, out, ref)
    # solve!'s loop on a Bluestein grid: energies of each state, then the step
    hist = gpu.steps_with_energy(m, 3)
    assert hist[0][3] == pytest.approx(ref[3], rel=1e-12)


@pytest.mark.parametrize("prec", ["f64", "f32x"])
@pytest.mark.parametrize("case", [CASES[0], CASES[3], CASES[-2], CASES[-1]], ids=_ids)
def test_imag_time_any_size(gpu, orc, case, prec):
    shape, bounds, Om, _ = case
    psi = smooth_state(shape, bounds, seed=9)
    V = random_potential(shape, bounds, amp=10.0)
    kw = dict(V=V, coeffD=-0.5, beta=20.0, Omega=Om, dt=5e-3, nsteps=25, order=2, imag=True)
    ref = orc.split_step(psi.copy(), bounds, **kw)
    out = gpu_run(gpu, psi, bounds, prec=prec, **kw)
    err = rel_l2(out, ref)
    print(f"imag {shape} {prec}: rel L2 = {err:.3e}")
    assert err < TOL[prec]


def test_adjoint_any_size(gpu, orc):
    shape, bounds = (29, 31), ((-3.0, 3.0), (-3.0, 3.0))
    psi0 = smooth_state(shape, bounds, seed=1)
    psi0 = psi0 / orc.norm(psi0, bounds)
    tgt = smooth_state(shape, bounds, seed=8)
    tgt = tgt / orc.norm(tgt, bounds)
    V = random_potential(shape, bounds, amp=3.0)
    loss, gV, gb = orc.adjoint(psi0, tgt, bounds, V=V, coeffD=-0.5, beta=6.0, Omega=0.4, dt=2e-3, nsteps=13)
    g = gpu.Grid(shape, bounds)
    f = gpu.Field(g)
    f.set(psi0)
    p = gpu.GrossPitaevskiiParameters(pot=gpu.PotentialArray(f, V), coeffΔ=-0.5, β=6.0, Ω=0.4)
    m = gpu.NumModelSplit2(f, p, 2e-3, 1, 1)
    l2, gV2, gb2 = gpu.adjoint_gradient(m, torch.as_tensor(tgt).to("cuda"), 13)
    assert l2 == pytest.approx(loss, rel=1e-12, abs=1e-14)
    assert gb2 == pytest.approx(gb, rel=1e-10)
    assert np.sqrt(np.sum((gV2.cpu().numpy() - gV) ** 2) / np.sum(gV ** 2)) < 1e-10


def test_variant_count_and_autotune_on_bluestein_axes(gpu):
    g = gpu.Grid((29, 64), ((-1, 1), (-1, 1)), FT=np.float32)
    f = gpu.Field(g)
    c = f.ctx()
    assert c.variant(0)[1] == 1  # one Bluestein variant on x
    assert c.variant(1)[1] >= 1
    tune = f.autotune()
    assert len(tune[0]) == 1
