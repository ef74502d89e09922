"""GPU parity: the HIP split step (through the C ABI) against the FP64 oracle.

Bars: FP64 relative L2 <= 1e-12 (the GPU=CPU bar of runtests.jl:1438-1452 is
1e-10; this suite holds the HIP path to 1e-12); FP32 relative L2 <= 2e-5
after the same steps (single-precision round-off of ~1e-7 per sweep,
accumulated).
"""
import math

import numpy as np
import pytest

from helpers import PRECS, TOLS, gpu_run, mesh, random_potential, rel_l2, smooth_state

pytestmark = pytest.mark.gpu

F64_TOL = 1e-12
F32_TOL = 2e-5


def test_reference_gpu_case(gpu, orc):
    """runtests.jl:1438-1452: 32² on [-6,6)², ψ0 = e^{-r²/2}(1+0.1x) normalised,
    coeffΔ=-0.5, β=10, Ω=0, quadratic V (γ=1), Δt=0.01, 20 Split2 steps."""
    bounds = ((-6, 6), (-6, 6))
    x, y = mesh((32, 32), bounds)
    psi = np.exp(-0.5 * (x ** 2 + y ** 2)) * (1 + 0.1 * x) + 0j
    psi = np.ascontiguousarray(psi / orc.norm(np.ascontiguousarray(psi + 0j), bounds))
    V = np.ascontiguousarray(np.broadcast_to(0.5 * (1.0 * x ** 2 + 1.0 * y ** 2), psi.shape))
    ref = orc.split_step(psi.copy(), bounds, V=V, beta=10.0, dt=0.01, nsteps=20, order=2)
    out = gpu_run(gpu, psi, bounds, V=V, beta=10.0, dt=0.01, nsteps=20)
    err = rel_l2(out, ref)
    print(f"reference GPU case rel L2 = {err:.3e}")
    assert err < 1e-10
    assert err < F64_TOL


CASES = [
    # shape (x-first), bounds, Omega, with V
    ((16, 16), ((-4, 4), (-4, 4)), 0.0, True),
    ((64, 32), ((-6, 6), (-3, 3)), 0.7, True),
    ((128, 128), ((-8, 8), (-8, 8)), 0.0, False),
    ((256, 8), ((-8, 8), (-1, 1)), 0.3, True),
    ((512, 16), ((-10, 10), (-2, 2)), 0.0, True),
    ((1024, 4), ((-10, 10), (-1, 1)), 0.0, False),
    ((2, 64), ((-1, 1), (-4, 4)), 0.0, True),
    ((8, 8, 8), ((-2, 2), (-2, 2), (-2, 2)), 0.4, True),
    ((32, 16, 64), ((-4, 4), (-3, 3), (-5, 5)), 0.5, True),
    ((16, 64, 32), ((-3, 3), (-6, 6), (-4, 4)), 0.0, False),
    ((64, 64, 64), ((0, 2 * math.pi),) * 3, 0.0, False),
]


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("prec", ["f64", "f32", "f32x"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])) if isinstance(c, tuple) else None)
def test_split_parity(gpu, orc, case, prec, order):
    shape, bounds, Om, withV = case
    psi = smooth_state(shape, bounds, seed=len(shape) + shape[0])
    V = random_potential(shape, bounds) if withV else None
    kw = dict(V=V, coeffD=-0.5, beta=3.0, Omega=Om, dt=2e-3, nsteps=10, order=order)
    ref = orc.split_step(psi.copy(), bounds, **kw)
    out = gpu_run(gpu, psi, bounds, prec=prec, **kw)
    err = rel_l2(out, ref)
    print(f"{shape} {prec} order {order}: rel L2 = {err:.3e}")
    assert err < TOLS[prec]


@pytest.mark.parametrize("prec", ["f64", "f32", "f32x"])
@pytest.mark.parametrize("case", [CASES[1], CASES[8]], ids=["2d", "3d"])
def test_imag_time_parity(gpu, orc, case, prec):
    shape, bounds, Om, _ = case
    psi = smooth_state(shape, bounds, seed=7)
    V = random_potential(shape, bounds, amp=10.0)
    kw = dict(V=V, coeffD=-0.5, beta=20.0, Omega=Om, dt=5e-3, nsteps=25, order=2, imag=True)
    ref = orc.split_step(psi.copy(), bounds, **kw)
    out = gpu_run(gpu, psi, bounds, prec=prec, **kw)
    err = rel_l2(out, ref)
    print(f"imag {shape} {prec}: rel L2 = {err:.3e}")
    assert err < TOLS[prec]


def test_thomas_fermi_1000_steps_f64(gpu, orc):
    """128² on [-8,8)², V = ½(x²+y²), β = 500, Δt = 1e-3, 1000 Split2 steps,
    Thomas-Fermi initial state with a seeded phase noise: FP64 HIP vs the
    oracle at <= 1e-12 relative L2."""
    from sfdhip.workloads import thomas_fermi_2d
    psi, V, bounds = thomas_fermi_2d()
    kw = dict(V=V, beta=500.0, dt=1e-3, nsteps=1000, order=2)
    ref = orc.split_step(psi.copy(), bounds, **kw)
    out = gpu_run(gpu, psi, bounds, **kw)
    err = rel_l2(out, ref)
    print(f"TF 128² 1000 steps FP64: rel L2 = {err:.3e}")
    assert err <= 1e-12


VARIANT_CASES = [
    ((64, 64), ((-6, 6), (-6, 6))),
    ((128, 64, 16), ((-4, 4), (-3, 3), (-2, 2))),
    ((256, 256), ((-8, 8), (-8, 8))),
    ((16, 256, 128), ((-2, 2), (-6, 6), (-5, 5))),
    ((512, 32), ((-8, 8), (-2, 2))),
    ((32, 1024), ((-2, 2), (-10, 10))),
]


@pytest.mark.parametrize("prec", ["f64", "f32", "f32x"])
@pytest.mark.parametrize("case", VARIANT_CASES, ids=lambda c: "x".join(map(str, c[0])))
def test_every_launch_variant(gpu, orc, case, prec):
    """Every compiled sweep variant (line split, workgroup shape) of every
    axis gives the oracle's answer: a Split2 step with V and rotation."""
    shape, bounds = case
    psi = smooth_state(shape, bounds, seed=3)
    V = random_potential(shape, bounds)
    kw = dict(V=V, coeffD=-0.5, beta=3.0, Omega=0.4, dt=2e-3, nsteps=2, order=2)
    ref = orc.split_step(psi.copy(), bounds, **kw)
    FT, compute = PRECS[prec]
    g = gpu.Grid(shape, bounds, FT=FT, compute=compute)
    f = gpu.Field(g)
    pot = gpu.PotentialArray(f, V)
    p = gpu.GrossPitaevskiiParameters(pot=pot, coeffΔ=-0.5, β=3.0, Ω=0.4)
    m = gpu.NumModelSplit2(f, p, 2e-3, 2, 2)
    mi = gpu.NumModelSplitImag(f, p, 2e-3, 2, 2, order=2)
    ref_imag = orc.split_step(psi.copy(), bounds, imag=True, **kw)
    ctx = f.ctx()
    counts = [ctx.variant(a)[1] for a in range(len(shape))]
    for a in range(len(shape)):
        for v in range(counts[a]):
            for b in range(len(shape)):
                ctx.set_variant(b, 0)
            ctx.set_variant(a, v)
            f.set(psi)
            m.steps(2)
            err = rel_l2(f.to_numpy(), ref)
            f.set(psi)
            mi.steps(2)
            erri = rel_l2(f.to_numpy(), ref_imag)
            print(f"{shape} {prec} axis {a} variant {v} {ctx.variant(a)[2]}: {err:.3e} imag {erri:.3e}")
            assert err < TOLS[prec]
            assert erri < TOLS[prec]


# Without rotation every axis has a single phase row, so the pipelined
# (persistent, prefetching) variants run; these shapes leave their tiles
# partly empty: 8 columns for 16-column strided tiles, 4 lines for 8- and
# 16-line contiguous tiles.
PIPE_CASES = [
    ((8, 256, 256), ((-2, 2), (-6, 6), (-5, 5))),
    ((256, 4), ((-8, 8), (-1, 1))),
    ((256, 256, 16), ((-8, 8), (-7, 7), (-2, 2))),
    ((512, 4), ((-8, 8), (-1, 1))),
    ((16, 512), ((-2, 2), (-8, 8))),
]


@pytest.mark.parametrize("prec", ["f64", "f32", "f32x"])
@pytest.mark.parametrize("case", PIPE_CASES, ids=lambda c: "x".join(map(str, c[0])))
def test_every_launch_variant_without_rotation(gpu, orc, case, prec):
    """Every variant of every axis at Ω = 0, real and imaginary time (the
    latter with the FP64 norm partials fused into the last sweep), against
    the oracle, tiles partly empty included."""
    shape, bounds = case
    psi = smooth_state(shape, bounds, seed=5)
    V = random_potential(shape, bounds)
    kw = dict(V=V, coeffD=-0.5, beta=3.0, Omega=0.0, dt=2e-3, nsteps=3, order=2)
    ref = orc.split_step(psi.copy(), bounds, **kw)
    ref_imag = orc.split_step(psi.copy(), bounds, imag=True, **kw)
    FT, compute = PRECS[prec]
    g = gpu.Grid(shape, bounds, FT=FT, compute=compute)
    f = gpu.Field(g)
    p = gpu.GrossPitaevskiiParameters(pot=gpu.PotentialArray(f, V), coeffΔ=-0.5, β=3.0)
    m = gpu.NumModelSplit2(f, p, 2e-3, 3, 3)
    mi = gpu.NumModelSplitImag(f, p, 2e-3, 3, 3, order=2)
    ctx = f.ctx()
    shape_x_first = tuple(reversed(psi.shape))
    for a in range(len(shape)):
        for v in range(ctx.variant(a)[1]):
            for b in range(len(shape)):
                ctx.set_variant(b, 0)
            ctx.set_variant(a, v)
            f.set(psi)
            m.steps(3)
            err = rel_l2(f.to_numpy(), ref)
            f.set(psi)
            mi.steps(3)
            erri = rel_l2(f.to_numpy(), ref_imag)
            print(f"{shape_x_first} {prec} axis {a} variant {v} {ctx.variant(a)[2]}: {err:.3e} imag {erri:.3e}")
            assert err < TOLS[prec]
            assert erri < TOLS[prec]


FOLD_CASES = [
    ((16, 8, 256), ((-2, 2), (-1, 1), (-6, 6))),
    ((8, 16, 512), ((-1, 1), (-2, 2), (-8, 8))),
    ((48, 512), ((-3, 3), (-8, 8))),
]


@pytest.mark.parametrize("prec", ["f64", "f32", "f32x"])
@pytest.mark.parametrize("case", FOLD_CASES, ids=lambda c: "x".join(map(str, c[0])))
def test_every_fold_variant_without_potential(gpu, orc, case, prec):
    """V = 0 (PotentialZero, a null potential): the double round trip
    [d N d] of the innermost axis runs kernels that hold no potential
    registers.  Every variant of that axis, real and imaginary time, against
    the oracle; 8 and 16 columns leave 16- and 32-column tiles partly empty,
    and 48 columns put a Bluestein x axis beside the folded y axis."""
    shape, bounds = case
    psi = smooth_state(shape, bounds, seed=8)
    kw = dict(coeffD=-0.5, beta=3.0, Omega=0.0, dt=2e-3, nsteps=3, order=2)
    ref = orc.split_step(psi.copy(), bounds, **kw)
    ref_imag = orc.split_step(psi.copy(), bounds, imag=True, **kw)
    FT, compute = PRECS[prec]
    f = gpu.Field(gpu.Grid(shape, bounds, FT=FT, compute=compute))
    p = gpu.GrossPitaevskiiParameters(pot=gpu.PotentialZero(f), coeffΔ=-0.5, β=3.0)
    m = gpu.NumModelSplit2(f, p, 2e-3, 3, 3)
    mi = gpu.NumModelSplitImag(f, p, 2e-3, 3, 3, order=2)
    ctx = f.ctx()
    ax = len(shape) - 1
    for v in range(ctx.variant(ax)[1]):
        ctx.set_variant(ax, v)
        f.set(psi)
        m.steps(3)
        err = rel_l2(f.to_numpy(), ref)
        f.set(psi)
        mi.steps(3)
        erri = rel_l2(f.to_numpy(), ref_imag)
        print(f"{shape} {prec} fold axis {ax} variant {v} {ctx.variant(ax)[2]}: {err:.3e} imag {erri:.3e}")
        assert err < TOLS[prec]
        assert erri < TOLS[prec]


def test_autotune_keeps_results(gpu, orc):
    shape, bounds = (128, 64, 32), ((-4, 4), (-3, 3), (-2, 2))
    psi = smooth_state(shape, bounds, seed=4)
    kw = dict(coeffD=-0.5, beta=2.0, dt=2e-3, nsteps=3, order=2)
    ref = orc.split_step(psi.copy(), bounds, **kw)
    g = gpu.Grid(shape, bounds, FT=np.float32)
    f = gpu.Field(g)
    f.set(psi)
    times = f.autotune()
    assert set(times) == {0, 1, 2}
    assert all(len(v) >= 1 and all(t > 0 for t in v) for v in times.values())
    p = gpu.GrossPitaevskiiParameters(pot=gpu.PotentialZero(f), coeffΔ=-0.5, β=2.0)
    gpu.NumModelSplit2(f, p, 2e-3, 3, 3).steps(3)
    assert rel_l2(f.to_numpy(), ref) < F32_TOL
    # the double round trips ([z N z] of one-step calls, [x x] of chained
    # imaginary-time steps) run the variants the tuner chose for them, which
    # may differ from the single round trips'
    m = gpu.NumModelSplit2(f, p, 2e-3, 1, 1)
    f.set(psi)
    for _ in range(3):
        m.steps(1)
    assert rel_l2(f.to_numpy(), ref) < F32_TOL
    ref_imag = orc.split_step(psi.copy(), bounds, imag=True, **kw)
    f.set(psi)
    gpu.NumModelSplitImag(f, p, 2e-3, 3, 3).steps(3)
    assert rel_l2(f.to_numpy(), ref_imag) < F32_TOL
