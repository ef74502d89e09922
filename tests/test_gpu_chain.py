"""GPU checks of chained Strang steps (splitting.jl:149-154 applies x, y, z,
N, x, y, z per step; every flow commutes with every other at Omega = 0, so
only rounding differs):
- real time: a call of several steps runs y(τ), z(τ), x(τ) + N, then y(Δt),
  z(Δt), x(Δt) + N per further step, then y(τ), z(τ), x(τ): the half-step
  flows that meet between two nonlinear phases are one round trip per axis,
  and the phase is the epilogue of the x round trip;
- imaginary time (the norm of the state between two steps is needed): x, y,
  [z N z], y, [x x], y, [z N z], ..., y, x, where [x x] is one pass holding
  the x round trip that ends a step and the one that starts the next.

Bars: FP64 relative L2 <= 1e-12 against the oracle and <= 1e-13 against the
same steps taken one call at a time; FP32 <= 2e-5 against the oracle.
"""
import ctypes

import numpy as np
import pytest

from helpers import PRECS, random_potential, rel_l2, smooth_state

pytestmark = pytest.mark.gpu

SLOTS = 6  # x, y, z round trips, [d N d], other sweeps, [x x]


def _model(sfd, psi0, bounds, V, prec, Omega=0.0, imag=False, dt=2e-3, beta=3.0):
    FT, compute = PRECS[prec]
    g = sfd.Grid(tuple(reversed(psi0.shape)), bounds, FT=FT, compute=compute)
    f = sfd.Field(g, sfd.ComplexField())
    f.set(psi0)
    pot = sfd.PotentialArray(f, V) if V is not None else sfd.PotentialZero(f)
    p = sfd.GrossPitaevskiiParameters(pot=pot, coeffΔ=-0.5, β=beta, Ω=Omega)
    if imag:
        return f, sfd.NumModelSplitImag(f, p, dt, 1, 1, order=2)
    return f, sfd.NumModelSplit2(f, p, dt, 1, 1)


def _launch_counts(sfd, f, run):
    lib, ctx = sfd.load(), f.ctx()
    sfd._lib.check(lib.sfd_ctx_time_sweeps(ctx.h, 1))
    sfd._lib.check(lib.sfd_ctx_sweep_times(ctx.h, None, None, 1))
    run()
    cnt = (ctypes.c_int64 * SLOTS)()
    sfd._lib.check(lib.sfd_ctx_sweep_times(ctx.h, None, cnt, 1))
    sfd._lib.check(lib.sfd_ctx_time_sweeps(ctx.h, 0))
    return list(cnt)


CASES = [
    # shape (x-first), bounds, with V
    ((64, 32), ((-6, 6), (-3, 3)), True),
    ((256, 16), ((-8, 8), (-2, 2)), False),
    ((32, 16, 64), ((-4, 4), (-3, 3), (-5, 5)), True),
    ((64, 64, 64), ((0, 2 * np.pi),) * 3, False),
    ((512, 8, 4), ((-8, 8), (-1, 1), (-1, 1)), True),
    ((48, 16, 32), ((-4, 4), (-3, 3), (-5, 5)), True),   # Bluestein x
    ((24, 40), ((-4, 4), (-5, 5)), False),                # Bluestein x and y
]


@pytest.mark.parametrize("prec", ["f64", "f32", "f32x"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])))
def test_chained_steps_match_oracle_and_single_steps(gpu, orc, case, prec):
    shape_x, bounds, withV = case
    shape = tuple(reversed(shape_x))
    psi0 = smooth_state(shape_x, bounds, seed=7 + shape_x[0])
    V = random_potential(shape_x, bounds) if withV else None
    K = 12
    ref = orc.split_step(psi0.copy(), bounds, V=V, coeffD=-0.5, beta=3.0, Omega=0.0, dt=2e-3, nsteps=K, order=2)
    f, m = _model(gpu, psi0, bounds, V, prec)
    counts = _launch_counts(gpu, f, lambda: m.steps(K))
    chained = f.to_numpy().astype(np.complex128)
    D = len(shape)
    # one round trip per axis and step, plus the closing ones; no double
    # round trips
    assert counts[3] == 0 and counts[5] == 0, counts
    assert counts[:D] == [K + 1] * D, counts
    f.set(psi0)
    for _ in range(K):
        assert gpu.timeStep(m) == 2
    single = f.to_numpy().astype(np.complex128)
    e_ref, e_single = rel_l2(chained, ref), rel_l2(chained, single)
    print(f"{shape_x} {prec}: chained vs oracle {e_ref:.3e}, vs single steps {e_single:.3e}")
    if prec == "f64":
        assert e_ref < 1e-12
        assert e_single < 1e-13
        assert rel_l2(single, ref) < 1e-12
    elif prec == "f32x":
        assert e_ref < 2e-6
    else:
        assert e_ref < 2e-5


@pytest.mark.parametrize("kind", ["rotating", "bluestein-x-imaginary", "one-step"])
def test_chain_only_where_every_flow_commutes(gpu, orc, kind):
    """Rotation (x and y flows do not commute), single-step calls and
    imaginary-time steps on a Bluestein x axis (no [x x] on Bluestein lines)
    keep the x, y, [z N z], x, y order; results still match the oracle."""
    shape_x, bounds = ((48, 16, 32), ((-4, 4), (-3, 3), (-5, 5))) if kind.startswith("bluestein") else \
        ((32, 16, 64), ((-4, 4), (-3, 3), (-5, 5)))
    psi0 = smooth_state(shape_x, bounds, seed=3)
    V = random_potential(shape_x, bounds)
    Om = 0.4 if kind == "rotating" else 0.0
    imag = kind.endswith("imaginary")
    K = 1 if kind == "one-step" else 6
    f, m = _model(gpu, psi0, bounds, V, "f64", Omega=Om, imag=imag)
    counts = _launch_counts(gpu, f, lambda: m.steps(K))
    assert counts[5] == 0, counts
    assert counts[3] == K, counts
    assert counts[0] == 2 * K, counts
    out = f.to_numpy().astype(np.complex128)
    if imag:
        ref = orc.split_step(psi0.copy(), bounds, V=V, coeffD=-0.5, beta=3.0, Omega=0.0, dt=2e-3, nsteps=K,
                             order=2, imag=True)
    else:
        ref = orc.split_step(psi0.copy(), bounds, V=V, coeffD=-0.5, beta=3.0, Omega=Om, dt=2e-3, nsteps=K,
                             order=2)
    err = rel_l2(out, ref)
    print(f"{kind}: rel L2 vs oracle {err:.3e}")
    assert err < 1e-12


@pytest.mark.parametrize("imag", [False, True], ids=["real", "imaginary"])
def test_every_x_variant_runs_the_chained_steps(gpu, orc, imag):
    """Each compiled x-axis variant gives the oracle's result over a chained
    call, 2D and 3D: the merged x(Δt) round trips in real time, the [x x]
    pass (which takes the x variant's shape) in imaginary time."""
    for shape_x, bounds in (((256, 16), ((-8, 8), (-2, 2))), ((128, 8, 16), ((-6, 6), (-2, 2), (-3, 3)))):
        psi0 = smooth_state(shape_x, bounds, seed=11)
        V = random_potential(shape_x, bounds)
        ref = orc.split_step(psi0.copy(), bounds, V=V, coeffD=-0.5, beta=3.0, Omega=0.0, dt=2e-3, nsteps=5,
                             order=2, imag=imag)
        for prec in ("f64", "f32"):
            f, m = _model(gpu, psi0, bounds, V, prec, imag=imag)
            ctx = f.ctx()
            for v in range(ctx.variant(0)[1]):
                ctx.set_variant(0, v)
                f.set(psi0)
                m.steps(5)
                err = rel_l2(f.to_numpy().astype(np.complex128), ref)
                print(f"{shape_x} {prec} x variant {v} {ctx.variant(0)[2]}: {err:.3e}")
                assert err < (1e-12 if prec == "f64" else 2e-5)



@pytest.mark.parametrize("prec", ["f64", "f32", "f32x"])
@pytest.mark.parametrize("case", [CASES[0], CASES[2]], ids=["64x32", "32x16x64"])
def test_chained_imaginary_time_steps(gpu, orc, case, prec):
    """Imaginary time (normalize! after every step) chains too: [x x] takes
    the norm partials of the state that ends a step, between its round trips,
    and the sweep after it applies the renormalisation (a scalar commutes with
    the next step's x flow).  Against the oracle's imaginary-time steps and
    the same steps taken one call at a time."""
    shape_x, bounds, withV = case
    psi0 = smooth_state(shape_x, bounds, seed=13)
    V = random_potential(shape_x, bounds)
    K = 9
    ref = orc.split_step(psi0.copy(), bounds, V=V, coeffD=-0.5, beta=3.0, Omega=0.0, dt=2e-3, nsteps=K, order=2,
                         imag=True)
    f, m = _model(gpu, psi0, bounds, V, prec, imag=True)
    counts = _launch_counts(gpu, f, lambda: m.steps(K))
    assert counts[5] == K - 1 and counts[3] == K and counts[0] == 2, counts
    chained = f.to_numpy().astype(np.complex128)
    f.set(psi0)
    for _ in range(K):
        m.steps(1)
    single = f.to_numpy().astype(np.complex128)
    e_ref, e_single = rel_l2(chained, ref), rel_l2(chained, single)
    print(f"imag {shape_x} {prec}: chained vs oracle {e_ref:.3e}, vs single steps {e_single:.3e}")
    if prec == "f64":
        assert e_ref < 1e-12 and e_single < 1e-13
    elif prec == "f32x":
        assert e_ref < 2e-6
    else:
        assert e_ref < 2e-5
