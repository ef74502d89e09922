"""Imaginary-time ground states of the HIP split step against the fixed
point of the reference's own imaginary-time model, NumModelBackwardEuler
(backward-euler.jl:72-90, restated in oracle/backward_euler.py), on the same
grids: the Thomas-Fermi case of runtests.jl:614-642 and a rotating
anisotropic trap.

The split's fixed point depends on dt: its energy converges to the
backward-Euler energy at second order, the state and mu at first order, and
the Richardson combination 2 psi(dt/2) - psi(dt) is far closer than either.
With rotation, the state of the opposite Omega is far away, so the sign of
the imaginary-time rotation multipliers is pinned too."""
import numpy as np
import pytest

from helpers import PRECS

pytestmark = pytest.mark.gpu

be = pytest.importorskip("oracle.backward_euler")

# the HIP sweeps take power-of-two line lengths: 64² instead of the 48² of
# runtests.jl:614-642, on the same box
CASES = {
    "thomas-fermi": ((64, 64), ((-10, 10), (-10, 10)), (1.0, 1.0), 400.0, 0.0),
    "rotating": ((64, 64), ((-8, 8), (-8, 8)), (1.0, 2.25), 30.0, 0.7),
}
_BE = {}


def _setup(case):
    shape, bounds, (gx, gy), beta, Om = CASES[case]
    g = be.Grid(shape, bounds)
    x, y = g.bcast(g.coord[0], 0), g.bcast(g.coord[1], 1)
    full = tuple(reversed(shape))
    V = np.ascontiguousarray(np.broadcast_to(0.5 * (gx * x ** 2 + gy * y ** 2), full))
    phi0 = np.ascontiguousarray(np.broadcast_to(np.exp(-0.5 * (x ** 2 + y ** 2)) * (1 + 0.1 * x) + 0j, full))
    if case not in _BE:
        gs, _, change = be.ground_state(shape, bounds, phi0, V, beta=beta, Omega=Om, dt=0.5)
        assert change < 1e-11
        _BE[case] = gs
    return g, V, phi0, _BE[case]


def _hip_ground_state(sfd, case, prec, dt, V, phi0, tol):
    shape, bounds, _, beta, Om = CASES[case]
    FT, compute = PRECS[prec]
    grid = sfd.Grid(shape, bounds, FT=FT, compute=compute)
    f = sfd.Field(grid)
    f.set(phi0)
    sfd.normalize(f)
    p = sfd.GrossPitaevskiiParameters(pot=sfd.PotentialArray(f, V), coeffΔ=-0.5, β=beta, Ω=Om)
    m = sfd.NumModelSplitImag(f, p, dt, 1, 1, order=2)
    g = be.Grid(shape, bounds)
    prev = f.to_numpy().astype(np.complex128)
    for _ in range(400):
        m.steps(100)
        cur = f.to_numpy().astype(np.complex128)
        if be.norm(g, cur - prev) / 100 < tol:
            return cur, sfd.energy_all(m)
        prev = cur
    raise AssertionError(f"{case} {prec} dt={dt}: no convergence")


@pytest.mark.parametrize("prec", ["f64", "f32x"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_split_ground_state_approaches_backward_euler(gpu, orc, case, prec):
    shape, bounds, _, beta, Om = CASES[case]
    g, V, phi0, gs = _setup(case)
    E = orc.energy(gs, bounds, V=V, beta=beta, Omega=Om)

    def aligned(p):
        ov = np.vdot(gs, p)
        return p * abs(ov) / ov

    tol = 1e-12 if prec == "f64" else 1e-9
    states, dE, dmu = [], [], []
    for dt in (0.02, 0.01):
        psi, Eg = _hip_ground_state(gpu, case, prec, dt, V, phi0, tol)
        states.append(aligned(psi))
        dE.append(Eg[3] - E[3])
        dmu.append(Eg[4] - E[4])
    err = [be.norm(g, s - gs) for s in states]
    rich = 2 * states[1] - states[0]
    rich /= be.norm(g, rich)
    erich = be.norm(g, rich - gs)
    print(f"{case} {prec}: state {err}, richardson {erich:.2e}, dE {dE}, dmu {dmu}")
    assert 1.8 < err[0] / err[1] < 2.2
    assert 1.8 < dmu[0] / dmu[1] < 2.2
    assert 3.6 < dE[0] / dE[1] < 4.4 and dE[1] > 0
    assert erich < err[1] / 20
    if Om:
        flipped, _, _ = be.ground_state(shape, bounds, phi0, V, beta=beta, Omega=-Om, dt=0.5)
        assert be.norm(g, aligned(flipped) - gs) > 100 * err[1]


def test_split_ground_state_matches_the_oracle_split(gpu, orc):
    """FP64: the HIP fixed point at dt = 0.02 is the C oracle's, run the same
    way (parity of the imaginary-time path over thousands of steps)."""
    case = "rotating"
    shape, bounds, _, beta, Om = CASES[case]
    g, V, phi0, _ = _setup(case)
    psi_gpu, _ = _hip_ground_state(gpu, case, "f64", 0.02, V, phi0, 1e-12)
    psi = phi0 / be.norm(g, phi0)
    for _ in range(400):
        prev = psi.copy()
        orc.split_step(psi, bounds, V=V, beta=beta, Omega=Om, dt=0.02, nsteps=100, order=2, imag=True)
        if be.norm(g, psi - prev) / 100 < 1e-12:
            break
    assert be.norm(g, psi_gpu - psi) < 1e-9
