"""Pins of the backward-Euler oracle (oracle/backward_euler.py, the
restatement of NumModelBackwardEuler, backward-euler.jl:72-90 and
krylov.jl:1-33), and the imaginary-time split-step oracle checked against
its fixed point.

Pins of the backward-Euler restatement:
* beta = 0, V = r^2/2: the ground state is the Gaussian with mu = 1 (closed
  form; the spectral discretisation is exact to round-off on this box).
* beta = 0 with an anisotropic trap and rotation: the fixed point is the
  lowest eigenvector of the Hamiltonian assembled as a dense matrix from DFT
  matrices (an independent route to the same discrete operator).
* The reference's own Thomas-Fermi test (runtests.jl:614-642).
* The fixed point does not depend on dt.

Cross-check of the imaginary-time split against the reference's own
imaginary-time model: the imaginary-time Strang split of the C oracle converges to the backward-Euler
ground state as dt -> 0, with energy errors of second order and state and mu
errors of first order (the nonlinear factor uses |psi|^2 after the first
half kinetic flow), and the Richardson combination 2 psi(dt/2) - psi(dt) far
closer.  With rotation, the state for -Omega is far from the one for +Omega,
so this also pins the sign of the imaginary-time rotation multipliers.
"""
import math

import numpy as np
import pytest

from helpers import mesh

be = pytest.importorskip("oracle.backward_euler")


def _grid_fields(shape, bounds, Vf):
    g = be.Grid(shape, bounds)
    x, y = g.bcast(g.coord[0], 0), g.bcast(g.coord[1], 1)
    full = tuple(reversed(shape))
    V = np.ascontiguousarray(np.broadcast_to(Vf(x, y), full))
    phi0 = np.ascontiguousarray(np.broadcast_to(np.exp(-0.5 * (x ** 2 + y ** 2)) * (1 + 0.1 * x) + 0j, full))
    return g, x, y, V, phi0


def test_harmonic_ground_state_closed_form():
    shape, bounds = (48, 48), ((-8, 8), (-8, 8))
    g, x, y, V, _ = _grid_fields(shape, bounds, lambda x, y: 0.5 * (x ** 2 + y ** 2))
    phi0 = np.ascontiguousarray(np.broadcast_to(np.exp(-0.3 * (x - 0.5) ** 2 - 0.2 * y ** 2) + 0j, V.shape))
    phi, steps, change = be.ground_state(shape, bounds, phi0, V, beta=0.0, dt=1.0)
    assert change < 1e-11
    mu, res = be.residual(shape, bounds, phi, V, beta=0.0)
    assert abs(mu - 1.0) < 1e-12 and res < 1e-9
    ex = np.broadcast_to(np.exp(-0.5 * (x ** 2 + y ** 2)) + 0j, V.shape)
    ex = ex / be.norm(g, ex)
    ov = np.vdot(ex, phi)
    assert be.norm(g, phi * abs(ov) / ov - ex) < 1e-9


def test_rotating_fixed_point_is_the_dense_ground_state():
    n, Om = 20, 0.5
    shape, bounds = (n, n), ((-6, 6), (-6, 6))
    g, x, y, V, phi0 = _grid_fields(shape, bounds, lambda x, y: 0.5 * (x ** 2 + 2.25 * y ** 2))
    phi0 = phi0 * (1 + 0.2j * y)
    # H = -lapRot + V as a dense matrix; flattened index iy * n + ix
    F = np.fft.fft(np.eye(n), axis=0)
    Fi = np.linalg.inv(F)
    I = np.eye(n)

    def op1(mult):
        return Fi @ np.diag(mult) @ F

    lap = np.kron(I, op1(-g.xi[0] ** 2)) + np.kron(op1(-g.xi[1] ** 2), I)
    Y, X = np.kron(np.diag(g.coord[1]), I), np.kron(I, np.diag(g.coord[0]))
    Lz_part = Y @ np.kron(I, op1(1j * g.xi[0])) - X @ np.kron(op1(1j * g.xi[1]), I)  # rx + ry
    H = -(0.5 * lap + 1j * Om * Lz_part) + np.diag(V.ravel())
    assert np.abs(H - H.conj().T).max() < 1e-12
    w, U = np.linalg.eigh(H)
    phi, steps, change = be.ground_state(shape, bounds, phi0, V, beta=0.0, Omega=Om, dt=1.0)
    mu, res = be.residual(shape, bounds, phi, V, beta=0.0, Omega=Om)
    assert abs(mu - w[0]) < 1e-10 * abs(w[0])
    u = U[:, 0].reshape(n, n)
    u = u / be.norm(g, u)
    assert abs(np.vdot(u, phi) * g.dv) > 1 - 1e-12


def test_reference_thomas_fermi_case(orc):
    """runtests.jl:614-642 with NumModelBackwardEuler: 48² on [-10,10)²,
    β = 400, Δt = 0.02, 1500 steps from the normalised Gaussian: the energy
    decreases, |E - 2μ_TF/3|/(2μ_TF/3) < 0.03 with μ_TF = √(β/π), and the
    mass stays 1 to 1e-8."""
    N, L, beta, dt = 48, 10.0, 400.0, 0.02
    bounds = ((-L, L), (-L, L))
    X, Y = mesh((N, N), bounds)
    g = be.Grid((N, N), bounds)
    phi = np.ascontiguousarray(np.exp(-0.5 * (X ** 2 + Y ** 2)) + 0j)
    phi /= be.norm(g, phi)
    V = np.ascontiguousarray(np.broadcast_to(0.5 * (X ** 2 + Y ** 2), phi.shape))
    e0 = orc.energy(phi, bounds, V=V, beta=beta)[3]
    for _ in range(1500):
        be.time_step(g, phi, V, beta=beta, dt=dt)
    e1 = orc.energy(phi, bounds, V=V, beta=beta)[3]
    mu_tf = math.sqrt(beta / math.pi)
    mass = float(np.sum(np.abs(phi) ** 2)) * g.dv
    assert e1 < e0
    assert abs(e1 - 2 * mu_tf / 3) / (2 * mu_tf / 3) < 0.03
    assert abs(mass - 1.0) < 1e-8


def test_fixed_point_does_not_depend_on_dt():
    shape, bounds = (32, 32), ((-8, 8), (-8, 8))
    g, x, y, V, phi0 = _grid_fields(shape, bounds, lambda x, y: 0.5 * (x ** 2 + 2.25 * y ** 2))
    a, _, _ = be.ground_state(shape, bounds, phi0, V, beta=30.0, Omega=0.4, dt=0.5)
    b, _, _ = be.ground_state(shape, bounds, phi0, V, beta=30.0, Omega=0.4, dt=2.0)
    ov = np.vdot(a, b)
    assert be.norm(g, b * abs(ov) / ov - a) < 1e-9


def _split_ground_state(orc, psi0, bounds, V, beta, Om, dt, tol=1e-12, chunk=100):
    g = be.Grid(tuple(reversed(psi0.shape)), bounds)
    psi = psi0 / be.norm(g, psi0)
    for _ in range(0, 400000, chunk):
        prev = psi.copy()
        orc.split_step(psi, bounds, V=V, beta=beta, Omega=Om, dt=dt, nsteps=chunk, order=2, imag=True)
        if be.norm(g, psi - prev) / chunk < tol:
            return psi
    raise AssertionError("split imaginary time did not converge")


CROSS = {
    "thomas-fermi": ((48, 48), ((-10, 10), (-10, 10)), lambda x, y: 0.5 * (x ** 2 + y ** 2), 400.0, 0.0),
    "rotating": ((48, 48), ((-8, 8), (-8, 8)), lambda x, y: 0.5 * (x ** 2 + 2.25 * y ** 2), 30.0, 0.7),
}


@pytest.mark.parametrize("case", sorted(CROSS))
def test_split_imag_time_converges_to_backward_euler(orc, case):
    shape, bounds, Vf, beta, Om = CROSS[case]
    g, x, y, V, phi0 = _grid_fields(shape, bounds, Vf)
    gs, _, change = be.ground_state(shape, bounds, phi0, V, beta=beta, Omega=Om, dt=0.5)
    assert change < 1e-11
    E = orc.energy(gs, bounds, V=V, beta=beta, Omega=Om)

    def aligned(p):
        ov = np.vdot(gs, p)
        return p * abs(ov) / ov

    states, dE, dmu = [], [], []
    for dt in (0.04, 0.02, 0.01):
        ps = aligned(_split_ground_state(orc, phi0, bounds, V, beta, Om, dt))
        Es = orc.energy(ps, bounds, V=V, beta=beta, Omega=Om)
        states.append(ps)
        dE.append(Es[3] - E[3])
        dmu.append(Es[4] - E[4])
    err = [be.norm(g, s - gs) for s in states]
    print(case, "state", err, "dE", dE, "dmu", dmu)
    for i in range(2):
        assert 1.8 < err[i] / err[i + 1] < 2.2          # first order in the state
        assert 1.8 < dmu[i] / dmu[i + 1] < 2.2          # and in mu
        assert 3.6 < dE[i] / dE[i + 1] < 4.4            # second order in the energy
    assert all(d > 0 for d in dE)                      # the ground state minimises E
    rich = 2 * states[2] - states[1]
    rich /= be.norm(g, rich)
    assert be.norm(g, rich - gs) < err[2] / 20
    if Om:
        flipped, _, _ = be.ground_state(shape, bounds, phi0, V, beta=beta, Omega=-Om, dt=0.5)
        assert be.norm(g, aligned(flipped) - gs) > 100 * err[2]


def test_bao_du_2d_case_1(orc):
    """Bao & Du 2D case I, the row of benchmarks/validate_expected.jl:127-140
    (β = 200, box [-8,8)×[-4,4), V = ½(x² + 16 y²): γy = 4 squared, the
    paper's convention).  The backward-Euler fixed point on 64² meets the
    literature E = 11.1563 within that file's tolerance 1e-3 (at 2.5e-5)
    and the rms widths (2.2734, 0.6074) within its 5e-3.  Its μ = 16.2972
    is 2.5e-3 below the listed 16.3377: the discrete ground state does not
    reproduce that μ at the file's 1e-3."""
    n = 64
    shape, bounds = (n, n), ((-8, 8), (-4, 4))
    g = be.Grid(shape, bounds)
    x, y = g.bcast(g.coord[0], 0), g.bcast(g.coord[1], 1)
    V = np.ascontiguousarray(np.broadcast_to(0.5 * (x ** 2 + 16.0 * y ** 2), (n, n)))
    phi0 = np.ascontiguousarray(np.broadcast_to(np.exp(-0.5 * (x ** 2 + y ** 2)) + 0j, (n, n)))
    gs, _, change = be.ground_state(shape, bounds, phi0, V, beta=200.0, dt=0.5)
    assert change < 1e-11
    E = orc.energy(gs, bounds, V=V, beta=200.0)
    rho = np.abs(gs) ** 2 * g.dv
    rms = (float(np.sqrt(np.sum(rho * x ** 2))), float(np.sqrt(np.sum(rho * y ** 2))))
    assert abs(E[3] - 11.1563) / 11.1563 < 1e-3
    assert abs(rms[0] - 2.2734) / 2.2734 < 5e-3 and abs(rms[1] - 0.6074) / 0.6074 < 5e-3
    assert abs(E[4] - 16.2972) < 1e-3 and abs(E[4] - 16.3377) / 16.3377 < 3e-3
