"""Pins of the FP64 CPU oracle (oracle/sfd_oracle.c).

Each check is one the mathematics or the reference's own tests fix, chosen so
that a plausible slip (sign of a phase, wrong wavenumber, coordinate instead
of index, dropped normalisation, wrong axis order) fails at least one:

* FFT against numpy/pocketfft and a brute-force DFT, all axes, odd sizes.
* Golden value of the reference's norm doctest (Fields.jl:166-170).
* norm / normalize! (runtests.jl:718-725).
* Plane-wave energies in closed form (runtests.jl:29-90 pins dx = i k ψ,
  rx = y dx, ry = -x dy; the energy sums follow in closed form).
* Kinetic flow exact for plane waves incl. the Nyquist mode (runtests.jl:525-557).
* Nonlinear flow exact on a uniform state.
* Split1/Split2 orders and mass conservation with the exact setup of
  runtests.jl:644-671.
* Strang time reversibility.
* Rotation: an L_z = +1 state evolves by e^{iΩt} times free evolution.
* Imaginary time reaches the Thomas-Fermi energy of runtests.jl:614-642.
"""
import math

import numpy as np
import pytest

from helpers import mesh, random_potential, rel_l2, smooth_state


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 3, 5, 6, 12, 24, 29, 31, 33, 48])
def test_fft_line_vs_numpy(orc, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))
    a = np.ascontiguousarray(a)
    f = orc.fft_axis(a.copy(), 0, -1)
    assert rel_l2(f, np.fft.fft(a, axis=1)) < 1e-13
    b = orc.fft_axis(a.copy(), 0, +1)
    assert rel_l2(b, np.fft.ifft(a, axis=1)) < 1e-13


@pytest.mark.parametrize("n", [2, 3, 8, 12, 16])
def test_fft_line_vs_brute_force_dft(orc, n):
    rng = np.random.default_rng(100 + n)
    a = np.ascontiguousarray(rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n)))
    j = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(j, j) / n)
    assert rel_l2(orc.fft_axis(a.copy(), 0, -1), a @ W.T) < 1e-14


@pytest.mark.parametrize("shape", [(16, 8, 4), (5, 12, 7), (32, 24, 40)])
def test_fft_axes_3d(orc, shape):
    rng = np.random.default_rng(7)
    a = np.ascontiguousarray(rng.standard_normal(tuple(reversed(shape))) + 1j * rng.standard_normal(tuple(reversed(shape))))
    for ax in range(3):
        npax = 2 - ax
        assert rel_l2(orc.fft_axis(a.copy(), ax, -1), np.fft.fft(a, axis=npax)) < 1e-13
        back = orc.fft_axis(orc.fft_axis(a.copy(), ax, -1), ax, +1)
        assert rel_l2(back, a) < 1e-14


# --------------------------------------------------------- geometry / norm
def test_coordinates_linrange(orc):
    c, k = orc.axis_tables(8, -2.0, 2.0)
    assert np.array_equal(c, np.array([-2.0, -1.5, -1.0, -0.5, 0.0, 0.5, 1.0, 1.5]))
    # fftfreq(8, 2π/0.5) = (2π/4)·[0,1,2,3,-4,-3,-2,-1]
    assert np.allclose(k, (2 * np.pi / 4) * np.array([0, 1, 2, 3, -4, -3, -2, -1]), rtol=0, atol=1e-15)
    c, k = orc.axis_tables(5, 0.0, 1.0)
    assert np.allclose(k, 2 * np.pi * np.array([0, 1, 2, -2, -1]), rtol=1e-15)


def test_norm_golden_3d(orc):
    """Fields.jl:166-170 (DocTestSetup of docs/make.jl:11-12)."""
    bounds = ((-12, 12),) * 3
    x, y, z = mesh((32, 32, 32), bounds)
    phi = (-12 - x) * (12 + x) + (-12 - y) * (12 + y) + (-12 - z) * (12 + z) + 0j
    phi = np.ascontiguousarray(np.broadcast_to(phi, (32, 32, 32)))
    assert orc.norm(phi, bounds) == pytest.approx(72951.55177239206, rel=1e-14)


def test_norm_constant(orc):
    """runtests.jl:718-725."""
    bounds = ((-12, 12), (-12, 12))
    phi = np.ones((64, 64), dtype=np.complex128)
    assert orc.norm(phi, bounds) == pytest.approx(24.0, rel=1e-14)


# ---------------------------------------------------------------- energy
@pytest.mark.parametrize("dims", [2, 3])
def test_energy_plane_wave_closed_form(orc, dims):
    """ψ = e^{i k·x}: |∇ψ|² = |k|², rx = y·i kx ψ, ry = -x·i ky ψ, so
    EΔ = -cD |k|² |box|, Eβ = β/2 |box|, EΩ = Σ Re(i ψ̄ Ω(rx+ry)) dV
       = -Ω Σ (y kx - x ky) dV."""
    n = (64, 64) if dims == 2 else (48, 48, 48)
    bounds = ((-12, 12),) * dims
    X = mesh(n, bounds)
    L = 24.0
    kx, ky = 2 * math.pi * 2 / L, 2 * math.pi * 3 / L
    kz = 2 * math.pi * 4 / L
    ph = kx * X[0] + ky * X[1] + (kz * X[2] if dims == 3 else 0)
    psi = np.ascontiguousarray(np.exp(1j * ph))
    cD, beta, Om = -0.5, 3.0, 0.7
    EOm, ED, Eb, E, mu = orc.energy(psi, bounds, coeffD=cD, beta=beta, Omega=Om)
    vol = L ** dims
    dv = (L / n[0]) ** dims
    k2 = kx ** 2 + ky ** 2 + (kz ** 2 if dims == 3 else 0)
    assert ED == pytest.approx(-cD * k2 * vol, rel=1e-10)
    assert Eb == pytest.approx(0.5 * beta * vol, rel=1e-10)
    x_full = np.broadcast_to(X[0], psi.shape)
    y_full = np.broadcast_to(X[1], psi.shape)
    EOm_exact = -Om * float(np.sum(y_full * kx - x_full * ky)) * dv
    assert abs(EOm - EOm_exact) <= 1e-10 * abs(ED)
    assert E == pytest.approx(-EOm + ED + Eb, rel=1e-15)
    assert mu == pytest.approx(-EOm + ED + 2 * Eb, rel=1e-15)


def test_energy_potential_term(orc):
    bounds = ((-4, 4), (-4, 4))
    psi = smooth_state((32, 32), bounds, seed=3)
    V = random_potential((32, 32), bounds)
    _, ED0, _, _, _ = orc.energy(psi, bounds, coeffD=-0.5)
    _, ED1, _, _, _ = orc.energy(psi, bounds, V=V, coeffD=-0.5)
    dv = (8 / 32) ** 2
    assert ED1 - ED0 == pytest.approx(float(np.sum(V * np.abs(psi) ** 2)) * dv, rel=1e-12)


# ----------------------------------------------------------- exact flows
@pytest.mark.parametrize("order", [1, 2])
def test_kinetic_flow_exact_plane_waves(orc, order):
    """β = V = Ω = 0: each Fourier mode picks up e^{i cD |k|² t}; both splits
    are exact for any Δt.  Includes the Nyquist mode (ξ = -n/2 · 2π/L)."""
    n = (32, 16)
    bounds = ((-3, 5), (-2, 2))
    X = mesh(n, bounds)
    Lx, Ly = 8.0, 4.0
    modes = [(1, 0, 1.0), (-3, 2, 0.5j), (16, -8, 0.25), (5, 7, 0.3 - 0.2j)]  # (mx, my, amp); 16/-8 = Nyquist
    psi = sum(a * np.exp(1j * (2 * np.pi * mx / Lx * X[0] + 2 * np.pi * my / Ly * X[1])) for mx, my, a in modes)
    psi = np.ascontiguousarray(psi + 0j)
    cD, dt, ns = -0.5, 0.013, 17
    t = dt * ns
    exact = np.zeros_like(psi)
    for mx, my, a in modes:
        # the grid cannot tell m from m-n: the Nyquist mode carries ξ = -n/2
        kx = 2 * np.pi * (mx if mx < n[0] // 2 else mx - n[0]) / Lx
        ky = 2 * np.pi * (my if my < n[1] // 2 else my - n[1]) / Ly
        exact = exact + a * np.exp(1j * cD * (kx ** 2 + ky ** 2) * t) * np.exp(
            1j * (2 * np.pi * mx / Lx * X[0] + 2 * np.pi * my / Ly * X[1]))
    out = orc.split_step(psi.copy(), bounds, coeffD=cD, beta=0.0, dt=dt, nsteps=ns, order=order)
    assert rel_l2(out, exact) < 1e-12


def test_nonlinear_flow_exact_uniform(orc):
    """A uniform state in a uniform V: ψ(t) = ψ0 e^{-i(V0 + β|ψ0|²)t}."""
    bounds = ((0, 1), (0, 1), (0, 1))
    psi0 = (0.8 - 0.3j) * np.ones((8, 8, 8))
    V = np.full((8, 8, 8), 1.7)
    beta, dt, ns = 2.5, 0.01, 30
    out = orc.split_step(np.ascontiguousarray(psi0.copy()), bounds, V=V, beta=beta, dt=dt, nsteps=ns, order=2)
    exact = psi0 * np.exp(-1j * (1.7 + beta * abs(0.8 - 0.3j) ** 2) * dt * ns)
    assert rel_l2(out, exact) < 1e-13


def test_imag_nonlinear_uniform(orc):
    """Imaginary time on a uniform state: the decay is removed by the
    renormalisation, so ψ stays uniform with norm 1."""
    bounds = ((0, 2), (0, 2))
    psi0 = np.ascontiguousarray((0.3 + 0.4j) * np.ones((8, 8)))
    out = orc.split_step(psi0.copy(), bounds, beta=5.0, dt=0.01, nsteps=5, imag=True)
    assert orc.norm(out, bounds) == pytest.approx(1.0, rel=1e-14)
    assert np.allclose(out, out.flat[0], rtol=1e-14)
    assert np.angle(out.flat[0]) == pytest.approx(np.angle(0.3 + 0.4j), abs=1e-14)


# ------------------------------------------------------------ whole step
def _run_split(orc, order, nsteps, N=32, L=2 * np.pi, beta=20.0, T=0.2):
    dt = T / nsteps
    bounds = ((-L, L), (-L, L))
    X, Y = mesh((N, N), bounds)
    psi = (1 + 0.3 * np.cos(X) + 0.2 * np.sin(2 * Y)) * np.exp(1j * 0.5 * X)
    psi = np.ascontiguousarray(psi + 0j)
    psi /= orc.norm(psi, bounds)
    orc.split_step(psi, bounds, coeffD=-0.5, beta=beta, Omega=0.0, dt=dt, nsteps=nsteps, order=order)
    return psi, float(np.sum(np.abs(psi) ** 2))


@pytest.mark.parametrize("order,lo,hi,steps", [(1, 0.85, 1.15, (50, 100, 200)), (2, 1.85, 2.15, (25, 50, 100))])
def test_split_order_and_mass(orc, order, lo, hi, steps):
    """runtests.jl:644-671, same setup and bounds."""
    runs = [_run_split(orc, order, k) for k in steps]
    d = [math.sqrt(np.sum(np.abs(runs[i][0] - runs[i + 1][0]) ** 2) / np.sum(np.abs(runs[i + 1][0]) ** 2))
         for i in range(2)]
    assert lo < math.log2(d[0] / d[1]) < hi
    masses = [r[1] for r in runs]
    assert max(abs(m - masses[0]) for m in masses) < 1e-12


@pytest.mark.parametrize("shape", [(32, 16), (16, 8, 12)])
def test_strang_time_reversible(orc, shape):
    """With Ω = 0 the directional flows commute and the Strang step is
    symmetric, so Split2(-Δt) inverts Split2(Δt).  (With Ω ≠ 0 it does not:
    splitting.jl:156-203 applies x then y in both halves.)"""
    bounds = tuple((-3.0, 3.0) for _ in shape)
    psi0 = smooth_state(shape, bounds, seed=5)
    V = random_potential(shape, bounds)
    kw = dict(V=V, coeffD=-0.5, beta=4.0, Omega=0.0, nsteps=20, order=2)
    out = orc.split_step(psi0.copy(), bounds, dt=3e-3, **kw)
    assert rel_l2(out, psi0) > 1e-3
    back = orc.split_step(out, bounds, dt=-3e-3, **kw)
    assert rel_l2(back, psi0) < 1e-13


def test_rotation_lz_eigenstate(orc):
    """V = β = 0, Ω ≠ 0: ψ0 = (x + iy)e^{-r²/2} has L_z = +1 and L_z commutes
    with Δ, so ψ(t) = e^{+iΩt} e^{-i cD Δ t}ψ0.  Because splitting.jl:188-201
    runs the x then the y round trip in both half steps, the rotating Strang
    step is first order when Ω ≠ 0 (the x and y flows do not commute): the
    error must halve with Δt and vanish.  A flipped rotation sign or a
    coordinate slip leaves an O(Ωt) error that does not converge."""
    n = 64
    bounds = ((-8, 8), (-8, 8))
    X, Y = mesh((n, n), bounds)
    psi0 = np.ascontiguousarray((X + 1j * Y) * np.exp(-0.5 * (X ** 2 + Y ** 2)) + 0j)
    cD, Om, T = -0.5, 0.9, 0.5
    kx = 2 * np.pi * np.fft.fftfreq(n, 16 / n)
    K2 = kx[None, :] ** 2 + kx[:, None] ** 2
    exact = np.exp(1j * Om * T) * np.fft.ifft2(np.exp(1j * cD * K2 * T) * np.fft.fft2(psi0))
    errs = []
    for ns in (20, 40, 80):
        out = orc.split_step(psi0.copy(), bounds, coeffD=cD, beta=0.0, Omega=Om, dt=T / ns, nsteps=ns, order=2)
        errs.append(rel_l2(out, exact))
    assert errs[-1] < 2e-3
    assert 0.9 < math.log2(errs[0] / errs[1]) < 1.1
    assert 0.9 < math.log2(errs[1] / errs[2]) < 1.1
    wrong = np.exp(-1j * Om * T) * np.fft.ifft2(np.exp(1j * cD * K2 * T) * np.fft.fft2(psi0))
    assert rel_l2(out, wrong) > 0.5


def test_imag_time_thomas_fermi(orc):
    """runtests.jl:614-642 setup (48², box ±10, β = 400, Δt = 0.02, 1500
    steps, V = ½r², Gaussian start) with the imaginary-time Strang split:
    energy decreases, |E - 2μ_TF/3|/(2μ_TF/3) < 0.03, mass = 1 to 1e-8."""
    N, L, beta, dt, ns = 48, 10.0, 400.0, 0.02, 1500
    bounds = ((-L, L), (-L, L))
    X, Y = mesh((N, N), bounds)
    psi = np.ascontiguousarray(np.exp(-0.5 * (X ** 2 + Y ** 2)) + 0j)
    psi /= orc.norm(psi, bounds)
    V = np.ascontiguousarray(np.broadcast_to(0.5 * (X ** 2 + Y ** 2), psi.shape))
    e0 = orc.energy(psi, bounds, V=V, beta=beta)[3]
    orc.split_step(psi, bounds, V=V, beta=beta, dt=dt, nsteps=ns, order=2, imag=True)
    e1 = orc.energy(psi, bounds, V=V, beta=beta)[3]
    mu_tf = math.sqrt(beta / math.pi)
    dx = 2 * L / N
    mass = float(np.sum(np.abs(psi) ** 2)) * dx * dx
    assert e1 < e0
    assert abs(e1 - 2 * mu_tf / 3) / (2 * mu_tf / 3) < 0.03
    assert abs(mass - 1.0) < 1e-8


def test_split_conserves_norm_3d(orc):
    """benchmarks/conservation.jl (gp): norm to machine precision, energy
    bounded; N = 32, coeffΔ = -0.075, β = 27, Δt = 0.005."""
    bounds = ((-math.pi, math.pi),) * 3
    psi = smooth_state((32, 32, 32), bounds, seed=9, kmax=2)
    psi /= orc.norm(psi, bounds)
    kw = dict(coeffD=-0.075, beta=27.0)
    E0 = orc.energy(psi, bounds, **kw)[3]
    n0 = orc.norm(psi, bounds)
    orc.split_step(psi, bounds, dt=0.005, nsteps=100, order=2, **kw)
    assert orc.norm(psi, bounds) == pytest.approx(n0, rel=1e-13)
    assert abs(orc.energy(psi, bounds, **kw)[3] / E0 - 1) < 1e-3
