"""Synthetic initial states of the benchmark and parity workloads.

All states are generated once in FP64 on the host with numpy's
``Generator(PCG64(seed))`` (the reference's MersenneTwister streams are not
reproducible outside Julia), in the x-fastest layout: numpy shape
``(ny, nx)`` or ``(nz, ny, nx)``.
"""
from __future__ import annotations

import hashlib
import math

import numpy as np


def coords(n, lo, hi):
    t = np.arange(n, dtype=np.float64) / float(n)
    return (1.0 - t) * float(lo) + t * float(hi)


def discrete_normalize(psi, d):
    """normalize! (Fields.jl:196-200): divide by sqrt(Σ|ψ|²)·sqrt(ΠΔ)."""
    nrm = math.sqrt(float(np.sum(np.abs(psi) ** 2))) * math.sqrt(math.prod(d))
    return psi / nrm


def thomas_fermi_2d(n=128, L=8.0, beta=500.0, seed=0, noise=0.1):
    """2D Thomas-Fermi state on [-L, L)² in V = ½(x²+y²):
    ψ0 = sqrt(max(0, (μ - V)/β)) e^{iθ}, μ = sqrt(β/π), θ ~ U[-noise, noise]
    per site, then a discrete normalize!.  Returns (psi, V, bounds)."""
    x = coords(n, -L, L)
    X, Y = x[None, :], x[:, None]
    V = 0.5 * (X ** 2 + Y ** 2)
    mu = math.sqrt(beta / math.pi)
    rho = np.maximum(0.0, (mu - V) / beta)
    rng = np.random.Generator(np.random.PCG64(seed))
    theta = rng.uniform(-noise, noise, size=(n, n))
    psi = np.sqrt(rho) * np.exp(1j * theta)
    d = (2 * L / n, 2 * L / n)
    return np.ascontiguousarray(discrete_normalize(psi, d)), np.ascontiguousarray(V), ((-L, L), (-L, L))


def random_phase_3d(n=256, kmax=8, seed=0, rms=1.0):
    """Homogeneous condensate ψ0 = e^{iθ} on [0, 2π)³: θ is a Gaussian random
    field on the integer modes 1 <= |k| <= kmax with modal variance ∝ |k|^-2
    (complex normal coefficients, real part of the inverse FFT), scaled to
    the given rms.  Returns psi (complex128, shape (n, n, n))."""
    rng = np.random.Generator(np.random.PCG64(seed))
    C = np.zeros((n, n, n), dtype=np.complex128)
    ks = range(-kmax, kmax + 1)
    for kz in ks:
        for ky in ks:
            for kx in ks:
                k2 = kx * kx + ky * ky + kz * kz
                if k2 < 1 or k2 > kmax * kmax:
                    continue
                a, b = rng.standard_normal(2)
                C[kz % n, ky % n, kx % n] = (a + 1j * b) / math.sqrt(2.0 * k2)
    theta = np.fft.ifftn(C).real
    theta *= rms / math.sqrt(float(np.mean(theta ** 2)))
    return np.exp(1j * theta)


def metric_params(n=256):
    """Parameters of the 256³ metric workload: box [0, 2π)³, ξ = 2Δ,
    β = g = 1/ξ², V = 0, coeffΔ = -0.5, Δt = 5e-4."""
    L = 2 * math.pi
    dx = L / n
    xi = 2 * dx
    return dict(bounds=((0.0, L),) * 3, beta=1.0 / xi ** 2, coeffD=-0.5, dt=5e-4, Omega=0.0)


def sha256(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def batch_ground_states(seed=0, counts=((128, 32), (256, 48), (512, 16)), L=10.0):
    """96 uneven 2D ground-state problems on [-L, L)²: V = ½(x² + γ²y²) with
    γ ~ U[1, 3], β ~ U[100, 2000], and a seeded Gaussian initial guess
    exp(-(x²/σx² + y²/σy²)/2) with σx, σy ~ U[1, 3], normalised.
    Returns a list of dicts (n, gamma, beta, psi0, V, bounds)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for n, count in counts:
        x = coords(n, -L, L)
        X, Y = x[None, :], x[:, None]
        d = (2 * L / n, 2 * L / n)
        for _ in range(count):
            gamma = rng.uniform(1.0, 3.0)
            beta = rng.uniform(100.0, 2000.0)
            sx, sy = rng.uniform(1.0, 3.0, size=2)
            psi = np.exp(-0.5 * (X ** 2 / sx ** 2 + Y ** 2 / sy ** 2)) + 0j
            psi = np.ascontiguousarray(discrete_normalize(psi, d))
            V = np.ascontiguousarray(0.5 * (X ** 2 + gamma ** 2 * Y ** 2) + 0 * X)
            out.append(dict(n=n, gamma=gamma, beta=beta, psi0=psi, V=V, bounds=((-L, L), (-L, L))))
    return out


def control_problem(n=256, L=8.0, seed=0, sigma2=0.5, lattice=8, a_std=0.1):
    """Differentiable-control setup on [-L, L)²: V(a) = ½|x|² +
    Σ_j a_j exp(-|x - c_j|²/σ²) with c_j on a lattice×lattice grid of cell
    centres and a_j ~ N(0, a_std²) from the seed.  Returns (basis (nb, n, n),
    a, V0 = ½|x|², bounds)."""
    x = coords(n, -L, L)
    X, Y = x[None, :], x[:, None]
    h = 2 * L / lattice
    centres = [(-L + (i + 0.5) * h, -L + (j + 0.5) * h) for j in range(lattice) for i in range(lattice)]
    basis = np.stack([np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / sigma2) for cx, cy in centres])
    rng = np.random.Generator(np.random.PCG64(seed))
    a = rng.normal(0.0, a_std, size=len(centres))
    V0 = np.ascontiguousarray(0.5 * (X ** 2 + Y ** 2) + 0 * X)
    return np.ascontiguousarray(basis), a, V0, ((-L, L), (-L, L))


def shift_x(psi, shift, L):
    """Band-limited shift ψ(x - shift) along x (the last numpy axis)."""
    n = psi.shape[-1]
    k = 2 * np.pi * np.fft.fftfreq(n, 2 * L / n)
    return np.fft.ifft(np.fft.fft(psi, axis=-1) * np.exp(-1j * k * shift), axis=-1)
