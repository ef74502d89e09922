"""Shared helpers of the test suite (initial states, error metrics, GPU runs)."""
from __future__ import annotations

import numpy as np


def rel_l2(a, b) -> float:
    """sqrt(Σ|a-b|²)/sqrt(Σ|b|²): the metric of runtests.jl:1444."""
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2)) / np.sqrt(np.sum(np.abs(b) ** 2)))


def coords(n, lo, hi):
    t = np.arange(n, dtype=np.float64) / float(n)
    return (1.0 - t) * float(lo) + t * float(hi)


def mesh(shape_xfirst, bounds):
    """Broadcast coordinates (x, y[, z]) for a numpy array of shape reversed(shape_xfirst)."""
    nd = len(shape_xfirst)
    out = []
    for a in range(nd):
        shp = [1] * nd
        shp[nd - 1 - a] = shape_xfirst[a]
        out.append(coords(shape_xfirst[a], *bounds[a]).reshape(shp))
    return out


def smooth_state(shape_xfirst, bounds, seed=0, kmax=3):
    """A smooth random complex state: a Gaussian envelope times a few random
    low modes (band-limited, so split steps stay well resolved)."""
    rng = np.random.default_rng(seed)
    X = mesh(shape_xfirst, bounds)
    shape = tuple(reversed(shape_xfirst))
    psi = np.zeros(shape, dtype=np.complex128)
    centre = [0.5 * (b[0] + b[1]) for b in bounds]
    width = [0.25 * (b[1] - b[0]) for b in bounds]
    env = np.ones(shape)
    for a, x in enumerate(X):
        env = env * np.exp(-((x - centre[a]) / width[a]) ** 2)
    for _ in range(4):
        ph = np.zeros(shape)
        for a, x in enumerate(X):
            L = bounds[a][1] - bounds[a][0]
            ph = ph + rng.integers(-kmax, kmax + 1) * 2 * np.pi * x / L
        psi = psi + (rng.standard_normal() + 1j * rng.standard_normal()) * np.exp(1j * ph)
    psi = psi * env + 0.05
    return np.ascontiguousarray(psi)


def random_potential(shape_xfirst, bounds, seed=1, amp=2.0):
    X = mesh(shape_xfirst, bounds)
    rng = np.random.default_rng(seed)
    V = np.zeros(tuple(reversed(shape_xfirst)))
    for a, x in enumerate(X):
        V = V + rng.uniform(0.2, 1.0) * amp * (x / (bounds[a][1] - bounds[a][0])) ** 2
    return np.ascontiguousarray(V)


PRECS = {"f64": (np.float64, None), "f32": (np.float32, None), "f32x": (np.float32, np.float64)}
TOLS = {"f64": 1e-12, "f32": 2e-5, "f32x": 1e-6}


def gpu_run(sfd, psi0, bounds, *, V=None, FT=np.float64, coeffD=-0.5, beta=1.0, Omega=0.0, dt=1e-3,
            nsteps=1, order=2, imag=False, prec=None):
    """Run the HIP split step through the host API on a numpy state."""
    compute = None
    if prec is not None:
        FT, compute = PRECS[prec]
    g = sfd.Grid(tuple(reversed(psi0.shape)), bounds, FT=FT, compute=compute)
    f = sfd.Field(g, sfd.ComplexField())
    f.set(psi0)
    pot = sfd.PotentialArray(f, V) if V is not None else sfd.PotentialZero(f)
    p = sfd.GrossPitaevskiiParameters(pot=pot, coeffΔ=coeffD, β=beta, Ω=Omega)
    if imag:
        m = sfd.NumModelSplitImag(f, p, dt, nsteps, nsteps, order=order)
    elif order == 1:
        m = sfd.NumModelSplit1(f, p, dt, nsteps, nsteps)
    else:
        m = sfd.NumModelSplit2(f, p, dt, nsteps, nsteps)
    m.steps(nsteps)
    return f.to_numpy().astype(np.complex128)
