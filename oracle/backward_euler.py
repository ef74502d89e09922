"""numpy restatement of the reference's imaginary-time ground-state model,
NumModelBackwardEuler.

TEST INFRASTRUCTURE ONLY: the checker for the split-step imaginary-time
solver, imported by ``tests/``.  The HIP product path never imports it.

What it restates (reference file:line):
  - timeStep!(::NumModelBackwardEuler)        backward-euler.jl:72-85
        M = 1 / (1/dt + V + beta |phi|^2)
        b = M phi / dt
        krylov!(n, phi); normalize!(f)
  - prodA: A phi = phi - M .* lapRot(phi)     backward-euler.jl:87-90
  - lapRot 2D/3D:                             GrossPitaevskii.jl:12-45
        -coeffD (ddx + ddy [+ ddz]) + i Omega (rx + ry)
  - spectral derivatives with rotation        derivatives.jl:27-54, 90-126
        ddx = ifft_x(-xi_x^2 fft_x phi), rx = ifft_x(i y xi_x fft_x phi)
        ddy = ifft_y(-xi_y^2 fft_y phi), ry = ifft_y(-i x xi_y fft_y phi)
  - krylov!: BiCGSTAB with the real parts of the complex inner products as
    scalars, no preconditioner              krylov.jl:1-33
  - normalize!, norm                          Fields.jl:172-200
  - wavenumbers fftfreq(n, 2pi/dx), coordinates LinRange(min, max, n+1)[1:n]
                                              Plans.jl:829-830, Grids.jl:78-82

The fixed point of the step solves H phi = mu phi with
H = -lapRot + V + beta |phi|^2, independently of dt (the ground state the
split imaginary-time solver approaches as dt -> 0).

Arrays are complex128 in the reference's memory order (x fastest): a numpy
array of shape (ny, nx) or (nz, ny, nx).
"""
from __future__ import annotations

import numpy as np


class Grid:
    """Coordinates and wavenumbers of a periodic box (x-first sizes)."""

    def __init__(self, shape_xfirst, bounds):
        self.nd = len(shape_xfirst)
        self.n = tuple(int(v) for v in shape_xfirst)
        self.d = tuple((float(b[1]) - float(b[0])) / n for b, n in zip(bounds, self.n))
        self.coord, self.xi = [], []
        for a, (n, b) in enumerate(zip(self.n, bounds)):
            t = np.arange(n, dtype=np.float64) / float(n)
            self.coord.append((1.0 - t) * float(b[0]) + t * float(b[1]))  # Grids.jl:78-82
            self.xi.append(np.fft.fftfreq(n, d=self.d[a]) * 2.0 * np.pi)  # Plans.jl:829-830

    def axis(self, a):
        """numpy axis of grid axis a (x is the last numpy axis)."""
        return self.nd - 1 - a

    def bcast(self, v, a):
        shp = [1] * self.nd
        shp[self.axis(a)] = v.size
        return v.reshape(shp)

    @property
    def dv(self):
        return float(np.prod(self.d))


def _dirderiv(g: Grid, phi, a, mult):
    """ldiv_a!(mult .* mul_a!(phi)): an unnormalised forward FFT along axis
    a, a multiplier, the normalised inverse (Plans.jl:1132-1250)."""
    ax = g.axis(a)
    return np.fft.ifft(mult * np.fft.fft(phi, axis=ax), axis=ax)


def lap_rot(g: Grid, phi, coeffD, Omega):
    """lapRot (GrossPitaevskii.jl:12-45) through computeDerivatives!
    (derivatives.jl:27-54, 90-126)."""
    out = np.zeros_like(phi)
    for a in range(g.nd):
        xi = g.bcast(g.xi[a], a)
        out += -coeffD * _dirderiv(g, phi, a, -xi ** 2)
    if Omega != 0.0 and g.nd >= 2:
        x, y = g.bcast(g.coord[0], 0), g.bcast(g.coord[1], 1)
        rx = _dirderiv(g, phi, 0, 1j * y * g.bcast(g.xi[0], 0))
        ry = _dirderiv(g, phi, 1, -1j * x * g.bcast(g.xi[1], 1))
        out += 1j * Omega * (rx + ry)
    return out


def _rdot(a, b):
    """mapreduce((x, y) -> real(x * conj(y)), +, a, b) (krylov.jl)."""
    return float(np.sum((a * np.conj(b)).real))


def krylov(prodA, b, phi, nkrylov=70, tol=1e-8):
    """krylov! (krylov.jl:1-33): BiCGSTAB on A phi = b from the current
    phi; returns the iterations done and the final relative residual."""
    normb = np.sqrt(_rdot(b, b))
    r = b - prodA(phi)
    r2 = r.copy()
    p = r.copy()
    rr = _rdot(r, r)
    for i in range(1, nkrylov + 1):
        Ap = prodA(p)
        rr2 = _rdot(r, r2)
        alpha = rr2 / _rdot(Ap, r2)
        s = r - alpha * Ap
        As = prodA(s)
        omega = _rdot(As, s) / _rdot(As, As)
        phi += alpha * p + omega * s
        rr2old = _rdot(r, r2)
        r = s - omega * As
        rr = _rdot(r, r)
        if np.sqrt(rr) / normb < tol:
            return i, np.sqrt(rr) / normb
        rr2 = _rdot(r, r2)
        beta = rr2 / rr2old * alpha / omega
        p = r + beta * (p - omega * Ap)
    return nkrylov, np.sqrt(rr) / normb


def norm(g: Grid, phi):
    """norm (Fields.jl:172-180)."""
    return float(np.sqrt(np.sum(np.abs(phi) ** 2) * g.dv))


def time_step(g: Grid, phi, V, *, coeffD=-0.5, beta=1.0, Omega=0.0, dt=0.01, nkrylov=70, tolkrylov=1e-8):
    """One timeStep!(::NumModelBackwardEuler) (backward-euler.jl:72-85), in
    place on phi; returns the Krylov iteration count."""
    Vv = 0.0 if V is None else V
    M = 1.0 / (1.0 / dt + Vv + beta * np.abs(phi) ** 2)
    b = M * phi / dt

    def prodA(x):
        return x - M * lap_rot(g, x, coeffD, Omega)

    it, _ = krylov(prodA, b, phi, nkrylov, tolkrylov)
    phi /= norm(g, phi)
    return it


def ground_state(shape_xfirst, bounds, phi0, V, *, coeffD=-0.5, beta=1.0, Omega=0.0, dt=1.0, tol=1e-11,
                 max_steps=5000, nkrylov=200, tolkrylov=1e-13):
    """Backward-Euler steps from phi0 until ||phi_{n+1} - phi_n|| (with the
    field norm) falls below tol.  Returns (phi, steps, last change)."""
    g = Grid(shape_xfirst, bounds)
    phi = np.array(phi0, dtype=np.complex128, copy=True)
    phi /= norm(g, phi)
    change = np.inf
    for s in range(1, max_steps + 1):
        prev = phi.copy()
        time_step(g, phi, V, coeffD=coeffD, beta=beta, Omega=Omega, dt=dt, nkrylov=nkrylov, tolkrylov=tolkrylov)
        change = norm(g, phi - prev)
        if change < tol:
            return phi, s, change
    return phi, max_steps, change


def residual(shape_xfirst, bounds, phi, V, *, coeffD=-0.5, beta=1.0, Omega=0.0):
    """(mu, ||H phi - mu phi||) with H = -lapRot + V + beta|phi|^2 and
    mu = <phi, H phi> for a normalised phi: the eigen-residual of the fixed
    point."""
    g = Grid(shape_xfirst, bounds)
    Vv = 0.0 if V is None else V
    Hphi = -lap_rot(g, phi, coeffD, Omega) + (Vv + beta * np.abs(phi) ** 2) * phi
    mu = float(np.sum(np.conj(phi) * Hphi).real * g.dv)
    return mu, norm(g, Hphi - mu * phi)
