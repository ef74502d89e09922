"""Pins of the oracle adjoint (orc_adjoint): central finite differences of
the forward loss, a directional derivative in V, and the dot-product
(linearisation) identity."""
import numpy as np
import pytest

from helpers import random_potential, smooth_state


def _setup(shape=(16, 16), seed=0, Om=0.0):
    bounds = ((-3.0, 3.0), (-3.0, 3.0))[:len(shape)]
    psi0 = smooth_state(shape, bounds, seed=seed)
    tgt = smooth_state(shape, bounds, seed=seed + 17)
    V = random_potential(shape, bounds, seed=seed + 3, amp=3.0)
    return bounds, psi0, tgt, V


@pytest.mark.parametrize("Om", [0.0, 0.4])
def test_adjoint_beta_matches_central_difference(orc, Om):
    bounds, psi0, tgt, V = _setup(Om=Om)
    kw = dict(coeffD=-0.5, Omega=Om, dt=2e-3, nsteps=12)
    psi0 = psi0 / orc.norm(np.ascontiguousarray(psi0), bounds)
    tgt = tgt / orc.norm(np.ascontiguousarray(tgt), bounds)
    beta = 7.0
    loss, gV, gb = orc.adjoint(psi0, tgt, bounds, V=V, beta=beta, **kw)
    h = 1e-4
    fd = (orc.overlap_loss(psi0, tgt, bounds, V=V, beta=beta + h, **kw)
          - orc.overlap_loss(psi0, tgt, bounds, V=V, beta=beta - h, **kw)) / (2 * h)
    assert loss == pytest.approx(orc.overlap_loss(psi0, tgt, bounds, V=V, beta=beta, **kw), rel=1e-14)
    assert gb == pytest.approx(fd, rel=1e-6)


@pytest.mark.parametrize("Om", [0.0, 0.4])
def test_adjoint_potential_directional_derivative(orc, Om):
    """<dL/dV, δV> equals the central difference of L along δV."""
    bounds, psi0, tgt, V = _setup(shape=(16, 8), seed=2, Om=Om)
    psi0 = psi0 / orc.norm(np.ascontiguousarray(psi0), bounds)
    tgt = tgt / orc.norm(np.ascontiguousarray(tgt), bounds)
    kw = dict(coeffD=-0.5, beta=4.0, Omega=Om, dt=3e-3, nsteps=10)
    _, gV, _ = orc.adjoint(psi0, tgt, bounds, V=V, **kw)
    rng = np.random.default_rng(9)
    for _ in range(3):
        dV = rng.standard_normal(V.shape)
        h = 1e-4
        fd = (orc.overlap_loss(psi0, tgt, bounds, V=V + h * dV, **kw)
              - orc.overlap_loss(psi0, tgt, bounds, V=V - h * dV, **kw)) / (2 * h)
        assert float(np.sum(gV * dV)) == pytest.approx(fd, rel=1e-6)


def test_adjoint_single_site_components(orc):
    """Individual components of dL/dV(x) against one-site perturbations."""
    bounds, psi0, tgt, V = _setup(shape=(8, 8), seed=4)
    psi0 = psi0 / orc.norm(np.ascontiguousarray(psi0), bounds)
    tgt = tgt / orc.norm(np.ascontiguousarray(tgt), bounds)
    kw = dict(coeffD=-0.5, beta=3.0, dt=4e-3, nsteps=6)
    _, gV, _ = orc.adjoint(psi0, tgt, bounds, V=V, **kw)
    for idx in [(0, 0), (3, 5), (7, 2)]:
        dV = np.zeros_like(V)
        dV[idx] = 1.0
        h = 1e-4
        fd = (orc.overlap_loss(psi0, tgt, bounds, V=V + h * dV, **kw)
              - orc.overlap_loss(psi0, tgt, bounds, V=V - h * dV, **kw)) / (2 * h)
        assert gV[idx] == pytest.approx(fd, rel=1e-5, abs=1e-12)
