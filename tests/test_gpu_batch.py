"""GPU checks of the batched imaginary-time solver (sfd_batch_imag_run)."""
import numpy as np
import pytest

from helpers import PRECS, TOLS, gpu_run, random_potential, rel_l2, smooth_state

pytestmark = pytest.mark.gpu

SHAPES = [(32, 32), (64, 32), (32, 32), (128, 64), (16, 256), (64, 32), (512, 16), (16, 512)]


def _make(gpu, prec, seed=0):
    FT, compute = PRECS[prec]
    rng = np.random.default_rng(seed)
    fields, params, inits = [], [], []
    for k, shp in enumerate(SHAPES):
        bounds = ((-4.0 - k, 4.0 + k), (-3.0, 3.0))
        psi = smooth_state(shp, bounds, seed=k)
        V = random_potential(shp, bounds, seed=k + 1, amp=8.0) if k % 3 else None
        beta = float(rng.uniform(5, 40))
        g = gpu.Grid(shp, bounds, FT=FT, compute=compute)
        f = gpu.Field(g)
        f.set(psi)
        pot = gpu.PotentialArray(f, V) if V is not None else gpu.PotentialZero(f)
        fields.append(f)
        params.append(gpu.GrossPitaevskiiParameters(pot=pot, coeffΔ=-0.5, β=beta))
        inits.append((psi, V, bounds, beta))
    return fields, params, inits


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("prec", ["f64", "f32", "f32x"])
def test_batch_equals_single_field_runs(gpu, prec, graph):
    """Each batched field ends where the single-field imaginary-time Split2
    (sfd_split_step, SFD_IMAG_TIME) ends after the same steps."""
    fields, params, inits = _make(gpu, prec)
    solver = gpu.BatchImagSolver(fields, params, 4e-3)
    steps, dpsi = solver.run(max_steps=37, tol=0.0, check_every=10, use_graph=graph)
    assert list(steps) == [37] * len(fields)
    assert np.all(dpsi > 0)
    for f, (psi0, V, bounds, beta) in zip(fields, inits):
        ref = gpu_run(gpu, psi0, bounds, V=V, beta=beta, dt=4e-3, nsteps=37, order=2, imag=True, prec=prec)
        err = rel_l2(f.to_numpy(), ref)
        assert err <= (1e-14 if prec == "f64" else 1e-6), err
        assert gpu.norm(f) == pytest.approx(1.0, rel=1e-6)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_batch_autotune_keeps_results(gpu, prec):
    """After sfd_batch_autotune picks a variant per size group and axis, the
    batch still ends where the single-field runs end, and the fields the
    tuner ran beside were not touched."""
    fields, params, inits = _make(gpu, prec, seed=7)
    solver = gpu.BatchImagSolver(fields, params, 4e-3)
    before = [f.to_numpy() for f in fields]
    tune = solver.autotune()
    # axes with one compiled variant are not timed
    assert tune and any(len(v) > 1 for row in tune for v in row.values())
    for f, b in zip(fields, before):
        assert np.array_equal(f.to_numpy(), b)
    steps, _ = solver.run(max_steps=23, tol=0.0, check_every=10, use_graph=True)
    assert list(steps) == [23] * len(fields)
    for f, (psi0, V, bounds, beta) in zip(fields, inits):
        ref = gpu_run(gpu, psi0, bounds, V=V, beta=beta, dt=4e-3, nsteps=23, order=2, imag=True, prec=prec)
        assert rel_l2(f.to_numpy(), ref) <= (1e-14 if prec == "f64" else 1e-6)


def test_batch_matches_oracle(gpu, orc):
    fields, params, inits = _make(gpu, "f64", seed=3)
    gpu.BatchImagSolver(fields, params, 5e-3).run(max_steps=20, tol=0.0, check_every=7)
    for f, (psi0, V, bounds, beta) in zip(fields, inits):
        ref = orc.split_step(psi0.copy(), bounds, V=V, beta=beta, dt=5e-3, nsteps=20, order=2, imag=True)
        assert rel_l2(f.to_numpy(), ref) < TOLS["f64"]


def test_batch_convergence_flags(gpu):
    """Fields stop individually once ||ψ_{n+1} - ψ_n|| < tol; the others go on."""
    fields, params, _ = _make(gpu, "f64", seed=5)
    steps, dpsi = gpu.BatchImagSolver(fields, params, 1e-2).run(max_steps=4000, tol=1e-7, check_every=50)
    assert np.all(steps <= 4000)
    done = steps < 4000
    assert done.any()
    assert np.all(dpsi[done] < 1e-7)
    assert np.all(steps[done] % 50 == 0)


def test_imag_time_bao_du_2d(gpu):
    """Bao & Du 2D case I (the values of benchmarks/validate_expected.jl:128-131):
    β = 200, box [-8,8)×[-4,4), 256², V = ½(x² + γy² y²) with γy = 4 in the
    γ² convention of the paper; E = 11.1563, μ = 16.3377.  The imaginary-time
    Strang split at Δt = 0.02 reaches E within 1e-4 and μ within 3e-3 (its
    fixed point carries a splitting error, O(Δt²) in E and O(Δt) in μ)."""
    g = gpu.Grid((256, 256), ((-8, 8), (-4, 4)))
    f = gpu.Field(g)
    x, y = f.coords()
    f.set(np.ascontiguousarray(np.broadcast_to(np.exp(-0.5 * (x ** 2 + y ** 2)) + 0j, g.shape)))
    gpu.normalize(f)
    pot = gpu.PotentialQuadratic(f, γx=1.0, γy=16.0)
    p = gpu.GrossPitaevskiiParameters(pot=pot, coeffΔ=-0.5, β=200.0)
    fields, params = [f], [p]
    steps, dpsi = gpu.BatchImagSolver(fields, params, 0.02).run(max_steps=6000, tol=1e-10, check_every=100)
    EOm, ED, Eb, E, mu = gpu.energy_all(gpu.NumModelSplitImag(f, p, 0.02, 1, 1))
    print(f"Bao-Du 2D_I: E = {E:.6f}, mu = {mu:.6f}, steps {steps[0]}, dpsi {dpsi[0]:.2e}")
    assert abs(E - 11.1563) / 11.1563 < 1e-4
    assert abs(mu - 16.3377) / 16.3377 < 3e-3
    assert gpu.norm(f) == pytest.approx(1.0, rel=1e-12)
