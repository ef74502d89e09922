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


MIXED = [  # (extents x-first, Omega, with V): rotating and plain 2D, rotating and plain 3D
    ((64, 32), 0.4, True),
    ((64, 32), 0.0, True),   # same shape as a rotating instance: its group runs the unfolded step
    ((32, 32), 0.0, False),
    ((32, 16, 8), 0.3, True),
    ((32, 16, 8), 0.0, False),
    ((16, 16, 16), 0.5, True),
]


def _make_mixed(gpu, prec, seed=11):
    FT, compute = PRECS[prec]
    rng = np.random.default_rng(seed)
    fields, params, inits = [], [], []
    for k, (shp, Om, with_v) in enumerate(MIXED):
        bounds = tuple((-3.0 - 0.5 * a, 3.0 + 0.5 * a) for a in range(len(shp)))
        psi = smooth_state(shp, bounds, seed=20 + k)
        V = random_potential(shp, bounds, seed=k + 2, amp=6.0) if with_v else None
        beta = float(rng.uniform(5, 30))
        f = gpu.Field(gpu.Grid(shp, bounds, FT=FT, compute=compute))
        f.set(psi)
        pot = gpu.PotentialArray(f, V) if V is not None else gpu.PotentialZero(f)
        fields.append(f)
        params.append(gpu.GrossPitaevskiiParameters(pot=pot, coeffΔ=-0.5, β=beta, Ω=Om))
        inits.append((psi, V, bounds, beta, Om))
    return fields, params, inits


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("prec", ["f64", "f32", "f32x"])
def test_batch_rotation_and_3d_equal_single_field_runs(gpu, prec, graph):
    """Batches holding rotating and non-rotating, 2D and 3D instances end
    where the single-field imaginary-time Split2 (SFD_IMAG_TIME) ends: the
    same sweeps, rotation multipliers and renormalisation per instance.  In
    FP64 they also equal the oracle."""
    fields, params, inits = _make_mixed(gpu, prec)
    steps, dpsi = gpu.BatchImagSolver(fields, params, 4e-3).run(max_steps=37, tol=0.0, check_every=10,
                                                                 use_graph=graph)
    assert list(steps) == [37] * len(fields)
    for f, (psi0, V, bounds, beta, Om) in zip(fields, inits):
        ref = gpu_run(gpu, psi0, bounds, V=V, beta=beta, Omega=Om, dt=4e-3, nsteps=37, order=2, imag=True,
                      prec=prec)
        err = rel_l2(f.to_numpy(), ref)
        assert err <= (1e-13 if prec == "f64" else 1e-6), (f.g.n, Om, err)
        assert gpu.norm(f) == pytest.approx(1.0, rel=1e-6)


@pytest.mark.parametrize("check_every", [1, 2, 3, 4, 37])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_batch_chunk_lengths_equal_single_field_runs(gpu, prec, check_every):
    """Non-rotating groups chain the steps of a chunk before its check step
    (x, [y N y], [x x], ..., x in 2D; x, y, [z N z], y, [x x], ... in 3D);
    chunks of 1 and 2 steps do not chain, 3 chains two steps, 4 three.  Every
    chunk length ends where the single-field runs end, rotating groups
    beside them included."""
    fields, params, inits = _make_mixed(gpu, prec, seed=17)
    steps, dpsi = gpu.BatchImagSolver(fields, params, 4e-3).run(max_steps=37, tol=0.0, check_every=check_every)
    assert list(steps) == [37] * len(fields)
    assert np.all(dpsi > 0)
    for f, (psi0, V, bounds, beta, Om) in zip(fields, inits):
        ref = gpu_run(gpu, psi0, bounds, V=V, beta=beta, Omega=Om, dt=4e-3, nsteps=37, order=2, imag=True,
                      prec=prec)
        err = rel_l2(f.to_numpy(), ref)
        assert err <= (1e-13 if prec == "f64" else 1e-6), (f.g.n, Om, check_every, err)
        assert gpu.norm(f) == pytest.approx(1.0, rel=1e-6)


def test_batch_rotation_and_3d_match_oracle(gpu, orc):
    fields, params, inits = _make_mixed(gpu, "f64", seed=13)
    gpu.BatchImagSolver(fields, params, 5e-3).run(max_steps=20, tol=0.0, check_every=7)
    for f, (psi0, V, bounds, beta, Om) in zip(fields, inits):
        ref = orc.split_step(psi0.copy(), bounds, V=V, beta=beta, Omega=Om, dt=5e-3, nsteps=20, order=2, imag=True)
        assert rel_l2(f.to_numpy(), ref) < TOLS["f64"], (f.g.n, Om)


def _cfg4(gpu, prec):
    """The workload's fields and parameters.  V is rounded to FP32 for every
    precision: an FP32 batch stores V in FP32, so its FP64 references solve
    that same problem (rounding V shifts the ground state by up to 1.7e-7)."""
    from sfdhip.workloads import batch_ground_states
    FT, compute = PRECS[prec]
    probs = batch_ground_states(seed=0)
    fields, params = [], []
    for pr in probs:
        f = gpu.Field(gpu.Grid((pr["n"], pr["n"]), pr["bounds"], FT=FT, compute=compute))
        f.set(pr["psi0"])
        V = pr["V"].astype(np.float32).astype(np.float64)
        params.append(gpu.GrossPitaevskiiParameters(pot=gpu.PotentialArray(f, V), coeffΔ=-0.5, β=pr["beta"]))
        fields.append(f)
    return probs, fields, params


def _energy64(gpu, pr, psi64):
    g = gpu.Grid((pr["n"], pr["n"]), pr["bounds"])
    f = gpu.Field(g)
    f.ϕ.copy_(psi64)
    V = pr["V"].astype(np.float32).astype(np.float64)
    p = gpu.GrossPitaevskiiParameters(pot=gpu.PotentialArray(f, V), coeffΔ=-0.5, β=pr["beta"])
    return gpu.energy_all(gpu.NumModelSplitImag(f, p, 5e-3, 1, 1))


@pytest.mark.parametrize("prec", ["f32", "f32x"])
def test_batch_cfg4_fp32_escalates_to_fp64_and_converges(gpu, prec):
    """The full batched ground-state workload (32 x 128², 48 x 256², 16 x 512²
    on [-10,10)², V = ½(x² + γ²y²), γ ~ U[1,3], β ~ U[100,2000], seeded
    Gaussian guesses, Δt = 5e-3) in FP32 storage with FP64 renormalisation:
    every instance runs in FP32 to its rounding floor, finishes in FP64 and
    meets ||Δψ|| < 1e-9 well within 20000 steps.  Its final state lies as
    close to the FP64 fixed point (an FP64 batch run to 1e-12) as the FP64
    batch stopped at 1e-9 does, and every instance's energy matches the FP64
    batch's to 1e-10 relative (the energy is stationary at the ground state;
    μ is not and is compared at the states' own accuracy, 1e-7).  Three sampled instances also equal
    the single-field FP64 imaginary-time fixed point (sfd_split_step), which
    the oracle pins in test_gpu_ground_state."""
    import time
    import torch
    probs, fields, params = _cfg4(gpu, prec)
    out64 = [torch.empty(f.g.shape, dtype=torch.complex128, device="cuda") for f in fields]
    solver = gpu.BatchImagSolver(fields, params, 5e-3)
    solver.autotune()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    steps, dpsi = solver.run(max_steps=20000, tol=1e-9, check_every=100, escalate=True, fp64_out=out64)
    wall = time.perf_counter() - t0
    print(f"cfg4 {prec} + FP64 finish: {wall:.3f} s, steps {steps.min()}..{steps.max()}, "
          f"dpsi max {dpsi.max():.2e}, FP64 levels {solver.levels.sum()}")
    assert np.all(steps < 20000) and np.all(dpsi < 1e-9) and np.all(dpsi >= 0)
    assert np.all(solver.levels == 1)
    # FP64 references: the batch at tol 1e-9 and its fixed point (tol 1e-12)
    _, f64, p64 = _cfg4(gpu, "f64")
    s64 = gpu.BatchImagSolver(f64, p64, 5e-3)
    t0 = time.perf_counter()
    st64, _ = s64.run(max_steps=20000, tol=1e-9, check_every=100)
    print(f"cfg4 FP64: {time.perf_counter() - t0:.3f} s, steps {st64.min()}..{st64.max()}")
    ref64 = [f.ϕ.clone() for f in f64]
    _, fx, px = _cfg4(gpu, "f64")
    stx, dx = gpu.BatchImagSolver(fx, px, 5e-3).run(max_steps=40000, tol=1e-12, check_every=100)
    assert np.all(dx < 1e-12)
    e_esc, e_64 = [], []
    for i, pr in enumerate(probs):
        fixed = fx[i].ϕ
        nrm = fixed.abs().pow(2).sum().sqrt()
        e_esc.append(float((out64[i] - fixed).abs().pow(2).sum().sqrt() / nrm))
        e_64.append(float((ref64[i] - fixed).abs().pow(2).sum().sqrt() / nrm))
        # the returned FP32 field is the FP64 state rounded once
        assert rel_l2(fields[i].to_numpy(), out64[i].cpu().numpy()) < 1e-7
        # every instance's energy and chemical potential against the FP64 batch's
        Ee, E6 = _energy64(gpu, pr, out64[i]), _energy64(gpu, pr, ref64[i])
        assert abs(Ee[3] - E6[3]) <= 1e-10 * abs(E6[3]), (i, Ee[3], E6[3])
        assert abs(Ee[4] - E6[4]) <= 1e-7 * abs(E6[4]), (i, Ee[4], E6[4])
    e_esc, e_64 = np.array(e_esc), np.array(e_64)
    print(f"cfg4 {prec}: distance to the FP64 fixed point: escalated median {np.median(e_esc):.2e} "
          f"max {e_esc.max():.2e}; FP64 batch median {np.median(e_64):.2e} max {e_64.max():.2e}")
    # both stop at their first check below 1e-9 (the escalated run on the
    # same steps: the FP32 stage only replaces the FP64 steps above its floor)
    assert np.all(e_esc <= 2.0 * e_64 + 1e-8)
    assert np.median(e_esc) <= 1.5 * np.median(e_64)
    for i in (0, 40, 95):  # sampled instances against the single-field FP64 fixed point
        pr = probs[i]
        g = gpu.Grid((pr["n"], pr["n"]), pr["bounds"])
        f = gpu.Field(g)
        f.set(pr["psi0"])
        V = pr["V"].astype(np.float32).astype(np.float64)
        p = gpu.GrossPitaevskiiParameters(pot=gpu.PotentialArray(f, V), coeffΔ=-0.5, β=pr["beta"])
        gpu.NumModelSplitImag(f, p, 5e-3, 1, 1).steps(3000)
        single = f.ϕ
        d = float((out64[i] - single).abs().pow(2).sum().sqrt() / single.abs().pow(2).sum().sqrt())
        assert d < 1e-7, (i, d)
        Ee = _energy64(gpu, pr, out64[i])
        Ef = _energy64(gpu, pr, single)
        assert abs(Ee[3] - Ef[3]) <= 1e-10 * abs(Ef[3]), (i, Ee[3], Ef[3])
        assert abs(Ee[4] - Ef[4]) <= 1e-7 * abs(Ef[4]), (i, Ee[4], Ef[4])
