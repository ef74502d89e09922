"""GPU checks of the directional FFTs, reductions and full-size invariants."""

import numpy as np
import pytest
import torch

from helpers import mesh, random_potential, rel_l2, smooth_state

pytestmark = pytest.mark.gpu


def _field(sfd, psi, bounds, FT=np.float64):
    g = sfd.Grid(tuple(reversed(psi.shape)), bounds, FT=FT)
    f = sfd.Field(g)
    f.set(psi)
    return f


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("shape", [(32, 16), (256, 64), (16, 32, 8), (64, 128, 32), (512, 8), (8, 16, 512)])
def test_fft_axis_matches_numpy(gpu, shape, prec):
    """mul_d! is the unnormalised forward FFT along d, ldiv_d! the normalised
    inverse (Plans.jl:1132-1250); both in pen_x layout here."""
    FT = np.float64 if prec == "f64" else np.float32
    rng = np.random.default_rng(3)
    a = rng.standard_normal(tuple(reversed(shape))) + 1j * rng.standard_normal(tuple(reversed(shape)))
    bounds = [(0, 1)] * len(shape)
    f = _field(gpu, a, bounds, FT)
    out = torch.empty_like(f.ϕ)
    tol = 1e-13 if prec == "f64" else 2e-6
    for ax in range(len(shape)):
        npax = len(shape) - 1 - ax
        [gpu.mul_x, gpu.mul_y, gpu.mul_z][ax](out, f, f.ϕ)
        assert rel_l2(out.cpu().numpy(), np.fft.fft(a, axis=npax)) < tol
        [gpu.ldiv_x, gpu.ldiv_y, gpu.ldiv_z][ax](out, f, f.ϕ)
        assert rel_l2(out.cpu().numpy(), np.fft.ifft(a, axis=npax)) < tol
    # in place round trip
    g = f.ϕ.clone()
    gpu.mul_y(g, f, g)
    gpu.ldiv_y(g, f, g)
    assert rel_l2(g.cpu().numpy(), a) < tol


def test_norm_golden_3d(gpu):
    """Fields.jl:166-170: on Grid((32,32,32), ((-12,12),)*3) the field
    (-12-x)(12+x)+(-12-y)(12+y)+(-12-z)(12+z) has norm 72951.55177239206."""
    bounds = ((-12, 12),) * 3
    x, y, z = mesh((32, 32, 32), bounds)
    phi = (-12 - x) * (12 + x) + (-12 - y) * (12 + y) + (-12 - z) * (12 + z) + 0j
    f = _field(gpu, np.ascontiguousarray(np.broadcast_to(phi, (32, 32, 32))), bounds)
    assert gpu.norm(f) == pytest.approx(72951.55177239206, rel=1e-14)


def test_norm_normalize(gpu):
    """runtests.jl:718-725: ϕ = 1 on 64² [-12,12)² has norm sqrt(Lx Ly)."""
    bounds = ((-12, 12), (-12, 12))
    f = _field(gpu, np.ones((64, 64), dtype=np.complex128), bounds)
    assert gpu.norm(f) == pytest.approx(24.0, rel=1e-14)
    gpu.normalize(f)
    assert gpu.norm(f) == pytest.approx(1.0, rel=1e-14)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("case", [((64, 32), ((-6, 6), (-3, 3)), 0.7),
                                  ((32, 16, 64), ((-4, 4), (-3, 3), (-5, 5)), 0.5),
                                  ((512, 16), ((-8, 8), (-2, 2)), 0.3)], ids=["2d", "3d", "2d-512"])
def test_energy_matches_oracle(gpu, orc, case, prec):
    shape, bounds, Om = case
    FT = np.float64 if prec == "f64" else np.float32
    psi = smooth_state(shape, bounds, seed=11)
    V = random_potential(shape, bounds, amp=5.0)
    ref = orc.energy(psi, bounds, V=V, coeffD=-0.5, beta=7.0, Omega=Om)
    f = _field(gpu, psi, bounds, FT)
    p = gpu.GrossPitaevskiiParameters(pot=gpu.PotentialArray(f, V), coeffΔ=-0.5, β=7.0, Ω=Om)
    m = gpu.NumModelSplit2(f, p, 1e-3, 1, 1)
    got = gpu.energy_all(m)
    scale = max(abs(v) for v in ref)
    tol = 1e-12 if prec == "f64" else 2e-6
    for g_, r_ in zip(got, ref):
        assert abs(g_ - r_) <= tol * scale, (got, ref)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_full_size_invariants(gpu, prec):
    """256³ metric workload: norm conserved and Split2(Δt) then Split2(-Δt)
    returns the initial state (the step is unitary and time-reversible)."""
    from sfdhip.workloads import metric_params, random_phase_3d
    FT = np.float64 if prec == "f64" else np.float32
    mp = metric_params(256)
    psi0 = random_phase_3d(256, seed=0)
    f = _field(gpu, psi0, mp["bounds"], FT)
    pot = gpu.PotentialZero(f)
    p = gpu.GrossPitaevskiiParameters(pot=pot, coeffΔ=-0.5, β=mp["beta"])
    n0 = gpu.norm(f)
    fwd = gpu.NumModelSplit2(f, p, mp["dt"], 10, 10)
    bwd = gpu.NumModelSplit2(f, p, -mp["dt"], 10, 10)
    fwd.steps(10)
    n1 = gpu.norm(f)
    moved = rel_l2(f.to_numpy(), psi0)
    bwd.steps(10)
    back = rel_l2(f.to_numpy(), psi0)
    print(f"256³ {prec}: norm drift {abs(n1 - n0) / n0:.2e}, moved {moved:.3e}, reversal error {back:.3e}")
    assert moved > 1e-3
    tol = 1e-12 if prec == "f64" else 1e-5
    assert abs(n1 - n0) / n0 < tol
    assert back < (1e-12 if prec == "f64" else 2e-5)


@pytest.mark.parametrize("autotune", [False, True], ids=["default-variants", "autotuned"])
def test_metric_workload_matches_oracle(gpu, orc, autotune):
    """The 256³ metric workload itself (V = 0, β = 1/ξ², Δt = 5e-4,
    random-phase ψ0): 5 FP64 Strang steps on the GPU against the oracle,
    with the default launch variants and with the autotuned (pipelined)
    ones, relative L2 within 1e-12 -- the FP64 bar, on the configuration the
    speed metric is quoted on."""
    from sfdhip.workloads import metric_params, random_phase_3d
    mp = metric_params(256)
    psi0 = random_phase_3d(256, seed=0)
    f = _field(gpu, psi0, mp["bounds"], np.float64)
    if autotune:
        f.autotune()
        f.set(psi0)
    p = gpu.GrossPitaevskiiParameters(pot=gpu.PotentialZero(f), coeffΔ=-0.5, β=mp["beta"])
    gpu.NumModelSplit2(f, p, mp["dt"], 5, 5).steps(5)
    ref = orc.split_step(psi0.copy(), mp["bounds"], coeffD=-0.5, beta=mp["beta"], dt=mp["dt"], nsteps=5, order=2)
    err = rel_l2(f.to_numpy(), ref)
    print(f"256³ FP64 5 steps vs oracle ({'autotuned' if autotune else 'default'}): {err:.3e}")
    assert err < 1e-12


def test_error_paths(gpu):
    g = gpu.Grid((32, 32), ((-1, 1), (-1, 1)))
    f = gpu.Field(g)
    lib = gpu.load()
    import ctypes
    c = f.ctx()
    rc = lib.sfd_split_step(c.h, None, None, -0.5, 1.0, 0.0, 1e-3, 1, 2, 0)
    assert rc == -1
    rc = lib.sfd_split_step(c.h, ctypes.c_void_p(f.ϕ.data_ptr()), None, -0.5, 1.0, 0.0, 1e-3, 1, 3, 0)
    assert rc == -3
    assert b"order" in lib.sfd_last_error()
    # any n in [3, 1024] and powers of two up to 2048 are accepted (Bluestein
    # lines for the others); larger lines are refused
    gpu.Field(gpu.Grid((48, 32), ((-1, 1), (-1, 1)))).ctx()
    for bad in ((1025, 8), (4096, 2), (8, 1)):
        with pytest.raises(gpu.ShapeError):
            gpu.Field(gpu.Grid(bad, ((-1, 1), (-1, 1)))).ctx()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_nonlinear_phase_accuracy(gpu, prec):
    """solveNL! on ψ = 1 with V = θ/Δt gives exp(-iθ): the device sine/cosine
    (Cody-Waite reduction + Taylor series, library fallback for huge |θ|)
    against numpy over |θ| up to 3e4, multiples of π/2 included.  FP32
    rotates by two lifting steps of -θ/2 (six rounded fused multiply-adds of
    unit size, worst case ~7 ulp, unbiased), hence 6e-7."""
    FT = np.float64 if prec == "f64" else np.float32
    rng = np.random.default_rng(0)
    th = np.concatenate([rng.uniform(-4, 4, 4096), rng.uniform(-200, 200, 4096), rng.uniform(-3e4, 3e4, 2048),
                         (np.arange(-1024, 1024) * (np.pi / 2)), np.linspace(-1e-3, 1e-3, 4096)])
    th = th.reshape(128, 128)
    th = th.astype(FT).astype(np.float64)  # representable in the field precision
    g = gpu.Grid((128, 128), ((0, 1), (0, 1)), FT=FT)
    f = gpu.Field(g)
    f.set(np.ones((128, 128), dtype=np.complex128))
    dt = 0.5
    pot = gpu.PotentialArray(f, th / dt)
    p = gpu.GrossPitaevskiiParameters(pot=pot, coeffΔ=-0.5, β=0.0)
    m = gpu.NumModelSplit2(f, p, dt, 1, 1)
    gpu.solveNL(m, dt)
    got = f.to_numpy()
    exact = np.exp(-1j * (th / dt).astype(FT).astype(np.float64) * dt)
    err = np.max(np.abs(got - exact))
    print(f"{prec}: max |error| {err:.3e}")
    assert err < (4e-16 if prec == "f64" else 6e-7)


def test_lap_rot_and_nl_compose_to_split2(gpu, orc):
    """solveLapRot!(Δt/2) ∘ solveNL!(Δt) ∘ solveLapRot!(Δt/2) equals one timeStep!."""
    shape, bounds = (64, 32, 16), ((-4, 4), (-3, 3), (-2, 2))
    psi = smooth_state(shape, bounds, seed=2)
    V = random_potential(shape, bounds)
    kw = dict(V=V, coeffD=-0.5, beta=4.0, Omega=0.3, dt=3e-3, nsteps=1, order=2)
    ref = orc.split_step(psi.copy(), bounds, **kw)
    f = _field(gpu, psi, bounds)
    p = gpu.GrossPitaevskiiParameters(pot=gpu.PotentialArray(f, V), coeffΔ=-0.5, β=4.0, Ω=0.3)
    m = gpu.NumModelSplit2(f, p, 3e-3, 1, 1)
    gpu.solveLapRot(m, 1.5e-3)
    gpu.solveNL(m, 3e-3)
    gpu.solveLapRot(m, 1.5e-3)
    assert rel_l2(f.to_numpy(), ref) < 1e-12


def test_dt_schedule_cycles_the_phase_cache(gpu, orc):
    """A Δt schedule of 90 distinct steps, with Ω switched on every third
    step, on one context: the phase-table cache (32 entries) fills and
    rebuilds its least recently used tables in place on the stream, and
    grows a table when a rotating step needs more rows.  Every step is
    checked against the oracle run of the same schedule (FP64, 1e-12)."""
    shape, bounds = (64, 32), ((-4, 4), (-3, 3))
    psi = smooth_state(shape, bounds, seed=21)
    V = random_potential(shape, bounds)
    f = _field(gpu, psi, bounds)
    pot = gpu.PotentialArray(f, V)
    ref = psi.copy()
    for k in range(90):
        dt = 1e-3 * (1.0 + k / 97.0)
        Om = 0.3 if k % 3 == 0 else 0.0
        p = gpu.GrossPitaevskiiParameters(pot=pot, coeffΔ=-0.5, β=4.0, Ω=Om)
        gpu.NumModelSplit2(f, p, dt, 1, 1).steps(1)
        orc.split_step(ref, bounds, V=V, coeffD=-0.5, beta=4.0, Omega=Om, dt=dt, nsteps=1, order=2)
        if k % 15 == 14:
            assert rel_l2(f.to_numpy(), ref) < 1e-12, k
    assert rel_l2(f.to_numpy(), ref) < 1e-12
