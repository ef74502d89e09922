"""GPU adjoint split step (sfd_adjoint): against the FP64 oracle adjoint,
independence of the checkpoint interval, and the 256² control problem
against central finite differences."""
import numpy as np
import pytest
import torch

from helpers import PRECS, random_potential, rel_l2, smooth_state

pytestmark = pytest.mark.gpu


def _model(gpu, psi0, bounds, V, beta, Om=0.0, dt=2e-3, prec="f64"):
    FT, compute = PRECS[prec]
    g = gpu.Grid(tuple(reversed(psi0.shape)), bounds, FT=FT, compute=compute)
    f = gpu.Field(g)
    f.set(psi0)
    pot = gpu.PotentialArray(f, V) if V is not None else gpu.PotentialZero(f)
    p = gpu.GrossPitaevskiiParameters(pot=pot, coeffΔ=-0.5, β=beta, Ω=Om)
    return f, gpu.NumModelSplit2(f, p, dt, 1, 1)


@pytest.mark.parametrize("case", [((32, 16), 0.0), ((32, 16), 0.4), ((16, 8, 8), 0.3)], ids=["2d", "2d-rot", "3d"])
def test_adjoint_matches_oracle(gpu, orc, case):
    shape, Om = case
    bounds = tuple((-3.0, 3.0) for _ in shape)
    psi0 = smooth_state(shape, bounds, seed=1)
    psi0 = psi0 / orc.norm(psi0, bounds)
    tgt = smooth_state(shape, bounds, seed=8)
    tgt = tgt / orc.norm(tgt, bounds)
    V = random_potential(shape, bounds, amp=3.0)
    kw = dict(V=V, coeffD=-0.5, beta=6.0, Omega=Om, dt=2e-3, nsteps=17)
    loss, gV, gb = orc.adjoint(psi0, tgt, bounds, **kw)
    f, m = _model(gpu, psi0, bounds, V, 6.0, Om)
    t = torch.as_tensor(np.ascontiguousarray(tgt)).to("cuda")
    l2, gV2, gb2 = gpu.adjoint_gradient(m, t, 17)
    assert l2 == pytest.approx(loss, rel=1e-12, abs=1e-14)
    assert gb2 == pytest.approx(gb, rel=1e-10)
    g2 = gV2.cpu().numpy()
    assert np.sqrt(np.sum((g2 - gV) ** 2) / np.sum(gV ** 2)) < 1e-10
    assert rel_l2(f.to_numpy(), psi0) < 1e-15  # the model's field is left unchanged


def test_checkpoint_interval_does_not_change_the_gradient(gpu):
    shape, bounds = (64, 32), ((-4.0, 4.0), (-3.0, 3.0))
    psi0 = smooth_state(shape, bounds, seed=2)
    tgt = smooth_state(shape, bounds, seed=3)
    V = random_potential(shape, bounds)
    f, m = _model(gpu, psi0, bounds, V, 9.0)
    t = torch.as_tensor(tgt).to("cuda")
    ref = gpu.adjoint_gradient(m, t, 30, checkpoint_every=30)
    for every in (1, 4, 7, 0):
        l, gV, gb = gpu.adjoint_gradient(m, t, 30, checkpoint_every=every)
        assert l == ref[0]
        assert gb == ref[2]
        assert torch.equal(gV, ref[1])


def _control(gpu, nsteps_gs=3000):
    """The 256² control problem: ψ0 = ground state of ½|x|² at β = 500 (FP64
    imaginary time), ψ_target = ψ0 shifted by +1.5 in x, learnable
    V(a) = ½|x|² + Σ_j a_j exp(-|x - c_j|²/0.5) on an 8×8 lattice."""
    from sfdhip.workloads import control_problem, shift_x
    basis, a, V0, bounds = control_problem()
    n = basis.shape[-1]
    g = gpu.Grid((n, n), bounds)
    f = gpu.Field(g)
    x, y = f.coords()
    f.set(np.ascontiguousarray(np.broadcast_to(np.exp(-0.1 * (x ** 2 + y ** 2)) + 0j, g.shape)))
    gpu.normalize(f)
    p0 = gpu.GrossPitaevskiiParameters(pot=gpu.PotentialArray(f, V0), coeffΔ=-0.5, β=500.0)
    gpu.BatchImagSolver([f], [p0], 5e-3).run(max_steps=nsteps_gs, tol=1e-12, check_every=100)
    psi0 = f.to_numpy()
    tgt = np.ascontiguousarray(shift_x(psi0, 1.5, 8.0))
    return basis, a, V0, bounds, psi0, tgt


def test_control_gradient_matches_finite_differences(gpu):
    """∂L/∂a_j (8 seeded coordinates) and ∂L/∂g of the 256² problem (β = 500,
    Δt = 1e-3, 800 steps) against central differences with h = 1e-4:
    relative agreement within 1e-6."""
    basis, a, V0, bounds, psi0, tgt = _control(gpu)
    Vd = torch.as_tensor(V0 + np.tensordot(a, basis, axes=1)).to("cuda")
    B = torch.as_tensor(basis).to("cuda")
    f, m = _model(gpu, psi0, bounds, Vd.cpu().numpy(), 500.0, dt=1e-3)
    t = torch.as_tensor(tgt).to("cuda")
    loss, gV, gb = gpu.adjoint_gradient(m, t, 800)
    ga = gpu.project(f, gV, B)
    pot = m.param.pot

    def L_of(Vnew, beta):
        pot.V.copy_(Vnew)
        m.param.β = beta
        return gpu.adjoint_gradient(m, t, 800, checkpoint_every=800)[0]

    # 8 seeded coordinates among the Gaussians centred inside the condensate
    # (|c_j| < 4.5, Thomas-Fermi radius ≈ 5): outside it dL/da_j ~ 1e-12,
    # below the resolution of a central difference of an FP64 loss.
    L_box, lat = 8.0, 8
    hc = 2 * L_box / lat
    centres = np.array([(-L_box + (i + 0.5) * hc, -L_box + (k + 0.5) * hc) for k in range(lat) for i in range(lat)])
    inside = np.flatnonzero(np.hypot(centres[:, 0], centres[:, 1]) < 4.5)
    rng = np.random.default_rng(11)
    js = rng.choice(inside, size=8, replace=False)
    h = 1e-4
    worst = 0.0
    for j in js:
        fd = (L_of(Vd + h * B[j], 500.0) - L_of(Vd - h * B[j], 500.0)) / (2 * h)
        rel = abs(fd - ga[j]) / max(abs(fd), 1e-300)
        worst = max(worst, rel)
        print(f"a_{j}: adjoint {ga[j]:+.10e} fd {fd:+.10e} rel {rel:.2e}")
        assert rel < 1e-6
    fdb = (L_of(Vd, 500.0 + h) - L_of(Vd, 500.0 - h)) / (2 * h)
    print(f"g: adjoint {gb:+.10e} fd {fdb:+.10e}; loss {loss:.6f}; worst rel {worst:.2e}")
    assert abs(fdb - gb) / abs(fdb) < 1e-6
    assert 0.0 < loss < 1.0


@pytest.mark.parametrize("prec", ["f32", "f32x"])
def test_reduced_precision_adjoint_tracks_fp64(gpu, orc, prec):
    """FP32 (and FP32-storage/FP64-arithmetic) adjoints against the FP64 one:
    relative loss error 1e-4 (1e-6), gradient relative L2 1e-3 (1e-5)."""
    shape, bounds = (128, 64), ((-6.0, 6.0), (-4.0, 4.0))
    psi0 = smooth_state(shape, bounds, seed=4)
    psi0 = psi0 / orc.norm(psi0, bounds)
    tgt = smooth_state(shape, bounds, seed=5)
    tgt = tgt / orc.norm(tgt, bounds)
    V = random_potential(shape, bounds, amp=4.0)
    res = {}
    for p in ("f64", prec):
        f, m = _model(gpu, psi0, bounds, V, 8.0, dt=2e-3, prec=p)
        t = torch.as_tensor(tgt).to(f.ϕ.dtype).to("cuda")
        l, gV, gb = gpu.adjoint_gradient(m, t, 100)
        res[p] = (l, gV.cpu().numpy(), gb)
    l64, g64, b64 = res["f64"]
    l, g, b = res[prec]
    tol = 1e-4 if prec == "f32" else 1e-6
    print(f"{prec}: loss {l:.8f} vs {l64:.8f}, dL/dbeta {b:.6e} vs {b64:.6e}")
    assert abs(l - l64) / abs(l64) < tol
    assert np.sqrt(np.sum((g - g64) ** 2) / np.sum(g64 ** 2)) < 10 * tol
    assert abs(b - b64) / abs(b64) < 10 * tol


def test_autograd_function_matches_adjoint_and_finite_differences(gpu, orc):
    """fidelity_loss (a torch.autograd.Function over sfd_adjoint):
    loss.backward() gives the adjoint's dL/dV and dL/dβ, and torch's chain
    rule through V(a) = V0 + Σ_j a_j B_j gives dL/da_j, which central
    differences (h = 1e-4) confirm to 1e-6 relative."""
    shape, bounds = (64, 32), ((-4.0, 4.0), (-3.0, 3.0))
    psi0 = smooth_state(shape, bounds, seed=2)
    psi0 = psi0 / orc.norm(psi0, bounds)
    tgt = smooth_state(shape, bounds, seed=3)
    tgt = tgt / orc.norm(tgt, bounds)
    V0 = random_potential(shape, bounds, amp=3.0)
    f, m = _model(gpu, psi0, bounds, V0, 9.0)
    t = torch.as_tensor(tgt).to("cuda")
    l_ref, gV_ref, gb_ref = gpu.adjoint_gradient(m, t, 30)

    V = torch.as_tensor(V0).to("cuda").requires_grad_(True)
    beta = torch.tensor(9.0, dtype=torch.float64, device="cuda", requires_grad=True)
    loss = gpu.fidelity_loss(m, V, beta, t, 30)
    loss.backward()
    assert loss.item() == l_ref
    assert torch.equal(V.grad, gV_ref)
    assert beta.grad.item() == gb_ref

    x = torch.linspace(bounds[0][0], bounds[0][1], shape[0] + 1, dtype=torch.float64, device="cuda")[:-1]
    y = torch.linspace(bounds[1][0], bounds[1][1], shape[1] + 1, dtype=torch.float64, device="cuda")[:-1]
    Y, X = torch.meshgrid(y, x, indexing="ij")
    B = torch.stack([torch.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / 0.5) for cx, cy in ((0.5, 0.0), (-1.0, 0.5), (0.0, -1.0))])
    V0d = torch.as_tensor(V0).to("cuda")
    a = torch.zeros(3, dtype=torch.float64, device="cuda", requires_grad=True)
    loss2 = gpu.fidelity_loss(m, V0d + torch.tensordot(a, B, dims=1), 9.0, t, 30)
    loss2.backward()
    h = 1e-4
    with torch.no_grad():
        for j in range(3):
            e = torch.zeros(3, dtype=torch.float64, device="cuda")
            e[j] = h
            lp = gpu.fidelity_loss(m, V0d + torch.tensordot(a + e, B, dims=1), 9.0, t, 30).item()
            lm = gpu.fidelity_loss(m, V0d + torch.tensordot(a - e, B, dims=1), 9.0, t, 30).item()
            fd = (lp - lm) / (2 * h)
            print(f"a_{j}: autograd {a.grad[j].item():+.10e} fd {fd:+.10e}")
            assert abs(a.grad[j].item() - fd) <= 1e-6 * abs(fd)
    assert rel_l2(f.to_numpy(), psi0) < 1e-15  # the model's field is left unchanged


def test_huge_checkpoint_interval_is_clamped_or_rejected(gpu):
    """checkpoint_every is clamped to nsteps (one segment, nsteps saved states,
    not three times that), and a step count whose checkpoint storage would
    overflow size_t is rejected with SFD_ERR_ARG before any allocation."""
    shape, bounds = (32, 16), ((-3.0, 3.0), (-3.0, 3.0))
    psi0 = smooth_state(shape, bounds, seed=4)
    tgt = smooth_state(shape, bounds, seed=5)
    f, m = _model(gpu, psi0, bounds, random_potential(shape, bounds), 5.0)
    t = torch.as_tensor(tgt).to("cuda")
    ref = gpu.adjoint_gradient(m, t, 12, checkpoint_every=12)
    for every in (13, 1 << 40, (1 << 62)):
        l, gV, gb = gpu.adjoint_gradient(m, t, 12, checkpoint_every=every)
        assert l == ref[0] and gb == ref[2] and torch.equal(gV, ref[1])
    from sfdhip import _lib as L
    with pytest.raises(L.ArgError, match="overflows"):
        gpu.adjoint_gradient(m, t, 1 << 62, checkpoint_every=1 << 62)


@pytest.mark.parametrize("nside", ["1", "2"])
def test_side_stream_count_does_not_change_the_gradient(gpu, orc, nside, monkeypatch):
    """The recompute runs on one or two side streams (forced through
    SFD_ADJOINT_SIDE_STREAMS): results are bitwise those of the default and
    equal the oracle adjoint, for several checkpoint intervals."""
    shape, bounds = (32, 16), ((-3.0, 3.0), (-3.0, 3.0))
    psi0 = smooth_state(shape, bounds, seed=1)
    psi0 = psi0 / orc.norm(psi0, bounds)
    tgt = smooth_state(shape, bounds, seed=8)
    tgt = tgt / orc.norm(tgt, bounds)
    V = random_potential(shape, bounds, amp=3.0)
    f, m = _model(gpu, psi0, bounds, V, 6.0, 0.4)
    t = torch.as_tensor(np.ascontiguousarray(tgt)).to("cuda")
    monkeypatch.delenv("SFD_ADJOINT_SIDE_STREAMS", raising=False)
    ref = {e: gpu.adjoint_gradient(m, t, 17, checkpoint_every=e) for e in (1, 3, 5, 0)}
    monkeypatch.setenv("SFD_ADJOINT_SIDE_STREAMS", nside)
    for e, r in ref.items():
        l, gV, gb = gpu.adjoint_gradient(m, t, 17, checkpoint_every=e)
        assert l == r[0] and gb == r[2] and torch.equal(gV, r[1])
    loss, gVo, gbo = orc.adjoint(psi0, tgt, bounds, V=V, coeffD=-0.5, beta=6.0, Omega=0.4, dt=2e-3, nsteps=17)
    assert gb == pytest.approx(gbo, rel=1e-10)
    assert np.sqrt(np.sum((gV.cpu().numpy() - gVo) ** 2) / np.sum(gVo ** 2)) < 1e-10


def test_full_phase_cache_does_not_change_the_gradient(gpu):
    """With the phase-table cache already holding 32 unrelated tables, the
    adjoint builds its own tables while the cache is pinned and the result is
    bitwise unchanged."""
    shape, bounds = (32, 32), ((-3.0, 3.0), (-3.0, 3.0))
    psi0 = smooth_state(shape, bounds, seed=6)
    tgt = smooth_state(shape, bounds, seed=7)
    f, m = _model(gpu, psi0, bounds, random_potential(shape, bounds), 4.0, 0.2)
    t = torch.as_tensor(tgt).to("cuda")
    ref = gpu.adjoint_gradient(m, t, 9, checkpoint_every=2)
    for k in range(40):  # 40 distinct taus x 2 axes on the same context: the cache is full of other tables
        gpu.solveLapRot(m, 1e-3 * (k + 1))
    f.set(psi0)
    l, gV, gb = gpu.adjoint_gradient(m, t, 9, checkpoint_every=2)
    assert l == ref[0] and gb == ref[2] and torch.equal(gV, ref[1])
    l, gV, gb = gpu.adjoint_gradient(m, t, 9, checkpoint_every=2)
    assert l == ref[0] and gb == ref[2] and torch.equal(gV, ref[1])
