"""solve!'s loop (NumModels.jl:32-37) in one call, sfd_split_step_energy:
the energy of the state before each step comes out of that step's own
first-half sweeps (real time, Omega = 0) or from a separate sfd_energy
(rotation, imaginary time).  In real time with Omega = 0 the Strang steps
are chained (x, y, [z N z], y, [x x], ...): the x sums of the state between
two steps come out of the second round trip of the [x x] pass.
Checked against sfd_energy + sfd_split_step called in turn on a second copy
of the state, and (FP64) against the oracle's energy of the oracle's state
after the same steps."""
import numpy as np
import pytest

from helpers import PRECS, random_potential, rel_l2, smooth_state

pytestmark = pytest.mark.gpu

CASES = {
    "2d": ((64, 32), ((-6, 6), (-3, 3))),
    "3d": ((32, 16, 64), ((-4, 4), (-3, 3), (-5, 5))),
    "3d-512": ((512, 8, 512), ((-8, 8), (-2, 2), (-8, 8))),  # 512-point lines on x and z (halved lines)
}


def _model(sfd, psi, bounds, V, prec, *, beta, Omega, dt, order, imag):
    FT, compute = PRECS[prec]
    g = sfd.Grid(tuple(reversed(psi.shape)), bounds, FT=FT, compute=compute)
    f = sfd.Field(g)
    f.set(psi)
    pot = sfd.PotentialArray(f, V) if V is not None else sfd.PotentialZero(f)
    p = sfd.GrossPitaevskiiParameters(pot=pot, coeffΔ=-0.5, β=beta, Ω=Omega)
    if imag:
        return sfd.NumModelSplitImag(f, p, dt, 5, 5, order=order)
    return (sfd.NumModelSplit1 if order == 1 else sfd.NumModelSplit2)(f, p, dt, 5, 5)


@pytest.mark.parametrize("prec", ["f64", "f32", "f32x"])
@pytest.mark.parametrize("mode", ["real2", "real1", "imag2", "imag1"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_fused_energy_matches_energy_then_step(gpu, case, mode, prec):
    shape, bounds = CASES[case]
    psi = smooth_state(shape, bounds, seed=5)
    V = random_potential(shape, bounds, amp=4.0)
    kw = dict(beta=20.0, Omega=0.0, dt=2e-3, order=int(mode[-1]), imag=mode.startswith("imag"))
    a = _model(gpu, psi, bounds, V, prec, **kw)
    b = _model(gpu, psi, bounds, V, prec, **kw)
    k = 7
    hist = gpu.steps_with_energy(a, k)
    ref = []
    for _ in range(k):
        ref.append(gpu.energy_all(b))
        b.steps(1)
    ref = np.array(ref)
    scale = np.abs(ref).max()
    err = np.abs(hist - ref).max() / scale
    dstate = rel_l2(a.f.to_numpy(), b.f.to_numpy())
    print(f"{case} {mode} {prec}: energy {err:.2e}, state {dstate:.2e}")
    assert hist.shape == (k, 5)
    if kw["imag"]:
        # the damped flows change the y/z spectra: one sfd_energy per step,
        # the same calls as the loop on `b`
        assert np.array_equal(hist, ref)
        assert dstate == 0.0
        return
    # FP32 fields: the y/z sums come from the rounded x-flowed state, and in
    # real2 the chained steps round the FP32 field at other points than the
    # single steps of `b` (a state difference of ~2e-7 after 7 steps)
    bar = {"f64": 1e-12, "f32": 3e-7, "f32x": 3e-7 if mode == "real2" else 3e-8}[prec]
    assert err < bar, (hist, ref)
    # the energy sweeps run the same round trips as the plain steps; chained
    # Strang steps (real2) order the commuting flows differently from the
    # single steps of `b`, a difference of rounding
    assert dstate < {"f64": 1e-13, "f32": 3e-6, "f32x": 3e-6}[prec]


@pytest.mark.parametrize("case", sorted(CASES))
def test_fused_energy_matches_oracle(gpu, orc, case):
    shape, bounds = CASES[case]
    psi = smooth_state(shape, bounds, seed=8)
    V = random_potential(shape, bounds, amp=4.0)
    m = _model(gpu, psi, bounds, V, "f64", beta=20.0, Omega=0.0, dt=2e-3, order=2, imag=False)
    hist = gpu.steps_with_energy(m, 4)
    x = psi.copy()
    for s in range(4):
        ref = np.array(orc.energy(x, bounds, V=V, coeffD=-0.5, beta=20.0, Omega=0.0))
        assert np.abs(hist[s] - ref).max() <= 1e-12 * np.abs(ref).max(), (s, hist[s], ref)
        orc.split_step(x, bounds, V=V, coeffD=-0.5, beta=20.0, Omega=0.0, dt=2e-3, nsteps=1, order=2)
    assert rel_l2(m.f.to_numpy(), x) < 1e-12


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_rotation_takes_separate_energies(gpu, orc, prec):
    """With Omega != 0 the y sums after the x flow would differ from the
    state's: the call then evaluates sfd_energy before each step."""
    shape, bounds = CASES["2d"]
    psi = smooth_state(shape, bounds, seed=9)
    V = random_potential(shape, bounds, amp=4.0)
    kw = dict(beta=20.0, Omega=0.6, dt=2e-3, order=2, imag=False)
    a = _model(gpu, psi, bounds, V, prec, **kw)
    b = _model(gpu, psi, bounds, V, prec, **kw)
    hist = gpu.steps_with_energy(a, 5)
    ref = []
    for _ in range(5):
        ref.append(gpu.energy_all(b))
        b.steps(1)
    assert np.array_equal(hist, np.array(ref))
    assert np.array_equal(a.f.to_numpy(), b.f.to_numpy())
    assert abs(hist[0, 0]) > 1e-6  # the rotation term is present


def test_solve_history(gpu):
    """solve! returns niter+1 rows (nkrylov, total, EΩ, EΔ, Eβ, E) (NumModels.jl:30-37);
    a Split2 step contributes 2."""
    shape, bounds = CASES["2d"]
    psi = smooth_state(shape, bounds, seed=3)
    V = random_potential(shape, bounds, amp=4.0)
    m = _model(gpu, psi, bounds, V, "f64", beta=20.0, Omega=0.0, dt=2e-3, order=2, imag=False)
    e0 = gpu.energy(m)
    res = gpu.solve(m)
    assert len(res) == m.niter + 1
    assert [r[:2] for r in res] == [(2, 2 * (i + 1)) for i in range(m.niter + 1)]
    assert np.allclose(res[0][2:], e0, rtol=1e-12, atol=1e-14)
    # real-time Split2 conserves the energy to O(dt^2) over these few steps
    E = np.array([r[5] for r in res])
    assert np.abs(E - E[0]).max() < 1e-3 * abs(E[0])


def test_zero_steps_and_null_output(gpu):
    import ctypes
    shape, bounds = CASES["2d"]
    m = _model(gpu, smooth_state(shape, bounds), bounds, None, "f64", beta=1.0, Omega=0.0, dt=1e-3, order=2,
               imag=False)
    assert gpu.steps_with_energy(m, 0).shape == (0, 5)
    c = m.f.ctx()
    rc = c.lib.sfd_split_step_energy(c.h, ctypes.c_void_p(m.f.ϕ.data_ptr()), None, -0.5, 1.0, 0.0, 1e-3, 3, 2, 0,
                                     None)
    assert rc != 0


def test_fused_energy_runs_chained_steps(gpu):
    """A real-time Split2 energy loop of k steps at Omega = 0 runs k - 1 [x x]
    passes (sweep-timer slot 5) and k [z N z] passes."""
    import ctypes
    shape, bounds = CASES["3d"]
    psi = smooth_state(shape, bounds, seed=5)
    V = random_potential(shape, bounds, amp=4.0)
    m = _model(gpu, psi, bounds, V, "f64", beta=20.0, Omega=0.0, dt=2e-3, order=2, imag=False)
    lib, ctx = gpu.load(), m.f.ctx()
    gpu._lib.check(lib.sfd_ctx_time_sweeps(ctx.h, 1))
    gpu._lib.check(lib.sfd_ctx_sweep_times(ctx.h, None, None, 1))
    gpu.steps_with_energy(m, 7)
    cnt = (ctypes.c_int64 * 6)()
    gpu._lib.check(lib.sfd_ctx_sweep_times(ctx.h, None, cnt, 1))
    gpu._lib.check(lib.sfd_ctx_time_sweeps(ctx.h, 0))
    assert cnt[5] == 6 and cnt[3] == 7 and cnt[0] == 2, list(cnt)
