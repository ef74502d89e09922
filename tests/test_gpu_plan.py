"""The FFT-plan plug on the HIP backend (Plan, computeDerivatives): the
reference's derivative fields (derivatives.jl:27-54, 90-126) built from
sfd_fft_axis in the pen_x layout, checked against numpy spectral derivatives,
and the reference's energy formula over them (GrossPitaevskii.jl:61-136)
against sfd_energy's one-sweep Parseval sums."""
import numpy as np
import pytest

from helpers import mesh, random_potential, rel_l2, smooth_state

pytestmark = pytest.mark.gpu

CASES = {
    "2d": ((64, 32), ((-6, 6), (-3, 3))),
    "3d": ((32, 16, 64), ((-4, 4), (-3, 3), (-5, 5))),
    "2d-512": ((512, 16), ((-8, 8), (-2, 2))),
}


def _numpy_derivs(psi, shape, bounds):
    X = mesh(shape, bounds)
    nd = len(shape)
    out = {}
    for a in range(nd):
        ax = nd - 1 - a
        L = bounds[a][1] - bounds[a][0]
        k = np.fft.fftfreq(shape[a], d=L / shape[a]) * 2 * np.pi
        kb = k.reshape([-1 if i == ax else 1 for i in range(nd)])
        hat = np.fft.fft(psi, axis=ax)
        d = "xyz"[a]
        out["d" + d] = np.fft.ifft(1j * kb * hat, axis=ax)
        out["dd" + d] = np.fft.ifft(-(kb ** 2) * hat, axis=ax)
        if a == 0:
            out["rx"] = np.fft.ifft(1j * X[1] * kb * hat, axis=ax)
        if a == 1:
            out["ry"] = np.fft.ifft(-1j * X[0] * kb * hat, axis=ax)
    return out


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_derivatives_and_energy_through_the_plan(gpu, case, prec):
    shape, bounds = CASES[case]
    FT = np.float64 if prec == "f64" else np.float32
    psi = smooth_state(shape, bounds, seed=21)
    V = random_potential(shape, bounds, amp=3.0)
    g = gpu.Grid(shape, bounds, FT=FT)
    f = gpu.Field(g)
    f.set(psi)
    plan = gpu.Plan(f)
    gf = gpu.GradientField(f)
    gpu.computeDerivatives(gf, plan, f.ϕ)
    # reference from the stored (rounded) field, in FP64
    ref = _numpy_derivs(f.to_numpy().astype(np.complex128), shape, bounds)
    for name, want in ref.items():
        got = getattr(gf, name).cpu().numpy()
        # rounding noise is white in k, and the k^2 of a second derivative
        # lifts it at the high modes, where the smooth field has no weight:
        # in FP32 on 512 points (k_max^2 ~ 1e4) that is ~1e-3
        second = name.startswith("dd")
        tol = (1e-11 if second else 1e-12) if prec == "f64" else (3e-3 if second else 5e-5)
        print(case, prec, name, rel_l2(got, want))
        assert rel_l2(got, want) < tol, (name, rel_l2(got, want))
    # GrossPitaevskii.jl:61-136 over the derivative fields
    beta, Om, cD = 7.0, 0.4, -0.5
    dv = float(np.prod([(b[1] - b[0]) / n for b, n in zip(bounds, shape)]))
    p_ = np.asarray(psi if prec == "f64" else psi.astype(np.complex64), dtype=np.complex128)
    grads = sum(np.sum(np.abs(getattr(gf, "d" + d).cpu().numpy()) ** 2) for d in "xyz"[:len(shape)])
    rot = gf.rx.cpu().numpy() + gf.ry.cpu().numpy()
    EOm = np.sum((1j * np.conj(p_) * (Om * rot)).real) * dv
    ED = (-cD * grads + np.sum(V * np.abs(p_) ** 2)) * dv
    Eb = np.sum(0.5 * beta * np.abs(p_) ** 4) * dv
    pot = gpu.PotentialArray(f, V)
    m = gpu.NumModelSplit2(f, gpu.GrossPitaevskiiParameters(pot=pot, coeffΔ=cD, β=beta, Ω=Om), 1e-3, 1, 1)
    got = gpu.energy_all(m)
    want = (EOm, ED, Eb, -EOm + ED + Eb, -EOm + ED + 2 * Eb)
    scale = max(abs(v) for v in want)
    for g_, w_ in zip(got, want):
        assert abs(g_ - w_) <= (1e-12 if prec == "f64" else 3e-6) * scale, (got, want)
