"""Golden fixtures (tests/golden/split_cases.*): the oracle reproduces them,
an independent numpy/pocketfft twin of the step agrees with them, and the
HIP path matches them."""
import json
import os

import numpy as np
import pytest

from helpers import TOLS, gpu_run, mesh, rel_l2

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _cases():
    meta = json.load(open(os.path.join(HERE, "split_cases.json")))
    data = np.load(os.path.join(HERE, "split_cases.npz"))
    out = []
    for name, m in meta.items():
        V = data[name + "/V"] if (name + "/V") in data.files else None
        out.append((name, tuple(m["shape"]), tuple(tuple(b) for b in m["bounds"]), m["params"],
                    data[name + "/psi0"], V, data[name + "/out"]))
    return out


CASES = _cases()
IDS = [c[0] for c in CASES]


def numpy_split(psi, bounds, V, coeffD, beta, Omega, dt, nsteps, order, imag):
    """Independent restatement with numpy.fft (pocketfft), axis by axis in the
    order x, y, z (splitting.jl:156-209)."""
    psi = psi.astype(np.complex128).copy()
    nd = psi.ndim
    shape = tuple(reversed(psi.shape))
    X = mesh(shape, bounds)
    d = [(b[1] - b[0]) / n for b, n in zip(bounds, shape)]
    xis = []
    for a in range(nd):
        mult = (6.283185307179586 / d[a]) / shape[a]
        k = np.arange(shape[a])
        k = np.where(k < (shape[a] + 1) // 2, k, k - shape[a]).astype(np.float64) * mult
        shp = [1] * nd
        shp[nd - 1 - a] = shape[a]
        xis.append(k.reshape(shp))
    fac = (lambda a: np.exp(a)) if imag else (lambda a: np.exp(1j * a))

    def lap_rot(p, tau):
        for a in range(nd):
            ax = nd - 1 - a
            xi = xis[a]
            if a == 0:
                arg = (coeffD * (xi * xi) - (Omega * X[1]) * xi) * tau
            elif a == 1:
                arg = (coeffD * (xi * xi) + (Omega * X[0]) * xi) * tau
            else:
                arg = (coeffD * (xi * xi)) * tau
            p = np.fft.ifft(fac(arg) * np.fft.fft(p, axis=ax), axis=ax)
        return p

    def nl(p):
        th = ((V if V is not None else 0.0) + beta * np.abs(p) ** 2) * dt
        return (np.exp(-th) if imag else np.exp(-1j * th)) * p

    dv = float(np.prod(d))
    for _ in range(nsteps):
        if order == 2:
            psi = lap_rot(nl(lap_rot(psi, dt * 0.5)), dt * 0.5)
        else:
            psi = nl(lap_rot(psi, dt))
        if imag:
            psi = psi / (np.sqrt(np.sum(np.abs(psi) ** 2)) * np.sqrt(dv))
    return psi


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_oracle_reproduces_fixture(orc, case):
    name, shape, bounds, prm, psi0, V, out = case
    got = orc.split_step(np.ascontiguousarray(psi0.copy()), bounds, V=V, **prm)
    assert rel_l2(got, out) < 1e-14


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_numpy_twin_agrees(case):
    name, shape, bounds, prm, psi0, V, out = case
    got = numpy_split(psi0, bounds, V, **prm)
    assert rel_l2(got, out) < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f64", "f32", "f32x"])
@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_hip_matches_fixture(gpu, case, prec):
    name, shape, bounds, prm, psi0, V, out = case
    got = gpu_run(gpu, np.ascontiguousarray(psi0), bounds, V=V, prec=prec, **prm)
    err = rel_l2(got, out)
    print(f"{name} {prec}: {err:.3e}")
    assert err < TOLS[prec]
