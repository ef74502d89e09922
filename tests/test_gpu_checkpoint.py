"""Checkpoint and resume on the HIP path: write! / read! of a device field
(Writers.jl:41-134) and solve!'s backups every freqbckp iterations
(NumModels.jl:29-41), from which a run resumes with istart."""
import os

import numpy as np
import pytest

from helpers import random_potential, smooth_state

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _io(gpu):
    from sfdhip import io
    if not os.path.exists(io.IO_PATH):
        pytest.skip("libsfdio.so not built (no HDF5 headers at build time)")


def _model(sfd, psi, bounds, FT, niter, freqbckp, V):
    g = sfd.Grid(tuple(reversed(psi.shape)), bounds, FT=FT)
    f = sfd.Field(g)
    f.set(psi)
    p = sfd.GrossPitaevskiiParameters(pot=sfd.PotentialArray(f, V), coeffΔ=-0.5, β=20.0)
    return sfd.NumModelSplit2(f, p, 2e-3, niter, freqbckp)


@pytest.mark.parametrize("FT", [np.float64, np.float32])
def test_resume_from_a_checkpoint_is_exact(gpu, tmp_path, monkeypatch, FT):
    monkeypatch.chdir(tmp_path)
    shape, bounds = (64, 32), ((-6, 6), (-3, 3))
    psi = smooth_state(shape, bounds, seed=3)
    V = random_potential(shape, bounds, amp=4.0)
    a = _model(gpu, psi, bounds, FT, 10, 10, V)
    a.steps(10)
    gpu.write(gpu.WriterSave(a.f), prefix="res", istep=10)
    a.steps(10)
    b = _model(gpu, np.zeros_like(psi), bounds, FT, 10, 10, V)
    gpu.read(gpu.WriterSave(b.f), prefix="res", istep=10)
    b.steps(10)
    assert np.array_equal(a.f.to_numpy(), b.f.to_numpy())


def test_solve_writes_backups_and_resumes(gpu, tmp_path, monkeypatch):
    """solve! writes res-{istart}.h5, then res-{it}.h5 after each iteration
    with it % freqbckp == 0; a second solve! from istart = 13 on the field
    read back from res-12.h5 continues the history of the first run."""
    monkeypatch.chdir(tmp_path)
    shape, bounds = (64, 32), ((-6, 6), (-3, 3))
    psi = smooth_state(shape, bounds, seed=6)
    V = random_potential(shape, bounds, amp=4.0)
    m = _model(gpu, psi, bounds, np.float64, 20, 4, V)
    res = gpu.solve(m, istart=1, writers=[gpu.WriterSave(m.f)])
    assert len(res) == 21
    assert sorted(os.listdir(tmp_path)) == sorted(f"res-{i}.h5" for i in (1, 4, 8, 12, 16, 20))
    final = m.f.to_numpy()
    # the same history without writers (the same calls between backup points)
    m2 = _model(gpu, psi, bounds, np.float64, 20, 4, V)
    res2 = gpu.solve(m2)
    assert np.array_equal(np.array(res), np.array(res2))
    # resume: the state after iteration 12, then iterations 13..21 with the
    # same backup frequency, hence the same calls
    m3 = _model(gpu, np.zeros_like(psi), bounds, np.float64, 8, 4, V)
    gpu.read(gpu.WriterSave(m3.f), prefix="res", istep=12)
    res3 = gpu.solve(m3, istart=13)
    assert np.allclose(np.array(res3)[:, 2:], np.array(res)[12:, 2:], rtol=0, atol=1e-13 * abs(res[0][5]))
    assert np.array_equal(m3.f.to_numpy(), final)
