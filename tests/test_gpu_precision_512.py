"""The 512³ precision workload (FP32 bulk arithmetic, FP64 reductions, on
the random-phase condensate of the 256³ workload with ξ = 2Δ, so
β = (512/4π)² ≈ 1660) at full size: invariants in FP32 and FP64, FP64
parity with the oracle, the FP32 and F32X error against FP64 at the stable
Δt, and the instability at the configured Δt = 2.5e-4."""
import math

import numpy as np
import pytest
import torch

from helpers import rel_l2

pytestmark = pytest.mark.gpu

N = 512


def _field(gpu, prec, psi0, dt):
    from sfdhip.workloads import metric_params
    from helpers import PRECS
    mp = metric_params(N)
    FT, compute = PRECS[prec]
    g = gpu.Grid((N, N, N), mp["bounds"], FT=FT, compute=compute)
    f = gpu.Field(g)
    f.set(psi0)
    p = gpu.GrossPitaevskiiParameters(pot=gpu.PotentialZero(f), coeffΔ=-0.5, β=mp["beta"])
    return f, gpu.NumModelSplit2(f, p, dt, 1, 1), mp


@pytest.fixture(scope="module")
def psi0():
    from sfdhip.workloads import random_phase_3d
    return random_phase_3d(N, seed=0)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_512_norm_and_time_reversal(gpu, psi0, prec):
    """20 Strang steps at the stable Δt = 2.5e-5 keep the norm (FP64 1e-12,
    FP32 1e-6) and 20 steps with -Δt return ψ0 (FP64 1e-11; FP32 5e-6, the
    rounding of 40 steps: 2.05e-6 measured)."""
    f, m, mp = _field(gpu, prec, psi0, 2.5e-5)
    n0 = gpu.norm(f)
    m.steps(20)
    n1 = gpu.norm(f)
    assert abs(n1 / n0 - 1) < (1e-12 if prec == "f64" else 1e-6)
    m.Δt = -2.5e-5
    m.steps(20)
    err = rel_l2(f.to_numpy(), psi0)
    assert err < (1e-11 if prec == "f64" else 5e-6), err


def test_512_fp64_matches_oracle(gpu, orc, psi0):
    """Two FP64 Strang steps at the configured Δt = 2.5e-4 against the CPU
    oracle (the reference's x, y, z, N, x, y, z order): 1e-12."""
    f, m, mp = _field(gpu, "f64", psi0, 2.5e-4)
    m.steps(2)
    ref = np.ascontiguousarray(psi0.copy())
    orc.split_step(ref, mp["bounds"], beta=mp["beta"], dt=2.5e-4, nsteps=2, order=2)
    err = rel_l2(f.to_numpy(), ref)
    assert err < 1e-12, err


def test_512_precision_study_at_the_stable_dt(gpu):
    """2000 Strang steps at Δt = 2.5e-5 (D|coeffΔ|k_max²Δt = 2.46): FP32 ends
    within 1.25e-5 of FP64 (9.9e-6 measured), its error growing as √t (unbiased
    rounding: 4x the steps give well under 4x the error); norm drift below
    1e-7; the FP32 energy drift within 1e-6 of FP64's own (splitting) drift;
    F32X within 4e-6."""
    from sfdhip.precision import run_study
    st = run_study(n=N, steps=2000, dt=2.5e-5, checkpoints=4, with_rounded=False)
    rows = st["rows"]
    print([(r["step"], r["f32"]["rel_l2_vs_f64"], r["f32x"]["rel_l2_vs_f64"]) for r in rows])
    first, end = rows[0], rows[-1]
    assert end["f32"]["rel_l2_vs_f64"] < 1.25e-5
    assert end["f32"]["rel_l2_vs_f64"] < 3.0 * first["f32"]["rel_l2_vs_f64"]  # sqrt(4) = 2 for a random walk
    assert end["f32"]["norm_drift"] < 1e-7
    assert abs(end["f32"]["energy_drift"] - end["f64"]["energy_drift"]) < 1e-6
    assert end["f32x"]["rel_l2_vs_f64"] < 4e-6
    assert end["f64"]["norm_drift"] < 1e-12


def test_512_configured_dt_is_unstable(gpu):
    """At the configured Δt = 2.5e-4 (D|coeffΔ|k_max²Δt = 24.6) an FP64 run
    started from ψ0 rounded to FP32 is O(1) away from the FP64 run after 300
    steps and the FP64 energy has grown several fold: no precision can be
    compared at this Δt."""
    from sfdhip.precision import run_study
    st = run_study(n=N, steps=300, dt=2.5e-4, checkpoints=3, with_rounded=True)
    end = st["rows"][-1]
    print({k: v for k, v in end.items() if isinstance(v, dict)})
    assert end["f64r"]["rel_l2_vs_f64"] > 0.5
    assert end["f64"]["energy_drift"] > 2.0
