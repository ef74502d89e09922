"""The stability limit of the unfiltered Strang split, pinned on the FP64 CPU
oracle at 64³ with the two dimensionless numbers of the 256³ metric
workload: D·|coeffΔ|·k_max²·Δt (the phase of the corner mode per step) and
βΔt.  The reference documents the same blow-up of its split without
dealiasing (splitting.jl:36-46; docs/src/grosspitaevskii/nummodel.md:50-60,
E/E0 = 338 at |coeffΔ|k_max²Δt = 2.021 in 2D).

At the metric's 12.3 two FP64 runs whose initial states differ by one FP32
rounding (3e-8) decorrelate to O(1) within a few hundred steps and the
energy grows by more than 10x; at 2.46 (the Δt used for the precision
gates) the same pair stays at its initial distance and the energy is kept
to 1e-3.  So at the metric Δt no precision can be compared with another
after 1000 steps, while the step itself costs the same."""
import math

import numpy as np
import pytest

from helpers import rel_l2


def _pair(orc, number, steps, dt256, n=64):
    """Two oracle runs, ψ0 and ψ0 rounded to FP32; dt256 is the 256³ step with
    the same D|coeffΔ|k_max²Δt, whose βΔt the 64³ run must share."""
    from sfdhip.workloads import metric_params, random_phase_3d
    mp = metric_params(n)
    kmax = n // 2
    dt = number / (3 * 0.5 * kmax ** 2)
    assert mp["beta"] * dt == pytest.approx(metric_params(256)["beta"] * dt256, rel=3e-3)
    a = random_phase_3d(n, seed=0)
    b = a.astype(np.complex64).astype(np.complex128)
    d0 = rel_l2(b, a)
    E0 = orc.energy(a, mp["bounds"], beta=mp["beta"])[3]
    orc.split_step(a, mp["bounds"], beta=mp["beta"], dt=dt, nsteps=steps, order=2)
    orc.split_step(b, mp["bounds"], beta=mp["beta"], dt=dt, nsteps=steps, order=2)
    E = orc.energy(a, mp["bounds"], beta=mp["beta"])[3]
    return d0, rel_l2(b, a), E / E0


def test_metric_dt_is_past_the_stability_limit(orc):
    d0, d, growth = _pair(orc, 12.3, 250, 5e-4)
    assert 1e-8 < d0 < 1e-7
    assert d > 0.5, d          # an FP32 rounding of ψ0 grows to O(1)
    assert growth > 10.0, growth


def test_stable_dt_keeps_the_perturbation(orc):
    d0, d, growth = _pair(orc, 2.46, 250, 1e-4)
    assert d < 1.2 * d0, (d0, d)
    assert abs(growth - 1.0) < 1e-3, growth
