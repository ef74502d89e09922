"""Precision study of the split step: complex FP32 bulk arithmetic (FFT
sweeps, phases) against an FP64 run of the same workload on the GPU.

Three runs advance from one seeded initial state:

* ``f64``  -- the FP64 reference trajectory;
* ``f32``  -- the FP32 field with FP32 arithmetic (FP64 geometry; every
  rotation -- DFT constants, twiddles, kinetic and nonlinear phases -- as
  lifting steps whose FP32 round trips carry no systematic error; FP64
  reductions);
* ``f32x`` -- the FP32 field with FP64 in-register arithmetic and tables
  (same HBM bytes as ``f32``);
* ``f64r`` -- FP64 started from the initial state rounded to FP32, which
  isolates how much of the FP32-vs-FP64 gap is the dynamics amplifying a
  3e-8 perturbation rather than FP32 arithmetic.

At every checkpoint the study reports the relative L2 distance to the FP64
run, and each run's norm drift |N(t)/N0 - 1| and energy drift
|E(t)/E0 - 1|.  Energies come from ``sfd_energy`` (FP64 sums).

Usage: ``python -m sfdhip.precision --n 256 --steps 1000 --dt 5e-4``.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np


def _rel_l2_dev(a, b):
    import torch
    d = (a.to(torch.complex128) - b.to(torch.complex128)).abs().pow(2).sum()
    return math.sqrt(float(d) / float(b.to(torch.complex128).abs().pow(2).sum()))


def run_study(n=256, steps=1000, dt=5e-4, beta=None, checkpoints=8, seed=0, bounds=None, order=2,
              with_rounded=True, log=None):
    import torch
    from . import api as sfd
    from .workloads import metric_params, random_phase_3d

    mp = metric_params(n)
    if beta is None:
        beta = mp["beta"]
    bounds = bounds or mp["bounds"]
    psi0 = random_phase_3d(n, seed=seed)
    runs = {}
    specs = [("f64", np.float64, None, psi0), ("f32", np.float32, None, psi0),
             ("f32x", np.float32, np.float64, psi0)]
    if with_rounded:
        specs.append(("f64r", np.float64, None, psi0.astype(np.complex64).astype(np.complex128)))
    for name, FT, compute, init in specs:
        g = sfd.Grid((n, n, n), bounds, FT=FT, compute=compute)
        f = sfd.Field(g)
        f.set(init)
        p = sfd.GrossPitaevskiiParameters(pot=sfd.PotentialZero(f), coeffΔ=-0.5, β=beta)
        model = (sfd.NumModelSplit2 if order == 2 else sfd.NumModelSplit1)(f, p, dt, steps, steps)
        e0 = sfd.energy_all(model)[3]
        runs[name] = dict(field=f, model=model, E0=e0, N0=sfd.norm(f))
    marks = sorted(set(int(round(steps * (k + 1) / checkpoints)) for k in range(checkpoints)))
    rows = []
    done = 0
    t0 = time.perf_counter()
    for m in marks:
        for r in runs.values():
            r["model"].steps(m - done)
        done = m
        torch.cuda.synchronize()
        row = {"step": m, "t": m * dt}
        ref = runs["f64"]["field"].ϕ
        for name, r in runs.items():
            E = sfd.energy_all(r["model"])[3]
            N = sfd.norm(r["field"])
            row[name] = {"norm_drift": abs(N / r["N0"] - 1.0), "energy_drift": abs(E / r["E0"] - 1.0)}
            if name != "f64":
                row[name]["rel_l2_vs_f64"] = _rel_l2_dev(r["field"].ϕ, ref)
        rows.append(row)
        if log:
            log(f"[precision] step {m}: " + ", ".join(
                f"{k} relL2 {v.get('rel_l2_vs_f64', 0):.3e} dN {v['norm_drift']:.1e} dE {v['energy_drift']:.1e}"
                for k, v in row.items() if isinstance(v, dict)))
    return {
        "workload": f"3D GP Split{order} {n}^3 [0,2pi)^3 V=0 beta={beta:.6g} dt={dt} random-phase seed {seed}",
        "n": n, "steps": steps, "dt": dt, "beta": beta,
        "stiffness": 0.5 * (math.pi * n / (bounds[0][1] - bounds[0][0])) ** 2 * dt,
        "rows": rows, "seconds": time.perf_counter() - t0,
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description="FP32-vs-FP64 precision study of the split step")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--dt", type=float, default=5e-4)
    ap.add_argument("--beta", type=float, default=None)
    ap.add_argument("--checkpoints", type=int, default=8)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args(argv)
    res = run_study(args.n, args.steps, args.dt, args.beta, args.checkpoints, args.seed,
                    log=lambda s: print(s, file=sys.stderr, flush=True))
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
