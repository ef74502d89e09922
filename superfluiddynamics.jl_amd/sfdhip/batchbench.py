"""Batched imaginary-time ground states: 96 uneven 2D instances (32 at 128²,
48 at 256², 16 at 512²) on [-10, 10)², V = ½(x² + γ²y²), γ ~ U[1,3],
β ~ U[100,2000], seeded Gaussian guesses, advanced together by
``sfd_batch_imag_solve`` until ||ψ_{n+1} - ψ_n|| < tol or max_steps; FP32
batches finish each instance in FP64 from its FP32 rounding floor unless
``--no-escalate``.

Reports the wall time of the solve, site-updates/s summed over instances,
the steps each instance took, and the energy/μ of a few instances.

Usage: ``python -m sfdhip.batchbench --prec f32 --steps 20000 --dt 5e-3``.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def run(prec="f32", max_steps=20000, dt=5e-3, tol=1e-9, check_every=100, use_graph=True, seed=0,
        counts=((128, 32), (256, 48), (512, 16)), autotune=True, escalate=True):
    import torch
    from . import api as sfd
    from .workloads import batch_ground_states

    FT, compute = {"f32": (np.float32, None), "f32x": (np.float32, np.float64), "f64": (np.float64, None)}[prec]
    probs = batch_ground_states(seed=seed, counts=counts)
    fields, params = [], []
    for pr in probs:
        g = sfd.Grid((pr["n"], pr["n"]), pr["bounds"], FT=FT, compute=compute)
        f = sfd.Field(g)
        f.set(pr["psi0"])
        params.append(sfd.GrossPitaevskiiParameters(pot=sfd.PotentialArray(f, pr["V"]), coeffΔ=-0.5, β=pr["beta"]))
        fields.append(f)
    solver = sfd.BatchImagSolver(fields, params, dt)
    tune = solver.autotune() if autotune else None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    esc = escalate and prec != "f64"
    steps, dpsi = solver.run(max_steps=max_steps, tol=tol, check_every=check_every, use_graph=use_graph,
                             escalate=esc)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    site_steps = float(sum(int(s) * pr["n"] ** 2 for s, pr in zip(steps, probs)))
    sample = []
    for k in (0, len(probs) // 2, len(probs) - 1):
        m = sfd.NumModelSplitImag(fields[k], params[k], dt, 1, 1)
        E = sfd.energy_all(m)
        sample.append({"instance": k, "n": probs[k]["n"], "gamma": probs[k]["gamma"], "beta": probs[k]["beta"],
                       "E": E[3], "mu": E[4], "steps": int(steps[k]), "dpsi": float(dpsi[k])})
    return {
        "workload": "96 2D imaginary-time ground states (32x128^2, 48x256^2, 16x512^2), [-10,10)^2",
        "precision": prec, "dt": dt, "tol": tol, "max_steps": max_steps, "check_every": check_every,
        "graph": use_graph, "escalate_to_fp64": esc, "finished_in_fp64": int(solver.levels.sum()),
        "wall_s": wall, "site_updates_per_s": site_steps / wall,
        "converged": int(np.sum(steps < max_steps)), "steps_min": int(steps.min()), "steps_max": int(steps.max()),
        "dpsi_median": float(np.median(dpsi)), "sample": sample, "autotune_ms": tune,
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description="batched imaginary-time ground states")
    ap.add_argument("--prec", default="f32", choices=["f32", "f32x", "f64"])
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--dt", type=float, default=5e-3)
    ap.add_argument("--tol", type=float, default=1e-9)
    ap.add_argument("--check-every", type=int, default=100)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-escalate", action="store_true", help="FP32 batches stay in FP32 to the end")
    args = ap.parse_args(argv)
    res = run(args.prec, args.steps, args.dt, args.tol, args.check_every, not args.no_graph,
              autotune=not args.no_autotune, escalate=not args.no_escalate)
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
