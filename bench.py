#!/usr/bin/env python3
"""Benchmark of the Gross-Pitaevskii Strang split step (NumModelSplit2) at 256³.

Workload: periodic box [0, 2π)³, 256³ sites, V = 0, coeffΔ = -0.5,
ξ = 2Δ, β = 1/ξ², Δt = 5e-4; ψ0 = e^{iθ} with θ a seeded Gaussian random
field on 1 <= |k| <= 8 (modal variance ∝ |k|^-2, rms 1).  A step is one full
Strang step (x, y, z round trips over Δt/2, the nonlinear phase over Δt,
x, y, z again) of the whole field, which stays resident in HBM; with Ω = 0
it runs as x, y, [z N z], x, y.  FP32 is the headline ``value``; FP32
storage with FP64 arithmetic and FP64 are measured side by side.

Prints one JSON line.  ``python bench.py --gpus N --steps K --warmup W``.
Drivers for profiling runs, each printing its own line: ``--traffic-driver``
(the launches of the rocprofv3 --pmc runs), ``--solve-driver PREC`` (the
solve! loop), ``--adjoint-driver PREC`` (Split2 steps and sfd_adjoint on
the 256² control problem).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "superfluiddynamics.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def bytes_per_site_update(prec: str, ndim: int = 3, with_v: bool = False, passes=None) -> int:
    """Bytes per site-update of a Strang step made of `passes` one-axis
    sweeps, each one read and one write of the field, the nonlinear phase
    fused (V read once when present).  2D passes is the reference's
    structure; the step here makes 2D - 1 (the innermost axis's two round
    trips share one sweep around the nonlinear phase)."""
    sc = 8 if prec == "f32" else 16
    sr = 4 if prec == "f32" else 8
    if passes is None:
        passes = 2 * ndim
    return passes * 2 * sc + (sr if with_v else 0)


def passes_per_step(ndim: int, steps: int, chained: bool = True) -> float:
    """Sweeps per Strang step of one call of `steps` steps (Omega = 0, real
    time): x, y, [z N z], x, y per step (2D - 1) when each step runs alone;
    y(τ), z(τ), x(τ) + N, y(Δt), z(Δt), x(Δt) + N, ..., y(τ), z(τ), x(τ) when
    the call chains them (D per step plus the D closing sweeps)."""
    if not chained or steps < 2:
        return float(2 * ndim - 1)
    return ndim * (steps + 1) / steps


def load_traffic(slot: str, n: int = 256):
    """HBM bytes per launch of the kernel of sweep slot `slot` on the n³ FP32
    workload from the committed rocprofv3 --pmc summary (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration recorded with it); None
    if there is no summary for n or it holds no figure for that kernel."""
    path = os.path.join(ROOT, "profiles", f"traffic_{n}cube_f32.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    k = d.get("kernels", {}).get(slot)
    return k.get("traffic_bytes_per_launch") if k else None


PRECISIONS = {"f32": ("float32", None), "f32x": ("float32", "float64"), "f64": ("float64", None)}


SWEEP_SLOTS = ("x round trip", "y round trip", "z round trip", "double round trip [d N d]", "other sweeps",
               "double round trip [x x]")


def gpu_leg(sfd, torch, psi0, mp, prec, steps, warmup, reps, launches, settle_ms=200.0, autotune=True):
    """One precision: autotune, warm up, then `reps` timed runs of `steps`
    Split2 steps (HIP events on the context's stream, which is torch's
    current stream); ms/step is their median.  A further run of `steps`
    steps with the library's sweep timer gives the average duration of each
    kernel of the step, measured inside the step sequence itself."""
    import ctypes
    import numpy as np
    FT, compute = PRECISIONS[prec]
    n = psi0.shape[0]
    g = sfd.Grid((n, n, n), mp["bounds"], FT=np.dtype(FT), compute=compute)
    f = sfd.Field(g)
    tune = f.autotune() if autotune else None
    variants = [f.ctx().variant(a)[2] for a in range(3)]
    f.set(psi0)
    p = sfd.GrossPitaevskiiParameters(pot=sfd.PotentialZero(f), coeffΔ=mp["coeffD"], β=mp["beta"])
    model = sfd.NumModelSplit2(f, p, mp["dt"], steps, steps)
    model.steps(max(warmup, 1))
    torch.cuda.synchronize()
    # the precision comparison takes every leg's state after the same steps
    state = f.to_numpy()
    # settle: further untimed steps until settle_ms of GPU work have run, so
    # that the clocks, lowered while the host prepared the field, are back
    # up before the timed runs
    settle_steps = 0
    t_settle = time.perf_counter()
    while (time.perf_counter() - t_settle) * 1e3 < settle_ms:
        model.steps(10)
        torch.cuda.synchronize()
        settle_steps += 10
    runs = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        model.steps(steps)
        e1.record()
        torch.cuda.synchronize()
        runs.append(e0.elapsed_time(e1) / steps)
    ms_step = float(np.median(runs))
    nrm = sfd.norm(f)
    solve = solve_loop(sfd, torch, model, min(steps, 50))
    lib, ctx = sfd.load(), f.ctx()
    sfd._lib.check(lib.sfd_ctx_time_sweeps(ctx.h, 1))
    sfd._lib.check(lib.sfd_ctx_sweep_times(ctx.h, None, None, 1))
    model.steps(steps)
    tot = (ctypes.c_double * len(SWEEP_SLOTS))()
    cnt = (ctypes.c_int64 * len(SWEEP_SLOTS))()
    sfd._lib.check(lib.sfd_ctx_sweep_times(ctx.h, tot, cnt, 1))
    sfd._lib.check(lib.sfd_ctx_time_sweeps(ctx.h, 0))
    sweeps = {}
    for k, name in enumerate(SWEEP_SLOTS):
        if cnt[k]:
            sweeps[name] = {"launches_per_step": cnt[k] / steps, "avg_ms": tot[k] / cnt[k],
                            "ms_per_step": tot[k] / steps}
    # The one-axis round trips of the chained step that carry no nonlinear
    # phase (y, z: the same kernels, variants and flags as in the step)
    # launched back to back on the context's stream: HIP events around
    # `launches` launches, median of 3 batches.  The in-step events above add
    # a few microseconds to every launch; these do not.
    back_to_back = {}
    psi = ctypes.c_void_p(f.ϕ.data_ptr())
    for ax, name in ((1, "y round trip"), (2, "z round trip")):
        reps = []
        for _ in range(3):
            k0 = torch.cuda.Event(enable_timing=True)
            k1 = torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            k0.record()
            for _ in range(launches):
                lib.sfd_lap_rot_axis(ctx.h, psi, ax, mp["coeffD"], 0.0, mp["dt"] * 0.5, 0)
            k1.record()
            torch.cuda.synchronize()
            reps.append(k0.elapsed_time(k1) / launches)
        back_to_back[name] = sorted(reps)[1]
    # A plain elementwise pass over the same field, timed the same way: torch's
    # in-place mul_ (one read and one write of every byte), the rate a kernel
    # that only streams the field reaches on this GPU at this size.
    flat = torch.view_as_real(f.ϕ)
    reps = []
    for _ in range(3):
        k0 = torch.cuda.Event(enable_timing=True)
        k1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        k0.record()
        for _ in range(launches):
            flat.mul_(1.0)
        k1.record()
        torch.cuda.synchronize()
        reps.append(k0.elapsed_time(k1) / launches)
    copy_ms = sorted(reps)[1]
    del model, f, flat
    torch.cuda.empty_cache()
    return dict(ms=ms_step, runs_ms=runs, state=state, norm=nrm, variants=variants, tune=tune, solve=solve,
                sweeps=sweeps, back_to_back=back_to_back, copy_ms=copy_ms, settle_steps=settle_steps)


def solve_loop(sfd, torch, model, k):
    """Wall time of one solve! iteration (energy, then a Split2 step): the
    energy as its own call before each step (sfd_energy + sfd_split_step),
    and fused into the steps' first-half sweeps (sfd_split_step_energy)."""
    sfd.steps_with_energy(model, 2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(k):
        sfd.energy_all(model)
        model.steps(1)
    torch.cuda.synchronize()
    sep = (time.perf_counter() - t0) / k
    t0 = time.perf_counter()
    sfd.steps_with_energy(model, k)
    fused = (time.perf_counter() - t0) / k
    return {"iterations": k, "separate_ms_per_iter": sep * 1e3, "fused_ms_per_iter": fused * 1e3}


def kernel_name(variants, slot):
    """Kernel of sweep slot `slot` (0-2: the round trip of axis x, y, z; 3: the
    double round trip of the innermost axis; 5: the double round trip of x)
    under the FP32 variants (n1, n2, ts, g, pipelined, waves, halves) per
    axis."""
    ax = {3: 2, 5: 0}.get(slot, slot)
    n1, n2, ts, g, pipe = variants[ax][:5]
    if slot in (3, 5):
        return f"k_sweep<float, float, {n1}, {n2}, {ts}, {g}, 7>"
    return f"{'k_sweep_pipe' if pipe else 'k_sweep'}<float, float, {n1}, {n2}, {ts}, {g}>"


def dominant_kernel(variants, sweeps):
    """(kernel, slots): the sweep kernel with the largest share of the step's
    time, summed over the timer slots it runs in (`sweeps`: slot name ->
    {"ms_per_step", ...}), and those slots in slot order."""
    share, slots = {}, {}
    for k, name in enumerate(SWEEP_SLOTS):
        if name == "other sweeps" or name not in sweeps:
            continue
        kn = kernel_name(variants, k)
        share[kn] = share.get(kn, 0.0) + sweeps[name]["ms_per_step"]
        slots.setdefault(kn, []).append(name)
    kn = max(share, key=share.get)
    return kn, slots[kn]


def cpu_leg(psi0, mp, cpu_steps):
    """The FP64 CPU oracle (C + OpenMP restatement of NumModelSplit2) on the
    same 256³ workload: one untimed step, then cpu_steps timed steps
    (the protocol of benchmarks/scalability_gp.jl:36-51)."""
    import numpy as np
    from oracle import oracle
    psi = np.ascontiguousarray(psi0.copy())
    kw = dict(coeffD=mp["coeffD"], beta=mp["beta"], dt=mp["dt"], order=2)
    oracle.split_step(psi, mp["bounds"], nsteps=1, **kw)
    t0 = time.perf_counter()
    oracle.split_step(psi, mp["bounds"], nsteps=cpu_steps, **kw)
    t = (time.perf_counter() - t0) / cpu_steps
    return t, oracle.num_threads()


def solve_driver(n=256, iters=20, prec="f32"):
    """Launches for a rocprofv3 kernel trace of solve!'s loop on the
    workload: `iters` iterations with the energy as its own call
    (sfd_energy + sfd_split_step), then `iters` fused iterations
    (sfd_split_step_energy)."""
    import numpy as np
    import torch
    import sfdhip as sfd
    from sfdhip.workloads import metric_params, random_phase_3d
    mp = metric_params(n)
    FT, compute = PRECISIONS[prec]
    g = sfd.Grid((n, n, n), mp["bounds"], FT=np.dtype(FT), compute=compute)
    f = sfd.Field(g)
    f.autotune()
    f.set(random_phase_3d(n, seed=0))
    p = sfd.GrossPitaevskiiParameters(pot=sfd.PotentialZero(f), coeffΔ=mp["coeffD"], β=mp["beta"])
    model = sfd.NumModelSplit2(f, p, mp["dt"], iters, iters)
    r = solve_loop(sfd, torch, model, iters)
    print(json.dumps({"solve_driver": prec, "n": n, **r}), flush=True)
    return 0


def traffic_driver(n=256, launches=5):
    """Launches for the rocprofv3 --pmc runs: `launches` round trips of each
    axis of the 256³ FP32 workload under the autotuned variants (x: the
    contiguous kernel, then y and z: the strided kernel), one call of
    `launches` chained Split2 steps (y, z, x + N), `launches` one-step calls
    (x, y, [z N z], x, y: the innermost double round trip), one call of
    `launches` chained imaginary-time steps (whose [x x] is the x double
    round trip), then `launches` in-place passes of a plain elementwise
    kernel over the same 128 MiB (torch mul_), the calibration whose
    algorithmic traffic is exactly one read and one write of the field."""
    import ctypes
    import numpy as np
    import torch
    import sfdhip as sfd
    from sfdhip.workloads import metric_params, random_phase_3d
    mp = metric_params(n)
    g = sfd.Grid((n, n, n), mp["bounds"], FT=np.float32)
    f = sfd.Field(g)
    f.autotune()  # the variants the bench's FP32 leg runs
    f.set(random_phase_3d(n, seed=0))
    lib, ctx = sfd.load(), f.ctx()
    psi = ctypes.c_void_p(f.ϕ.data_ptr())
    for ax in range(3):
        for _ in range(launches):
            rc = lib.sfd_lap_rot_axis(ctx.h, psi, ax, mp["coeffD"], 0.0, mp["dt"] * 0.5, 0)
            assert rc == 0, lib.sfd_last_error()
        torch.cuda.synchronize()
        print(f"[traffic] axis {ax}: {launches} launches, variant {ctx.variant(ax)[2]}", flush=True)
    rc = lib.sfd_split_step(ctx.h, psi, None, mp["coeffD"], mp["beta"], 0.0, mp["dt"], launches, 2, 0)
    assert rc == 0, lib.sfd_last_error()
    torch.cuda.synchronize()
    print(f"[traffic] {launches} chained Split2 steps", flush=True)
    for _ in range(launches):
        rc = lib.sfd_split_step(ctx.h, psi, None, mp["coeffD"], mp["beta"], 0.0, mp["dt"], 1, 2, 0)
        assert rc == 0, lib.sfd_last_error()
    torch.cuda.synchronize()
    print(f"[traffic] {launches} one-step Split2 calls", flush=True)
    rc = lib.sfd_split_step(ctx.h, psi, None, mp["coeffD"], mp["beta"], 0.0, mp["dt"], launches, 2, 1)
    assert rc == 0, lib.sfd_last_error()
    torch.cuda.synchronize()
    print(f"[traffic] {launches} chained imaginary-time Split2 steps", flush=True)
    flat = torch.view_as_real(f.ϕ)
    for _ in range(launches):
        flat.mul_(1.0)
    torch.cuda.synchronize()
    print(f"[traffic] calibration: {launches} in-place passes over {flat.numel() * 4} bytes", flush=True)
    return 0


def adjoint_driver(prec="f64", nsteps=800, reps=3):
    """The 256² control problem (V = ½|x|² + Σ_j a_j exp(-|x - c_j|²/0.5) on an
    8×8 lattice, β = 500, Δt = 1e-3, the Thomas-Fermi state as ψ0 and its
    shift by +1.5 in x as the target): wall time of `nsteps` Split2 steps,
    and of sfd_adjoint over the same steps (loss, dL/dV, dL/dβ, √N
    checkpoints), best of `reps` each."""
    import numpy as np
    import torch
    import sfdhip as sfd
    from sfdhip.workloads import control_problem, shift_x, thomas_fermi_2d
    basis, a, V0, bounds = control_problem()
    n = basis.shape[-1]
    V = np.ascontiguousarray(V0 + np.tensordot(a, basis, axes=1))
    psi0, _, _ = thomas_fermi_2d(n=n, L=8.0, beta=500.0, seed=0)
    FT, compute = PRECISIONS[prec]
    g = sfd.Grid((n, n), bounds, FT=np.dtype(FT), compute=compute)
    f = sfd.Field(g)
    f.autotune()
    tgt = sfd.Field(g)
    tgt.set(np.ascontiguousarray(shift_x(psi0, 1.5, 8.0)))
    p = sfd.GrossPitaevskiiParameters(pot=sfd.PotentialArray(f, V), coeffΔ=-0.5, β=500.0)
    model = sfd.NumModelSplit2(f, p, 1e-3, nsteps, nsteps)
    fwd, adj = [], []
    for _ in range(reps):
        f.set(psi0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        model.steps(nsteps)
        torch.cuda.synchronize()
        fwd.append(time.perf_counter() - t0)
        f.set(psi0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        loss, gV, gb = sfd.adjoint_gradient(model, tgt, nsteps)
        torch.cuda.synchronize()
        adj.append(time.perf_counter() - t0)
    r = {"adjoint_driver": prec, "n": n, "nsteps": nsteps, "forward_ms": 1e3 * min(fwd),
         "adjoint_ms": 1e3 * min(adj), "adjoint_over_forward": min(adj) / min(fwd), "loss": loss,
         "dL_dbeta": gb, "dL_dV_l2": float(torch.linalg.vector_norm(gV))}
    print(json.dumps(r), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--reps", type=int, default=5, help="timed runs of --steps steps; ms/step is their median")
    ap.add_argument("--settle-ms", type=float, default=200.0,
                    help="untimed steps after the warm-up until this much GPU time has passed (clocks settle)")
    ap.add_argument("--kernel-launches", type=int, default=40,
                    help="back-to-back launches timed per batch for the dominant kernel's duration")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the 1000-step FP32/F32X-vs-FP64 accuracy run at the stable dt")
    ap.add_argument("--traffic-driver", action="store_true",
                    help="only launch the kernels profiled by the rocprofv3 --pmc traffic runs")
    ap.add_argument("--solve-driver", choices=["f32", "f32x", "f64"],
                    help="only run solve!'s loop (separate and fused energies) for a kernel trace")
    ap.add_argument("--adjoint-driver", choices=["f32", "f32x", "f64"],
                    help="only time Split2 steps and sfd_adjoint on the 256² control problem")
    args = ap.parse_args(argv)
    if args.adjoint_driver:
        return adjoint_driver(prec=args.adjoint_driver)
    if args.traffic_driver:
        return traffic_driver(args.n)
    if args.solve_driver:
        return solve_driver(args.n, prec=args.solve_driver)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the whole {args.n}³ domain is resident on a single MI355X; "
              f"--gpus {args.gpus} is not supported (the slab-decomposed multi-GPU path is not built).",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch
    import sfdhip as sfd
    from sfdhip.workloads import metric_params, random_phase_3d

    n = args.n
    mp = metric_params(n)
    t0 = time.perf_counter()
    psi0 = random_phase_3d(n, seed=0)
    print(f"[bench] initial state {n}³ generated in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)

    res = {}
    states = {}
    for prec in ("f32", "f32x", "f64"):
        r = gpu_leg(sfd, torch, psi0, mp, prec, args.steps, args.warmup, args.reps, args.kernel_launches,
                    args.settle_ms, not args.no_autotune)
        states[prec] = r.pop("state")
        res[prec] = r
        solve = r["solve"]
        print(f"[bench] {prec}: solve! iteration {solve['separate_ms_per_iter']:.4f} ms with a separate energy, "
              f"{solve['fused_ms_per_iter']:.4f} ms fused", file=sys.stderr, flush=True)
        print(f"[bench] {prec}: variants (n1,n2,ts,g,pipe,waves,halves) per axis {r['variants']}; "
              f"autotune ms {r['tune']}", file=sys.stderr, flush=True)
        print(f"[bench] {prec}: {r['ms']:.4f} ms/step (median of {args.reps} runs of {args.steps} steps: "
              f"{['%.4f' % v for v in r['runs_ms']]}); kernels in the step: "
              + ", ".join(f"{k} {v['avg_ms'] * 1e3:.1f} us x {v['launches_per_step']:.0f}"
                          for k, v in r["sweeps"].items()), file=sys.stderr, flush=True)

    def rel(a, b):
        return float(np.sqrt(np.sum(np.abs(a.astype(np.complex128) - b) ** 2) / np.sum(np.abs(b) ** 2)))
    err32 = rel(states["f32"], states["f64"])
    err32x = rel(states["f32x"], states["f64"])
    del states

    sites = n ** 3
    out = {}
    for prec in ("f32", "f32x", "f64"):
        r = res[prec]
        sups = sites / (r["ms"] * 1e-3)
        p_ = "f64" if prec == "f64" else "f32"
        b = bytes_per_site_update(p_, passes=passes_per_step(3, args.steps))
        b6 = bytes_per_site_update(p_)
        out[prec] = dict(site_updates_per_s=sups, ms_per_step=r["ms"], runs_ms=r["runs_ms"],
                         bytes_per_site_update=b, step_hbm_frac=sups * b / HBM_PEAK,
                         step_hbm_frac_6_pass_bytes=sups * b6 / HBM_PEAK, kernels=r["sweeps"],
                         variants=r["variants"])

    # Dominant kernel: the sweep kernel with the largest share of the FP32
    # step's time (in-step timer, the slots of one kernel summed: the y and z
    # round trips of a chained step are one kernel when they share a
    # variant), one read and one write of the complex FP32 field per launch
    # (V = 0: no potential bytes).  Its average duration is the back-to-back
    # figure of its slots where there is one (y, z), else the in-step one,
    # weighted by launches per step.
    k32 = res["f32"]["sweeps"]
    dom_kernel, dom_slots = dominant_kernel(res["f32"]["variants"], k32)
    b2b = res["f32"]["back_to_back"]
    nl_ = sum(k32[sl]["launches_per_step"] for sl in dom_slots)
    dom_ms = sum(b2b.get(sl, k32[sl]["avg_ms"]) * k32[sl]["launches_per_step"] for sl in dom_slots) / nl_
    dom_ms_in_step = sum(k32[sl]["ms_per_step"] for sl in dom_slots) / nl_
    dom_share = sum(k32[sl]["ms_per_step"] for sl in dom_slots)
    launch_bytes = 2 * 8 * sites
    achieved = launch_bytes / (dom_ms * 1e-3)
    traffic = next((t for t in (load_traffic(sl, n) for sl in dom_slots) if t is not None), None)
    copy_ms = res["f32"]["copy_ms"]

    # Accuracy at the largest stable step: at dt = 5e-4 the unfiltered split is
    # past its stability limit (D |coeffD| kmax^2 dt = 12.3 > pi) and any two
    # precisions, or FP64 runs one ulp apart, decorrelate within ~100 steps.
    accuracy = None
    if not args.no_accuracy:
        from sfdhip.precision import run_study
        st = run_study(n=n, steps=1000, dt=1e-4, checkpoints=1, with_rounded=False)
        row = st["rows"][-1]
        accuracy = {"dt": 1e-4, "steps": 1000, "vs": "FP64 GPU run, same ψ0",
                    "f32_rel_l2": row["f32"]["rel_l2_vs_f64"], "f32x_rel_l2": row["f32x"]["rel_l2_vs_f64"],
                    "f32_norm_drift": row["f32"]["norm_drift"], "f32_energy_drift": row["f32"]["energy_drift"],
                    "f64_norm_drift": row["f64"]["norm_drift"], "f64_energy_drift": row["f64"]["energy_drift"]}
        print(f"[bench] accuracy dt=1e-4 1000 steps: f32 relL2 {accuracy['f32_rel_l2']:.3e}, "
              f"f32x relL2 {accuracy['f32x_rel_l2']:.3e}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        t_cpu, cores = cpu_leg(psi0, mp, args.cpu_steps)
        cpu = {"value": sites / t_cpu, "unit": "site-updates/s", "cores": cores, "kind": "port",
               "sample": f"FP64 CPU oracle (C+OpenMP restatement of NumModelSplit2), {n}³ metric workload, "
                         f"1 untimed + {args.cpu_steps} timed steps, {t_cpu:.3f} s/step"}

    line = {
        "metric": "GP split-step site-updates/s at 256³ (FP32 vs FP64) + % HBM roofline",
        "value": out["f32"]["site_updates_per_s"],
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": out["f32"]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"3D GP Strang split (NumModelSplit2), {n}³ periodic [0,2π)³, V=0, "
                               f"β=1/ξ² (ξ=2Δ), Δt=5e-4, random-phase ψ0 seed 0",
                   "grid": [n, n, n], "parallelism": "single GPU, whole domain resident"},
        "fp64": {"value": out["f64"]["site_updates_per_s"], "ms_per_step": out["f64"]["ms_per_step"],
                 "runs_ms": out["f64"]["runs_ms"], "hbm_frac": out["f64"]["step_hbm_frac"],
                 "kernels": out["f64"]["kernels"], "variants": out["f64"]["variants"]},
        "f32x": {"value": out["f32x"]["site_updates_per_s"], "ms_per_step": out["f32x"]["ms_per_step"],
                 "runs_ms": out["f32x"]["runs_ms"], "hbm_frac": out["f32x"]["step_hbm_frac"],
                 "kernels": out["f32x"]["kernels"],
                 "variants": out["f32x"]["variants"],
                 "note": "complex FP32 field in HBM, FP64 arithmetic and tables in registers/LDS"},
        "f32_over_f64_speedup": out["f64"]["ms_per_step"] / out["f32"]["ms_per_step"],
        "f32x_over_f64_speedup": out["f64"]["ms_per_step"] / out["f32x"]["ms_per_step"],
        "f32_step_hbm_frac": out["f32"]["step_hbm_frac"],
        "f32_step_hbm_frac_note": f"a call of {args.steps} chained steps (y, z, x + N per step: the half-step "
                                  "flows between two nonlinear phases merged into one round trip per axis) "
                                  f"moves {passes_per_step(3, args.steps):.3f} passes x 16 B = "
                                  f"{out['f32']['bytes_per_site_update']:.2f} B per site-update; priced at the "
                                  "6-pass 96 B of the reference's structure it would read "
                                  f"{out['f32']['step_hbm_frac_6_pass_bytes']:.3f}",
        "f32_kernels": out["f32"]["kernels"],
        "f32_back_to_back_ms": res["f32"]["back_to_back"],
        "f32_runs_ms": out["f32"]["runs_ms"],
        "settle": {"ms": args.settle_ms, "f32_steps": res["f32"]["settle_steps"],
                   "note": "untimed steps after the warm-up steps, until the clocks lowered during host-side "
                           "set-up are back up; then --reps timed runs of exactly --steps steps each"},
        "f32_variants": out["f32"]["variants"],
        "f32_vs_f64_rel_l2": err32,
        "f32x_vs_f64_rel_l2": err32x,
        "f32_vs_f64_steps": max(args.warmup, 1),
        "f32_vs_f64_note": "at the metric dt the unfiltered split is past its stability limit "
                           "(D |coeffD| kmax^2 dt = 12.3 > pi): any two precisions, or FP64 runs one rounding "
                           "apart, decorrelate to ~sqrt(2); accuracy_stable_dt holds the comparison at dt = 1e-4",
        "accuracy_stable_dt": accuracy,
        "solve_loop": {"note": "wall time of one solve! iteration (energy of the state, then one Split2 "
                               "step) at the workload size: separate = sfd_energy + sfd_split_step per "
                               "iteration; fused = sfd_split_step_energy (energy sums taken in the step's "
                               "first-half sweeps)",
                       **{p_: res[p_]["solve"] for p_ in ("f32", "f32x", "f64")}},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "kernel": f"{dom_kernel}: {' and '.join(dom_slots)}s, FP32, {n}³ (the step's largest "
                               f"share of time, {dom_share * 1e3:.1f} us of {out['f32']['ms_per_step'] * 1e3:.1f} "
                               f"us per step)",
                     "avg_launch_ms": dom_ms, "avg_launch_ms_in_step": dom_ms_in_step,
                     "bytes_per_launch": launch_bytes,
                     "timing": "avg_launch_ms: HIP events around back-to-back launches on the context's stream; "
                               "avg_launch_ms_in_step: events around each launch inside the timed steps "
                               "(adds a few us per launch)",
                     "copy_reference": {
                         "kernel": "torch in-place mul_ over the same complex FP32 field (one read + one write)",
                         "avg_launch_ms": copy_ms, "GB/s": launch_bytes / (copy_ms * 1e-3) / 1e9,
                         "frac": launch_bytes / (copy_ms * 1e-3) / HBM_PEAK,
                         "dominant_over_copy": copy_ms / dom_ms}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
