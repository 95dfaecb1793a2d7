#!/usr/bin/env python3
"""Benchmark: grid-cell updates/s of the Chorin projection rollout on one
MI355X at N=4096, float32.

Workload (the "plain method on GPU, large grids" configuration at its
largest size): a 4096 x 4096 periodic field on [0, 2 pi)^2, Taylor-Green
plus a seeded divergence-free perturbation (modes |k| <= 8, amplitude 0.1),
central advection, nu = 1e-4, dt from CFL 0.3, Gaussian forcing with theta
drawn from U[0.5, 1.5].  One step = one full Chorin step (predictor,
pressure projection, objective accumulation) of that field; the timed
region is one rollout of --steps steps with every input already in HBM.

Also reported (one JSON line):
  roofline      the dominant kernel's algorithmic bytes per launch over its
                average launch time, measured here with events on the launch
                stream; traffic from profiles/traffic_N4096.json if present
  cpu_baseline  the FP32 NumPy/SciPy oracle timed on this host for a bounded
                number of steps of the same workload
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--N", type=int, default=4096)
    ap.add_argument("--scheme", default="central", choices=["central", "upwind"])
    ap.add_argument("--nu", type=float, default=1e-4)
    ap.add_argument("--reps", type=int, default=50, help="launches per pass for the roofline timing")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    return ap.parse_args()


def cpu_baseline(u0, theta, N, L, nu, dt, scheme, budget_s):
    """FP32 oracle rollout on the host cores for about budget_s seconds."""
    import numpy as np
    from oracle import ns_oracle as O
    cores = len(os.sched_getaffinity(0))
    O.set_fft_workers(cores)
    try:
        ph = O.PhysParams(nu)
        x = O.grid_x(N, L, np.float32)
        u = u0.copy()
        dx = L / N
        t0 = time.perf_counter()
        steps = 0
        while True:
            f = np.zeros_like(u)
            f[0] = O.forcing_x(theta, O.step_time(steps, dt, np.float32), x, ph)[:, None]
            _ = np.sum(u * f, dtype=np.float32)
            u = O.step(u, f, dx, nu, dt, scheme)
            steps += 1
            el = time.perf_counter() - t0
            if el >= budget_s or steps >= 1000:
                break
    finally:
        O.set_fft_workers(None)
    return {"value": N * N * steps / el, "unit": "cell-updates/s", "cores": cores, "kind": "port",
            "sample": f"FP32 NumPy/SciPy oracle rollout, N={N}, {steps} steps "
                      f"({el:.1f} s), scipy.fft workers={cores}"}


def pass_times(solver, reps):
    """Average launch time (ms) of the column pass and of the fused row pass,
    with torch events on the current stream (the stream the kernels run on)."""
    import torch
    from ns_solver_jax_differential_distributed_amd import _lib
    lib = _lib.lib()
    p = solver._plan()
    s0, s1 = solver._ws["spec0"], solver._ws["spec1"]
    slots = _lib.make_slots(solver.phys.nu, solver.time.dt, 1.0, 0.5, 0, 1 << 30)
    st = _lib.stream_ptr()
    scheme = _lib.SCHEME[solver.scheme]
    out = {}
    for name in ("column", "row"):
        def launch():
            if name == "column":
                _lib.check(lib.nsk_column_pass(ctypes.byref(p), s0.data_ptr(), 0, st), "column")
            else:
                _lib.check(lib.nsk_row_pass(ctypes.byref(p), s0.data_ptr(), s1.data_ptr(), slots.data_ptr(),
                                            0, scheme, 0, None, 0, st), "row")
        for _ in range(5):
            launch()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            launch()
        e1.record()
        torch.cuda.synchronize()
        out[name] = e0.elapsed_time(e1) / reps
    return out


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}, WORLD_SIZE {world}); "
              "multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from ns_solver_jax_differential_distributed_amd import NS_solver as S
    from ns_solver_jax_differential_distributed_amd import workloads as W

    torch.cuda.set_device(0)
    N, L, nu = args.N, 2 * math.pi, args.nu
    u0 = W.initial_velocity(N, seed=0, kmax=8)
    dt = W.cfl_dt(u0, L, 0.3, nu)
    theta = W.theta_batch(1, seed=0)[0]
    solver = S.SingleCoreNSSolver(S.make_grid(N, L), S.Phys(nu, 0.5, 0.1, 1.0), S.Time(dt, args.steps),
                                  advection_scheme=args.scheme, use_graph=not args.no_graph)
    u0_dev = torch.from_numpy(u0).cuda()
    if args.warmup > 0:
        solver.rollout(u0_dev, theta, steps=args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J, uf = solver.rollout(u0_dev, theta, steps=args.steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    value = N * N * args.steps / el
    finite = bool(torch.isfinite(uf).all().item()) and math.isfinite(float(J))

    pt = pass_times(solver, args.reps)
    alg = 16.0 * N * N          # bytes per launch: read + write of one spectrum buffer
    dom = max(pt, key=pt.get)
    achieved = alg / (pt[dom] * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"traffic_N{N}.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get(f"{dom}_pass_bytes")
    roofline = {"bound": "hbm", "kernel": f"{dom} pass (k_{'col' if dom == 'column' else 'row'})",
                "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "alg_bytes_per_launch": alg,
                "pass_us": {k: round(v * 1e3, 2) for k, v in pt.items()}}
    step_us = el / args.steps * 1e6
    line = {
        "metric": "grid-cell updates/s at N=4096 fp32",
        "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": step_us / 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Taylor-Green + divergence-free perturbation)",
        "config": {"workload": f"Chorin rollout, single field N={N} on [0,2pi)^2, {args.scheme} advection, "
                               f"nu={nu:g}, dt={dt:.4g} (CFL 0.3), Gaussian forcing",
                   "N": N, "scheme": args.scheme, "nu": nu, "dt": dt, "parallelism": "single GPU"},
        "roofline": roofline,
        "step_hbm_frac_40B": round(40.0 * N * N / (step_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
        "finite": finite,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(u0, theta, N, L, nu, dt, args.scheme, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
