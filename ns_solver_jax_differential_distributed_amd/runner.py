"""CLI entry point (the reference's ``runner.py``).

Same arguments (``--config``, ``--profile``, ``--profile-warmup``,
``--profile-runs``; runner.py:31-56) and the same single-device behaviour
(runner.py:104-113, 119-170): load the config, build the single-core
solver, optionally profile its stages, roll out from u = 0 with
theta0, and write the speed field sqrt(ux^2 + uy^2) as
``result/velocity_magnitude_single_N{N}.npy`` and a ``.png`` of it
(2nd/98th-percentile colour range, x along the horizontal axis).  The
reference's distributed comparison needs several GPUs; like the reference
on one device it is skipped with a log line.

    python -m ns_solver_jax_differential_distributed_amd.runner --config sample_config/config_N_512.json
"""
from __future__ import annotations

import argparse
import os
import time
from typing import Optional, Sequence

import numpy as np
import torch

from .NS_solver import Phys, SingleCoreNSSolver, Time, load_config, make_grid, setup_logger

logger = setup_logger(__name__)


def build_parser() -> argparse.ArgumentParser:
    """The CLI parser (runner.py:31-56)."""
    parser = argparse.ArgumentParser(description="Navier-Stokes CFD Solver")
    parser.add_argument("--config", type=str, default=None,
                        help="Path to a JSON config file. Missing fields fall back to defaults.")
    parser.add_argument("--profile", action="store_true",
                        help="Run stage-wise profiling of the solver.")
    parser.add_argument("--profile-warmup", type=int, default=1,
                        help="Number of warmup runs per profiled kernel (excluded from statistics).")
    parser.add_argument("--profile-runs", type=int, default=5,
                        help="Number of measured runs per profiled kernel.")
    parser.add_argument("--out", type=str, default="result",
                        help="Output folder (not in the reference, which always writes to result/).")
    return parser


def output_paths(out: str, N: int):
    """(npy, png) paths of the single-device speed field (runner.py:162-164)."""
    stem = os.path.join(out, f"velocity_magnitude_single_N{N}")
    return stem + ".npy", stem + ".png"


def _save_png(speed: np.ndarray, path: str, L: float, title: str) -> bool:
    try:
        import matplotlib
        matplotlib.use("Agg")
        import matplotlib.pyplot as plt
    except ImportError:
        logger.warning("matplotlib not available: no .png written")
        return False
    vmin, vmax = float(np.percentile(speed, 2)), float(np.percentile(speed, 98))
    if not np.isfinite(vmin) or not np.isfinite(vmax) or vmax <= vmin:
        c = float(np.mean(speed))
        vmin, vmax = c - 1e-12, c + 1e-12
    fig, ax = plt.subplots(1, 1, figsize=(7, 6))
    im = ax.imshow(speed.T, origin="lower", extent=[0, L, 0, L], cmap="viridis", vmin=vmin, vmax=vmax)
    ax.set_title(title)
    ax.set_xlabel("x")
    ax.set_ylabel("y")
    fig.colorbar(im, ax=ax, label="Speed")
    fig.suptitle("Velocity Magnitude — Single-core", fontsize=13)
    fig.tight_layout()
    fig.savefig(path, dpi=300)
    plt.close(fig)
    return True


def main(argv: Optional[Sequence[str]] = None) -> dict:
    args = build_parser().parse_args(argv)
    logger.info("=" * 80)
    logger.info("NAVIER-STOKES CFD SOLVER (MI355X)")
    logger.info("=" * 80)
    config = load_config(args.config)
    grid = make_grid(config.N, L=config.L)
    phys = Phys(nu=config.nu, mu0=config.mu0, sigma0=config.sigma0, C=config.C_force)
    solver = SingleCoreNSSolver(grid, phys, Time(dt=config.dt, steps=config.steps),
                                advection_scheme=config.advection_scheme)
    theta = config.theta0
    logger.info(f"Advection scheme: {config.advection_scheme}")
    logger.info("Single GPU: distributed comparison disabled; running the single-core solver only.")
    if args.profile:
        from .solver_profiler import profile_solver_kernels
        profile_solver_kernels(solver, None, theta=theta, warmup=args.profile_warmup,
                               runs=args.profile_runs, logger=logger)
    u0 = torch.zeros((2, config.N, config.N), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J, u = solver.rollout(u0, theta)
    torch.cuda.synchronize()
    t_single = time.perf_counter() - t0
    logger.info(f"Total time used for single-core solver: {t_single:.3f}s")
    speed = torch.sqrt(u[0] ** 2 + u[1] ** 2).cpu().numpy()
    os.makedirs(args.out, exist_ok=True)
    npy, png = output_paths(args.out, config.N)
    np.save(npy, speed)
    t_label = f"t={config.steps * config.dt:.3f}"
    wrote_png = _save_png(speed, png, config.L,
                          f"Single-core ({config.advection_scheme}) {t_label}\ntime: {t_single:.3f}s")
    logger.info(f"Saved: {npy}" + (f", {png}" if wrote_png else ""))
    return {"J": float(J), "seconds": t_single, "npy": npy, "png": png if wrote_png else None,
            "steps": config.steps, "N": config.N}


if __name__ == "__main__":
    main()
