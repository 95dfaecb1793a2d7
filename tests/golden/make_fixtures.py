"""Extract golden fixtures from the reference's committed result files.

Reads only data files (``.npy`` / ``.npz``, loaded with ``allow_pickle=False``)
under ``/root/reference/result``; no reference code is imported or run.
Writes ``reference_results.json`` and ``speed_single_N128.npy`` next to this
script.  Run once in a container where the reference is mounted:

    python tests/golden/make_fixtures.py
"""
import json
import os
import sys

import numpy as np

REF = os.environ.get("NS_REFERENCE_ROOT", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))


def _scalars(npz_path):
    d = np.load(npz_path, allow_pickle=False)
    return {k: float(d[k]) for k in d.files if d[k].shape == ()}


def main():
    res = os.path.join(REF, "result")
    if not os.path.isdir(res):
        sys.exit(f"reference result folder not found: {res}")
    out = {
        "source": "helmhlotz/ns_solver_jax_differential_distributed result/ (JAX 0.4.30, CPU)",
        # scheme comparison / viscosity study: N=128, L=1, t=0.5, u0=0,
        # theta=[1,0.5], dt and steps from SimConfig (studies.py:545-606, 105-150)
        "scheme_comparison": _scalars(os.path.join(res, "scheme_comparison_results.npz")),
        "viscosity_study": _scalars(os.path.join(res, "viscosity_study_results.npz")),
    }
    opt = np.load(os.path.join(res, "optimization_study_results.npz"), allow_pickle=False)
    # SGD(lr=0.1) trajectory: theta_{k+1} = theta_k - 0.1 * jax.grad(J)(theta_k)
    out["sgd_lr0.1"] = {
        "parameters": opt["SGD (lr=0.1)_parameters"][:6].astype(float).tolist(),
        "objectives": opt["SGD (lr=0.1)_objectives"][:6].astype(float).tolist(),
        "lr": 0.1,
    }
    # all four optimizer runs of the optimization study (studies.py:864-947,
    # 1120-1127): default config, theta0 = [1, 0.5], 100 steps each
    out["optimization_study"] = {}
    for name in ("Adam (lr=0.01)", "Adam (lr=0.001)", "SGD (lr=0.1)", "RMSprop (lr=0.01)"):
        out["optimization_study"][name] = {
            "parameters": opt[f"{name}_parameters"].astype(float).tolist(),
            "objectives": opt[f"{name}_objectives"].astype(float).tolist(),
            "obj_initial": float(opt[f"{name}_obj_initial"]),
            "obj_final": float(opt[f"{name}_obj_final"]),
            "improvement": float(opt[f"{name}_improvement"]),
        }
    gb = np.load(os.path.join(res, "gradient_benchmark_results.npz"), allow_pickle=False)
    out["gradient_benchmark"] = {
        k: (gb[k].astype(float).tolist() if gb[k].shape else float(gb[k]))
        for k in gb.files
    }
    # runner output at N=512 (config_N_512.json: nu=1e-4, t=0.5, central):
    # keep the per-row mean and spread of the 512x512 speed field
    s512 = np.load(os.path.join(res, "velocity_magnitude_distributed_P8_N512.npy"), allow_pickle=False)
    out["speed_N512"] = {
        "row_mean": s512.astype(np.float64).mean(axis=1).tolist(),
        "mean": float(s512.astype(np.float64).mean()),
        "min": float(s512.min()),
        "max": float(s512.max()),
    }
    with open(os.path.join(HERE, "reference_results.json"), "w") as f:
        json.dump(out, f, indent=1)
    s128 = np.load(os.path.join(res, "velocity_magnitude_single_N128.npy"), allow_pickle=False)
    np.save(os.path.join(HERE, "speed_single_N128.npy"), s128.astype(np.float32))
    print("wrote", os.path.join(HERE, "reference_results.json"))


if __name__ == "__main__":
    main()
