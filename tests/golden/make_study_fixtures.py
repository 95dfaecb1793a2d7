"""Optimizer trajectories driven by the FP64 oracle gradient.

The reference's committed optimisation study (Adam, SGD, RMSprop from
theta0 = [1, 0.5] at the default config) used JAX's float32 reverse-mode
gradient, whose tiny first component (dJ/dA ~ -2.6e-5) is about 1 % off the
FP64 value.  Adam and RMSprop normalise that component, so their x
trajectories follow that error.  This script reruns the four optimizers
(the package's float32 Optax restatements) on the FP64 discrete adjoint of
``oracle/adjoint_oracle.py`` and writes ``optimization_fp64_oracle.json``
next to it: theta and J at every step, the path a correct float32 gradient
must track, and the FP64 value and gradient at the gradient benchmark's
three parameter vectors.  Takes a few minutes on 4 cores:

    python tests/golden/make_study_fixtures.py
"""
import json
import os
import sys
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

STEPS = 100
CONFIGS = [("Adam (lr=0.01)", "adam", 0.01), ("Adam (lr=0.001)", "adam", 0.001),
           ("SGD (lr=0.1)", "sgd", 0.1), ("RMSprop (lr=0.01)", "rmsprop", 0.01)]


def _run(cfg):
    from ns_solver_jax_differential_distributed_amd import studies as ST
    from ns_solver_jax_differential_distributed_amd.NS_solver import SimConfig
    from oracle import adjoint_oracle as AO
    from oracle import ns_oracle as O
    name, kind, lr = cfg
    c = SimConfig()
    phys = O.PhysParams(c.nu, c.mu0, c.sigma0, c.C_force)
    u0 = np.zeros((2, c.N, c.N))

    def value_and_grad(th):
        J, _, g, _ = AO.rollout_gradient(u0, th.astype(np.float64), c.N, c.L, phys, c.dt, c.steps)
        return float(J), np.asarray(g[:2], np.float64)

    opt = getattr(ST, kind)(lr)
    th = np.array(c.theta0, np.float32)
    state = opt.init(th)
    params, objectives = [th.tolist()], []
    for _ in range(STEPS):
        J, g = value_and_grad(th)
        objectives.append(J)
        u, state = opt.update(g.astype(np.float32), state)
        th = ST.apply_updates(th, u)
        params.append(th.tolist())
    objectives.append(value_and_grad(th)[0])
    return name, {"parameters": params, "objectives": objectives}


GRAD_THETAS = ([1.0, 0.5], [0.5, 0.5], [2.0, 1.0])


def _grad(theta):
    from ns_solver_jax_differential_distributed_amd.NS_solver import SimConfig
    from oracle import adjoint_oracle as AO
    from oracle import ns_oracle as O
    c = SimConfig()
    phys = O.PhysParams(c.nu, c.mu0, c.sigma0, c.C_force)
    J, _, g, _ = AO.rollout_gradient(np.zeros((2, c.N, c.N)), np.array(theta, np.float64), c.N, c.L, phys,
                                     c.dt, c.steps)
    return {"theta": list(theta), "J": float(J), "grad": [float(g[0]), float(g[1])]}


def main():
    with Pool(len(CONFIGS)) as pool:
        out = dict(pool.map(_run, CONFIGS))
        # FP64 value and gradient at the gradient benchmark's parameter vectors
        out["gradients"] = pool.map(_grad, GRAD_THETAS)
    out = {"source": "oracle/adjoint_oracle.py FP64 gradient, default SimConfig, theta0=[1, 0.5], "
                     f"{STEPS} steps, float32 optimizer state", **out}
    with open(os.path.join(HERE, "optimization_fp64_oracle.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", os.path.join(HERE, "optimization_fp64_oracle.json"))


if __name__ == "__main__":
    main()
