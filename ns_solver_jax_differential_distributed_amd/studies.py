"""Gradient and parameter-optimisation studies on the GPU adjoint.

Counterparts of the reference's ``studies.py`` classes that call the
single-core solver's objective and gradient:

  GradientBenchmark           adjoint vs central finite differences over a
                              sweep of step sizes (studies.py:271-521)
  ParameterOptimizationStudy  several optimizers from the same theta0
                              (studies.py:864-1002)

Same constructor arguments, ``run_all``/``save_results`` and result keys as
the reference, so saved ``.npz`` files have the reference's layout.  The
reference drives Optax; Optax is not a dependency here, so ``sgd``,
``adam`` and ``rmsprop`` restate Optax 0.2.4's update rules (the version
``requirements.txt:16`` of the reference pins) in float32, with its
defaults:

  sgd(lr):      u = -lr g
  adam(lr):     m = (1-b1) g + b1 m,  v = (1-b2) g^2 + b2 v,  k += 1,
                u = -lr (m / (1-b1^k)) / (sqrt(v / (1-b2^k)) + eps)
                (b1 = 0.9, b2 = 0.999, eps = 1e-8, outside the root)
  rmsprop(lr):  v = (1-d) g^2 + d v,  u = -lr g / sqrt(v + eps)
                (d = 0.9, eps = 1e-8 inside the root, v starts at 0)

What differs, by design:
  * When the objective / gradient callables offer ``batch`` (as
    ``SingleCoreNSSolver.jit_objective`` / ``jit_gradient`` do), a whole
    finite-difference sweep of one theta is one batched rollout, and all
    optimizers advance in lockstep, one batched rollout + adjoint per step
    whose forward pass also yields J at the current theta.  Every value is
    bit-identical to the one-theta-at-a-time path.
  * The objective returns J accumulated in FP64 (the reference's is
    float32), so finite differences at the smallest step sizes are less
    noisy than the reference's.
  * ``time_fd`` of a batched sweep is the sweep's wall time divided by the
    number of step sizes.
  * Plots are not produced.
"""
from __future__ import annotations

import logging
import os
import time
from dataclasses import dataclass
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np
import torch

logger = logging.getLogger(__name__)

_F32 = np.float32


def _f32(x) -> np.ndarray:
    if isinstance(x, torch.Tensor):
        x = x.detach().cpu().numpy()
    return np.asarray(x, dtype=_F32)


def _scalar(x) -> float:
    if isinstance(x, torch.Tensor):
        return float(x.detach().cpu())
    return float(x)


# ==========================
# Optimizers (Optax 0.2.4 update rules, float32)
# ==========================
class GradientTransformation:
    """``init(params) -> state`` and ``update(grads, state) -> (updates,
    state)``, like an Optax GradientTransformation on a float32 vector."""

    def init(self, params):
        raise NotImplementedError

    def update(self, grads, state, params=None):
        raise NotImplementedError


@dataclass
class _Sgd(GradientTransformation):
    learning_rate: float

    def init(self, params):
        return ()

    def update(self, grads, state, params=None):
        return _f32(grads) * _F32(-self.learning_rate), state


@dataclass
class _Adam(GradientTransformation):
    learning_rate: float
    b1: float = 0.9
    b2: float = 0.999
    eps: float = 1e-8
    eps_root: float = 0.0

    def init(self, params):
        z = np.zeros_like(_f32(params))
        return {"count": 0, "mu": z, "nu": z.copy()}

    def update(self, grads, state, params=None):
        g = _f32(grads)
        mu = _F32(1.0 - self.b1) * g + _F32(self.b1) * state["mu"]
        nu = _F32(1.0 - self.b2) * (g * g) + _F32(self.b2) * state["nu"]
        k = state["count"] + 1
        mu_hat = mu / _F32(1.0 - _F32(self.b1) ** _F32(k))
        nu_hat = nu / _F32(1.0 - _F32(self.b2) ** _F32(k))
        u = mu_hat / (np.sqrt(nu_hat + _F32(self.eps_root)) + _F32(self.eps))
        return u * _F32(-self.learning_rate), {"count": k, "mu": mu, "nu": nu}


@dataclass
class _RmsProp(GradientTransformation):
    learning_rate: float
    decay: float = 0.9
    eps: float = 1e-8
    initial_scale: float = 0.0

    def init(self, params):
        return {"nu": np.full_like(_f32(params), _F32(self.initial_scale))}

    def update(self, grads, state, params=None):
        g = _f32(grads)
        nu = _F32(1.0 - self.decay) * (g * g) + _F32(self.decay) * state["nu"]
        u = g / np.sqrt(nu + _F32(self.eps))
        return u * _F32(-self.learning_rate), {"nu": nu}


def sgd(learning_rate: float) -> GradientTransformation:
    """Plain gradient descent (optax.sgd without momentum)."""
    return _Sgd(float(learning_rate))


def adam(learning_rate: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8,
         eps_root: float = 0.0) -> GradientTransformation:
    """Adam with bias correction (optax.adam)."""
    return _Adam(float(learning_rate), b1, b2, eps, eps_root)


def rmsprop(learning_rate: float, decay: float = 0.9, eps: float = 1e-8,
            initial_scale: float = 0.0) -> GradientTransformation:
    """RMSprop, eps inside the square root, no centring or momentum
    (optax.rmsprop defaults)."""
    return _RmsProp(float(learning_rate), decay, eps, initial_scale)


def apply_updates(params, updates) -> np.ndarray:
    """params + updates in the params' dtype (optax.apply_updates)."""
    p = _f32(params)
    return (p + _f32(updates)).astype(p.dtype)


# ==========================
# Gradient benchmark (studies.py:271-386, 505-521)
# ==========================
class GradientBenchmark:
    """Adjoint gradient vs central finite differences
    g_i = (J(theta + eps e_i) - J(theta - eps e_i)) / (2 eps) over a sweep of
    eps, per parameter vector (studies.py:271-386)."""

    def __init__(self, objective: Callable, grad_jax: Callable, param_vectors: Optional[list] = None,
                 epsilon_range=None, jax_timing_runs: int = 5, config=None):
        self.objective = objective
        self.grad_jax = grad_jax
        self.config = config
        if param_vectors is None:
            param_vectors = [[1.0, 0.5], [0.5, 1.0], [2.0, 0.2]]
        self.param_vectors = [_f32(p) for p in param_vectors]
        if epsilon_range is None:
            epsilon_range = np.logspace(-7, -1, 20)
        self.epsilon_range = _f32(epsilon_range).reshape(-1)
        self.jax_timing_runs = max(1, int(jax_timing_runs))
        self.results: Dict[int, dict] = {}
        logger.info(f"GradientBenchmark initialized: {len(self.epsilon_range)} epsilon values, "
                    f"{len(self.param_vectors)} param vectors")

    def _fd_thetas(self, theta: np.ndarray, eps: float) -> np.ndarray:
        rows = []
        for i in range(len(theta)):
            tp, tm = theta.copy(), theta.copy()
            tp[i] = tp[i] + _F32(eps)
            tm[i] = tm[i] - _F32(eps)
            rows += [tp, tm]
        return np.stack(rows)

    @staticmethod
    def _fd_from_values(J: np.ndarray, n: int, eps: float) -> np.ndarray:
        return np.array([(J[2 * i] - J[2 * i + 1]) / (2.0 * float(eps)) for i in range(n)])

    def _compute_fd_gradient(self, theta, epsilon: float) -> np.ndarray:
        """Central differences with one objective call per point."""
        th = _f32(theta)
        rows = self._fd_thetas(th, float(epsilon))
        J = np.array([_scalar(self.objective(r)) for r in rows])
        return self._fd_from_values(J, len(th), float(epsilon))

    def _sweep_fd(self, theta: np.ndarray):
        """{eps: (grad_fd, seconds)}: one batched call for the whole sweep when
        the objective offers ``batch``, else point by point."""
        batch = getattr(self.objective, "batch", None)
        out = {}
        if batch is None:
            for eps in self.epsilon_range:
                t0 = time.perf_counter()
                g = self._compute_fd_gradient(theta, float(eps))
                out[float(eps)] = (g, time.perf_counter() - t0)
            return out
        rows = np.concatenate([self._fd_thetas(theta, float(e)) for e in self.epsilon_range])
        t0 = time.perf_counter()
        J = batch(rows)
        J = J.detach().cpu().numpy() if isinstance(J, torch.Tensor) else np.asarray(J)
        per = (time.perf_counter() - t0) / len(self.epsilon_range)
        m = 2 * len(theta)
        for j, eps in enumerate(self.epsilon_range):
            out[float(eps)] = (self._fd_from_values(J[j * m:(j + 1) * m], len(theta), float(eps)), per)
        return out

    def sweep_epsilon(self, theta) -> dict:
        th = _f32(theta)
        logger.info(f"Sweeping epsilon for theta={th}...")
        # first call (engine set-up) and steady-state calls timed apart; the
        # copy to the host inside _grad64 waits for the GPU
        t0 = time.perf_counter()
        g = _grad64(self.grad_jax, th)
        time_first = time.perf_counter() - t0
        samples = []
        for _ in range(self.jax_timing_runs):
            t1 = time.perf_counter()
            g = _grad64(self.grad_jax, th)
            samples.append(time.perf_counter() - t1)
        exec_mean = float(np.mean(samples))
        exec_std = float(np.std(samples))
        sweep = {}
        for eps, (gfd, t_fd) in self._sweep_fd(th).items():
            abs_err = float(np.linalg.norm(g - gfd))
            sweep[eps] = {"grad_fd": gfd, "abs_error": abs_err,
                          "rel_error": abs_err / (float(np.linalg.norm(gfd)) + 1e-12), "time_fd": t_fd}
        return {"grad_jax": g, "time_jax": exec_mean, "time_jax_compile_first": time_first,
                "time_jax_exec_mean": exec_mean, "time_jax_exec_std": exec_std, "epsilon_sweep": sweep}

    def run_all(self) -> dict:
        logger.info(f"Running gradient benchmark on {len(self.param_vectors)} parameter vectors...")
        self.results = {i: self.sweep_epsilon(th) for i, th in enumerate(self.param_vectors)}
        return self.results

    def save_results(self, filename: str = "result/gradient_benchmark_results.npz"):
        """Same keys as the reference's file (studies.py:505-521)."""
        if not self.results:
            logger.warning("No results to save")
            return
        data = {}
        for i, r in self.results.items():
            data[f"param_{i}_time_jax"] = r["time_jax"]
            data[f"param_{i}_time_jax_compile_first"] = r["time_jax_compile_first"]
            data[f"param_{i}_time_jax_exec_mean"] = r["time_jax_exec_mean"]
            data[f"param_{i}_time_jax_exec_std"] = r["time_jax_exec_std"]
            eps = sorted(r["epsilon_sweep"])
            data[f"param_{i}_epsilons"] = np.array(eps)
            data[f"param_{i}_rel_errors"] = np.array([r["epsilon_sweep"][e]["rel_error"] for e in eps])
            data[f"param_{i}_abs_errors"] = np.array([r["epsilon_sweep"][e]["abs_error"] for e in eps])
            data[f"param_{i}_times_fd"] = np.array([r["epsilon_sweep"][e]["time_fd"] for e in eps])
        _savez(filename, data)


def _grad64(grad, theta) -> np.ndarray:
    g = grad(theta)
    if isinstance(g, torch.Tensor):
        g = g.detach().cpu().numpy()
    return np.asarray(g, dtype=np.float64)


def _savez(filename: str, data: dict) -> None:
    d = os.path.dirname(filename)
    if d:
        os.makedirs(d, exist_ok=True)
    np.savez(filename, **data)
    logger.info(f"Saved to {filename}")


# ==========================
# Parameter optimisation (studies.py:864-947)
# ==========================
class ParameterOptimizationStudy:
    """Runs each optimizer for ``optim_steps`` steps from theta0, recording J
    and theta after every update (studies.py:864-947)."""

    def __init__(self, objective: Callable, grad_jax: Callable, theta0, optim_configs: Optional[list] = None,
                 optim_steps: int = 100):
        self.objective = objective
        self.grad_jax = grad_jax
        self.theta0 = _f32(theta0)
        self.optim_steps = int(optim_steps)
        self.optim_configs = optim_configs or [
            {"name": "Adam (lr=0.01)", "optimizer": adam(0.01)},
            {"name": "Adam (lr=0.001)", "optimizer": adam(0.001)},
            {"name": "SGD (lr=0.1)", "optimizer": sgd(0.1)},
            {"name": "RMSprop (lr=0.01)", "optimizer": rmsprop(0.01)},
        ]
        self.results: Dict[str, dict] = {}
        logger.info(f"ParameterOptimizationStudy initialized: {len(self.optim_configs)} optimizers, "
                    f"{self.optim_steps} steps")

    @staticmethod
    def _summary(name: str, objectives: List[float], params: List[np.ndarray]) -> dict:
        obj0, obj_final = float(objectives[0]), float(objectives[-1])
        improvement = float((obj0 - obj_final) / (obj0 + 1e-20))
        logger.info(f"{name}: initial={obj0:.6e}, final={obj_final:.6e}, improvement={improvement * 100:.3f}%")
        return {"obj_initial": obj0, "obj_final": obj_final, "improvement": improvement,
                "history": {"steps": np.arange(len(objectives), dtype=np.int32),
                            "objectives": np.array(objectives, dtype=np.float64),
                            "parameters": np.stack(params).astype(np.float32)}}

    def _run_single_optimizer(self, name: str, optimizer: GradientTransformation) -> dict:
        theta = self.theta0.copy()
        state = optimizer.init(theta)
        objectives, params = [_scalar(self.objective(theta))], [theta.copy()]
        for _ in range(self.optim_steps):
            updates, state = optimizer.update(_f32(_grad64(self.grad_jax, theta)), state)
            theta = apply_updates(theta, updates)
            objectives.append(_scalar(self.objective(theta)))
            params.append(theta.copy())
        return self._summary(name, objectives, params)

    def _run_lockstep(self) -> Dict[str, dict]:
        """All optimizers together: per step one batched rollout + adjoint at
        every optimizer's current theta (its forward pass gives J there), then
        one batched objective after the last update."""
        vag = self.grad_jax.value_and_grad_batch
        names = [c["name"] for c in self.optim_configs]
        opts = [c["optimizer"] for c in self.optim_configs]
        thetas = [self.theta0.copy() for _ in opts]
        states = [o.init(t) for o, t in zip(opts, thetas)]
        objectives = [[] for _ in opts]
        params = [[t.copy()] for t in thetas]
        for _ in range(self.optim_steps):
            J, G = vag(np.stack(thetas))
            J = J.detach().cpu().numpy()
            G = G.detach().cpu().numpy()
            for b, o in enumerate(opts):
                objectives[b].append(float(J[b]))
                updates, states[b] = o.update(_f32(G[b]), states[b])
                thetas[b] = apply_updates(thetas[b], updates)
                params[b].append(thetas[b].copy())
        J = self.objective.batch(np.stack(thetas))
        J = J.detach().cpu().numpy() if isinstance(J, torch.Tensor) else np.asarray(J)
        for b in range(len(opts)):
            objectives[b].append(float(J[b]))
        return {n: self._summary(n, objectives[b], params[b]) for b, n in enumerate(names)}

    def run_all(self) -> dict:
        logger.info(f"Testing {len(self.optim_configs)} optimizer configurations over {self.optim_steps} steps")
        if (hasattr(self.grad_jax, "value_and_grad_batch") and hasattr(self.objective, "batch")
                and self.optim_steps > 0):
            self.results = self._run_lockstep()
        else:
            self.results = {c["name"]: self._run_single_optimizer(c["name"], c["optimizer"])
                            for c in self.optim_configs}
        return self.results

    def save_results(self, filename: str = "result/optimization_study_results.npz"):
        """Same keys as the reference's file (studies.py:931-947)."""
        if not self.results:
            logger.warning("No optimization results to save")
            return
        data = {}
        for name, r in self.results.items():
            data[f"{name}_obj_initial"] = r["obj_initial"]
            data[f"{name}_obj_final"] = r["obj_final"]
            data[f"{name}_improvement"] = r["improvement"]
            data[f"{name}_steps"] = r["history"]["steps"]
            data[f"{name}_objectives"] = r["history"]["objectives"]
            data[f"{name}_parameters"] = r["history"]["parameters"]
        _savez(filename, data)


def main(argv: Optional[Sequence[str]] = None) -> None:
    """The gradient benchmark and the optimisation study of the reference's
    studies main (studies.py:1005-1127) on the GPU solver."""
    import argparse
    from . import NS_solver as S
    ap = argparse.ArgumentParser(description="Gradient and optimisation studies on the GPU adjoint")
    ap.add_argument("--config", type=str, default=None, help="JSON config (defaults for missing fields)")
    ap.add_argument("--out", type=str, default="result")
    args = ap.parse_args(argv)
    cfg = S.load_config(args.config)
    solver = S.SingleCoreNSSolver(S.make_grid(cfg.N, L=cfg.L),
                                  S.Phys(nu=cfg.nu, mu0=cfg.mu0, sigma0=cfg.sigma0, C=cfg.C_force),
                                  S.Time(dt=cfg.dt, steps=cfg.steps), advection_scheme=cfg.advection_scheme)
    gb = GradientBenchmark(solver.jit_objective, solver.jit_gradient,
                           param_vectors=[cfg.theta0, [0.5, 0.5], [2.0, 1.0]],
                           epsilon_range=np.logspace(-7, -1, 20), config=cfg)
    gb.run_all()
    gb.save_results(os.path.join(args.out, "gradient_benchmark_results.npz"))
    st = ParameterOptimizationStudy(solver.jit_objective, solver.jit_gradient, cfg.theta0, optim_steps=100)
    st.run_all()
    st.save_results(os.path.join(args.out, "optimization_study_results.npz"))


if __name__ == "__main__":
    main()
