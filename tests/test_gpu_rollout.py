"""Rollout, objective and adjoint gradient of the HIP path against the
oracle and the reference's committed results (NS_solver.py:732-756)."""
import json
import math
import os

import numpy as np
import pytest
import torch

from oracle import adjoint_oracle as AO
from oracle import ns_oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def S():
    from ns_solver_jax_differential_distributed_amd import NS_solver
    return NS_solver


def gold():
    with open(os.path.join(GOLD, "reference_results.json")) as f:
        return json.load(f)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def solver(N, L, nu, dt, steps, scheme="central", **kw):
    grid = S().make_grid(N, L)
    phys = S().Phys(nu=nu, mu0=0.5, sigma0=0.1, C=1.0)
    return S().SingleCoreNSSolver(grid, phys, S().Time(dt=dt, steps=steps), advection_scheme=scheme, **kw)


def test_golden_speed_field_N128():
    """Runner output for sample_config/config_N_128.json: N=128, nu=1e-4,
    t=0.5 (160 steps), central, theta=[1,0.5], u0=0 (runner.py:115-125)."""
    N, nu = 128, 1e-4
    dt = O.config_dt(N, 1.0, nu)
    steps = O.config_steps(0.5, dt)
    s = solver(N, 1.0, nu, dt, steps)
    J, u = s.rollout(torch.zeros(2, N, N), [1.0, 0.5])
    speed = torch.sqrt(u[0] ** 2 + u[1] ** 2).cpu().numpy()
    ref = np.load(os.path.join(GOLD, "speed_single_N128.npy"))
    assert np.max(np.abs(speed - ref)) < 2e-6


@pytest.mark.parametrize("scheme", ["central", "upwind"])
@pytest.mark.parametrize("nu", [0.005, 0.01, 0.05])
def test_golden_scheme_comparison(scheme, nu):
    """J, final kinetic energy and max |u| of the reference's scheme study
    (studies.py:545-606): N=128, L=1, t=0.5, theta=[1,0.5], u0=0."""
    g = gold()["scheme_comparison"]
    key = f"{scheme}_nu_{nu:.4f}"
    N = 128
    dt = g[key + "_dt"]
    steps = O.config_steps(0.5, dt)
    s = solver(N, 1.0, nu, dt, steps, scheme)
    J, u = s.rollout(torch.zeros(2, N, N), [1.0, 0.5])
    un = u.cpu().numpy().astype(np.float64)
    # The state crosses four float32 FFTs per step (~1e-7 relative each
    # step); over 164..1639 steps that random walk reaches a few 1e-6 in J.
    assert abs(float(J) - g[key + "_objective"]) / g[key + "_objective"] < 1e-5
    assert abs(O.kinetic_energy(un, 1.0 / N) - g[key + "_kinetic_energy"]) / g[key + "_kinetic_energy"] < 1e-5
    assert abs(O.speed(un).max() - g[key + "_max_velocity"]) / g[key + "_max_velocity"] < 1e-5


def test_golden_speed_field_N512():
    """config_N_512.json runner output (640 steps), per-row mean of |u|."""
    N, nu = 512, 1e-4
    dt = O.config_dt(N, 1.0, nu)
    steps = O.config_steps(0.5, dt)
    s = solver(N, 1.0, nu, dt, steps)
    _, u = s.rollout(torch.zeros(2, N, N), [1.0, 0.5])
    speed = torch.sqrt(u[0] ** 2 + u[1] ** 2).double().mean(dim=1).cpu().numpy()
    ref = np.array(gold()["speed_N512"]["row_mean"])
    assert np.max(np.abs(speed - ref)) < 2e-6


@pytest.mark.parametrize("N,steps", [(16, 30), (64, 25), (256, 12)])
@pytest.mark.parametrize("scheme", ["central", "upwind"])
@pytest.mark.parametrize("objective", ["work", "mean_ke"])
def test_rollout_matches_fp64_oracle(N, steps, scheme, objective):
    from ns_solver_jax_differential_distributed_amd import workloads as W
    L = 2 * math.pi
    u0 = W.initial_velocity(N, seed=3, kmax=4)
    nu = 1e-3
    dt = W.cfl_dt(u0, L, 0.3, nu)
    theta = [1.2, 0.8, 0.6]
    s = solver(N, L, nu, dt, steps, scheme, objective=objective)
    J, u = s.rollout(torch.from_numpy(u0), theta)
    Jo, uo = O.rollout(u0.astype(np.float64), theta, N, L, O.PhysParams(nu), dt, steps, scheme, objective)
    assert rel_l2(u.cpu().numpy(), uo) < 1e-5
    assert abs(float(J) - Jo) / abs(Jo) < 1e-5


def test_rollout_graph_and_eager_agree_bitwise():
    N, steps = 64, 70
    rng = np.random.default_rng(0)
    u0 = (0.2 * rng.standard_normal((2, N, N))).astype(np.float32)
    a = solver(N, 1.0, 0.01, 1e-3, steps, use_graph=True)
    b = solver(N, 1.0, 0.01, 1e-3, steps, use_graph=False)
    Ja, ua = a.rollout(torch.from_numpy(u0), [0.9, 0.4])
    Jb, ub = b.rollout(torch.from_numpy(u0), [0.9, 0.4])
    Ja2, ua2 = a.rollout(torch.from_numpy(u0), [0.9, 0.4])
    assert torch.equal(ua, ub) and torch.equal(ua, ua2)
    assert float(Ja) == float(Jb) == float(Ja2)


def test_zero_steps_returns_input():
    s = solver(16, 1.0, 0.01, 1e-3, 0)
    u0 = torch.randn(2, 16, 16)
    J, u = s.rollout(u0, [1.0, 0.5])
    assert float(J) == 0.0 and torch.equal(u.cpu(), u0)


@pytest.mark.parametrize("scheme", ["central", "upwind"])
@pytest.mark.parametrize("objective", ["work", "mean_ke"])
def test_gradient_matches_adjoint_oracle(scheme, objective):
    N, L, nu, dt, steps = 32, 1.0, 0.01, 1e-3, 23
    rng = np.random.default_rng(4)
    u0 = (0.4 * rng.standard_normal((2, N, N))).astype(np.float32)
    theta = [0.9, 0.4, 1.7]
    s = solver(N, L, nu, dt, steps, scheme, objective=objective, ckpt_every=5)
    J, u, g, du0 = s.rollout_and_grad(torch.from_numpy(u0), theta)
    Jo, uo, go, lo = AO.rollout_gradient(u0.astype(np.float64), theta, N, L, O.PhysParams(nu), dt, steps,
                                         scheme, objective)
    gg = g.cpu().numpy()
    assert gg[2] == 0.0
    assert np.max(np.abs(gg[:2] - go[:2])) / np.max(np.abs(go[:2])) < 1e-4
    assert rel_l2(du0.cpu().numpy(), lo) < 1e-4
    assert abs(float(J) - Jo) / abs(Jo) < 1e-5


@pytest.mark.parametrize("scheme", ["central", "upwind"])
def test_gradient_without_recomputation_is_bitwise_equal(scheme):
    """ckpt_every=1 (every state kept, reverse sweep reads the checkpoints)
    and ckpt_every=5 (segments recomputed from checkpoints) give identical
    bits: recomputed states are bit-identical to the forward's, and "auto"
    picks the first for a trajectory this small."""
    N, L, nu, dt, steps = 64, 1.0, 0.01, 1e-3, 17
    rng = np.random.default_rng(7)
    u0 = torch.from_numpy((0.4 * rng.standard_normal((2, N, N))).astype(np.float32))
    theta = [0.8, 1.3, 0.2]
    outs = []
    for k in (1, 5, "auto"):
        s = solver(N, L, nu, dt, steps, scheme, ckpt_every=k)
        outs.append([t.cpu() for t in s.rollout_and_grad(u0, theta)])
    assert s.engine.resolve_ckpt("auto", steps) == 1
    for a, b in zip(outs[0], outs[1]):
        assert torch.equal(a, b)
    for a, b in zip(outs[0], outs[2]):
        assert torch.equal(a, b)


@pytest.mark.parametrize("scheme", ["central", "upwind"])
def test_gradient_at_the_largest_grid_is_independent_of_checkpointing(scheme):
    """N=4096 (the largest supported grid), 12 steps: every state kept and
    the reverse sweep replayed from graphs, every state kept and launched
    directly, and segments of 5 recomputed from checkpoints give the same
    bits for J, u_final, dJ/dtheta and dJ/du0."""
    N, L, nu, steps = 4096, 2 * math.pi, 1e-4, 12
    from ns_solver_jax_differential_distributed_amd import workloads as W
    u0 = W.initial_velocity(N, seed=1, kmax=8)
    dt = W.cfl_dt(u0, L, 0.3, nu, stable=True, steps=steps)
    u0t = torch.from_numpy(u0)
    theta = [1.1, 0.7, 0.9]
    outs = []
    for k, graph in ((1, True), (1, False), (5, False)):
        s = solver(N, L, nu, dt, steps, scheme, ckpt_every=k, use_graph=graph)
        outs.append([t.cpu() for t in s.rollout_and_grad(u0t, theta)])
        del s
        torch.cuda.empty_cache()
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            assert torch.equal(a, b)
    J, uf, g, du0 = outs[0]
    assert torch.isfinite(uf).all() and torch.isfinite(du0).all() and float(g[2]) == 0.0


@pytest.mark.parametrize("scheme", ["central", "upwind"])
def test_graph_replayed_gradient_equals_eager(scheme):
    """75 steps with every state kept: the forward in graph-replayed chunks and
    the reverse sweep as two 32-step chunks replayed from one captured graph
    plus 11 steps launched directly.  Equal bits to the eager forward and
    sweep and to the recomputing sweep (ckpt_every=5), on three calls in a
    row with eager work between them."""
    N, L, nu, dt, steps = 64, 1.0, 0.01, 1e-3, 75
    rng = np.random.default_rng(11)
    u0 = torch.from_numpy((0.4 * rng.standard_normal((2, N, N))).astype(np.float32))
    theta = [0.9, 1.1]
    ref = [t.cpu() for t in solver(N, L, nu, dt, steps, scheme, ckpt_every=5,
                                   use_graph=False).rollout_and_grad(u0, theta)]
    eager = [t.cpu() for t in solver(N, L, nu, dt, steps, scheme, ckpt_every=1,
                                     use_graph=False).rollout_and_grad(u0, theta)]
    for a, b in zip(ref, eager):
        assert torch.equal(a, b)
    g = solver(N, L, nu, dt, steps, scheme, ckpt_every=1)
    other = solver(N, L, nu, dt, 9, scheme, use_graph=False)
    for _ in range(3):
        out = [t.cpu() for t in g.rollout_and_grad(u0, theta)]
        for a, b in zip(ref, out):
            assert torch.equal(a, b)
        other.rollout(u0, theta)
        torch.cuda.synchronize()


def test_rollouts_of_different_lengths_share_graphs_and_agree():
    """One solver, rollouts of 40, 100 and again 40 steps (the step graphs are
    cached per partial-sum capacity, not per length): each equals a fresh
    solver's rollout of that length bit for bit, and its J the sum over its
    own steps."""
    N, L, nu, dt = 64, 1.0, 0.01, 1e-3
    rng = np.random.default_rng(5)
    u0 = torch.from_numpy((0.3 * rng.standard_normal((2, N, N))).astype(np.float32))
    theta = [1.2, 0.7]
    s = solver(N, L, nu, dt, 40)
    for steps in (40, 100, 40):
        J, uf = s.rollout(u0, theta, steps=steps)
        J1, uf1 = solver(N, L, nu, dt, steps).rollout(u0, theta)
        assert torch.equal(uf, uf1) and float(J) == float(J1)
    assert s.engine._jcap == 1024


def test_jit_gradient_matches_jax_reverse_mode():
    """JAX's gradient at theta=[1,0.5] from the reference's SGD(lr=0.1)
    trajectory (studies.py:890-918): g = (theta_0 - theta_1)/0.1;
    default config N=128, nu=0.01, 328 steps, u0=0."""
    p = np.array(gold()["sgd_lr0.1"]["parameters"])
    g_jax = (p[0] - p[1]) / 0.1
    N, nu = 128, 0.01
    dt = O.config_dt(N, 1.0, nu)
    steps = O.config_steps(0.5, dt)
    s = solver(N, 1.0, nu, dt, steps)
    g = s.jit_gradient(np.array([1.0, 0.5], np.float32)).cpu().numpy()
    # theta_1 is stored in float32: |g_jax error| <= ulp(1)/2/0.1 = 6e-7
    assert abs(g[1] - g_jax[1]) < 2e-6
    assert abs(g[0] - g_jax[0]) < 2e-6
