"""The gradient configuration at its full size: N=512 on [0, 2 pi)^2,
2000 steps, checkpoint every 50 steps with recomputation, 16 seeded theta
vectors batched (the reverse mode of NS_solver.py:752-756 at the size the
gradient benchmark of studies.py:271-521 is quoted on).

The FP64 adjoint oracle pins the gradient to 1e-4 at small sizes
(test_gpu_rollout.py, test_gpu_batched.py); at this size the oracle is too
slow, so the check is against central differences of the same float32 GPU
objective, all 64 perturbed rollouts in one batch.  Differences with the
reference's h = 1e-4 are dominated by float32 rounding of J (noise / h: up
to 7 % of |grad|); h = 1e-3 balances rounding against truncation (measured
worst case 0.7 %, median 0.1 %), h = 1e-2 is truncation-dominated.  The
bounds below sit about three times above the measured errors: they catch a
wrong term or sign at full size, not a 1e-4 bias."""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_full_size_gradient_matches_central_differences():
    from ns_solver_jax_differential_distributed_amd import NS_solver as S
    from ns_solver_jax_differential_distributed_amd import workloads as W
    N, steps, k, nth, h = 512, 2000, 50, 16, 1e-3
    L, nu = 2 * math.pi, 1e-3
    u0 = W.initial_velocity(N, seed=0, kmax=8)
    dt = W.cfl_dt(u0, L, 0.3, nu, stable=True, steps=steps)
    th = W.theta_batch(nth, seed=0)
    bs = S.BatchedNSSolver(N, L, nth, ckpt_every=k)
    u0d = torch.from_numpy(np.broadcast_to(u0, (nth,) + u0.shape).copy()).cuda()
    J, uf, g, du0 = bs.rollout_and_grad(u0d, th, nu, dt, steps)
    g = g.cpu().numpy()
    assert np.isfinite(g).all() and g.shape == (nth, 2)

    # theta_i +- h e_A, +- h e_B for every theta, one batch of 64
    thp = np.repeat(th, 4, axis=0).astype(np.float64)
    thp[0::4, 0] += h
    thp[1::4, 0] -= h
    thp[2::4, 1] += h
    thp[3::4, 1] -= h
    thp = thp.astype(np.float32)
    bf = S.BatchedNSSolver(N, L, 4 * nth)
    ub = torch.from_numpy(np.broadcast_to(u0, (4 * nth,) + u0.shape).copy()).cuda()
    Jp, _ = bf.rollout(ub, thp, nu, dt, steps)
    Jp = Jp.cpu().numpy().astype(np.float64)
    # the actual float32 steps taken
    fdA = (Jp[0::4] - Jp[1::4]) / (thp[0::4, 0].astype(np.float64) - thp[1::4, 0])
    fdB = (Jp[2::4] - Jp[3::4]) / (thp[2::4, 1].astype(np.float64) - thp[3::4, 1])
    scale = np.abs(g).max(axis=1)
    err = np.concatenate([np.abs(fdA - g[:, 0]) / scale, np.abs(fdB - g[:, 1]) / scale])
    assert err.max() < 2e-2, err.max()
    assert np.median(err) < 3e-3, np.median(err)
    # J of the unperturbed batch equals the objective of the same rollout alone
    J1, _ = S.BatchedNSSolver(N, L, 1).rollout(u0d[:1], th[:1], nu, dt, steps)
    assert float(J1[0]) == float(J[0])
