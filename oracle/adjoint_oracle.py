"""Discrete adjoint of the Chorin rollout, restated in NumPy (FP64).

TEST INFRASTRUCTURE ONLY (see ``oracle/ns_oracle.py``).

The reference differentiates ``SingleCoreNSSolver.objective`` with
``jax.jit(jax.grad(...))`` (NS_solver.py:725), i.e. reverse mode through the
``lax.scan`` of NS_solver.py:742-749.  This module writes that reverse sweep
out by hand:

  u_{n+1} = P(u*_n),  u*_n = u_n + dt (nu L u_n + A(u_n) + f_n)
  J = sum_n dt dx^2 <u_n, f_n>                      (NS_solver.py:746)

  lambda_S = 0
  mu_n     = P^T lambda_{n+1} = P lambda_{n+1}      (P = I - G L^-1 D is
                                                      symmetric: D^T = -G)
  lambda_n = dJ/du_n|direct + mu_n + dt (nu L mu_n + (dA/du_n)^T mu_n)
  dJ/df_n  = dJ/df_n|direct + dt mu_n
  dJ/dtheta = sum_n sum_cells (dJ/df_n)_x * d f_x/d theta (x, t_n)

``round`` in the forcing wrap contributes no derivative and ``where`` in the
upwind stencil does not differentiate its condition (JAX semantics).
Pinned by FD on the FP64 rollout, dot-product tests and a dense Jacobian on
an 8x8 grid (tests/test_oracle.py), and by JAX's own gradient at
theta=[1,0.5] recovered from the reference's committed SGD trajectory.
"""
from __future__ import annotations

import numpy as np

from . import ns_oracle as O


def central_advect_vjp(u, mu, dx):
    """(dA/du)^T mu for A_c = -(u_x G_x u_c + u_y G_y u_c) (NS_solver.py:474-477):
    (.)_c = G_x(u_x mu_c) + G_y(u_y mu_c) - sum_d mu_d G_c u_d."""
    ux, uy = u
    out = np.empty_like(mu)
    gxx, gxy = O.grad(ux, dx)   # d_x ux, d_y ux
    gyx, gyy = O.grad(uy, dx)   # d_x uy, d_y uy
    for c in range(2):
        a, _ = O.grad(ux * mu[c], dx)
        _, b = O.grad(uy * mu[c], dx)
        out[c] = a + b
    out[0] -= mu[0] * gxx + mu[1] * gyx
    out[1] -= mu[0] * gxy + mu[1] * gyy
    return out


def _upwind_D(v, s, axis, dx):
    """Upwind difference with the side fixed by s (True: backward)."""
    back = v - np.roll(v, 1, axis=axis)
    fwd = np.roll(v, -1, axis=axis) - v
    return np.where(s, back, fwd) / dx


def _upwind_DT(w, s, axis, dx):
    """Transpose of ``_upwind_D`` for the fixed side pattern s."""
    own = np.where(s, w, -w)
    from_next = np.where(np.roll(s, -1, axis=axis), -np.roll(w, -1, axis=axis), 0.0)
    from_prev = np.where(np.roll(s, 1, axis=axis), 0.0, np.roll(w, 1, axis=axis))
    return (own + from_next + from_prev) / dx


def upwind_advect_vjp(u, mu, dx):
    """(dA/du)^T mu for the upwind scheme with the sides chosen by the forward
    signs of u (NS_solver.py:257-265)."""
    ux, uy = u
    sx = ux >= 0
    sy = uy >= 0
    out = np.empty_like(mu)
    Dxu = [_upwind_D(u[c], sx, 0, dx) for c in range(2)]
    Dyu = [_upwind_D(u[c], sy, 1, dx) for c in range(2)]
    for c in range(2):
        out[c] = -_upwind_DT(ux * mu[c], sx, 0, dx) - _upwind_DT(uy * mu[c], sy, 1, dx)
    out[0] -= mu[0] * Dxu[0] + mu[1] * Dxu[1]
    out[1] -= mu[0] * Dyu[0] + mu[1] * Dyu[1]
    return out


def ustar_vjp(u, mu, dx, nu, dt, scheme):
    """(du*/du)^T mu = mu + dt (nu L mu + (dA/du)^T mu)."""
    adv = central_advect_vjp(u, mu, dx) if scheme == "central" else upwind_advect_vjp(u, mu, dx)
    lap = np.stack([O.laplacian(mu[0], dx), O.laplacian(mu[1], dx)])
    return mu + dt * (nu * lap + adv)


def projection(v, dx):
    """P v = v - G L^-1 D v (NS_solver.py:691-697 with dt cancelled)."""
    return O.project(v, dx, 1.0)


def forcing_theta_partials(theta, t, x, phys: O.PhysParams):
    """d f_x / dA and d f_x / dB at coordinates x (NS_solver.py:621-627).
    ``round`` contributes nothing; theta[2:] is never read."""
    A, B = float(theta[0]), float(theta[1])
    mu_t = phys.mu0 + np.sin(2 * np.pi * A * t)
    sig_t = phys.sigma0 + np.cos(2 * np.pi * B * t) ** 2
    d = (x - mu_t) - np.round(x - mu_t)
    f = phys.C * np.exp(-(d * d) / (2.0 * sig_t * sig_t))
    dA = f * d / sig_t ** 2 * (2 * np.pi * t * np.cos(2 * np.pi * A * t))
    dB = f * d * d / sig_t ** 3 * (-2 * np.pi * t * np.sin(4 * np.pi * B * t))
    return dA, dB


def rollout_gradient(u0, theta, N, L, phys: O.PhysParams, dt, steps,
                     scheme="central", objective="work", jseed=1.0, lam_final=None):
    """Returns (J, u_final, dL/dtheta (len(theta)), dL/du0) in FP64 for
    L = jseed * J + <lam_final, u_final> (jseed = 1, lam_final = 0: dJ)."""
    u0 = np.asarray(u0, dtype=np.float64)
    dx = L / N
    x = O.grid_x(N, L, np.float64)
    J, u_final, states = O.rollout(u0, theta, N, L, phys, dt, steps, scheme,
                                   objective, keep_states=True)
    theta = np.asarray(theta, dtype=np.float64)
    g_theta = np.zeros(len(theta))
    lam = np.zeros_like(u0) if lam_final is None else np.asarray(lam_final, dtype=np.float64).copy()
    for n in range(steps - 1, -1, -1):
        t = float(n) * dt
        un = states[n]
        fx = O.forcing_x(theta, t, x, phys)
        mu = projection(lam, dx)
        if objective == "work":
            direct_u = np.zeros_like(un)
            direct_u[0] = jseed * dt * dx * dx * fx[:, None]
            gf = jseed * dt * dx * dx * un[0] + dt * mu[0]
        else:
            direct_u = jseed * un / (steps * N * N)
            gf = dt * mu[0]
        dA, dB = forcing_theta_partials(theta, t, x, phys)
        row = gf.sum(axis=1)
        g_theta[0] += np.dot(row, dA)
        g_theta[1] += np.dot(row, dB)
        lam = direct_u + ustar_vjp(un, mu, dx, phys.nu, dt, scheme)
    return J, u_final, g_theta, lam
