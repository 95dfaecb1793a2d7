"""Seeded synthetic inputs for the benchmark workloads.

u0 = Taylor-Green (sin x cos y, -cos x sin y) on [0, 2 pi)^2 plus a random
perturbation u' = (d_y psi, -d_x psi) built from a streamfunction
psi = sum_{0<|k|<=K} Re[(a_k + i b_k) e^{i k.x}], a, b ~ N(0, 1) drawn from
numpy.random.default_rng(seed), scaled so that max|u'| = amplitude.  The
derivatives are the solver's own central differences, so the discrete
divergence of u0 is zero up to round-off.  theta ~ U[0.5, 1.5]^3 and
nu log-uniform are drawn from the same generator.
"""
from __future__ import annotations

import math

import numpy as np


def _central(f, axis, dx):
    return (np.roll(f, -1, axis=axis) - np.roll(f, 1, axis=axis)) / (2.0 * dx)


def streamfunction(N: int, rng, kmax: int) -> np.ndarray:
    """psi(x_i, y_j) = sum_{0<|k|<=kmax} Re[(a_k + i b_k) e^{i k.x}] on the
    N x N grid, evaluated with one inverse FFT (modes placed at their
    integer wavenumbers; kmax < N/2)."""
    ks = [(i, j) for i in range(-kmax, kmax + 1) for j in range(-kmax, kmax + 1)
          if 0 < i * i + j * j <= kmax * kmax]
    coef = rng.standard_normal((len(ks), 2))
    C = np.zeros((N, N), dtype=np.complex128)
    for (i, j), (a, b) in zip(ks, coef):
        C[i % N, j % N] += a + 1j * b
    return (np.fft.ifft2(C) * (N * N)).real


def initial_velocity(N: int, seed: int = 0, kmax: int = 4, amplitude: float = 0.1,
                     L: float = 2.0 * math.pi, dtype=np.float32) -> np.ndarray:
    """(2, N, N) Taylor-Green + divergence-free perturbation (see module doc)."""
    if 2 * kmax >= N:
        raise ValueError(f"kmax={kmax} needs N > {2 * kmax}")
    rng = np.random.default_rng(seed)
    dx = L / N
    kk = 2.0 * math.pi / L
    psi = streamfunction(N, rng, kmax)
    up = np.stack([_central(psi, 1, dx), -_central(psi, 0, dx)])
    m = np.abs(up).max()
    if m > 0:
        up *= amplitude / m
    x = np.arange(N, dtype=np.float64) * dx
    sx, cx = np.sin(kk * x), np.cos(kk * x)
    tg = np.stack([np.outer(sx, cx), -np.outer(cx, sx)])
    return (tg + up).astype(dtype)


def theta_batch(n: int, seed: int = 0, dim: int = 3) -> np.ndarray:
    """theta ~ U[0.5, 1.5]^dim per trajectory."""
    return np.random.default_rng(seed + 7919).uniform(0.5, 1.5, size=(n, dim)).astype(np.float32)


def nu_batch(n: int, seed: int = 0, lo: float = 1e-4, hi: float = 1e-2) -> np.ndarray:
    """Log-uniform viscosities."""
    r = np.random.default_rng(seed + 104729).uniform(math.log(lo), math.log(hi), size=n)
    return np.exp(r).astype(np.float32)


def stable_dt(dx: float, umax0: float, nu: float | None, cfl: float = 0.3, steps: int = 0,
              C: float = 1.0, stable: bool = True) -> float:
    """Largest dt with cfl*dx/u >= dt, dx^2/(4 nu) >= dt and (stable) 0.9 * 2 nu / u^2 >= dt,
    where u = umax0 + C*steps*dt bounds the speed over the whole run: the Gaussian forcing
    (amplitude C, NS_solver.py:621-640) can add at most C per unit time.  The last condition
    is the von Neumann limit of explicit Euler with central advection ((u dt/dx)^2 <= 2 nu
    dt/dx^2); without it a long central rollout at small nu grows until it overflows."""
    def ok(dt):
        u = umax0 + C * steps * dt
        if dt > cfl * dx / u:
            return False
        if nu is not None:
            if dt > dx * dx / (4.0 * nu):
                return False
            if stable and dt > 0.9 * 2.0 * nu / (u * u):
                return False
        return True
    hi = cfl * dx / umax0
    if ok(hi):
        return hi
    lo = 0.0
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if ok(mid) else (lo, mid)
    return lo


def cfl_dt(u0: np.ndarray, L: float, cfl: float = 0.3, nu: float | None = None,
           stable: bool = False, steps: int = 0, C: float = 1.0) -> float:
    """dt = cfl*dx/max|u0|, also <= dx^2/(4 nu) when nu is given.  With
    stable=True also dt <= 0.9 * 2 nu / max|u|^2 (explicit Euler with central
    advection amplifies every mode unless the diffusion keeps up).  With
    steps > 0 the speed bound is max|u0| + C*steps*dt (forcing amplitude C),
    so the conditions hold for the whole rollout (stable_dt)."""
    N = u0.shape[-1]
    dx = L / N
    umax = float(np.abs(u0).max())
    if steps > 0:
        return float(np.float32(stable_dt(dx, umax, nu, cfl, steps, C, stable)))
    dt = cfl * dx / umax
    if nu is not None:
        dt = min(dt, dx * dx / (4.0 * nu))
        if stable:
            dt = min(dt, 0.9 * 2.0 * nu / (umax * umax))
    return float(np.float32(dt))
