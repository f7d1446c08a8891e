"""Test infrastructure: FP64 CPU restatement of the finite-horizon discrete
Riccati recursion (LQR / MPC backward pass) that the GPU composite computes.

The reference has no Riccati routine; its syrk+potrf benchmarks
(bench/blas/SyrkPotrf.cpp:15-57) and BLASFEO wrapper
(include/blasfeo/blasfeo_api/SyrkPotrf.hpp:14-39) are the kernel sequence
of one square-root step.  This module states the textbook recursion

    P_N = Q
    K_k = -(R + B^T P B)^{-1} B^T P A
    P   = Q + A^T P A + A^T P B K_k          k = N-1 .. 0

and the instance generator of the uneven-batch workload (20000 instances,
nx uniform in 4..200, nu uniform in 1..nx/2, horizon 10, A and B entries
N(0, 1/nx), Q = I + 0.1 W W^T, R = I).

Pinning (tests/test_oracle_riccati.py): the scalar case against its closed
form, the long-horizon limit against scipy's discrete algebraic Riccati
solver, and the cost identity x0^T P_0 x0 = optimal cost of the LQR problem
obtained by simulating the closed loop with the gains K_k.
"""

from __future__ import annotations

import numpy as np


def riccati(A, B, Q, R, horizon):
    """Returns (K, P0): K[k] (nu x nx) for k = 0..horizon-1 and P_0."""
    P = np.array(Q, dtype=np.float64)
    K = [None] * horizon
    for k in range(horizon - 1, -1, -1):
        S = R + B.T @ P @ B
        H = B.T @ P @ A
        Kk = -np.linalg.solve(S, H)
        P = Q + A.T @ P @ A + H.T @ Kk
        P = 0.5 * (P + P.T)
        K[k] = Kk
    return np.stack(K), P


def make_instances(rng, count, nx_range=(4, 200)):
    """Sizes and matrices of the uneven workload (per-instance lists)."""
    nx = rng.integers(nx_range[0], nx_range[1] + 1, size=count).astype(np.int32)
    nu = np.array([rng.integers(1, max(1, n // 2) + 1) for n in nx], dtype=np.int32)
    A, B, Q, R = [], [], [], []
    for n, m in zip(nx, nu):
        s = 1.0 / np.sqrt(n)
        A.append(rng.normal(0.0, s, (n, n)))
        B.append(rng.normal(0.0, s, (n, m)))
        W = rng.standard_normal((n, n))
        Q.append(np.eye(n) + 0.1 * W @ W.T)
        R.append(np.eye(m))
    return nx, nu, A, B, Q, R


def step_flops(nx, nu):
    """FP64 flops of one square-root step as the GPU composite performs it:
    M = L^T [B A] (triangular, nx^2 (nx+nu)), Lam += M^T M (lower,
    (nx+nu)^2 nx), chol(Lam) ((nx+nu)^3 / 3), K = -Ls^{-T} Lks^T (nu^2 nx)."""
    nt = nx + nu
    return nx * nx * nt + nt * nt * nx + nt ** 3 / 3.0 + nu * nu * nx
