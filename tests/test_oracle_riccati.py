"""Pins the FP64 Riccati recursion of oracle/riccati.py against facts that do
not restate it: the scalar closed form, the infinite-horizon limit (scipy's
DARE solver) and the LQR cost identity."""

import numpy as np
import pytest
import scipy.linalg

from oracle import riccati


def test_scalar_closed_form():
    a, b, q, r = 1.3, 0.7, 2.0, 0.5
    K, P0 = riccati.riccati(np.array([[a]]), np.array([[b]]), np.array([[q]]), np.array([[r]]), 6)
    p = q
    for k in range(5, -1, -1):
        kk = -a * b * p / (r + b * b * p)
        assert K[k, 0, 0] == pytest.approx(kk, rel=1e-14)
        p = q + a * a * p * r / (r + b * b * p)
    assert P0[0, 0] == pytest.approx(p, rel=1e-14)


def test_long_horizon_matches_dare():
    rng = np.random.default_rng(3)
    nx, nu = 7, 3
    A = rng.normal(0, 1 / np.sqrt(nx), (nx, nx))
    B = rng.normal(0, 1 / np.sqrt(nx), (nx, nu))
    W = rng.standard_normal((nx, nx))
    Q, R = np.eye(nx) + 0.1 * W @ W.T, np.eye(nu)
    K, P0 = riccati.riccati(A, B, Q, R, 400)
    P = scipy.linalg.solve_discrete_are(A, B, Q, R)
    np.testing.assert_allclose(P0, P, rtol=1e-10)
    np.testing.assert_allclose(K[0], -np.linalg.solve(R + B.T @ P @ B, B.T @ P @ A), rtol=1e-9, atol=1e-12)


def test_cost_identity():
    # x0^T P_0 x0 = sum_k (x_k^T Q x_k + u_k^T R u_k) + x_N^T Q x_N along the
    # closed loop u_k = K_k x_k (terminal cost P_N = Q).
    rng = np.random.default_rng(4)
    nx, nu, N = 9, 4, 10
    A = rng.normal(0, 1 / np.sqrt(nx), (nx, nx))
    B = rng.normal(0, 1 / np.sqrt(nx), (nx, nu))
    W = rng.standard_normal((nx, nx))
    Q, R = np.eye(nx) + 0.1 * W @ W.T, np.eye(nu) + 0.2 * np.diag(rng.uniform(size=nu))
    K, P0 = riccati.riccati(A, B, Q, R, N)
    x0 = rng.standard_normal(nx)
    x, cost = x0.copy(), 0.0
    for k in range(N):
        u = K[k] @ x
        cost += x @ Q @ x + u @ R @ u
        x = A @ x + B @ u
    cost += x @ Q @ x
    assert x0 @ P0 @ x0 == pytest.approx(cost, rel=1e-12)
    # and the gains are optimal: perturbing u_0 raises the cost
    for _ in range(3):
        d = 1e-3 * rng.standard_normal(nu)
        x, c2 = x0.copy(), 0.0
        for k in range(N):
            u = K[k] @ x + (d if k == 0 else 0.0)
            c2 += x @ Q @ x + u @ R @ u
            x = A @ x + B @ u
        c2 += x @ Q @ x
        assert c2 > cost


def test_generator_ranges():
    rng = np.random.default_rng(5)
    nx, nu, A, B, Q, R = riccati.make_instances(rng, 200)
    assert nx.min() >= 4 and nx.max() <= 200
    assert np.all(nu >= 1) and np.all(nu <= np.maximum(1, nx // 2))
    for i in range(5):
        assert A[i].shape == (nx[i], nx[i]) and B[i].shape == (nx[i], nu[i])
        assert np.all(np.linalg.eigvalsh(Q[i]) >= 1.0 - 1e-12)
