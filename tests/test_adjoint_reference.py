"""Pins the FP64 reference adjoints (oracle/adjoint.py) with central finite
differences: <grad, d> == (phi(x + h d) - phi(x - h d)) / 2h for a random
direction d, phi = <cotangent, f(x)>.  CPU only."""

import numpy as np
import pytest

from oracle import adjoint
from tests.numerics import spd

H = 1e-6


def _fd(phi, x, d):
    return (phi(x + H * d) - phi(x - H * d)) / (2 * H)


def _close(a, b, rel=1e-6):
    assert abs(a - b) <= rel * max(abs(a), abs(b), 1e-12), (a, b)


@pytest.mark.parametrize("n", [1, 5, 17])
def test_potrf_backward_fd(n):
    rng = np.random.default_rng(n)
    a = np.tril(spd(rng, 1, n)[0])
    lb = np.tril(rng.standard_normal((n, n)))
    d = np.tril(rng.standard_normal((n, n)))

    def phi(al):
        return np.sum(lb * np.linalg.cholesky(adjoint.sym_from_lower(al)))

    _close(np.sum(adjoint.potrf_backward(np.linalg.cholesky(adjoint.sym_from_lower(a)), lb) * d), _fd(phi, a, d))


@pytest.mark.parametrize("n,r", [(1, 1), (6, 3), (20, 4)])
def test_posv_backward_fd(n, r):
    rng = np.random.default_rng(n + r)
    a = np.tril(spd(rng, 1, n)[0])
    b = rng.standard_normal((n, r))
    xb = rng.standard_normal((n, r))
    x = np.linalg.solve(adjoint.sym_from_lower(a), b)
    ab, bb = adjoint.posv_backward(a, x, xb)
    da, db = np.tril(rng.standard_normal((n, n))), rng.standard_normal((n, r))
    _close(np.sum(ab * da), _fd(lambda al: np.sum(xb * np.linalg.solve(adjoint.sym_from_lower(al), b)), a, da))
    _close(np.sum(bb * db), _fd(lambda bv: np.sum(xb * np.linalg.solve(adjoint.sym_from_lower(a), bv)), b, db))


@pytest.mark.parametrize("upper", [0, 1])
def test_trsm_backward_fd(upper):
    rng = np.random.default_rng(3 + upper)
    n, r = 9, 3
    a = rng.uniform(-1, 1, (n, n)) + 4 * np.eye(n)
    a = np.triu(a) if upper else np.tril(a)
    b, xb = rng.standard_normal((n, r)), rng.standard_normal((n, r))
    x = np.linalg.solve(a, b)
    ab, bb = adjoint.trsm_backward(upper, a, x, xb)
    da = rng.standard_normal((n, n))
    da = np.triu(da) if upper else np.tril(da)
    _close(np.sum(ab * da), _fd(lambda av: np.sum(xb * np.linalg.solve(av, b)), a, da))
    db = rng.standard_normal((n, r))
    _close(np.sum(bb * db), _fd(lambda bv: np.sum(xb * np.linalg.solve(a, bv)), b, db))


def test_getrs_and_gemm_backward_fd():
    rng = np.random.default_rng(5)
    n, r = 8, 2
    a = rng.uniform(-1, 1, (n, n)) + 3 * np.eye(n)
    b, xb = rng.standard_normal((n, r)), rng.standard_normal((n, r))
    x = np.linalg.solve(a, b)
    ab, bb = adjoint.getrs_backward(a, x, xb)
    da = rng.standard_normal((n, n))
    _close(np.sum(ab * da), _fd(lambda av: np.sum(xb * np.linalg.solve(av, b)), a, da))
    m, k, nn = 4, 5, 3
    A, B, C, Db = (rng.standard_normal(s) for s in ((m, k), (k, nn), (m, nn), (m, nn)))
    gA, gB, gC = adjoint.gemm_backward(1.5, A, B, -0.5, Db)
    for g, x0, f in ((gA, A, lambda v: 1.5 * v @ B - 0.5 * C), (gB, B, lambda v: 1.5 * A @ v - 0.5 * C),
                     (gC, C, lambda v: 1.5 * A @ B - 0.5 * v)):
        d = rng.standard_normal(x0.shape)
        _close(np.sum(g * d), _fd(lambda v: np.sum(Db * f(v)), x0, d))
