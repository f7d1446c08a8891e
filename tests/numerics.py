"""Shared input generators and error measures for the tests and the benchmark."""

from __future__ import annotations

import numpy as np

EPS64 = np.finfo(np.float64).eps
EPS32 = float(np.finfo(np.float32).eps)


def spd(rng, batch, n, dtype=np.float64):
    """A = G G^T + n I with G ~ N(0,1) (the reference's makePositiveDefinite
    construction, algorithm/MakePositiveDefinite.hpp:33-40)."""
    g = rng.standard_normal((batch, n, n))
    return (g @ g.transpose(0, 2, 1) + n * np.eye(n)).astype(dtype)


def dyadic(rng, shape, bits=8):
    """Entries k / 2^bits in [-1, 1]: products and sums of up to 64 terms are
    exact in float64, so results do not depend on summation order."""
    return rng.integers(-(2 ** bits), 2 ** bits + 1, size=shape).astype(np.float64) / 2 ** bits


def fro(x, axes=(-2, -1)):
    return np.sqrt(np.sum(np.square(x), axis=axes))


def potrf_backward_error(a, l):
    """||L L^T - A||_F / ||A||_F per instance (lower triangle of L used)."""
    l = np.tril(l)
    return fro(l @ np.swapaxes(l, -1, -2) - a) / fro(a)


def solve_backward_error(a, x, b):
    """Normwise backward error ||b - A x|| / (||A|| ||x|| + ||b||) per instance."""
    r = b - a @ x
    return fro(r) / (fro(a) * fro(x) + fro(b))


def gemm_backward_error(alpha, a, b, beta, c, d):
    k = a.shape[-1]
    return fro(d - alpha * (a @ b) - beta * c) / (max(k, 1) * fro(a) * fro(b) + fro(c) + 1e-300)


def lu_backward_error(a, lu, ipiv):
    """||P A - L U|| / ||A|| for one instance, ipiv in LAPACK sequential form."""
    m, n = a.shape
    k = min(m, n)
    l = np.tril(lu[:, :k], -1) + np.eye(m, k)
    u = np.triu(lu[:k, :])
    pa = a.copy()
    for i in range(k):
        p = int(ipiv[i])
        if p != i:
            pa[[i, p]] = pa[[p, i]]
    return fro(pa - l @ u) / fro(a)


def haar(rng, batch, n):
    """Haar-distributed orthogonal matrices: QR of a Gaussian matrix with the
    signs of R's diagonal folded into Q (Mezzadri 2007)."""
    q, r = np.linalg.qr(rng.standard_normal((batch, n, n)))
    return q * np.sign(np.diagonal(r, axis1=-2, axis2=-1))[:, None, :]


def conditioned(rng, batch, n, kappa, spd=True):
    """A = U diag(s) V^T (V = U when spd) with s log-uniform in [1/kappa, 1]
    and both ends present, so cond_2(A) = kappa exactly (kappa: scalar or one
    value per instance)."""
    kappa = np.broadcast_to(np.asarray(kappa, dtype=np.float64), (batch,))
    u = haar(rng, batch, n)
    v = u if spd else haar(rng, batch, n)
    t = rng.uniform(0.0, 1.0, (batch, n))
    t[:, 0], t[:, -1] = 0.0, 1.0
    s = np.exp(-t * np.log(kappa)[:, None])
    a = (u * s[:, None, :]) @ np.swapaxes(v, -1, -2)
    if spd:
        a = 0.5 * (a + np.swapaxes(a, -1, -2))
    return a
