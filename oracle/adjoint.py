"""FP64 CPU reference adjoints (numpy).  TEST INFRASTRUCTURE ONLY.

Closed forms (Giles 2008, "An extended collection of matrix derivative
results"; Murray 2016, "Differentiation of the Cholesky decomposition") for
the reverse-mode derivatives the HIP library implements.  BLAST has no
adjoints; these functions are pinned by central finite differences in
tests/test_adjoint_reference.py.

Symmetric inputs are parameterised by their lower triangle: a stored entry
a_ij (i > j) sets both A_ij and A_ji, so a full gradient G folds to
G_ij + G_ji below the diagonal and G_ii on it (upper part 0).
"""

import numpy as np


def fold_lower(g):
    """Full gradient -> gradient w.r.t. the stored lower entries."""
    return np.tril(g + np.swapaxes(g, -1, -2), -1) + np.eye(g.shape[-1]) * g


def sym_from_lower(a):
    """Symmetric matrix whose lower triangle is a's."""
    return np.tril(a) + np.swapaxes(np.tril(a, -1), -1, -2)


def gemm_backward(alpha, a, b, beta, db):
    """D = alpha A B + beta C."""
    return alpha * db @ np.swapaxes(b, -1, -2), alpha * np.swapaxes(a, -1, -2) @ db, beta * db


def trsm_backward(upper, a, x, xb):
    """A X = B, A triangular (lower or upper), non-unit."""
    at = np.triu(a) if upper else np.tril(a)
    bb = np.linalg.solve(np.swapaxes(at, -1, -2), xb)
    ab = -bb @ np.swapaxes(x, -1, -2)
    return (np.triu(ab) if upper else np.tril(ab)), bb


def potrf_backward(l, lb):
    """L = chol(A), A read from its lower triangle."""
    l = np.tril(l)
    p = np.tril(np.swapaxes(l, -1, -2) @ np.tril(lb))
    p = p - 0.5 * np.eye(l.shape[-1]) * p
    linv = np.linalg.inv(l)
    g = np.swapaxes(linv, -1, -2) @ p @ linv
    return fold_lower(g)


def posv_backward(a_lower, x, xb):
    """X = A^{-1} B, A SPD read from its lower triangle."""
    a = sym_from_lower(a_lower)
    bb = np.linalg.solve(a, xb)
    return fold_lower(-bb @ np.swapaxes(x, -1, -2)), bb


def getrs_backward(a, x, xb):
    """X = A^{-1} B, general A."""
    bb = np.linalg.solve(np.swapaxes(a, -1, -2), xb)
    return -bb @ np.swapaxes(x, -1, -2), bb
