"""Writes tests/golden/lapack_small.npz: known answers from LAPACK (via scipy).

Run in the development container (scipy's bundled OpenBLAS LAPACK):
    python tests/golden/gen_golden.py
The fixture holds inputs and LAPACK outputs only (no code): dpotrf lower
factors, dgetrf pivots (0-based) and factors, dtrsm solutions and a dgemm
product, for sizes up to 16.  The oracle and the HIP kernels are checked
against it to tolerance (LAPACK's blocking orders operations differently), and
bit-exactly for getrf pivots on these well-separated random inputs.
"""

import os

import numpy as np
import scipy
import scipy.linalg as sl
from scipy.linalg import lapack

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lapack_small.npz")


def main():
    rng = np.random.default_rng(20240611)
    data = {"scipy_version": np.array(scipy.__version__)}
    for n in (1, 2, 3, 5, 8, 13, 16):
        g = rng.standard_normal((n, n))
        a = g @ g.T + n * np.eye(n)
        data[f"potrf_a_{n}"] = a
        data[f"potrf_l_{n}"] = sl.cholesky(a, lower=True)
    for m, n in ((5, 5), (8, 8), (16, 16), (12, 7), (7, 12), (13, 13)):
        a = rng.uniform(-1.0, 1.0, (m, n))
        lu, piv, info = lapack.dgetrf(a)
        assert info == 0
        data[f"getrf_a_{m}x{n}"] = a
        data[f"getrf_lu_{m}x{n}"] = lu
        data[f"getrf_piv_{m}x{n}"] = piv.astype(np.int64)
    for n, r in ((4, 3), (11, 5), (16, 2)):
        l = np.tril(rng.uniform(-1.0, 1.0, (n, n))) + 2.0 * np.eye(n)
        b = rng.uniform(-1.0, 1.0, (n, r))
        data[f"trsm_l_{n}x{r}"] = l
        data[f"trsm_b_{n}x{r}"] = b
        data[f"trsm_xlower_{n}x{r}"] = sl.solve_triangular(l, b, lower=True)
        data[f"trsm_xupper_{n}x{r}"] = sl.solve_triangular(l.T, b, lower=False)
        data[f"trsm_xunit_{n}x{r}"] = sl.solve_triangular(l, b, lower=True, unit_diagonal=True)
    a = rng.uniform(-1.0, 1.0, (5, 3))
    b = rng.uniform(-1.0, 1.0, (3, 7))
    c = rng.uniform(-1.0, 1.0, (5, 7))
    data["gemm_a"], data["gemm_b"], data["gemm_c"] = a, b, c
    data["gemm_d"] = 1.5 * (a @ b) - 0.5 * c
    np.savez_compressed(OUT, **data)
    print("wrote", OUT, len(data), "arrays")


if __name__ == "__main__":
    main()
