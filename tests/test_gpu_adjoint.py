"""Adjoint kernels on the GPU vs the FP64 reference adjoints (oracle/adjoint.py,
pinned by finite differences) and vs central finite differences of the GPU
forward solve itself (h = 1e-6)."""

import numpy as np
import pytest

from blast_amd import capi
from blast_amd.layout import Layout
from oracle import adjoint
from tests.numerics import spd

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-10, np.float32: 2e-3}


def dev(layout, host):
    return capi.Operand.from_host(layout, host)


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(autouse=True)
def auto_algorithm():
    capi.set_algorithm(capi.ALGO_AUTO)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,r", [(32, 1), (96, 4), (192, 2)])
@pytest.mark.parametrize("recompute", [0, 1])
def test_posv_backward(dtype, n, r, recompute):
    import torch

    rng = np.random.default_rng(n + r + recompute)
    batch = 3
    a = spd(rng, batch, n)
    b, xb = rng.standard_normal((batch, n, r)), rng.standard_normal((batch, n, r))
    la, lb = Layout.make(n, n, dtype), Layout.make(n, r, dtype)
    A, B = dev(la, la.pack(a)), dev(lb, lb.pack(b))
    L, X = capi.Operand.empty(la, batch), capi.Operand.empty(lb, batch)
    if dtype == np.float64:
        capi.posv(A, L, B, X, batch)  # forward keeps L
    else:
        capi.potrf(A, L, batch)
        capi.trsm(0, 0, 0, 1.0, L, B, X, batch)
        capi.trsm(0, 1, 0, 1.0, L.t(), X, X, batch)
    Ab, Bb = capi.Operand.empty(la, batch, fill=np.nan), capi.Operand.empty(lb, batch, fill=np.nan)
    info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    capi.posv_backward(recompute, A if recompute else L, X, dev(lb, lb.pack(xb)), Ab, Bb, batch, info)
    x = X.logical(batch).astype(np.float64)
    for k in range(batch):
        ab_ref, bb_ref = adjoint.posv_backward(np.tril(a[k]), x[k], xb[k])
        assert _rel(Ab.logical(batch)[k], ab_ref) < TOL[dtype] * n
        assert _rel(Bb.logical(batch)[k], bb_ref) < TOL[dtype] * n


def test_posv_backward_matches_gpu_finite_differences():
    rng = np.random.default_rng(11)
    n, r, batch, h = 64, 3, 2, 1e-6
    a = spd(rng, batch, n)
    b, xb = rng.standard_normal((batch, n, r)), rng.standard_normal((batch, n, r))
    da = np.tril(rng.standard_normal((batch, n, n)))
    la, lb = Layout.make(n, n), Layout.make(n, r)

    def solve(amat):
        A, B = dev(la, la.pack(amat)), dev(lb, lb.pack(b))
        L, X = capi.Operand.empty(la, batch), capi.Operand.empty(lb, batch)
        capi.posv(A, L, B, X, batch)
        return X.logical(batch)

    x = solve(a)
    Ab, Bb = capi.Operand.empty(la, batch), capi.Operand.empty(lb, batch)
    capi.posv_backward(1, dev(la, la.pack(a)), dev(lb, lb.pack(x)), dev(lb, lb.pack(xb)), Ab, Bb, batch)
    sym = lambda m: np.tril(m) + np.swapaxes(np.tril(m, -1), -1, -2)  # noqa: E731
    fd = (np.sum(xb * solve(sym(np.tril(a) + h * da)), axis=(1, 2))
          - np.sum(xb * solve(sym(np.tril(a) - h * da)), axis=(1, 2))) / (2 * h)
    got = np.sum(Ab.logical(batch) * da, axis=(1, 2))
    np.testing.assert_allclose(got, fd, rtol=1e-6)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n", [32, 80, 192])
def test_potrf_backward(dtype, n):
    rng = np.random.default_rng(n)
    batch = 3
    a = spd(rng, batch, n)
    la = Layout.make(n, n, dtype)
    L = capi.Operand.empty(la, batch)
    capi.potrf(dev(la, la.pack(a)), L, batch)
    lbar = np.tril(rng.standard_normal((batch, n, n)))
    lbar_host = la.pack(lbar + np.triu(rng.standard_normal((batch, n, n)), 1))  # upper part must be ignored
    Ab = capi.Operand.empty(la, batch, fill=np.nan)
    capi.potrf_backward(L, dev(la, lbar_host), Ab, batch)
    lg = L.logical(batch).astype(np.float64)
    for k in range(batch):
        ref = adjoint.potrf_backward(lg[k], lbar[k])
        assert _rel(Ab.logical(batch)[k], ref) < TOL[dtype] * n


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("upper", [0, 1])
def test_trsm_backward(dtype, upper):
    rng = np.random.default_rng(5 + upper)
    n, r, batch = 40, 3, 2
    a = rng.uniform(-1, 1, (batch, n, n)) + 8 * np.eye(n)
    a = np.triu(a) if upper else np.tril(a)
    x, xb = rng.standard_normal((batch, n, r)), rng.standard_normal((batch, n, r))
    la, lb = Layout.make(n, n, dtype), Layout.make(n, r, dtype)
    Ab, Bb = capi.Operand.empty(la, batch, fill=np.nan), capi.Operand.empty(lb, batch, fill=np.nan)
    capi.trsm_backward(upper, dev(la, la.pack(a)), dev(lb, lb.pack(x)), dev(lb, lb.pack(xb)), Ab, Bb, batch)
    for k in range(batch):
        ab_ref, bb_ref = adjoint.trsm_backward(upper, a[k], x[k], xb[k])
        assert _rel(Ab.logical(batch)[k], ab_ref) < TOL[dtype] * n
        assert _rel(Bb.logical(batch)[k], bb_ref) < TOL[dtype] * n


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_getrs_forward_and_backward(dtype):
    import torch

    rng = np.random.default_rng(9)
    n, r, batch = 48, 3, 4
    a = rng.uniform(-1, 1, (batch, n, n)) + 2 * np.eye(n)
    b, xb = rng.standard_normal((batch, n, r)), rng.standard_normal((batch, n, r))
    la, lb = Layout.make(n, n, dtype, panel=False), Layout.make(n, r, dtype)
    LU = dev(la, la.pack(a))
    ipiv = torch.zeros(batch * n, dtype=torch.int64, device="cuda")
    capi.getrf(LU, ipiv, n, batch)
    X = capi.Operand.empty(lb, batch)
    capi.getrs(LU, ipiv, n, dev(lb, lb.pack(b)), X, batch)
    x = X.logical(batch).astype(np.float64)
    assert _rel(x, np.linalg.solve(a, b)) < TOL[dtype] * n
    Ab, Bb = capi.Operand.empty(la, batch, fill=np.nan), capi.Operand.empty(lb, batch, fill=np.nan)
    capi.getrs_backward(LU, ipiv, n, X, dev(lb, lb.pack(xb)), Ab, Bb, batch)
    for k in range(batch):
        ab_ref, bb_ref = adjoint.getrs_backward(a[k], x[k], xb[k])
        assert _rel(Ab.logical(batch)[k], ab_ref) < TOL[dtype] * n
        assert _rel(Bb.logical(batch)[k], bb_ref) < TOL[dtype] * n


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_gemm_backward(dtype):
    rng = np.random.default_rng(12)
    m, k, n, batch = 33, 20, 17, 3
    A, B, Db = (rng.standard_normal((batch,) + s) for s in ((m, k), (k, n), (m, n)))
    la, lb, ld = Layout.make(m, k, dtype), Layout.make(k, n, dtype), Layout.make(m, n, dtype)
    Ab, Bb, Cb = (capi.Operand.empty(l, batch, fill=np.nan) for l in (la, lb, ld))
    capi.gemm_backward(1.5, dev(la, la.pack(A)), dev(lb, lb.pack(B)), -0.5, dev(ld, ld.pack(Db)), Ab, Bb, Cb, batch)
    gA, gB, gC = adjoint.gemm_backward(1.5, A, B, -0.5, Db)
    assert _rel(Ab.logical(batch), gA) < TOL[dtype] * k
    assert _rel(Bb.logical(batch), gB) < TOL[dtype] * m
    assert _rel(Cb.logical(batch), gC) < TOL[dtype]
