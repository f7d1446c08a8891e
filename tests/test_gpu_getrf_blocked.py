"""AUTO getrf / dgesv for 256 < n <= 1024: the blocked LU (32-column panels
factored in registers with partial pivoting, the interchanges and U12 in the
same launch, the trailing update on the gemm tile kernel) against the
oracle.  FP64: pivots equal BLAST's on well-conditioned inputs and the factor
agrees within rounding; info of a zero pivot in the first and in a later
panel; a row chosen as pivot twice within one panel; the anti-diagonal
closed form (every interchange reaches beyond the panel).  FP32: backward
error and a valid pivot sequence (BLAST's FP32 blocking rounds differently,
so near-ties may pick other rows)."""

import numpy as np
import pytest

from blast_amd import capi
from blast_amd.layout import Layout
from oracle import oracle
from tests.numerics import EPS32, EPS64, lu_backward_error, solve_backward_error

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def auto_algorithm():
    capi.set_algorithm(capi.ALGO_AUTO)


def dev(layout, host):
    return capi.Operand.from_host(layout, host)


def i32(batch, fill):
    import torch

    return torch.full((batch,), fill, dtype=torch.int32, device="cuda")


def run_getrf(la, ah, batch):
    import torch

    n = la.rows
    A = dev(la, ah)
    ipiv = torch.full((batch * n,), -1, dtype=torch.int64, device="cuda")
    info = i32(batch, -1)
    capi.getrf(A, ipiv, n, batch, info)
    return A.logical(batch), ipiv.cpu().numpy().reshape(batch, n), info.cpu().numpy()


@pytest.mark.parametrize("n,panel", [(257, False), (257, True), (300, True), (512, False), (700, True), (1024, False)])
def test_getrf_blocked_fp64_vs_oracle(n, panel):
    rng = np.random.default_rng(n)
    batch = 2
    a = rng.uniform(-1, 1, (batch, n, n)) + 0.5 * np.eye(n)
    la = Layout.make(n, n, panel=panel)
    ah = la.pack(a)
    ref = ah.copy()
    piv_ref, inf_ref = oracle.getrf(la, ref, batch)
    lu, piv, inf = run_getrf(la, ah, batch)
    assert inf.tolist() == [0] * batch
    np.testing.assert_array_equal(inf, inf_ref)
    np.testing.assert_array_equal(piv, piv_ref)  # well-conditioned: the same pivots
    for k in range(batch):
        assert lu_backward_error(a[k], lu[k], piv[k]) < 4 * n * EPS64
    lu_ref = la.unpack(ref, batch)
    np.testing.assert_allclose(lu, lu_ref, rtol=0, atol=16 * n * EPS64 * np.abs(lu_ref).max())


def test_getrf_blocked_zero_pivot_info():
    n, batch = 320, 3
    rng = np.random.default_rng(7)
    a = rng.uniform(-1, 1, (batch, n, n)) + 0.5 * np.eye(n)
    a[1][:, 37] = 0.0   # zero column: every update keeps it zero, pivot 0 at column 38 (first panel)
    a[2][:, 200] = 0.0  # column 201, the seventh panel
    la = Layout.make(n, n, panel=False)
    ah = la.pack(a)
    piv_ref, inf_ref = oracle.getrf(la, ah.copy(), batch)
    lu, piv, inf = run_getrf(la, ah, batch)
    assert inf.tolist() == [0, 38, 201]
    np.testing.assert_array_equal(inf, inf_ref)
    np.testing.assert_array_equal(piv[0], piv_ref[0])
    assert lu_backward_error(a[0], lu[0], piv[0]) < 4 * n * EPS64
    # the pivots chosen before the zero column of instance 2 (earlier panels)
    np.testing.assert_array_equal(piv[2][:192], piv_ref[2][:192])


@pytest.mark.parametrize("panel", [False, True])
def test_getrf_blocked_anti_diagonal(panel):
    n = 300
    a = np.fliplr(np.eye(n))[None]
    la = Layout.make(n, n, panel=panel)
    lu, piv, inf = run_getrf(la, la.pack(a), 1)
    k = np.arange(n)
    np.testing.assert_array_equal(piv[0], np.where(k < n // 2, n - 1 - k, k))
    np.testing.assert_array_equal(lu[0], np.eye(n))  # L = U = I
    assert inf.tolist() == [0]


def test_getrf_blocked_same_pivot_row_twice_in_a_panel():
    # column 0's pivot is row 280; the row swapped into 280 then holds column
    # 1's largest entry, so row 280 is chosen again (the row map must follow
    # the content, not the row index)
    n = 288
    rng = np.random.default_rng(3)
    a = 0.1 * rng.uniform(-1, 1, (1, n, n)) + 0.5 * np.eye(n)
    a[0][280, 0] = 10.0
    a[0][0, 1] = 10.0
    la = Layout.make(n, n, panel=False)
    ah = la.pack(a)
    ref = ah.copy()
    piv_ref, _ = oracle.getrf(la, ref, 1)
    lu, piv, inf = run_getrf(la, ah, 1)
    assert piv[0][:2].tolist() == [280, 280]
    np.testing.assert_array_equal(piv, piv_ref)
    assert inf.tolist() == [0]
    assert lu_backward_error(a[0], lu[0], piv[0]) < 4 * n * EPS64
    np.testing.assert_allclose(lu, la.unpack(ref, 1), rtol=0, atol=1e-13 * np.abs(lu).max())


@pytest.mark.parametrize("n,panel", [(384, False), (300, True), (520, True)])
def test_getrf_blocked_fp32(n, panel):
    rng = np.random.default_rng(n + 1)
    batch = 2
    a = (rng.uniform(-1, 1, (batch, n, n)) + 0.5 * np.eye(n)).astype(np.float32)
    la = Layout.make(n, n, np.float32, panel=panel)
    lu, piv, inf = run_getrf(la, la.pack(a), batch)
    assert inf.tolist() == [0] * batch
    k = np.arange(n)
    assert np.all(piv >= k) and np.all(piv < n)  # sequential interchanges: row k with a row >= k
    for b in range(batch):
        assert lu_backward_error(a[b].astype(np.float64), lu[b].astype(np.float64), piv[b]) < 4 * n * EPS32


def test_gesv_blocked_vs_oracle():
    import torch

    n, nrhs, batch = 400, 3, 2
    rng = np.random.default_rng(11)
    a = rng.uniform(-1, 1, (batch, n, n)) + 0.5 * np.eye(n)
    b = rng.uniform(-1, 1, (batch, n, nrhs))
    la, lb = Layout.make(n, n, panel=False), Layout.make(n, nrhs)
    ah = la.pack(a)
    piv_ref, _ = oracle.getrf(la, ah.copy(), batch)
    A, B, X = dev(la, ah), dev(lb, lb.pack(b)), capi.Operand.empty(lb, batch)
    ipiv = torch.zeros(batch * n, dtype=torch.int64, device="cuda")
    info = i32(batch, -1)
    capi.gesv(A, ipiv, n, B, X, batch, info)
    assert info.cpu().numpy().tolist() == [0] * batch
    np.testing.assert_array_equal(ipiv.cpu().numpy().reshape(batch, n), piv_ref)
    x = X.logical(batch)
    assert np.all(solve_backward_error(a, x, b) < 4 * n * EPS64)
    np.testing.assert_allclose(x, np.linalg.solve(a, b), rtol=0, atol=1e-10 * np.abs(x).max())
