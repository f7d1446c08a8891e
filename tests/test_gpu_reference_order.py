"""Reference-order HIP kernels vs the oracle: bit-exact (same operation
sequence, IEEE division and square root).  Runs on the GPU, through the C ABI."""

import numpy as np
import pytest

from blast_amd import capi
from blast_amd.layout import Layout
from oracle import oracle
from tests.numerics import spd

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def reference_algorithm():
    capi.set_algorithm(capi.ALGO_REFERENCE)
    yield
    capi.set_algorithm(capi.ALGO_AUTO)


def dev(layout, host):
    return capi.Operand.from_host(layout, host)


def _lay(rows, cols, dtype, panel, trans=False):
    if trans:
        return Layout.make(cols, rows, dtype, panel=panel).t()
    return Layout.make(rows, cols, dtype, panel=panel)


DTYPES = [np.float64, np.float32]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (5, 7, 3), (20, 20, 20), (64, 64, 64), (33, 17, 9), (8, 3, 0)])
@pytest.mark.parametrize("ta,tb,panel", [(0, 0, 1), (0, 1, 1), (1, 0, 0), (1, 1, 0), (0, 1, 0)])
def test_gemm_bitexact(dtype, m, n, k, ta, tb, panel):
    rng = np.random.default_rng(m * 100 + n + k)
    batch = 3
    a = rng.uniform(-1, 1, (batch, m, k))
    b = rng.uniform(-1, 1, (batch, k, n))
    c = rng.uniform(-1, 1, (batch, m, n))
    la, lb = _lay(m, k, dtype, panel, ta), _lay(k, n, dtype, not panel, tb)
    lc, ld = _lay(m, n, dtype, panel), _lay(m, n, dtype, panel)
    ah, bh, ch = la.pack(a), lb.pack(b), lc.pack(c)
    dh = np.zeros(batch * ld.batch_stride, dtype)
    oracle.gemm(1.5, la, ah, lb, bh, -0.5, lc, ch, ld, dh, batch)
    D = capi.Operand.empty(ld, batch)
    capi.gemm(1.5, dev(la, ah), dev(lb, bh), -0.5, dev(lc, ch), D, batch)
    np.testing.assert_array_equal(D.logical(batch), ld.unpack(dh, batch))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("m,k", [(1, 1), (9, 4), (20, 20), (37, 11)])
def test_syrk_bitexact(dtype, m, k):
    rng = np.random.default_rng(m + k)
    batch = 2
    a, c = rng.uniform(-1, 1, (batch, m, k)), rng.uniform(-1, 1, (batch, m, m))
    la, lc = _lay(m, k, dtype, False), _lay(m, m, dtype, False)
    ah, ch = la.pack(a), lc.pack(c)
    dh = np.full(batch * lc.batch_stride, 3.0, dtype)
    oracle.syrk_lower(0.5, la, ah, -2.0, lc, ch, lc, dh, batch)
    D = capi.Operand.empty(lc, batch, fill=3.0)
    capi.syrk_lower(0.5, dev(la, ah), -2.0, dev(lc, ch), D, batch)
    np.testing.assert_array_equal(D.host(), dh)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("side_right", [0, 1])
@pytest.mark.parametrize("m,n", [(1, 1), (7, 5), (20, 13)])
def test_trmm_bitexact(dtype, side_right, m, n):
    rng = np.random.default_rng(m * n + side_right)
    batch = 2
    na = n if side_right else m
    a, b = rng.uniform(-1, 1, (batch, na, na)), rng.uniform(-1, 1, (batch, m, n))
    la, lb, lc = _lay(na, na, dtype, True), _lay(m, n, dtype, True), _lay(m, n, dtype, True)
    ah, bh = la.pack(a), lb.pack(b)
    ch = np.zeros(batch * lc.batch_stride, dtype)
    if side_right:
        oracle.trmm_right_lower(0.75, lb, bh, la, ah, lc, ch, batch)
    else:
        oracle.trmm_left_upper(0.75, la, ah, lb, bh, lc, ch, batch)
    C = capi.Operand.empty(lc, batch)
    capi.trmm(side_right, not side_right, 0.75, dev(la, ah), dev(lb, bh), C, batch)
    np.testing.assert_array_equal(C.logical(batch), lc.unpack(ch, batch))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("side_right,upper,unit", [(0, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (1, 0, 0)])
@pytest.mark.parametrize("m,n", [(1, 1), (4, 3), (50, 7), (128, 4)])
def test_trsm_bitexact(dtype, side_right, upper, unit, m, n):
    rng = np.random.default_rng(m + 7 * n)
    batch = 3
    na = n if side_right else m
    a = rng.uniform(-1, 1, (batch, na, na)) + 4 * np.eye(na)
    b = rng.uniform(-1, 1, (batch, m, n))
    la, lb = _lay(na, na, dtype, True), _lay(m, n, dtype, False)
    ah, bh = la.pack(a), lb.pack(b)
    side = "right" if side_right else "left"
    xh = bh.copy()
    oracle.trsm(side, not upper, unit, 2.0, la, ah, lb, xh, lb, xh, batch)  # X aliases B
    X = dev(lb, bh)
    capi.trsm(side_right, upper, unit, 2.0, dev(la, ah), X, X, batch)
    np.testing.assert_array_equal(X.host(), xh)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n", [0, 1, 2, 5, 13, 50, 128])
@pytest.mark.parametrize("panel", [True, False])
def test_potrf_bitexact(dtype, n, panel):
    rng = np.random.default_rng(n + 1000)
    batch = 4
    a = spd(rng, batch, n) if n else np.zeros((batch, 0, 0))
    la = _lay(n, n, dtype, panel)
    ah = la.pack(a)
    lh = np.full(batch * la.batch_stride, 5.0, dtype)
    info_ref = oracle.potrf(la, ah, la, lh, batch)
    L = capi.Operand.empty(la, batch, fill=5.0)
    import torch

    info = torch.full((batch,), -7, dtype=torch.int32, device="cuda")
    capi.potrf(dev(la, ah), L, batch, info)
    np.testing.assert_array_equal(L.host(), lh)
    np.testing.assert_array_equal(info.cpu().numpy(), info_ref)


def test_potrf_isolates_bad_instance():
    import torch

    rng = np.random.default_rng(5)
    n, batch = 20, 5
    a = spd(rng, batch, n)
    a[2, 7, 7] = -100.0  # not SPD
    la = _lay(n, n, np.float64, True)
    A = dev(la, la.pack(a))
    info = torch.zeros(batch, dtype=torch.int32, device="cuda")
    capi.potrf(A, A, batch, info)  # in place
    inf = info.cpu().numpy()
    assert inf[2] > 0 and np.all(inf[[0, 1, 3, 4]] == 0)
    got = A.logical(batch)
    assert np.all(np.isfinite(np.tril(got[[0, 1, 3, 4]])))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("m,n", [(0, 0), (1, 1), (5, 5), (16, 16), (12, 7), (7, 12), (64, 64), (130, 130)])
def test_getrf_bitexact(dtype, m, n):
    import torch

    rng = np.random.default_rng(m * 3 + n)
    batch = 3
    a = rng.uniform(-1, 1, (batch, m, n))
    la = _lay(m, n, dtype, False)
    ah = la.pack(a)
    ref = ah.copy()
    piv_ref, info_ref = oracle.getrf(la, ref, batch)
    A = dev(la, ah)
    k = max(min(m, n), 1)
    ipiv = torch.full((batch * k,), -1, dtype=torch.int64, device="cuda")
    info = torch.full((batch,), -7, dtype=torch.int32, device="cuda")
    capi.getrf(A, ipiv, k, batch, info)
    np.testing.assert_array_equal(A.host(), ref)
    np.testing.assert_array_equal(ipiv.cpu().numpy().reshape(batch, k)[:, : min(m, n)], piv_ref)
    np.testing.assert_array_equal(info.cpu().numpy(), info_ref)


def test_getrf_rowmajor_and_singular():
    import torch

    rng = np.random.default_rng(9)
    lr = Layout.make(9, 9, np.float64, panel=False).t()
    a = rng.uniform(-1, 1, (2, 9, 9)) + 9 * np.eye(9)
    a[1] = 1.0  # rank one: singular for the pivoted path, irrelevant here
    ah = lr.pack(a)
    ref = ah.copy()
    piv_ref, _ = oracle.getrf(lr, ref, 2)
    A = dev(lr, ah)
    ipiv = torch.zeros(18, dtype=torch.int64, device="cuda")
    capi.getrf(A, ipiv, 9, 2)
    np.testing.assert_array_equal(A.host(), ref)
    np.testing.assert_array_equal(ipiv.cpu().numpy().reshape(2, 9), piv_ref)
    # column-major singular instance is isolated
    lc = Layout.make(6, 6, np.float64, panel=False)
    a = rng.uniform(-1, 1, (3, 6, 6))
    a[1] = 1.0
    A = dev(lc, lc.pack(a))
    ipiv = torch.zeros(18, dtype=torch.int64, device="cuda")
    info = torch.zeros(3, dtype=torch.int32, device="cuda")
    capi.getrf(A, ipiv, 6, 3, info)
    np.testing.assert_array_equal(info.cpu().numpy(), [0, 2, 0])


@pytest.mark.parametrize("n", [20, 150])
def test_getrf_pivot_skips_nan_after_first(n):
    # BLAST's iamax scans from element 0 with a strict '>' (dense/Iamax.hpp:82-95):
    # a NaN after the first element is never the pivot.  Element 1 is a NaN
    # and element 2 the maximum, so the first pivot is row 2, as in the oracle.
    import torch

    rng = np.random.default_rng(n)
    a = rng.uniform(-1, 1, (1, n, n))
    a[0, 1, 0], a[0, 2, 0] = np.nan, 10.0
    la = Layout.make(n, n, np.float64, panel=False)
    ah = la.pack(a)
    piv_ref, _ = oracle.getrf(la, ah.copy(), 1)
    assert piv_ref[0, 0] == 2
    for algo in (capi.ALGO_REFERENCE, capi.ALGO_AUTO):
        capi.set_algorithm(algo)
        ipiv = torch.full((n,), -1, dtype=torch.int64, device="cuda")
        capi.getrf(dev(la, ah), ipiv, n, 1)
        assert int(ipiv[0]) == 2, algo
    capi.set_algorithm(capi.ALGO_REFERENCE)


@pytest.mark.parametrize("n,r", [(1, 1), (17, 3), (128, 4)])
def test_posv_bitexact(n, r):
    rng = np.random.default_rng(n)
    batch = 5
    a = spd(rng, batch, n)
    b = rng.uniform(-1, 1, (batch, n, r))
    la, lb = _lay(n, n, np.float64, True), _lay(n, r, np.float64, True)
    ah, bh = la.pack(a), lb.pack(b)
    lh, xh = np.zeros_like(ah), np.zeros_like(bh)
    oracle.posv(la, ah, la, lh, lb, bh, lb, xh, batch)
    L, X = capi.Operand.empty(la, batch), capi.Operand.empty(lb, batch)
    capi.posv(dev(la, ah), L, dev(lb, bh), X, batch)
    np.testing.assert_array_equal(X.host(), xh)
    np.testing.assert_array_equal(L.host(), lh)
