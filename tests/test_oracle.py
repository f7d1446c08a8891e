"""Pins the CPU oracle: LAPACK golden vectors, closed forms, exact-arithmetic
cases and brute force.  Runs on the CPU only."""

import numpy as np
import pytest

from blast_amd.layout import Layout
from oracle import oracle
from tests.numerics import dyadic, lu_backward_error, potrf_backward_error, solve_backward_error, spd

LAYOUTS = [dict(panel=True), dict(panel=False)]


def _pack(x, panel=True, trans=False, dtype=np.float64):
    x = np.asarray(x, dtype=dtype)
    if x.ndim == 2:
        x = x[None]
    lay = Layout.make(x.shape[1], x.shape[2], dtype, panel=panel, trans=trans)
    return lay, lay.pack(x)


def _empty(rows, cols, batch=1, panel=True, trans=False, dtype=np.float64):
    lay = Layout.make(rows, cols, dtype, panel=panel, trans=trans)
    return lay, np.zeros(batch * lay.batch_stride, dtype=dtype)


# ------------------------------------------------------------------ layout
def test_panel_index_formula():
    # panel/DynamicPanelMatrix.hpp:60,187: spacing = PS*N, idx = i/PS*spacing + i%PS + j*PS
    lay = Layout.make(9, 5, np.float64, panel=True)
    assert lay.spacing == 4 * 5
    idx = lay.index_map()
    assert idx[0, 0] == 0 and idx[3, 0] == 3 and idx[0, 1] == 4 and idx[4, 0] == 20 and idx[8, 4] == 56
    assert Layout.make(9, 5, np.float32, panel=True).spacing == 8 * 5
    dense = Layout.make(9, 5, np.float64, panel=False)
    assert dense.spacing == 12 and dense.index_map()[2, 3] == 2 + 12 * 3


@pytest.mark.parametrize("panel", [True, False])
@pytest.mark.parametrize("trans", [True, False])
def test_pack_roundtrip(panel, trans):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((3, 7, 6))
    lay = Layout.make(7, 6, np.float64, panel=panel, trans=trans)
    buf = lay.pack(x)
    assert buf.size == 3 * lay.batch_stride and lay.batch_stride % 8 == 0
    np.testing.assert_array_equal(lay.unpack(buf, 3), x)
    # transposed view of the same storage
    np.testing.assert_array_equal(lay.t().unpack(buf, 3), x.transpose(0, 2, 1))


# ------------------------------------------------------------------ gemm
def test_gemm_golden(golden):
    la, a = _pack(golden["gemm_a"])
    lb, b = _pack(golden["gemm_b"], panel=False)
    lc, c = _pack(golden["gemm_c"])
    ld, d = _empty(5, 7)
    oracle.gemm(1.5, la, a, lb, b, -0.5, lc, c, ld, d, 1)
    np.testing.assert_allclose(ld.unpack(d, 1)[0], golden["gemm_d"], rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("ta,tb", [(False, False), (False, True), (True, False), (True, True)])
@pytest.mark.parametrize("panel", [True, False])
def test_gemm_exact(ta, tb, panel):
    rng = np.random.default_rng(2)
    m, n, k, batch = 13, 9, 17, 3
    a, b, c = dyadic(rng, (batch, m, k)), dyadic(rng, (batch, k, n)), dyadic(rng, (batch, m, n))
    la = Layout.make(k, m, panel=panel).t() if ta else Layout.make(m, k, panel=panel)
    lb = Layout.make(n, k, panel=not panel).t() if tb else Layout.make(k, n, panel=not panel)
    lc, cb = _pack(c, panel=panel)
    ld, d = _empty(m, n, batch, panel=panel)
    oracle.gemm(1.5, la, la.pack(a), lb, lb.pack(b), -0.5, lc, cb, ld, d, batch)
    np.testing.assert_array_equal(ld.unpack(d, batch), 1.5 * (a @ b) - 0.5 * c)


def test_gemm_inplace_c_is_d():
    rng = np.random.default_rng(3)
    a, b, c = dyadic(rng, (2, 8, 5)), dyadic(rng, (2, 5, 6)), dyadic(rng, (2, 8, 6))
    la, ab = _pack(a)
    lb, bb = _pack(b)
    lc, cb = _pack(c)
    oracle.gemm(1.0, la, ab, lb, bb, 1.0, lc, cb, lc, cb, 2)
    np.testing.assert_array_equal(lc.unpack(cb, 2), a @ b + c)


# ------------------------------------------------------------------ syrk / trmm / ger
def test_syrk_lower_exact():
    rng = np.random.default_rng(4)
    a, c = dyadic(rng, (2, 11, 7)), dyadic(rng, (2, 11, 11))
    la, ab = _pack(a, panel=False)
    lc, cb = _pack(c, panel=False)
    ld, d = _pack(np.full((2, 11, 11), 7.0), panel=False)
    oracle.syrk_lower(0.5, la, ab, 2.0, lc, cb, ld, d, 2)
    got = ld.unpack(d, 2)
    ref = 0.5 * a @ a.transpose(0, 2, 1) + 2.0 * c
    low = np.tril(np.ones((11, 11), bool))
    np.testing.assert_array_equal(got[:, low], ref[:, low])
    assert np.all(got[:, ~low] == 7.0)  # strict upper untouched


def test_trmm_exact():
    rng = np.random.default_rng(5)
    a, b = dyadic(rng, (2, 9, 9)), dyadic(rng, (2, 9, 6))
    la, ab = _pack(a)
    lb, bb = _pack(b)
    lc, cb = _empty(9, 6, 2)
    oracle.trmm_left_upper(0.5, la, ab, lb, bb, lc, cb, 2)
    np.testing.assert_array_equal(lc.unpack(cb, 2), 0.5 * np.triu(a) @ b)
    b2 = dyadic(rng, (2, 6, 9))
    lb2, bb2 = _pack(b2)
    lc2, cb2 = _empty(6, 9, 2)
    oracle.trmm_right_lower(2.0, lb2, bb2, la, ab, lc2, cb2, 2)
    np.testing.assert_array_equal(lc2.unpack(cb2, 2), 2.0 * b2 @ np.tril(a))


def test_ger_exact():
    rng = np.random.default_rng(6)
    x, y, a = dyadic(rng, 7), dyadic(rng, 5), dyadic(rng, (7, 5))
    la, ab = _pack(a, panel=False)
    lb, bb = _empty(7, 5, panel=False)
    oracle.ger(-1.0, x, y, la, ab, lb, bb)
    np.testing.assert_array_equal(lb.unpack(bb, 1)[0], a - np.outer(x, y))


# ------------------------------------------------------------------ trsm
@pytest.mark.parametrize("n,r", [(4, 3), (11, 5), (16, 2)])
def test_trsm_golden(golden, n, r):
    l, b = golden[f"trsm_l_{n}x{r}"], golden[f"trsm_b_{n}x{r}"]
    ll, lbuf = _pack(l)
    lb, bb = _pack(b)
    for lower, unit, key, lay in ((1, 0, "xlower", ll), (0, 0, "xupper", ll.t()), (1, 1, "xunit", ll)):
        lx, x = _empty(n, r)
        oracle.trsm("left", lower, unit, 1.0, lay, lbuf, lb, bb, lx, x, 1)
        np.testing.assert_allclose(lx.unpack(x, 1)[0], golden[f"trsm_{key}_{n}x{r}"], rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("lower", [1, 0])
@pytest.mark.parametrize("unit", [1, 0])
def test_trsm_exact_and_alias(lower, unit):
    # integer triangular A with power-of-two diagonal and B = A X0 exact => X = X0 exactly
    rng = np.random.default_rng(7)
    n, r = 12, 4
    a = rng.integers(-2, 3, (n, n)).astype(np.float64)
    a = np.tril(a, -1) if lower else np.triu(a, 1)
    diag = np.ones(n) if unit else 2.0 ** rng.integers(-1, 2, n)
    a += np.diag(diag)
    x0 = rng.integers(-3, 4, (n, r)).astype(np.float64)
    b = a @ x0
    la, ab = _pack(a)
    lb, bb = _pack(b, panel=False)
    oracle.trsm("left", lower, unit, 1.0, la, ab, lb, bb, lb, bb, 1)  # X aliases B
    np.testing.assert_array_equal(lb.unpack(bb, 1)[0], x0)
    # right side: X A = alpha B
    x1 = rng.integers(-3, 4, (r, n)).astype(np.float64)
    b1 = 0.5 * (x1 @ a)
    lb1, bb1 = _pack(b1)
    lx1, xx1 = _empty(r, n)
    oracle.trsm("right", lower, unit, 2.0, la, ab, lb1, bb1, lx1, xx1, 1)
    np.testing.assert_array_equal(lx1.unpack(xx1, 1)[0], x1)


# ------------------------------------------------------------------ potrf
@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 13, 16])
@pytest.mark.parametrize("panel", [True, False])
def test_potrf_golden(golden, n, panel):
    a, lref = golden[f"potrf_a_{n}"], golden[f"potrf_l_{n}"]
    la, ab = _pack(a, panel=panel)
    ll, lbuf = _empty(n, n, panel=panel)
    info = oracle.potrf(la, ab, ll, lbuf, 1)
    assert info[0] == 0
    np.testing.assert_allclose(np.tril(ll.unpack(lbuf, 1)[0]), lref, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("panel", [True, False])
def test_potrf_closed_forms(panel):
    n = 11
    # identity and perfect-square diagonals
    for d in (np.ones(n), np.arange(1, n + 1, dtype=np.float64) ** 2):
        la, ab = _pack(np.diag(d), panel=panel)
        ll, lb = _empty(n, n, panel=panel)
        oracle.potrf(la, ab, ll, lb, 1)
        np.testing.assert_array_equal(np.tril(ll.unpack(lb, 1)[0]), np.diag(np.sqrt(d)))
    # A = L0 L0^T with small-integer L0 and power-of-two diagonal: every step is exact
    rng = np.random.default_rng(8)
    l0 = np.tril(rng.integers(-2, 3, (n, n)).astype(np.float64), -1) + np.diag(2.0 ** rng.integers(0, 3, n))
    la, ab = _pack(l0 @ l0.T, panel=panel)
    ll, lb = _pack(np.full((n, n), 9.0), panel=panel)
    oracle.potrf(la, ab, ll, lb, 1)
    got = ll.unpack(lb, 1)[0]
    np.testing.assert_array_equal(np.tril(got), l0)
    up = np.triu(np.ones((n, n), bool), 1)
    if panel:
        # strict upper of each 4x4 diagonal block written as 0, rest untouched
        blk = (np.arange(n)[:, None] // 4) == (np.arange(n)[None, :] // 4)
        assert np.all(got[up & blk] == 0.0) and np.all(got[up & ~blk] == 9.0)
    else:
        assert np.all(got[up] == 9.0)


@pytest.mark.parametrize("n", [0, 1, 4, 7, 33, 50])
def test_potrf_reconstruction_and_inplace(n):
    rng = np.random.default_rng(9)
    a = spd(rng, 2, n) if n else np.zeros((2, 0, 0))
    la, ab = _pack(a)
    info = oracle.potrf(la, ab, la, ab, 2)  # L aliases A
    assert np.all(info == 0)
    if n:
        assert np.all(potrf_backward_error(a, la.unpack(ab, 2)) < 4 * n * np.finfo(float).eps)


def test_potrf_not_spd_reports_column_and_nan():
    a = np.eye(6)
    a[3, 3] = -1.0
    la, ab = _pack(a, panel=False)
    ll, lb = _empty(6, 6, panel=False)
    info = oracle.potrf(la, ab, ll, lb, 1)
    assert info[0] == 4
    assert np.isnan(ll.unpack(lb, 1)[0][3, 3])


def test_potrf_tall():
    rng = np.random.default_rng(10)
    a = spd(rng, 1, 9)[0]
    la, ab = _pack(a[:, :6])  # 9 x 6: tall potrf (M >= N)
    ll, lb = _empty(9, 6)
    oracle.potrf(la, ab, ll, lb, 1)
    full = np.linalg.cholesky(a)
    np.testing.assert_allclose(np.tril(ll.unpack(lb, 1)[0]), full[:, :6], rtol=1e-13, atol=1e-13)


# ------------------------------------------------------------------ getrf
@pytest.mark.parametrize("shape", [(5, 5), (8, 8), (16, 16), (12, 7), (7, 12), (13, 13)])
def test_getrf_golden(golden, shape):
    m, n = shape
    key = f"{m}x{n}"
    la, ab = _pack(golden[f"getrf_a_{key}"], panel=False)
    ipiv, info = oracle.getrf(la, ab, 1)
    assert info[0] == 0
    np.testing.assert_array_equal(ipiv[0], golden[f"getrf_piv_{key}"])
    np.testing.assert_allclose(la.unpack(ab, 1)[0], golden[f"getrf_lu_{key}"], rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("n", [1, 4, 5, 9, 16])
def test_getrf_closed_forms(n):
    la, ab = _pack(np.eye(n), panel=False)
    ipiv, info = oracle.getrf(la, ab, 1)
    np.testing.assert_array_equal(ipiv[0], np.arange(n))
    # anti-diagonal (bench/blast/math/dense/StaticGetrf.cpp:39-58)
    la, ab = _pack(np.eye(n)[::-1], panel=False)
    ipiv, info = oracle.getrf(la, ab, 1)
    k = np.arange(n)
    np.testing.assert_array_equal(ipiv[0], np.where(k < n // 2, n - 1 - k, k))
    np.testing.assert_array_equal(la.unpack(ab, 1)[0], np.eye(n))


@pytest.mark.parametrize("shape", [(40, 40), (64, 33), (33, 64), (4, 4), (0, 3)])
def test_getrf_backward_error(shape):
    m, n = shape
    rng = np.random.default_rng(11)
    a = rng.uniform(-1, 1, (2, m, n))
    la, ab = _pack(a, panel=False)
    ipiv, info = oracle.getrf(la, ab, 2)
    assert np.all(info == 0)
    if m and n:
        lu = la.unpack(ab, 2)
        for b in range(2):
            assert lu_backward_error(a[b], lu[b], ipiv[b]) < 8 * max(m, n) * np.finfo(float).eps


def test_getrf_singular_and_rowmajor():
    a = np.ones((5, 5))
    la, ab = _pack(a, panel=False)
    _, info = oracle.getrf(la, ab, 1)
    assert info[0] == 2  # column 1 is exactly zero after eliminating column 0
    # row-major storage: no pivoting (dense/Getrf.hpp:141-162)
    rng = np.random.default_rng(12)
    a = rng.uniform(-1, 1, (6, 6)) + 6 * np.eye(6)
    lr = Layout.make(6, 6, panel=False).t()
    buf = lr.pack(a)
    ipiv, info = oracle.getrf(lr, buf, 1)
    np.testing.assert_array_equal(ipiv[0], np.arange(6))
    lu = lr.unpack(buf, 1)[0]
    np.testing.assert_allclose((np.tril(lu, -1) + np.eye(6)) @ np.triu(lu), a, rtol=1e-13, atol=1e-13)


def test_getf2_matches_golden_pivots(golden):
    a = golden["getrf_a_5x5"]
    la, ab = _pack(a, panel=False)
    ipiv, r = oracle.getf2(la, ab)
    assert r == 0
    np.testing.assert_array_equal(ipiv, golden["getrf_piv_5x5"])


def test_iamax_brute_force():
    rng = np.random.default_rng(13)
    for n in range(1, 50):
        x = rng.integers(-5, 6, n).astype(np.float64)
        assert oracle.iamax(x) == int(np.argmax(np.abs(x)))  # numpy argmax = first max
    x = np.array([np.nan, 1.0, 2.0])
    assert oracle.iamax(x) == 0  # strict '>' never moves off a leading NaN
    x = np.array([1.0, np.nan, -3.0, 3.0])
    assert oracle.iamax(x) == 2


def test_laswp_matches_sequential_swaps():
    rng = np.random.default_rng(14)
    a = rng.standard_normal((7, 4))
    ipiv = np.array([3, 1, 6, 5, 4, 6, 6])
    la, ab = _pack(a, panel=False)
    oracle.laswp(la, ab, 1, 5, ipiv)
    ref = a.copy()
    for k in range(1, 5):
        ref[[k, ipiv[k]]] = ref[[ipiv[k], k]]
    np.testing.assert_array_equal(la.unpack(ab, 1)[0], ref)


# ------------------------------------------------------------------ posv (the composed solve)
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-15), (np.float32, 1e-6)])
def test_posv(dtype, tol):
    rng = np.random.default_rng(15)
    n, r, batch = 40, 4, 3
    a = spd(rng, batch, n)
    b = rng.uniform(-1, 1, (batch, n, r))
    la, ab = _pack(a, dtype=dtype)
    ll, lb = _empty(n, n, batch, dtype=dtype)
    lbb, bb = _pack(b, dtype=dtype)
    lx, xb = _empty(n, r, batch, dtype=dtype)
    info = oracle.posv(la, ab, ll, lb, lbb, bb, lx, xb, batch)
    assert np.all(info == 0)
    x = lx.unpack(xb, batch).astype(np.float64)
    assert np.all(solve_backward_error(a, x, b) < tol)
    info, used = oracle.posv_mt(la, ab, ll, lb, lbb, bb, lx, xb.copy(), batch, threads=2) if dtype == np.float64 else (info, 2)
    assert used == 2
