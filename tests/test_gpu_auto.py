"""Fast (AUTO) kernels vs the oracle: agreement to backward error c*n*eps, plus
the storage contract (what is and is not written).  GPU, through the C ABI."""

import numpy as np
import pytest

from blast_amd import capi
from blast_amd.layout import Layout
from oracle import oracle
from tests.numerics import EPS64, potrf_backward_error, solve_backward_error, spd

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def auto_algorithm():
    capi.set_algorithm(capi.ALGO_AUTO)
    yield


def dev(layout, host):
    return capi.Operand.from_host(layout, host)


@pytest.mark.parametrize("n", [1, 3, 16, 17, 40, 64, 100, 127, 128])
@pytest.mark.parametrize("nrhs", [1, 4])
@pytest.mark.parametrize("panel", [True, False])
def test_posv_fused_vs_oracle(n, nrhs, panel):
    import torch

    rng = np.random.default_rng(n * 10 + nrhs)
    batch = 7
    a = spd(rng, batch, n)
    b = rng.uniform(-1, 1, (batch, n, nrhs))
    la, lb = Layout.make(n, n, panel=panel), Layout.make(n, nrhs, panel=not panel)
    ah, bh = la.pack(a), lb.pack(b)
    lh = np.full_like(ah, 3.0)
    xh = np.zeros_like(bh)
    oracle.posv(la, ah, la, lh, lb, bh, lb, xh, batch)
    L, X = capi.Operand.empty(la, batch, fill=3.0), capi.Operand.empty(lb, batch)
    info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    capi.posv(dev(la, ah), L, dev(lb, bh), X, batch, info)
    assert np.all(info.cpu().numpy() == 0)
    x, x_ref = X.logical(batch), lb.unpack(xh, batch)
    lg, l_ref = L.logical(batch), la.unpack(lh, batch)
    assert np.all(solve_backward_error(a, x, b) < 4 * n * EPS64)
    assert np.all(potrf_backward_error(a, lg) < 4 * n * EPS64)
    np.testing.assert_allclose(x, x_ref, rtol=0, atol=1e-13 * np.abs(x_ref).max())
    low = np.tril(np.ones((n, n), bool))
    np.testing.assert_allclose(lg[:, low], l_ref[:, low], rtol=0, atol=1e-13 * np.abs(l_ref).max())
    # storage contract: identical to the reference-order path outside the lower triangle
    np.testing.assert_array_equal(lg[:, ~low], l_ref[:, ~low])


@pytest.mark.parametrize("n", [8, 128])
def test_potrf_fused_inplace_and_info(n):
    import torch

    rng = np.random.default_rng(3)
    batch = 6
    a = spd(rng, batch, n)
    a[4, n // 2, n // 2] = -1.0e4  # not SPD
    la = Layout.make(n, n)
    A = dev(la, la.pack(a))
    info = torch.zeros(batch, dtype=torch.int32, device="cuda")
    capi.potrf(A, A, batch, info)  # L aliases A
    inf = info.cpu().numpy()
    assert inf[4] > 0 and np.all(np.delete(inf, 4) == 0)
    good = [0, 1, 2, 3, 5]
    assert np.all(potrf_backward_error(a[good], A.logical(batch)[good]) < 4 * n * EPS64)


def test_posv_alias_x_b():
    rng = np.random.default_rng(4)
    n, batch = 96, 5
    a = spd(rng, batch, n)
    b = rng.uniform(-1, 1, (batch, n, 3))
    la, lb = Layout.make(n, n), Layout.make(n, 3)
    A, B = dev(la, la.pack(a)), dev(lb, lb.pack(b))
    capi.posv(A, A, B, B, batch)  # L aliases A, X aliases B
    assert np.all(solve_backward_error(a, B.logical(batch), b) < 4 * n * EPS64)


# ------------------------------------------------------------------ mixed precision + FP64 refinement
def _haar(rng, n):
    q, r = np.linalg.qr(rng.standard_normal((n, n)))
    return q * np.sign(np.diag(r))


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("n", [1, 7, 16, 33, 64, 128])
@pytest.mark.parametrize("nrhs", [1, 4])
def test_sposv_ir_converges_to_fp64(precision, n, nrhs):
    import torch

    rng = np.random.default_rng(100 * n + nrhs + precision)
    batch = 9
    a = spd(rng, batch, n)
    b = rng.uniform(-1, 1, (batch, n, nrhs))
    la, lb = Layout.make(n, n), Layout.make(n, nrhs)
    ah = la.pack(a)
    A, B, X = dev(la, ah), dev(lb, lb.pack(b)), capi.Operand.empty(lb, batch)
    it = torch.full((batch,), -9, dtype=torch.int32, device="cuda")
    info = torch.full((batch,), -9, dtype=torch.int32, device="cuda")
    capi.sposv(precision, A, B, X, batch, it, info)
    its, inf = it.cpu().numpy(), info.cpu().numpy()
    assert np.all(its >= 0) and np.all(inf == 0), (its, inf)
    assert np.all(its <= (3 if precision == 0 else 30)), its
    x = X.logical(batch)
    assert np.all(solve_backward_error(a, x, b) < 1e-15)
    np.testing.assert_allclose(x, np.linalg.solve(a, b), rtol=0, atol=1e-12 * np.abs(x).max())
    np.testing.assert_array_equal(A.host(), ah)  # A untouched when refinement converges


def test_sposv_ir_falls_back_for_hard_instances():
    import torch

    rng = np.random.default_rng(7)
    n, batch = 64, 4
    a = spd(rng, batch, n)
    u = _haar(rng, n)
    a[1] = (u * np.logspace(0, -13, n)) @ u.T  # SPD, cond 1e13: beyond FP32
    a[1] = 0.5 * (a[1] + a[1].T)
    b = rng.uniform(-1, 1, (batch, n, 2))
    la, lb = Layout.make(n, n), Layout.make(n, 2)
    A, B, X = dev(la, la.pack(a)), dev(lb, lb.pack(b)), capi.Operand.empty(lb, batch)
    it = torch.zeros(batch, dtype=torch.int32, device="cuda")
    info = torch.full((batch,), -9, dtype=torch.int32, device="cuda")
    capi.sposv(0, A, B, X, batch, it, info)
    its = it.cpu().numpy()
    assert its[1] < 0 and np.all(its[[0, 2, 3]] >= 0), its
    x = X.logical(batch)
    assert np.all(solve_backward_error(a, x, b) < 1e-14)
    lfac = A.logical(batch)[1]  # the fallback wrote the FP64 factor into A
    assert potrf_backward_error(a[1], lfac) < 1e-14


def test_sposv_uncovered_shape_uses_fp64():
    import torch

    rng = np.random.default_rng(8)
    n, batch = 150, 3
    a = spd(rng, batch, n)
    b = rng.uniform(-1, 1, (batch, n, 5))
    la, lb = Layout.make(n, n), Layout.make(n, 5)
    A, B, X = dev(la, la.pack(a)), dev(lb, lb.pack(b)), capi.Operand.empty(lb, batch)
    it = torch.zeros(batch, dtype=torch.int32, device="cuda")
    capi.sposv(0, A, B, X, batch, it)
    assert np.all(it.cpu().numpy() == -3)
    assert np.all(solve_backward_error(a, X.logical(batch), b) < 1e-15)


# ------------------------------------------------------------------ gemm / syrk / trmm on MFMA tiles
def _lay(rows, cols, dtype, panel, trans=False):
    return Layout.make(cols, rows, dtype, panel=panel).t() if trans else Layout.make(rows, cols, dtype, panel=panel)


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-14), (np.float32, 2e-6)])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (5, 7, 3), (16, 16, 16), (64, 64, 64), (33, 70, 9), (100, 20, 129),
                                   (48, 64, 40), (64, 17, 33), (20, 64, 64)])
@pytest.mark.parametrize("ta,tb,panel", [(0, 1, 1), (0, 0, 0), (1, 1, 1), (1, 0, 0)])
def test_gemm_tiles(dtype, tol, m, n, k, ta, tb, panel):
    rng = np.random.default_rng(m + 31 * n + 7 * k)
    batch = 5
    a, b, c = (rng.uniform(-1, 1, s) for s in ((batch, m, k), (batch, k, n), (batch, m, n)))
    la, lb, lc = _lay(m, k, dtype, panel, ta), _lay(k, n, dtype, not panel, tb), _lay(m, n, dtype, panel)
    D = capi.Operand.empty(lc, batch)
    capi.gemm(1.5, dev(la, la.pack(a)), dev(lb, lb.pack(b)), -0.5, dev(lc, lc.pack(c)), D, batch)
    ref = 1.5 * (a @ b) - 0.5 * c
    err = np.abs(D.logical(batch) - ref).max() / (k * max(np.abs(a).max() * np.abs(b).max(), 1e-30) + 1)
    assert err < tol


def test_gemm_tiles_exact_dyadic_and_inplace():
    from tests.numerics import dyadic

    rng = np.random.default_rng(1)
    batch, m = 6, 64
    a, b, c = dyadic(rng, (batch, m, m)), dyadic(rng, (batch, m, m)), dyadic(rng, (batch, m, m))
    la = Layout.make(m, m)
    C = dev(la, la.pack(c))
    capi.gemm(1.5, dev(la, la.pack(a)), dev(la.t(), la.t().pack(b)), -0.5, C, C, batch)  # NT, C aliases D
    np.testing.assert_array_equal(C.logical(batch), 1.5 * (a @ b) - 0.5 * c)


def test_gemm_nn_all_panel_64_config():
    """The dgemm benchmark shape: D = 1.5 A B - 0.5 C, m = n = k = 64, every
    operand in StaticPanelMatrix storage (NN), U(-1,1) entries; against the
    oracle (BLAST's per-element FMA order) within the gemm backward-error
    bound, and bit-exact on dyadic inputs (every sum exact)."""
    from tests.numerics import dyadic

    rng = np.random.default_rng(64)
    batch, m = 8, 64
    la = Layout.make(m, m)  # panel
    a, b, c = (rng.uniform(-1, 1, (batch, m, m)) for _ in range(3))
    ah, bh, ch = la.pack(a), la.pack(b), la.pack(c)
    dh = np.zeros_like(ah)
    oracle.gemm(1.5, la, ah, la, bh, -0.5, la, ch, la, dh, batch)
    D = capi.Operand.empty(la, batch)
    capi.gemm(1.5, dev(la, ah), dev(la, bh), -0.5, dev(la, ch), D, batch)
    got, ref = D.logical(batch), la.unpack(dh, batch)
    err = np.abs(got - ref).max() / (m * np.abs(a).max() * np.abs(b).max())
    assert err < 4 * EPS64, err
    # dyadic entries: every product and partial sum is exact, so any
    # summation order gives the same bits
    a, b, c = dyadic(rng, (batch, m, m)), dyadic(rng, (batch, m, m)), dyadic(rng, (batch, m, m))
    D = capi.Operand.empty(la, batch)
    capi.gemm(1.5, dev(la, la.pack(a)), dev(la, la.pack(b)), -0.5, dev(la, la.pack(c)), D, batch)
    np.testing.assert_array_equal(D.logical(batch), 1.5 * (a @ b) - 0.5 * c)


@pytest.mark.parametrize("m,k", [(1, 1), (9, 4), (37, 11), (64, 64)])
def test_syrk_tiles(m, k):
    rng = np.random.default_rng(m * k)
    batch = 4
    a, c = rng.uniform(-1, 1, (batch, m, k)), rng.uniform(-1, 1, (batch, m, m))
    la, lc = Layout.make(m, k, panel=False), Layout.make(m, m, panel=False)
    D = capi.Operand.empty(lc, batch, fill=7.0)
    capi.syrk_lower(0.5, dev(la, la.pack(a)), 2.0, dev(lc, lc.pack(c)), D, batch)
    got = D.logical(batch)
    ref = 0.5 * a @ a.transpose(0, 2, 1) + 2.0 * c
    low = np.tril(np.ones((m, m), bool))
    np.testing.assert_allclose(got[:, low], ref[:, low], rtol=0, atol=1e-13 * max(k, 1))
    assert np.all(got[:, ~low] == 7.0)


@pytest.mark.parametrize("side_right", [0, 1])
@pytest.mark.parametrize("m,n", [(1, 1), (7, 5), (20, 33)])
def test_trmm_tiles(side_right, m, n):
    rng = np.random.default_rng(m * n + 3 * side_right)
    batch = 3
    na = n if side_right else m
    a, b = rng.uniform(-1, 1, (batch, na, na)), rng.uniform(-1, 1, (batch, m, n))
    la, lb = Layout.make(na, na), Layout.make(m, n)
    C = capi.Operand.empty(lb, batch)
    capi.trmm(side_right, not side_right, 0.75, dev(la, la.pack(a)), dev(lb, lb.pack(b)), C, batch)
    ref = 0.75 * (b @ np.tril(a) if side_right else np.triu(a) @ b)
    np.testing.assert_allclose(C.logical(batch), ref, rtol=0, atol=1e-13 * na)


# ------------------------------------------------------------------ trsm on MFMA tiles
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-13), (np.float32, 2e-5)])
@pytest.mark.parametrize("side_right,upper,unit", [(0, 0, 0), (0, 1, 0), (0, 0, 1), (0, 1, 1), (1, 1, 0), (1, 0, 1),
                                                   (1, 0, 0)])
@pytest.mark.parametrize("m,n", [(1, 1), (17, 3), (64, 33), (130, 5), (256, 8), (5, 200)])
@pytest.mark.parametrize("trans_a", [False, True])
def test_trsm_tiles_vs_oracle(dtype, tol, side_right, upper, unit, m, n, trans_a):
    rng = np.random.default_rng(m * 7 + n + 3 * side_right + upper)
    batch = 3
    na = n if side_right else m
    if na > 256:
        return
    # well conditioned with or without the unit diagonal: off-diagonal entries U(-1,1)/n
    a = rng.uniform(-1, 1, (batch, na, na)) / na + np.eye(na) * (1 + rng.uniform(0, 1, (batch, 1, na)))
    a = np.triu(a) if upper else np.tril(a)
    a += np.triu(rng.standard_normal((batch, na, na)), 1) * 50 if not upper else np.tril(
        rng.standard_normal((batch, na, na)), -1) * 50  # garbage in the other triangle: must not be read
    b = rng.uniform(-1, 1, (batch, m, n))
    la = (Layout.make(na, na, dtype, panel=True).t() if trans_a else Layout.make(na, na, dtype, panel=True))
    lb = Layout.make(m, n, dtype, panel=False)
    ah, bh = la.pack(a.astype(dtype)), lb.pack(b.astype(dtype))
    side = "right" if side_right else "left"
    xh = np.zeros_like(bh)
    oracle.trsm(side, not upper, unit, 1.5, la, ah, lb, bh, lb, xh, batch)
    X = capi.Operand.empty(lb, batch)
    capi.trsm(side_right, upper, unit, 1.5, dev(la, ah), dev(lb, bh), X, batch)
    x, x_ref = X.logical(batch), lb.unpack(xh, batch)
    np.testing.assert_allclose(x, x_ref, rtol=0, atol=tol * max(1.0, np.abs(x_ref).max()))
    # X may alias B
    B = dev(lb, bh)
    capi.trsm(side_right, upper, unit, 1.5, dev(la, ah), B, B, batch)
    np.testing.assert_allclose(B.logical(batch), x_ref, rtol=0, atol=tol * max(1.0, np.abs(x_ref).max()))


@pytest.mark.parametrize("n", [1, 7, 16, 33, 100, 128, 192, 256])
@pytest.mark.parametrize("panel", [True, False])
def test_float_posv_and_potrf_auto(n, panel):
    """FP32 storage on the AUTO path (FP64 arithmetic inside, results rounded
    to FP32): backward errors at FP32 rounding level, X against the FP64
    solution, L only in its lower triangle (zeros above it inside the 4x4
    diagonal blocks of panel storage, as the reference-order kernel writes)."""
    import torch

    from tests.numerics import EPS32

    rng = np.random.default_rng(n + 1000 * panel)
    batch, nrhs = 5, min(8, max(1, n // 16))
    a = spd(rng, batch, n).astype(np.float32)
    b = rng.uniform(-1, 1, (batch, n, nrhs)).astype(np.float32)
    la, lb = Layout.make(n, n, np.float32, panel=panel), Layout.make(n, nrhs, np.float32, panel=not panel)
    ah, bh = la.pack(a), lb.pack(b)
    L, X = capi.Operand.empty(la, batch, fill=3.0), capi.Operand.empty(lb, batch)
    info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    capi.posv(dev(la, ah), L, dev(lb, bh), X, batch, info)
    assert np.all(info.cpu().numpy() == 0)
    a64, b64 = a.astype(np.float64), b.astype(np.float64)
    lg, x = L.logical(batch).astype(np.float64), X.logical(batch).astype(np.float64)
    assert np.all(potrf_backward_error(a64, lg) < 4 * EPS32)
    assert np.all(solve_backward_error(a64, x, b64) < 4 * EPS32)
    x_ref = np.linalg.solve(a64, b64)
    assert np.all(np.abs(x - x_ref).max(axis=(1, 2)) <= 8 * EPS32 * np.linalg.cond(a64) * np.abs(x_ref).max(axis=(1, 2)))
    low = np.tril(np.ones((n, n), bool))
    blk = np.zeros((n, n), bool)
    for i in range(n):
        for j in range(i + 1, n):
            blk[i, j] = panel and i // 4 == j // 4
    assert np.all(lg[:, blk] == 0.0)
    assert np.all(lg[:, ~low & ~blk] == 3.0)  # untouched
    # potrf alone, in place
    A = dev(la, ah)
    capi.potrf(A, A, batch, info)
    assert np.all(info.cpu().numpy() == 0)
    np.testing.assert_array_equal(A.logical(batch)[:, low], L.logical(batch)[:, low])
