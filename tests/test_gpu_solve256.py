"""Workgroup-per-instance factor + solve kernels (n <= 256, nrhs <= 8) vs the
oracle: AUTO potrf / posv / getrf for 128 < n <= 256, the LAPACK-style
dgesv, and the mixed-precision dsposv / dsgesv with FP64 refinement on
matrices of prescribed condition number 1e2 .. 1e8.  GPU, through the C ABI."""

import numpy as np
import pytest

from blast_amd import capi
from blast_amd.layout import Layout
from oracle import oracle
from tests.numerics import EPS64, conditioned, lu_backward_error, potrf_backward_error, solve_backward_error, spd

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def auto_algorithm():
    capi.set_algorithm(capi.ALGO_AUTO)
    yield


def dev(layout, host):
    return capi.Operand.from_host(layout, host)


def i32(batch, fill):
    import torch

    return torch.full((batch,), fill, dtype=torch.int32, device="cuda")


@pytest.mark.parametrize("n", [129, 144, 200, 256])
@pytest.mark.parametrize("panel", [True, False])
def test_potrf_and_posv_vs_oracle(n, panel):
    rng = np.random.default_rng(n + panel)
    batch, nrhs = 4, 8
    a = spd(rng, batch, n)
    b = rng.uniform(-1, 1, (batch, n, nrhs))
    la, lb = Layout.make(n, n, panel=panel), Layout.make(n, nrhs, panel=not panel)
    ah, bh = la.pack(a), lb.pack(b)
    lh, xh = np.full_like(ah, 3.0), np.zeros_like(bh)
    oracle.posv(la, ah, la, lh, lb, bh, lb, xh, batch)
    L, X = capi.Operand.empty(la, batch, fill=3.0), capi.Operand.empty(lb, batch)
    info = i32(batch, -1)
    capi.posv(dev(la, ah), L, dev(lb, bh), X, batch, info)
    assert np.all(info.cpu().numpy() == 0)
    x, lg = X.logical(batch), L.logical(batch)
    assert np.all(solve_backward_error(a, x, b) < 4 * n * EPS64)
    assert np.all(potrf_backward_error(a, lg) < 4 * n * EPS64)
    l_ref = la.unpack(lh, batch)
    low = np.tril(np.ones((n, n), bool))
    np.testing.assert_allclose(lg[:, low], l_ref[:, low], rtol=0, atol=1e-13 * np.abs(l_ref).max())
    # storage contract: same as the oracle outside the lower triangle (untouched /
    # zeros in the 4x4 diagonal panel blocks)
    np.testing.assert_array_equal(lg[:, ~low], l_ref[:, ~low])
    # potrf alone, in place
    A = dev(la, ah)
    info2 = i32(batch, -1)
    capi.potrf(A, A, batch, info2)
    assert np.all(info2.cpu().numpy() == 0)
    assert np.all(potrf_backward_error(a, A.logical(batch)) < 4 * n * EPS64)


@pytest.mark.parametrize("n", [1, 17, 64, 100, 128])
@pytest.mark.parametrize("nrhs", [5, 8])
def test_posv_many_rhs_small_n(n, nrhs):
    # 5..8 right-hand sides take the multi-wavefront register kernel at any n
    rng = np.random.default_rng(n * 100 + nrhs)
    batch = 3
    a = spd(rng, batch, n)
    b = rng.uniform(-1, 1, (batch, n, nrhs))
    la, lb = Layout.make(n, n), Layout.make(n, nrhs)
    ah, bh = la.pack(a), lb.pack(b)
    lh, xh = np.full_like(ah, 3.0), np.zeros_like(bh)
    oracle.posv(la, ah, la, lh, lb, bh, lb, xh, batch)
    L, X = capi.Operand.empty(la, batch, fill=3.0), capi.Operand.empty(lb, batch)
    info = i32(batch, -1)
    capi.posv(dev(la, ah), L, dev(lb, bh), X, batch, info)
    assert np.all(info.cpu().numpy() == 0)
    assert np.all(solve_backward_error(a, X.logical(batch), b) < 4 * max(n, 8) * EPS64)
    lg, l_ref = L.logical(batch), la.unpack(lh, batch)
    low = np.tril(np.ones((n, n), bool))
    np.testing.assert_allclose(lg[:, low], l_ref[:, low], rtol=0, atol=1e-13 * np.abs(l_ref).max())
    np.testing.assert_array_equal(lg[:, ~low], l_ref[:, ~low])


def test_potrf_not_spd_reports_first_column():
    rng = np.random.default_rng(2)
    n, batch = 200, 3
    a = spd(rng, batch, n)
    a[1, 150, 150] = -1e5
    la = Layout.make(n, n)
    ah = la.pack(a)
    lh = np.zeros_like(ah)
    inf_ref = oracle.potrf(la, ah, la, lh, batch)
    L = capi.Operand.empty(la, batch)
    info = i32(batch, -1)
    capi.potrf(dev(la, ah), L, batch, info)
    np.testing.assert_array_equal(info.cpu().numpy(), inf_ref)
    assert inf_ref[1] == 151


@pytest.mark.parametrize("n", [129, 176, 256])
@pytest.mark.parametrize("panel", [False, True])
def test_getrf_pivots_bit_exact_and_backward_error(n, panel):
    import torch

    rng = np.random.default_rng(n)
    batch = 4
    a = rng.uniform(-1, 1, (batch, n, n)) + 0.5 * np.eye(n)
    la = Layout.make(n, n, panel=panel)
    ah = la.pack(a)
    ref = ah.copy()
    piv_ref, inf_ref = oracle.getrf(la, ref, batch)
    A = dev(la, ah)
    ipiv = torch.full((batch * n,), -1, dtype=torch.int64, device="cuda")
    info = i32(batch, -1)
    capi.getrf(A, ipiv, n, batch, info)
    piv = ipiv.cpu().numpy().reshape(batch, n)
    np.testing.assert_array_equal(piv, piv_ref)  # well-conditioned: the same pivots
    np.testing.assert_array_equal(info.cpu().numpy(), inf_ref)
    lu = A.logical(batch)
    for k in range(batch):
        assert lu_backward_error(a[k], lu[k], piv[k]) < 4 * n * EPS64
    np.testing.assert_allclose(lu, la.unpack(ref, batch), rtol=0, atol=1e-12 * np.abs(lu).max())


def test_getrf_singular_and_anti_diagonal():
    import torch

    n, batch = 160, 2
    a = np.zeros((batch, n, n))
    a[0] = np.fliplr(np.eye(n))  # pivots n-1-k for k < n/2, then k; L = U = I
    a[1] = np.random.default_rng(1).uniform(-1, 1, (n, n))
    a[1][:, 37] = 0.0  # exact zero column: singular at column 38
    la = Layout.make(n, n, panel=False)
    A = dev(la, la.pack(a))
    ipiv = torch.zeros(batch * n, dtype=torch.int64, device="cuda")
    info = i32(batch, -1)
    capi.getrf(A, ipiv, n, batch, info)
    piv = ipiv.cpu().numpy().reshape(batch, n)
    k = np.arange(n)
    np.testing.assert_array_equal(piv[0], np.where(k < n // 2, n - 1 - k, k))
    np.testing.assert_array_equal(A.logical(batch)[0], np.eye(n))
    assert info.cpu().numpy().tolist() == [0, 38]


@pytest.mark.parametrize("n,nrhs", [(256, 8), (130, 1), (64, 5), (17, 8)])
def test_gesv(n, nrhs):
    import torch

    rng = np.random.default_rng(n * nrhs)
    batch = 5
    a = rng.uniform(-1, 1, (batch, n, n)) + 0.5 * np.eye(n)
    b = rng.uniform(-1, 1, (batch, n, nrhs))
    la, lb = Layout.make(n, n, panel=False), Layout.make(n, nrhs)
    ah = la.pack(a)
    piv_ref, _ = oracle.getrf(la, ah.copy(), batch)
    A, B, X = dev(la, ah), dev(lb, lb.pack(b)), capi.Operand.empty(lb, batch)
    ipiv = torch.zeros(batch * n, dtype=torch.int64, device="cuda")
    info = i32(batch, -1)
    capi.gesv(A, ipiv, n, B, X, batch, info)
    assert np.all(info.cpu().numpy() == 0)
    np.testing.assert_array_equal(ipiv.cpu().numpy().reshape(batch, n), piv_ref)
    x = X.logical(batch)
    assert np.all(solve_backward_error(a, x, b) < 4 * n * EPS64)


KAPPAS = [1e2, 1e4, 1e6, 1e8]


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("spd_variant", [True, False])
def test_mixed_256_conditioned(precision, spd_variant):
    """n = 256, 8 right-hand sides, cond 1e2 .. 1e8: every instance ends at
    backward error <= 1e-12, by refinement or by the FP64 fallback; A is
    unchanged exactly for the instances that converged."""
    import torch

    rng = np.random.default_rng(40 + precision + 2 * spd_variant)
    n, nrhs = 256, 8
    kap = np.repeat(KAPPAS, 2)
    batch = len(kap)
    a = conditioned(rng, batch, n, kap, spd=spd_variant)
    b = rng.uniform(-1, 1, (batch, n, nrhs))
    la, lb = Layout.make(n, n, panel=spd_variant), Layout.make(n, nrhs)
    ah = la.pack(a)
    A, B, X = dev(la, ah), dev(lb, lb.pack(b)), capi.Operand.empty(lb, batch)
    it, info = i32(batch, -9), i32(batch, -9)
    work = capi.mixed_work(n, batch)
    if spd_variant:
        capi.sposv_ws(precision, A, B, X, work, batch, it, info, tol=1e-13)
    else:
        ipiv = torch.zeros(batch * n, dtype=torch.int64, device="cuda")
        capi.sgesv(precision, A, ipiv, n, B, X, work, batch, it, info, tol=1e-13)
    its, inf = it.cpu().numpy(), info.cpu().numpy()
    assert np.all(inf == 0), inf
    x = X.logical(batch)
    bwd = solve_backward_error(a, x, b)
    assert np.all(bwd <= 1e-12), (bwd, its)
    conv = its >= 0
    assert np.all((its >= -2) & (its <= 30)), its
    if precision == 0:
        assert np.all(conv[kap <= 1e4]), its  # kappa * u_32 << 1: refinement converges
    host = A.host().reshape(batch, -1)
    np.testing.assert_array_equal(host[conv], ah.reshape(batch, -1)[conv])
    if not spd_variant and np.any(conv):
        piv = ipiv.cpu().numpy().reshape(batch, n)
        assert np.all((piv >= np.arange(n)) & (piv < n))
    # agreement with the FP64 solution where the conditioning allows it: the
    # forward error is at most about kappa times the backward error
    well = kap <= 1e4
    x_ref = np.linalg.solve(a[well], b[well])
    ferr = np.abs(x[well] - x_ref).max(axis=(1, 2)) / np.abs(x_ref).max(axis=(1, 2))
    assert np.all(ferr <= kap[well] * 1e-12), ferr


def test_mixed_small_shapes_and_fallback_flags():
    """n = 40, 8 RHS (outside the n <= 128 / nrhs <= 4 LDS kernel): the
    workgroup kernel; a non-SPD instance breaks the FP32 factor (iter -1),
    and the FP64 fallback reports it in info."""
    rng = np.random.default_rng(5)
    n, nrhs, batch = 40, 8, 3
    a = spd(rng, batch, n)
    a[2, 10, 10] = -1e4
    b = rng.uniform(-1, 1, (batch, n, nrhs))
    la, lb = Layout.make(n, n), Layout.make(n, nrhs)
    A, B, X = dev(la, la.pack(a)), dev(lb, lb.pack(b)), capi.Operand.empty(lb, batch)
    it, info = i32(batch, -9), i32(batch, -9)
    capi.sposv_ws(0, A, B, X, capi.mixed_work(n, batch), batch, it, info)
    its, inf = it.cpu().numpy(), info.cpu().numpy()
    assert its[2] == -1 and np.all(its[:2] >= 0), its
    assert inf[2] == 11 and np.all(inf[:2] == 0), inf
    assert np.all(solve_backward_error(a[:2], X.logical(batch)[:2], b[:2]) < 1e-15)


def test_mixed_uncovered_shape_and_missing_work():
    import torch

    rng = np.random.default_rng(6)
    n, batch = 270, 2
    a = spd(rng, batch, n)
    b = rng.uniform(-1, 1, (batch, n, 2))
    la, lb = Layout.make(n, n), Layout.make(n, 2)
    A, B, X = dev(la, la.pack(a)), dev(lb, lb.pack(b)), capi.Operand.empty(lb, batch)
    it = i32(batch, 0)
    assert capi.mixed_work_elems(n) == 0
    capi.sposv_ws(0, A, B, X, None, batch, it)
    assert np.all(it.cpu().numpy() == -3)
    assert np.all(solve_backward_error(a, X.logical(batch), b) < 1e-14)
    la2, lb2 = Layout.make(200, 200), Layout.make(200, 2)
    A2 = capi.Operand.empty(la2, 1)
    with pytest.raises(capi.BlastError) as e:
        capi.sposv_ws(0, A2, capi.Operand.empty(lb2, 1), capi.Operand.empty(lb2, 1), None, 1,
                      torch.zeros(1, dtype=torch.int32, device="cuda"))
    assert e.value.code == -8


@pytest.mark.parametrize("n", [5, 48, 129, 256])
@pytest.mark.parametrize("panel", [False, True])
def test_float_getrf_auto(n, panel):
    """FP32 storage on the AUTO path (FP64 arithmetic inside, factors rounded
    to FP32): pivots equal the FP64 oracle's on the widened input
    (well-conditioned, so no near-ties), backward error at FP32 rounding."""
    import torch

    from tests.numerics import EPS32

    rng = np.random.default_rng(n + 7)
    batch = 3
    a = (rng.uniform(-1, 1, (batch, n, n)) + 0.5 * np.eye(n)).astype(np.float32)
    la32, la64 = Layout.make(n, n, np.float32, panel=panel), Layout.make(n, n, np.float64, panel=panel)
    ref = la64.pack(a.astype(np.float64))
    piv_ref, inf_ref = oracle.getrf(la64, ref, batch)
    A = dev(la32, la32.pack(a))
    ipiv = torch.full((batch * n,), -1, dtype=torch.int64, device="cuda")
    info = i32(batch, -1)
    capi.getrf(A, ipiv, n, batch, info)
    piv = ipiv.cpu().numpy().reshape(batch, n)
    np.testing.assert_array_equal(piv, piv_ref)
    np.testing.assert_array_equal(info.cpu().numpy(), inf_ref)
    lu = A.logical(batch).astype(np.float64)
    for k in range(batch):
        assert lu_backward_error(a[k].astype(np.float64), lu[k], piv[k]) < 8 * EPS32
