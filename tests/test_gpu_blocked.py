"""AUTO potrf / posv / trsm for n > 256 (the blocked path: 256-row blocks on
the n <= 256 MFMA kernels, trsm_tiles and the syrk / gemm tile kernels)
against the oracle, FP64 and FP32, panel and dense storage; info of a
non-SPD matrix whose first bad pivot lies beyond the first block."""

import numpy as np
import pytest

from blast_amd import capi
from blast_amd.layout import Layout
from oracle import oracle
from tests.numerics import EPS32, EPS64, potrf_backward_error, solve_backward_error, spd

pytestmark = pytest.mark.gpu

EPS = {np.float64: EPS64, np.float32: EPS32}


@pytest.fixture(autouse=True)
def auto_algorithm():
    capi.set_algorithm(capi.ALGO_AUTO)


def dev(layout, host):
    return capi.Operand.from_host(layout, host)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,panel", [(257, True), (300, False), (512, True), (520, False)])
def test_potrf_and_posv_blocked_vs_oracle(dtype, n, panel):
    import torch

    rng = np.random.default_rng(n + panel)
    batch, nrhs = 2, 3
    a = spd(rng, batch, n)
    b = rng.uniform(-1, 1, (batch, n, nrhs))
    la, lb = Layout.make(n, n, dtype, panel=panel), Layout.make(n, nrhs, dtype, panel=not panel)
    ah, bh = la.pack(a), lb.pack(b)
    lref = np.full_like(ah, 3.0)
    oracle.potrf(la, ah.copy(), la, lref, batch)
    L = capi.Operand.empty(la, batch, fill=3.0)
    info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    capi.potrf(dev(la, ah), L, batch, info)
    assert np.all(info.cpu().numpy() == 0)
    lg, lr = L.logical(batch).astype(np.float64), la.unpack(lref, batch).astype(np.float64)
    eps = EPS[dtype]
    assert np.all(potrf_backward_error(a, lg) < 8 * n * eps)
    low = np.tril(np.ones((n, n), bool))
    assert np.abs(lg[:, low] - lr[:, low]).max() <= 64 * n * eps * np.abs(lr).max()
    # storage contract: as the oracle outside the lower triangle (untouched,
    # zeros in the 4x4 diagonal panel blocks)
    np.testing.assert_array_equal(lg[:, ~low], lr[:, ~low])
    # posv: the factor and the two solves
    L2, X = capi.Operand.empty(la, batch, fill=3.0), capi.Operand.empty(lb, batch)
    info2 = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    capi.posv(dev(la, ah), L2, dev(lb, bh), X, batch, info2)
    assert np.all(info2.cpu().numpy() == 0)
    x = X.logical(batch).astype(np.float64)
    assert np.all(solve_backward_error(a, x, b) < 8 * n * eps)
    np.testing.assert_array_equal(L2.logical(batch), L.logical(batch))


def test_potrf_blocked_in_place_and_info_beyond_first_block():
    import torch

    n, batch = 400, 3
    rng = np.random.default_rng(4)
    a = spd(rng, batch, n)
    a[1][300, 300] = -5.0  # not SPD: first bad pivot at column 301 (the second block)
    a[2][10, 10] = -1.0    # first bad pivot at column 11
    la = Layout.make(n, n)
    A = dev(la, la.pack(a))
    info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    capi.potrf(A, A, batch, info)
    inf = info.cpu().numpy()
    ref_info = oracle.potrf(la, la.pack(a), la, np.zeros(batch * la.batch_stride), batch)
    assert inf.tolist() == [0, 301, 11]
    np.testing.assert_array_equal(inf, ref_info)
    assert potrf_backward_error(a[:1], A.logical(batch)[:1])[0] < 8 * n * EPS64


@pytest.mark.parametrize("side_right,upper", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_trsm_blocked_vs_oracle(side_right, upper):
    rng = np.random.default_rng(10 + 2 * side_right + upper)
    n, m, batch, alpha = 300, 5, 2, 0.75
    t = rng.uniform(-1, 1, (batch, n, n)) / np.sqrt(n) + 2 * np.eye(n)
    t = np.triu(t) if upper else np.tril(t)
    shape = (batch, m, n) if side_right else (batch, n, m)
    b = rng.uniform(-1, 1, shape)
    la, lb = Layout.make(n, n), Layout.make(*shape[1:])
    ah, bh = la.pack(t), lb.pack(b)
    xref = np.zeros_like(bh)
    oracle.trsm("right" if side_right else "left", not upper, 0, alpha, la, ah, lb, bh, lb, xref, batch)
    X = capi.Operand.empty(lb, batch)
    capi.trsm(side_right, upper, 0, alpha, dev(la, ah), dev(lb, bh), X, batch)
    x, xr = X.logical(batch), lb.unpack(xref, batch)
    assert np.abs(x - xr).max() <= 1e-12 * np.abs(xr).max()
