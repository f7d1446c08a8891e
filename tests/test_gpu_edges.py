"""Edge cases of the AUTO entry points on the GPU: empty batches and 0 x 0
matrices (return 0, write nothing), and a NaN entry of A, which every
Cholesky path reports as the first pivot that is not positive (LAPACK's
info rule: the NaN at (i, j), i > j, makes l_ij and then pivot i NaN)."""

import numpy as np
import pytest

from blast_amd import capi
from blast_amd.layout import Layout
from tests.numerics import spd

pytestmark = pytest.mark.gpu


def dev(layout, host):
    return capi.Operand.from_host(layout, host)


def test_empty_batch_and_zero_size():
    import torch

    la, lb = Layout.make(16, 16), Layout.make(16, 4, panel=False)
    A = dev(la, la.pack(spd(np.random.default_rng(1), 2, 16)))
    B = dev(lb, lb.pack(np.ones((2, 16, 4))))
    L, X = capi.Operand.empty(la, 2, fill=7.0), capi.Operand.empty(lb, 2, fill=7.0)
    info = torch.full((2,), -5, dtype=torch.int32, device="cuda")
    # batch 0: nothing is read or written
    capi.posv(A, L, B, X, 0, info)
    capi.potrf(A, L, 0, info)
    capi.gemm(1.0, A, A, 1.0, A, L, 0)
    capi.trsm(0, 0, 0, 1.0, A, B, X, 0)
    it = torch.zeros(2, dtype=torch.int32, device="cuda")
    capi.sposv(0, A, B, X, 0, it, info)
    torch.cuda.synchronize()
    assert np.all(L.host() == 7.0) and np.all(X.host() == 7.0)
    assert np.all(info.cpu().numpy() == -5)
    # 0 x 0 matrices, batch 3: every call succeeds and leaves info at 0
    z, zb = Layout.make(0, 0), Layout.make(0, 4, panel=False)
    Z, ZB = capi.Operand.empty(z, 3), capi.Operand.empty(zb, 3)
    info3 = torch.full((3,), -5, dtype=torch.int32, device="cuda")
    capi.potrf(Z, capi.Operand.empty(z, 3), 3, info3)
    torch.cuda.synchronize()
    assert np.all(info3.cpu().numpy() == 0)
    capi.posv(Z, capi.Operand.empty(z, 3), ZB, capi.Operand.empty(zb, 3), 3, info3)
    capi.gemm(1.0, Z, Z, 0.0, Z, capi.Operand.empty(z, 3), 3)
    torch.cuda.synchronize()
    # 64 x 0 result with K = 40 (the single-pass gemm's shape range)
    la40, lb0 = Layout.make(64, 40), Layout.make(40, 0)
    capi.gemm(1.0, capi.Operand.empty(la40, 2), capi.Operand.empty(lb0, 2), 0.0,
              capi.Operand.empty(Layout.make(64, 0), 2), capi.Operand.empty(Layout.make(64, 0), 2), 2)
    torch.cuda.synchronize()


@pytest.mark.parametrize("n,panel,i,j", [(128, True, 70, 40), (100, False, 60, 20), (200, True, 150, 100),
                                         (16, True, 9, 2)])
def test_nan_entry_reported_as_first_bad_pivot(n, panel, i, j):
    import torch

    rng = np.random.default_rng(n + i)
    batch = 4
    a = spd(rng, batch, n)
    a[2, i, j] = a[2, j, i] = np.nan
    la, lb = Layout.make(n, n, panel=panel), Layout.make(n, 2, panel=not panel)
    b = rng.uniform(-1, 1, (batch, n, 2))
    for solve in (False, True):
        L = capi.Operand.empty(la, batch)
        info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
        if solve:
            X = capi.Operand.empty(lb, batch)
            capi.posv(dev(la, la.pack(a)), L, dev(lb, lb.pack(b)), X, batch, info)
        else:
            capi.potrf(dev(la, la.pack(a)), L, batch, info)
        got = info.cpu().numpy()
        assert got[2] == i + 1, (solve, got)
        assert np.all(got[[0, 1, 3]] == 0), (solve, got)


@pytest.mark.parametrize("n,i,j", [(128, 70, 40), (64, 33, 5)])
def test_nan_entry_mixed_path_falls_back_and_reports(n, i, j):
    """dsposv: the FP32 factor of a NaN instance fails, the FP64 fallback
    reports the first NaN pivot; the other instances refine normally."""
    import torch

    rng = np.random.default_rng(n * i)
    batch = 4
    a = spd(rng, batch, n)
    a[1, i, j] = a[1, j, i] = np.nan
    la, lb = Layout.make(n, n), Layout.make(n, 2, panel=False)
    b = rng.uniform(-1, 1, (batch, n, 2))
    A, B, X = dev(la, la.pack(a)), dev(lb, lb.pack(b)), capi.Operand.empty(lb, batch)
    it = torch.zeros(batch, dtype=torch.int32, device="cuda")
    info = torch.full((batch,), -9, dtype=torch.int32, device="cuda")
    capi.sposv(0, A, B, X, batch, it, info)
    its, inf = it.cpu().numpy(), info.cpu().numpy()
    assert its[1] < 0 and inf[1] == i + 1, (its, inf)
    assert np.all(its[[0, 2, 3]] >= 0) and np.all(inf[[0, 2, 3]] == 0), (its, inf)
