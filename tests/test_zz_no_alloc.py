"""The allocation-free mixed-precision LU entry point
(blast_hip_dsgesv_noalloc_batched) against blast_hip_dsgesv_batched with a
caller work buffer: the same kernels on the same inputs, so X, the pivots,
the iteration counts, info and A must be identical.  Two calls in a row, the
second smaller (the device scratch keeps the first call's contents), then a
batch beyond the scratch is rejected before anything runs.  GPU, through the
C ABI."""

import numpy as np
import pytest

from blast_amd import capi
from blast_amd.layout import Layout
from tests.numerics import conditioned, solve_backward_error

pytestmark = pytest.mark.gpu


def _run_pair(rng, n, nrhs, batch, kappas, precision):
    import torch

    a = conditioned(rng, batch, n, kappas, spd=False)
    b = rng.uniform(-1, 1, (batch, n, nrhs))
    la, lb = Layout.make(n, n, panel=False), Layout.make(n, nrhs)
    ah, bh = la.pack(a), lb.pack(b)
    out = []
    for noalloc in (False, True):
        A, B, X = capi.Operand.from_host(la, ah), capi.Operand.from_host(lb, bh), capi.Operand.empty(lb, batch)
        ipiv = torch.full((batch * n,), -7, dtype=torch.int64, device="cuda")
        it = torch.full((batch,), -9, dtype=torch.int32, device="cuda")
        info = torch.full((batch,), -9, dtype=torch.int32, device="cuda")
        if noalloc:
            capi.sgesv_noalloc(precision, A, ipiv, n, B, X, batch, it, info, tol=1e-13)
        else:
            work = capi.mixed_work(n, batch)
            capi.sgesv(precision, A, ipiv, n, B, X, work, batch, it, info, tol=1e-13)
        torch.cuda.synchronize()
        out.append((X.host(), ipiv.cpu().numpy(), it.cpu().numpy(), info.cpu().numpy(), A.host()))
    for ref, got in zip(out[0], out[1]):
        np.testing.assert_array_equal(got, ref)
    x = X.logical(batch)
    assert np.all(solve_backward_error(a, x, b) <= 1e-12)
    return out[1][2]


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_equals_work_buffer_entry_on_two_calls(precision):
    rng = np.random.default_rng(77 + precision)
    its = _run_pair(rng, 200, 8, 32, np.geomspace(1e2, 1e6, 32), precision)
    assert np.all(its >= -2)
    # second call, smaller: the scratch holds the first call's factors
    _run_pair(rng, 96, 3, 8, np.geomspace(1e2, 1e4, 8), precision)


def test_noalloc_rejects_a_batch_beyond_the_scratch():
    import torch

    n = 256
    per = capi.mixed_work_elems(n)
    batch = capi.NOALLOC_SCRATCH_FLOATS // per + 1
    la, lb = Layout.make(n, n, panel=False), Layout.make(n, 1)
    A, B, X = capi.Operand.empty(la, batch), capi.Operand.empty(lb, batch), capi.Operand.empty(lb, batch)
    ipiv = torch.zeros(batch * n, dtype=torch.int64, device="cuda")
    it = torch.zeros(batch, dtype=torch.int32, device="cuda")
    with pytest.raises(capi.BlastError) as e:
        capi.sgesv_noalloc(0, A, ipiv, n, B, X, batch, it)
    assert e.value.code == -10
    assert "scratch" in str(e.value)
    # one instance less fits
    assert (batch - 1) * per <= capi.NOALLOC_SCRATCH_FLOATS
