"""The dense configurations at their full batch sizes, through the C ABI:
dgemm 64^3 NN in panel storage (batch 4096), dposv n = 128 with 4 RHS
(batch 8192), dposv and dgesv n = 256 with 8 RHS (batch 2048).  Inputs are
generated on the GPU (torch Philox, no torch linear algebra); every instance
is checked through a size-independent property (normwise backward error,
agreement with an FP64 product), and a sample of instances against the oracle
(the L factor, the pivot sequence)."""

import numpy as np
import pytest

from blast_amd import capi
from blast_amd.layout import Layout
from oracle import oracle
from tests.numerics import EPS64

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def auto_algorithm():
    capi.set_algorithm(capi.ALGO_AUTO)


def _pack(layout, x):
    """Logical (batch, rows, cols) device tensor -> flat device buffer in layout's storage."""
    import torch

    idx = torch.from_numpy(layout.index_map().reshape(-1)).to(x.device)
    buf = torch.zeros(x.shape[0] * layout.batch_stride, dtype=x.dtype, device=x.device)
    buf.view(x.shape[0], layout.batch_stride).index_copy_(1, idx, x.reshape(x.shape[0], -1))
    return buf


def _unpack(layout, buf, batch):
    import torch

    idx = torch.from_numpy(layout.index_map().reshape(-1)).to(buf.device)
    return buf.view(batch, layout.batch_stride)[:, idx].reshape(batch, layout.rows, layout.cols)


def _fro(x):
    import torch

    return torch.linalg.vector_norm(x, dim=(-2, -1))


def _spd(gen, batch, n):
    """A = G G^T + n I, G ~ N(0, 1), FP64 on the GPU."""
    import torch

    g = torch.randn(batch, n, n, dtype=torch.float64, device="cuda", generator=gen)
    return torch.bmm(g, g.transpose(1, 2).contiguous()) + n * torch.eye(n, dtype=torch.float64, device="cuda")


def _solve_backward_error(a, x, b):
    import torch

    return _fro(b - torch.bmm(a, x)) / (_fro(a) * _fro(x) + _fro(b))


def test_gemm_64_nn_panel_batch_4096():
    import torch

    batch, n = 4096, 64
    gen = torch.Generator(device="cuda").manual_seed(101)
    a, b, c = (torch.randn(batch, n, n, dtype=torch.float64, device="cuda", generator=gen) for _ in range(3))
    lay = Layout.make(n, n)
    A, B, C = (capi.Operand(lay, _pack(lay, m)) for m in (a, b, c))
    D = capi.Operand.empty(lay, batch, fill=np.nan)
    capi.gemm(1.5, A, B, -0.5, C, D, batch)
    torch.cuda.synchronize()
    d = _unpack(lay, D.tensor, batch)
    ref = 1.5 * torch.bmm(a, b) - 0.5 * c
    # each entry is a 64-term FP64 dot product: error <= ~64 eps |a||b| per term
    scale = 1.5 * torch.bmm(a.abs(), b.abs()) + 0.5 * c.abs()
    assert bool(torch.all((d - ref).abs() <= 2 * n * EPS64 * scale))


def test_posv_128_nrhs4_batch_8192():
    import torch

    batch, n, nrhs = 8192, 128, 4
    gen = torch.Generator(device="cuda").manual_seed(102)
    a = _spd(gen, batch, n)
    b = torch.rand(batch, n, nrhs, dtype=torch.float64, device="cuda", generator=gen) * 2 - 1
    la, lb = Layout.make(n, n), Layout.make(n, nrhs)
    A, B = capi.Operand(la, _pack(la, a)), capi.Operand(lb, _pack(lb, b))
    L, X = capi.Operand.empty(la, batch, fill=3.0), capi.Operand.empty(lb, batch, fill=np.nan)
    info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    capi.posv(A, L, B, X, batch, info)
    torch.cuda.synchronize()
    assert int((info != 0).sum()) == 0
    bwd = _solve_backward_error(a, _unpack(lb, X.tensor, batch), b)
    assert float(bwd.max()) <= 1e-12
    # L of a sample of instances vs the oracle's factor
    pick = [0, 1, 4095, 8191]
    ah = A.tensor.view(batch, la.batch_stride)[pick].reshape(-1).cpu().numpy()
    lh = np.full_like(ah, 3.0)
    oracle.potrf(la, ah, la, lh, len(pick))
    lg = L.tensor.view(batch, la.batch_stride)[pick].reshape(-1).cpu().numpy()
    l_ref, l_got = la.unpack(lh, len(pick)), la.unpack(lg, len(pick))
    low = np.tril(np.ones((n, n), bool))
    np.testing.assert_allclose(l_got[:, low], l_ref[:, low], rtol=0, atol=1e-13 * np.abs(l_ref).max())
    np.testing.assert_array_equal(l_got[:, ~low], l_ref[:, ~low])


def test_posv_256_nrhs8_batch_2048():
    import torch

    batch, n, nrhs = 2048, 256, 8
    gen = torch.Generator(device="cuda").manual_seed(103)
    a = _spd(gen, batch, n)
    b = torch.rand(batch, n, nrhs, dtype=torch.float64, device="cuda", generator=gen) * 2 - 1
    la, lb = Layout.make(n, n), Layout.make(n, nrhs)
    A, B = capi.Operand(la, _pack(la, a)), capi.Operand(lb, _pack(lb, b))
    L, X = capi.Operand.empty(la, batch), capi.Operand.empty(lb, batch, fill=np.nan)
    info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    capi.posv(A, L, B, X, batch, info)
    torch.cuda.synchronize()
    assert int((info != 0).sum()) == 0
    bwd = _solve_backward_error(a, _unpack(lb, X.tensor, batch), b)
    assert float(bwd.max()) <= 1e-12


def test_gesv_256_nrhs8_batch_2048():
    import torch

    batch, n, nrhs = 2048, 256, 8
    gen = torch.Generator(device="cuda").manual_seed(104)
    a = torch.rand(batch, n, n, dtype=torch.float64, device="cuda", generator=gen) * 2 - 1
    a += 0.5 * torch.eye(n, dtype=torch.float64, device="cuda")
    b = torch.rand(batch, n, nrhs, dtype=torch.float64, device="cuda", generator=gen) * 2 - 1
    la, lb = Layout.make(n, n, panel=False), Layout.make(n, nrhs)
    A, B = capi.Operand(la, _pack(la, a)), capi.Operand(lb, _pack(lb, b))
    X = capi.Operand.empty(lb, batch, fill=np.nan)
    pick = [0, 777, 2047]
    ah = A.tensor.view(batch, la.batch_stride)[pick].reshape(-1).cpu().numpy()
    ipiv = torch.zeros(batch * n, dtype=torch.int64, device="cuda")
    info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    capi.gesv(A, ipiv, n, B, X, batch, info)
    torch.cuda.synchronize()
    assert int((info != 0).sum()) == 0
    bwd = _solve_backward_error(a, _unpack(lb, X.tensor, batch), b)
    # partial pivoting: the backward error grows with the pivot growth, which
    # is modest for these matrices
    assert float(bwd.max()) <= 1e-12
    piv_ref, _ = oracle.getrf(la, ah, len(pick))
    np.testing.assert_array_equal(ipiv.view(batch, n)[pick].cpu().numpy(), piv_ref)
