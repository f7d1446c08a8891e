"""Uneven batched Cholesky solve: instances of different n (1 .. 192), any
storage, 0 .. 8 right-hand sides, in one persistent launch, against the
oracle's posv per instance; bad instances isolated; graph replay."""

import numpy as np
import pytest

from blast_amd import capi
from blast_amd.layout import Layout
from oracle import oracle
from tests.numerics import EPS64, potrf_backward_error, solve_backward_error, spd

pytestmark = pytest.mark.gpu


def _problem(seed, sizes, nrhs_of, panel_of):
    rng = np.random.default_rng(seed)
    la = [Layout.make(n, n, panel=panel_of(k)) for k, n in enumerate(sizes)]
    lb = [Layout.make(n, nrhs_of(k), panel=not panel_of(k)) for k, n in enumerate(sizes)]
    a = [spd(rng, 1, n)[0] for n in sizes]
    b = [rng.uniform(-1, 1, (n, lb[k].cols)) for k, n in enumerate(sizes)]
    return la, lb, a, b


def test_uneven_posv_vs_oracle():
    sizes = [1, 5, 16, 17, 31, 40, 64, 100, 128, 129, 144, 160, 177, 192, 33, 8]
    la, lb, a, b = _problem(1, sizes, lambda k: 1 + k % 8, lambda k: k % 2 == 0)
    u = capi.UnevenBatch(la, lb)
    u.pack(a, b)
    u.run()
    info = u.info.cpu().numpy()
    assert np.all(info[:len(sizes)] == 0), info
    # largest block count first
    nb = [(sizes[i] + 15) // 16 for i in u.order]
    assert nb == sorted(nb, reverse=True)
    for k, n in enumerate(sizes):
        l, x = u.result("l", k), u.result("x", k)
        ah, bh = la[k].pack(a[k]), lb[k].pack(b[k])
        lref, xref = np.zeros_like(ah), np.zeros_like(bh)
        oracle.posv(la[k], ah, la[k], lref, lb[k], bh, lb[k], xref, 1)
        low = np.tril(np.ones((n, n), bool))
        lr = la[k].unpack(lref, 1)[0]
        assert np.abs(l[low] - lr[low]).max() <= 1e-13 * np.abs(lr).max(), (k, n)
        assert potrf_backward_error(a[k][None], l[None])[0] < 4 * n * EPS64
        assert solve_backward_error(a[k][None], x[None], b[k][None])[0] < 4 * n * EPS64
        assert np.abs(x - lb[k].unpack(xref, 1)[0]).max() <= 1e-11 * max(1.0, np.abs(x).max())


def test_uneven_potrf_only_in_place_and_bad_instances():
    sizes = [48, 200, 20, 96]
    rng = np.random.default_rng(2)
    la = [Layout.make(n, n) for n in sizes]
    a = [spd(rng, 1, n)[0] for n in sizes]
    a[3] = a[3].copy()
    a[3][50, 50] = -1.0  # not SPD: first bad pivot at column 51
    u = capi.UnevenBatch(la, None, alias_l=True)
    u.pack(a)
    u.run()
    info = u.info.cpu().numpy()
    assert info[1] == -1  # n = 200 > 192: skipped
    assert info[3] == 51
    assert info[0] == 0 and info[2] == 0
    for k in (0, 2):
        l = u.result("l", k)
        assert potrf_backward_error(a[k][None], l[None])[0] < 4 * sizes[k] * EPS64
    assert capi.lib().blast_hip_uneven_max_n() == 192


def test_uneven_graph_replay():
    import torch

    sizes = [70, 12, 130, 64, 3]
    la, lb, a, b = _problem(3, sizes, lambda k: 4, lambda k: True)
    u = capi.UnevenBatch(la, lb)
    u.pack(a, b)
    u.run()
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        u.run(stream=s.cuda_stream)
        s.synchronize()
        with torch.cuda.graph(g, stream=s):
            u.run(stream=s.cuda_stream)
    torch.cuda.current_stream().wait_stream(s)
    rng = np.random.default_rng(4)
    for rep in range(3):
        b2 = [rng.uniform(-1, 1, (n, 4)) for n in sizes]
        u.pack(a, b2)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for k, n in enumerate(sizes):
            assert solve_backward_error(a[k][None], u.result("x", k)[None], b2[k][None])[0] < 4 * n * EPS64, (rep, k)


def test_uneven_posv_backward_vs_reference_adjoint():
    # the uneven adjoint (one launch) from the forward's L and X, against the
    # FP64 reference adjoint (oracle/adjoint.py, pinned by central
    # differences) and against the uniform batched backward of each instance
    from oracle import adjoint

    sizes = [1, 2, 15, 16, 17, 33, 64, 100, 128, 129, 150, 191, 192, 47, 80, 9]
    la, lb, a, b = _problem(7, sizes, lambda k: 1 + (3 * k) % 8, lambda k: k % 3 == 0)
    rng = np.random.default_rng(8)
    xb = [rng.standard_normal((n, lb[k].cols)) for k, n in enumerate(sizes)]
    u = capi.UnevenBatch(la, lb)
    u.pack(a, b)
    u.run()
    u.backward_buffers()
    u.pack_xbar(xb)
    u.backward()
    info = u.adj_info.cpu().numpy()
    assert np.all(info[:len(sizes)] == 0), info
    for k, n in enumerate(sizes):
        x = u.result("x", k)
        ab_ref, bb_ref = adjoint.posv_backward(np.tril(a[k]), x, xb[k])
        ab, bb = u.result("ab", k), u.result("bb", k)
        tol = 1e-10 * n
        assert np.abs(bb - bb_ref).max() <= tol * max(np.abs(bb_ref).max(), 1e-300), (k, n)
        assert np.abs(ab - ab_ref).max() <= tol * max(np.abs(ab_ref).max(), 1e-300), (k, n)
        assert np.all(ab[np.triu_indices(n, 1)] == 0.0)
        # the uniform entry point on this instance alone
        A1 = capi.Operand.from_host(la[k], la[k].pack(u.result("l", k)))
        X1 = capi.Operand.from_host(lb[k], lb[k].pack(x))
        XB1 = capi.Operand.from_host(lb[k], lb[k].pack(xb[k]))
        AB1, BB1 = capi.Operand.empty(la[k], 1), capi.Operand.empty(lb[k], 1)
        capi.posv_backward(0, A1, X1, XB1, AB1, BB1, 1)
        np.testing.assert_allclose(bb, BB1.logical(1)[0], rtol=0, atol=1e-12 * max(np.abs(bb).max(), 1e-300))
        np.testing.assert_allclose(ab, AB1.logical(1)[0], rtol=0, atol=1e-12 * max(np.abs(ab).max(), 1e-300))


def test_uneven_posv_backward_bad_descriptor_and_graph_replay():
    import torch

    sizes = [40, 200, 130, 7]
    la = [Layout.make(n, n) for n in sizes]
    lb = [Layout.make(n, 2) for n in sizes]
    rng = np.random.default_rng(9)
    a = [spd(rng, 1, n)[0] for n in sizes]
    b = [rng.uniform(-1, 1, (n, 2)) for n in sizes]
    xb = [rng.standard_normal((n, 2)) for n in sizes]
    u = capi.UnevenBatch(la, lb)
    u.pack(a, b)
    u.run()
    u.backward_buffers()
    u.pack_xbar(xb)
    u.backward()
    torch.cuda.synchronize()
    first = u.ab.clone()
    info = u.adj_info.cpu().numpy()
    assert info[1] == -1 and info[0] == 0 and info[2] == 0 and info[3] == 0
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        u.backward(stream=s.cuda_stream)
        s.synchronize()
        with torch.cuda.graph(g, stream=s):
            u.backward(stream=s.cuda_stream)
    torch.cuda.current_stream().wait_stream(s)
    for _ in range(3):
        u.ab.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(u.ab, first)
