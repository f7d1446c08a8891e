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
