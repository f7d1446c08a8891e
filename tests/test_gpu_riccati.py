"""Uneven batched Riccati recursion on the GPU (one persistent launch) vs the
FP64 recursion of oracle/riccati.py: gains K_k and P_0 = L_0 L_0^T to
relative error 1e-10 (FP64 square-root form; the textbook recursion and the
square-root recursion differ by rounding only, O(nx * eps * cond))."""

import numpy as np
import pytest

from blast_amd import capi
from oracle import riccati

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _check(batch, A, B, Q, R, idx=None):
    K, L0 = batch.results()
    info = batch.info.cpu().numpy()
    for i in range(batch.count) if idx is None else idx:
        Kr, Pr = riccati.riccati(A[i], B[i], Q[i], R[i], batch.horizon)
        assert info[i] == 0, (i, info[i])
        assert np.abs(K[i] - Kr).max() <= TOL * max(1.0, np.abs(Kr).max()), (i, batch.inst[i])
        assert np.allclose(L0[i], np.tril(L0[i]), rtol=0, atol=0)
        P = L0[i] @ L0[i].T
        assert np.abs(P - Pr).max() <= TOL * np.abs(Pr).max(), (i, batch.inst[i])


def _batch(nx, nu, horizon, seed):
    rng = np.random.default_rng(seed)
    A, B, Q, R = [], [], [], []
    for n, m in zip(nx, nu):
        s = 1.0 / np.sqrt(n)
        A.append(rng.normal(0, s, (n, n)))
        B.append(rng.normal(0, s, (n, m)))
        W = rng.standard_normal((n, n))
        Q.append(np.eye(n) + 0.1 * W @ W.T)
        R.append(np.eye(m) + 0.3 * np.diag(rng.uniform(size=m)))
    b = capi.RiccatiBatch(nx, nu, horizon)
    # the kernel reads only the lower triangles of Q and R: garbage above
    Qg = [q + np.triu(rng.standard_normal(q.shape), 1) * 1e3 for q in Q]
    Rg = [r + np.triu(rng.standard_normal(r.shape), 1) * 1e3 for r in R]
    b.pack(A, B, Qg, Rg)
    return b, A, B, Q, R


@pytest.mark.parametrize("tn", [1, 2, 4])
def test_riccati_uneven_sizes(tn):
    old = capi.riccati_get_tuning()
    capi.riccati_set_tuning([tn] * 16)
    try:
        nx = [1, 2, 3, 4, 5, 15, 16, 17, 31, 33, 48, 64, 100, 129, 200, 200, 7, 63]
        nu = [1, 1, 2, 2, 1, 7, 8, 1, 15, 16, 24, 32, 50, 64, 100, 1, 3, 31]
        b, A, B, Q, R = _batch(nx, nu, 4, seed=tn)
        b.run()
        _check(b, A, B, Q, R)
    finally:
        capi.riccati_set_tuning(old)


def test_riccati_workload_sample():
    rng = np.random.default_rng(20240611)
    nx, nu, A, B, Q, R = riccati.make_instances(rng, 96)
    b = capi.RiccatiBatch(nx, nu, 10)
    b.pack(A, B, Q, R)
    b.run()
    _check(b, A, B, Q, R)


def test_riccati_bad_instance_isolated():
    nx, nu = [12, 20, 9], [3, 5, 4]
    b, A, B, Q, R = _batch(nx, nu, 3, seed=7)
    Qbad = list(Q)
    Qbad[1] = Q[1].copy()
    Qbad[1][4, 4] = -5.0  # Q not positive definite: chol(Q) fails at column 5
    b.pack(A, B, Qbad, R)
    b.run()
    info = b.info.cpu().numpy()
    assert info[1] == 5
    _check(b, A, B, Q, R, idx=[0, 2])


def test_riccati_info_names_the_step_of_a_failed_s():
    """S = R + B^T P B not positive definite (R = -100 I) in instance 1: the
    Lam factorisation of the first step (k = horizon - 1) fails at column 1,
    info = 65536 (k + 1) + 1; the other instances are unaffected."""
    nx, nu = [12, 20, 9], [3, 5, 4]
    horizon = 3
    b, A, B, Q, R = _batch(nx, nu, horizon, seed=8)
    Rbad = list(R)
    Rbad[1] = -100.0 * np.eye(nu[1])
    b.pack(A, B, Q, Rbad)
    b.run()
    info = b.info.cpu().numpy()
    assert info[1] == 65536 * horizon + 1, info
    assert info[0] == 0 and info[2] == 0
    _check(b, A, B, Q, R, idx=[0, 2])


def test_riccati_graph_replay():
    import torch

    nx, nu = [40, 8, 130, 64], [20, 3, 10, 32]
    b, A, B, Q, R = _batch(nx, nu, 5, seed=9)
    b.run()
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        b.run(stream=s.cuda_stream)  # warm-up on the capture stream
        s.synchronize()
        with torch.cuda.graph(g, stream=s):
            b.run(stream=s.cuda_stream)
    torch.cuda.current_stream().wait_stream(s)
    rng = np.random.default_rng(10)
    for rep in range(3):
        Qr = [q * (1.0 + 0.5 * rep) for q in Q]
        Ar = [a * (1.0 - 0.1 * rep) for a in A]
        b.output.fill_(np.nan)
        b.pack(Ar, B, Qr, R)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        _ = rng.standard_normal(4)  # host work between replays
        _check(b, Ar, B, Qr, R)


def test_riccati_autotune_and_tuning_file(tmp_path):
    rng = np.random.default_rng(12)
    nx, nu, A, B, Q, R = riccati.make_instances(rng, 64)
    b = capi.RiccatiBatch(nx, nu, 3)
    b.pack(A, B, Q, R)
    old = capi.riccati_get_tuning()
    try:
        ms = b.autotune(reps=2)
        t = capi.riccati_get_tuning()
        assert set(np.unique(t)) <= {1, 2, 4}
        buckets = {min(int(n) // 16, 15) for n in nx}
        for k in buckets:
            assert np.all(ms[k] > 0)
            assert t[k] == (1, 2, 4)[int(np.argmin(ms[k]))]
        path = tmp_path / "tuning.txt"
        capi.save_tuning(path)
        capi.riccati_set_tuning([1] * 16)
        capi.load_tuning(path)
        assert np.array_equal(capi.riccati_get_tuning(), t)
        b.run()
        _check(b, A, B, Q, R)
    finally:
        capi.riccati_set_tuning(old)
