"""BLASFEO-format entry points on the GPU, through the C ABI, against the
mathematics of each BLASFEO routine (numpy in FP64): gemm_{nn,nt,tn,tt},
syrk_ln, potrf_l, syrk_potrf_ln, trmm_rlnn / trmm_rutn, on sub-blocks at
row offsets that are multiples of the panel height and odd column offsets,
with padded cn; plus the storage contract (nothing outside the written part
changes, the unread triangle of a trmm operand is never read)."""

import numpy as np
import pytest

from blast_amd import blasfeo

pytestmark = pytest.mark.gpu

SENTINEL = 5.0


def tol(dtype, n):
    return (200 if dtype == np.float32 else 4) * n * np.finfo(dtype).eps


def big(rng, batch, rows, cols, dtype, cn_extra=3, fill=None):
    lay = blasfeo.BlasfeoLayout(rows, cols, dtype, cn=cols + cn_extra)
    a = rng.uniform(-1, 1, (batch, rows, cols)) if fill is None else np.full((batch, rows, cols), fill)
    return lay, a.astype(dtype)


def dev(lay, a):
    return blasfeo.Operand.from_host(lay, a, pad=-3.0)


def spd(rng, batch, n):
    m = rng.uniform(-1, 1, (batch, n, n))
    return m @ m.transpose(0, 2, 1) + n * np.eye(n)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ta,tb", [(False, False), (False, True), (True, False), (True, True)])
@pytest.mark.parametrize("m,n,k", [(13, 10, 7), (40, 36, 20)])
def test_gemm(dtype, ta, tb, m, n, k):
    rng = np.random.default_rng(m * 100 + n + 7 * ta + 3 * tb)
    batch, bs = 3, blasfeo.panel_size(dtype)
    ai, aj, bi, bj, ci, cj, di, dj = bs, 3, 0, 1, bs, 2, 2 * bs, 5
    sa = (k, m) if ta else (m, k)
    sb = (n, k) if tb else (k, n)
    la, a = big(rng, batch, ai + sa[0] + 1, aj + sa[1] + 2, dtype)
    lb, b = big(rng, batch, bi + sb[0] + 3, bj + sb[1], dtype)
    lc, c = big(rng, batch, ci + m, cj + n + 1, dtype)
    ld, d0 = big(rng, batch, di + m + 2, dj + n + 2, dtype, fill=SENTINEL)
    A, B, C, D = dev(la, a), dev(lb, b), dev(lc, c), dev(ld, d0)
    alpha, beta = 0.75, -1.25
    blasfeo.gemm(ta, tb, m, n, k, alpha, A, ai, aj, B, bi, bj, beta, C, ci, cj, D, di, dj, batch)
    opa = a[:, ai:ai + sa[0], aj:aj + sa[1]].astype(np.float64)
    opb = b[:, bi:bi + sb[0], bj:bj + sb[1]].astype(np.float64)
    opa = opa.transpose(0, 2, 1) if ta else opa
    opb = opb.transpose(0, 2, 1) if tb else opb
    ref = beta * c[:, ci:ci + m, cj:cj + n] + alpha * opa @ opb
    got = D.logical(batch)
    blk = got[:, di:di + m, dj:dj + n]
    scale = np.abs(alpha) * np.abs(opa) @ np.abs(opb) + np.abs(beta * c[:, ci:ci + m, cj:cj + n])
    assert np.all(np.abs(blk - ref) <= tol(dtype, k) * (scale + 1e-30))
    outside = np.ones(got.shape[1:], bool)
    outside[di:di + m, dj:dj + n] = False
    assert np.all(got[:, outside] == SENTINEL)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("m,k", [(13, 9), (33, 17)])
def test_syrk_ln(dtype, m, k):
    rng = np.random.default_rng(m + k)
    batch, bs = 3, blasfeo.panel_size(dtype)
    la, a = big(rng, batch, bs + m, 2 + k, dtype)
    lb, b = big(rng, batch, m, 1 + k, dtype)
    lc, c = big(rng, batch, m + 1, m + 3, dtype)
    ld, d0 = big(rng, batch, bs + m, m + 4, dtype, fill=SENTINEL)
    A, B, C, D = dev(la, a), dev(lb, b), dev(lc, c), dev(ld, d0)
    blasfeo.syrk_ln(m, k, 2.0, A, bs, 2, B, 0, 1, 0.5, C, 0, 3, D, bs, 4, batch)
    opa = a[:, bs:bs + m, 2:2 + k].astype(np.float64)
    opb = b[:, 0:m, 1:1 + k].astype(np.float64)
    ref = 0.5 * c[:, 0:m, 3:3 + m] + 2.0 * opa @ opb.transpose(0, 2, 1)
    scale = 2.0 * np.abs(opa) @ np.abs(opb).transpose(0, 2, 1) + 0.5 * np.abs(c[:, 0:m, 3:3 + m])
    got = D.logical(batch)
    blk = got[:, bs:bs + m, 4:4 + m]
    low = np.tril(np.ones((m, m), bool))
    assert np.all(np.abs(blk - ref)[:, low] <= tol(dtype, k) * (scale[:, low] + 1e-30))
    assert np.all(blk[:, ~low] == SENTINEL)  # strict upper part of D not written
    outside = np.ones(got.shape[1:], bool)
    outside[bs:bs + m, 4:4 + m] = False
    assert np.all(got[:, outside] == SENTINEL)


def check_cholesky(dtype, s, lg, m):
    """lg: the D block after potrf (lower part = L); s: the symmetric input."""
    low = np.tril(np.ones((m, m), bool))
    L = np.where(low, lg, 0).astype(np.float64)
    err = np.abs(L @ L.transpose(0, 2, 1) - s).max(axis=(1, 2)) / np.abs(s).max(axis=(1, 2))
    assert np.all(err < tol(dtype, m)), err
    assert np.all(np.diagonal(L, axis1=1, axis2=2) > 0)


def upper_contract(blk, m, bs):
    """Strict upper part of the D block: untouched, except 0 allowed inside
    the bs x bs diagonal blocks (the panel potrf contract)."""
    r, c = np.meshgrid(np.arange(m), np.arange(m), indexing="ij")
    up = r < c
    diag_blk = up & (r // bs == c // bs)
    assert np.all(blk[:, up & ~diag_blk] == SENTINEL)
    assert np.all((blk[:, diag_blk] == SENTINEL) | (blk[:, diag_blk] == 0))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("m", [13, 64, 100, 128, 150])
def test_potrf_l(dtype, m):
    import torch

    rng = np.random.default_rng(m)
    batch, bs = 4, blasfeo.panel_size(dtype)
    s = spd(rng, batch, m)
    lc = blasfeo.BlasfeoLayout(bs + m, m + 3, dtype, cn=m + 5)
    c = np.full((batch, bs + m, m + 3), -9.0)
    c[:, bs:, 3:] = np.where(np.tril(np.ones((m, m), bool)), s, 77.0)  # upper part of C is not read
    ld, d0 = big(rng, batch, m, m + 1, dtype, fill=SENTINEL)
    C, D = dev(lc, c.astype(dtype)), dev(ld, d0)
    info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    blasfeo.potrf_l(m, C, bs, 3, D, 0, 1, batch, info)
    assert np.all(info.cpu().numpy() == 0)
    got = D.logical(batch)
    check_cholesky(dtype, s, got[:, :, 1:1 + m], m)
    upper_contract(got[:, :, 1:1 + m], m, bs)
    assert np.all(got[:, :, 0] == SENTINEL)


@pytest.mark.parametrize("m", [16, 128])
def test_potrf_l_in_place_and_info(m):
    import torch

    rng = np.random.default_rng(5)
    batch = 3
    s = spd(rng, batch, m)
    s[1, m // 2, m // 2] = -1e3  # instance 1 is not positive definite at column m/2
    lay = blasfeo.BlasfeoLayout(m, m, np.float64, cn=m + 4)
    C = dev(lay, s)
    info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    blasfeo.potrf_l(m, C, 0, 0, C, 0, 0, batch, info)
    inf = info.cpu().numpy()
    assert inf[0] == 0 and inf[2] == 0 and inf[1] == m // 2 + 1
    got = C.logical(batch)
    check_cholesky(np.float64, s[[0, 2]], got[[0, 2]], m)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("m,k,same", [(20, 12, True), (64, 30, False)])
def test_syrk_potrf_ln(dtype, m, k, same):
    import torch

    rng = np.random.default_rng(m * k)
    batch, bs = 3, blasfeo.panel_size(dtype)
    la, a = big(rng, batch, m, k + 1, dtype)
    b = a if same else (0.2 * rng.uniform(-1, 1, a.shape)).astype(dtype)
    lb = la
    c = spd(rng, batch, m)
    lc = blasfeo.BlasfeoLayout(m, m, dtype, cn=m + 2)
    ld, d0 = big(rng, batch, bs + m, m, dtype, fill=SENTINEL)
    A, C, D = dev(la, a), dev(lc, c.astype(dtype)), dev(ld, d0)
    B = A if same else dev(lb, b)
    info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    blasfeo.syrk_potrf_ln(m, k, A, 0, 1, B, 0, 1, C, 0, 0, D, bs, 0, batch, info)
    assert np.all(info.cpu().numpy() == 0)
    t = c.astype(dtype).astype(np.float64) + a[:, :, 1:].astype(np.float64) @ b[:, :, 1:].astype(np.float64).transpose(0, 2, 1)
    low = np.tril(t)
    s = low + np.tril(t, -1).transpose(0, 2, 1)  # potrf reads the lower triangle only
    got = D.logical(batch)
    check_cholesky(dtype, s, got[:, bs:, :], m)
    upper_contract(got[:, bs:, :], m, bs)
    assert np.all(got[:, :bs, :] == SENTINEL)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("upper_t", [False, True])
@pytest.mark.parametrize("m,n", [(13, 9), (33, 20)])
def test_trmm(dtype, upper_t, m, n):
    rng = np.random.default_rng(m * n + upper_t)
    batch, bs = 3, blasfeo.panel_size(dtype)
    la, a = big(rng, batch, bs + n, n + 2, dtype)
    tri = np.triu(np.ones((n, n), bool)) if upper_t else np.tril(np.ones((n, n), bool))
    blk = a[:, bs:bs + n, 2:2 + n]
    blk[:, ~tri] = np.nan  # the other triangle must not be read
    lb, b = big(rng, batch, m, n + 1, dtype)
    ld, d0 = big(rng, batch, m + 1, n + 3, dtype, fill=SENTINEL)
    A, B, D = dev(la, a), dev(lb, b), dev(ld, d0)
    blasfeo.trmm(upper_t, m, n, -1.5, A, bs, 2, B, 0, 1, D, 0, 3, batch)
    at = np.where(tri, a[:, bs:bs + n, 2:2 + n], 0).astype(np.float64)
    at = at.transpose(0, 2, 1) if upper_t else at
    bb = b[:, :m, 1:1 + n].astype(np.float64)
    ref = -1.5 * bb @ at
    scale = 1.5 * np.abs(bb) @ np.abs(at)
    got = D.logical(batch)
    assert np.all(np.abs(got[:, :m, 3:3 + n] - ref) <= tol(dtype, n) * (scale + 1e-30))
    outside = np.ones(got.shape[1:], bool)
    outside[:m, 3:3 + n] = False
    assert np.all(got[:, outside] == SENTINEL)
