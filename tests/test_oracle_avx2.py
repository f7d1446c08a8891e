"""The BLAST-restated AVX2 Cholesky solve (oracle/blast_avx2.cpp, the
benchmark's CPU baseline) against the scalar oracle: the same per-element fma
sequence, so L, X and info agree bit for bit, for every row-tile split of
BLAST's panel potrf (12 / 8 / 4-row tiles) and both storages of B."""

import numpy as np
import pytest

from blast_amd.layout import Layout
from oracle import oracle
from tests.numerics import spd


@pytest.mark.parametrize("n", [4, 8, 12, 16, 20, 24, 28, 36, 64, 100, 128])
@pytest.mark.parametrize("bpanel", [True, False])
def test_avx2_posv_equals_scalar_oracle(n, bpanel):
    rng = np.random.default_rng(n + 1000 * bpanel)
    batch, nrhs = 6, 4
    a = spd(rng, batch, n)
    b = rng.uniform(-1, 1, (batch, n, nrhs))
    la, lb = Layout.make(n, n), Layout.make(n, nrhs, panel=bpanel)
    ah, bh = la.pack(a), lb.pack(b)
    l0, x0 = np.zeros_like(ah), np.zeros_like(bh)
    info0 = oracle.posv(la, ah, la, l0, lb, bh, lb, x0, batch)
    l1, x1 = np.zeros_like(ah), np.zeros_like(bh)
    info1, used = oracle.posv_avx2_mt(la, ah, la, l1, lb, bh, lb, x1, batch, threads=3)
    assert used == 3
    np.testing.assert_array_equal(l1, l0)
    np.testing.assert_array_equal(x1, x0)
    np.testing.assert_array_equal(info1, info0)


def test_avx2_reports_the_first_bad_pivot():
    rng = np.random.default_rng(7)
    n, batch = 32, 3
    a = spd(rng, batch, n)
    a[1, 13, 13] = -1.0
    la, lb = Layout.make(n, n), Layout.make(n, 1)
    ah, bh = la.pack(a), lb.pack(rng.uniform(-1, 1, (batch, n, 1)))
    l, x = np.zeros_like(ah), np.zeros_like(bh)
    info, _ = oracle.posv_avx2_mt(la, ah, la, l, lb, bh, lb, x, batch)
    ref = oracle.potrf(la, ah, la, np.zeros_like(ah), batch)
    assert info[1] == ref[1] and info[1] > 0 and info[0] == 0 and info[2] == 0
