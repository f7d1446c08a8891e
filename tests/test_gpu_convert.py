"""Storage conversion D := S between every pair of layouts (dense / panel,
plain / transposed view), FP64 and FP32: a pure element copy, so the result
must equal the host packing of the same logical matrix bit for bit, and D's
storage outside the logical matrix must be untouched."""

import itertools

import numpy as np
import pytest

from blast_amd import capi
from blast_amd.layout import Layout

pytestmark = pytest.mark.gpu

KINDS = list(itertools.product([False, True], [False, True]))  # (panel, trans)


def lay(rows, cols, dtype, panel, trans):
    return Layout.make(cols, rows, dtype, panel=panel).t() if trans else Layout.make(rows, cols, dtype, panel=panel)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("src", KINDS)
@pytest.mark.parametrize("dst", KINDS)
@pytest.mark.parametrize("m,n", [(1, 1), (37, 53), (64, 64), (130, 7)])
def test_copy_all_layout_pairs(dtype, src, dst, m, n):
    rng = np.random.default_rng(m * 1000 + n)
    batch = 3
    x = rng.standard_normal((batch, m, n)).astype(dtype)
    ls, ld = lay(m, n, dtype, *src), lay(m, n, dtype, *dst)
    S = capi.Operand.from_host(ls, ls.pack(x))
    D = capi.Operand.empty(ld, batch, fill=7.0)
    capi.copy(S, D, batch)
    np.testing.assert_array_equal(D.logical(batch), x)
    expect = np.full(batch * ld.batch_stride, 7.0, dtype=dtype)
    idx = ld.index_map()
    for k in range(batch):
        expect[k * ld.batch_stride + idx] = x[k]
    np.testing.assert_array_equal(D.host(), expect)


def test_copy_rejects_size_mismatch():
    la, lb = Layout.make(8, 8), Layout.make(8, 9)
    with pytest.raises(capi.BlastError) as e:
        capi.copy(capi.Operand.empty(la, 1), capi.Operand.empty(lb, 1), 1)
    assert e.value.code == -3
