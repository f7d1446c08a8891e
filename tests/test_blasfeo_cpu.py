"""BLASFEO-format layer without a GPU: the panel-major index map, packing,
and the argument checks of the blast_hip_blasfeo_* entry points (every call
here is rejected before any device work)."""

import ctypes

import numpy as np
import pytest

from blast_amd import blasfeo, capi


def test_index_map_matches_panel_major_definition():
    # BLASFEO panel-major: panels of bs rows, each panel bs*cn elements,
    # column j of a panel at offset j*bs (blasfeo_dmat: BLASFEO_DMATEL)
    lay = blasfeo.BlasfeoLayout(11, 6, np.float64, cn=8)
    assert (lay.bs, lay.pm, lay.cn) == (4, 12, 8)
    seen = set()
    for i in range(11):
        for j in range(6):
            panel, r = divmod(i, 4)
            expect = panel * 4 * 8 + r + 4 * j
            assert lay.index(i, j) == expect
            seen.add(expect)
    assert len(seen) == 66 and max(seen) < lay.pm * lay.cn <= lay.batch_stride


@pytest.mark.parametrize("dtype,bs", [(np.float64, 4), (np.float32, 8)])
def test_pack_unpack_round_trip_and_padding(dtype, bs):
    rng = np.random.default_rng(1)
    lay = blasfeo.BlasfeoLayout(13, 5, dtype, cn=7)
    assert lay.bs == bs and lay.batch_stride % (64 // np.dtype(dtype).itemsize) == 0
    a = rng.standard_normal((3, 13, 5)).astype(dtype)
    buf = lay.pack(a, pad=-7.0)
    assert buf.size == 3 * lay.batch_stride
    np.testing.assert_array_equal(lay.unpack(buf, 3), a)
    # everything outside the m x n elements is padding
    assert np.count_nonzero(buf == -7.0) == buf.size - a.size


def _fake(lay, ptr=0x1000):
    return ctypes.byref(lay.descriptor(ptr))


def _gemm_nn(L, A, ai, B, bi, C, D, m=8, n=8, k=8, batch=2):
    return L.blast_hip_blasfeo_dgemm_nn(None, m, n, k, 1.0, A, ai, 0, B, bi, 0, 0.0, C, 0, 0, D, 0, 0, batch)


def test_arguments_are_checked_before_any_device_work():
    L = blasfeo._bind()
    good = blasfeo.BlasfeoLayout(8, 8, np.float64)
    g = _fake(good)
    # panel height other than 4 rows for double
    bad_bs = blasfeo.BlasfeoLayout(8, 8, np.float64, bs=8)
    assert _gemm_nn(L, _fake(bad_bs), 0, g, 0, g, g) == -6
    assert "panel height" in capi.last_error()
    # row offset not a multiple of the panel height
    big = blasfeo.BlasfeoLayout(16, 8, np.float64)
    assert _gemm_nn(L, _fake(big), 2, g, 0, g, g) == -6
    assert "multiple" in capi.last_error()
    # block beyond the matrix (B is 8 x 8, offset 4 leaves 4 rows)
    assert _gemm_nn(L, g, 0, g, 4, g, g) == -9
    # negative size, negative batch, null descriptor
    assert _gemm_nn(L, g, 0, g, 0, g, g, m=-1) == -2
    assert _gemm_nn(L, g, 0, g, 0, g, g, batch=-1) == -19
    assert _gemm_nn(L, g, 0, g, 0, None, g) == -13
    # cn smaller than n
    d = good.descriptor(0x1000)
    d.cn = 4
    assert _gemm_nn(L, g, 0, g, 0, g, ctypes.byref(d)) == -16


def test_float_panel_height_and_trmm_overlap():
    L = blasfeo._bind()
    f4 = blasfeo.BlasfeoLayout(8, 8, np.float32, bs=4)
    f8 = blasfeo.BlasfeoLayout(8, 8, np.float32)
    r = L.blast_hip_blasfeo_spotrf_l(None, 8, _fake(f4), 0, 0, _fake(f8), 0, 0, 1, None)
    assert r == -3 and "8 rows" in capi.last_error()
    d8 = blasfeo.BlasfeoLayout(8, 8, np.float64)
    # D may not be B for trmm (the kernel reads B while writing D)
    r = L.blast_hip_blasfeo_dtrmm_rlnn(None, 8, 8, 1.0, _fake(d8, 0x2000), 0, 0, _fake(d8), 0, 0, _fake(d8), 0, 0, 1)
    assert r == -11 and "overlap" in capi.last_error()
    # syrk_potrf: sD block outside its matrix
    r = L.blast_hip_blasfeo_dsyrk_dpotrf_ln(None, 8, 8, _fake(d8), 0, 0, _fake(d8), 0, 0, _fake(d8), 0, 0,
                                            _fake(d8), 4, 0, 1, None)
    assert r == -13
