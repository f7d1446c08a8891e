"""C ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and rejects bad arguments before touching the device."""

import ctypes

import numpy as np
import pytest

from blast_amd import capi
from blast_amd.layout import Layout


def test_exports_every_declared_symbol():
    lib = capi.lib()
    names = capi.declared_symbols()
    assert len(names) >= 17
    for name in names:
        assert hasattr(lib, name), name


def test_version_and_algorithm_switch():
    assert b"gfx950" in capi.lib().blast_hip_version()
    capi.set_algorithm(capi.ALGO_REFERENCE)
    assert capi.lib().blast_hip_get_algorithm() == capi.ALGO_REFERENCE
    capi.set_algorithm(capi.ALGO_AUTO)
    with pytest.raises(capi.BlastError):
        capi.set_algorithm(7)


def _d(rows, cols, ptr=0x1000, panel=True, spacing=None, trans=False):
    lay = Layout.make(rows, cols, np.float64, panel=panel, trans=trans)
    d = lay.descriptor(ptr)
    if spacing is not None:
        d.spacing = spacing
    return ctypes.byref(d)


def test_gemm_size_mismatch_is_rejected():
    lib = capi.lib()
    # B rows != A cols  (algorithm/Gemm.hpp:87-88)
    r = lib.blast_hip_dgemm_batched(None, 1.0, _d(4, 3), _d(4, 5), 1.0, _d(4, 5), _d(4, 5), 1)
    assert r == -4 and "do not match" in capi.last_error()
    r = lib.blast_hip_dgemm_batched(None, 1.0, _d(4, 3), _d(3, 5), 1.0, _d(4, 6), _d(4, 5), 1)
    assert r == -6
    r = lib.blast_hip_dgemm_batched(None, 1.0, _d(4, 3), _d(3, 5), 1.0, _d(4, 5), _d(5, 5), 1)
    assert r == -7
    r = lib.blast_hip_dgemm_batched(None, 1.0, _d(4, 3), _d(3, 5), 1.0, _d(4, 5), _d(4, 5), -1)
    assert r == -8


def test_bad_storage_is_rejected():
    lib = capi.lib()
    # panel spacing must be >= PS * columns
    r = lib.blast_hip_dpotrf_batched(None, _d(8, 8, spacing=8), _d(8, 8), 1, None)
    assert r == -2 and "spacing" in capi.last_error()
    # null data with a nonempty matrix
    r = lib.blast_hip_dpotrf_batched(None, _d(8, 8, ptr=0), _d(8, 8), 1, None)
    assert r == -2
    # potrf: columns > rows (dense/Potrf.hpp:78-79)
    r = lib.blast_hip_dpotrf_batched(None, _d(4, 8), _d(4, 8), 1, None)
    assert r == -2 and "Invalid matrix size" in capi.last_error()
    # trmm: unimplemented side/triangle combination (algorithm/Trmm.hpp:124)
    r = lib.blast_hip_dtrmm_batched(None, 0, 0, 1.0, _d(4, 4), _d(4, 4), _d(4, 4), 1)
    assert r == -3
    # trsm size check (dense/Trsm.hpp:244-251)
    r = lib.blast_hip_dtrsm_batched(None, 0, 0, 0, 1.0, _d(5, 5), _d(4, 2), _d(4, 2), 1)
    assert r == -6
    # getrf needs room for min(M, N) pivots per instance
    r = lib.blast_hip_dgetrf_batched(None, _d(6, 6, panel=False), ctypes.c_void_p(0x2000), 5, 1, None)
    assert r == -4


def test_empty_batches_are_noops():
    lib = capi.lib()
    assert lib.blast_hip_dgemm_batched(None, 1.0, _d(4, 3), _d(3, 5), 1.0, _d(4, 5), _d(4, 5), 0) == 0
