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


def test_riccati_arguments_and_tuning_file_are_checked(tmp_path):
    """Riccati entry limits (horizon, count: the int32 work counter) and the
    tuning table: null tables and out-of-range strip widths in a tuning file
    are rejected (257 must not narrow to 1)."""
    lib = capi.lib()
    assert lib.blast_hip_riccati_set_tuning(None) == -1
    assert lib.blast_hip_riccati_get_tuning(None) == -1
    fake = ctypes.c_void_p(0x1000)
    assert lib.blast_hip_driccati_batched(None, 40000, fake, 1, fake, fake, fake, fake, None) == -2
    assert lib.blast_hip_driccati_batched(None, 10, fake, 0x7fffffff - 10, fake, fake, fake, fake, None) == -4
    saved = (ctypes.c_int8 * 16)()
    assert lib.blast_hip_riccati_get_tuning(saved) == 0
    path = tmp_path / "t.tuning"
    assert lib.blast_hip_save_tuning(str(path).encode()) == 0
    text = path.read_text().split("\n")
    widths = text[1].split()
    assert widths[0] == "riccati_strip_width" and len(widths) == 17
    bad = tmp_path / "bad.tuning"
    bad.write_text(text[0] + "\nriccati_strip_width " + " ".join(["257"] + widths[2:]) + "\n")
    assert lib.blast_hip_load_tuning(str(bad).encode()) == -1
    after = (ctypes.c_int8 * 16)()
    assert lib.blast_hip_riccati_get_tuning(after) == 0
    assert bytes(after) == bytes(saved)
    assert lib.blast_hip_load_tuning(str(path).encode()) == 0


def test_subblock_alignment_rule():
    from blast_amd.layout import Layout

    la = Layout.make(40, 30)
    with pytest.raises(ValueError):
        la.submatrix(2, 0, 4, 4)  # row 2 splits a panel
    with pytest.raises(ValueError):
        la.submatrix(36, 0, 8, 4)
    la.t().submatrix(0, 4, 5, 8)  # transposed view: the storage row is the logical column
    with pytest.raises(ValueError):
        la.t().submatrix(0, 6, 5, 8)
