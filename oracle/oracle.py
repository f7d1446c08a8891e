"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's CPU-baseline leg, always as the checker.  The product library never
routes through it.

All functions take numpy arrays in the flat storage format described by a
``blast_amd.layout.Layout`` and operate on ``batch`` instances.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from blast_amd.layout import BlastMatrix, Layout

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        M = ctypes.POINTER(BlastMatrix)
        i64, i32p = ctypes.c_int64, ctypes.POINTER(ctypes.c_int32)
        for p, real in (("d", ctypes.c_double), ("s", ctypes.c_float)):
            f = getattr(_lib, f"oracle_{p}gemm"); f.argtypes = [real, M, M, real, M, M, i64]
            f = getattr(_lib, f"oracle_{p}syrk_lower"); f.argtypes = [real, M, real, M, M, i64]
            f = getattr(_lib, f"oracle_{p}trmm_left_upper"); f.argtypes = [real, M, M, M, i64]
            f = getattr(_lib, f"oracle_{p}trmm_right_lower"); f.argtypes = [real, M, M, M, i64]
            for side in ("left", "right"):
                f = getattr(_lib, f"oracle_{p}trsm_{side}")
                f.argtypes = [ctypes.c_int, ctypes.c_int, real, M, M, M, i64]
            f = getattr(_lib, f"oracle_{p}potrf"); f.argtypes = [M, M, i64, i32p]
            f = getattr(_lib, f"oracle_{p}getrf")
            f.argtypes = [M, ctypes.POINTER(ctypes.c_int64), i64, i64, i32p]
            f = getattr(_lib, f"oracle_{p}getf2"); f.argtypes = [M, ctypes.POINTER(ctypes.c_int64)]
            f = getattr(_lib, f"oracle_{p}iamax"); f.argtypes = [i64, ctypes.c_void_p, i64]; f.restype = i64
            f = getattr(_lib, f"oracle_{p}posv"); f.argtypes = [M, M, M, M, i64, i32p]
        _lib.oracle_dposv_avx2_mt.argtypes = [M, M, M, M, i64, i32p, ctypes.c_int]
        _lib.oracle_dlaswp.argtypes = [M, i64, i64, ctypes.POINTER(ctypes.c_int64)]
        _lib.oracle_dger.argtypes = [ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p, M, M]
        _lib.oracle_dposv_mt.argtypes = [M, M, M, M, i64, i32p, ctypes.c_int]
        _lib.oracle_dgemm_mt.argtypes = [ctypes.c_double, M, M, ctypes.c_double, M, M, i64, ctypes.c_int]
    return _lib


def _p(dtype):
    return "d" if np.dtype(dtype) == np.float64 else "s"


def _desc(layout: Layout, buf: np.ndarray):
    assert buf.dtype == layout.dtype and buf.flags.c_contiguous
    return ctypes.byref(layout.descriptor(buf.ctypes.data))


def _i32(n):
    return np.zeros(n, dtype=np.int32)


def gemm(alpha, la, a, lb, b, beta, lc, c, ld, d, batch):
    r = getattr(lib(), f"oracle_{_p(la.dtype)}gemm")(alpha, _desc(la, a), _desc(lb, b), beta,
                                                      _desc(lc, c), _desc(ld, d), batch)
    assert r == 0


def syrk_lower(alpha, la, a, beta, lc, c, ld, d, batch):
    r = getattr(lib(), f"oracle_{_p(la.dtype)}syrk_lower")(alpha, _desc(la, a), beta, _desc(lc, c),
                                                            _desc(ld, d), batch)
    assert r == 0


def trmm_left_upper(alpha, la, a, lb, b, lc, c, batch):
    r = getattr(lib(), f"oracle_{_p(la.dtype)}trmm_left_upper")(alpha, _desc(la, a), _desc(lb, b),
                                                                 _desc(lc, c), batch)
    assert r == 0


def trmm_right_lower(alpha, lb, b, la, a, lc, c, batch):
    r = getattr(lib(), f"oracle_{_p(la.dtype)}trmm_right_lower")(alpha, _desc(lb, b), _desc(la, a),
                                                                  _desc(lc, c), batch)
    assert r == 0


def trsm(side, lower, unit, alpha, la, a, lb, b, lx, x, batch):
    f = getattr(lib(), f"oracle_{_p(la.dtype)}trsm_{side}")
    r = f(int(lower), int(unit), alpha, _desc(la, a), _desc(lb, b), _desc(lx, x), batch)
    assert r == 0


def potrf(la, a, ll, l, batch):
    info = _i32(batch)
    r = getattr(lib(), f"oracle_{_p(la.dtype)}potrf")(_desc(la, a), _desc(ll, l), batch,
                                                       info.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    assert r == 0
    return info


def getrf(la, a, batch):
    n = min(la.rows, la.cols)
    ipiv = np.zeros(batch * max(n, 1), dtype=np.int64)
    info = _i32(batch)
    r = getattr(lib(), f"oracle_{_p(la.dtype)}getrf")(
        _desc(la, a), ipiv.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), max(n, 1), batch,
        info.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    assert r == 0
    return ipiv.reshape(batch, max(n, 1))[:, :n], info


def getf2(la, a):
    ipiv = np.zeros(max(min(la.rows, la.cols), 1), dtype=np.int64)
    r = getattr(lib(), f"oracle_{_p(la.dtype)}getf2")(_desc(la, a),
                                                       ipiv.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    return ipiv[: min(la.rows, la.cols)], r


def iamax(x: np.ndarray, incx: int = 1):
    return getattr(lib(), f"oracle_{_p(x.dtype)}iamax")(len(x[::incx]), x.ctypes.data, incx)


def laswp(la, a, k0, k1, ipiv):
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int64)
    lib().oracle_dlaswp(_desc(la, a), k0, k1, ipiv.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))


def ger(alpha, x, y, la, a, lb, b):
    r = lib().oracle_dger(alpha, x.ctypes.data, y.ctypes.data, _desc(la, a), _desc(lb, b))
    assert r == 0


def posv(la, a, ll, l, lb, b, lx, x, batch):
    info = _i32(batch)
    r = getattr(lib(), f"oracle_{_p(la.dtype)}posv")(_desc(la, a), _desc(ll, l), _desc(lb, b), _desc(lx, x),
                                                      batch, info.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    assert r == 0
    return info


def posv_mt(la, a, ll, l, lb, b, lx, x, batch, threads=0):
    info = _i32(batch)
    used = lib().oracle_dposv_mt(_desc(la, a), _desc(ll, l), _desc(lb, b), _desc(lx, x), batch,
                                 info.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), threads)
    return info, used


def posv_avx2_mt(la, a, ll, l, lb, b, lx, x, batch, threads=0):
    """BLAST-restated AVX2 Cholesky solve (blast_avx2.cpp), one instance per
    worker; returns (info, workers)."""
    info = _i32(batch)
    used = lib().oracle_dposv_avx2_mt(_desc(la, a), _desc(ll, l), _desc(lb, b), _desc(lx, x), batch,
                                      info.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), threads)
    assert used > 0, "unsupported shape for the AVX2 restatement"
    return info, used


def gemm_mt(alpha, la, a, lb, b, beta, lc, c, ld, d, batch, threads=0):
    return lib().oracle_dgemm_mt(alpha, _desc(la, a), _desc(lb, b), beta, _desc(lc, c), _desc(ld, d),
                                 batch, threads)
