"""BLASFEO-format entry points of libblast_hip.so (``blast_hip_blasfeo_*``).

BLAST wraps BLASFEO's C API in ``include/blasfeo/blasfeo_api/`` (gemm_nn /
gemm_nt / gemm_tn, syrk_ln, potrf, syrk_potrf, trmm_rlnn / trmm_rutn).  These
bindings run the same routines batched on the GPU, on matrices in BLASFEO's
panel-major format: element (i, j) of an instance at
``pA[(i // bs) * bs * cn + i % bs + j * bs]``, ``cn >= n`` the padded column
count.  That is BLAST's panel storage with spacing ``bs * cn``; the library
turns every operand into such a view and runs its AUTO kernels.

Operands are :class:`Operand` objects: a :class:`BlasfeoLayout` plus a flat
torch tensor holding ``batch`` instances ``batch_stride`` elements apart.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import capi


class BlasfeoMat(ctypes.Structure):
    """``blast_blasfeo_mat_t`` (include/blast_hip.h)."""

    _fields_ = [
        ("pA", ctypes.c_void_p),
        ("m", ctypes.c_int64),
        ("n", ctypes.c_int64),
        ("cn", ctypes.c_int64),
        ("bs", ctypes.c_int64),
        ("batch_stride", ctypes.c_int64),
    ]


def panel_size(dtype) -> int:
    """Panel height the library accepts: 4 rows for double, 8 for float."""
    return 4 if np.dtype(dtype) == np.float64 else 8


class BlasfeoLayout:
    """Storage of a batch of m x n BLASFEO matrices (panel height bs, padded
    column count cn, pm = m rounded up to bs rows, instances 64-byte aligned)."""

    def __init__(self, m: int, n: int, dtype=np.float64, cn: int | None = None, bs: int | None = None):
        self.m, self.n = int(m), int(n)
        self.dtype = np.dtype(dtype)
        self.bs = int(bs) if bs is not None else panel_size(dtype)
        self.cn = int(cn) if cn is not None else -(-self.n // 4) * 4
        if self.cn < self.n:
            raise ValueError("cn must be >= n")
        self.pm = -(-self.m // self.bs) * self.bs
        align = 64 // self.dtype.itemsize
        self.batch_stride = max(align, -(-(self.pm * self.cn) // align) * align)

    def index(self, i, j):
        """Element offset of (i, j) inside one instance."""
        i = np.asarray(i)
        j = np.asarray(j)
        return (i // self.bs) * self.bs * self.cn + i % self.bs + j * self.bs

    def _grid(self):
        i, j = np.meshgrid(np.arange(self.m), np.arange(self.n), indexing="ij")
        return self.index(i, j)

    def pack(self, a: np.ndarray, pad=0.0) -> np.ndarray:
        """[batch, m, n] -> flat buffer; padding elements hold ``pad``."""
        a = np.asarray(a, dtype=self.dtype)
        batch = a.shape[0]
        buf = np.full((batch, self.batch_stride), pad, dtype=self.dtype)
        buf[:, self._grid()] = a
        return buf.reshape(-1)

    def unpack(self, buf: np.ndarray, batch: int) -> np.ndarray:
        b = np.asarray(buf).reshape(-1)[: batch * self.batch_stride].reshape(batch, self.batch_stride)
        return b[:, self._grid()]

    def descriptor(self, ptr: int) -> BlasfeoMat:
        return BlasfeoMat(ptr, self.m, self.n, self.cn, self.bs, self.batch_stride)


class Operand:
    """A batch of BLASFEO matrices on the device."""

    def __init__(self, layout: BlasfeoLayout, tensor):
        self.layout = layout
        self.tensor = tensor

    @staticmethod
    def from_host(layout: BlasfeoLayout, a: np.ndarray, pad=0.0, device="cuda") -> "Operand":
        import torch

        return Operand(layout, torch.from_numpy(layout.pack(a, pad)).to(device))

    def desc(self) -> BlasfeoMat:
        return self.layout.descriptor(self.tensor.data_ptr())

    def host(self) -> np.ndarray:
        return self.tensor.cpu().numpy()

    def logical(self, batch: int) -> np.ndarray:
        return self.layout.unpack(self.host(), batch)


_bound = False


def _bind():
    global _bound
    if _bound:
        return capi.lib()
    L = capi.lib()
    vp, i64, c_int = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
    P = ctypes.POINTER(BlasfeoMat)
    for p, real in (("d", ctypes.c_double), ("s", ctypes.c_float)):
        for t in ("nn", "nt", "tn", "tt"):
            getattr(L, f"blast_hip_blasfeo_{p}gemm_{t}").argtypes = [
                vp, c_int, c_int, c_int, real, P, c_int, c_int, P, c_int, c_int, real, P, c_int, c_int, P, c_int,
                c_int, i64]
        getattr(L, f"blast_hip_blasfeo_{p}syrk_ln").argtypes = [
            vp, c_int, c_int, real, P, c_int, c_int, P, c_int, c_int, real, P, c_int, c_int, P, c_int, c_int, i64]
        getattr(L, f"blast_hip_blasfeo_{p}potrf_l").argtypes = [vp, c_int, P, c_int, c_int, P, c_int, c_int, i64, vp]
        for t in ("rlnn", "rutn"):
            getattr(L, f"blast_hip_blasfeo_{p}trmm_{t}").argtypes = [
                vp, c_int, c_int, real, P, c_int, c_int, P, c_int, c_int, P, c_int, c_int, i64]
    for name in ("blast_hip_blasfeo_dsyrk_dpotrf_ln", "blast_hip_blasfeo_ssyrk_spotrf_ln"):
        getattr(L, name).argtypes = [vp, c_int, c_int, P, c_int, c_int, P, c_int, c_int, P, c_int, c_int, P, c_int,
                                     c_int, i64, vp]
    _bound = True
    return L


def _p(op: Operand):
    return ctypes.byref(op.desc())


def _prefix(op: Operand) -> str:
    return "d" if op.layout.dtype == np.float64 else "s"


def _info(info):
    return None if info is None else info.data_ptr()


def gemm(trans_a: bool, trans_b: bool, m, n, k, alpha, A, ai, aj, B, bi, bj, beta, C, ci, cj, D, di, dj, batch,
         stream=None):
    """blasfeo_?gemm_{n,t}{n,t}: D = beta*C + alpha*op(A)*op(B)."""
    L = _bind()
    name = f"blast_hip_blasfeo_{_prefix(D)}gemm_{'t' if trans_a else 'n'}{'t' if trans_b else 'n'}"
    return capi.check(name, getattr(L, name)(capi._stream(stream), m, n, k, alpha, _p(A), ai, aj, _p(B), bi, bj,
                                             beta, _p(C), ci, cj, _p(D), di, dj, batch))


def syrk_ln(m, k, alpha, A, ai, aj, B, bi, bj, beta, C, ci, cj, D, di, dj, batch, stream=None):
    """blasfeo_?syrk_ln: lower triangle of D = beta*C + alpha*A*B^T."""
    L = _bind()
    name = f"blast_hip_blasfeo_{_prefix(D)}syrk_ln"
    return capi.check(name, getattr(L, name)(capi._stream(stream), m, k, alpha, _p(A), ai, aj, _p(B), bi, bj, beta,
                                             _p(C), ci, cj, _p(D), di, dj, batch))


def potrf_l(m, C, ci, cj, D, di, dj, batch, info=None, stream=None):
    """blasfeo_?potrf_l: D = chol(C), lower."""
    L = _bind()
    name = f"blast_hip_blasfeo_{_prefix(D)}potrf_l"
    return capi.check(name, getattr(L, name)(capi._stream(stream), m, _p(C), ci, cj, _p(D), di, dj, batch,
                                             _info(info)))


def syrk_potrf_ln(m, k, A, ai, aj, B, bi, bj, C, ci, cj, D, di, dj, batch, info=None, stream=None):
    """blasfeo_dsyrk_dpotrf_ln / blasfeo_ssyrk_spotrf_ln: D = chol(C + A*B^T)."""
    L = _bind()
    name = ("blast_hip_blasfeo_dsyrk_dpotrf_ln" if D.layout.dtype == np.float64
            else "blast_hip_blasfeo_ssyrk_spotrf_ln")
    return capi.check(name, getattr(L, name)(capi._stream(stream), m, k, _p(A), ai, aj, _p(B), bi, bj, _p(C), ci,
                                             cj, _p(D), di, dj, batch, _info(info)))


def trmm(upper_t: bool, m, n, alpha, A, ai, aj, B, bi, bj, D, di, dj, batch, stream=None):
    """blasfeo_?trmm_rlnn (D = alpha*B*A, A lower) or, upper_t,
    blasfeo_?trmm_rutn (D = alpha*B*A^T, A upper)."""
    L = _bind()
    name = f"blast_hip_blasfeo_{_prefix(D)}trmm_{'rutn' if upper_t else 'rlnn'}"
    return capi.check(name, getattr(L, name)(capi._stream(stream), m, n, alpha, _p(A), ai, aj, _p(B), bi, bj, _p(D),
                                             di, dj, batch))
