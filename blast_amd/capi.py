"""ctypes binding of libblast_hip.so (the C ABI declared in include/blast_hip.h).

Device buffers are torch tensors on ``cuda:0`` (PyTorch is used for device
memory and streams only; every computation runs in the HIP kernels of
libblast_hip.so).  If the library is missing, every call raises: there is no
CPU fallback.
"""

from __future__ import annotations

import ctypes
import os
import re

import numpy as np

from .layout import BlastMatrix, Layout

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libblast_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "blast_hip.h")

ALGO_AUTO = 0
ALGO_REFERENCE = 1
ERR_RUNTIME = -1000

_lib = None


class BlastError(RuntimeError):
    def __init__(self, func, code, msg):
        super().__init__(f"{func} returned {code}: {msg}")
        self.code = code


def declared_symbols() -> list[str]:
    """Every function declared in include/blast_hip.h."""
    with open(HEADER_PATH) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(blast_hip_\w+)\s*\(", text)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc, gfx950)")
        # torch bundles its own HIP runtime and loads it as "libamdhip64.so";
        # loading torch first lets libblast_hip.so's libamdhip64.so.7 /
        # libhsa-runtime64.so.1 dependencies bind to that same runtime by
        # SONAME, so the process holds one HIP runtime, not two.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER(BlastMatrix)
        vp, i64, i32p, c_int = ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int
        for p, real in (("d", ctypes.c_double), ("s", ctypes.c_float)):
            getattr(L, f"blast_hip_{p}gemm_batched").argtypes = [vp, real, P, P, real, P, P, i64]
            getattr(L, f"blast_hip_{p}syrk_lower_batched").argtypes = [vp, real, P, real, P, P, i64]
            getattr(L, f"blast_hip_{p}trmm_batched").argtypes = [vp, c_int, c_int, real, P, P, P, i64]
            getattr(L, f"blast_hip_{p}trsm_batched").argtypes = [vp, c_int, c_int, c_int, real, P, P, P, i64]
            getattr(L, f"blast_hip_{p}potrf_batched").argtypes = [vp, P, P, i64, i32p]
            getattr(L, f"blast_hip_{p}getrf_batched").argtypes = [vp, P, vp, i64, i64, i32p]
        L.blast_hip_dposv_batched.argtypes = [vp, P, P, P, P, i64, i32p]
        L.blast_hip_sposv_batched.argtypes = [vp, P, P, P, P, i64, i32p]
        L.blast_hip_dsposv_batched.argtypes = [vp, c_int, c_int, ctypes.c_double, P, P, P, i64, i32p, i32p]
        for p, real in (("d", ctypes.c_double), ("s", ctypes.c_float)):
            getattr(L, f"blast_hip_{p}gemm_backward_batched").argtypes = [vp, real, P, P, real, P, P, P, P, i64]
            getattr(L, f"blast_hip_{p}trsm_backward_batched").argtypes = [vp, c_int, P, P, P, P, P, i64]
            getattr(L, f"blast_hip_{p}potrf_backward_batched").argtypes = [vp, P, P, P, i64]
            getattr(L, f"blast_hip_{p}posv_backward_batched").argtypes = [vp, c_int, P, P, P, P, P, i64, i32p]
            getattr(L, f"blast_hip_{p}getrs_batched").argtypes = [vp, P, vp, i64, P, P, i64]
            getattr(L, f"blast_hip_{p}getrs_backward_batched").argtypes = [vp, P, vp, i64, P, P, P, P, i64]
        L.blast_hip_dcopy_batched.argtypes = [vp, P, P, i64]
        L.blast_hip_scopy_batched.argtypes = [vp, P, P, i64]
        L.blast_hip_mixed_work_elems.argtypes = [i64]
        L.blast_hip_mixed_work_elems.restype = i64
        L.blast_hip_dsposv_ws_batched.argtypes = [vp, c_int, c_int, ctypes.c_double, P, P, P, vp, i64, i32p, i32p]
        L.blast_hip_dgesv_batched.argtypes = [vp, P, vp, i64, P, P, i64, i32p]
        L.blast_hip_dsgesv_batched.argtypes = [vp, c_int, c_int, ctypes.c_double, P, vp, i64, P, P, vp, i64, i32p,
                                               i32p]
        L.blast_hip_uneven_order.argtypes = [i64, vp, vp]
        L.blast_hip_dposv_uneven_batched.argtypes = [vp, vp, i64, vp, vp]
        L.blast_hip_dposv_backward_uneven_batched.argtypes = [vp, vp, i64, vp, vp]
        L.blast_hip_dtrsv_batched.argtypes = [vp, c_int, P, P, P, i64]
        L.blast_hip_strsv_batched.argtypes = [vp, c_int, P, P, P, i64]
        L.blast_hip_dsgesv_noalloc_batched.argtypes = [vp, c_int, c_int, ctypes.c_double, P, vp, i64, P, P, i64,
                                                       i32p, i32p]
        L.blast_hip_riccati_plan.argtypes = [i64, vp, vp, c_int, vp, vp, vp, vp, vp]
        L.blast_hip_driccati_batched.argtypes = [vp, c_int, vp, i64, vp, vp, vp, vp, vp]
        L.blast_hip_riccati_set_tuning.argtypes = [vp]
        L.blast_hip_riccati_get_tuning.argtypes = [vp]
        L.blast_hip_riccati_autotune.argtypes = [vp, c_int, vp, vp, i64, vp, vp, vp, vp, vp, c_int, vp]
        L.blast_hip_save_tuning.argtypes = [ctypes.c_char_p]
        L.blast_hip_load_tuning.argtypes = [ctypes.c_char_p]
        L.blast_hip_set_algorithm.argtypes = [c_int]
        L.blast_hip_last_error.restype = ctypes.c_char_p
        L.blast_hip_version.restype = ctypes.c_char_p
        _lib = L
    return _lib


def last_error() -> str:
    return lib().blast_hip_last_error().decode()


def set_algorithm(algo: int) -> None:
    r = lib().blast_hip_set_algorithm(algo)
    if r:
        raise BlastError("blast_hip_set_algorithm", r, "bad algorithm")


def check(name, code):
    if code != 0:
        raise BlastError(name, code, last_error())
    return code


def _prefix(dtype):
    return "d" if np.dtype(dtype) == np.float64 else "s"


def _stream(stream):
    if stream is not None:
        return stream
    import torch

    return torch.cuda.current_stream().cuda_stream


class Operand:
    """A batch operand on the device: a torch tensor plus its storage Layout."""

    def __init__(self, layout: Layout, tensor):
        self.layout = layout
        self.tensor = tensor

    @staticmethod
    def from_host(layout: Layout, host_buf: np.ndarray, device="cuda"):
        import torch

        return Operand(layout, torch.from_numpy(np.ascontiguousarray(host_buf)).to(device))

    @staticmethod
    def empty(layout: Layout, batch: int, device="cuda", fill=0.0):
        import torch

        tdt = torch.float64 if layout.dtype == np.float64 else torch.float32
        return Operand(layout, torch.full((batch * layout.batch_stride,), fill, dtype=tdt, device=device))

    def t(self) -> "Operand":
        return Operand(self.layout.t(), self.tensor)

    def sub(self, i, j, m, n) -> "Operand":
        """Rows i..i+m-1, columns j..j+n-1 of every instance (submatrix view)."""
        lay, off = self.layout.submatrix(i, j, m, n)
        return Operand(lay, self.tensor[off:])

    def desc(self) -> BlastMatrix:
        return self.layout.descriptor(self.tensor.data_ptr())

    def host(self) -> np.ndarray:
        return self.tensor.cpu().numpy()

    def logical(self, batch) -> np.ndarray:
        return self.layout.unpack(self.host(), batch)


def _ref(op: Operand):
    return ctypes.byref(op.desc())


def gemm(alpha, A, B, beta, C, D, batch, stream=None):
    p = _prefix(D.layout.dtype)
    name = f"blast_hip_{p}gemm_batched"
    return check(name, getattr(lib(), name)(_stream(stream), alpha, _ref(A), _ref(B), beta, _ref(C), _ref(D), batch))


def syrk_lower(alpha, A, beta, C, D, batch, stream=None):
    p = _prefix(D.layout.dtype)
    name = f"blast_hip_{p}syrk_lower_batched"
    return check(name, getattr(lib(), name)(_stream(stream), alpha, _ref(A), beta, _ref(C), _ref(D), batch))


def trmm(side_right, upper, alpha, A, B, C, batch, stream=None):
    p = _prefix(C.layout.dtype)
    name = f"blast_hip_{p}trmm_batched"
    return check(name, getattr(lib(), name)(_stream(stream), int(side_right), int(upper), alpha, _ref(A), _ref(B),
                                            _ref(C), batch))


def trsm(side_right, upper, unit, alpha, A, B, X, batch, stream=None):
    p = _prefix(X.layout.dtype)
    name = f"blast_hip_{p}trsm_batched"
    return check(name, getattr(lib(), name)(_stream(stream), int(side_right), int(upper), int(unit), alpha, _ref(A),
                                            _ref(B), _ref(X), batch))


def trsv(upper, A, b, x, batch, stream=None):
    """A x = b, one right-hand side per instance (b, x: n x 1 operands)."""
    p = _prefix(x.layout.dtype)
    name = f"blast_hip_{p}trsv_batched"
    return check(name, getattr(lib(), name)(_stream(stream), int(upper), _ref(A), _ref(b), _ref(x), batch))


class SolveInstance(ctypes.Structure):
    """blast_solve_instance_t: descriptors of one instance of an uneven batch."""
    _fields_ = [("a", BlastMatrix), ("l", BlastMatrix), ("b", BlastMatrix), ("x", BlastMatrix)]


class AdjointInstance(ctypes.Structure):
    """blast_adjoint_instance_t: descriptors of one instance of an uneven adjoint."""
    _fields_ = [("l", BlastMatrix), ("x", BlastMatrix), ("xb", BlastMatrix), ("ab", BlastMatrix),
                ("bb", BlastMatrix)]


class UnevenBatch:
    """Instances of different n packed back to back in one FP64 buffer per
    operand (A, L, B, X), each in its own Layout; the device descriptor array
    and the queue (order by blast_hip_uneven_order) are built once."""

    def __init__(self, layouts_a, layouts_b=None, alias_l=False, device="cuda"):
        import torch

        self.la = list(layouts_a)
        self.lb = list(layouts_b) if layouts_b is not None else None
        count = len(self.la)
        self.count = count

        def offsets(lays):
            off, o = [], 0
            for lay in lays:
                off.append(o)
                o += lay.batch_stride
            return off, max(o, 1)

        self.oa, na = offsets(self.la)
        self.a = torch.zeros(na, dtype=torch.float64, device=device)
        self.l = self.a if alias_l else torch.zeros(na, dtype=torch.float64, device=device)
        if self.lb is not None:
            self.ob, nb = offsets(self.lb)
            self.b = torch.zeros(nb, dtype=torch.float64, device=device)
            self.x = torch.zeros(nb, dtype=torch.float64, device=device)
        inst = (SolveInstance * max(count, 1))()
        for k in range(count):
            inst[k].a = self.la[k].descriptor(self.a.data_ptr() + 8 * self.oa[k])
            inst[k].l = self.la[k].descriptor(self.l.data_ptr() + 8 * self.oa[k])
            if self.lb is not None:
                inst[k].b = self.lb[k].descriptor(self.b.data_ptr() + 8 * self.ob[k])
                inst[k].x = self.lb[k].descriptor(self.x.data_ptr() + 8 * self.ob[k])
        raw = np.frombuffer(bytes(inst), dtype=np.uint8)
        self.inst = torch.from_numpy(raw.copy()).to(device)
        n = np.array([lay.rows for lay in self.la], dtype=np.int32)
        order = np.zeros(max(count, 1), dtype=np.int32)
        check("blast_hip_uneven_order", lib().blast_hip_uneven_order(count, n.ctypes.data, order.ctypes.data))
        self.order = order[:count]
        self.queue = torch.from_numpy(np.concatenate([[0], self.order]).astype(np.int32)).to(device)
        self.info = torch.zeros(max(count, 1), dtype=torch.int32, device=device)

    def pack(self, a_list, b_list=None):
        host = np.zeros(self.a.numel())
        for k, lay in enumerate(self.la):
            host[self.oa[k]:self.oa[k] + lay.batch_stride] = lay.pack(a_list[k])
        self.a.copy_(__import__("torch").from_numpy(host))
        if b_list is not None:
            hb = np.zeros(self.b.numel())
            for k, lay in enumerate(self.lb):
                hb[self.ob[k]:self.ob[k] + lay.batch_stride] = lay.pack(b_list[k])
            self.b.copy_(__import__("torch").from_numpy(hb))

    def run(self, stream=None):
        return check("blast_hip_dposv_uneven_batched",
                     lib().blast_hip_dposv_uneven_batched(_stream(stream), self.inst.data_ptr(), self.count,
                                                          self.queue.data_ptr(), self.info.data_ptr()))

    def backward_buffers(self, device="cuda"):
        """Xbar, Abar and Bbar buffers (laid out as B, A and B) and the adjoint
        descriptor array over this batch's L and X (the forward's outputs)."""
        import torch

        assert self.lb is not None, "the adjoint needs the right-hand sides' layouts"
        self.xb = torch.zeros_like(self.b)
        self.bb = torch.zeros_like(self.b)
        self.ab = torch.zeros_like(self.a)
        inst = (AdjointInstance * max(self.count, 1))()
        for k in range(self.count):
            inst[k].l = self.la[k].descriptor(self.l.data_ptr() + 8 * self.oa[k])
            inst[k].x = self.lb[k].descriptor(self.x.data_ptr() + 8 * self.ob[k])
            inst[k].xb = self.lb[k].descriptor(self.xb.data_ptr() + 8 * self.ob[k])
            inst[k].ab = self.la[k].descriptor(self.ab.data_ptr() + 8 * self.oa[k])
            inst[k].bb = self.lb[k].descriptor(self.bb.data_ptr() + 8 * self.ob[k])
        raw = np.frombuffer(bytes(inst), dtype=np.uint8)
        self.adj_inst = torch.from_numpy(raw.copy()).to(device)
        self.adj_queue = self.queue.clone()
        self.adj_info = torch.full((max(self.count, 1),), -2, dtype=torch.int32, device=device)

    def pack_xbar(self, xb_list):
        hb = np.zeros(self.xb.numel())
        for k, lay in enumerate(self.lb):
            hb[self.ob[k]:self.ob[k] + lay.batch_stride] = lay.pack(xb_list[k])
        self.xb.copy_(__import__("torch").from_numpy(hb))

    def backward(self, stream=None):
        return check("blast_hip_dposv_backward_uneven_batched",
                     lib().blast_hip_dposv_backward_uneven_batched(_stream(stream), self.adj_inst.data_ptr(),
                                                                   self.count, self.adj_queue.data_ptr(),
                                                                   self.adj_info.data_ptr()))

    def result(self, which, k):
        bufs = {"l": (self.l, self.la, self.oa), "x": (getattr(self, "x", None), self.lb, getattr(self, "ob", None)),
                "ab": (getattr(self, "ab", None), self.la, self.oa),
                "bb": (getattr(self, "bb", None), self.lb, getattr(self, "ob", None))}
        buf, lays, offs = bufs[which]
        lay = lays[k]
        host = buf[offs[k]:offs[k] + lay.batch_stride].cpu().numpy()
        return lay.unpack(host, 1)[0]


def _info_ptr(info):
    return None if info is None else info.data_ptr()


def potrf(A, L, batch, info=None, stream=None):
    p = _prefix(L.layout.dtype)
    name = f"blast_hip_{p}potrf_batched"
    return check(name, getattr(lib(), name)(_stream(stream), _ref(A), _ref(L), batch, _info_ptr(info)))


def getrf(A, ipiv, ipiv_stride, batch, info=None, stream=None):
    p = _prefix(A.layout.dtype)
    name = f"blast_hip_{p}getrf_batched"
    return check(name, getattr(lib(), name)(_stream(stream), _ref(A), ipiv.data_ptr() if ipiv is not None else None,
                                            ipiv_stride, batch, _info_ptr(info)))


def posv(A, L, B, X, batch, info=None, stream=None):
    name = f"blast_hip_{_prefix(L.layout.dtype)}posv_batched"
    return check(name, getattr(lib(), name)(_stream(stream), _ref(A), _ref(L), _ref(B), _ref(X), batch,
                                            _info_ptr(info)))


def sposv(precision, A, B, X, batch, iter_, info=None, itmax=30, tol=0.0, stream=None):
    """Mixed precision (0 = FP32, 1 = BF16 products) + FP64 refinement."""
    name = "blast_hip_dsposv_batched"
    return check(name, lib().blast_hip_dsposv_batched(_stream(stream), int(precision), int(itmax), float(tol), _ref(A),
                                                      _ref(B), _ref(X), batch, _info_ptr(info), _info_ptr(iter_)))


def copy(S, D, batch, stream=None):
    """D := S (any two storage layouts of the same logical size)."""
    name = f"blast_hip_{_prefix(D.layout.dtype)}copy_batched"
    return check(name, getattr(lib(), name)(_stream(stream), _ref(S), _ref(D), batch))


def mixed_work_elems(n):
    """FP32 factor buffer (floats per instance) of sposv_ws / sgesv."""
    return int(lib().blast_hip_mixed_work_elems(int(n)))


def mixed_work(n, batch, device="cuda"):
    import torch

    return torch.empty(max(1, mixed_work_elems(n) * batch), dtype=torch.float32, device=device)


def sposv_ws(precision, A, B, X, work, batch, iter_, info=None, itmax=30, tol=0.0, stream=None):
    """Mixed precision Cholesky solve, n <= 256, nrhs <= 8 (LAPACK dsposv)."""
    name = "blast_hip_dsposv_ws_batched"
    return check(name, lib().blast_hip_dsposv_ws_batched(_stream(stream), int(precision), int(itmax), float(tol),
                                                         _ref(A), _ref(B), _ref(X), _info_ptr(work), batch,
                                                         _info_ptr(info), _info_ptr(iter_)))


def gesv(A, ipiv, ipiv_stride, B, X, batch, info=None, stream=None):
    """FP64 LU with partial pivoting in place of A, then X = A^{-1} B (LAPACK dgesv)."""
    name = "blast_hip_dgesv_batched"
    return check(name, lib().blast_hip_dgesv_batched(_stream(stream), _ref(A), _info_ptr(ipiv), ipiv_stride, _ref(B),
                                                     _ref(X), batch, _info_ptr(info)))


def sgesv(precision, A, ipiv, ipiv_stride, B, X, work, batch, iter_, info=None, itmax=30, tol=0.0, stream=None):
    """Mixed precision LU + FP64 refinement (LAPACK dsgesv)."""
    name = "blast_hip_dsgesv_batched"
    return check(name, lib().blast_hip_dsgesv_batched(_stream(stream), int(precision), int(itmax), float(tol),
                                                      _ref(A), _info_ptr(ipiv), ipiv_stride, _ref(B), _ref(X),
                                                      _info_ptr(work), batch, _info_ptr(info), _info_ptr(iter_)))


def sgesv_noalloc(precision, A, ipiv, ipiv_stride, B, X, batch, iter_, info=None, itmax=30, tol=0.0, stream=None):
    """sgesv without a work buffer: the FP32 factor lives in the library's
    device scratch (NOALLOC_SCRATCH_FLOATS floats per GPU)."""
    name = "blast_hip_dsgesv_noalloc_batched"
    return check(name, lib().blast_hip_dsgesv_noalloc_batched(_stream(stream), int(precision), int(itmax), float(tol),
                                                              _ref(A), _info_ptr(ipiv), ipiv_stride, _ref(B), _ref(X),
                                                              batch, _info_ptr(info), _info_ptr(iter_)))


NOALLOC_SCRATCH_FLOATS = 16 << 20


def _oref(op):
    return None if op is None else _ref(op)


def gemm_backward(alpha, A, B, beta, Db, Ab, Bb, Cb, batch, stream=None):
    name = f"blast_hip_{_prefix(Db.layout.dtype)}gemm_backward_batched"
    return check(name, getattr(lib(), name)(_stream(stream), alpha, _ref(A), _ref(B), beta, _ref(Db), _oref(Ab),
                                            _oref(Bb), _oref(Cb), batch))


def trsm_backward(upper, A, X, Xb, Ab, Bb, batch, stream=None):
    name = f"blast_hip_{_prefix(X.layout.dtype)}trsm_backward_batched"
    return check(name, getattr(lib(), name)(_stream(stream), int(upper), _ref(A), _ref(X), _ref(Xb), _oref(Ab),
                                            _ref(Bb), batch))


def potrf_backward(L, Lb, Ab, batch, stream=None):
    name = f"blast_hip_{_prefix(L.layout.dtype)}potrf_backward_batched"
    return check(name, getattr(lib(), name)(_stream(stream), _ref(L), _ref(Lb), _ref(Ab), batch))


def posv_backward(recompute, F, X, Xb, Ab, Bb, batch, info=None, stream=None):
    name = f"blast_hip_{_prefix(X.layout.dtype)}posv_backward_batched"
    return check(name, getattr(lib(), name)(_stream(stream), int(recompute), _ref(F), _ref(X), _ref(Xb), _ref(Ab),
                                            _ref(Bb), batch, _info_ptr(info)))


def getrs(LU, ipiv, ipiv_stride, B, X, batch, stream=None):
    name = f"blast_hip_{_prefix(X.layout.dtype)}getrs_batched"
    return check(name, getattr(lib(), name)(_stream(stream), _ref(LU), ipiv.data_ptr(), ipiv_stride, _ref(B),
                                            _ref(X), batch))


def getrs_backward(LU, ipiv, ipiv_stride, X, Xb, Ab, Bb, batch, stream=None):
    name = f"blast_hip_{_prefix(X.layout.dtype)}getrs_backward_batched"
    return check(name, getattr(lib(), name)(_stream(stream), _ref(LU), ipiv.data_ptr(), ipiv_stride, _ref(X),
                                            _ref(Xb), _oref(Ab), _ref(Bb), batch))


# ---- uneven batched Riccati recursion -------------------------------------

RICCATI_INSTANCE = np.dtype([("nx", np.int32), ("nu", np.int32), ("a", np.int64), ("b", np.int64),
                             ("q", np.int64), ("r", np.int64), ("k", np.int64), ("l", np.int64),
                             ("work", np.int64)], align=True)


def riccati_plan(nx, nu, horizon):
    """Host plan (blast_hip_riccati_plan): descriptors, queue order and the
    element counts of the input, output and work buffers."""
    nx = np.ascontiguousarray(nx, dtype=np.int32)
    nu = np.ascontiguousarray(nu, dtype=np.int32)
    count = len(nx)
    inst = np.zeros(count, dtype=RICCATI_INSTANCE)
    order = np.zeros(count, dtype=np.int32)
    sizes = np.zeros(3, dtype=np.int64)
    check("blast_hip_riccati_plan",
          lib().blast_hip_riccati_plan(count, nx.ctypes.data, nu.ctypes.data, int(horizon), inst.ctypes.data,
                                       order.ctypes.data, sizes[0:].ctypes.data, sizes[1:].ctypes.data,
                                       sizes[2:].ctypes.data))
    return inst, order, [int(v) for v in sizes]


class RiccatiBatch:
    """Device buffers of one uneven Riccati batch, laid out by riccati_plan."""

    def __init__(self, nx, nu, horizon, device="cuda"):
        import torch

        self.horizon = int(horizon)
        self.inst, self.order, (n_in, n_out, n_work) = riccati_plan(nx, nu, horizon)
        self.count = len(self.inst)
        self.d_inst = torch.from_numpy(self.inst.view(np.uint8).copy()).to(device)
        queue = np.concatenate([np.zeros(1, np.int32), self.order])
        self.queue = torch.from_numpy(queue).to(device)
        self.input = torch.zeros(max(n_in, 1), dtype=torch.float64, device=device)
        self.output = torch.zeros(max(n_out, 1), dtype=torch.float64, device=device)
        self.work = torch.zeros(max(n_work, 1), dtype=torch.float64, device=device)
        self.info = torch.full((max(self.count, 1),), -1, dtype=torch.int32, device=device)

    def pack(self, A, B, Q, R):
        """Host lists of per-instance matrices -> the device input buffer."""
        import torch

        host = np.zeros(self.input.numel(), dtype=np.float64)
        for i, d in enumerate(self.inst):
            for key, m in (("a", A[i]), ("b", B[i]), ("q", Q[i]), ("r", R[i])):
                m = np.asarray(m, dtype=np.float64)
                host[d[key]:d[key] + m.size] = m.ravel(order="F")
        self.input.copy_(torch.from_numpy(host))

    def run(self, stream=None):
        return check("blast_hip_driccati_batched",
                     lib().blast_hip_driccati_batched(_stream(stream), self.horizon, self.d_inst.data_ptr(),
                                                      self.count, self.input.data_ptr(), self.output.data_ptr(),
                                                      self.work.data_ptr(), self.queue.data_ptr(),
                                                      self.info.data_ptr()))

    def autotune(self, reps=3, stream=None):
        ms = np.zeros(48, dtype=np.float32)
        check("blast_hip_riccati_autotune",
              lib().blast_hip_riccati_autotune(_stream(stream), self.horizon, self.d_inst.data_ptr(),
                                               self.inst.ctypes.data, self.count, self.input.data_ptr(),
                                               self.output.data_ptr(), self.work.data_ptr(), self.queue.data_ptr(),
                                               self.order.ctypes.data, int(reps), ms.ctypes.data))
        return ms.reshape(16, 3)

    def results(self):
        """Host copies: list of K (horizon x nu x nx) and list of L0 (nx x nx)."""
        out = self.output.cpu().numpy()
        K, L0 = [], []
        for d in self.inst:
            nx, nu = int(d["nx"]), int(d["nu"])
            k = out[d["k"]:d["k"] + self.horizon * nu * nx].reshape(self.horizon, nx, nu).transpose(0, 2, 1)
            K.append(k)
            L0.append(out[d["l"]:d["l"] + nx * nx].reshape(nx, nx).T)
        return K, L0


def riccati_get_tuning():
    t = np.zeros(16, dtype=np.int8)
    check("blast_hip_riccati_get_tuning", lib().blast_hip_riccati_get_tuning(t.ctypes.data))
    return t


def riccati_set_tuning(t):
    t = np.ascontiguousarray(t, dtype=np.int8)
    return check("blast_hip_riccati_set_tuning", lib().blast_hip_riccati_set_tuning(t.ctypes.data))


def save_tuning(path):
    return check("blast_hip_save_tuning", lib().blast_hip_save_tuning(os.fsencode(path)))


def load_tuning(path):
    return check("blast_hip_load_tuning", lib().blast_hip_load_tuning(os.fsencode(path)))
