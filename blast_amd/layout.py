"""Host-side description of BLAST matrix storage and batches.

A batch of matrices is a flat buffer holding ``batch`` instances ``batch_stride``
elements apart.  Each instance is stored column-major in one of BLAST's two
formats:

* dense  (``dense/DynamicMatrix.hpp:36-39,149``): element (i, j) at
  ``i + spacing*j`` with ``spacing = nextMultiple(rows, SS)``;
* panel  (``panel/DynamicPanelMatrix.hpp:60-61,184-188``): element (i, j) at
  ``(i // PS)*spacing + i % PS + j*PS`` with ``spacing = PS*cols`` and
  ``ceil(rows/PS)`` panels.

PS (= SS, the SIMD width BLAST builds for on AVX2) is 4 for float64 and 8 for
float32 (``panel/PanelSize.hpp:25-26``).  Row-major storage of a matrix is the
``trans=1`` view of the column-major storage of its transpose, which is how
``trans()`` of a panel pointer works in BLAST (it flips the storage-order
flag only, ``expressions/PMatTransExpr.hpp:232-236``).

``BlastMatrix`` has the field order of ``blast_matrix_t`` in
``include/blast_hip.h``.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

PANEL_SIZE = {np.dtype(np.float64): 4, np.dtype(np.float32): 8}
ALIGN_BYTES = 64  # system/CacheLine.hpp:16


class BlastMatrix(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("rows", ctypes.c_int64),
        ("cols", ctypes.c_int64),
        ("spacing", ctypes.c_int64),
        ("batch_stride", ctypes.c_int64),
        ("panel", ctypes.c_int32),
        ("trans", ctypes.c_int32),
    ]


def panel_size(dtype) -> int:
    return PANEL_SIZE[np.dtype(dtype)]


def _next_multiple(x: int, m: int) -> int:
    return (x + m - 1) // m * m


@dataclass(frozen=True)
class Layout:
    """Storage of one batch operand: logical rows x cols, format, transpose."""

    rows: int
    cols: int
    dtype: np.dtype
    panel: bool = True
    trans: bool = False
    spacing: int = 0
    batch_stride: int = 0

    @staticmethod
    def make(rows, cols, dtype=np.float64, panel=True, trans=False):
        dtype = np.dtype(dtype)
        ps = panel_size(dtype)
        srows, scols = (cols, rows) if trans else (rows, cols)
        if panel:
            spacing = ps * scols
            size = _next_multiple(srows, ps) * scols
        else:
            spacing = _next_multiple(srows, ps)
            size = spacing * scols
        stride = _next_multiple(max(size, 1), ALIGN_BYTES // dtype.itemsize)
        return Layout(rows, cols, dtype, panel, trans, spacing, stride)

    @property
    def storage_shape(self):
        return (self.cols, self.rows) if self.trans else (self.rows, self.cols)

    def t(self) -> "Layout":
        """Transposed view of the same storage (BLAST trans())."""
        return Layout(self.cols, self.rows, self.dtype, self.panel, not self.trans,
                      self.spacing, self.batch_stride)

    def submatrix(self, i, j, m, n):
        """Sub-block rows i..i+m-1, columns j..j+n-1 of every instance: its
        Layout (same storage, spacing and batch stride) and the element offset
        of its (0, 0) from the instance base.  In panel storage the block's
        first storage row must lie on a panel boundary, as BLAST's
        PanelSubmatrix requires (views/submatrix/Panel.hpp:96-102)."""
        if min(i, j, m, n) < 0 or i + m > self.rows or j + n > self.cols:
            raise ValueError("Invalid submatrix specification")
        ps = panel_size(self.dtype)
        r, c = (j, i) if self.trans else (i, j)
        if self.panel and m > 0 and n > 0 and r % ps:
            raise ValueError("submatrix of a panel matrix not aligned on a panel boundary")
        off = (r // ps) * self.spacing + c * ps if self.panel else r + self.spacing * c
        return Layout(m, n, self.dtype, self.panel, self.trans, self.spacing, self.batch_stride), off

    def descriptor(self, ptr: int) -> BlastMatrix:
        return BlastMatrix(ptr, self.rows, self.cols, self.spacing, self.batch_stride,
                           int(self.panel), int(self.trans))

    def index_map(self) -> np.ndarray:
        """Storage offset of every logical element (rows x cols int64)."""
        ps = panel_size(self.dtype)
        i = np.arange(self.rows)[:, None]
        j = np.arange(self.cols)[None, :]
        if self.trans:
            i, j = j, i
        if self.panel:
            return ((i // ps) * self.spacing + (i % ps) + j * ps).astype(np.int64)
        return (i + self.spacing * j).astype(np.int64)

    def pack(self, x: np.ndarray, out: np.ndarray | None = None) -> np.ndarray:
        """Logical (batch, rows, cols) array -> flat storage buffer."""
        x = np.asarray(x, dtype=self.dtype)
        if x.ndim == 2:
            x = x[None]
        batch = x.shape[0]
        assert x.shape[1:] == (self.rows, self.cols), (x.shape, self.rows, self.cols)
        if out is None:
            out = np.zeros(batch * self.batch_stride, dtype=self.dtype)
        idx = self.index_map()
        view = out.reshape(batch, self.batch_stride)
        view[:, idx] = x
        return out

    def unpack(self, buf: np.ndarray, batch: int) -> np.ndarray:
        """Flat storage buffer -> logical (batch, rows, cols) array."""
        view = np.asarray(buf).reshape(-1)[: batch * self.batch_stride].reshape(batch, self.batch_stride)
        return view[:, self.index_map()]
