"""Host-side ELLPACK-R / ELLBLOCK containers and their file format.

The layouts are bml's (reference ``src/macros.h:13``,
``ellpack/bml_types_ellpack.h:11-46``, ``ellblock/bml_types_ellblock.h:11-56``):

* ELLPACK-R: ``value[N, M]``, ``index[N, M]`` (int32), ``nnz[N]`` (int32),
  row-major; row ``i`` holds ``nnz[i]`` (column, value) pairs in its first
  slots.
* ELLBLOCK: block-row ``ib`` holds ``nnzb[ib]`` blocks; slot ``(ib, jp)``
  has block column ``indexb[ib, jp]``.  The values are held here as one
  ``bmax x bmax`` tile per slot (``value[NB, MB, bmax, bmax]``), the
  ``bsize[ib] x bsize[jb]`` block in its top-left corner and zeros elsewhere.
  The reference packs each block with leading dimension ``bsize[jb]``
  (``ellblock/bml_allocate_ellblock_typed.c:364-371``); the numbers are the
  same.

File formats ``ELL1`` / ``BLK1`` are the ones ``oracle/ref_driver.c`` reads.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

# bml_matrix_precision_t (reference src/C-interface/bml_types.h:24-36)
SINGLE_REAL, DOUBLE_REAL, SINGLE_COMPLEX, DOUBLE_COMPLEX = 1, 2, 3, 4
DTYPES = {
    SINGLE_REAL: np.float32,
    DOUBLE_REAL: np.float64,
    SINGLE_COMPLEX: np.complex64,
    DOUBLE_COMPLEX: np.complex128,
}
PREC_OF = {np.dtype(v): k for k, v in DTYPES.items()}
PREC_NAMES = {
    SINGLE_REAL: "single_real",
    DOUBLE_REAL: "double_real",
    SINGLE_COMPLEX: "single_complex",
    DOUBLE_COMPLEX: "double_complex",
}


def precision_of(dtype) -> int:
    return PREC_OF[np.dtype(dtype)]


@dataclass
class Ell:
    """An ELLPACK-R matrix in host memory."""

    value: np.ndarray
    index: np.ndarray
    nnz: np.ndarray

    @property
    def N(self) -> int:
        return self.value.shape[0]

    @property
    def M(self) -> int:
        return self.value.shape[1]

    @property
    def dtype(self):
        return self.value.dtype

    @property
    def precision(self) -> int:
        return precision_of(self.value.dtype)

    @classmethod
    def zeros(cls, N: int, M: int, dtype=np.float64) -> "Ell":
        return cls(np.zeros((N, M), dtype), np.zeros((N, M), np.int32),
                   np.zeros(N, np.int32))

    def copy(self) -> "Ell":
        return Ell(self.value.copy(), self.index.copy(), self.nnz.copy())

    def with_m(self, M: int) -> "Ell":
        """Same matrix stored with M slots per row."""
        if int(self.nnz.max(initial=0)) > M:
            raise ValueError("row does not fit M")
        out = Ell.zeros(self.N, M, self.dtype)
        k = min(M, self.M)
        out.value[:, :k] = self.value[:, :k]
        out.index[:, :k] = self.index[:, :k]
        out.nnz[:] = self.nnz
        return out

    def astype(self, dtype) -> "Ell":
        return Ell(self.value.astype(dtype), self.index.copy(), self.nnz.copy())

    def live_mask(self) -> np.ndarray:
        return np.arange(self.M)[None, :] < self.nnz[:, None]

    def to_dense(self) -> np.ndarray:
        """Dense N x N copy; duplicate columns in a row add up."""
        d = np.zeros((self.N, self.N), self.dtype)
        m = self.live_mask()
        rows = np.broadcast_to(np.arange(self.N)[:, None], m.shape)[m]
        np.add.at(d, (rows, self.index[m]), self.value[m])
        return d

    def to_coo(self):
        """(rows, cols, values) of the live entries, in storage order."""
        m = self.live_mask()
        rows = np.broadcast_to(np.arange(self.N)[:, None], m.shape)[m]
        return rows.astype(np.int64), self.index[m].astype(np.int64), self.value[m]

    def sorted_rows(self) -> "Ell":
        """Same entries with each row sorted by column (canonical form)."""
        out = Ell.zeros(self.N, self.M, self.dtype)
        m = self.live_mask()
        key = np.where(m, self.index, np.iinfo(np.int32).max).astype(np.int64)
        order = np.argsort(key, axis=1, kind="stable")
        out.index[:] = np.take_along_axis(self.index, order, 1)
        out.value[:] = np.take_along_axis(self.value, order, 1)
        out.nnz[:] = self.nnz
        dead = ~out.live_mask()
        out.index[dead] = 0
        out.value[dead] = 0
        return out

    @classmethod
    def from_dense(cls, d: np.ndarray, M: int, threshold: float = 0.0) -> "Ell":
        """Keep entries with |d| > threshold, in column order."""
        N = d.shape[0]
        out = cls.zeros(N, M, d.dtype)
        for i in range(N):
            cols = np.nonzero(np.abs(d[i]) > threshold)[0]
            if len(cols) > M:
                raise ValueError("row does not fit M")
            out.index[i, : len(cols)] = cols
            out.value[i, : len(cols)] = d[i, cols]
            out.nnz[i] = len(cols)
        return out

    @classmethod
    def from_coo(cls, N: int, M: int, rows, cols, vals, dtype=None) -> "Ell":
        """Build from COO triplets (no duplicates), rows sorted by column."""
        rows = np.asarray(rows, np.int64)
        cols = np.asarray(cols, np.int64)
        vals = np.asarray(vals)
        dtype = dtype or vals.dtype
        order = np.lexsort((cols, rows))
        rows, cols, vals = rows[order], cols[order], vals[order]
        counts = np.bincount(rows, minlength=N).astype(np.int64)
        if N and counts.max(initial=0) > M:
            raise ValueError(f"row needs {counts.max()} slots, M={M}")
        start = np.concatenate([[0], np.cumsum(counts)[:-1]])
        slot = np.arange(len(rows)) - start[rows]
        out = cls.zeros(N, M, dtype)
        out.index[rows, slot] = cols
        out.value[rows, slot] = vals
        out.nnz[:] = counts
        return out

    # -- ELL1 file format ---------------------------------------------------
    def write(self, path: str) -> None:
        with open(path, "wb") as f:
            f.write(b"ELL1")
            f.write(struct.pack("<3i", self.precision, self.N, self.M))
            f.write(np.ascontiguousarray(self.value).tobytes())
            f.write(np.ascontiguousarray(self.index, np.int32).tobytes())
            f.write(np.ascontiguousarray(self.nnz, np.int32).tobytes())

    @classmethod
    def read(cls, path: str) -> "Ell":
        with open(path, "rb") as f:
            if f.read(4) != b"ELL1":
                raise ValueError("not an ELL1 file")
            prec, N, M = struct.unpack("<3i", f.read(12))
            dt = np.dtype(DTYPES[prec])
            value = np.frombuffer(f.read(dt.itemsize * N * M), dt).reshape(N, M).copy()
            index = np.frombuffer(f.read(4 * N * M), np.int32).reshape(N, M).copy()
            nnz = np.frombuffer(f.read(4 * N), np.int32).copy()
        return cls(value, index, nnz)


@dataclass
class Blk:
    """An ELLBLOCK matrix in host memory (one bmax x bmax tile per slot)."""

    value: np.ndarray  # [NB, MB, bmax, bmax]
    indexb: np.ndarray  # [NB, MB] int32
    nnzb: np.ndarray  # [NB] int32
    bsize: np.ndarray = field(default=None)  # [NB] int32

    @property
    def NB(self) -> int:
        return self.value.shape[0]

    @property
    def MB(self) -> int:
        return self.value.shape[1]

    @property
    def bmax(self) -> int:
        return self.value.shape[2]

    @property
    def N(self) -> int:
        return int(self.bsize.sum())

    @property
    def dtype(self):
        return self.value.dtype

    @property
    def precision(self) -> int:
        return precision_of(self.value.dtype)

    @property
    def offsets(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(self.bsize)]).astype(np.int64)

    @classmethod
    def zeros(cls, bsize, MB: int, dtype=np.float64, bmax: int | None = None) -> "Blk":
        bsize = np.asarray(bsize, np.int32)
        NB = len(bsize)
        bmax = int(bmax or bsize.max())
        return cls(np.zeros((NB, MB, bmax, bmax), dtype),
                   np.full((NB, MB), -1, np.int32), np.zeros(NB, np.int32), bsize)

    def copy(self) -> "Blk":
        return Blk(self.value.copy(), self.indexb.copy(), self.nnzb.copy(),
                   self.bsize.copy())

    def astype(self, dtype) -> "Blk":
        return Blk(self.value.astype(dtype), self.indexb.copy(), self.nnzb.copy(),
                   self.bsize.copy())

    def to_dense(self) -> np.ndarray:
        off = self.offsets
        d = np.zeros((self.N, self.N), self.dtype)
        for ib in range(self.NB):
            for jp in range(int(self.nnzb[ib])):
                jb = int(self.indexb[ib, jp])
                bi, bj = int(self.bsize[ib]), int(self.bsize[jb])
                d[off[ib]: off[ib] + bi, off[jb]: off[jb] + bj] += \
                    self.value[ib, jp, :bi, :bj]
        return d

    def block_set(self) -> set:
        return {(ib, int(self.indexb[ib, jp])) for ib in range(self.NB)
                for jp in range(int(self.nnzb[ib]))}

    @classmethod
    def from_dense(cls, d: np.ndarray, bsize, MB: int, threshold: float = 0.0,
                   bmax: int | None = None) -> "Blk":
        """Keep blocks whose max |element| exceeds threshold, plus diagonal blocks."""
        out = cls.zeros(bsize, MB, d.dtype, bmax)
        off = out.offsets
        for ib in range(out.NB):
            n = 0
            for jb in range(out.NB):
                blk = d[off[ib]: off[ib + 1], off[jb]: off[jb + 1]]
                if ib == jb or np.abs(blk).max(initial=0) > threshold:
                    if n >= MB:
                        raise ValueError("block row does not fit MB")
                    out.value[ib, n, : blk.shape[0], : blk.shape[1]] = blk
                    out.indexb[ib, n] = jb
                    n += 1
            out.nnzb[ib] = n
        return out

    def write(self, path: str) -> None:
        with open(path, "wb") as f:
            f.write(b"BLK1")
            f.write(struct.pack("<4i", self.precision, self.NB, self.MB, self.bmax))
            f.write(np.ascontiguousarray(self.bsize, np.int32).tobytes())
            f.write(np.ascontiguousarray(self.value).tobytes())
            f.write(np.ascontiguousarray(self.indexb, np.int32).tobytes())
            f.write(np.ascontiguousarray(self.nnzb, np.int32).tobytes())

    @classmethod
    def read(cls, path: str) -> "Blk":
        with open(path, "rb") as f:
            if f.read(4) != b"BLK1":
                raise ValueError("not a BLK1 file")
            prec, NB, MB, bmax = struct.unpack("<4i", f.read(16))
            dt = np.dtype(DTYPES[prec])
            bsize = np.frombuffer(f.read(4 * NB), np.int32).copy()
            value = np.frombuffer(f.read(dt.itemsize * NB * MB * bmax * bmax), dt)
            value = value.reshape(NB, MB, bmax, bmax).copy()
            indexb = np.frombuffer(f.read(4 * NB * MB), np.int32).reshape(NB, MB).copy()
            nnzb = np.frombuffer(f.read(4 * NB), np.int32).copy()
        return cls(value, indexb, nnzb, bsize)
