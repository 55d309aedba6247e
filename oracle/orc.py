"""Python face of the parity oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY -- imported by tests/, by __graft_entry__.smoke()
and by bench.py's cpu_baseline leg as the checker; the product library never
imports it.  Each function takes and returns host matrices
(:class:`bml_amd.ell.Ell` / :class:`bml_amd.ell.Blk`) and raises
:class:`OracleError` where the reference would exit (a row overflowing M).
The C functions and their reference citations are in oracle/orc_typed.inc.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from bml_amd.ell import Blk, Ell

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")
SFX = {np.dtype(np.float32): "f32", np.dtype(np.float64): "f64",
       np.dtype(np.complex64): "c64", np.dtype(np.complex128): "c128"}

ERR_M_OVERFLOW = 1
ERR_ALLOC = 2


class OracleError(RuntimeError):
    pass


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _fn(name, dtype, restype=C.c_int):
    f = getattr(lib(), f"orc_{name}_{SFX[np.dtype(dtype)]}")
    f.restype = restype
    return f


def _check(err, what):
    if err == ERR_M_OVERFLOW:
        raise OracleError(f"{what}: number of non-zeroes per row > M")
    if err:
        raise OracleError(f"{what}: error {err}")


def _ell_args(A: Ell):
    return [C.c_int(A.M), _p(A.value), _p(A.index), _p(A.nnz)]


def _d(x):
    return C.c_double(float(x))


# ---- ELLPACK ---------------------------------------------------------------

def x2(X: Ell, M2: int, threshold: float):
    """X2 = X*X thresholded; returns (X2, [tr X, tr X^2])."""
    X2 = Ell.zeros(X.N, M2, X.dtype)
    tr = np.zeros(2)
    err = _fn("x2", X.dtype)(C.c_int(X.N), *_ell_args(X), *_ell_args(X2), _d(threshold), _p(tr))
    _check(err, "x2")
    return X2, tr


def ab(A: Ell, B: Ell, MC: int, threshold: float) -> Ell:
    Cm = Ell.zeros(A.N, MC, A.dtype)
    err = _fn("ab", A.dtype)(C.c_int(A.N), *_ell_args(A), *_ell_args(B), *_ell_args(Cm),
                             _d(threshold))
    _check(err, "multiply_AB")
    return Cm


def multiply(A: Ell, B: Ell | None, Cm: Ell, alpha: float, beta: float,
             threshold: float) -> Ell:
    """C <- alpha*A*B + beta*C (B None means B is A itself); returns new C."""
    Cm = Cm.copy()
    same = B is None
    B = A if same else B
    err = _fn("multiply", A.dtype)(C.c_int(A.N), *_ell_args(A), *_ell_args(B), *_ell_args(Cm),
                                   _d(alpha), _d(beta), _d(threshold), C.c_int(int(same)))
    _check(err, "multiply")
    return Cm


def adjust_ab(A: Ell, B: Ell, MC: int, threshold: float):
    Cm = Ell.zeros(A.N, MC, A.dtype)
    used = np.zeros(1)
    err = _fn("adjust_ab", A.dtype)(C.c_int(A.N), *_ell_args(A), *_ell_args(B), *_ell_args(Cm),
                                    _d(threshold), _p(used))
    _check(err, "multiply_adjust_AB")
    return Cm, float(used[0])


def add(A: Ell, B: Ell, alpha: float, beta: float, threshold: float) -> Ell:
    A = A.copy()
    err = _fn("add", A.dtype)(C.c_int(A.N), *_ell_args(A), *_ell_args(B), _d(alpha), _d(beta),
                              _d(threshold))
    _check(err, "add")
    return A


def add_norm(A: Ell, B: Ell, alpha: float, beta: float, threshold: float):
    A = A.copy()
    n = np.zeros(1)
    err = _fn("add_norm", A.dtype)(C.c_int(A.N), *_ell_args(A), *_ell_args(B), _d(alpha),
                                   _d(beta), _d(threshold), _p(n))
    _check(err, "add_norm")
    return A, float(n[0])


def add_identity(A: Ell, beta: float, threshold: float) -> Ell:
    A = A.copy()
    _check(_fn("add_identity", A.dtype)(C.c_int(A.N), *_ell_args(A), _d(beta), _d(threshold)),
           "add_identity")
    return A


def scale_add_identity(A: Ell, alpha: float, beta: float, threshold: float) -> Ell:
    A = A.copy()
    _check(_fn("scale_add_identity", A.dtype)(C.c_int(A.N), *_ell_args(A), _d(alpha), _d(beta),
                                              _d(threshold)), "scale_add_identity")
    return A


def threshold(A: Ell, thr: float) -> Ell:
    A = A.copy()
    _fn("threshold", A.dtype)(C.c_int(A.N), *_ell_args(A), _d(thr))
    return A


def trace(A: Ell) -> float:
    return float(_fn("trace", A.dtype, C.c_double)(C.c_int(A.N), *_ell_args(A)))


def gershgorin(A: Ell):
    out = np.zeros(2)
    _fn("gershgorin", A.dtype)(C.c_int(A.N), *_ell_args(A), _p(out))
    return float(out[0]), float(out[1])


def normalize(A: Ell, mineval: float, maxeval: float) -> Ell:
    A = A.copy()
    _check(_fn("normalize", A.dtype)(C.c_int(A.N), *_ell_args(A), _d(mineval), _d(maxeval)),
           "normalize")
    return A


def sp2(H: Ell, M: int, nocc: float, threshold: float, tol: float = 1e-8,
        maxit: int = 60):
    """SP2 density matrix; returns (X, iterations, history[it, 4])."""
    X = Ell.zeros(H.N, M, H.dtype)
    hist = np.zeros((maxit, 4))
    it = _fn("sp2", H.dtype)(C.c_int(H.N), *_ell_args(H), *_ell_args(X), _d(nocc),
                             _d(threshold), _d(tol), C.c_int(maxit), _p(hist))
    if it < 0:
        _check(-it, "sp2")
    return X, it, hist[:it]


# ---- ELLBLOCK --------------------------------------------------------------

def _blk_args(A: Blk):
    return [C.c_int(A.MB), _p(A.value), _p(A.indexb), _p(A.nnzb)]


def _blk_head(A: Blk):
    return [C.c_int(A.NB), _p(A.bsize), C.c_int(A.bmax)]


def eb_x2(X: Blk, MB2: int, threshold: float):
    X2 = Blk.zeros(X.bsize, MB2, X.dtype, X.bmax)
    tr = np.zeros(2)
    err = _fn("eb_x2", X.dtype)(*_blk_head(X), *_blk_args(X), *_blk_args(X2), _d(threshold),
                                _p(tr))
    _check(err, "x2 (ellblock)")
    return X2, tr


def eb_ab(A: Blk, B: Blk, MBC: int, threshold: float) -> Blk:
    Cm = Blk.zeros(A.bsize, MBC, A.dtype, A.bmax)
    err = _fn("eb_ab", A.dtype)(*_blk_head(A), *_blk_args(A), *_blk_args(B), *_blk_args(Cm),
                                _d(threshold))
    _check(err, "multiply_AB (ellblock)")
    return Cm


def eb_multiply(A: Blk, B: Blk | None, Cm: Blk, alpha: float, beta: float,
                threshold: float) -> Blk:
    Cm = Cm.copy()
    same = B is None
    B = A if same else B
    err = _fn("eb_multiply", A.dtype)(*_blk_head(A), *_blk_args(A), *_blk_args(B),
                                      *_blk_args(Cm), _d(alpha), _d(beta), _d(threshold),
                                      C.c_int(int(same)))
    _check(err, "multiply (ellblock)")
    return Cm


def eb_add(A: Blk, B: Blk, alpha: float, beta: float, threshold: float) -> Blk:
    A = A.copy()
    err = _fn("eb_add", A.dtype)(*_blk_head(A), *_blk_args(A), *_blk_args(B), _d(alpha),
                                 _d(beta), _d(threshold))
    _check(err, "add (ellblock)")
    return A


def eb_threshold(A: Blk, thr: float) -> Blk:
    A = A.copy()
    _fn("eb_threshold", A.dtype)(*_blk_head(A), *_blk_args(A), _d(thr))
    return A


def eb_trace(A: Blk) -> float:
    return float(_fn("eb_trace", A.dtype, C.c_double)(*_blk_head(A), *_blk_args(A)))


def convert(A, M: int) -> Ell:
    """bml_convert to ellpack (ellpack/bml_convert_ellpack_typed.c:13-40):
    bml_set_element(B, i, j, bml_get_element(A, i, j)) for every (i, j) in
    row-major order.  bml_get_element yields A(i, j) or zero
    (ellpack/bml_getters_ellpack_typed.c:19-55), and bml_set_element appends
    a column not yet in the row (ellpack/bml_setters_ellpack_typed.c:63-107),
    so row i holds all N columns in order, zeros included.  A is an Ell or a
    Blk of B's precision.  N > M raises (the reference writes past the row)."""
    D = A.to_dense()
    N = D.shape[0]
    if N > M:
        raise OracleError("convert: Number of non-zeroes per row > M")
    value = np.zeros((N, M), D.dtype)
    index = np.zeros((N, M), np.int32)
    value[:, :N] = D
    index[:, :N] = np.arange(N, dtype=np.int32)
    return Ell(value, index, np.full(N, N, np.int32))


def eb_convert(A, bsize, bmax: int | None = None) -> Blk:
    """bml_convert to ellblock (ellblock/bml_convert_ellblock_typed.c:16-72):
    the block grid of bsize (bml_get_block_sizes(N, N)) with MB = NB and
    every block stored, zeros included, in block-column order, each filled
    element by element from bml_get_element(A, i, j)."""
    D = A.to_dense()
    bsize = np.asarray(bsize, np.int32)
    NB = len(bsize)
    if int(bsize.sum()) != D.shape[0]:
        raise OracleError("eb_convert: block sizes do not cover N")
    out = Blk.zeros(bsize, NB, D.dtype, bmax or int(bsize.max()))
    off = np.concatenate([[0], np.cumsum(bsize)])
    for ib in range(NB):
        out.nnzb[ib] = NB
        for jb in range(NB):
            out.indexb[ib, jb] = jb
            out.value[ib, jb, :bsize[ib], :bsize[jb]] = D[off[ib]:off[ib + 1], off[jb]:off[jb + 1]]
    return out
