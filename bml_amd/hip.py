"""ctypes binding of libbml_hip.so and a thin Python mirror of bml's C API.

The functions here keep bml's names and argument meaning (reference
``src/C-interface/bml_multiply.h``, ``bml_add.h``, ``bml_threshold.h``,
``bml_trace.h``, ``bml_normalize.h``): matrices are handles to device-resident
``ellpack_hip`` / ``ellblock_hip`` matrices, outputs are preallocated by the
caller and overwritten in place, ``multiply_x2`` returns ``(tr X, tr X^2)``.
bml's fatal errors surface as :class:`BmlError` (the library runs in its
record-errors mode under Python).

There is no fallback: if the shared library is missing or the device is not
usable, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .ell import Blk, DTYPES, Ell, precision_of

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbml_hip.so")

ELLPACK_HIP = 6
ELLBLOCK_HIP = 7
ERRORS_EXIT, ERRORS_RECORD = 0, 1
SYNC_EACH_CALL, SYNC_DEFERRED = 0, 1
ERR_M_OVERFLOW = 1
ERR_ARGUMENT = 2
ERR_UNSUPPORTED = 5


class BmlError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class _Ell(C.Structure):
    _fields_ = [("matrix_type", C.c_int), ("matrix_precision", C.c_int),
                ("distribution_mode", C.c_int), ("N", C.c_int), ("M", C.c_int),
                ("value", C.c_void_p), ("index", C.c_void_p), ("nnz", C.c_void_p),
                ("aux", C.c_void_p)]


class _Blk(C.Structure):
    _fields_ = [("matrix_type", C.c_int), ("matrix_precision", C.c_int),
                ("distribution_mode", C.c_int), ("N", C.c_int), ("M", C.c_int),
                ("NB", C.c_int), ("MB", C.c_int), ("bt", C.c_int),
                ("value", C.c_void_p), ("indexb", C.c_void_p), ("nnzb", C.c_void_p),
                ("bsize", C.c_void_p), ("bsize_host", C.c_void_p), ("aux", C.c_void_p)]


class Sp2Options(C.Structure):
    _fields_ = [("threshold", C.c_double), ("tol", C.c_double), ("maxit", C.c_int),
                ("mixed", C.c_int), ("refine_iters", C.c_int), ("poll", C.c_int),
                ("graph", C.c_int), ("mixed_switch", C.c_int), ("resume", C.c_int)]


class Sp2Result(C.Structure):
    _fields_ = [("iterations", C.c_int), ("single_iterations", C.c_int),
                ("converged", C.c_int), ("trace_x", C.c_double), ("trace_x2", C.c_double),
                ("idempotency", C.c_double), ("emin", C.c_double), ("emax", C.c_double)]


class Profile(C.Structure):
    """Per precision (index 1..4): row-product launches, summed kernel ms and
    the work they did."""
    _fields_ = [("launches", C.c_int * 5), ("kernel_ms", C.c_double * 5),
                ("products", C.c_longlong * 5), ("rows", C.c_longlong * 5),
                ("in_nnz", C.c_longlong * 5), ("out_nnz", C.c_longlong * 5),
                ("ovf_rows", C.c_longlong * 5), ("touched", C.c_longlong * 5),
                ("streamed", C.c_longlong * 5)]


_lib = None
_P = C.c_void_p
_I = C.c_int
_D = C.c_double

_SIGS = {
    "bml_hip_init": (_I, [_I]),
    "bml_hip_shutdown": (None, []),
    "bml_hip_set_error_mode": (None, [_I]),
    "bml_hip_set_sync_mode": (None, [_I]),
    "bml_hip_last_error": (_I, [C.c_char_p, _I]),
    "bml_hip_clear_error": (None, []),
    "bml_hip_synchronize": (None, []),
    "bml_hip_stream": (_P, []),
    "bml_hip_version": (C.c_char_p, []),
    "bml_hip_set_autotune": (None, [_I, C.c_char_p]),
    "bml_hip_get_autotune": (_I, []),
    "bml_hip_autotune_entries": (_I, []),
    "bml_hip_autotune_clear": (None, []),
    "bml_zero_matrix_ellpack_hip": (_P, [_I, _I, _I, _I]),
    "bml_noinit_matrix_ellpack_hip": (_P, [_I, _I, _I, _I]),
    "bml_identity_matrix_ellpack_hip": (_P, [_I, _I, _I, _I]),
    "bml_deallocate_ellpack_hip": (None, [_P]),
    "bml_clear_ellpack_hip": (None, [_P]),
    "bml_copy_ellpack_hip_new": (_P, [_P]),
    "bml_copy_ellpack_hip": (None, [_P, _P]),
    "bml_import_from_dense_ellpack_hip": (_P, [_I, _I, _I, _P, _D, _I, _I]),
    "bml_export_to_dense_ellpack_hip": (_P, [_P, _I]),
    "bml_convert_ellpack_hip": (_P, [_P, _I, _I, _I]),
    "bml_multiply_x2_ellpack_hip_noalloc": (None, [_P, _P, _D, _P]),
    "bml_write_bml_matrix_ellpack_hip": (None, [_P, C.c_char_p]),
    "bml_read_bml_matrix_ellpack_hip": (None, [_P, C.c_char_p]),
    "bml_write_bml_matrix_ellblock_hip": (None, [_P, C.c_char_p]),
    "bml_read_bml_matrix_ellblock_hip": (None, [_P, C.c_char_p]),
    "bml_ellpack_hip_set_arrays": (None, [_P, _P, _P, _P, _I]),
    "bml_ellpack_hip_get_arrays": (None, [_P, _P, _P, _P]),
    "bml_multiply_x2_ellpack_hip": (_P, [_P, _P, _D]),
    "bml_multiply_ellpack_hip": (None, [_P, _P, _P, _D, _D, _D]),
    "bml_multiply_AB_ellpack_hip": (None, [_P, _P, _P, _D]),
    "bml_multiply_adjust_AB_ellpack_hip": (None, [_P, _P, _P, _D]),
    "bml_add_ellpack_hip": (None, [_P, _P, _D, _D, _D]),
    "bml_add_norm_ellpack_hip": (_D, [_P, _P, _D, _D, _D]),
    "bml_add_identity_ellpack_hip": (None, [_P, _D, _D]),
    "bml_scale_add_identity_ellpack_hip": (None, [_P, _D, _D, _D]),
    "bml_scale_inplace_ellpack_hip": (None, [_P, _P]),
    "bml_threshold_ellpack_hip": (None, [_P, _D]),
    "bml_threshold_new_ellpack_hip": (_P, [_P, _D]),
    "bml_trace_ellpack_hip": (_D, [_P]),
    "bml_gershgorin_ellpack_hip": (_P, [_P]),
    "bml_normalize_ellpack_hip": (None, [_P, _D, _D]),
    "bml_ellpack_hip_row_stats": (None, [_P, C.POINTER(C.c_int), C.POINTER(C.c_longlong)]),
    "bml_hip_sp2_defaults": (None, [C.POINTER(Sp2Options)]),
    "bml_hip_sp2_ellpack": (None, [_P, _P, _D, C.POINTER(Sp2Options), C.POINTER(Sp2Result), _P]),
    "bml_hip_sp2_ellpack_batch": (None, [_I, _P, _P, _P, C.POINTER(Sp2Options), _P]),
    "bml_hip_profile_begin": (None, []),
    "bml_hip_profile_end": (None, [C.POINTER(Profile)]),
}
_SIGS.update({
    "bml_hip_set_row_engine": (None, [_I]),
    "bml_hip_get_row_engine": (_I, []),
    "bml_hip_set_block_engine": (None, [_I]),
    "bml_hip_set_block_sizes": (None, [_P, _I, _I]),
    "bml_hip_get_block_sizes": (_P, [_I, _I]),
    "bml_hip_get_nb": (_I, []),
    "bml_hip_get_mb": (_I, []),
    "bml_block_matrix_ellblock_hip": (_P, [_I, _I, _I, _I, _P, _I]),
    "bml_zero_matrix_ellblock_hip": (_P, [_I, _I, _I, _I]),
    "bml_deallocate_ellblock_hip": (None, [_P]),
    "bml_clear_ellblock_hip": (None, [_P]),
    "bml_copy_ellblock_hip_new": (_P, [_P]),
    "bml_copy_ellblock_hip": (None, [_P, _P]),
    "bml_ellblock_hip_set_blocks": (None, [_P, _P, _P, _P, _I, _I]),
    "bml_ellblock_hip_get_blocks": (None, [_P, _P, _P, _P, _I]),
    "bml_export_to_dense_ellblock_hip": (_P, [_P, _I]),
    "bml_convert_ellblock_hip": (_P, [_P, _I, _I, _I]),
    "bml_import_from_dense_ellblock_hip": (_P, [_I, _I, _I, _P, C.c_double, _I, _I]),
    "bml_multiply_x2_ellblock_hip": (_P, [_P, _P, _D]),
    "bml_multiply_ellblock_hip": (None, [_P, _P, _P, _D, _D, _D]),
    "bml_multiply_AB_ellblock_hip": (None, [_P, _P, _P, _D]),
    "bml_add_ellblock_hip": (None, [_P, _P, _D, _D, _D]),
    "bml_threshold_ellblock_hip": (None, [_P, _D]),
    "bml_trace_ellblock_hip": (_D, [_P]),
    "bml_gershgorin_ellblock_hip": (_P, [_P]),
    "bml_normalize_ellblock_hip": (None, [_P, _D, _D]),
    "bml_scale_inplace_ellblock_hip": (None, [_P, _P]),
    "bml_add_identity_ellblock_hip": (None, [_P, _D, _D]),
    "bml_hip_sp2_ellblock": (None, [_P, _P, _D, C.POINTER(Sp2Options), C.POINTER(Sp2Result), _P]),
})
_OPTIONAL: dict = {}

BLOCK_ENGINE_AUTO, BLOCK_ENGINE_EXACT, BLOCK_ENGINE_MFMA, BLOCK_ENGINE_MFMA_BF16 = -1, 0, 1, 2
ROW_ENGINE_EXACT, ROW_ENGINE_UNORDERED, ROW_ENGINE_WINDOW = 0, 1, 2


def header_symbols() -> list[str]:
    """Entry points declared in include/*.h (parsed from the headers)."""
    import re
    inc = os.path.join(os.path.dirname(HERE), "include")
    names = []
    for f in sorted(os.listdir(inc)):
        if f.endswith(".h"):
            txt = open(os.path.join(inc, f)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names += re.findall(r"\b(bml_\w+)\s*\(", txt)
    return sorted(set(names))


def lib():
    """Load libbml_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in {**_SIGS, **_OPTIONAL}.items():
            if not hasattr(L, name):
                if name in _SIGS:
                    raise ImportError(f"{LIB_PATH} lacks {name}")
                continue
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        L.bml_hip_set_error_mode(ERRORS_RECORD)
        _libc.free.argtypes = [_P]
        _lib = L
    return _lib


_libc = C.CDLL(None)


def _check():
    buf = C.create_string_buffer(1024)
    code = lib().bml_hip_last_error(buf, 1024)
    if code:
        lib().bml_hip_clear_error()
        raise BmlError(code, buf.value.decode())


def _call(name, *args):
    r = getattr(lib(), name)(*args)
    _check()
    return r


def init(device: int = 0) -> None:
    _call("bml_hip_init", device)


def synchronize() -> None:
    _call("bml_hip_synchronize")


def stream() -> int:
    return _call("bml_hip_stream")


def set_sync_mode(mode: int) -> None:
    lib().bml_hip_set_sync_mode(mode)


class EllpackHip:
    """Handle to a device-resident ellpack_hip matrix."""

    def __init__(self, ptr):
        if not ptr:
            _check()
            raise BmlError(2, "allocation returned NULL")
        self.ptr = C.c_void_p(ptr)

    @property
    def _s(self) -> _Ell:
        return C.cast(self.ptr, C.POINTER(_Ell)).contents

    @property
    def N(self) -> int:
        return self._s.N

    @property
    def M(self) -> int:
        return self._s.M

    @property
    def precision(self) -> int:
        return self._s.matrix_precision

    @property
    def dtype(self):
        return np.dtype(DTYPES[self.precision])

    def free(self) -> None:
        if self.ptr:
            lib().bml_deallocate_ellpack_hip(self.ptr)
            self.ptr = C.c_void_p(None)

    def __del__(self):
        try:
            if _lib is not None and self.ptr:
                self.free()
        except Exception:
            pass

    # host <-> device
    def set(self, A: Ell) -> "EllpackHip":
        if A.N != self.N or A.dtype != self.dtype:
            raise BmlError(2, "shape/precision mismatch")
        v = np.ascontiguousarray(A.value)
        ix = np.ascontiguousarray(A.index, np.int32)
        nn = np.ascontiguousarray(A.nnz, np.int32)
        _call("bml_ellpack_hip_set_arrays", self.ptr, v.ctypes.data, ix.ctypes.data,
              nn.ctypes.data, A.M)
        return self

    def get(self) -> Ell:
        out = Ell.zeros(self.N, self.M, self.dtype)
        _call("bml_ellpack_hip_get_arrays", self.ptr, out.value.ctypes.data,
              out.index.ctypes.data, out.nnz.ctypes.data)
        dead = ~out.live_mask()
        out.value[dead] = 0
        out.index[dead] = 0
        return out

    def row_stats(self):
        mx = C.c_int(0)
        tot = C.c_longlong(0)
        _call("bml_ellpack_hip_row_stats", self.ptr, C.byref(mx), C.byref(tot))
        return mx.value, tot.value


def zero_matrix(precision: int, N: int, M: int) -> EllpackHip:
    return EllpackHip(_call("bml_zero_matrix_ellpack_hip", precision, N, M, 0))


def noinit_matrix(precision: int, N: int, M: int) -> EllpackHip:
    return EllpackHip(_call("bml_noinit_matrix_ellpack_hip", precision, N, M, 0))


def identity_matrix(precision: int, N: int, M: int) -> EllpackHip:
    return EllpackHip(_call("bml_identity_matrix_ellpack_hip", precision, N, M, 0))


def from_ell(A: Ell, M: int | None = None) -> EllpackHip:
    h = zero_matrix(A.precision, A.N, M or A.M)
    return h.set(A)


def import_from_dense(dense: np.ndarray, M: int, threshold: float = 0.0,
                      order: int = 0) -> EllpackHip:
    d = np.ascontiguousarray(dense)
    return EllpackHip(_call("bml_import_from_dense_ellpack_hip", precision_of(d.dtype), order,
                            d.shape[0], d.ctypes.data, threshold, M, 0))


def write_matrix_market(A, path: str) -> None:
    """bml_write_bml_matrix for ellpack_hip / ellblock_hip: a Matrix Market
    coordinate file."""
    kind = "ellblock" if isinstance(A, EllblockHip) else "ellpack"
    _call(f"bml_write_bml_matrix_{kind}_hip", A.ptr, os.fsencode(path))


def read_matrix_market(A, path: str):
    """bml_read_bml_matrix for ellpack_hip / ellblock_hip: adds the file's
    entries to A."""
    kind = "ellblock" if isinstance(A, EllblockHip) else "ellpack"
    _call(f"bml_read_bml_matrix_{kind}_hip", A.ptr, os.fsencode(path))
    return A


def export_to_dense(A: EllpackHip, order: int = 0) -> np.ndarray:
    p = _call("bml_export_to_dense_ellpack_hip", A.ptr, order)
    n = A.N * A.N
    buf = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_ubyte)), shape=(n * A.dtype.itemsize,))
    out = np.frombuffer(buf.tobytes(), A.dtype).reshape(A.N, A.N)
    _libc.free(p)
    return out


def convert(A, precision: int, M: int) -> EllpackHip:
    """bml_convert to ellpack_hip: every element of A (ellpack_hip or
    ellblock_hip) stored, row by row in column order."""
    return EllpackHip(_call("bml_convert_ellpack_hip", A.ptr, precision, M, 0))


def eb_convert(A, precision: int, M: int) -> "EllblockHip":
    """bml_convert to ellblock_hip: every block of the registered block grid
    stored."""
    return EllblockHip(_call("bml_convert_ellblock_hip", A.ptr, precision, M, 0))


def copy_new(A: EllpackHip) -> EllpackHip:
    return EllpackHip(_call("bml_copy_ellpack_hip_new", A.ptr))


def copy(A: EllpackHip, B: EllpackHip) -> None:
    _call("bml_copy_ellpack_hip", A.ptr, B.ptr)


def clear(A: EllpackHip) -> None:
    _call("bml_clear_ellpack_hip", A.ptr)


def _take2(p):
    arr = C.cast(p, C.POINTER(C.c_double))
    out = (arr[0], arr[1])
    _libc.free(p)
    return out


def multiply_x2(X: EllpackHip, X2: EllpackHip, threshold: float):
    """X2 <- X*X; returns (tr X, tr X^2)."""
    return _take2(_call("bml_multiply_x2_ellpack_hip", X.ptr, X2.ptr, threshold))


def multiply_x2_noalloc(X: EllpackHip, X2: EllpackHip, threshold: float):
    """bml_multiply_x2_ellpack_hip_noalloc: the product on the library's
    fixed device workspace; returns (tr X, tr X^2)."""
    tr = np.zeros(2)
    _call("bml_multiply_x2_ellpack_hip_noalloc", X.ptr, X2.ptr, threshold, tr.ctypes.data)
    return float(tr[0]), float(tr[1])


def multiply(A: EllpackHip, B: EllpackHip, Cm: EllpackHip, alpha: float = 1.0,
             beta: float = 0.0, threshold: float = 0.0) -> None:
    _call("bml_multiply_ellpack_hip", A.ptr, B.ptr, Cm.ptr, alpha, beta, threshold)


def multiply_AB(A: EllpackHip, B: EllpackHip, Cm: EllpackHip, threshold: float) -> None:
    _call("bml_multiply_AB_ellpack_hip", A.ptr, B.ptr, Cm.ptr, threshold)


def multiply_adjust_AB(A: EllpackHip, B: EllpackHip, Cm: EllpackHip, threshold: float) -> None:
    _call("bml_multiply_adjust_AB_ellpack_hip", A.ptr, B.ptr, Cm.ptr, threshold)


def add(A: EllpackHip, B: EllpackHip, alpha: float, beta: float, threshold: float) -> None:
    _call("bml_add_ellpack_hip", A.ptr, B.ptr, alpha, beta, threshold)


def add_norm(A: EllpackHip, B: EllpackHip, alpha: float, beta: float,
             threshold: float) -> float:
    return _call("bml_add_norm_ellpack_hip", A.ptr, B.ptr, alpha, beta, threshold)


def add_identity(A: EllpackHip, beta: float, threshold: float) -> None:
    _call("bml_add_identity_ellpack_hip", A.ptr, beta, threshold)


def scale_add_identity(A: EllpackHip, alpha: float, beta: float, threshold: float) -> None:
    _call("bml_scale_add_identity_ellpack_hip", A.ptr, alpha, beta, threshold)


def scale_inplace(scale, A: EllpackHip) -> None:
    s = np.array([scale], A.dtype)
    _call("bml_scale_inplace_ellpack_hip", s.ctypes.data, A.ptr)


def threshold(A: EllpackHip, thr: float) -> None:
    _call("bml_threshold_ellpack_hip", A.ptr, thr)


def threshold_new(A: EllpackHip, thr: float) -> EllpackHip:
    return EllpackHip(_call("bml_threshold_new_ellpack_hip", A.ptr, thr))


def trace(A: EllpackHip) -> float:
    return _call("bml_trace_ellpack_hip", A.ptr)


def gershgorin(A: EllpackHip):
    return _take2(_call("bml_gershgorin_ellpack_hip", A.ptr))


def normalize(A: EllpackHip, mineval: float, maxeval: float) -> None:
    _call("bml_normalize_ellpack_hip", A.ptr, mineval, maxeval)


def sp2(H: EllpackHip, X: EllpackHip, nocc: float, threshold: float = 1e-5,
        tol: float = 1e-8, maxit: int = 60, mixed: bool = False, refine_iters: int = 4,
        poll: int = 1, mixed_switch: bool = True, resume: bool = False, graph: bool = True):
    """X <- SP2 density matrix of H; returns (Sp2Result, history[it, 4]).
    resume: continue from the iterate X holds; graph (with poll > 1): the
    iterations between two polls replay a captured graph."""
    o = Sp2Options()
    lib().bml_hip_sp2_defaults(C.byref(o))
    o.threshold, o.tol, o.maxit = threshold, tol, maxit
    o.mixed, o.refine_iters, o.poll, o.graph = int(mixed), refine_iters, poll, int(graph)
    o.mixed_switch, o.resume = int(mixed_switch), int(resume)
    r = Sp2Result()
    hist = np.zeros((maxit, 4))
    _call("bml_hip_sp2_ellpack", H.ptr, X.ptr, nocc, C.byref(o), C.byref(r), hist.ctypes.data)
    return r, hist[: r.iterations]


def sp2_batch(Hs, Xs, nocc, threshold: float = 1e-5, tol: float = 1e-8, maxit: int = 60,
              mixed: bool = False, refine_iters: int = 4, poll: int = 1, graph: bool = True):
    """Independent SP2 solves of the ellpack_hip pairs (Hs[s], Xs[s]) in one
    batch; returns the list of Sp2Result."""
    n = len(Hs)
    o = Sp2Options()
    lib().bml_hip_sp2_defaults(C.byref(o))
    o.threshold, o.tol, o.maxit = threshold, tol, maxit
    o.mixed, o.refine_iters, o.poll, o.graph = int(mixed), refine_iters, poll, int(graph)
    hp = (C.c_void_p * n)(*[h.ptr.value for h in Hs])
    xp = (C.c_void_p * n)(*[x.ptr.value for x in Xs])
    oc = np.ascontiguousarray(np.broadcast_to(np.asarray(nocc, np.float64), (n,)))
    res = (Sp2Result * n)()
    _call("bml_hip_sp2_ellpack_batch", n, C.cast(hp, C.c_void_p), C.cast(xp, C.c_void_p),
          oc.ctypes.data, C.byref(o), C.cast(res, C.c_void_p))
    return list(res)


def profile_begin() -> None:
    _call("bml_hip_profile_begin")


def profile_end() -> Profile:
    p = Profile()
    _call("bml_hip_profile_end", C.byref(p))
    return p


# ---- ELLBLOCK ---------------------------------------------------------------

class EllblockHip:
    """Handle to a device-resident ellblock_hip matrix."""

    def __init__(self, ptr):
        if not ptr:
            _check()
            raise BmlError(2, "allocation returned NULL")
        self.ptr = C.c_void_p(ptr)

    @property
    def _s(self) -> _Blk:
        return C.cast(self.ptr, C.POINTER(_Blk)).contents

    @property
    def N(self) -> int:
        return self._s.N

    @property
    def NB(self) -> int:
        return self._s.NB

    @property
    def MB(self) -> int:
        return self._s.MB

    @property
    def bt(self) -> int:
        return self._s.bt

    @property
    def precision(self) -> int:
        return self._s.matrix_precision

    @property
    def dtype(self):
        return np.dtype(DTYPES[self.precision])

    @property
    def bsize(self) -> np.ndarray:
        return np.ctypeslib.as_array(C.cast(self._s.bsize_host, C.POINTER(C.c_int)),
                                     shape=(self.NB,)).copy()

    def free(self) -> None:
        if self.ptr:
            lib().bml_deallocate_ellblock_hip(self.ptr)
            self.ptr = C.c_void_p(None)

    def __del__(self):
        try:
            if _lib is not None and self.ptr:
                self.free()
        except Exception:
            pass

    def set(self, A: Blk) -> "EllblockHip":
        if A.NB != self.NB or A.dtype != self.dtype:
            raise BmlError(2, "shape/precision mismatch")
        v = np.ascontiguousarray(A.value)
        ix = np.ascontiguousarray(A.indexb, np.int32)
        nn = np.ascontiguousarray(A.nnzb, np.int32)
        _call("bml_ellblock_hip_set_blocks", self.ptr, v.ctypes.data, ix.ctypes.data,
              nn.ctypes.data, A.MB, A.bmax)
        return self

    def get(self, bmax: int | None = None) -> Blk:
        bsize = self.bsize
        bmax = int(bmax or bsize.max())
        out = Blk.zeros(bsize, self.MB, self.dtype, bmax)
        _call("bml_ellblock_hip_get_blocks", self.ptr, out.value.ctypes.data,
              out.indexb.ctypes.data, out.nnzb.ctypes.data, bmax)
        for ib in range(self.NB):
            out.indexb[ib, out.nnzb[ib]:] = 0
        return out


def set_block_engine(engine: int) -> None:
    _call("bml_hip_set_block_engine", engine)


def set_row_engine(engine: int) -> None:
    """bml_hip_set_row_engine: ROW_ENGINE_EXACT, ROW_ENGINE_UNORDERED or ROW_ENGINE_WINDOW."""
    _call("bml_hip_set_row_engine", engine)


def get_row_engine() -> int:
    return int(lib().bml_hip_get_row_engine())


def set_autotune(on: bool, cache_path: str | None = None) -> None:
    """bml_hip_set_autotune: measured launch shapes, cached in cache_path."""
    lib().bml_hip_set_autotune(int(on), cache_path.encode() if cache_path else None)


def get_autotune() -> bool:
    return bool(lib().bml_hip_get_autotune())


def autotune_entries() -> int:
    return int(lib().bml_hip_autotune_entries())


def autotune_clear() -> None:
    lib().bml_hip_autotune_clear()


def block_matrix(precision: int, bsize, MB: int, M: int = 0) -> EllblockHip:
    b = np.ascontiguousarray(bsize, np.int32)
    return EllblockHip(_call("bml_block_matrix_ellblock_hip", precision, len(b), MB,
                             M or int(b.sum()), b.ctypes.data, 0))


def from_blk(A: Blk, MB: int | None = None) -> EllblockHip:
    return block_matrix(A.precision, A.bsize, MB or A.MB).set(A)


def eb_import_from_dense(dense: np.ndarray, M: int, threshold: float = 0.0,
                         order: int = 0) -> EllblockHip:
    """bml_import_from_dense for ellblock_hip (the registered block sizes)."""
    d = np.ascontiguousarray(dense)
    return EllblockHip(_call("bml_import_from_dense_ellblock_hip", precision_of(d.dtype), order,
                             d.shape[0], d.ctypes.data, threshold, M, 0))


def block_sizes(N: int, M: int) -> np.ndarray:
    """The registered block sizes (registering the default for N, M when
    none are set)."""
    p = _call("bml_hip_get_block_sizes", N, M)
    nb = _call("bml_hip_get_nb")
    return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_int)), shape=(nb,)).copy()


def eb_export_to_dense(A: EllblockHip, order: int = 0) -> np.ndarray:
    p = _call("bml_export_to_dense_ellblock_hip", A.ptr, order)
    n = A.N * A.N
    buf = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_ubyte)), shape=(n * A.dtype.itemsize,))
    out = np.frombuffer(buf.tobytes(), A.dtype).reshape(A.N, A.N)
    _libc.free(p)
    return out


def eb_copy_new(A: EllblockHip) -> EllblockHip:
    return EllblockHip(_call("bml_copy_ellblock_hip_new", A.ptr))


def eb_copy(A: EllblockHip, B: EllblockHip) -> None:
    _call("bml_copy_ellblock_hip", A.ptr, B.ptr)


def eb_clear(A: EllblockHip) -> None:
    _call("bml_clear_ellblock_hip", A.ptr)


def eb_multiply_x2(X: EllblockHip, X2: EllblockHip, threshold: float):
    """X2 <- X*X blockwise; returns (tr X, tr X^2)."""
    return _take2(_call("bml_multiply_x2_ellblock_hip", X.ptr, X2.ptr, threshold))


def eb_multiply(A: EllblockHip, B: EllblockHip, Cm: EllblockHip, alpha: float = 1.0,
                beta: float = 0.0, threshold: float = 0.0) -> None:
    _call("bml_multiply_ellblock_hip", A.ptr, B.ptr, Cm.ptr, alpha, beta, threshold)


def eb_multiply_AB(A: EllblockHip, B: EllblockHip, Cm: EllblockHip, threshold: float) -> None:
    _call("bml_multiply_AB_ellblock_hip", A.ptr, B.ptr, Cm.ptr, threshold)


def eb_add(A: EllblockHip, B: EllblockHip, alpha: float, beta: float, threshold: float) -> None:
    _call("bml_add_ellblock_hip", A.ptr, B.ptr, alpha, beta, threshold)


def eb_threshold(A: EllblockHip, thr: float) -> None:
    _call("bml_threshold_ellblock_hip", A.ptr, thr)


def eb_trace(A: EllblockHip) -> float:
    return _call("bml_trace_ellblock_hip", A.ptr)


def eb_gershgorin(A: EllblockHip):
    return _take2(_call("bml_gershgorin_ellblock_hip", A.ptr))


def eb_normalize(A: EllblockHip, mineval: float, maxeval: float) -> None:
    _call("bml_normalize_ellblock_hip", A.ptr, mineval, maxeval)


def eb_scale_inplace(scale, A: EllblockHip) -> None:
    s = np.array([scale], A.dtype)
    _call("bml_scale_inplace_ellblock_hip", s.ctypes.data, A.ptr)


def eb_add_identity(A: EllblockHip, beta: float, threshold: float) -> None:
    _call("bml_add_identity_ellblock_hip", A.ptr, beta, threshold)


def eb_sp2(H: EllblockHip, X: EllblockHip, nocc: float, threshold: float = 1e-5,
           tol: float = 1e-8, maxit: int = 60, mixed: bool = False, refine_iters: int = 4,
           poll: int = 1, mixed_switch: bool = True):
    """X <- SP2 density matrix of the ELLBLOCK H; returns (Sp2Result, history[it, 4])."""
    o = Sp2Options()
    lib().bml_hip_sp2_defaults(C.byref(o))
    o.threshold, o.tol, o.maxit = threshold, tol, maxit
    o.mixed, o.refine_iters, o.poll = int(mixed), refine_iters, poll
    o.mixed_switch = int(mixed_switch)
    r = Sp2Result()
    hist = np.zeros((maxit, 4))
    _call("bml_hip_sp2_ellblock", H.ptr, X.ptr, nocc, C.byref(o), C.byref(r), hist.ctypes.data)
    return r, hist[: r.iterations]
