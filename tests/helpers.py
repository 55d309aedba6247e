"""Shared comparison helpers for the parity tests."""
import os

import numpy as np

from bml_amd.ell import Blk, Ell

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DTYPES = [np.float32, np.float64, np.complex64, np.complex128]
TAG = {np.float32: "f32", np.float64: "f64", np.complex64: "c64", np.complex128: "c128"}


def load_cases(fname):
    z = np.load(os.path.join(GOLDEN, fname), allow_pickle=False)
    cases = {}
    for k in z.files:
        case, key = k.split("/", 1)
        cases.setdefault(case, {})[key] = z[k]
    return cases


def ell_of(d, prefix):
    return Ell(d[prefix + "_value"].copy(), d[prefix + "_index"].copy(),
               d[prefix + "_nnz"].copy())


def blk_of(d, prefix):
    return Blk(d[prefix + "_value"].copy(), d[prefix + "_indexb"].copy(),
               d[prefix + "_nnzb"].copy(), d[prefix + "_bsize"].copy())


def live(A: Ell):
    """(nnz, per-row index lists, per-row value lists) of the live entries."""
    m = A.live_mask()
    return A.nnz, [A.index[i, : A.nnz[i]] for i in range(A.N)], \
        [A.value[i, : A.nnz[i]] for i in range(A.N)]


def assert_ell_identical(got: Ell, ref: Ell, what=""):
    """Same nnz, same columns in the same order, bit-identical values."""
    assert got.N == ref.N, what
    np.testing.assert_array_equal(got.nnz, ref.nnz, err_msg=f"{what}: nnz")
    for i in range(ref.N):
        n = int(ref.nnz[i])
        np.testing.assert_array_equal(got.index[i, :n], ref.index[i, :n],
                                      err_msg=f"{what}: row {i} columns")
        a = np.ascontiguousarray(got.value[i, :n])
        b = np.ascontiguousarray(ref.value[i, :n])
        if not np.array_equal(a.view(np.uint8), b.view(np.uint8)):
            bad = np.nonzero(a != b)[0]
            raise AssertionError(f"{what}: row {i} values differ at slots {bad[:5]}: "
                                 f"{a[bad[:3]]} vs {b[bad[:3]]}")


def assert_ell_close(got: Ell, ref: Ell, rtol_max: float, thr: float = 0.0, what=""):
    """Values within rtol_max * max|ref|; patterns equal except entries within
    rtol_max * max|ref| of the threshold (compared as sets per row)."""
    scale = max(float(np.abs(ref.value[ref.live_mask()]).max(initial=0.0)), 1e-300)
    tol = rtol_max * scale
    gd = dict_rows(got)
    rd = dict_rows(ref)
    for i in range(ref.N):
        g, r = gd[i], rd[i]
        for c in set(g) | set(r):
            gv, rv = g.get(c, 0.0), r.get(c, 0.0)
            if c in g and c in r:
                assert abs(gv - rv) <= tol, f"{what}: ({i},{c}) {gv} vs {rv}"
            else:
                v = gv if c in g else rv
                assert abs(abs(v) - thr) <= tol, \
                    f"{what}: ({i},{c}) only in {'got' if c in g else 'ref'}: {v}"


def dict_rows(A: Ell):
    return [dict(zip(A.index[i, : A.nnz[i]].tolist(), A.value[i, : A.nnz[i]].tolist()))
            for i in range(A.N)]


def assert_blk_identical(got: Blk, ref: Blk, what=""):
    np.testing.assert_array_equal(got.nnzb, ref.nnzb, err_msg=f"{what}: nnzb")
    for ib in range(ref.NB):
        n = int(ref.nnzb[ib])
        np.testing.assert_array_equal(got.indexb[ib, :n], ref.indexb[ib, :n],
                                      err_msg=f"{what}: block row {ib}")
        b = ref.bmax
        a = np.ascontiguousarray(got.value[ib, :n, :b, :b])
        r = np.ascontiguousarray(ref.value[ib, :n, :b, :b])
        assert np.array_equal(a.view(np.uint8), r.view(np.uint8)), f"{what}: block row {ib} values"


def rel_close(a, b, rtol):
    return abs(a - b) <= rtol * max(abs(a), abs(b), 1e-300)
