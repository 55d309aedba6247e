"""Entry points with a single call site in bml (no other test reaches them):
bml_threshold_new_ellpack (ellpack/bml_threshold_ellpack_typed.c:28-78),
bml_scale_inplace_ellpack (ellpack/bml_scale_ellpack_typed.c), the ELLBLOCK
scale_inplace (ellblock/bml_scale_ellblock_typed.c:60-87) and
add_identity (ellblock/bml_add_ellblock_typed.c:354-370), each against the
oracle or the elementwise closed form."""
import numpy as np
import pytest

import orc
from bml_amd import generators as G
from bml_amd.ell import Blk
from helpers import DTYPES, assert_blk_identical, assert_ell_identical

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", DTYPES)
def test_threshold_new(hip, dtype):
    """A new matrix with the entries |v| > thr of A, in A's order (the
    diagonal is not protected); A is left unchanged."""
    A = G.random_ell(64, 24, 6, 12, dtype)
    dA = hip.from_ell(A)
    B = hip.threshold_new(dA, 0.5)
    assert B.M == A.M
    assert_ell_identical(B.get(), orc.threshold(A, 0.5), "threshold_new")
    assert_ell_identical(dA.get(), A, "A unchanged")


def _scaled(v, s):
    """s * v elementwise, one product per element (BLAS ?scal); complex as
    (ar xr - ai xi, ar xi + ai xr) with each product rounded."""
    if np.iscomplexobj(v):
        a = np.asarray(s, v.dtype)
        re = a.real * v.real - a.imag * v.imag
        im = a.real * v.imag + a.imag * v.real
        return (re + 1j * im).astype(v.dtype)
    return (np.asarray(s, v.dtype) * v).astype(v.dtype)


@pytest.mark.parametrize("dtype", DTYPES)
def test_scale_inplace_ellpack(hip, dtype):
    A = G.random_ell(64, 24, 6, 13, dtype)
    s = 0.7 - 0.3j if np.iscomplexobj(A.value) else -0.7
    dA = hip.from_ell(A)
    hip.scale_inplace(s, dA)
    got = dA.get()
    want = A.copy()
    want.value = _scaled(A.value, s)
    np.testing.assert_array_equal(got.nnz, A.nnz)
    m = A.live_mask()
    np.testing.assert_array_equal(got.index[m], A.index[m])
    if np.iscomplexobj(A.value):  # the reference's zscal may fuse; one ulp
        eps = np.finfo(A.value.real.dtype).eps
        assert np.abs(got.value[m] - want.value[m]).max() <= 2 * eps * np.abs(want.value[m]).max()
    else:
        np.testing.assert_array_equal(got.value[m], want.value[m])


@pytest.mark.parametrize("dtype", DTYPES)
def test_scale_inplace_ellblock(hip, dtype):
    A = G.random_blk([3, 5, 2, 4, 1, 6], 4, 2, 14, dtype)
    s = 0.7 - 0.3j if np.iscomplexobj(A.value) else -0.7
    dA = hip.from_blk(A)
    hip.eb_scale_inplace(s, dA)
    got = dA.get(A.bmax)
    np.testing.assert_array_equal(got.nnzb, A.nnzb)
    D, W = got.to_dense(), A.to_dense()
    W = _scaled(W, s)
    if np.iscomplexobj(D):
        eps = np.finfo(D.real.dtype).eps
        assert np.abs(D - W).max() <= 2 * eps * np.abs(W).max()
    else:
        np.testing.assert_array_equal(D, W)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("thr", [0.0, 0.3])
def test_add_identity_ellblock(hip, dtype, thr):
    """A <- A + beta I as bml_add_ellblock(A, Id, 1, beta, thr) with Id the
    identity of A's block sizes: the oracle's blockwise add, bit for bit;
    block rows without a diagonal block gain one."""
    bs = [3, 5, 2, 4, 1, 6]
    A = G.random_blk(bs, 6, 2, 15, dtype)
    for ib in range(0, len(bs), 2):  # drop the diagonal block of some rows
        keep = [jp for jp in range(int(A.nnzb[ib])) if int(A.indexb[ib, jp]) != ib]
        n = len(keep)
        A.value[ib, :n] = A.value[ib, keep]
        A.indexb[ib, :n] = A.indexb[ib, keep]
        A.value[ib, n:] = 0
        A.indexb[ib, n:] = -1
        A.nnzb[ib] = n
    Id = Blk.zeros(bs, 1, A.dtype, A.bmax)
    for ib, b in enumerate(bs):
        Id.nnzb[ib] = 1
        Id.indexb[ib, 0] = ib
        Id.value[ib, 0, :b, :b] = np.eye(b)
    ref = orc.eb_add(A, Id, 1.0, 0.75, thr)
    dA = hip.from_blk(A)
    hip.eb_add_identity(dA, 0.75, thr)
    assert_blk_identical(dA.get(A.bmax), ref, "add_identity")
    if thr == 0.0:  # the oracle itself against the dense A + 0.75 I
        np.testing.assert_allclose(ref.to_dense(), A.to_dense() + 0.75 * np.eye(A.N),
                                   rtol=0, atol=1e-6)
