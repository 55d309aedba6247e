"""bml_convert for ellpack_hip and ellblock_hip (ellpack/bml_convert_ellpack_typed.c:
13-40, ellblock/bml_convert_ellblock_typed.c:16-72) against the reference's
output (tests/golden/convert.npz) bit for bit, and against the oracle where
the block-size registry of this process decides the block grid."""
import numpy as np
import pytest

import orc
from helpers import (TAG, DTYPES, assert_blk_identical, assert_ell_identical, blk_of, ell_of,
                     load_cases)

pytestmark = pytest.mark.gpu
CONV = load_cases("convert.npz")


@pytest.mark.parametrize("dtype", DTYPES)
def test_convert_to_ellpack_matches_reference(hip, dtype):
    d = CONV[f"cv_{TAG[dtype]}"]
    A, B = ell_of(d, "A"), blk_of(d, "B")
    got = hip.convert(hip.from_ell(A), A.precision, 24).get()
    assert got.M == 24
    assert_ell_identical(got, ell_of(d, "ell"), "ellpack -> ellpack")
    got = hip.convert(hip.from_blk(B), B.precision, 26).get()
    assert_ell_identical(got, ell_of(d, "blk"), "ellblock -> ellpack")


@pytest.mark.parametrize("dtype", DTYPES)
def test_convert_to_ellblock(hip, dtype):
    """Every block of the registered grid stored; with the registry unset
    the default 4-row blocks of N = 22 (the reference's output itself),
    otherwise the oracle on the registered sizes."""
    d = CONV[f"cv_{TAG[dtype]}"]
    A = ell_of(d, "A")
    if hip._call("bml_hip_get_nb") == 0:
        ref = blk_of(d, "eb")
        got = hip.eb_convert(hip.from_ell(A), A.precision, 22)
        assert_blk_identical(got.get(ref.bmax), ref, "ellpack -> ellblock")
        return
    bs = hip.block_sizes(22, 22)
    N = int(bs.sum())
    from bml_amd import generators as G
    A = G.random_ell(N, 8, 4, 57, dtype)
    ref = orc.eb_convert(A, bs)
    got = hip.eb_convert(hip.from_ell(A), A.precision, N).get(ref.bmax)
    assert_blk_identical(got, ref, "ellpack -> ellblock (registered sizes)")


def test_convert_casts_precision_and_checks_m(hip):
    """Another precision: each element converted as C converts it (the
    reference reinterprets the element's bytes there); N > M is the M
    overflow."""
    d = CONV["cv_f64"]
    A = ell_of(d, "A")
    dA = hip.from_ell(A)
    got = hip.convert(dA, 1, 24).get()  # to single real
    np.testing.assert_array_equal(got.to_dense(), A.to_dense().astype(np.float32))
    got = hip.convert(dA, 4, 24).get()  # to double complex
    np.testing.assert_array_equal(got.to_dense(), A.to_dense().astype(np.complex128))
    with pytest.raises(hip.BmlError) as e:
        hip.convert(dA, A.precision, A.N - 1)
    assert e.value.code == hip.ERR_M_OVERFLOW
