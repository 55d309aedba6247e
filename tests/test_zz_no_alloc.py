"""bml_multiply_x2_ellpack_hip_noalloc (no counterpart in bml: for callers
that cannot allocate) against bml_multiply_x2_ellpack_hip, the entry point
it varies, on two calls in a row, the second on a smaller input, through the
C ABI.  Both run the same engine, so the results agree to the engine's own
bar: bit for bit on the exact engine (the reference's sums), within
1e-12 x max|X^2| on the unordered engine (sums in arrival order)."""
import numpy as np
import pytest

from bml_amd import generators as G
from helpers import assert_ell_close, assert_ell_identical, rel_close

pytestmark = pytest.mark.gpu


@pytest.fixture
def engine(hip):
    yield hip
    hip.set_row_engine(hip.ROW_ENGINE_EXACT)


@pytest.mark.parametrize("row_engine", ["exact", "unordered"])
@pytest.mark.parametrize("dtype", [np.float64, np.complex128, np.float32])
def test_noalloc_matches_x2_two_calls(engine, row_engine, dtype):
    hip = engine
    hip.set_row_engine(hip.ROW_ENGINE_EXACT if row_engine == "exact"
                       else hip.ROW_ENGINE_UNORDERED)
    for N, M, fill, seed in ((3000, 128, 100, 61), (700, 64, 40, 62)):
        A = G.random_ell(N, M, fill, seed, dtype)
        dA = hip.from_ell(A)
        ref = hip.zero_matrix(A.precision, N, N)
        rtr = hip.multiply_x2(dA, ref, 0.01)
        got = hip.zero_matrix(A.precision, N, N)
        tr = hip.multiply_x2_noalloc(dA, got, 0.01)
        if row_engine == "exact":
            assert_ell_identical(got.get(), ref.get(), f"N={N}")
            assert tr[0] == rtr[0] and tr[1] == rtr[1]
        else:
            tol = 1e-12 if np.dtype(dtype) in (np.float64, np.complex128) else 2e-5
            assert_ell_close(got.get(), ref.get(), tol, 0.01, f"N={N}")
            assert rel_close(tr[0], rtr[0], tol) and rel_close(tr[1], rtr[1], tol)


def test_noalloc_second_pass_rows_and_limits(engine):
    """Rows that pass the LDS table (a tiny table forces the global-table
    pass, whose tables also sit in the workspace), then an input whose row
    records alone pass the 64 MiB workspace: refused, the library usable
    afterwards."""
    hip = engine
    A = G.random_ell(2000, 128, 120, 63, np.float64)
    dA = hip.from_ell(A)
    ref = hip.zero_matrix(A.precision, A.N, A.N)
    rtr = hip.multiply_x2(dA, ref, 0.0)
    import os
    os.environ["BML_HIP_TABLE"] = "64"
    try:
        got = hip.zero_matrix(A.precision, A.N, A.N)
        tr = hip.multiply_x2_noalloc(dA, got, 0.0)
    finally:
        del os.environ["BML_HIP_TABLE"]
    assert_ell_identical(got.get(), ref.get(), "global-table pass")
    assert tr == rtr
    big = hip.zero_matrix(A.precision, 3_000_000, 1)
    with pytest.raises(hip.BmlError) as e:
        hip.multiply_x2_noalloc(big, hip.zero_matrix(A.precision, 3_000_000, 1), 0.0)
    assert e.value.code == hip.ERR_UNSUPPORTED
    again = hip.zero_matrix(A.precision, A.N, A.N)
    assert hip.multiply_x2_noalloc(dA, again, 0.0) == rtr
    assert_ell_identical(again.get(), ref.get(), "after the refusal")
