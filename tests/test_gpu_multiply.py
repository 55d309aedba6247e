"""bml_multiply (C <- alpha*A*B + beta*C) on the fused path: the product's
epilogue merges C's row as the reference's bml_add does
(bml_multiply_ellpack_typed.c:55-118 forms the thresholded product in a
temporary, then bml_add_ellpack_typed.c:162-196 adds it into C), with no
temporary.  Results are required to be bit-identical to the oracle's (the
reference's sums and orders) in every precision, over the scalar cases the
reference's code branches on (alpha or beta zero, A = B), every launch
shape of the exact engine (groups of 8..64 lanes, several wavefronts per
row, rows that overflow the LDS table into the global-table pass), and when
C is A itself (the temporary path).
"""
import numpy as np
import pytest

import orc
from bml_amd import generators as G
from helpers import DTYPES, assert_ell_identical

pytestmark = pytest.mark.gpu

SCALARS = [(1.0, 0.0), (0.5, 1.0), (-1.25, 2.0), (2.0, -1.0), (0.0, 0.75), (1.0, 1.0)]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("alpha,beta", SCALARS)
@pytest.mark.parametrize("same", [True, False])
def test_multiply_matches_oracle(hip, dtype, alpha, beta, same):
    N, M = 500, 48
    A = G.random_ell(N, M, 16, 11, dtype)
    B = G.random_ell(N, M, 16, 12, dtype)
    Cm = G.random_ell(N, M, 20, 13, dtype)
    MO = 400
    thr = 0.01
    ref = orc.multiply(A, None if same else B, Cm.with_m(MO), alpha, beta, thr)
    dA = hip.from_ell(A)
    dC = hip.from_ell(Cm, MO)
    hip.multiply(dA, dA if same else hip.from_ell(B), dC, alpha, beta, thr)
    assert_ell_identical(dC.get(), ref, f"multiply a={alpha} b={beta} same={same}")


@pytest.mark.parametrize("group,table,bucket,waves", [
    (8, 64, 2, 1), (16, 128, 4, 1), (32, 256, 8, 1), (64, 64, 4, 1), (64, 1000, 4, 1),
    (64, 512, 4, 2), (64, 1024, 4, 3)])
@pytest.mark.parametrize("same", [True, False])
def test_multiply_launch_shapes(hip, monkeypatch, group, table, bucket, waves, same):
    """Every exact-engine launch shape: small tables send rows to the
    global-table pass (its epilogue merges C too), waves > 1 runs the
    several-wavefront kernel."""
    A = G.random_banded(N=1024, M=128, seed=7)
    B = G.random_banded(N=1024, M=128, seed=8)
    Cm = G.random_banded(N=1024, M=128, seed=9)
    ref = orc.multiply(A, None if same else B, Cm.with_m(1024), -0.5, 2.0, 1e-5)
    monkeypatch.setenv("BML_HIP_GROUP", str(group))
    monkeypatch.setenv("BML_HIP_TABLE", str(table))
    monkeypatch.setenv("BML_HIP_BUCKET", str(bucket))
    monkeypatch.setenv("BML_HIP_WAVES", str(waves))
    dA = hip.from_ell(A)
    dC = hip.from_ell(Cm, 1024)
    hip.multiply(dA, dA if same else hip.from_ell(B), dC, -0.5, 2.0, 1e-5)
    assert_ell_identical(dC.get(), ref, f"group {group} table {table} waves {waves}")


@pytest.mark.parametrize("engine", ["UNORDERED", "WINDOW"])
def test_multiply_runs_exact_whatever_the_row_engine(hip, engine):
    """The alpha/beta form always runs the exact engine, so it stays
    bit-identical when the unordered or window engine is selected."""
    A = G.random_ell(600, 64, 24, 21, np.float64)
    Cm = G.random_ell(600, 64, 24, 22, np.float64)
    ref = orc.multiply(A, None, Cm.with_m(600), 1.5, -0.5, 1e-3)
    hip.set_row_engine(getattr(hip, "ROW_ENGINE_" + engine))
    try:
        dA = hip.from_ell(A)
        dC = hip.from_ell(Cm, 600)
        hip.multiply(dA, dA, dC, 1.5, -0.5, 1e-3)
    finally:
        hip.set_row_engine(hip.ROW_ENGINE_EXACT)
    assert_ell_identical(dC.get(), ref, f"{engine} setting")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_multiply_into_operand(hip, dtype):
    """C = A (C <- alpha*A*A + beta*A): the reference's temporary keeps the
    operand intact while the product is formed; so does this path."""
    A = G.random_ell(300, 64, 12, 31, dtype)
    ref = orc.multiply(A.with_m(300), None, A.with_m(300), 0.5, -1.0, 1e-4)
    dA = hip.from_ell(A, 300)
    hip.multiply(dA, dA, dA, 0.5, -1.0, 1e-4)
    assert_ell_identical(dA.get(), ref, "C = A")
    B = G.random_ell(300, 64, 12, 32, dtype)
    ref = orc.multiply(A.with_m(300), B.with_m(300), B.with_m(300), 2.0, 0.5, 1e-4)
    dA, dB = hip.from_ell(A, 300), hip.from_ell(B, 300)
    hip.multiply(dA, dB, dB, 2.0, 0.5, 1e-4)
    assert_ell_identical(dB.get(), ref, "C = B")


def test_multiply_empty_rows_and_cancellation(hip):
    """Rows of A or C with no entries; beta*C + alpha*P cancelling exactly
    (C = -P with alpha = beta = 1 leaves nothing above the threshold); the
    diagonal of an x2 product kept even when below the threshold."""
    A = G.random_ell(400, 32, 10, 41, np.float64)
    for i in range(0, 400, 5):
        A.nnz[i] = 0
    P, _ = orc.x2(A, 400, 0.0)
    Cm = P.copy()
    Cm.value = -Cm.value
    ref = orc.multiply(A, None, Cm.with_m(400), 1.0, 1.0, 0.0)
    dA = hip.from_ell(A)
    dC = hip.from_ell(Cm, 400)
    hip.multiply(dA, dA, dC, 1.0, 1.0, 0.0)
    got = dC.get()
    assert_ell_identical(got, ref, "cancellation")
    assert got.nnz.sum() == 0
    Z = G.random_ell(400, 32, 10, 42, np.float64)
    for i in range(0, 400, 3):
        Z.nnz[i] = 0
    ref = orc.multiply(A, None, Z.with_m(400), 1.0, 0.25, 0.5)
    dC = hip.from_ell(Z, 400)
    hip.multiply(dA, dA, dC, 1.0, 0.25, 0.5)
    assert_ell_identical(dC.get(), ref, "empty rows, large threshold")


def test_multiply_m_overflow_raises(hip):
    """A product row with more entries than C's M is the reference's fatal
    "Number of non-zeroes per row > M" (bml_multiply_ellpack_typed.c:325-330);
    so is a merged row that outgrows C (the reference's bml_add would write
    past the row).  The library stays usable afterwards."""
    A = G.random_ell(64, 24, 10, 99, np.float64)
    Cm = G.random_ell(64, 24, 4, 98, np.float64)
    dA = hip.from_ell(A)
    with pytest.raises(hip.BmlError) as e:
        hip.multiply(dA, dA, hip.from_ell(Cm, 8), 0.5, 1.0, 0.0)
    assert e.value.code == 1
    # C as wide as the widest product row: the product fits, its union with
    # C's other columns does not
    P, _ = orc.x2(A, 64, 0.0)
    mp = int(P.nnz.max())
    W = G.random_ell(64, 64, 40, 97, np.float64)
    with pytest.raises(hip.BmlError) as e:
        hip.multiply(dA, dA, hip.from_ell(W, mp), 0.5, 1.0, 0.0)
    assert e.value.code == 1
    ref = orc.multiply(A, None, Cm.with_m(64), 0.5, 1.0, 0.0)
    dC = hip.from_ell(Cm, 64)
    hip.multiply(dA, dA, dC, 0.5, 1.0, 0.0)
    assert_ell_identical(dC.get(), ref, "after overflow")
