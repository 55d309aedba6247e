"""The unordered ELLPACK row engine (bml_hip_set_row_engine(UNORDERED)).

Several wavefronts share each row's table and add products with LDS atomics, so
a column's sum is formed in arrival order and the row comes out in table
order.  Each result is checked against the oracle (the reference's sums) as
sets per row: values within 1e-12 x max|ref| in double precision (2e-5 in
single), sparsity patterns equal except entries that close to the
threshold, traces to the same relative tolerance.  Rows that overflow the
LDS table are finished by the exact global-table pass.
"""
import numpy as np
import pytest

import orc
from bml_amd import generators as G
from helpers import DTYPES, assert_ell_close, rel_close

pytestmark = pytest.mark.gpu


def vtol(dtype):
    return 1e-12 if np.dtype(dtype) in (np.float64, np.complex128) else 2e-5


@pytest.fixture
def unordered(hip):
    hip.set_row_engine(hip.ROW_ENGINE_UNORDERED)
    assert hip.get_row_engine() == hip.ROW_ENGINE_UNORDERED
    yield hip
    hip.set_row_engine(hip.ROW_ENGINE_EXACT)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(600, 160, 120), (1500, 320, 300), (257, 128, 100)])
def test_x2_unordered_matches_oracle(unordered, dtype, shape):
    """Rows of B of 100-300 entries (64-lane rows; the longest need the
    chunks beyond the first three)."""
    hip = unordered
    N, M, fill = shape
    A = G.random_ell(N, M, fill, 300 + N, dtype)
    thr = 0.05
    ref, rtr = orc.x2(A, N, thr)
    X2 = hip.zero_matrix(A.precision, N, N)
    tr = hip.multiply_x2(hip.from_ell(A), X2, thr)
    assert_ell_close(X2.get(), ref, vtol(dtype), thr, f"x2 {shape}")
    assert rel_close(tr[0], rtr[0], vtol(dtype)) and rel_close(tr[1], rtr[1], vtol(dtype))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("table", [192, 704])
def test_unordered_table_overflow_goes_to_exact_pass(unordered, monkeypatch, dtype, table):
    """A table too small for the rows: reservations fail, the rows go to the
    global-table pass (exact engine) and the result is still the product."""
    hip = unordered
    A = G.random_ell(800, 160, 120, 5, dtype)
    ref, rtr = orc.x2(A, 800, 0.05)
    monkeypatch.setenv("BML_HIP_TABLE", str(table))
    X2 = hip.zero_matrix(A.precision, A.N, 800)
    tr = hip.multiply_x2(hip.from_ell(A), X2, 0.05)
    assert_ell_close(X2.get(), ref, vtol(dtype), 0.05, f"table {table}")
    assert rel_close(tr[1], rtr[1], vtol(dtype))


@pytest.mark.parametrize("waves,bucket", [(1, 2), (2, 2), (4, 2), (8, 2), (4, 4), (8, 4)])
def test_unordered_wavefronts_per_row(unordered, monkeypatch, waves, bucket):
    hip = unordered
    A = G.random_ell(1200, 256, 220, 9, np.float64)
    ref, rtr = orc.x2(A, 1200, 0.1)
    monkeypatch.setenv("BML_HIP_WAVES", str(waves))
    monkeypatch.setenv("BML_HIP_BUCKET", str(bucket))
    X2 = hip.zero_matrix(A.precision, A.N, 1200)
    tr = hip.multiply_x2(hip.from_ell(A), X2, 0.1)
    assert_ell_close(X2.get(), ref, 1e-12, 0.1, f"waves {waves} bucket {bucket}")
    assert rel_close(tr[1], rtr[1], 1e-12)


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_multiply_ab_unordered(unordered, dtype):
    """C = A B for two different matrices (bml_multiply with A != B), then
    alpha, beta through the add."""
    hip = unordered
    A = G.random_ell(700, 160, 110, 21, dtype)
    B = G.random_ell(700, 160, 120, 22, dtype)
    ref = orc.ab(A, B, 700, 0.02)
    C = hip.zero_matrix(A.precision, 700, 700)
    hip.multiply_AB(hip.from_ell(A), hip.from_ell(B), C, 0.02)
    assert_ell_close(C.get(), ref, 1e-12, 0.02, "AB")


def test_sp2_unordered_matches_exact(hip):
    """SP2 on the 12^3 lattice: the unordered engine's solve lands on the
    exact engine's density matrix (its sums round in arrival order, so the
    stopping iteration may differ by one when the idempotency meets the
    tolerance within rounding)."""
    H = G.lattice_hamiltonian(12, 12, 12, M=16, hopping=-0.1, disorder=0.01)
    dH = hip.from_ell(H)
    out = {}
    for eng in (hip.ROW_ENGINE_EXACT, hip.ROW_ENGINE_UNORDERED):
        hip.set_row_engine(eng)
        try:
            X = hip.zero_matrix(H.precision, H.N, H.N)
            r, _ = hip.sp2(dH, X, H.N / 2, threshold=1e-7, tol=1e-9, maxit=60)
        finally:
            hip.set_row_engine(hip.ROW_ENGINE_EXACT)
        out[eng] = (r, X.get().to_dense())
    (re_, De), (ru, Du) = out[hip.ROW_ENGINE_EXACT], out[hip.ROW_ENGINE_UNORDERED]
    assert abs(ru.iterations - re_.iterations) <= 1 and ru.converged == re_.converged
    assert abs(ru.trace_x - re_.trace_x) <= 1e-9
    assert np.abs(Du - De).max() <= 1e-9


def test_batch_unordered_matches_exact(hip):
    """Batched SP2 on both engines.  The unordered engine's sums round in
    arrival order, so a system whose idempotency reaches the tolerance within
    rounding of it may stop one iteration apart; the density matrices agree
    to 1e-9 either way (one more step moves a converged X by about the
    tolerance)."""
    shapes = [(4, 4, 4), (6, 6, 4), (8, 8, 8), (6, 6, 6), (10, 8, 6)]
    Hs = [G.lattice_hamiltonian(*L, M=16, hopping=-0.1, disorder=0.01, seed=40 + k)
          for k, L in enumerate(shapes)]
    dHs = [hip.from_ell(H) for H in Hs]
    nocc = [H.N / 2 for H in Hs]
    res = {}
    for eng in (hip.ROW_ENGINE_EXACT, hip.ROW_ENGINE_UNORDERED):
        hip.set_row_engine(eng)
        try:
            Xs = [hip.zero_matrix(H.precision, H.N, 512) for H in Hs]
            r = hip.sp2_batch(dHs, Xs, nocc, threshold=1e-6, tol=1e-9, maxit=50)
        finally:
            hip.set_row_engine(hip.ROW_ENGINE_EXACT)
        res[eng] = (r, [X.get().to_dense() for X in Xs])
    (re_, De), (ru, Du) = res[hip.ROW_ENGINE_EXACT], res[hip.ROW_ENGINE_UNORDERED]
    for a, b, da, db in zip(re_, ru, De, Du):
        assert abs(a.iterations - b.iterations) <= 1 and a.converged == b.converged
        assert np.abs(da - db).max() <= 1e-9


def test_unordered_empty_rows_no_diagonal_zero_threshold(unordered):
    """Rows of ~110 entries (64-lane rows) without diagonals, every third row
    empty, threshold 0."""
    hip = unordered
    A = G.random_ell(400, 128, 110, 23, np.float64, diag=False)
    A.nnz[::3] = 0
    ref, rtr = orc.x2(A, 400, 0.0)
    X2 = hip.zero_matrix(A.precision, 400, 400)
    tr = hip.multiply_x2(hip.from_ell(A), X2, 0.0)
    got = X2.get()
    np.testing.assert_array_equal(got.nnz, ref.nnz)
    assert_ell_close(got, ref, 1e-12, 0.0, "empty rows")
    # traces: per-row values summed in a fixed tree (the reference sums in
    # row order)
    assert rel_close(tr[0], rtr[0], 1e-13) and rel_close(tr[1], rtr[1], 1e-12)


def test_unordered_m_overflow_raises(unordered):
    hip = unordered
    A = G.random_ell(300, 128, 110, 99, np.float64)
    X2 = hip.zero_matrix(A.precision, 300, 64)
    with pytest.raises(hip.BmlError) as e:
        hip.multiply_x2(hip.from_ell(A), X2, 0.0)
    assert e.value.code == hip.ERR_M_OVERFLOW
    X2b = hip.zero_matrix(A.precision, 300, 300)
    hip.multiply_x2(hip.from_ell(A), X2b, 0.0)


def test_sp2_complex_unordered_matches_exact(hip):
    """Hermitian (Peierls-phase) 8^3 lattice: the two engines' complex
    double SP2 density matrices agree."""
    H = G.peierls_hamiltonian(8, M=16)
    dH = hip.from_ell(H)
    out = {}
    for eng in (hip.ROW_ENGINE_EXACT, hip.ROW_ENGINE_UNORDERED):
        hip.set_row_engine(eng)
        try:
            X = hip.zero_matrix(H.precision, H.N, 512)
            r, _ = hip.sp2(dH, X, H.N / 2, threshold=1e-8, tol=1e-9, maxit=60)
        finally:
            hip.set_row_engine(hip.ROW_ENGINE_EXACT)
        out[eng] = (r, X.get().to_dense())
    (re_, De), (ru, Du) = out[hip.ROW_ENGINE_EXACT], out[hip.ROW_ENGINE_UNORDERED]
    assert abs(ru.iterations - re_.iterations) <= 1  # rounding at the tolerance
    assert np.abs(Du - De).max() <= 1e-9


def test_batch_unordered_graph_matches_eager(hip):
    """Graph replays of the unordered engine's batched iteration give the
    eager launches' density matrices (to rounding: the atomics' order)."""
    shapes = [(6, 6, 4), (8, 8, 8), (10, 8, 6)]
    Hs = [G.lattice_hamiltonian(*L, M=16, hopping=-0.1, disorder=0.01, seed=60 + k)
          for k, L in enumerate(shapes)]
    dHs = [hip.from_ell(H) for H in Hs]
    nocc = [H.N / 2 for H in Hs]
    outs = {}
    hip.set_row_engine(hip.ROW_ENGINE_UNORDERED)
    try:
        for graph in (False, True):
            Xs = [hip.zero_matrix(H.precision, H.N, 512) for H in Hs]
            res = hip.sp2_batch(dHs, Xs, nocc, threshold=1e-6, tol=1e-9, maxit=50, graph=graph)
            outs[graph] = (res, [X.get().to_dense() for X in Xs])
    finally:
        hip.set_row_engine(hip.ROW_ENGINE_EXACT)
    for a, b, da, db in zip(outs[False][0], outs[True][0], outs[False][1], outs[True][1]):
        assert abs(a.iterations - b.iterations) <= 1  # rounding at the tolerance
        assert np.abs(da - db).max() <= 1e-9


def _symmetric_ell(N, fill, seed, dtype):
    """Random symmetric (Hermitian) ELLPACK matrix with ~2 fill entries per
    row and the diagonal."""
    rng = np.random.default_rng(seed)
    D = np.zeros((N, N), dtype=dtype)
    for i in range(N):
        cols = rng.choice(N, size=fill, replace=False)
        v = rng.uniform(-1, 1, fill)
        if np.iscomplexobj(D):
            v = v + 1j * rng.uniform(-1, 1, fill)
        D[i, cols] = v
    D = D + D.conj().T
    D[np.arange(N), np.arange(N)] = rng.uniform(0.5, 1.0, N)
    from bml_amd.ell import Ell
    M = int((np.abs(D) > 0).sum(axis=1).max())
    return Ell.from_dense(D, M)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("table", [None, 320])
def test_x2_symmetric_path_matches_oracle(unordered, monkeypatch, dtype, table):
    """The symmetric X^2 (upper triangle formed, lower mirrored, products
    packed into full chunks) against the oracle; the small table sends rows
    through the global-table pass, which then forms the upper part too."""
    hip = unordered
    A = _symmetric_ell(700, 60, 5, dtype)
    ref, rtr = orc.x2(A, 700, 0.05)
    monkeypatch.setenv("BML_HIP_ASSUME_SYMMETRIC", "1")
    if table:
        monkeypatch.setenv("BML_HIP_TABLE", str(table))
    X2 = hip.zero_matrix(A.precision, A.N, 700)
    tr = hip.multiply_x2(hip.from_ell(A), X2, 0.05)
    got = X2.get()
    assert_ell_close(got, ref, vtol(dtype), 0.05, f"symmetric {np.dtype(dtype)} table {table}")
    assert rel_close(tr[0], rtr[0], vtol(dtype)) and rel_close(tr[1], rtr[1], vtol(dtype))
    # the result is exactly symmetric (Hermitian)
    D = got.to_dense()
    assert np.array_equal(D, D.conj().T)


def test_x2_symmetric_path_m_overflow(unordered, monkeypatch):
    hip = unordered
    A = _symmetric_ell(400, 60, 7, np.float64)
    monkeypatch.setenv("BML_HIP_ASSUME_SYMMETRIC", "1")
    X2 = hip.zero_matrix(A.precision, A.N, 64)
    with pytest.raises(hip.BmlError) as e:
        hip.multiply_x2(hip.from_ell(A), X2, 0.0)
    assert e.value.code == hip.ERR_M_OVERFLOW


def test_sp2_nonsymmetric_h_uses_full_product(hip):
    """An H that is not exactly symmetric keeps the full X^2 on the unordered
    engine: the iterate is not mirrored into an exactly symmetric matrix, and
    the solve stays near the exact engine's (the non-normal iteration
    amplifies rounding, so only to 1e-5)."""
    H = G.lattice_hamiltonian(8, 8, 8, M=16, hopping=-0.1, disorder=0.01)
    e = int(np.nonzero(H.index[3, :H.nnz[3]] != 3)[0][0])
    H.value[3, e] += 1e-9  # break the symmetry of one off-diagonal entry
    dH = hip.from_ell(H)
    out = {}
    for eng in (hip.ROW_ENGINE_EXACT, hip.ROW_ENGINE_UNORDERED):
        hip.set_row_engine(eng)
        try:
            X = hip.zero_matrix(H.precision, H.N, H.N)
            r, _ = hip.sp2(dH, X, H.N / 2, threshold=1e-7, tol=1e-9, maxit=40)
        finally:
            hip.set_row_engine(hip.ROW_ENGINE_EXACT)
        out[eng] = X.get().to_dense()
    assert np.abs(out[0] - out[1]).max() <= 1e-5
    assert not np.array_equal(out[1], out[1].T)
