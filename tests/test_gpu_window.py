"""The window row engine (bml_hip_set_row_engine(WINDOW), and the unordered
setting's choice for squares whose rows touch more columns than an LDS hash
table holds).

Each output row adds its products into a dense LDS window of 2^10..2^13
columns per window it touches, and rows come out sorted by column.  The
one-row kernel sums doubles in arrival order; the row blocks (2 or 4 rows)
sum 64-bit fixed-point units, so their results do not depend on the order
the adds arrive in.  Results are compared with the oracle (the
reference's sums, bml_multiply_ellpack_typed.c:130-349 and the SP2 steps a
bml client runs) as sets per row: values within 1e-12 x max|ref| in double
precision (2e-5 in single), patterns equal except entries that close to the
threshold, traces to the same relative tolerance.
"""
import numpy as np
import pytest

import orc
from bml_amd import generators as G
from bml_amd.ell import Ell
from helpers import DTYPES, assert_ell_close, rel_close

pytestmark = pytest.mark.gpu


def vtol(dtype):
    return 1e-12 if np.dtype(dtype) in (np.float64, np.complex128) else 2e-5


@pytest.fixture
def window(hip):
    hip.set_row_engine(hip.ROW_ENGINE_WINDOW)
    assert hip.get_row_engine() == hip.ROW_ENGINE_WINDOW
    yield hip
    hip.set_row_engine(hip.ROW_ENGINE_EXACT)


@pytest.fixture(params=[1, 2, 4])
def rows(request, monkeypatch):
    """Output rows per block: 1 (k_spgemm_win) or blocks of 2 and 4 rows
    sharing each streamed row of X (k_win_union + k_spgemm_winb)."""
    monkeypatch.setenv("BML_HIP_WIN_ROWS", str(request.param))
    return request.param


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(600, 160, 120), (2500, 128, 100), (257, 64, 40)])
def test_x2_window_matches_oracle(window, rows, dtype, shape):
    hip = window
    N, M, fill = shape
    A = G.random_ell(N, M, fill, 700 + N, dtype)
    thr = 0.05
    ref, rtr = orc.x2(A, N, thr)
    X2 = hip.zero_matrix(A.precision, N, N)
    tr = hip.multiply_x2(hip.from_ell(A), X2, thr)
    got = X2.get()
    assert_ell_close(got, ref, vtol(dtype), thr, f"x2 {shape}")
    for i in range(N):  # the window engine writes each row sorted by column
        cols = got.index[i, : got.nnz[i]]
        assert np.all(np.diff(cols) > 0), f"row {i} not sorted"
    assert rel_close(tr[0], rtr[0], vtol(dtype)) and rel_close(tr[1], rtr[1], vtol(dtype))


@pytest.mark.parametrize("rows_per_block", [2, 4])
@pytest.mark.parametrize("dtype", DTYPES)
def test_row_block_sums_are_order_independent(window, monkeypatch, rows_per_block, dtype):
    """Row blocks sum in 64-bit fixed point: integer adds are associative, so
    the same product repeated gives the same bits whatever order the
    wavefronts' LDS adds arrive in (the double-precision one-row kernel's
    last bits may change between runs)."""
    hip = window
    monkeypatch.setenv("BML_HIP_WIN_ROWS", str(rows_per_block))
    A = G.random_ell(3000, 200, 160, 91, dtype)
    dA = hip.from_ell(A)
    outs = []
    for _ in range(3):
        X2 = hip.zero_matrix(A.precision, A.N, 3000)
        tr = hip.multiply_x2(dA, X2, 1e-3)
        outs.append((X2.get(), tuple(tr)))
    for got, tr in outs[1:]:
        assert tr == outs[0][1]
        np.testing.assert_array_equal(got.nnz, outs[0][0].nnz)
        np.testing.assert_array_equal(got.index, outs[0][0].index)
        assert got.value.tobytes() == outs[0][0].value.tobytes()
    ref, _ = orc.x2(A, 3000, 1e-3)
    assert_ell_close(outs[0][0], ref, vtol(dtype), 1e-3, "fixed-point x2")


@pytest.mark.parametrize("rows_per_block", [2, 4])
@pytest.mark.parametrize("dtype", DTYPES)
def test_union_builders_give_the_same_bits(window, monkeypatch, rows_per_block, dtype):
    """The row blocks' column union comes from an LDS hash (its slot order)
    or, for N <= 2^16, from a bitmap of the columns (column order).  The
    fixed-point sums do not depend on the order the union lists the
    streamed rows in, so both builders give the same bits, on a product
    whose rows span several windows and on one with empty rows."""
    hip = window
    monkeypatch.setenv("BML_HIP_WIN_ROWS", str(rows_per_block))
    for N, M, fill, seed in ((5000, 192, 150, 17), (700, 96, 60, 18)):
        A = G.random_ell(N, M, fill, seed, dtype)
        A.nnz[::7] = 0  # empty rows
        dA = hip.from_ell(A)
        outs = {}
        for mode in ("0", "1"):
            monkeypatch.setenv("BML_HIP_WIN_UNION", mode)
            X2 = hip.zero_matrix(A.precision, N, N)
            tr = hip.multiply_x2(dA, X2, 1e-3)
            outs[mode] = (X2.get(), tuple(tr))
        (g0, t0), (g1, t1) = outs["0"], outs["1"]
        assert t0 == t1
        np.testing.assert_array_equal(g0.nnz, g1.nnz)
        np.testing.assert_array_equal(g0.index, g1.index)
        assert g0.value.tobytes() == g1.value.tobytes()
        ref, _ = orc.x2(A, N, 1e-3)
        assert_ell_close(g1, ref, vtol(dtype), 1e-3, f"bitmap union x2 N={N}")


@pytest.mark.parametrize("rows,wbits,waves", [(1, 10, 8), (1, 11, 8), (1, 12, 8), (1, 13, 8),
                                               (1, 10, 4), (1, 12, 4), (2, 10, 8), (2, 11, 8),
                                               (4, 10, 8), (4, 11, 8), (4, 12, 8)])
def test_window_sizes(window, monkeypatch, rows, wbits, waves):
    """Windows of 2^10..2^13 columns over N = 9000 (2..9 windows), four or
    eight wavefronts per block, blocks of 1, 2 or 4 rows (4 rows with
    2^12-column windows run the one-row kernel: the row blocks are built
    for windows up to 2^11)."""
    hip = window
    A = G.random_ell(9000, 96, 80, 31, np.float64)
    ref, rtr = orc.x2(A, 9000, 0.02)
    monkeypatch.setenv("BML_HIP_WIN_ROWS", str(rows))
    monkeypatch.setenv("BML_HIP_WINDOW_BITS", str(wbits))
    monkeypatch.setenv("BML_HIP_WAVES", str(waves))
    X2 = hip.zero_matrix(A.precision, A.N, 4096)
    tr = hip.multiply_x2(hip.from_ell(A), X2, 0.02)
    assert_ell_close(X2.get(), ref, 1e-12, 0.02, f"window 2^{wbits} waves {waves}")
    assert rel_close(tr[0], rtr[0], 1e-12) and rel_close(tr[1], rtr[1], 1e-12)


def test_window_empty_rows_and_no_diagonal(window, rows):
    """Rows with no entries, rows whose product never reaches the diagonal
    (no diagonal entry is stored then, as the reference's first-touch list
    never holds it), N = 1."""
    hip = window
    A = G.random_ell(700, 64, 30, 41, np.float64, diag=False)
    for i in range(0, 700, 7):
        A.nnz[i] = 0
    ref, rtr = orc.x2(A, 700, 0.01)
    X2 = hip.zero_matrix(A.precision, 700, 700)
    tr = hip.multiply_x2(hip.from_ell(A), X2, 0.01)
    assert_ell_close(X2.get(), ref, 1e-12, 0.01, "empty rows")
    assert rel_close(tr[1], rtr[1], 1e-12)
    one = G.random_ell(1, 4, 1, 42, np.float64)
    ref1, _ = orc.x2(one, 1, 0.0)
    Y = hip.zero_matrix(one.precision, 1, 4)
    hip.multiply_x2(hip.from_ell(one), Y, 0.0)
    assert_ell_close(Y.get(), ref1, 1e-12, 0.0, "N=1")


def test_window_m_overflow_raises(window, rows):
    """A product row with more kept entries than X2's M is the reference's
    fatal "Number of non-zeroes per row > M" (bml_multiply_ellpack_typed.c:
    325-330); the library stays usable afterwards."""
    hip = window
    A = G.random_ell(500, 64, 60, 43, np.float64)
    X2 = hip.zero_matrix(A.precision, 500, 32)
    with pytest.raises(hip.BmlError) as e:
        hip.multiply_x2(hip.from_ell(A), X2, 0.0)
    assert e.value.code == 1
    ref, _ = orc.x2(A, 500, 0.0)
    Y = hip.zero_matrix(A.precision, 500, 500)
    hip.multiply_x2(hip.from_ell(A), Y, 0.0)
    assert_ell_close(Y.get(), ref, 1e-12, 0.0, "after overflow")


@pytest.mark.parametrize("dtype", [np.float64, np.complex128, np.float32])
def test_window_symmetric_x2(window, rows, monkeypatch, dtype):
    """The symmetric path the SP2 drivers take (upper triangle formed, the
    rest mirrored): a lattice SP2 iterate, squared."""
    hip = window
    if np.dtype(dtype).kind == "c":
        H = G.peierls_hamiltonian(L=8, M=64, dtype=dtype)
    else:
        H = G.lattice_hamiltonian(10, 10, 10, M=64, hopping=-0.5, disorder=0.05, dtype=dtype)
    # X = the normalized H, squared twice (a symmetric matrix with wide
    # rows), then made exactly symmetric (Hermitian): the oracle's squares
    # agree with their transposes only to rounding, and the mirrored path
    # takes X to be exact, as the SP2 iterates are
    lo, hi = orc.gershgorin(H)
    X = orc.normalize(H.with_m(H.N), lo, hi)
    X, _ = orc.x2(X, X.N, 1e-6)
    X, _ = orc.x2(X, X.N, 1e-6)
    D = X.to_dense()
    X = Ell.from_dense((D + D.conj().T) / 2, X.N)
    ref, rtr = orc.x2(X, X.N, 1e-6)
    monkeypatch.setenv("BML_HIP_ASSUME_SYMMETRIC", "1")
    X2 = hip.zero_matrix(X.precision, X.N, X.N)
    tr = hip.multiply_x2(hip.from_ell(X), X2, 1e-6)
    got = X2.get()
    assert_ell_close(got, ref, vtol(dtype), 1e-6, "symmetric x2")
    D = got.to_dense()
    assert np.array_equal(D, D.conj().T), "mirrored product not exactly Hermitian"
    assert rel_close(tr[1], rtr[1], vtol(dtype))


def test_sp2_window_matches_oracle_hop05(window, rows):
    """SP2 on the 10^3 lattice with the hopping of the 32^3 SP2 workload
    (-0.5 +- 0.05 disorder, onsite +-1, threshold 1e-5): the fused window
    step (X^2's upper triangle, the update, the norm, in place) against the
    oracle's SP2 (bml_multiply_x2 + bml_add_norm per iteration, the
    reference's sums).  The branch history agrees exactly; X, traces and the
    idempotency measure agree to rounding (1e-12 x max|X|; measured on the
    oracle alone, a 1e-15 perturbation of H moves this solve's X by 1e-15)."""
    hip = window
    H = G.lattice_hamiltonian(10, 10, 10, M=64, hopping=-0.5, disorder=0.05)
    ref, rit, rh = orc.sp2(H, 1000, 500.0, 1e-5, tol=1e-8, maxit=40)
    X = hip.zero_matrix(H.precision, H.N, 1000)
    r, h = hip.sp2(hip.from_ell(H), X, 500.0, threshold=1e-5, tol=1e-8, maxit=40)
    assert r.iterations == rit
    np.testing.assert_array_equal(h[:, 3], rh[:, 3], err_msg="branch history")
    np.testing.assert_allclose(h[:, :2], rh[:, :2], rtol=1e-12)
    np.testing.assert_allclose(h[:, 2], rh[:, 2], rtol=1e-9, atol=1e-13)
    assert_ell_close(X.get(), ref, 1e-12, 1e-5, "SP2 density matrix")


def test_sp2_window_two_windows(window, rows, monkeypatch):
    """SP2 on the 16^3 lattice (hopping -0.1, threshold 1e-7: ~620 entries
    per row) with windows of 2^11 columns (two windows per row range), 30
    iterations, against the oracle: the same branches, X within
    1e-9 x max|X| (both stop at 30 iterations; the threshold holds the
    idempotency near 1e-6)."""
    hip = window
    monkeypatch.setenv("BML_HIP_WINDOW_BITS", "11")
    H = G.lattice_hamiltonian(16, 16, 16, M=16, hopping=-0.1, disorder=0.01)
    ref, rit, rh = orc.sp2(H, 1024, H.N / 2, 1e-7, tol=1e-9, maxit=30)
    X = hip.zero_matrix(H.precision, H.N, 1024)
    r, h = hip.sp2(hip.from_ell(H), X, H.N / 2, threshold=1e-7, tol=1e-9, maxit=30)
    assert r.iterations == rit == 30
    np.testing.assert_array_equal(h[:, 3], rh[:, 3], err_msg="branch history")
    assert_ell_close(X.get(), ref, 1e-9, 1e-7, "SP2 two windows")


def test_sp2_window_mixed_against_oracle(window, rows):
    """Mixed precision on the window engine (single-precision iterations,
    FP64 refinement) against the oracle's FP64 SP2 on the 10^3 lattice at
    hopping -0.5: the density matrix within 2e-5 of the FP64 one (the
    single-precision phase leaves about eps_f32 x ||H|| / gap in the occupied
    subspace, which FP64 polynomial steps cannot rotate back) and the
    threshold keeping or dropping entries near 1e-5 differently in the two
    precisions: X within 5 x threshold of the FP64 one (measured 1.8e-5),
    the trace within 1e-6 of the oracle's."""
    hip = window
    H = G.lattice_hamiltonian(10, 10, 10, M=64, hopping=-0.5, disorder=0.05)
    ref, rit, rh = orc.sp2(H, 1000, 500.0, 1e-5, tol=1e-8, maxit=40)
    X = hip.zero_matrix(H.precision, H.N, 1000)
    r, h = hip.sp2(hip.from_ell(H), X, 500.0, threshold=1e-5, tol=1e-8, maxit=40, mixed=True)
    assert r.single_iterations > 0 and r.iterations == 40
    D, R = X.get().to_dense(), ref.to_dense()
    assert np.abs(D - R).max() <= 5e-5
    assert abs(np.trace(D) - np.trace(R)) <= 1e-6 * 500
