"""Measured launch shapes (bml_hip_set_autotune): the first product of each
kind of operand times candidate shapes and keeps the fastest, in memory and
in a cache file.  The exact engine gives the same bits in every shape, so a
tuned product or SP2 solve is required to equal the oracle's (x2) and the
untuned solve (SP2) bit for bit; the window engine's sums follow the
atomics' order, so its tuned solve is compared to rounding."""
import numpy as np
import pytest

import orc
from bml_amd import generators as G
from helpers import assert_ell_identical

pytestmark = pytest.mark.gpu


@pytest.fixture
def tuned(hip, tmp_path):
    path = tmp_path / "tune.txt"
    hip.autotune_clear()
    hip.set_autotune(True, str(path))
    yield hip, path
    hip.set_autotune(False, None)
    hip.autotune_clear()


def test_autotuned_x2_is_bit_identical_and_cached(tuned):
    hip, path = tuned
    A = G.random_banded(N=2048, M=128, seed=5)
    ref, rtr = orc.x2(A, 512, 1e-5)
    X2 = hip.zero_matrix(A.precision, A.N, 512)
    tr = hip.multiply_x2(hip.from_ell(A), X2, 1e-5)
    assert_ell_identical(X2.get(), ref, "tuned x2")
    assert hip.autotune_entries() == 1
    assert len(path.read_text().splitlines()) == 1
    # the same kind of operand again: the cached shape, no search
    X2 = hip.zero_matrix(A.precision, A.N, 512)
    hip.multiply_x2(hip.from_ell(A), X2, 1e-5)
    assert_ell_identical(X2.get(), ref, "cached shape")
    assert hip.autotune_entries() == 1
    # a later process reads the file back
    hip.autotune_clear()
    hip.set_autotune(True, str(path))
    assert hip.autotune_entries() == 1
    X2 = hip.zero_matrix(A.precision, A.N, 512)
    hip.multiply_x2(hip.from_ell(A), X2, 1e-5)
    assert_ell_identical(X2.get(), ref, "shape from the file")
    assert len(path.read_text().splitlines()) == 1


def test_autotuned_ab_matches_oracle(tuned):
    hip, _ = tuned
    A = G.random_ell(3000, 64, 30, 3, np.float32)
    B = G.random_ell(3000, 64, 30, 4, np.float32)
    ref = orc.ab(A, B, 1200, 0.01)
    C = hip.zero_matrix(A.precision, A.N, 1200)
    hip.multiply_AB(hip.from_ell(A), hip.from_ell(B), C, 0.01)
    assert_ell_identical(C.get(), ref, "tuned AB")
    assert hip.autotune_entries() == 1


def test_autotuned_sp2_exact_engine_same_bits(hip, tmp_path):
    """SP2 on the exact engine, every iteration's shape measured (the fill
    grows, so new kinds of operand keep coming), against the untuned solve."""
    H = G.lattice_hamiltonian(8, 8, 8, M=16, hopping=-0.1, disorder=0.01)
    dH = hip.from_ell(H)
    X0 = hip.zero_matrix(H.precision, H.N, 512)
    r0, h0 = hip.sp2(dH, X0, H.N / 2, threshold=1e-7, tol=1e-9, maxit=20)
    hip.autotune_clear()
    hip.set_autotune(True, str(tmp_path / "tune.txt"))
    try:
        X1 = hip.zero_matrix(H.precision, H.N, 512)
        r1, h1 = hip.sp2(dH, X1, H.N / 2, threshold=1e-7, tol=1e-9, maxit=20)
        assert hip.autotune_entries() >= 2
    finally:
        hip.set_autotune(False, None)
        hip.autotune_clear()
    assert r1.iterations == r0.iterations
    np.testing.assert_array_equal(h1, h0)
    assert_ell_identical(X1.get(), X0.get(), "tuned SP2")


def test_autotuned_sp2_window_engine(hip, tmp_path):
    """The fused window step with measured window widths and row blocks:
    the same branches as the untuned solve, X to rounding."""
    H = G.lattice_hamiltonian(10, 10, 10, M=64, hopping=-0.5, disorder=0.05)
    dH = hip.from_ell(H)
    hip.set_row_engine(hip.ROW_ENGINE_WINDOW)
    try:
        X0 = hip.zero_matrix(H.precision, H.N, 1000)
        r0, h0 = hip.sp2(dH, X0, 500.0, threshold=1e-5, tol=1e-8, maxit=25)
        hip.autotune_clear()
        hip.set_autotune(True, str(tmp_path / "tune.txt"))
        X1 = hip.zero_matrix(H.precision, H.N, 1000)
        r1, h1 = hip.sp2(dH, X1, 500.0, threshold=1e-5, tol=1e-8, maxit=25)
        assert hip.autotune_entries() >= 2
    finally:
        hip.set_autotune(False, None)
        hip.autotune_clear()
        hip.set_row_engine(hip.ROW_ENGINE_EXACT)
    assert r1.iterations == r0.iterations
    np.testing.assert_array_equal(h1[:, 3], h0[:, 3])
    assert np.abs(X1.get().to_dense() - X0.get().to_dense()).max() <= 1e-9
