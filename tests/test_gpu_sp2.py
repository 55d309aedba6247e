"""SP2 purification on the GPU (device-side branch and convergence test)
against the oracle's SP2, the reference-driven SP2 golden vectors and the
spectral projector."""
import numpy as np
import pytest

import orc
from bml_amd import generators as G
from helpers import assert_ell_identical, ell_of, load_cases

pytestmark = pytest.mark.gpu

SP2 = load_cases("sp2.npz")


def run_sp2(hip, H, M, nocc, **kw):
    dH = hip.from_ell(H)
    X = hip.zero_matrix(H.precision, H.N, M)
    r, hist = hip.sp2(dH, X, nocc, **kw)
    return X.get(), r, hist


def test_sp2_matches_reference_driven_golden(hip):
    d = SP2["sp2_f64"]
    H = ell_of(d, "H")
    X, r, hist = run_sp2(hip, H, int(d["M"]), float(d["nocc"]), threshold=float(d["threshold"]),
                         tol=float(d["tol"]), maxit=int(d["maxit"]))
    ref_hist = d["hist"]
    assert r.iterations == len(ref_hist)
    np.testing.assert_array_equal(hist[:, 3], ref_hist[:, 3])
    assert_ell_identical(X, ell_of(d, "X"), "sp2 X")
    np.testing.assert_allclose(hist[:, :3], ref_hist[:, :3], rtol=1e-10, atol=1e-13)


@pytest.mark.parametrize("L,hop", [(8, -0.1), (10, -0.25)])
def test_sp2_matches_oracle(hip, L, hop):
    H = G.lattice_hamiltonian(L, L, L, M=16, hopping=hop, disorder=0.1 * abs(hop))
    M = H.N
    ref, it, rhist = orc.sp2(H, M, H.N / 2, 1e-5, 1e-8, 25)
    X, r, hist = run_sp2(hip, H, M, H.N / 2, threshold=1e-5, tol=1e-8, maxit=25)
    assert r.iterations == it
    np.testing.assert_array_equal(hist[:, 3], rhist[:, 3])
    assert_ell_identical(X, ref, "sp2 vs oracle")
    assert abs(r.emin - orc.gershgorin(H)[0]) == 0


def test_sp2_converges_to_projector(hip):
    H = G.lattice_hamiltonian(6, 6, 6, M=216, hopping=-0.1, disorder=0.01)
    N = H.N
    X, r, hist = run_sp2(hip, H, N, N / 2, threshold=0.0, tol=1e-10, maxit=60)
    w, V = np.linalg.eigh(H.to_dense())
    P = V[:, : N // 2] @ V[:, : N // 2].T
    assert r.converged and r.idempotency <= 1e-10
    assert np.abs(X.to_dense() - P).max() < 1e-9


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_sp2_mixed_precision_refines_to_double(hip, dtype):
    if dtype == np.float64:
        H = G.lattice_hamiltonian(8, 8, 8, M=16, hopping=-0.1, disorder=0.01)
    else:
        H = G.peierls_hamiltonian(6, M=16, amplitude=0.1)
    N = H.N
    M = 256
    Xd, rd, hd = run_sp2(hip, H, M, N / 2, threshold=1e-5, tol=1e-8, maxit=30)
    Xm, rm, hm = run_sp2(hip, H, M, N / 2, threshold=1e-5, tol=1e-8, maxit=30, mixed=True,
                         refine_iters=4)
    assert rm.single_iterations == 26 and rm.iterations == 30
    # the double-precision refinement brings the trace back to nocc
    assert abs(rm.trace_x - N / 2) <= 1e-6 * N
    assert abs(rd.trace_x - N / 2) <= 1e-6 * N
    assert rm.idempotency <= 10 * max(rd.idempotency, 1e-7)
    diff = np.abs(Xm.to_dense() - Xd.to_dense()).max()
    assert diff < 1e-4
