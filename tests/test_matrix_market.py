"""Matrix Market files (bml_write_bml_matrix / bml_read_bml_matrix for
ellpack, bml_utilities.h:38-45, ellpack/bml_utilities_ellpack_typed.c:22-190).

The fixtures in tests/golden/mm.npz hold the reference's own output: the text
it writes for a random matrix, what it reads back from that text, and what it
reads from a "symmetric" file.  The CPU tests pin the fixture's format
against the C formatting rules it follows; the GPU tests compare the
library's writer byte for byte and its reader bit for bit."""
import os

import numpy as np
import pytest

from helpers import assert_ell_identical, ell_of, load_cases

CASES = load_cases("mm.npz")


def text(d, key):
    return bytes(d[key]).decode()


@pytest.mark.parametrize("case", sorted(CASES))
def test_reference_text_format(case):
    """Header with three '%' (the reference prints "%%%%%%"), "N N nnz",
    then one line per stored entry in row and storage order, 1-based, each
    value as C's %20.15g (real and imaginary parts for complex)."""
    d = CASES[case]
    A = ell_of(d, "A")
    lines = text(d, "general_text").splitlines()
    kind = "complex" if np.iscomplexobj(A.value) else "real"
    assert lines[0] == f"%%%MatrixMarket matrix coordinate {kind} general"
    assert lines[1] == f"{A.N} {A.N} {int(A.nnz.sum())}"
    want = []
    for i in range(A.N):
        for k in range(int(A.nnz[i])):
            v = A.value[i, k]
            vals = [v.real, v.imag] if kind == "complex" else [v]
            want.append(f"{i + 1} {A.index[i, k] + 1} " +
                        " ".join("%20.15g" % float(x) for x in vals))
    assert lines[2:] == want


@pytest.mark.parametrize("case", sorted(CASES))
def test_reference_read_is_the_parsed_text(case):
    """What the reference reads back is the text parsed as strtod (%lf) or
    strtof (%f, single real), rounded to the precision, appended row by row;
    a symmetric file also stores each off-diagonal entry at its mirror."""
    d = CASES[case]
    A = ell_of(d, "A")
    dt = A.value.dtype
    for key, sym in (("general", False), ("symmetric", True)):
        R = ell_of(d, f"{key}_read")
        rows = [[] for _ in range(A.N)]
        for ln in text(d, f"{key}_text").splitlines()[2:]:
            t = ln.split()
            i, j = int(t[0]) - 1, int(t[1]) - 1
            v = complex(float(t[2]), float(t[3])) if len(t) == 4 else float(t[2])
            v = np.array(v).astype(dt)
            rows[i].append((j, v))
            if sym and i != j:
                rows[j].append((i, v))
        for i in range(A.N):
            n = int(R.nnz[i])
            assert n == len(rows[i])
            assert [c for c, _ in rows[i]] == R.index[i, :n].tolist()
            np.testing.assert_array_equal(np.array([v for _, v in rows[i]], dt), R.value[i, :n])


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_write_matches_reference_bytes(hip, tmp_path, case):
    d = CASES[case]
    A = ell_of(d, "A")
    path = os.path.join(tmp_path, "a.mtx")
    hip.write_matrix_market(hip.from_ell(A), path)
    assert open(path, "rb").read() == bytes(d["general_text"])


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_read_matches_reference(hip, tmp_path, case):
    d = CASES[case]
    N, M = int(d["N"]), int(d["M"])
    prec = ell_of(d, "A").precision
    for key in ("general", "symmetric"):
        path = os.path.join(tmp_path, f"{key}.mtx")
        open(path, "wb").write(bytes(d[f"{key}_text"]))
        X = hip.read_matrix_market(hip.zero_matrix(prec, N, M), path)
        assert_ell_identical(X.get(), ell_of(d, f"{key}_read"), f"{case} {key}")


@pytest.mark.gpu
def test_read_errors(hip, tmp_path):
    """A row past M is the M overflow (the reference would overrun the
    row); a missing file is an argument error; the library stays usable."""
    d = CASES["mm_f64"]
    N = int(d["N"])
    path = os.path.join(tmp_path, "s.mtx")
    open(path, "wb").write(bytes(d["symmetric_text"]))
    with pytest.raises(hip.BmlError) as e:
        hip.read_matrix_market(hip.zero_matrix(2, N, 2), path)
    assert e.value.code == hip.ERR_M_OVERFLOW
    with pytest.raises(hip.BmlError) as e:
        hip.read_matrix_market(hip.zero_matrix(2, N, 12), os.path.join(tmp_path, "none.mtx"))
    assert e.value.code == hip.ERR_ARGUMENT
    X = hip.read_matrix_market(hip.zero_matrix(2, N, 12), path)
    assert_ell_identical(X.get(), ell_of(d, "symmetric_read"), "after errors")
