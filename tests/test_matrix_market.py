"""Matrix Market files (bml_write_bml_matrix / bml_read_bml_matrix for
ellpack, bml_utilities.h:38-45, ellpack/bml_utilities_ellpack_typed.c:22-190).

The fixtures in tests/golden/mm.npz hold the reference's own output: the text
it writes for a random matrix, what it reads back from that text, and (for
ellpack) what it reads from a "symmetric" file.  The ellblock "symmetric"
case is not taken from the reference, whose symmetric branch is broken
(ellblock/bml_utilities_ellblock_typed.c:131-160); it is checked against the
dense mirror.  The CPU tests pin the fixture's format
against the C formatting rules it follows; the GPU tests compare the
library's writer byte for byte and its reader bit for bit."""
import os

import numpy as np
import pytest

from helpers import assert_blk_identical, assert_ell_identical, blk_of, ell_of, load_cases

ALL = load_cases("mm.npz")
CASES = {k: v for k, v in ALL.items() if k.startswith("mm_")}
EB_CASES = {k: v for k, v in ALL.items() if k.startswith("ebmm_")}


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


def _offsets(bsize):
    return np.concatenate([[0], np.cumsum(bsize)]).astype(int)


@pytest.mark.parametrize("case", sorted(EB_CASES))
def test_ellblock_reference_text_format(case):
    """Every element of every stored block (block rows in order, blocks in
    storage order, row-major within a block, zeros included), 1-based global
    indices, each value as C's %20.15e."""
    d = EB_CASES[case]
    A = blk_of(d, "A")
    off = _offsets(A.bsize)
    lines = text(d, "general_text").splitlines()
    kind = "complex" if np.iscomplexobj(A.value) else "real"
    assert lines[0] == f"%%%MatrixMarket matrix coordinate {kind} general"
    want = []
    for ib in range(A.NB):
        for jp in range(int(A.nnzb[ib])):
            jb = int(A.indexb[ib, jp])
            for ii in range(int(A.bsize[ib])):
                for jj in range(int(A.bsize[jb])):
                    v = A.value[ib, jp, ii, jj]
                    vals = [v.real, v.imag] if kind == "complex" else [v]
                    want.append(f"{off[ib] + 1 + ii} {off[jb] + 1 + jj} " +
                                " ".join("%20.15e" % float(x) for x in vals))
    assert lines[1] == f"{int(off[-1])} {int(off[-1])} {len(want)}"
    assert lines[2:] == want


@pytest.mark.parametrize("case", sorted(EB_CASES))
def test_ellblock_reference_read_is_the_parsed_text(case):
    """The reference's read: each entry goes to its block, appended (zeroed)
    to its block row the first time one of its elements appears."""
    d = EB_CASES[case]
    A = blk_of(d, "A")
    R = blk_of(d, "general_read")
    off = _offsets(A.bsize)
    blk = np.searchsorted(off, np.arange(off[-1]), side="right") - 1
    dt = A.value.dtype
    order = [[] for _ in range(A.NB)]
    vals = {}
    for ln in text(d, "general_text").splitlines()[2:]:
        t = ln.split()
        r, c = int(t[0]) - 1, int(t[1]) - 1
        v = complex(float(t[2]), float(t[3])) if len(t) == 4 else float(t[2])
        ib, jb = blk[r], blk[c]
        if jb not in order[ib]:
            order[ib].append(jb)
        vals[(ib, jb, r - off[ib], c - off[jb])] = np.array(v).astype(dt)
    for ib in range(A.NB):
        assert R.nnzb[ib] == len(order[ib]) and R.indexb[ib, :len(order[ib])].tolist() == order[ib]
        for jp, jb in enumerate(order[ib]):
            for ii in range(int(A.bsize[ib])):
                for jj in range(int(A.bsize[jb])):
                    assert R.value[ib, jp, ii, jj].tobytes() == vals[(ib, jb, ii, jj)].tobytes()


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(EB_CASES))
def test_ellblock_write_and_read_match_reference(hip, tmp_path, case):
    d = EB_CASES[case]
    A = blk_of(d, "A")
    path = os.path.join(tmp_path, "a.mtx")
    hip.write_matrix_market(hip.from_blk(A), path)
    assert open(path, "rb").read() == bytes(d["general_text"])
    X = hip.read_matrix_market(hip.block_matrix(A.precision, A.bsize, A.MB), path)
    assert_blk_identical(X.get(A.bmax), blk_of(d, "general_read"), case)


@pytest.mark.gpu
def test_ellblock_read_symmetric_file(hip, tmp_path):
    """A "symmetric" file: each off-diagonal element also lands at its
    mirror (the dense matrix is the file's lower triangle mirrored)."""
    d = EB_CASES["ebmm_f64"]
    A = blk_of(d, "A")
    off = _offsets(A.bsize)
    N = int(off[-1])
    rng = np.random.default_rng(5)
    D = np.zeros((N, N))
    lines = []
    for i in range(N):
        for j in sorted(rng.choice(i + 1, size=min(i + 1, 2), replace=False)):
            v = rng.uniform(-1, 1)
            D[i, j] = D[j, i] = v
            lines.append(f"{i + 1} {j + 1} {v:.17e}")
    path = os.path.join(tmp_path, "s.mtx")
    open(path, "w").write(f"%%MatrixMarket matrix coordinate real symmetric\n{N} {N} {len(lines)}\n"
                          + "\n".join(lines) + "\n")
    X = hip.read_matrix_market(hip.block_matrix(2, A.bsize, A.NB), path)
    np.testing.assert_array_equal(hip.eb_export_to_dense(X), D)
