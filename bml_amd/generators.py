"""Seeded synthetic inputs for the thresholded-multiply and SP2 workloads.

The generator code is the specification of each input: numpy's PCG64
(``np.random.default_rng(seed)``), draws taken in the order written here.
Every matrix comes back as an :class:`~bml_amd.ell.Ell` (ELLPACK-R) or
:class:`~bml_amd.ell.Blk` (ELLBLOCK) with rows sorted by column.

Workloads
---------
* :func:`random_banded` -- symmetric random sparse X, N=4096, ~32 nnz/row at
  random columns within |i-j| <= 256, values U[-1,1] exp(-|i-j|/16),
  diagonal U[0,1]; one thresholded X^2 (1e-5) and one multiply.
* :func:`lattice_hamiltonian` -- 3D cubic lattice (periodic), one orbital per
  site, staggered onsite +-1 by the parity of x+y+z, nearest-neighbour
  hopping t plus optional U[-d,d] disorder per bond (32^3 for SP2).
* :func:`block_lattice_hamiltonian` -- 16x16x8 sites x 32 orbitals (ELLBLOCK,
  32x32 blocks): dense random symmetric onsite blocks U[-1,1] + (+-2) I by
  sublattice, 6 neighbour blocks 0.1 U[-1,1], H_JI = H_IJ^T.
* :func:`batch_sizes` / :func:`near_cubic` -- 512 system sizes log-uniform in
  [256, 8192] and their near-cubic lattice shapes.
* :func:`peierls_hamiltonian` -- complex Hermitian 3D lattice with hopping
  0.5 e^{i phi}, phi ~ U[0, 2pi) per bond, onsite +-1 (64^3).
"""
from __future__ import annotations

import math

import numpy as np

from .ell import Blk, Ell


def random_banded(N: int = 4096, M: int = 128, draws: int = 16, band: int = 256,
                  decay: float = 16.0, seed: int = 1, dtype=np.float64) -> Ell:
    """Symmetric random sparse matrix (~1 + 2*draws nnz per row).

    For each row i (in order), ``draws`` offsets d ~ U{1..band} name upper
    partners j = i + d (dropped when j >= N or already drawn for this row);
    each pair gets U[-1,1] * exp(-d/decay) and is mirrored.  Diagonal U[0,1].
    """
    rng = np.random.default_rng(seed)
    d = rng.integers(1, band + 1, size=(N, draws))
    u = rng.uniform(-1.0, 1.0, size=(N, draws))
    diag = rng.uniform(0.0, 1.0, size=N)
    i = np.repeat(np.arange(N), draws)
    j = i + d.ravel()
    v = u.ravel() * np.exp(-d.ravel() / decay)
    keep = j < N
    i, j, v = i[keep], j[keep], v[keep]
    # first occurrence of each (i, j) pair wins
    key = i.astype(np.int64) * N + j
    _, first = np.unique(key, return_index=True)
    first.sort()
    i, j, v = i[first], j[first], v[first]
    rows = np.concatenate([i, j, np.arange(N)])
    cols = np.concatenate([j, i, np.arange(N)])
    vals = np.concatenate([v, v, diag])
    return Ell.from_coo(N, M, rows, cols, vals.astype(dtype), dtype)


def near_cubic(n: int) -> tuple[int, int, int]:
    """Lattice shape for about n sites: L = round(n^(1/3)), Lx = Ly = L,
    Lz = max(1, round(n / L^2)); the system has Lx*Ly*Lz sites."""
    L = max(1, int(round(n ** (1.0 / 3.0))))
    Lz = max(1, int(round(n / (L * L))))
    return L, L, Lz


def _lattice_bonds(Lx: int, Ly: int, Lz: int):
    """Site index s = x + Lx*(y + Ly*z); +x, +y, +z bonds with periodic wrap
    (a direction of length 1 has no bonds, of length 2 one bond per pair)."""
    x, y, z = np.meshgrid(np.arange(Lx), np.arange(Ly), np.arange(Lz), indexing="ij")
    x, y, z = x.ravel(order="F"), y.ravel(order="F"), z.ravel(order="F")
    s = x + Lx * (y + Ly * z)
    order = np.argsort(s)
    x, y, z, s = x[order], y[order], z[order], s[order]
    src, dst = [], []
    for L, coord, stride in ((Lx, x, 1), (Ly, y, Lx), (Lz, z, Lx * Ly)):
        if L < 2:
            continue
        nxt = np.where(coord + 1 < L, s + stride, s - (L - 1) * stride)
        if L == 2:  # the wrap bond would repeat the forward bond
            m = coord == 0
            src.append(s[m])
            dst.append(nxt[m])
        else:
            src.append(s)
            dst.append(nxt)
    parity = (x + y + z) % 2
    return (np.concatenate(src) if src else np.zeros(0, np.int64),
            np.concatenate(dst) if dst else np.zeros(0, np.int64), parity)


def lattice_hamiltonian(Lx: int = 32, Ly: int = 32, Lz: int = 32, M: int = 256,
                        hopping: float = -0.5, disorder: float = 0.05,
                        onsite: float = 1.0, seed: int = 2,
                        dtype=np.float64) -> Ell:
    """Staggered-onsite tight-binding H on an Lx x Ly x Lz periodic lattice.

    Onsite +onsite on sites with x+y+z even, -onsite on odd.  One draw
    U[-disorder, disorder] per bond (+x bonds of all sites, then +y, then +z,
    in site order) is added to ``hopping``; H is symmetric.
    """
    rng = np.random.default_rng(seed)
    src, dst, parity = _lattice_bonds(Lx, Ly, Lz)
    N = Lx * Ly * Lz
    t = hopping + (rng.uniform(-disorder, disorder, size=len(src)) if disorder else 0.0)
    t = np.broadcast_to(np.asarray(t, np.float64), (len(src),))
    eps = np.where(parity == 0, onsite, -onsite).astype(np.float64)
    rows = np.concatenate([src, dst, np.arange(N)])
    cols = np.concatenate([dst, src, np.arange(N)])
    vals = np.concatenate([t, t, eps])
    return _coalesce(N, M, rows, cols, vals, dtype)


def peierls_hamiltonian(L: int = 64, M: int = 512, amplitude: float = 0.5,
                        onsite: float = 1.0, seed: int = 5,
                        dtype=np.complex128) -> Ell:
    """Complex Hermitian L^3 periodic lattice: H_ij = 0.5 e^{i phi_ij} on
    each bond (one phi ~ U[0, 2pi) per bond, in bond order), H_ji = conj,
    onsite +-1 by sublattice."""
    rng = np.random.default_rng(seed)
    src, dst, parity = _lattice_bonds(L, L, L)
    N = L ** 3
    phi = rng.uniform(0.0, 2.0 * np.pi, size=len(src))
    h = amplitude * np.exp(1j * phi)
    eps = np.where(parity == 0, onsite, -onsite).astype(np.complex128)
    rows = np.concatenate([src, dst, np.arange(N)])
    cols = np.concatenate([dst, src, np.arange(N)])
    vals = np.concatenate([h, np.conj(h), eps])
    return _coalesce(N, M, rows, cols, vals, dtype)


def _coalesce(N, M, rows, cols, vals, dtype) -> Ell:
    key = rows.astype(np.int64) * N + cols
    uk, inv = np.unique(key, return_inverse=True)
    acc = np.zeros(len(uk), np.result_type(vals.dtype, np.float64))
    np.add.at(acc, inv, vals)
    return Ell.from_coo(N, M, uk // N, uk % N, acc.astype(dtype), dtype)


def batch_sizes(count: int = 512, lo: int = 256, hi: int = 8192,
                seed: int = 4) -> np.ndarray:
    """System sizes log-uniform in [lo, hi] (before the near-cubic rounding)."""
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size=count)).round().astype(int)


def batch_hamiltonians(count: int = 512, M: int = 128, seed: int = 4,
                       dtype=np.float64) -> list[Ell]:
    """The batched-SP2 systems: near-cubic lattices of :func:`batch_sizes`,
    hopping -0.5, onsite +-1, no disorder."""
    out = []
    for n in batch_sizes(count, seed=seed):
        Lx, Ly, Lz = near_cubic(int(n))
        out.append(lattice_hamiltonian(Lx, Ly, Lz, M=M, hopping=-0.5, disorder=0.0,
                                       dtype=dtype))
    return out


def block_lattice_hamiltonian(Lx: int = 16, Ly: int = 16, Lz: int = 8, b: int = 32,
                              MB: int = 24, onsite_shift: float = 2.0,
                              coupling: float = 0.1, seed: int = 3,
                              dtype=np.float64) -> Blk:
    """ELLBLOCK tight-binding H: one b x b block per site.

    Onsite block of site s: A = U[-1,1]^(b x b) (drawn per site in site
    order), symmetrised (A + A^T)/2, plus (+shift) I on even-parity sites and
    (-shift) I on odd.  Each +x, +y, +z bond (bond order as
    :func:`lattice_hamiltonian`) gets C = coupling * U[-1,1]^(b x b) as block
    (I, J) and C^T as block (J, I).  Blocks of a block row sorted by column.
    """
    rng = np.random.default_rng(seed)
    src, dst, parity = _lattice_bonds(Lx, Ly, Lz)
    NB = Lx * Ly * Lz
    on = rng.uniform(-1.0, 1.0, size=(NB, b, b))
    on = 0.5 * (on + on.transpose(0, 2, 1))
    shift = np.where(parity == 0, onsite_shift, -onsite_shift)
    on = on + shift[:, None, None] * np.eye(b)[None]
    cpl = coupling * rng.uniform(-1.0, 1.0, size=(len(src), b, b))
    rows = np.concatenate([np.arange(NB), src, dst])
    cols = np.concatenate([np.arange(NB), dst, src])
    blocks = np.concatenate([on, cpl, cpl.transpose(0, 2, 1)])
    out = Blk.zeros(np.full(NB, b, np.int32), MB, dtype)
    order = np.lexsort((cols, rows))
    rows, cols, blocks = rows[order], cols[order], blocks[order]
    counts = np.bincount(rows, minlength=NB)
    if counts.max() > MB:
        raise ValueError("block row does not fit MB")
    start = np.concatenate([[0], np.cumsum(counts)[:-1]])
    slot = np.arange(len(rows)) - start[rows]
    out.value[rows, slot] = blocks.astype(dtype)
    out.indexb[rows, slot] = cols
    out.nnzb[:] = counts
    return out


def random_ell(N: int, M: int, fill: int, seed: int, dtype=np.float64,
               diag: bool = True, scale: float = 1.0) -> Ell:
    """Unstructured random ELLPACK matrix for tests: each row gets up to
    ``fill`` distinct random columns (plus the diagonal when ``diag``),
    values U[-1,1]*scale (complex: independent real and imaginary parts)."""
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for i in range(N):
        k = int(rng.integers(0, fill + 1))
        c = rng.choice(N, size=min(k, N), replace=False)
        if diag and i not in c:
            c = np.append(c, i)
        rows.append(np.full(len(c), i))
        cols.append(c)
    rows = np.concatenate(rows) if rows else np.zeros(0, int)
    cols = np.concatenate(cols) if cols else np.zeros(0, int)
    vals = rng.uniform(-1.0, 1.0, size=len(rows)) * scale
    if np.issubdtype(np.dtype(dtype), np.complexfloating):
        vals = vals + 1j * rng.uniform(-1.0, 1.0, size=len(rows)) * scale
    # scramble the in-row order so first-touch order is exercised
    out = Ell.from_coo(N, M, rows, cols, vals.astype(dtype), dtype)
    for i in range(N):
        n = int(out.nnz[i])
        p = rng.permutation(n)
        out.index[i, :n] = out.index[i, :n][p]
        out.value[i, :n] = out.value[i, :n][p]
    return out


def random_blk(bsize, MB: int, fill: int, seed: int, dtype=np.float64,
               bmax: int | None = None) -> Blk:
    """Random ELLBLOCK test matrix: diagonal block plus up to ``fill`` random
    block columns per block row, elements U[-1,1]."""
    rng = np.random.default_rng(seed)
    bsize = np.asarray(bsize, np.int32)
    out = Blk.zeros(bsize, MB, dtype, bmax)
    NB = len(bsize)
    for ib in range(NB):
        k = int(rng.integers(0, fill + 1))
        cols = list(rng.choice(NB, size=min(k, NB), replace=False))
        if ib not in cols:
            cols.append(ib)
        cols = cols[:MB]
        for jp, jb in enumerate(cols):
            bi, bj = int(bsize[ib]), int(bsize[jb])
            v = rng.uniform(-1.0, 1.0, size=(bi, bj))
            if np.issubdtype(np.dtype(dtype), np.complexfloating):
                v = v + 1j * rng.uniform(-1.0, 1.0, size=(bi, bj))
            out.value[ib, jp, :bi, :bj] = v
            out.indexb[ib, jp] = jb
        out.nnzb[ib] = len(cols)
    return out
