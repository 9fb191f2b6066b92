"""Matrix-Market assembly with BlockSp's insertion semantics.

``entries_to_bsr`` reproduces mtxMatrix_to_BlockSp (blocksp_tests.hpp:595-649):
every entry (i, j, v) goes through setFromGlobal (blocksp_utility.hpp:614-626),
which appends block column j//bs to block row i//bs on first touch
(blocksp_containers.hpp:147-164) and then assigns element (i%bs, j%bs); for a
symmetric file the mirrored entry (j, i) is written right after (i, j) when
i != j.  The resulting per-row column order (first-touch order) is the block
structure the solvers see, so it is reproduced exactly.

``vector_to_solarray`` reads a vector in the solArray flat layout
(ie, dim, ib) -> ie*B*D + dim*B + ib.  The reference reader computes the
in-block index as i % B instead of i % (B*D) (blocksp_tests.hpp:679), which
misplaces every value when D > 1; this reader uses the documented layout.
"""
from __future__ import annotations

import numpy as np

__all__ = ["entries_to_bsr", "read_mtx_entries", "read_mtx_vector", "vector_to_solarray"]


def _goto_line(f, start_line):
    """Skip start_line - 1 lines, as GotoLine does (blocksp_tests.hpp:582-590): the size line
    is line start_line (1-based; 0 and 1 both mean the first line)."""
    for _ in range(max(int(start_line) - 1, 0)):
        f.readline()


def read_mtx_entries(path, start_line=0):
    """Parse a BlockSp-style .mtx matrix (blocksp_tests.hpp:595-616): "M N nnz" on line
    start_line, then nnz lines "i j value", 1-based.  Returns M, N and 0-based i, j, values
    in file order."""
    with open(path) as f:
        _goto_line(f, start_line)
        head = f.readline().split()
        if len(head) < 3:
            raise ValueError(f"{path}: no 'M N nnz' line at line {max(int(start_line), 1)}")
        M, N, nnz = (int(t) for t in head[:3])
        data = np.loadtxt(f, dtype=np.float64, max_rows=nnz, ndmin=2) if nnz else np.zeros((0, 3))
    if data.shape[0] != nnz:
        raise ValueError(f"{path}: {data.shape[0]} entries, the size line says {nnz}")
    return M, N, data[:, 0].astype(np.int64) - 1, data[:, 1].astype(np.int64) - 1, data[:, 2].copy()


def read_mtx_vector(path, start_line=0):
    """Parse a BlockSp-style .mtx vector (blocksp_tests.hpp:652-675): "M 1" on line
    start_line, then M values."""
    with open(path) as f:
        _goto_line(f, start_line)
        head = f.readline().split()
        if not head:
            raise ValueError(f"{path}: no 'M N' line at line {max(int(start_line), 1)}")
        M = int(head[0])
        v = np.loadtxt(f, dtype=np.float64, max_rows=M, ndmin=1) if M else np.zeros(0)
    if v.shape[0] != M:
        raise ValueError(f"{path}: {v.shape[0]} values, the size line says {M}")
    return v


def entries_to_bsr(M, N, i, j, v, bs, symmetric=False):
    """Build (rowptr, colidx, vals) with the reference's block insertion order."""
    if M % bs or N % bs:
        raise ValueError(f"block size {bs} must divide the matrix size {M}x{N}")
    i = np.asarray(i, np.int64)
    j = np.asarray(j, np.int64)
    v = np.asarray(v, np.float64)
    if symmetric:
        off = i != j
        # interleave (i,j) then (j,i) per entry, in file order
        order = np.arange(i.shape[0])
        ti = np.concatenate([i, j[off]])
        tj = np.concatenate([j, i[off]])
        tv = np.concatenate([v, v[off]])
        seq = np.concatenate([2 * order, 2 * order[off] + 1])
        perm = np.argsort(seq, kind="stable")
        ti, tj, tv = ti[perm], tj[perm], tv[perm]
    else:
        ti, tj, tv = i, j, v
    nbr, nbc = M // bs, N // bs
    bi = ti // bs
    bj = tj // bs
    key = bi * nbc + bj
    uniq, first = np.unique(key, return_index=True)
    brow = uniq // nbc
    # blocks ordered by (row, first touch)
    order = np.lexsort((first, brow))
    uniq = uniq[order]
    brow = brow[order]
    counts = np.bincount(brow, minlength=nbr)
    rowptr = np.zeros(nbr + 1, np.int64)
    np.cumsum(counts, out=rowptr[1:])
    colidx = (uniq % nbc).astype(np.int32)
    # block id of every touch
    pos = np.empty(uniq.shape[0], np.int64)
    pos[:] = np.arange(uniq.shape[0])
    sorter = np.argsort(uniq)
    bid = pos[sorter][np.searchsorted(uniq, key, sorter=sorter)]
    vals = np.zeros((uniq.shape[0], bs, bs))
    # last write wins (setFromGlobal assigns)
    flat = bid * bs * bs + (ti % bs) * bs + (tj % bs)
    rev_u, rev_first = np.unique(flat[::-1], return_index=True)
    last = tv.shape[0] - 1 - rev_first
    vals.reshape(-1)[rev_u] = tv[last]
    return rowptr.astype(np.int32), colidx, vals


def vector_to_solarray(v, B, D=1):
    """Place a global vector into the solArray flat layout (identity for this layout)."""
    v = np.asarray(v, np.float64)
    if v.shape[0] % (B * D):
        raise ValueError("B*D must divide the vector length")
    return v.copy()
