"""BSR binary cache files (the layout of csrc/bsp_io.cpp), readable without a GPU.

Layout (little-endian): magic b"BSPBSR1\\0"; int32 n_rows, n_cols, bs, 0; int64 nnzb;
int32 rowptr[n_rows + 1]; int32 colidx[nnzb]; float64 values[nnzb, bs, bs].
"""
from __future__ import annotations

import numpy as np

MAGIC = b"BSPBSR1\0"


def save(path, n_rows, n_cols, bs, rowptr, colidx, vals):
    rp = np.ascontiguousarray(rowptr, "<i4")
    ci = np.ascontiguousarray(colidx, "<i4")
    v = np.ascontiguousarray(vals, "<f8")
    if rp.shape != (n_rows + 1,) or v.size != ci.size * bs * bs or rp[-1] != ci.size:
        raise ValueError("inconsistent BSR arrays")
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(np.array([n_rows, n_cols, bs, 0], "<i4").tobytes())
        f.write(np.array([ci.size], "<i8").tobytes())
        f.write(rp.tobytes())
        f.write(ci.tobytes())
        f.write(v.tobytes())


def load(path, mmap=True):
    """(n_rows, n_cols, bs, rowptr, colidx, vals); the arrays are memory-mapped by default."""
    with open(path, "rb") as f:
        if f.read(8) != MAGIC:
            raise ValueError(f"{path} is not a BSR cache file")
        n, m, bs, _ = np.frombuffer(f.read(16), "<i4")
        nnzb = int(np.frombuffer(f.read(8), "<i8")[0])
    off = 32
    mode = "r" if mmap else None
    if mmap:
        rp = np.memmap(path, "<i4", mode, off, (int(n) + 1,))
        ci = np.memmap(path, "<i4", mode, off + 4 * (int(n) + 1), (nnzb,))
        v = np.memmap(path, "<f8", mode, off + 4 * (int(n) + 1 + nnzb), (nnzb, int(bs), int(bs)))
    else:
        raw = open(path, "rb").read()
        rp = np.frombuffer(raw, "<i4", int(n) + 1, off)
        ci = np.frombuffer(raw, "<i4", nnzb, off + 4 * (int(n) + 1))
        v = np.frombuffer(raw, "<f8", nnzb * int(bs) ** 2, off + 4 * (int(n) + 1 + nnzb)).reshape(nnzb, bs, bs)
    if rp[0] != 0 or rp[-1] != nnzb:
        raise ValueError(f"inconsistent row pointers in {path}")
    return int(n), int(m), int(bs), rp, ci, v
