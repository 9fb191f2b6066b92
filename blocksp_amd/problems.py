"""Synthetic block-sparse problems for the five benchmark workloads.

Every generator returns a :class:`Problem` in BSR form with each block row's
columns in ascending order (south, west, self, east, north [, back, front]),
which is also the reference's insertion order when a user assembles the
stencil neighbour by neighbour (blocksp_containers.hpp:147-164).

All randomness comes from ``numpy.random.Generator(PCG64(seed))`` drawn in a
fixed order, so the CPU oracle and the GPU path see identical matrices.

Workloads (sizes as quoted by the benchmark definitions):

* ``stencil2d_b2``   64x64 5-point block stencil, B=2: diagonal 4I + 0.1 sym N(0,1),
  off-diagonal -I + 0.05 N(0,1), seed 0, RHS U[-1,1].
* ``dg2d_veclap``    512x512 quads, symmetric-interior-penalty (SWIP weights) Q1 DG
  Laplacian, 4 dofs/element x D=2 components -> bs=8, coefficient lognormal(0,1) per
  element (seed 1), RHS N(0,1).
* ``dg3d_lap``       64^3 hexes, the same Q1 SWIP operator in 3D, D=1 -> bs=8.
* ``dg3d_grf``       48^3 hexes with a lognormal Gaussian random field coefficient of
  correlation length 4 cells (seed 3).
* ``batch_blocklap`` 512 SPD 5-point block Laplacians, B=4.
"""
from __future__ import annotations

import dataclasses
import itertools

import numpy as np

__all__ = [
    "Problem",
    "stencil_pattern",
    "stencil2d_b2",
    "dg_laplacian",
    "dg2d_veclap",
    "dg3d_lap",
    "dg3d_grf",
    "batch_blocklap",
    "to_dense",
]


@dataclasses.dataclass
class Problem:
    n: int                 # block rows (= block columns)
    bs: int                # block size B*D
    B: int
    D: int
    rowptr: np.ndarray     # int32 [n+1]
    colidx: np.ndarray     # int32 [nnzb]
    vals: np.ndarray       # float64 [nnzb, bs, bs] row-major blocks
    b: np.ndarray          # float64 [n*bs] right-hand side (solArray flat layout)
    name: str = ""
    meta: dict = dataclasses.field(default_factory=dict)

    @property
    def nnzb(self) -> int:
        return int(self.colidx.shape[0])

    def bsmv(self, x: np.ndarray) -> np.ndarray:
        """Dense-free numpy product y = A x (an independent check, not the reference order)."""
        xb = x.reshape(self.n, self.bs)
        rows = np.repeat(np.arange(self.n), np.diff(self.rowptr))
        contrib = np.einsum("kij,kj->ki", self.vals, xb[self.colidx])
        y = np.zeros((self.n, self.bs))
        np.add.at(y, rows, contrib)
        return y.reshape(-1)


def stencil_pattern(dims):
    """Ascending-column BSR pattern of the 2D 5-point / 3D 7-point grid stencil.

    Returns (rowptr, colidx, slot) where slot[k] names the neighbour of block k:
    0 = self, 1+2*axis = minus side, 2+2*axis = plus side; axis 0 is the fastest
    index.  Non-periodic boundaries.
    """
    dims = tuple(int(d) for d in dims)
    nd = len(dims)
    n = int(np.prod(dims))
    idx = np.arange(n)
    coords = np.unravel_index(idx, dims[::-1])[::-1]   # coords[axis], axis 0 fastest
    strides = [int(np.prod(dims[:a])) for a in range(nd)]
    # candidate neighbours in ascending column order: slowest axis minus first
    cand = []
    for a in reversed(range(nd)):
        cand.append((1 + 2 * a, -strides[a], a, -1))
    cand.append((0, 0, None, 0))
    for a in range(nd):
        cand.append((2 + 2 * a, strides[a], a, +1))
    present = []
    for slot, off, a, side in cand:
        if a is None:
            ok = np.ones(n, bool)
        elif side < 0:
            ok = coords[a] > 0
        else:
            ok = coords[a] < dims[a] - 1
        present.append(ok)
    present = np.stack(present, axis=1)                      # [n, ncand]
    counts = present.sum(1)
    rowptr = np.zeros(n + 1, np.int64)
    np.cumsum(counts, out=rowptr[1:])
    offs = np.array([c[1] for c in cand])
    slots = np.array([c[0] for c in cand])
    cols = idx[:, None] + offs[None, :]
    colidx = cols[present].astype(np.int32)
    slot = np.broadcast_to(slots, present.shape)[present].astype(np.int8)
    return rowptr.astype(np.int32), colidx, slot


def stencil2d_b2(nx: int = 64, ny: int = 64, seed: int = 0) -> Problem:
    """64x64 5-point block stencil with B=2 (the CPU-runnable SOR-pGMRES workload)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    rowptr, colidx, slot = stencil_pattern((nx, ny))
    n = nx * ny
    bs = 2
    nnzb = colidx.shape[0]
    S = rng.standard_normal((n, bs, bs))
    S = np.triu(S) + np.transpose(np.triu(S, 1), (0, 2, 1))
    diag = 4.0 * np.eye(bs)[None] + 0.1 * S
    off = -np.eye(bs)[None] + 0.05 * rng.standard_normal((nnzb, bs, bs))
    vals = np.where((slot == 0)[:, None, None], 0.0, off)
    vals[slot == 0] = diag
    b = rng.uniform(-1.0, 1.0, n * bs)
    return Problem(n, bs, bs, 1, rowptr, colidx, np.ascontiguousarray(vals), b,
                   name=f"stencil2d_b2_{nx}x{ny}", meta={"seed": seed, "dims": (nx, ny)})


# ---------------------------------------------------------------------------
# Symmetric interior penalty (SWIP) Q1 discontinuous Galerkin Laplacian
# ---------------------------------------------------------------------------

_M1 = np.array([[1.0 / 3.0, 1.0 / 6.0], [1.0 / 6.0, 1.0 / 3.0]])     # int L_j L_k
_K1 = np.array([[1.0, -1.0], [-1.0, 1.0]])                           # int L_j' L_k'
_S = np.array([-1.0, 1.0])                                           # L_i' (constant)


def _q1_reference(d: int, sigma: float):
    """Reference-element matrices of the Q1 SWIP form in d dimensions.

    Nodes a = (i_0, ..., i_{d-1}) with axis 0 fastest.  All terms are
    h-independent after scaling on a uniform cube mesh with d=2; for d=3 the
    volume and face terms scale with h, which is a common factor dropped here.
    Returns (K_vol, C) where C[(axis, side)] = (self_self, cross) matrices:
    the face on the `side` (0 minus, 1 plus) of the element along `axis`.
    """
    nodes = list(itertools.product((0, 1), repeat=d))
    nodes = [tuple(reversed(t)) for t in nodes]          # axis 0 fastest
    nn = len(nodes)
    K = np.zeros((nn, nn))
    for a, na in enumerate(nodes):
        for b_, nb in enumerate(nodes):
            tot = 0.0
            for ax in range(d):
                prod = _K1[na[ax], nb[ax]]
                for o in range(d):
                    if o != ax:
                        prod *= _M1[na[o], nb[o]]
                tot += prod
            K[a, b_] = tot
    C = {}
    for ax in range(d):
        for side in (0, 1):
            # minus element = this element; its face sits at xi_ax = side.
            # outward normal sign: +1 for side 1, -1 for side 0.
            nsign = 1.0 if side == 1 else -1.0
            own_face = side
            nbr_face = 1 - side

            def tangential(na, nb):
                p = 1.0
                for o in range(d):
                    if o != ax:
                        p *= _M1[na[o], nb[o]]
                return p

            ss = np.zeros((nn, nn))
            cross = np.zeros((nn, nn))
            bnd = np.zeros((nn, nn))
            for a, na in enumerate(nodes):
                tr_a = 1.0 if na[ax] == own_face else 0.0
                dn_a = nsign * _S[na[ax]]
                for b_, nb in enumerate(nodes):
                    t = tangential(na, nb)
                    tr_b = 1.0 if nb[ax] == own_face else 0.0
                    dn_b = nsign * _S[nb[ax]]
                    # self-self: [u] = u_self on this side (sign +), {k du/dn}_w = k_h/2 du_self/dn
                    ss[a, b_] = (-0.5 * dn_b * tr_a - 0.5 * dn_a * tr_b + sigma * tr_a * tr_b) * t
                    # boundary (Nitsche, weight k_e): full normal derivative, no 1/2
                    bnd[a, b_] = (-dn_b * tr_a - dn_a * tr_b + sigma * tr_a * tr_b) * t
                    # cross: test on self (sign +), trial on neighbour (sign -)
                    trn_b = 1.0 if nb[ax] == nbr_face else 0.0
                    dnn_b = nsign * _S[nb[ax]]          # neighbour's derivative along the same normal
                    cross[a, b_] = (-0.5 * dnn_b * tr_a + 0.5 * dn_a * trn_b - sigma * tr_a * trn_b) * t
            C[(ax, side)] = (ss, cross, bnd)
    return K, C


def dg_laplacian(dims, kappa: np.ndarray, D: int = 1, sigma: float = 6.0):
    """Assemble the SWIP Q1 DG operator on a uniform grid of `dims` elements.

    kappa: per-element coefficient (len = prod(dims)), axis 0 fastest.
    Returns (rowptr, colidx, vals) with bs = 2^d * D and blocks kron(I_D, K).
    """
    d = len(dims)
    K, C = _q1_reference(d, sigma)
    nn = K.shape[0]
    rowptr, colidx, slot = stencil_pattern(dims)
    n = int(np.prod(dims))
    kappa = np.asarray(kappa, np.float64).reshape(n)
    rows = np.repeat(np.arange(n), np.diff(rowptr))
    cols = colidx.astype(np.int64)
    blocks = np.zeros((colidx.shape[0], nn, nn))
    # diagonal: volume + self parts of all faces (interior or boundary)
    diag_mask = slot == 0
    ke = kappa[rows[diag_mask]]
    blocks[diag_mask] = ke[:, None, None] * K[None]
    idx = np.arange(n)
    coords = np.unravel_index(idx, tuple(dims)[::-1])[::-1]
    strides = [int(np.prod(dims[:a])) for a in range(d)]
    diag_pos = np.flatnonzero(diag_mask)            # diag_pos[e] = block index of (e,e)
    for ax in range(d):
        for side in (0, 1):
            ss, cross, bnd = C[(ax, side)]
            if side == 1:
                has = coords[ax] < dims[ax] - 1
                nbr = idx + strides[ax]
            else:
                has = coords[ax] > 0
                nbr = idx - strides[ax]
            e_int = idx[has]
            kh = 2.0 * kappa[e_int] * kappa[nbr[has]] / (kappa[e_int] + kappa[nbr[has]])
            blocks[diag_pos[e_int]] += kh[:, None, None] * ss[None]
            e_bnd = idx[~has]
            blocks[diag_pos[e_bnd]] += kappa[e_bnd][:, None, None] * bnd[None]
            # off-diagonal block (e, nbr): slot for this side/axis
            sl = 2 + 2 * ax if side == 1 else 1 + 2 * ax
            m = slot == sl
            er = rows[m]
            en = cols[m]
            khm = 2.0 * kappa[er] * kappa[en] / (kappa[er] + kappa[en])
            blocks[m] = khm[:, None, None] * cross[None]
    if D > 1:
        big = np.zeros((blocks.shape[0], nn * D, nn * D))
        for dd in range(D):
            big[:, dd * nn:(dd + 1) * nn, dd * nn:(dd + 1) * nn] = blocks
        blocks = big
    return rowptr, colidx, np.ascontiguousarray(blocks)


def dg2d_veclap(nx: int = 512, ny: int = 512, seed: int = 1, D: int = 2) -> Problem:
    """2D LDG-style vector Laplacian: 512x512 Q1 SWIP DG x D=2 -> BlockSp<double,8>."""
    rng = np.random.Generator(np.random.PCG64(seed))
    kappa = np.exp(rng.standard_normal(nx * ny))
    rowptr, colidx, vals = dg_laplacian((nx, ny), kappa, D=D)
    n = nx * ny
    bs = vals.shape[1]
    b = rng.standard_normal(n * bs)
    return Problem(n, bs, 4, D, rowptr, colidx, vals, b, name=f"dg2d_veclap_{nx}x{ny}_D{D}",
                   meta={"seed": seed, "dims": (nx, ny)})


def dg3d_lap(nx: int = 64, ny: int = 64, nz: int = 64, seed: int = 4) -> Problem:
    """3D Q1 SWIP DG Laplacian on nx*ny*nz hexes, D=1 -> bs=8, coefficient lognormal(0,1)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = nx * ny * nz
    kappa = np.exp(rng.standard_normal(n))
    rowptr, colidx, vals = dg_laplacian((nx, ny, nz), kappa, D=1)
    b = rng.standard_normal(n * 8)
    return Problem(n, 8, 8, 1, rowptr, colidx, vals, b, name=f"dg3d_lap_{nx}x{ny}x{nz}",
                   meta={"seed": seed, "dims": (nx, ny, nz)})


def _gaussian_field(dims, corr_len: float, rng) -> np.ndarray:
    """Stationary Gaussian field with Gaussian covariance exp(-r^2 / (2 l^2)), unit variance,
    by spectral filtering of white noise (periodic embedding)."""
    white = rng.standard_normal(tuple(dims[::-1]))
    freqs = np.meshgrid(*[np.fft.fftfreq(s) for s in dims[::-1]], indexing="ij")
    k2 = sum(f * f for f in freqs)
    filt = np.exp(-((2 * np.pi) ** 2) * k2 * corr_len ** 2 / 4.0)
    g = np.real(np.fft.ifftn(np.fft.fftn(white) * filt))
    g = (g - g.mean()) / g.std()
    return g.reshape(-1)


def dg3d_grf(nx: int = 48, ny: int = 48, nz: int = 48, seed: int = 3, corr_len: float = 4.0) -> Problem:
    """3D random-coefficient diffusion: lognormal Gaussian random field (correlation 4 cells)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = nx * ny * nz
    kappa = np.exp(_gaussian_field((nx, ny, nz), corr_len, rng))
    rowptr, colidx, vals = dg_laplacian((nx, ny, nz), kappa, D=1)
    b = rng.standard_normal(n * 8)
    return Problem(n, 8, 8, 1, rowptr, colidx, vals, b, name=f"dg3d_grf_{nx}x{ny}x{nz}",
                   meta={"seed": seed, "dims": (nx, ny, nz), "corr_len": corr_len})


def batch_blocklap(nsys: int = 512, lo: float = 1e3, hi: float = 5e4, seed: int = 2, B: int = 4,
                   nrhs: int = 4):
    """512 independent SPD block Laplacians (5-point stencil, B=4).

    System s: grid nx = floor(sqrt(n_s)), ny = floor(n_s / nx) with n_s log-uniform in
    [lo, hi]; A_s = c_s (L_5pt (x) I_B + 0.25 I (x) R_s R_s^T / B) with c_s lognormal(0, 0.5)
    and R_s ~ N(0,1) (B x B), which is SPD.  Returns a list of Problems whose `b`
    holds nrhs right-hand sides as an [nrhs, n*bs] array in meta["rhs"].
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), nsys))
    coefs = np.exp(0.5 * rng.standard_normal(nsys))
    out = []
    for s in range(nsys):
        ns = int(sizes[s])
        nx = int(np.floor(np.sqrt(ns)))
        ny = ns // nx
        R = rng.standard_normal((B, B))
        K = 0.25 * (R @ R.T) / B
        rowptr, colidx, slot = stencil_pattern((nx, ny))
        n = nx * ny
        vals = np.zeros((colidx.shape[0], B, B))
        vals[slot == 0] = coefs[s] * (4.0 * np.eye(B) + K)
        vals[slot != 0] = -coefs[s] * np.eye(B)
        rhs = rng.standard_normal((nrhs, n * B))
        p = Problem(n, B, B, 1, rowptr, colidx, vals, rhs[0].copy(), name=f"blocklap_{s}",
                    meta={"dims": (nx, ny), "coef": float(coefs[s]), "rhs": rhs})
        out.append(p)
    return out


def parity_colours(dims) -> np.ndarray:
    """2^d-colouring of a structured grid by coordinate parity (axis 0 fastest):
    colour = sum_a (i_a mod 2) 2^a.  Neighbours along an axis differ in exactly one
    parity bit, so no two cells of a colour are coupled by the 2D 5-point / 3D
    7-point block stencil, and an ILU0 eliminated colour by colour has triangular
    sweeps of d+1 levels."""
    dims = tuple(int(d) for d in dims)
    n = int(np.prod(dims))
    coords = np.unravel_index(np.arange(n), dims[::-1])[::-1]
    col = np.zeros(n, np.int32)
    for a, c in enumerate(coords):
        col += ((c % 2) << a).astype(np.int32)
    return col


def to_dense(p: Problem) -> np.ndarray:
    """Dense n*bs x n*bs matrix of a (small) problem."""
    bs = p.bs
    A = np.zeros((p.n * bs, p.n * bs))
    for i in range(p.n):
        for k in range(p.rowptr[i], p.rowptr[i + 1]):
            j = p.colidx[k]
            A[i * bs:(i + 1) * bs, j * bs:(j + 1) * bs] += p.vals[k]
    return A
