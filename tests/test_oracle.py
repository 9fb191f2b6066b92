"""Pins the CPU oracle (oracle/blocksp_oracle.cpp) before it is trusted as the
parity checker: closed forms, bitwise order checks, numpy brute force on tiny
inputs, and the reference's own fixture with its known answer."""
import numpy as np
import pytest

import pyoracle as po
from blocksp_amd import problems as pr
from blocksp_amd import mtx


def rand_bsr(n, bs, seed, maxblk=4, square=True):
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n):
        k = int(rng.integers(0, maxblk + 1))
        cols = list(rng.choice(n, size=min(k, n), replace=False))
        if square and i not in cols:
            cols.append(i)
        rng.shuffle(cols)
        rows.append(cols)
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.array([c for r in rows for c in r], np.int32)
    v = rng.standard_normal((ci.shape[0], bs, bs))
    return rp, ci, v


def dense_of(n, bs, rp, ci, v):
    A = np.zeros((n * bs, n * bs))
    for i in range(n):
        for p in range(rp[i], rp[i + 1]):
            A[i * bs:(i + 1) * bs, ci[p] * bs:(ci[p] + 1) * bs] += v[p]
    return A


# ---------------------------------------------------------------- dense blocks

@pytest.mark.parametrize("bs", [1, 2, 3, 4, 5, 8])
def test_lu_solve_and_rsolve(bs):
    rng = np.random.default_rng(bs)
    A = rng.standard_normal((bs, bs)) + bs * np.eye(bs)
    lu, piv = po.lu_factor(A)
    x = rng.standard_normal(bs)
    assert np.allclose(po.lu_solve(lu, piv, x), np.linalg.solve(A, x), rtol=1e-13, atol=1e-13)
    X = rng.standard_normal((bs, bs))
    assert np.allclose(po.lu_rsolve(lu, piv, X), X @ np.linalg.inv(A), rtol=1e-12, atol=1e-12)


def test_lu_partial_pivoting_picks_first_max():
    A = np.array([[1.0, 2, 3, 4], [-4, 1, 0, 2], [4, 0, 1, 1], [2, 2, 2, 9]])
    lu, piv = po.lu_factor(A)
    assert piv[0] == 1    # |-4| and |4| tie: the first maximum (row 1) is the pivot
    P = np.eye(4)
    for j, p in enumerate(piv):
        P[[j, p]] = P[[p, j]]
    L = np.tril(lu, -1) + np.eye(4)
    U = np.triu(lu)
    assert np.allclose(P @ A, L @ U, atol=1e-14)


def test_small_solve_2x2_bitwise():
    """bs=2 uses the explicit formula of dense_solvers.hpp:1036-1041 (no pivoting)."""
    A = np.array([[0.3, -1.7], [2.9, 0.11]])
    b = np.array([0.7, -0.2])
    lu, piv = po.lu_factor(A)
    det = A[0, 0] * A[1, 1] - A[0, 1] * A[1, 0]
    y0 = (A[1, 1] * b[0] - A[0, 1] * b[1]) / det
    y1 = (-A[1, 0] * b[0] + A[0, 0] * b[1]) / det
    out = po.lu_solve(lu, piv, b)
    assert out[0] == y0 and out[1] == y1


def test_small_inverse_3x3():
    rng = np.random.default_rng(1)
    A = rng.standard_normal((3, 3)) + 3 * np.eye(3)
    lu, piv = po.lu_factor(A)
    x = rng.standard_normal(3)
    assert np.allclose(po.lu_solve(lu, piv, x), np.linalg.solve(A, x), rtol=1e-13)


@pytest.mark.parametrize("bs", [1, 2, 3, 4, 8])
def test_cholesky_solve(bs):
    rng = np.random.default_rng(10 + bs)
    R = rng.standard_normal((bs, bs))
    A = R @ R.T + bs * np.eye(bs)
    L = po.chol_factor(A)
    x = rng.standard_normal(bs)
    assert np.allclose(po.chol_solve(L, x), np.linalg.solve(A, x), rtol=1e-12)


def test_gels_matches_lstsq():
    rng = np.random.default_rng(2)
    for m, n in ((5, 4), (31, 30), (2, 1)):
        H = np.triu(rng.standard_normal((m, n)), -1)
        y = rng.standard_normal(m)
        ref = np.linalg.lstsq(H, y, rcond=None)[0]
        assert np.allclose(po.gels(H, y), ref, rtol=1e-10, atol=1e-12)


def test_eigvals_match_numpy():
    rng = np.random.default_rng(3)
    for n in (1, 2, 5, 17, 30):
        A = rng.standard_normal((n, n))
        ev = po.eigvals(A)
        ref = np.linalg.eigvals(A)
        for e in ref:
            assert np.min(np.abs(ev - e)) <= 1e-9 * max(1.0, abs(e))


def _reference_selection(ev, k, n):
    """deflated_restart.hpp:124-165 on numpy eigenvalues (stable sort by modulus)."""
    zt = 10 * n * n * np.finfo(float).eps
    perm = np.argsort(np.abs(ev), kind="stable")
    sel = np.zeros(n, bool)
    cnt = 0
    for ip in perm:
        if cnt >= k:
            break
        if sel[ip]:
            continue
        sel[ip] = True
        cnt += 1
        if abs(ev[ip].imag) > zt:
            for j in range(n):
                if j != ip and abs(ev[j].real - ev[ip].real) < zt and abs(ev[j].imag + ev[ip].imag) < zt:
                    if not sel[j]:
                        sel[j] = True
                        cnt += 1
                    break
    return sel


@pytest.mark.parametrize("seed", range(6))
def test_harmonic_subspace_invariants(seed):
    """P spans the invariant subspace of H~ = H_n + h^2 H_n^{-T} e_n e_n^T for the k
    harmonic Ritz values of smallest modulus (conjugate pairs together)."""
    rng = np.random.default_rng(seed)
    n, k = 30, 10
    H = np.triu(rng.standard_normal((n + 1, n)), -1)
    for i in range(n):
        H[i + 1, i] = abs(H[i + 1, i]) + 0.1
    P, ev = po.harmonic_subspace(H, k)
    Hn = H[:n, :n]
    w = np.linalg.solve(Hn.T, np.eye(n)[:, -1])
    Ht = Hn.copy()
    Ht[:, -1] += H[n, n - 1] ** 2 * w
    kout = P.shape[1]
    assert kout in (k, k + 1)
    assert np.allclose(P.T @ P, np.eye(kout), atol=1e-12)
    T = P.T @ Ht @ P
    assert np.linalg.norm(Ht @ P - P @ T) <= 1e-9 * np.linalg.norm(Ht)
    ref_ev = np.linalg.eigvals(Ht)
    sel = _reference_selection(ref_ev, k, n)
    assert sel.sum() == kout
    got = np.sort_complex(np.linalg.eigvals(T))
    want = np.sort_complex(ref_ev[sel])
    assert np.allclose(got, want, rtol=1e-8, atol=1e-10)


# ---------------------------------------------------------------- bsmv

@pytest.mark.parametrize("bs", [1, 2, 3, 4, 8, 16])
def test_bsmv_matches_dense(bs):
    n = 40
    rp, ci, v = rand_bsr(n, bs, seed=bs)
    x = np.random.default_rng(0).standard_normal(n * bs)
    y0 = np.random.default_rng(1).standard_normal(n * bs)
    A = dense_of(n, bs, rp, ci, v)
    assert np.allclose(po.bsmv_arrays(n, bs, rp, ci, v, x), A @ x, rtol=1e-13, atol=1e-13)
    assert np.allclose(po.bsmv_arrays(n, bs, rp, ci, v, x, y0, accumulate=True), y0 + A @ x, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("bs", [2, 8])
def test_bsmv_reference_summation_order_bitwise(bs):
    """y_r = ((0 + blk_0) + blk_1) + ..., blk = sum_c a_rc x_c from val=0 in column
    order (bs=2: a_r0 x_0 + a_r1 x_1); blocksp_multiply.hpp:289-325,
    blocksp_dense_multiply.hpp:29-48."""
    n = 12
    rp, ci, v = rand_bsr(n, bs, seed=5)
    x = np.random.default_rng(6).standard_normal(n * bs)
    y = po.bsmv_arrays(n, bs, rp, ci, v, x)
    ref = np.zeros(n * bs)
    for i in range(n):
        for r in range(bs):
            acc = 0.0
            for p in range(rp[i], rp[i + 1]):
                a = v[p, r]
                xc = x[ci[p] * bs:(ci[p] + 1) * bs]
                if bs == 2:
                    blk = float(a[0]) * float(xc[0]) + float(a[1]) * float(xc[1])
                else:
                    blk = 0.0
                    for c in range(bs):
                        blk = blk + float(a[c]) * float(xc[c])
                acc = acc + blk
            ref[i * bs + r] = acc
    assert np.array_equal(y, ref)


def test_bsmv_dim_and_transpose():
    n, B, D = 15, 3, 2
    rp, ci, v = rand_bsr(n, B, seed=9)
    A = dense_of(n, B, rp, ci, v)
    x = np.random.default_rng(2).standard_normal(n * B * D)
    y = po.bsmv_dim(n, B, D, -1, rp, ci, v, x)
    X = x.reshape(n, D, B)
    for d in range(D):
        assert np.allclose(y.reshape(n, D, B)[:, d, :].reshape(-1), A @ X[:, d, :].reshape(-1), atol=1e-13)
    y1 = po.bsmv_dim(n, B, D, 1, rp, ci, v, x)
    assert np.all(y1.reshape(n, D, B)[:, 0, :] == 0)
    xt = np.random.default_rng(3).standard_normal(n * B)
    assert np.allclose(po.bsmv_transpose(n, n, B, rp, ci, v, xt), A.T @ xt, atol=1e-13)


# ---------------------------------------------------------------- preconditioners

def _lu_from_factor(n, bs, rp, ci, lu):
    L = np.eye(n * bs)
    U = np.zeros((n * bs, n * bs))
    for i in range(n):
        for p in range(rp[i], rp[i + 1]):
            j = ci[p]
            blk = lu[p]
            if j < i:
                L[i * bs:(i + 1) * bs, j * bs:(j + 1) * bs] = blk
            else:
                U[i * bs:(i + 1) * bs, j * bs:(j + 1) * bs] = blk
    return L, U


@pytest.mark.parametrize("case", ["stencil", "dg2d", "dg3d"])
def test_ilu0_defining_property(case):
    """(LU)_ij = A_ij on the block pattern (the definition of ILU(0))."""
    p = {"stencil": lambda: pr.stencil2d_b2(6, 5), "dg2d": lambda: pr.dg2d_veclap(5, 4),
         "dg3d": lambda: pr.dg3d_lap(3, 3, 2)}[case]()
    lu, dlu, dpiv = po.ilu0_factor(p)
    L, U = _lu_from_factor(p.n, p.bs, p.rowptr, p.colidx, lu)
    LU = L @ U
    A = pr.to_dense(p)
    bs = p.bs
    for i in range(p.n):
        for k in range(p.rowptr[i], p.rowptr[i + 1]):
            j = p.colidx[k]
            blk = (slice(i * bs, (i + 1) * bs), slice(j * bs, (j + 1) * bs))
            assert np.allclose(LU[blk], A[blk], rtol=1e-11, atol=1e-11 * np.abs(A).max())
    # the apply solves L U z = r
    r = np.random.default_rng(0).standard_normal(p.n * bs)
    z = po.ilu0_apply(p, lu, dlu, dpiv, r)
    assert np.allclose(LU @ z, r, rtol=1e-10, atol=1e-10)


def test_ilu0_exact_on_block_tridiagonal():
    """A block-tridiagonal pattern is closed under elimination: ILU0 = LU, M^{-1} A x = x."""
    p = pr.dg2d_veclap(7, 1)      # a 1D chain of elements
    lu, dlu, dpiv = po.ilu0_factor(p)
    x = np.random.default_rng(1).standard_normal(p.n * p.bs)
    z = po.ilu0_apply(p, lu, dlu, dpiv, po.bsmv(p, x))
    assert np.allclose(z, x, rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("case", ["dg2d", "dg3d"])
def test_ichol0_defining_property(case):
    p = {"dg2d": lambda: pr.dg2d_veclap(5, 4), "dg3d": lambda: pr.dg3d_lap(3, 2, 3)}[case]()
    lrp, lci, lv = po.ichol0_factor(p)
    bs = p.bs
    L = dense_of(p.n, bs, lrp, lci, lv)
    LLt = L @ L.T
    A = pr.to_dense(p)
    for i in range(p.n):
        for k in range(lrp[i], lrp[i + 1]):
            j = lci[k]
            blk = (slice(i * bs, (i + 1) * bs), slice(j * bs, (j + 1) * bs))
            assert np.allclose(LLt[blk], A[blk], rtol=1e-11, atol=1e-11 * np.abs(A).max())
    r = np.random.default_rng(2).standard_normal(p.n * bs)
    z = po.ichol0_apply(p.n, bs, lrp, lci, lv, r)
    assert np.allclose(LLt @ z, r, rtol=1e-10, atol=1e-10)


def _dense_sor(A, b, bs, w, nu, typ):
    n = A.shape[0] // bs
    x = np.zeros_like(b)

    def sweep(order):
        for i in order:
            sl = slice(i * bs, (i + 1) * bs)
            ph = b[sl] - (A[sl] @ x - A[sl, sl] @ x[sl])
            ph = np.linalg.solve(A[sl, sl], ph)
            x[sl] = x[sl] + (-w * x[sl] + w * ph)

    fw, bw = range(n), range(n - 1, -1, -1)
    if typ < 0:
        for _ in range(nu):
            sweep(bw)
    elif typ == 0:
        for _ in range(nu):
            sweep(fw)
        for _ in range(nu):
            sweep(bw)
    else:
        for _ in range(nu):
            sweep(fw)
    return x


@pytest.mark.parametrize("typ", [-1, 0, 1])
def test_sor_matches_dense_gauss_seidel(typ):
    p = pr.stencil2d_b2(5, 4)
    A = pr.to_dense(p)
    r = np.random.default_rng(3).standard_normal(p.n * p.bs)
    for w, nu in ((1.2, 3), (0.7, 1)):
        assert np.allclose(po.sor_apply(p, r, w, nu, typ), _dense_sor(A, r, p.bs, w, nu, typ), rtol=1e-11, atol=1e-12)


def test_sor_exact_for_block_lower_triangular():
    """w = 1, one forward sweep on a block lower-triangular matrix is the exact solve."""
    n, bs = 20, 3
    rp, ci, v = rand_bsr(n, bs, seed=4)
    keep = np.array([ci[k] <= i for i in range(n) for k in range(rp[i], rp[i + 1])])
    rows = np.repeat(np.arange(n), np.diff(rp))[keep]
    ci2 = ci[keep]
    v2 = v[keep]
    for k in range(len(ci2)):
        if ci2[k] == rows[k]:
            v2[k] += 4 * np.eye(bs)
    rp2 = np.zeros(n + 1, np.int32)
    rp2[1:] = np.cumsum(np.bincount(rows, minlength=n))
    p = pr.Problem(n, bs, bs, 1, rp2, ci2, v2, np.zeros(n * bs))
    A = pr.to_dense(p)
    r = np.random.default_rng(5).standard_normal(n * bs)
    assert np.allclose(po.sor_apply(p, r, 1.0, 1, 1), np.linalg.solve(A, r), rtol=1e-10)


def test_levels_brute_force():
    n = 60
    rp, ci, _ = rand_bsr(n, 1, seed=7, maxblk=5)
    for lower in (True, False):
        nl, lev = po.levels(n, rp, ci, lower)
        ref = np.zeros(n, int)
        order = range(n) if lower else range(n - 1, -1, -1)
        for i in order:
            deps = [c for c in ci[rp[i]:rp[i + 1]] if (c < i if lower else c > i)]
            ref[i] = 1 + max(ref[d] for d in deps) if deps else 0
        assert np.array_equal(lev, ref) and nl == ref.max() + 1


def test_levels_grid_closed_form():
    """Lexicographic nx x ny 5-point grid: level(ix, iy) = ix + iy, nx + ny - 1 levels."""
    nx, ny = 13, 7
    rp, ci, _ = pr.stencil_pattern((nx, ny))
    nl, lev = po.levels(nx * ny, rp, ci, True)
    ix, iy = np.arange(nx * ny) % nx, np.arange(nx * ny) // nx
    assert nl == nx + ny - 1 and np.array_equal(lev, ix + iy)


def test_diag_apply():
    p = pr.dg2d_veclap(4, 3)
    r = np.random.default_rng(6).standard_normal(p.n * p.bs)
    A = pr.to_dense(p)
    ref = np.concatenate([np.linalg.solve(A[i * 8:(i + 1) * 8, i * 8:(i + 1) * 8], r[i * 8:(i + 1) * 8])
                          for i in range(p.n)])
    assert np.allclose(po.diag_apply(p, r, 0), ref, rtol=1e-12)
    assert np.allclose(po.diag_apply(p, r, 1), ref, rtol=1e-12)


# ---------------------------------------------------------------- Krylov

def test_pgmres_krylov_optimality():
    """Without restart, after j iterations the GMRES iterate minimises
    ||M^{-1}(b - A x)|| over x in K_j(M^{-1}A, M^{-1}b) (here M = I)."""
    p = pr.stencil2d_b2(4, 3, seed=3)
    A = pr.to_dense(p)
    b = p.b
    for j in (3, 6):
        x, st = po.pgmres(p, b, 1e-30, po.PRE_NONE, maxIters=j - 1)
        assert st.iters == j
        K = np.stack([np.linalg.matrix_power(A, i) @ b for i in range(j)], axis=1)
        Q, _ = np.linalg.qr(K)
        c = np.linalg.lstsq(A @ Q, b, rcond=None)[0]
        xref = Q @ c
        assert np.linalg.norm(x - xref) <= 1e-8 * np.linalg.norm(xref)


def test_pgmres_exact_after_n_steps_and_deflated_restart():
    p = pr.stencil2d_b2(3, 3, seed=4)       # 18 unknowns
    x, st = po.pgmres(p, p.b, 1e-13, po.PRE_NONE, maxIters=100)
    assert st.true_relres < 1e-11 and st.iters <= 18
    x2, st2 = po.pgmres(p, p.b, 1e-12, po.PRE_NONE, maxIters=500, restart=6, k=3)
    assert st2.true_relres < 1e-10 and st2.restarts >= 1


def test_pcg_matches_numpy_solution():
    p = pr.dg2d_veclap(4, 4)
    A = pr.to_dense(p)
    x, st = po.pcg(p, p.b, 1e-12, po.PRE_NONE, maxIters=1000)
    assert np.allclose(x, np.linalg.solve(A, p.b), rtol=1e-8, atol=1e-10)


@pytest.fixture(scope="module")
def ldg():
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "ldg_2d_veclap.npz"))
    return z


def test_ldg_fixture_structure(ldg):
    """Blocking of the reference's fixture with its insertion order (the fixture table of the reference tests)."""
    counts = {2: 116610, 4: 42210, 8: 12706, 32: 1666}
    for bs, nnzb in counts.items():
        rp, ci, v = mtx.entries_to_bsr(int(ldg["M"]), int(ldg["N"]), ldg["i"], ldg["j"], ldg["v"], bs, symmetric=True)
        assert ci.shape[0] == nnzb


def test_mtx_insertion_order_brute_force():
    rng = np.random.default_rng(8)
    M = 12
    ent = [(int(rng.integers(0, M)), int(rng.integers(0, M)), float(rng.standard_normal())) for _ in range(40)]
    i = np.array([e[0] for e in ent])
    j = np.array([e[1] for e in ent])
    v = np.array([e[2] for e in ent])
    bs = 3
    rp, ci, vals = mtx.entries_to_bsr(M, M, i, j, v, bs, symmetric=True)
    rows = [[] for _ in range(M // bs)]
    blocks = {}
    for (a, b_, val) in ent:
        for (r, c) in ((a, b_), (b_, a)) if a != b_ else ((a, b_),):
            R, Cc = r // bs, c // bs
            if Cc not in rows[R]:
                rows[R].append(Cc)
                blocks[(R, Cc)] = np.zeros((bs, bs))
            blocks[(R, Cc)][r % bs, c % bs] = val
    assert np.array_equal(ci, np.array([c for r in rows for c in r]))
    k = 0
    for R in range(M // bs):
        for Cc in rows[R]:
            assert np.array_equal(vals[k], blocks[(R, Cc)])
            k += 1


@pytest.mark.parametrize("bs,pre", [(2, po.PRE_ILU0), (8, po.PRE_ILU0), (4, po.PRE_SOR), (8, po.PRE_ICHOL0)])
def test_ldg_known_answer(ldg, bs, pre):
    """The reference's known answer: ||x - x_sol||_inf near 0.00029 (blocksp_tests.hpp:318)."""
    rp, ci, v = mtx.entries_to_bsr(int(ldg["M"]), int(ldg["N"]), ldg["i"], ldg["j"], ldg["v"], bs, symmetric=True)
    p = pr.Problem(rp.shape[0] - 1, bs, bs, 1, rp, ci, v, ldg["b"])
    if pre == po.PRE_ICHOL0:
        x, st = po.pcg(p, p.b, 1e-8, pre, maxIters=5000)
    else:
        x, st = po.pgmres(p, p.b, 1e-8, pre, maxIters=5000, restart=30, k=10, w=1.6, nu=3)
    err = np.abs(x - ldg["x_sol"]).max()
    assert abs(err - 2.946e-4) < 2e-6, err


# ---------------------------------------------------------------- bsmm

def test_bsmm_first_touch_order_and_values():
    n, bs = 10, 2
    rpa, cia, va = rand_bsr(n, bs, seed=11)
    rpb, cib, vb = rand_bsr(n, bs, seed=12)
    crp, cci, cv = po.bsmm(n, bs, rpa, cia, va, rpb, cib, vb)
    for i in range(n):
        order = []
        for pa in range(rpa[i], rpa[i + 1]):
            for pb in range(rpb[cia[pa]], rpb[cia[pa] + 1]):
                if cib[pb] not in order:
                    order.append(cib[pb])
        assert list(cci[crp[i]:crp[i + 1]]) == order
    Ad, Bd = dense_of(n, bs, rpa, cia, va), dense_of(n, bs, rpb, cib, vb)
    assert np.allclose(dense_of(n, bs, crp, cci, cv), Ad @ Bd, atol=1e-12)


def _rect_bsr(n, m, bs, seed, maxblk=4):
    rng = np.random.default_rng(seed)
    rows = [list(rng.permutation(rng.choice(m, size=int(rng.integers(0, maxblk + 1)), replace=False)))
            for _ in range(n)]
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    ci = np.array([c for r in rows for c in r], np.int32)
    return rp, ci, rng.standard_normal((len(ci), bs, bs))


def _dense_rect(n, m, bs, rp, ci, v):
    A = np.zeros((n * bs, m * bs))
    for i in range(n):
        for p in range(rp[i], rp[i + 1]):
            A[i * bs:(i + 1) * bs, ci[p] * bs:(ci[p] + 1) * bs] += v[p]
    return A


def test_bsmm_transA_hand_example_order():
    # A (3 x 2 blocks, bs 1): row 0 = {1: 2}, row 1 = {0: 3, 1: 5}, row 2 = {0: 7}
    # B (3 x 3 blocks):        row 0 = {2: 1}, row 1 = {1: 1, 0: 2}, row 2 = {2: 4}
    # row 0 of C = A^T B is touched by i = 1 (cols 1, 0 of B row 1) then i = 2 (col 2):
    #   order [1, 0, 2], values [3, 6, 28]; row 1 by i = 0 (col 2) then i = 1 (cols 1, 0):
    #   order [2, 1, 0], values [2, 5, 10]
    arp, aci, av = np.array([0, 1, 3, 4], np.int32), np.array([1, 0, 1, 0], np.int32), np.array([2., 3., 5., 7.])
    brp, bci, bv = np.array([0, 1, 3, 4], np.int32), np.array([2, 1, 0, 2], np.int32), np.array([1., 1., 2., 4.])
    crp, cci, cv = po.bsmm_transA(3, 2, 1, arp, aci, av.reshape(-1, 1, 1), brp, bci, bv.reshape(-1, 1, 1))
    assert list(crp) == [0, 3, 6]
    assert list(cci) == [1, 0, 2, 2, 1, 0]
    assert list(cv.ravel()) == [3., 6., 28., 2., 5., 10.]


@pytest.mark.parametrize("bs", [1, 3])
def test_bsmm_transA_dense_and_explicit_transpose(bs):
    n, m, k = 11, 7, 9
    arp, aci, av = _rect_bsr(n, m, bs, seed=31 + bs)
    brp, bci, bv = _rect_bsr(n, k, bs, seed=41 + bs)
    crp, cci, cv = po.bsmm_transA(n, m, bs, arp, aci, av, brp, bci, bv)
    Ad, Bd = _dense_rect(n, m, bs, arp, aci, av), _dense_rect(n, k, bs, brp, bci, bv)
    np.testing.assert_allclose(_dense_rect(m, k, bs, crp, cci, cv), Ad.T @ Bd, atol=1e-12)
    # the same product through an explicit transpose (rows i ascending) and the plain bsmm
    tr = [[] for _ in range(m)]
    for i in range(n):
        for p in range(arp[i], arp[i + 1]):
            tr[aci[p]].append((i, av[p].T))
    trp = np.concatenate([[0], np.cumsum([len(r) for r in tr])]).astype(np.int32)
    tci = np.array([i for r in tr for i, _ in r], np.int32)
    tv = np.array([b for r in tr for _, b in r]).reshape(-1, bs, bs)
    orp, oci, ov = po.bsmm(m, bs, trp, tci, tv, brp, bci, bv)
    np.testing.assert_array_equal(crp, orp)
    np.testing.assert_array_equal(cci, oci)
    np.testing.assert_array_equal(cv, ov)


def test_bsmm_transA_trace_is_squared_frobenius_norm():
    # frobeniusNorm(A) = sqrt(trace(bsmm(A, A, true))) (utility.hpp:383-388)
    n, bs = 13, 4
    rp, ci, v = rand_bsr(n, bs, seed=5)
    crp, cci, cv = po.bsmm_transA(n, n, bs, rp, ci, v, rp, ci, v)
    tr = sum(np.trace(cv[p]) for i in range(n) for p in range(crp[i], crp[i + 1]) if cci[p] == i)
    assert abs(tr - np.sum(v * v)) <= 1e-12 * np.sum(v * v)
