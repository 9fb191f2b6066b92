"""Pins of the complex oracle (oracle/zoracle.py) on the CPU.

* Real data through the complex route reproduces the real oracle (liboracle.so, itself
  pinned elsewhere) wherever the two routes share their arithmetic: pGMRES without restart
  and with pure restarts (k = 0), for every preconditioner.
* Complex-only properties that a missing or misplaced conjugate breaks: scaling b by a
  unit complex number c scales every iterate by c (x -> c x, same iteration count), the
  residual estimate equals the true preconditioned residual, and the solution solves the
  dense complex system.
* ILU0's defining property ((LU)_ij = A_ij on the block pattern) and its exactness on a
  block-tridiagonal pattern; the deflated restart's invariant subspace, orthonormality and
  projections, with H~ built independently from the formula.
"""
import numpy as np
import pytest

import pyoracle as po
import zoracle as zo
from blocksp_amd import problems as pr


def complexify(p, seed=0, scale=0.3):
    rp, ci, v = zo.complexify(p, seed, scale)
    return zo.ZBsr(p.n, p.bs, rp, ci, v)


def rhs(n, seed=1):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


@pytest.fixture(scope="module")
def small():
    return pr.dg2d_veclap(6, 5, seed=3)          # 30 block rows, bs 8


def test_bsmv_and_block_lu(small):
    A = complexify(small)
    x = rhs(A.n * A.bs)
    np.testing.assert_allclose(zo.bsmv(A, x), A.dense() @ x, rtol=1e-13, atol=1e-13)
    blk = A.v[0] + 2 * np.eye(8)
    f = zo.LU(blk)
    y = rhs(8, 4)
    np.testing.assert_allclose(f.solve(y), np.linalg.solve(blk, y), rtol=1e-12)
    X = rhs(64, 5).reshape(8, 8)
    np.testing.assert_allclose(f.rsolve(X), X @ np.linalg.inv(blk), rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("kind,kw", [(po.PRE_NONE, {}), (po.PRE_DIAG, {}), (po.PRE_SOR, dict(w=1.2, nu=2)),
                                     (po.PRE_ILU0, {})])
@pytest.mark.parametrize("restart,k", [(-1, -1), (12, 0)])
def test_real_data_matches_real_oracle(small, kind, kw, restart, k):
    p = small
    A = zo.ZBsr(p.n, p.bs, p.rowptr, p.colidx, p.vals.astype(np.complex128))
    M = zo.Precond(A, kind, **kw)
    xz, sz = zo.pgmres(A, p.b.astype(np.complex128), 1e-10, M, maxIters=400, restart=restart, k=k)
    xr, sr = po.pgmres(p, p.b, 1e-10, kind, maxIters=400, restart=restart, k=k, **kw)
    assert sz.iters == sr.iters, (sz.iters, sr.iters)
    assert np.abs(xz.imag).max() == 0.0
    np.testing.assert_allclose(xz.real, xr, rtol=0, atol=1e-9 * np.abs(xr).max())


def property_failures(A, b, M):
    """The complex-only properties of a pGMRES solve; returns the names of those that fail."""
    bad = []
    x, st = zo.pgmres(A, b, 1e-10, M, maxIters=600, restart=8, k=3)
    if st.restarts < 1:
        bad.append("no restart")
    # solves the dense complex system
    xd = np.linalg.solve(A.dense(), b)
    if not np.linalg.norm(x - xd) <= 1e-7 * np.linalg.norm(xd):
        bad.append("dense solution")
    # the least-squares residual estimate is the true preconditioned residual
    true = zo.vnorm(M(b - zo.bsmv(A, x)))
    if not (abs(true - st.res_est) <= 1e-6 * st.res0 + 1e-3 * true and st.res_est <= 1e-10 * st.res0):
        bad.append("residual estimate")
    # b -> c b with |c| = 1 gives x -> c x in the same number of iterations
    c = np.exp(0.7j)
    xc, stc = zo.pgmres(A, c * b, 1e-10, M, maxIters=600, restart=8, k=3)
    if stc.iters != st.iters or not np.allclose(xc, c * x, rtol=0, atol=1e-10 * np.abs(x).max()):
        bad.append("unit-modulus scaling")
    return bad


@pytest.mark.parametrize("kind,kw", [(zo.PRE_ILU0, {}), (zo.PRE_SOR, dict(w=1.0, nu=2)), (zo.PRE_DIAG, {})])
def test_complex_solve_properties(small, kind, kw):
    A = complexify(small, seed=2)
    b = rhs(A.n * A.bs, 3)
    assert property_failures(A, b, zo.Precond(A, kind, **kw)) == []


@pytest.mark.parametrize("wrong", ["no conjugate", "conjugate on the other operand"])
def test_planted_conjugation_mistakes_fail_the_pins(small, monkeypatch, wrong):
    """A plausible mistake planted in the oracle's inner product is caught by the properties."""
    fn = (lambda a, b: np.dot(a, b)) if wrong == "no conjugate" else (lambda a, b: np.vdot(b, a))
    monkeypatch.setattr(zo, "vdot", fn)
    A = complexify(small, seed=2)
    b = rhs(A.n * A.bs, 3)
    assert property_failures(A, b, zo.Precond(A, zo.PRE_ILU0)) != []


def test_ilu0_defining_property_and_tridiagonal_exactness():
    p = pr.dg2d_veclap(5, 4, seed=1)
    A = complexify(p, seed=5)
    LU_, dlu = zo.ilu0_factor(A)
    bs = A.bs
    L = np.eye(A.n * bs, dtype=complex)
    U = np.zeros((A.n * bs, A.n * bs), complex)
    for i in range(A.n):
        for q in range(A.rp[i], A.rp[i + 1]):
            c = A.ci[q]
            tgt = L if c < i else U
            tgt[i * bs:(i + 1) * bs, c * bs:(c + 1) * bs] = LU_[q]
    # block-unit lower L (identity diagonal blocks): (L U) equals A on every stored block
    prod = L @ U
    for i in range(A.n):
        for q in range(A.rp[i], A.rp[i + 1]):
            c = A.ci[q]
            np.testing.assert_allclose(prod[i * bs:(i + 1) * bs, c * bs:(c + 1) * bs], A.v[q],
                                       rtol=0, atol=1e-11 * np.abs(A.v).max())
    # block tridiagonal: no fill is dropped, so ILU0 is the exact inverse
    n, bs = 12, 3
    rng = np.random.default_rng(7)
    rp, ci, vals = [0], [], []
    for i in range(n):
        for c in (i, i - 1, i + 1):               # stored order: diagonal first
            if 0 <= c < n:
                ci.append(c)
                blk = rng.standard_normal((bs, bs)) + 1j * rng.standard_normal((bs, bs))
                vals.append(blk + (6 * np.eye(bs) if c == i else 0))
        rp.append(len(ci))
    T = zo.ZBsr(n, bs, rp, ci, np.array(vals))
    M = zo.Precond(T, zo.PRE_ILU0)
    x = rhs(n * bs, 8)
    np.testing.assert_allclose(M(zo.bsmv(T, x)), x, rtol=0, atol=1e-11 * np.abs(x).max())


@pytest.mark.parametrize("k", [1, 4, 7])
def test_deflated_restart_identities(k):
    rng = np.random.default_rng(k)
    n = 12
    H = np.triu(rng.standard_normal((n + 1, n)) + 1j * rng.standard_normal((n + 1, n)), -1)
    res_vec = rng.standard_normal(n + 1) + 1j * rng.standard_normal(n + 1)
    Qn = [rng.standard_normal(20) + 1j * rng.standard_normal(20) for _ in range(n + 1)]
    Hnew, Qnew, c_r = zo.deflated_restart(H, Qn, res_vec, k)
    assert Hnew.shape == (k + 1, k) and len(Qnew) == k + 1 and c_r.shape == (k + 1,)
    # H~ = H_n + |h|^2 H_n^{-H} e_n e_n^T, formed here with an explicit inverse
    Hn = H[:n, :n]
    Ht = Hn.copy()
    Ht[:, n - 1] += abs(H[n, n - 1]) ** 2 * np.linalg.inv(Hn.conj().T)[:, n - 1]
    ev = np.linalg.eigvals(Ht)
    smallest = np.sort(np.abs(ev))[:k]
    # recover P_k and P_{k+1} from Qn_new = Qn P_{k+1} (Qn here has full column rank)
    Qm = np.array(Qn).T
    P = np.linalg.lstsq(Qm, np.array(Qnew).T, rcond=None)[0]
    np.testing.assert_allclose(P.conj().T @ P, np.eye(k + 1), atol=1e-10)
    Pk = P[:n, :k]
    assert np.abs(P[n, :k]).max() < 1e-10
    R = Pk.conj().T @ Ht @ Pk
    assert np.linalg.norm(Ht @ Pk - Pk @ R) < 1e-9 * np.linalg.norm(Ht)
    np.testing.assert_allclose(np.sort(np.abs(np.linalg.eigvals(R))), smallest, rtol=1e-8)
    np.testing.assert_allclose(c_r, P.conj().T @ res_vec, atol=1e-10)
    np.testing.assert_allclose(Hnew, P.conj().T @ H @ Pk, atol=1e-9)
