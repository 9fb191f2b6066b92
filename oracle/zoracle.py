"""Complex (std::complex<double>) restatement of the block solve path: bsmv, the LU block
factor, SOR and ILU0 preconditioners and left-preconditioned pGMRES with Morgan's deflated
restart, as the reference runs them for T = std::complex<double>.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the blocksp_amd product.

The reference's complex path is its generic template code with the complex branches
taken (is_complex_v<T>): conjugated dot products (blocksp_containers.hpp:974-996), the
LAPACK z-routines behind LUFactor, gels, gees/trsen and geqrf/ungqr
(blocksp_lapack.hpp), and the complex branch of deflated_restart
(blocksp_deflated_restart.hpp:84-122, 186-245).  This restatement calls the same LAPACK
routines through scipy (zgetrf/zgetrs, zgels, zgees, ztrsen, zgeqrf/zungqr) and follows
the reference's loop order for the sparse parts.  Block matrices are (n, bs, rowptr,
colidx, vals[nnzb, bs, bs]) with row-major blocks in stored (insertion) order.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla
from scipy.linalg import lapack

PRE_NONE, PRE_DIAG, PRE_SOR, PRE_ILU0 = 0, 1, 2, 3

_EPS = np.finfo(np.float64).eps


class ZBsr:
    def __init__(self, n, bs, rowptr, colidx, vals):
        self.n, self.bs = int(n), int(bs)
        self.rp = np.asarray(rowptr, np.int64)
        self.ci = np.asarray(colidx, np.int64)
        self.v = np.asarray(vals, np.complex128).reshape(-1, bs, bs)

    def find(self, i, j):
        """position of block (i, j) in row i, or -1 (containers.hpp:206-212)"""
        for p in range(self.rp[i], self.rp[i + 1]):
            if self.ci[p] == j:
                return p
        return -1

    def dense(self):
        bs = self.bs
        D = np.zeros((self.n * bs, self.n * bs), np.complex128)
        for i in range(self.n):
            for p in range(self.rp[i], self.rp[i + 1]):
                c = self.ci[p]
                D[i * bs:(i + 1) * bs, c * bs:(c + 1) * bs] = self.v[p]
        return D


def complexify(p, seed=0, scale=0.3):
    """A complex test matrix on the block pattern of the real problem p: every block times
    (1 + i s u), u ~ U(-1, 1) per entry, plus i s/2 (mean |diagonal|) on the diagonal blocks --
    a non-Hermitian, shifted operator.  Returns (rowptr, colidx, complex vals)."""
    rng = np.random.default_rng(seed)
    v = p.vals * (1.0 + 1j * scale * rng.uniform(-1, 1, p.vals.shape))
    for i in range(p.n):
        for q in range(p.rowptr[i], p.rowptr[i + 1]):
            if p.colidx[q] == i:
                v[q] = v[q] + 0.5j * scale * np.abs(np.diag(p.vals[q])).mean() * np.eye(p.bs)
    return np.asarray(p.rowptr), np.asarray(p.colidx), v


def bsmv(A: ZBsr, x):
    """y_i = sum over row i's blocks, in stored order, of A_ij x_j (multiply.hpp:365-390)."""
    bs = A.bs
    x = np.asarray(x, np.complex128).reshape(A.n, bs)
    y = np.zeros((A.n, bs), np.complex128)
    for i in range(A.n):
        acc = np.zeros(bs, np.complex128)
        for p in range(A.rp[i], A.rp[i + 1]):
            acc += A.v[p] @ x[A.ci[p]]
        y[i] = acc
    return y.reshape(-1)


def vdot(a, b):
    """dot(x1, x2) with the complex conjugate on x1 (containers.hpp:982-986)."""
    return np.vdot(a, b)


def vnorm(a):
    return float(np.sqrt(abs(np.vdot(a, a))))


class LU:
    """dense::LUFactor on one block: zgetrf, then zgetrs for x <- A^{-1} x and X <- X A^{-1}
    (dense_solvers.hpp:80-108)."""

    def __init__(self, a):
        self.lu, self.piv, info = lapack.zgetrf(np.array(a, np.complex128))
        if info > 0:
            raise ZeroDivisionError("singular block")

    def solve(self, x):
        out, info = lapack.zgetrs(self.lu, self.piv, np.asarray(x, np.complex128))
        return out

    def rsolve(self, X):
        # X A^{-1}: rows r of X solve A^T y = X[r]^T (the plain transpose)
        out, info = lapack.zgetrs(self.lu, self.piv, np.asarray(X, np.complex128).T, trans=1)
        return out.T


class Precond:
    """preConditioners::NoPre / Diag<LUFactor> / SOR<LUFactor> / ILU0
    (blocksp_preconditioners.hpp)."""

    def __init__(self, A: ZBsr, kind=PRE_ILU0, w=1.0, nu=1, sor_type=0):
        self.A, self.kind, self.w, self.nu, self.type = A, kind, w, nu, sor_type
        if kind in (PRE_DIAG, PRE_SOR):
            self.di = [LU(A.v[A.find(i, i)]) for i in range(A.n)]
        if kind == PRE_ILU0:
            self.lu, self.dlu = ilu0_factor(A)

    def __call__(self, r):
        A, bs = self.A, self.A.bs
        r = np.asarray(r, np.complex128)
        if self.kind == PRE_NONE:
            return r.copy()
        if self.kind == PRE_DIAG:
            x = r.reshape(A.n, bs).copy()
            for i in range(A.n):
                x[i] = self.di[i].solve(x[i])
            return x.reshape(-1)
        if self.kind == PRE_SOR:
            x = np.zeros((A.n, bs), np.complex128)
            b = r.reshape(A.n, bs)
            if self.type < 0:
                dirs = [False] * self.nu
            elif self.type == 0:
                dirs = [True] * self.nu + [False] * self.nu
            else:
                dirs = [True] * self.nu
            for fwd in dirs:
                sor_sweep(A, self.di, b, x, self.w, fwd)
            return x.reshape(-1)
        return ilu0_apply(A, self.lu, self.dlu, r)


def sor_sweep(A: ZBsr, di, b, x, w, fwd):
    """One SOR sweep (preconditioners.hpp:135-202): x_i += w (D_i^{-1}(b_i - sum_{j!=i} A_ij x_j) - x_i)."""
    order = range(A.n) if fwd else range(A.n - 1, -1, -1)
    for i in order:
        ph = np.zeros(A.bs, np.complex128)
        for p in range(A.rp[i], A.rp[i + 1]):
            c = A.ci[p]
            if c != i:
                ph += A.v[p] @ x[c]
        ph = di[i].solve(b[i] - ph)
        x[i] = x[i] + (-w * x[i] + w * ph)


def ilu0_factor(A: ZBsr):
    """ILU0 factor (preconditioners.hpp:274-291), right-looking over the block pattern."""
    LU_ = A.v.copy()
    rows_of_col = [[] for _ in range(A.n)]
    for i in range(A.n):
        for p in range(A.rp[i], A.rp[i + 1]):
            rows_of_col[A.ci[p]].append(i)
    for i in range(A.n):
        f = LU(LU_[A.find(i, i)])
        for k in rows_of_col[i]:
            if k > i:
                q = A.find(k, i)
                LU_[q] = f.rsolve(LU_[q])
        for k in rows_of_col[i]:
            if k <= i:
                continue
            lki = LU_[A.find(k, i)]
            for p in range(A.rp[k], A.rp[k + 1]):
                j = A.ci[p]
                if j <= i:
                    continue
                pij = A.find(i, j)
                if pij < 0:
                    continue
                LU_[p] = LU_[p] - lki @ LU_[pij]
    dlu = [LU(LU_[A.find(i, i)]) for i in range(A.n)]
    return LU_, dlu


def ilu0_apply(A: ZBsr, LU_, dlu, r):
    bs = A.bs
    x = np.asarray(r, np.complex128).reshape(A.n, bs).copy()
    for i in range(1, A.n):
        for p in range(A.rp[i], A.rp[i + 1]):
            c = A.ci[p]
            if c < i:
                x[i] -= LU_[p] @ x[c]
    for i in range(A.n - 1, -1, -1):
        for p in range(A.rp[i], A.rp[i + 1]):
            c = A.ci[p]
            if c > i:
                x[i] -= LU_[p] @ x[c]
        x[i] = dlu[i].solve(x[i])
    return x.reshape(-1)


def gels(H, c):
    """zgels 'N' on the m x n Hessenberg (solvers.hpp:198, 220): least-squares y."""
    m, n = H.shape
    out = lapack.zgels(np.array(H, np.complex128), np.array(c, np.complex128).reshape(m, 1))
    return out[1][:n, 0].copy()


def harmonic_schur(Ht, k):
    """gees + trsen of the complex branch (deflated_restart.hpp:97-122): Schur form of H~,
    the k eigenvalues of smallest modulus (std::sort on |lambda|, ties by index) moved to
    the front in their Schur order.  Returns (Z, selected eigenvalues)."""
    n = Ht.shape[0]
    T, Z = sla.schur(np.array(Ht, np.complex128), output="complex")
    ev = np.diag(T).copy()
    perm = sorted(range(n), key=lambda i: (abs(ev[i]), i))
    select = np.zeros(n, np.int32)
    for i in range(k):
        select[perm[i]] = 1
    ts, qs, w, m, s, sep, info = lapack.ztrsen(select, T, Z, job="N")
    assert info == 0 and m == k
    return qs, w[:k]


def deflated_restart(H, Qn, res_vec, k):
    """Complex deflated_restart (deflated_restart.hpp:38-269).  H is (n+1) x n, Qn the n+1
    basis vectors (rows), res_vec the n+1 least-squares residual coefficients.  Returns
    (H_new (k+1) x k, Qn_new (k+1 rows), c_r), or a pure restart's (None, [], [0]) for k = 0."""
    m, n = H.shape
    k = min(k, n)
    if k == 0:
        return None, [], np.zeros(1, np.complex128)
    Hn = H[:n, :n].copy()
    h_mn = H[n, n - 1]
    e = np.zeros(n, np.complex128)
    e[n - 1] = 1.0
    w = LU(Hn.conj().T).solve(e)                    # H_n^{-H} e_n (transpose = conjugate transpose)
    Ht = Hn.copy()
    Ht[:, n - 1] += (np.conj(h_mn) * h_mn) * w
    Z, _ = harmonic_schur(Ht, k)
    Pk, _ = np.linalg.qr(Z[:, :k])                  # geqrf + ungqr
    P = np.zeros((n + 1, k + 1), np.complex128)
    P[:n, :k] = Pk
    P[:, k] = res_vec
    for i in range(k):
        d = np.vdot(P[:, i], P[:, k])
        P[:, k] -= d * P[:, i]
    P[:, k] /= np.sqrt(np.vdot(P[:, k], P[:, k]))
    c_r = P.conj().T @ res_vec
    Hnew = P.conj().T @ (H @ Pk)
    Qnew = P.T @ np.asarray(Qn)                     # Qn_new[ik] = sum_im Qn[im] P[im, ik]
    return Hnew, list(Qnew), c_r


class Stats:
    iters = 0
    restarts = 0
    res0 = 0.0
    res_est = 0.0
    true_relres = 0.0


def pgmres(A: ZBsr, b, tol, M: Precond, maxIters=500, restart=-1, k=-1):
    """Left-preconditioned pGMRES (solvers.hpp:156-295) with one_arnoldi (109-136, modified
    Gram-Schmidt with conjugated dots), gels on the whole Hessenberg every step and the
    deflated restart; x_0 = 0.  Returns (x, Stats)."""
    b = np.asarray(b, np.complex128)
    x0 = np.zeros_like(b)
    r0 = M(b)
    res_mag = [vnorm(r0)]
    Qn = [r0 / res_mag[-1]]
    zero_thresh = min(res_mag[-1] * _EPS, _EPS)
    tolerance = res_mag[-1] * tol
    st = Stats()
    st.res0 = res_mag[-1]
    H = np.zeros((1, 0), np.complex128)
    it = tot = 0
    y = np.zeros(0, np.complex128)
    c_r = None
    restarted = False
    while res_mag[-1] > tolerance:
        m, n = it + 2, it + 1
        Hn = np.zeros((m, n), np.complex128)
        Hn[:H.shape[0], :H.shape[1]] = H[:m, :n]
        H = Hn
        v = M(bsmv(A, Qn[it]))
        for j in range(it + 1):
            h = vdot(Qn[j], v)
            H[j, it] = h
            v = v - h * Qn[j]
        h = vnorm(v)
        H[it + 1, it] = h
        breakdown = abs(h) < zero_thresh
        if not breakdown:
            Qn.append(v / h)
        c = np.zeros(m, np.complex128)
        if not restarted:
            c[0] = res_mag[0]
        else:
            c[:min(len(c_r), m)] = c_r[:m]
        y = gels(H, c)
        res_vec = c - H @ y
        res_mag.append(float(np.sqrt(np.sum(np.abs(np.conj(res_vec) * res_vec)))))
        it += 1
        tot += 1
        if tot > maxIters or breakdown or not np.isfinite(res_mag[-1]):
            break
        if it == restart and res_mag[-1] > tolerance:
            for iy in range(len(y)):
                x0 = x0 + y[iy] * Qn[iy]
            rr = M(b - bsmv(A, x0))
            H, Qn, c_r = deflated_restart(H, Qn[:it + 1], res_vec, k)
            st.restarts += 1
            res_mag = [res_mag[-1]]
            if k == 0:
                res_mag = [vnorm(rr)]
                c_r = np.array([res_mag[-1]], np.complex128)
                Qn = [rr / res_mag[-1]]
                H = np.zeros((1, 0), np.complex128)
                it = 0
                restarted = False
            else:
                it = H.shape[1]
                restarted = True
            y = np.zeros(0, np.complex128)
    x = x0.copy()
    for iy in range(len(y)):
        x = x + y[iy] * Qn[iy]
    st.iters = tot
    st.res_est = res_mag[-1]
    st.true_relres = vnorm(b - bsmv(A, x)) / vnorm(b)
    return x, st
