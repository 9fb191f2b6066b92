"""ctypes bindings of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the blocksp_amd product.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
# the same oracle built with the reference's release flags (-O3 -funroll-loops -fstrict-aliasing
# -fomit-frame-pointer -ffast-math); timing only, never a parity checker
_LIB_FAST = os.path.join(_HERE, "liboracle_fast.so")
_use_fast = False

PRE_NONE, PRE_DIAG, PRE_SOR, PRE_ILU0, PRE_ICHOL0 = 0, 1, 2, 3, 4


class Precond(C.Structure):
    _fields_ = [("kind", C.c_int), ("inv_kind", C.c_int), ("w", C.c_double), ("nu", C.c_int),
                ("type", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("iters", C.c_int), ("restarts", C.c_int), ("res0", C.c_double),
                ("res_est", C.c_double), ("true_relres", C.c_double)]


def build(force: bool = False) -> str:
    if force or not (os.path.exists(_LIB) and os.path.exists(_LIB_FAST)):
        subprocess.check_call(["make", "-s", "-B" if force else "-s", "-C", _HERE])
    return _LIB


_lib = None


def use_fast_build(on: bool = True) -> None:
    """Route the following calls to liboracle_fast.so (True) or the parity build (False)."""
    global _use_fast, _lib
    if on != _use_fast:
        _use_fast = on
        _lib = None


def lib():
    global _lib
    if _lib is None:
        path = _LIB_FAST if _use_fast else _LIB
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        for name in ("orc_pcg", "orc_pgmres", "orc_harmonic_subspace", "orc_ichol0_factor",
                     "orc_levels", "orc_bsmm", "orc_bsmm_transA"):
            getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


def _f64(a):
    return np.ascontiguousarray(a, np.float64)


def bsmv(p, x, y=None, accumulate=False):
    x = _f64(x)
    y = np.zeros(p.n * p.bs) if y is None else _f64(y).copy()
    lib().orc_bsmv(p.n, p.bs, _p(_i32(p.rowptr), C.c_int), _p(_i32(p.colidx), C.c_int),
                   _p(_f64(p.vals), C.c_double), _p(x, C.c_double), _p(y, C.c_double),
                   int(accumulate))
    return y


def bsmv_arrays(n, bs, rowptr, colidx, vals, x, y=None, accumulate=False):
    x = _f64(x)
    y = np.zeros(n * bs) if y is None else _f64(y).copy()
    lib().orc_bsmv(n, bs, _p(_i32(rowptr), C.c_int), _p(_i32(colidx), C.c_int),
                   _p(_f64(vals), C.c_double), _p(x, C.c_double), _p(y, C.c_double), int(accumulate))
    return y


def bsmv_dim(n, B, D, dim, rowptr, colidx, vals, x, y=None, accumulate=False):
    x = _f64(x)
    y = np.zeros(n * B * D) if y is None else _f64(y).copy()
    lib().orc_bsmv_dim(n, B, D, dim, _p(_i32(rowptr), C.c_int), _p(_i32(colidx), C.c_int),
                       _p(_f64(vals), C.c_double), _p(x, C.c_double), _p(y, C.c_double),
                       int(accumulate))
    return y


def bsmv_transpose(nrows, ncols, bs, rowptr, colidx, vals, x, y=None, accumulate=False):
    x = _f64(x)
    y = np.zeros(ncols * bs) if y is None else _f64(y).copy()
    lib().orc_bsmv_transpose(nrows, ncols, bs, _p(_i32(rowptr), C.c_int), _p(_i32(colidx), C.c_int),
                             _p(_f64(vals), C.c_double), _p(x, C.c_double), _p(y, C.c_double),
                             int(accumulate))
    return y


def lu_factor(a):
    a = _f64(a)
    bs = a.shape[0]
    lu = np.zeros((bs, bs))
    piv = np.zeros(bs, np.int32)
    lib().orc_lu_factor(bs, _p(a, C.c_double), _p(lu, C.c_double), _p(piv, C.c_int))
    return lu, piv


def lu_solve(lu, piv, x):
    x = _f64(x).copy()
    lib().orc_lu_solve(lu.shape[0], _p(_f64(lu), C.c_double), _p(_i32(piv), C.c_int), _p(x, C.c_double))
    return x


def lu_rsolve(lu, piv, X):
    X = _f64(X).copy()
    lib().orc_lu_rsolve(lu.shape[0], _p(_f64(lu), C.c_double), _p(_i32(piv), C.c_int), _p(X, C.c_double))
    return X


def chol_factor(a):
    a = _f64(a)
    bs = a.shape[0]
    L = np.zeros(bs * (bs + 1) // 2)
    lib().orc_chol_factor(bs, _p(a, C.c_double), _p(L, C.c_double))
    return L


def chol_solve(L, x):
    x = _f64(x).copy()
    bs = x.shape[0]
    lib().orc_chol_solve(bs, _p(_f64(L), C.c_double), _p(x, C.c_double))
    return x


def ilu0_factor(p):
    lu = np.zeros_like(_f64(p.vals))
    dlu = np.zeros((p.n, p.bs, p.bs))
    dpiv = np.zeros((p.n, p.bs), np.int32)
    lib().orc_ilu0_factor(p.n, p.bs, _p(_i32(p.rowptr), C.c_int), _p(_i32(p.colidx), C.c_int),
                          _p(_f64(p.vals), C.c_double), _p(lu, C.c_double), _p(dlu, C.c_double),
                          _p(dpiv, C.c_int))
    return lu, dlu, dpiv


def ilu0_apply(p, lu, dlu, dpiv, x):
    x = _f64(x).copy()
    lib().orc_ilu0_apply(p.n, p.bs, _p(_i32(p.rowptr), C.c_int), _p(_i32(p.colidx), C.c_int),
                         _p(_f64(lu), C.c_double), _p(_f64(dlu), C.c_double), _p(_i32(dpiv), C.c_int),
                         _p(x, C.c_double))
    return x


def ichol0_factor(p):
    lrp = np.zeros(p.n + 1, np.int32)
    lci = np.zeros(p.nnzb, np.int32)
    lv = np.zeros((p.nnzb, p.bs, p.bs))
    nz = lib().orc_ichol0_factor(p.n, p.bs, _p(_i32(p.rowptr), C.c_int), _p(_i32(p.colidx), C.c_int),
                                 _p(_f64(p.vals), C.c_double), _p(lrp, C.c_int), _p(lci, C.c_int),
                                 _p(lv, C.c_double))
    return lrp, lci[:nz].copy(), lv[:nz].copy()


def ichol0_apply(n, bs, lrp, lci, lv, x):
    x = _f64(x).copy()
    lib().orc_ichol0_apply(n, bs, _p(_i32(lrp), C.c_int), _p(_i32(lci), C.c_int), _p(_f64(lv), C.c_double),
                           _p(x, C.c_double))
    return x


def sor_apply(p, x, w, nu, type_=0, inv_kind=0):
    x = _f64(x).copy()
    lib().orc_sor_apply(p.n, p.bs, _p(_i32(p.rowptr), C.c_int), _p(_i32(p.colidx), C.c_int),
                        _p(_f64(p.vals), C.c_double), C.c_double(w), nu, type_, inv_kind, _p(x, C.c_double))
    return x


def diag_apply(p, x, inv_kind=0):
    x = _f64(x).copy()
    lib().orc_diag_apply(p.n, p.bs, _p(_i32(p.rowptr), C.c_int), _p(_i32(p.colidx), C.c_int),
                         _p(_f64(p.vals), C.c_double), inv_kind, _p(x, C.c_double))
    return x


def levels(n, rowptr, colidx, lower=True):
    lev = np.zeros(n, np.int32)
    nl = lib().orc_levels(n, _p(_i32(rowptr), C.c_int), _p(_i32(colidx), C.c_int), int(lower),
                          _p(lev, C.c_int))
    return nl, lev


def _pre(kind, w=1.6, nu=3, type_=0, inv_kind=0):
    return Precond(kind, inv_kind, w, nu, type_)


def pcg(p, b, tol, pre_kind=PRE_ICHOL0, maxIters=5000, **kw):
    x = np.zeros(p.n * p.bs)
    st = Stats()
    pr = _pre(pre_kind, **kw)
    lib().orc_pcg(p.n, p.bs, _p(_i32(p.rowptr), C.c_int), _p(_i32(p.colidx), C.c_int),
                  _p(_f64(p.vals), C.c_double), C.byref(pr), _p(_f64(b), C.c_double), _p(x, C.c_double),
                  C.c_double(tol), maxIters, C.byref(st))
    return x, st


def pgmres(p, b, tol, pre_kind=PRE_ILU0, maxIters=5000, restart=-1, k=-1, **kw):
    x = np.zeros(p.n * p.bs)
    st = Stats()
    pr = _pre(pre_kind, **kw)
    lib().orc_pgmres(p.n, p.bs, _p(_i32(p.rowptr), C.c_int), _p(_i32(p.colidx), C.c_int),
                     _p(_f64(p.vals), C.c_double), C.byref(pr), _p(_f64(b), C.c_double), _p(x, C.c_double),
                     C.c_double(tol), maxIters, restart, k, C.byref(st))
    return x, st


def harmonic_subspace(H, k):
    H = _f64(H)
    n = H.shape[1]
    P = np.zeros(n * n)
    re = np.zeros(n)
    im = np.zeros(n)
    kout = lib().orc_harmonic_subspace(n, _p(H, C.c_double), k, _p(P, C.c_double), _p(re, C.c_double),
                                       _p(im, C.c_double))
    return P[: n * kout].reshape(n, kout), re + 1j * im


def eigvals(A):
    A = _f64(A)
    n = A.shape[0]
    re = np.zeros(n)
    im = np.zeros(n)
    lib().orc_eigvals(n, _p(A, C.c_double), _p(re, C.c_double), _p(im, C.c_double))
    return re + 1j * im


def gels(H, y):
    H = _f64(H)
    y = _f64(y).copy()
    m, n = H.shape
    lib().orc_gels(m, n, _p(H, C.c_double), _p(y, C.c_double))
    return y[:n]


def bsmm(pa_n, bs, arp, aci, av, brp, bci, bv):
    cap = int(sum((brp[aci[k] + 1] - brp[aci[k]]) for k in range(len(aci))))
    crp = np.zeros(pa_n + 1, np.int32)
    cci = np.zeros(max(cap, 1), np.int32)
    cv = np.zeros((max(cap, 1), bs, bs))
    nz = lib().orc_bsmm(pa_n, bs, _p(_i32(arp), C.c_int), _p(_i32(aci), C.c_int), _p(_f64(av), C.c_double),
                        _p(_i32(brp), C.c_int), _p(_i32(bci), C.c_int), _p(_f64(bv), C.c_double),
                        _p(crp, C.c_int), _p(cci, C.c_int), _p(cv, C.c_double), cap)
    assert nz >= 0
    return crp, cci[:nz].copy(), cv[:nz].copy()


def bsmm_transA(n, m, bs, arp, aci, av, brp, bci, bv):
    """C = A^T B (bsmm with Atran, blocksp_multiply.hpp:99-134); A is n x m blocks, B has n rows."""
    cap = int(sum((arp[i + 1] - arp[i]) * (brp[i + 1] - brp[i]) for i in range(n)))
    crp = np.zeros(m + 1, np.int32)
    cci = np.zeros(max(cap, 1), np.int32)
    cv = np.zeros((max(cap, 1), bs, bs))
    nz = lib().orc_bsmm_transA(n, m, bs, _p(_i32(arp), C.c_int), _p(_i32(aci), C.c_int), _p(_f64(av), C.c_double),
                               _p(_i32(brp), C.c_int), _p(_i32(bci), C.c_int), _p(_f64(bv), C.c_double),
                               _p(crp, C.c_int), _p(cci, C.c_int), _p(cv, C.c_double), cap)
    assert nz >= 0
    return crp, cci[:nz].copy(), cv[:nz].copy()
