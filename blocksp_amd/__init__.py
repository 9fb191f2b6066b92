"""blocksp_amd -- MI355X-native block-sparse solve path (BlockSp drop-in).

Python bindings over the C ABI of libblocksp_amd.so (include/blocksp_c.h).
The names follow the reference's surface (blocksp_solvers.hpp,
blocksp_preconditioners.hpp): ``bsmv``, ``pGMRES``, ``pCG``, and the
preconditioners ``NoPre``, ``Diag``, ``SOR``, ``ILU0``, ``IChol0`` whose
``__call__`` overwrites a device vector r with M^{-1} r.

All compute runs in the HIP library; loading fails loudly if it is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

__all__ = [
    "BspError", "lib", "Context", "DeviceArray", "BSR", "NoPre", "Diag", "SOR", "ILU0", "IChol0",
    "bsmv", "bsmv_dim", "bsmm_rhs", "pGMRES", "pCG", "Stats", "TOL_REFERENCE", "TOL_TRUE_RESIDUAL",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libblocksp_amd.so")

TOL_REFERENCE = 0
TOL_TRUE_RESIDUAL = 1

E_INVALID, E_HIP, E_NOMEM, E_UNSUPPORTED, E_NUMERIC, E_TIMEOUT = -1, -2, -3, -4, -5, -6
NOT_CONVERGED = 1

_ERRORS = {-1: "invalid argument", -2: "HIP error", -3: "out of memory", -4: "unsupported",
           -5: "numerical failure", -6: "sync-free wait timed out"}


class BspError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"blocksp_amd error {code} ({_ERRORS.get(code, '?')}): {msg}")
        self.code = code


class Stats(C.Structure):
    _fields_ = [("iters", C.c_int), ("restarts", C.c_int), ("converged", C.c_int),
                ("precond_applies", C.c_int), ("res0", C.c_double), ("res_est", C.c_double),
                ("true_relres", C.c_double), ("setup_ms", C.c_double), ("solve_ms", C.c_double),
                ("fallback", C.c_int), ("reorth", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None

_VP = C.c_void_p
_SIGS = {
    "bsp_ctx_create": [C.c_int, _VP, C.POINTER(_VP)],
    "bsp_ctx_destroy": [_VP],
    "bsp_ctx_sync": [_VP],
    "bsp_ctx_set_option": [_VP, C.c_char_p, C.c_int64],
    "bsp_malloc": [_VP, C.c_size_t, C.POINTER(_VP)],
    "bsp_free": [_VP, _VP],
    "bsp_memcpy_h2d": [_VP, _VP, _VP, C.c_size_t],
    "bsp_memcpy_d2h": [_VP, _VP, _VP, C.c_size_t],
    "bsp_memcpy_d2d": [_VP, _VP, _VP, C.c_size_t],
    "bsp_memset_zero": [_VP, _VP, C.c_size_t],
    "bsp_event_record": [_VP, C.c_int],
    "bsp_event_elapsed_ms": [_VP, C.c_int, C.c_int, C.POINTER(C.c_float)],
    "bsp_bsr_create": [_VP, C.c_int, C.c_int, C.c_int, _VP, _VP, _VP, C.POINTER(_VP)],
    "bsp_bsr_create_complex": [_VP, C.c_int, C.c_int, C.c_int, _VP, _VP, _VP, C.POINTER(_VP)],
    "bsp_bsr_complex_bs": [_VP, C.POINTER(C.c_int)],
    "bsp_bsr_download_complex": [_VP, _VP, _VP, _VP, _VP],
    "bsp_bsr_destroy": [_VP],
    "bsp_bsr_info": [_VP, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int64)],
    "bsp_bsr_download": [_VP, _VP, _VP, _VP, _VP],
    "bsp_bsr_set_values_dev": [_VP, _VP, _VP],
    "bsp_bsmv": [_VP, _VP, _VP, _VP, C.c_int],
    "bsp_bsmv_dim": [_VP, _VP, C.c_int, C.c_int, _VP, _VP, C.c_int],
    "bsp_bsmv_transpose": [_VP, _VP, _VP, _VP, C.c_int],
    "bsp_bsmm_rhs": [_VP, _VP, C.c_int, _VP, _VP],
    "bsp_bsmm": [_VP, _VP, _VP, C.POINTER(_VP)],
    "bsp_bsmm_trans_a": [_VP, _VP, _VP, C.POINTER(_VP)],
    "bsp_bsmm_numeric": [_VP, _VP, _VP, _VP],
    "bsp_pgmres_ritz": [_VP, _VP, _VP, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "bsp_bsr_save": [_VP, _VP, C.c_char_p],
    "bsp_orthogonalize": [_VP, C.c_int64, _VP, C.c_int, _VP, _VP, _VP],
    "bsp_orthogonalize_noalloc": [_VP, C.c_int64, _VP, C.c_int, _VP, _VP, _VP],
    "bsp_bsr_load": [_VP, C.c_char_p, C.POINTER(_VP)],
    "bsp_arnoldi_eigenvalues": [_VP, _VP, _VP, C.c_int, C.c_int, _VP, _VP, C.c_int, C.POINTER(C.c_int)],
    "bsp_bsmv_dim_transpose": [_VP, _VP, C.c_int, C.c_int, _VP, _VP, C.c_int],
    "bsp_dot": [_VP, C.c_int64, _VP, _VP, C.POINTER(C.c_double)],
    "bsp_nrm2": [_VP, C.c_int64, _VP, C.POINTER(C.c_double)],
    "bsp_infnorm": [_VP, C.c_int64, _VP, C.POINTER(C.c_double)],
    "bsp_axpy": [_VP, C.c_int64, C.c_double, _VP, _VP],
    "bsp_scal": [_VP, C.c_int64, C.c_double, _VP],
    "bsp_precond_create_none": [_VP, _VP, C.POINTER(_VP)],
    "bsp_precond_create_diag": [_VP, _VP, C.c_int, C.POINTER(_VP)],
    "bsp_precond_create_sor": [_VP, _VP, C.c_double, C.c_int, C.c_int, C.c_int, C.POINTER(_VP)],
    "bsp_precond_create_sor_ex": [_VP, _VP, C.c_double, C.c_int, C.c_int, C.c_int, _VP, C.POINTER(_VP)],
    "bsp_precond_create_ilu0": [_VP, _VP, C.POINTER(_VP)],
    "bsp_precond_create_ichol0": [_VP, _VP, C.POINTER(_VP)],
    "bsp_precond_create_ichol0_ex": [_VP, _VP, _VP, C.POINTER(_VP)],
    "bsp_precond_create_host": [_VP, _VP, _VP, _VP, C.POINTER(_VP)],
    "bsp_precond_destroy": [_VP],
    "bsp_precond_apply": [_VP, _VP, _VP],
    "bsp_precond_levels": [_VP, C.c_int, _VP, C.POINTER(C.c_int)],
    "bsp_precond_ell_width": [_VP, C.c_int],
    "bsp_precond_factor_info": [_VP, C.POINTER(C.c_int64)],
    "bsp_precond_download_factor": [_VP, _VP, _VP, _VP, _VP],
    "bsp_pgmres": [_VP, _VP, _VP, _VP, _VP, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(Stats)],
    "bsp_pcg": [_VP, _VP, _VP, _VP, _VP, C.c_double, C.c_int, C.c_int, C.POINTER(Stats)],
    "bsp_precond_create_ilu0_ex": [_VP, _VP, _VP, C.c_int, C.POINTER(_VP)],
    "bsp_fgmres_mixed": [_VP, _VP, _VP, _VP, _VP, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                         C.POINTER(Stats)],
    "bsp_fgmres_fallback": [_VP, _VP, _VP, _VP, _VP, _VP, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                            C.c_double, C.POINTER(Stats)],
    "bsp_mpnn_create": [_VP, C.c_int, C.c_uint64, C.POINTER(_VP)],
    "bsp_mpnn_destroy": [_VP],
    "bsp_mpnn_num_params": [_VP, C.POINTER(C.c_int64)],
    "bsp_mpnn_get_params": [_VP, _VP, _VP],
    "bsp_mpnn_set_params": [_VP, _VP, _VP],
    "bsp_mpnn_get_grads": [_VP, _VP, _VP],
    "bsp_mpnn_graph_create": [_VP, _VP, _VP, C.POINTER(_VP)],
    "bsp_mpnn_graph_destroy": [_VP],
    "bsp_mpnn_infer": [_VP, _VP, _VP, _VP, _VP],
    "bsp_mpnn_loss_grad": [_VP, _VP, _VP, _VP, C.c_int, C.POINTER(C.c_double)],
    "bsp_mpnn_adam_step": [_VP, _VP, C.c_float],
    "bsp_mpnn_train": [_VP, _VP, C.c_int, _VP, _VP, _VP, C.c_int, C.c_int, C.c_float, _VP],
    "bsp_precond_create_learned": [_VP, _VP, _VP, _VP, C.c_int, C.POINTER(_VP)],
    "bsp_ctx_record_basis": [_VP, _VP, C.c_int64, C.c_int, C.c_int],
    "bsp_ctx_recorded": [_VP, C.POINTER(C.c_int)],
    "bsp_batch_create": [_VP, _VP, C.c_int, _VP, C.c_int, C.POINTER(_VP)],
    "bsp_batch_create_ex": [_VP, _VP, C.c_int, _VP, C.c_int, _VP, C.POINTER(_VP)],
    "bsp_batch_destroy": [_VP],
    "bsp_batch_set_tiles": [_VP, _VP, C.c_int, C.c_int],
    "bsp_batch_autotune": [_VP, _VP, _VP, _VP, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "bsp_batch_autotune_solve": [_VP, _VP, _VP, _VP, C.c_double, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int),
                                 C.POINTER(C.c_int)],
    "bsp_batch_pcg": [_VP, _VP, _VP, _VP, C.c_double, C.c_int, C.c_int, _VP, _VP, C.POINTER(C.c_double)],
    "bsp_batch_info": [_VP, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "bsp_batch_ell_width": [_VP],
    "bsp_bsmm_rhs_bf16": [_VP, _VP, C.c_int, _VP, _VP],
    "bsp_rhs_pack_bf16": [_VP, C.c_int64, C.c_int, C.c_int, _VP, _VP],
    "bsp_rhs_unpack_f32": [_VP, C.c_int64, C.c_int, C.c_int, _VP, _VP],
}


def lib():
    """Load libblocksp_amd.so (build it with `python -m blocksp_amd.build`)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m blocksp_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.bsp_last_error.argtypes = [_VP]
        L.bsp_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def exported_symbols():
    return sorted(_SIGS) + ["bsp_last_error"]


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Context:
    """A device + stream (bsp_ctx)."""

    def __init__(self, device: int = 0, stream=None):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.bsp_ctx_create(device, stream, C.byref(h))
        if rc != 0:
            raise BspError(rc, f"bsp_ctx_create(device={device}) failed")
        self.h = h

    def check(self, rc, what=""):
        if rc < 0:
            raise BspError(rc, f"{what}: {self.L.bsp_last_error(self.h).decode()}")
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.L.bsp_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        self.check(self.L.bsp_ctx_sync(self.h), "sync")

    def set_option(self, name: str, value: int):
        self.check(self.L.bsp_ctx_set_option(self.h, name.encode(), int(value)), "set_option")

    def empty(self, shape, dtype=np.float64):
        return DeviceArray(self, shape, dtype)

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.shape, a.dtype)
        d.upload(a)
        return d

    def zeros(self, shape, dtype=np.float64):
        d = DeviceArray(self, shape, dtype)
        d.zero()
        return d

    def record(self, slot):
        self.check(self.L.bsp_event_record(self.h, slot), "event_record")

    def elapsed_ms(self, a, b):
        ms = C.c_float()
        self.check(self.L.bsp_event_elapsed_ms(self.h, a, b, C.byref(ms)), "event_elapsed")
        return ms.value


class DeviceArray:
    """Device memory owned by Python (bsp_malloc / bsp_free)."""

    def __init__(self, ctx: Context, shape, dtype=np.float64):
        self.ctx = ctx
        self.shape = tuple(np.atleast_1d(shape).tolist()) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        ctx.check(ctx.L.bsp_malloc(ctx.h, max(self.nbytes, 1), C.byref(p)), "malloc")
        self.ptr = p

    @property
    def size(self):
        return int(np.prod(self.shape))

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        self.ctx.check(self.ctx.L.bsp_memcpy_h2d(self.ctx.h, self.ptr, _ptr(a), self.nbytes), "h2d")

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        self.ctx.check(self.ctx.L.bsp_memcpy_d2h(self.ctx.h, _ptr(out), self.ptr, self.nbytes), "d2h")
        return out

    def zero(self):
        self.ctx.check(self.ctx.L.bsp_memset_zero(self.ctx.h, self.ptr, self.nbytes), "memset")

    def copy_from(self, other: "DeviceArray"):
        assert other.nbytes == self.nbytes
        self.ctx.check(self.ctx.L.bsp_memcpy_d2d(self.ctx.h, self.ptr, other.ptr, self.nbytes), "d2d")

    def __del__(self):
        try:
            if self.ptr and self.ctx.h:
                self.ctx.L.bsp_free(self.ctx.h, self.ptr)
                self.ptr = None
        except Exception:
            pass


class BSR:
    """BlockSp<T,B*D> on the device (bsp_bsr).  A complex matrix (BSR.from_complex) has
    cbs = its complex block size and bs = 2 cbs (the real-equivalent blocks the device holds);
    its vectors are complex128 device arrays of n_rows * cbs entries."""

    cbs = 0

    def __init__(self, ctx: Context, n_rows, n_cols, bs, rowptr, colidx, vals):
        self.ctx = ctx
        self.n_rows, self.n_cols, self.bs = int(n_rows), int(n_cols), int(bs)
        rp = np.ascontiguousarray(rowptr, np.int32)
        ci = np.ascontiguousarray(colidx, np.int32)
        v = np.ascontiguousarray(vals, np.float64)
        h = C.c_void_p()
        ctx.check(ctx.L.bsp_bsr_create(ctx.h, self.n_rows, self.n_cols, self.bs, _ptr(rp), _ptr(ci), _ptr(v),
                                       C.byref(h)), "bsr_create")
        self.h = h
        self.nnzb = int(ci.shape[0])

    @classmethod
    def from_problem(cls, ctx, p):
        return cls(ctx, p.n, p.n, p.bs, p.rowptr, p.colidx, p.vals)

    @classmethod
    def from_complex(cls, ctx, n_rows, n_cols, bs, rowptr, colidx, vals):
        """BlockSp<std::complex<double>, bs>: vals complex, (nnzb, bs, bs) row-major blocks in
        stored order (bsp_bsr_create_complex)."""
        rp = np.ascontiguousarray(rowptr, np.int32)
        ci = np.ascontiguousarray(colidx, np.int32)
        v = np.ascontiguousarray(vals, np.complex128)
        h = C.c_void_p()
        ctx.check(ctx.L.bsp_bsr_create_complex(ctx.h, int(n_rows), int(n_cols), int(bs), _ptr(rp), _ptr(ci), _ptr(v),
                                               C.byref(h)), "bsr_create_complex")
        return cls._wrap(ctx, h)

    @property
    def is_complex(self):
        return self.cbs > 0

    @classmethod
    def _wrap(cls, ctx, h, keep=()):
        """A BSR object around a handle the library created (e.g. by bsp_bsmm)."""
        self = cls.__new__(cls)
        self.ctx, self.h, self._keep = ctx, h, keep
        n, m, bs, nz = C.c_int(), C.c_int(), C.c_int(), C.c_int64()
        ctx.check(ctx.L.bsp_bsr_info(h, C.byref(n), C.byref(m), C.byref(bs), C.byref(nz)), "bsr_info")
        self.n_rows, self.n_cols, self.bs, self.nnzb = n.value, m.value, bs.value, nz.value
        cbs = C.c_int()
        ctx.check(ctx.L.bsp_bsr_complex_bs(h, C.byref(cbs)), "bsr_complex_bs")
        self.cbs = cbs.value
        return self

    def save(self, path):
        """Write the matrix to a BSR cache file (blocksp_amd.bsrio layout)."""
        self.ctx.check(self.ctx.L.bsp_bsr_save(self.ctx.h, self.h, str(path).encode()), "bsr_save")

    @classmethod
    def load(cls, ctx, path):
        h = C.c_void_p()
        ctx.check(ctx.L.bsp_bsr_load(ctx.h, str(path).encode(), C.byref(h)), "bsr_load")
        return cls._wrap(ctx, h)

    def download(self):
        rp = np.empty(self.n_rows + 1, np.int32)
        ci = np.empty(self.nnzb, np.int32)
        if self.cbs:
            v = np.empty((self.nnzb, self.cbs, self.cbs), np.complex128)
            self.ctx.check(self.ctx.L.bsp_bsr_download_complex(self.ctx.h, self.h, _ptr(rp), _ptr(ci), _ptr(v)),
                           "bsr_download_complex")
            return rp, ci, v
        v = np.empty((self.nnzb, self.bs, self.bs))
        self.ctx.check(self.ctx.L.bsp_bsr_download(self.ctx.h, self.h, _ptr(rp), _ptr(ci), _ptr(v)), "bsr_download")
        return rp, ci, v

    def __del__(self):
        try:
            if self.h:
                self.ctx.L.bsp_bsr_destroy(self.h)
                self.h = None
        except Exception:
            pass


def bsmv(ctx, A: BSR, x: DeviceArray, y: DeviceArray, accumulate=False):
    """y (+)= A x  (bsmv, blocksp_multiply.hpp:365-390)."""
    ctx.check(ctx.L.bsp_bsmv(ctx.h, A.h, x.ptr, y.ptr, int(accumulate)), "bsmv")


def bsmv_dim(ctx, A: BSR, D, dim, x: DeviceArray, y: DeviceArray, accumulate=False):
    ctx.check(ctx.L.bsp_bsmv_dim(ctx.h, A.h, int(D), int(dim), x.ptr, y.ptr, int(accumulate)), "bsmv_dim")


def bsmv_transpose(ctx, A: BSR, x: DeviceArray, y: DeviceArray, accumulate=False):
    """y (+)= A^T x  (bsmv(..., transpose=true), blocksp_multiply.hpp:196-204, 236-245)."""
    ctx.check(ctx.L.bsp_bsmv_transpose(ctx.h, A.h, x.ptr, y.ptr, int(accumulate)), "bsmv_transpose")


def bsmv_dim_transpose(ctx, A: BSR, D, dim, x: DeviceArray, y: DeviceArray, accumulate=False):
    ctx.check(ctx.L.bsp_bsmv_dim_transpose(ctx.h, A.h, int(D), int(dim), x.ptr, y.ptr, int(accumulate)),
              "bsmv_dim_transpose")


def bsmm(ctx, A: BSR, B: BSR, Atran=False) -> BSR:
    """C = A B, or A^T B with Atran, on the device (bsmm, blocksp_multiply.hpp:51-136);
    C keeps A and B alive."""
    h = C.c_void_p()
    if Atran:
        ctx.check(ctx.L.bsp_bsmm_trans_a(ctx.h, A.h, B.h, C.byref(h)), "bsmm_trans_a")
    else:
        ctx.check(ctx.L.bsp_bsmm(ctx.h, A.h, B.h, C.byref(h)), "bsmm")
    return BSR._wrap(ctx, h, keep=(A, B))


def bsmm_numeric(ctx, A: BSR, B: BSR, Cm: BSR):
    ctx.check(ctx.L.bsp_bsmm_numeric(ctx.h, A.h, B.h, Cm.h), "bsmm_numeric")


def bsmm_rhs(ctx, A: BSR, nrhs, X: DeviceArray, Y: DeviceArray):
    ctx.check(ctx.L.bsp_bsmm_rhs(ctx.h, A.h, int(nrhs), X.ptr, Y.ptr), "bsmm_rhs")


def rhs_pack_bf16(ctx, rows, bs, nrhs, X: DeviceArray, Xb: DeviceArray):
    """RHS-minor FP64 X -> block-major BF16 (uint16) Xb."""
    ctx.check(ctx.L.bsp_rhs_pack_bf16(ctx.h, int(rows), int(bs), int(nrhs), X.ptr, Xb.ptr), "rhs_pack_bf16")


def rhs_unpack_f32(ctx, rows, bs, nrhs, Yb: DeviceArray, Y: DeviceArray):
    ctx.check(ctx.L.bsp_rhs_unpack_f32(ctx.h, int(rows), int(bs), int(nrhs), Yb.ptr, Y.ptr), "rhs_unpack_f32")


def bsmm_rhs_bf16(ctx, A: BSR, nrhs, Xb: DeviceArray, Yb: DeviceArray):
    """Yb = A Xb on the matrix cores (BF16 operands, FP32 accumulation, block-major layouts)."""
    ctx.check(ctx.L.bsp_bsmm_rhs_bf16(ctx.h, A.h, int(nrhs), Xb.ptr, Yb.ptr), "bsmm_rhs_bf16")


class _Precond:
    kind = None

    def __call__(self, r: DeviceArray):
        self.ctx.check(self.ctx.L.bsp_precond_apply(self.ctx.h, self.h, r.ptr), "precond_apply")

    def levels(self, lower=True):
        n = self.A.n_rows
        lev = np.empty(n, np.int32)
        nl = C.c_int()
        self.ctx.check(self.ctx.L.bsp_precond_levels(self.h, int(lower), _ptr(lev), C.byref(nl)), "levels")
        return nl.value, lev

    def ell_width(self, lower=True):
        """Entries per slot of the sweep's ELL copy (0: the sweep reads the CSR factor)."""
        return int(self.ctx.L.bsp_precond_ell_width(self.h, int(lower)))

    def factor(self):
        nz = C.c_int64()
        self.ctx.check(self.ctx.L.bsp_precond_factor_info(self.h, C.byref(nz)), "factor_info")
        rp = np.empty(self.A.n_rows + 1, np.int32)
        ci = np.empty(nz.value, np.int32)
        v = np.empty((nz.value, self.A.bs, self.A.bs))
        self.ctx.check(self.ctx.L.bsp_precond_download_factor(self.ctx.h, self.h, _ptr(rp), _ptr(ci), _ptr(v)),
                       "download_factor")
        return rp, ci, v

    def __del__(self):
        try:
            if self.h:
                self.ctx.L.bsp_precond_destroy(self.h)
                self.h = None
        except Exception:
            pass


class NoPre(_Precond):
    """preConditioners::NoPre (blocksp_preconditioners.hpp:66-71)."""

    def __init__(self, ctx, A):
        self.ctx, self.A = ctx, A
        h = C.c_void_p()
        ctx.check(ctx.L.bsp_precond_create_none(ctx.h, A.h, C.byref(h)), "NoPre")
        self.h = h


class Diag(_Precond):
    """preConditioners::Diag<Inv> (blocksp_preconditioners.hpp:75-111); cholesky=True for CholeskyFactor."""

    def __init__(self, ctx, A, cholesky=False):
        self.ctx, self.A = ctx, A
        h = C.c_void_p()
        ctx.check(ctx.L.bsp_precond_create_diag(ctx.h, A.h, int(cholesky), C.byref(h)), "Diag")
        self.h = h


class SOR(_Precond):
    """preConditioners::SOR<Inv>(A, w, nu, type) (blocksp_preconditioners.hpp:116-240); with
    `colour`, multicolour SOR (rows swept colour by colour, natural order within a colour)."""

    def __init__(self, ctx, A, w, nu, type=0, cholesky=False, colour=None):
        self.ctx, self.A = ctx, A
        h = C.c_void_p()
        if colour is None:
            ctx.check(ctx.L.bsp_precond_create_sor(ctx.h, A.h, float(w), int(nu), int(type), int(cholesky),
                                                   C.byref(h)), "SOR")
        else:
            col = np.ascontiguousarray(colour, dtype=np.int32)
            assert col.shape == (A.n_rows,)
            self._colour = col
            ctx.check(ctx.L.bsp_precond_create_sor_ex(ctx.h, A.h, float(w), int(nu), int(type), int(cholesky),
                                                      _ptr(col), C.byref(h)), "SOR")
        self.h = h


class ILU0(_Precond):
    """preConditioners::ILU0(A) (blocksp_preconditioners.hpp:246-357)."""

    def __init__(self, ctx, A):
        self.ctx, self.A = ctx, A
        h = C.c_void_p()
        ctx.check(ctx.L.bsp_precond_create_ilu0(ctx.h, A.h, C.byref(h)), "ILU0")
        self.h = h


PREC_F64 = 0
PREC_F32 = 1
PREC_BF16 = 2   # ILU0 factor stored in BF16, applied in FP32


class ILU0Ex(_Precond):
    """ILU0 with an elimination order given by a colouring and FP64/FP32 storage
    (bsp_precond_create_ilu0_ex).  colour=None is the reference's natural order."""

    def __init__(self, ctx, A, colour=None, precision=PREC_F64):
        self.ctx, self.A = ctx, A
        h = C.c_void_p()
        col = None if colour is None else np.ascontiguousarray(colour, np.int32)
        self._colour = col
        ctx.check(ctx.L.bsp_precond_create_ilu0_ex(ctx.h, A.h, None if col is None else _ptr(col), int(precision),
                                                   C.byref(h)), "ILU0Ex")
        self.h = h


_HOST_FN = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.c_int64, C.c_void_p)


class HostPrecond(_Precond):
    """User preconditioner on the host (bsp_precond_create_host): fn(r) overwrites the
    numpy view r with M^{-1} r, as a reference PreCond callable does (tests.hpp:113)."""

    def __init__(self, ctx, A, fn):
        self.ctx, self.A = ctx, A

        def tramp(ptr, length, user):
            fn(np.ctypeslib.as_array(ptr, shape=(length,)))

        self._cb = _HOST_FN(tramp)
        h = C.c_void_p()
        ctx.check(ctx.L.bsp_precond_create_host(ctx.h, A.h, C.cast(self._cb, C.c_void_p), None, C.byref(h)),
                  "HostPrecond")
        self.h = h


class IChol0(_Precond):
    """preConditioners::IChol0(A) (blocksp_preconditioners.hpp:364-538); with `colour`, the
    IChol0 of A eliminated colour by colour (natural order within a colour)."""

    def __init__(self, ctx, A, colour=None):
        self.ctx, self.A = ctx, A
        h = C.c_void_p()
        if colour is None:
            ctx.check(ctx.L.bsp_precond_create_ichol0(ctx.h, A.h, C.byref(h)), "IChol0")
        else:
            col = np.ascontiguousarray(colour, dtype=np.int32)
            assert col.shape == (A.n_rows,)
            ctx.check(ctx.L.bsp_precond_create_ichol0_ex(ctx.h, A.h, _ptr(col), C.byref(h)), "IChol0")
        self.h = h


def pGMRES(ctx, A: BSR, b: DeviceArray, tol, M, maxIters=500, restart=-1, k=-1, tol_mode=TOL_REFERENCE, x=None):
    """pGMRES (blocksp_solvers.hpp:156-295). Returns (x DeviceArray, Stats, return code).
    For a complex A, b and x are complex128 arrays."""
    x = x if x is not None else ctx.empty(b.shape, b.dtype)
    st = Stats()
    rc = ctx.L.bsp_pgmres(ctx.h, A.h, M.h, b.ptr, x.ptr, float(tol), int(maxIters), int(restart), int(k),
                          int(tol_mode), C.byref(st))
    ctx.check(rc, "pgmres")
    return x, st, rc


def FGMRES_mixed(ctx, A: BSR, b: DeviceArray, tol, M, maxIters=20000, restart=30, k=10, spmv_precision=PREC_F32,
                 inner_reduction=1e-6, x=None, fallback=None):
    """Mixed-precision flexible GMRES-DR with FP64 refinement (bsp_fgmres_mixed, or
    bsp_fgmres_fallback when a fallback preconditioner is given)."""
    x = x if x is not None else ctx.empty(b.shape)
    st = Stats()
    if fallback is None:
        rc = ctx.L.bsp_fgmres_mixed(ctx.h, A.h, M.h, b.ptr, x.ptr, float(tol), int(maxIters), int(restart), int(k),
                                    int(spmv_precision), float(inner_reduction), C.byref(st))
    else:
        rc = ctx.L.bsp_fgmres_fallback(ctx.h, A.h, M.h, fallback.h, b.ptr, x.ptr, float(tol), int(maxIters),
                                       int(restart), int(k), int(spmv_precision), float(inner_reduction), C.byref(st))
    ctx.check(rc, "fgmres_mixed")
    return x, st, rc


class MPNN:
    """Learned ILU0 correction network (bsp_mpnn_*): 3 message-passing layers, hidden 64."""

    def __init__(self, ctx, bs=8, seed=0):
        self.ctx, self.bs = ctx, int(bs)
        h = C.c_void_p()
        ctx.check(ctx.L.bsp_mpnn_create(ctx.h, self.bs, C.c_uint64(seed), C.byref(h)), "mpnn_create")
        self.h = h
        cnt = C.c_int64()
        ctx.check(ctx.L.bsp_mpnn_num_params(self.h, C.byref(cnt)), "mpnn_num_params")
        self.nparams = cnt.value

    def params(self):
        out = np.empty(self.nparams, np.float32)
        self.ctx.check(self.ctx.L.bsp_mpnn_get_params(self.ctx.h, self.h, _ptr(out)), "mpnn_get_params")
        return out

    def set_params(self, theta):
        th = np.ascontiguousarray(theta, np.float32)
        assert th.size == self.nparams
        self.ctx.check(self.ctx.L.bsp_mpnn_set_params(self.ctx.h, self.h, _ptr(th)), "mpnn_set_params")

    def grads(self):
        out = np.empty(self.nparams, np.float32)
        self.ctx.check(self.ctx.L.bsp_mpnn_get_grads(self.ctx.h, self.h, _ptr(out)), "mpnn_get_grads")
        return out

    def infer(self, graph):
        n, E = graph.n, graph.E
        dD = np.empty((n, self.bs * self.bs), np.float32)
        g = np.empty(max(E, 1), np.float32)
        self.ctx.check(self.ctx.L.bsp_mpnn_infer(self.ctx.h, self.h, graph.h, _ptr(dD), _ptr(g)), "mpnn_infer")
        return dD, g[:E]

    def loss_grad(self, graph, Z: DeviceArray, K):
        L = C.c_double()
        self.ctx.check(self.ctx.L.bsp_mpnn_loss_grad(self.ctx.h, self.h, graph.h, Z.ptr, int(K), C.byref(L)),
                       "mpnn_loss_grad")
        return L.value

    def adam_step(self, lr):
        self.ctx.check(self.ctx.L.bsp_mpnn_adam_step(self.ctx.h, self.h, float(lr)), "mpnn_adam_step")

    def train(self, graphs, Zs, Ks, epochs, batch=16, lr=1e-3):
        ng = len(graphs)
        gh = (C.c_void_p * ng)(*[g.h.value for g in graphs])
        zp = (C.c_void_p * ng)(*[z.ptr.value for z in Zs])
        ks = (C.c_int * ng)(*[int(k) for k in Ks])
        out = np.zeros(max(int(epochs), 1))
        self.ctx.check(self.ctx.L.bsp_mpnn_train(self.ctx.h, self.h, ng, gh, zp, ks, int(epochs), int(batch),
                                                 float(lr), _ptr(out)), "mpnn_train")
        return out[:int(epochs)]

    def __del__(self):
        try:
            if self.h:
                self.ctx.L.bsp_mpnn_destroy(self.h)
                self.h = None
        except Exception:
            pass


class MPNNGraph:
    """Block graph of A with the ILU0 factor of M to correct (bsp_mpnn_graph_create)."""

    def __init__(self, ctx, A, M):
        self.ctx, self.A, self.M = ctx, A, M
        h = C.c_void_p()
        ctx.check(ctx.L.bsp_mpnn_graph_create(ctx.h, A.h, M.h, C.byref(h)), "mpnn_graph_create")
        self.h = h
        self.n = A.n_rows
        self.E = int(A.nnzb) - A.n_rows

    def __del__(self):
        try:
            if self.h:
                self.ctx.L.bsp_mpnn_graph_destroy(self.h)
                self.h = None
        except Exception:
            pass


class LearnedILU0(_Precond):
    """ILU0 (elimination order from `colour`) corrected by a trained MPNN (bsp_precond_create_learned)."""

    def __init__(self, ctx, A, net: MPNN, colour=None, precision=PREC_F64):
        self.ctx, self.A = ctx, A
        h = C.c_void_p()
        col = None if colour is None else np.ascontiguousarray(colour, np.int32)
        self._colour = col
        ctx.check(ctx.L.bsp_precond_create_learned(ctx.h, A.h, net.h, None if col is None else _ptr(col),
                                                   int(precision), C.byref(h)), "LearnedILU0")
        self.h = h


def record_basis(ctx, buf: DeviceArray, length, cap, every=1):
    """Record Krylov basis vectors of the following pGMRES solves into buf (cap x length doubles)."""
    ctx.check(ctx.L.bsp_ctx_record_basis(ctx.h, None if buf is None else buf.ptr, int(length), int(cap),
                                         int(every)), "record_basis")


def recorded(ctx):
    c = C.c_int()
    ctx.check(ctx.L.bsp_ctx_recorded(ctx.h, C.byref(c)), "recorded")
    return c.value


def orthogonalize(ctx, Q, w: DeviceArray, qnew: DeviceArray):
    """Classical Gram-Schmidt step with conditional reorthogonalisation (bsp_orthogonalize).
    Q: list of orthonormal DeviceArrays.  Returns the m + 1 coefficients (last: ||w||)."""
    m = len(Q)
    ptrs = (C.c_void_p * m)(*[q.ptr.value for q in Q])
    h = np.zeros(m + 1)
    ctx.check(ctx.L.bsp_orthogonalize(ctx.h, int(w.size), ptrs, m, w.ptr, _ptr(h), qnew.ptr), "orthogonalize")
    return h


def orthogonalize_noalloc(ctx, Q, w: DeviceArray, h: DeviceArray, qnew: DeviceArray):
    """bsp_orthogonalize_noalloc: the same step on the library's static scratch, no allocation and
    no synchronisation; h (device, m + 2 doubles) receives the coefficients, ||w|| and the
    second-pass indicator."""
    m = len(Q)
    ptrs = (C.c_void_p * m)(*[q.ptr.value for q in Q])
    ctx.check(ctx.L.bsp_orthogonalize_noalloc(ctx.h, int(w.size), ptrs, m, w.ptr, h.ptr, qnew.ptr),
              "orthogonalize_noalloc")


def pgmres_ritz(ctx, cap=256):
    """(Harmonic) Ritz values of the last pGMRES run on ctx: (complex array, harmonic flag)."""
    re, im = np.zeros(cap), np.zeros(cap)
    cnt, harm = C.c_int(), C.c_int()
    ctx.check(ctx.L.bsp_pgmres_ritz(ctx.h, _ptr(re), _ptr(im), cap, C.byref(cnt), C.byref(harm)), "pgmres_ritz")
    k = min(cnt.value, cap)
    return re[:k] + 1j * im[:k], bool(harm.value)


def arnoldi_eigenvalues(ctx, A: BSR, iters=30, M=None, harmonic=False):
    """arnoldi_eigenvalues (blocksp_solvers.hpp:298-333) on the device: Ritz values as a complex array."""
    re, im = np.zeros(iters), np.zeros(iters)
    cnt = C.c_int()
    ctx.check(ctx.L.bsp_arnoldi_eigenvalues(ctx.h, A.h, None if M is None else M.h, int(iters), int(harmonic),
                                            _ptr(re), _ptr(im), int(iters), C.byref(cnt)), "arnoldi_eigenvalues")
    return re[:cnt.value] + 1j * im[:cnt.value]


def pCG(ctx, A: BSR, b: DeviceArray, tol, M, maxIters=1000, tol_mode=TOL_REFERENCE, x=None):
    """pCG (blocksp_solvers.hpp:51-105). Returns (x DeviceArray, Stats, return code)."""
    x = x if x is not None else ctx.empty(b.shape)
    st = Stats()
    rc = ctx.L.bsp_pcg(ctx.h, A.h, M.h, b.ptr, x.ptr, float(tol), int(maxIters), int(tol_mode), C.byref(st))
    ctx.check(rc, "pcg")
    return x, st, rc


class Batch:
    """Independent systems (diagonal blocks of the block-diagonal A) solved by IChol0-pCG in
    one persistent launch (bsp_batch_*).  sys_rowptr: first block row of each system, plus n."""

    def __init__(self, ctx, A: BSR, sys_rowptr, nrhs, colour=None):
        """colour: optional colour per block row (each system's IChol0 eliminated colour by colour)."""
        self.ctx, self.A, self.nrhs = ctx, A, int(nrhs)
        self.sys_rowptr = np.ascontiguousarray(sys_rowptr, np.int32)
        self.nsys = len(self.sys_rowptr) - 1
        h = C.c_void_p()
        if colour is None:
            ctx.check(ctx.L.bsp_batch_create(ctx.h, A.h, self.nsys, _ptr(self.sys_rowptr), self.nrhs, C.byref(h)),
                      "batch_create")
        else:
            col = np.ascontiguousarray(colour, np.int32)
            assert col.shape == (A.n_rows,)
            ctx.check(ctx.L.bsp_batch_create_ex(ctx.h, A.h, self.nsys, _ptr(self.sys_rowptr), self.nrhs, _ptr(col),
                                                C.byref(h)), "batch_create")
        self.h = h

    def set_tiles(self, wg, rhs_per_item):
        self.ctx.check(self.ctx.L.bsp_batch_set_tiles(self.ctx.h, self.h, int(wg), int(rhs_per_item)), "set_tiles")

    def autotune(self, B: DeviceArray, X: DeviceArray, probe_iters=8):
        wg, r = C.c_int(), C.c_int()
        self.ctx.check(self.ctx.L.bsp_batch_autotune(self.ctx.h, self.h, B.ptr, X.ptr, int(probe_iters), C.byref(wg),
                                                     C.byref(r)), "autotune")
        return wg.value, r.value

    def autotune_solve(self, B: DeviceArray, X: DeviceArray, tol, maxIters=5000, tol_mode=TOL_REFERENCE,
                       probe_iters=8):
        """Short probes, then complete solves of the finalists; returns the kept (wg, rhs_per_item)."""
        wg, r = C.c_int(), C.c_int()
        self.ctx.check(self.ctx.L.bsp_batch_autotune_solve(self.ctx.h, self.h, B.ptr, X.ptr, float(tol), int(maxIters),
                                                           int(tol_mode), int(probe_iters), C.byref(wg), C.byref(r)),
                       "autotune_solve")
        return wg.value, r.value

    def info(self):
        wg, r, ni = C.c_int(), C.c_int(), C.c_int()
        self.ctx.check(self.ctx.L.bsp_batch_info(self.h, C.byref(wg), C.byref(r), C.byref(ni)), "batch_info")
        return wg.value, r.value, ni.value

    def ell_width(self):
        """Entries per slot of the ELL sweep copies (0: the sweeps read the CSR factor)."""
        return int(self.ctx.L.bsp_batch_ell_width(self.h))

    def pcg(self, B: DeviceArray, X: DeviceArray, tol, maxIters=5000, tol_mode=TOL_REFERENCE):
        its = np.zeros(self.nsys * self.nrhs, np.int32)
        rel = np.zeros(self.nsys * self.nrhs)
        ms = C.c_double()
        rc = self.ctx.L.bsp_batch_pcg(self.ctx.h, self.h, B.ptr, X.ptr, float(tol), int(maxIters), int(tol_mode),
                                      _ptr(its), _ptr(rel), C.byref(ms))
        self.ctx.check(rc, "batch_pcg")
        return its.reshape(self.nsys, self.nrhs), rel.reshape(self.nsys, self.nrhs), ms.value, rc

    def __del__(self):
        try:
            if self.h:
                self.ctx.L.bsp_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass


def block_diagonal(problems):
    """Concatenate independent problems into one block-diagonal (rowptr, colidx, vals, sys_rowptr)."""
    rps, cis, vs, sys = [], [], [], [0]
    off_r, off_nz = 0, 0
    for p in problems:
        rps.append(np.asarray(p.rowptr[:-1], np.int64) + off_nz)
        cis.append(np.asarray(p.colidx, np.int64) + off_r)
        vs.append(p.vals)
        off_r += p.n
        off_nz += len(p.colidx)
        sys.append(off_r)
    rowptr = np.concatenate(rps + [np.array([off_nz])]).astype(np.int32)
    return rowptr, np.concatenate(cis).astype(np.int32), np.concatenate(vs), np.array(sys, np.int32)
