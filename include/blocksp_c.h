/*
 * blocksp_c.h -- C ABI of the MI355X block-sparse solve library (libblocksp_amd.so).
 *
 * This is the drop-in boundary for BlockSp's solve path.  The reference is a
 * header-only C++20 template library with no FFI of its own; each entry point
 * below names the reference function or type it replaces (file:line, relative
 * to the BlockSp v1.0 checkout).  The C++20 headers in include/blocksp/ mirror
 * the reference's template surface on top of these calls.
 *
 * Conventions
 *  - Every function returns int: 0 (BSP_OK) on success, < 0 on error, and
 *    BSP_NOT_CONVERGED (1) when a solver stops at maxIters.  The message of the
 *    last error on a context is returned by bsp_last_error().
 *  - Pointers named *_dev are device memory owned by the caller (allocate with
 *    bsp_malloc or any HIP allocator on the context's device).  Host arrays are
 *    named *_host.  Handles are owned by the library until destroyed.
 *  - A context is bound to one device and one HIP stream and is not
 *    thread-safe; use one context per host thread.  All work is ordered on the
 *    context's stream; functions that return host results synchronise it.
 *  - BSR layout: rowptr[n_rows+1], colidx[nnzb] (int32) in the caller's per-row
 *    order, which is kept (the reference keeps insertion order,
 *    blocksp_containers.hpp:147-164); values are nnzb blocks of bs*bs doubles,
 *    row-major within a block.  Vectors use the solArray flat layout
 *    (ie, dim, ib) -> ie*B*D + dim*B + ib (blocksp_containers.hpp:572-736).
 *  - No exceptions cross this boundary.
 */
#ifndef BLOCKSP_C_H
#define BLOCKSP_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BSP_OK 0
#define BSP_NOT_CONVERGED 1
#define BSP_E_INVALID (-1)
#define BSP_E_HIP (-2)
#define BSP_E_NOMEM (-3)
#define BSP_E_UNSUPPORTED (-4)
#define BSP_E_NUMERIC (-5)
#define BSP_E_TIMEOUT (-6)

typedef struct bsp_ctx bsp_ctx;
typedef struct bsp_bsr bsp_bsr;
typedef struct bsp_precond bsp_precond;

/* ---- context and memory (no reference counterpart: device plumbing) ---- */

/* Create a context on `device`.  `stream` is a hipStream_t to borrow, or NULL to
 * let the library create its own (default priority). */
int bsp_ctx_create(int device, void* stream, bsp_ctx** out);
int bsp_ctx_destroy(bsp_ctx* ctx);
const char* bsp_last_error(const bsp_ctx* ctx);
int bsp_ctx_sync(bsp_ctx* ctx);
/* Set a tuning option by name: "trsv" (0 level launches, default / 1 single-launch
 * sync-free sweeps), "sweep_graphs" (1: replay many-level sweeps from HIP graphs),
 * "sweep_wg_per_cu", "spin_limit" (polls before a sync-free wait gives up), "verbose",
 * "bsmv_layout" (2: solver products for even block sizes up to 8 read an ELL copy of chunk-major
 * values when no row holds more than 8 blocks, else as 1 / 1: a chunk-major copy of the values /
 * 0: the values as stored; the same results), "fgmres_ortho" (0 classical / 1 modified
 * Gram-Schmidt), "cgs_pythagoras" (1: classical Gram-Schmidt of an FP64 basis takes the new norm
 * as ||w||^2 - ||h||^2 and normalises in the update pass / 0: sums it after the update),
 * "fgmres_basis_f32", "sweep_ell" (1: sweeps read their ELL copy where
 * built / 0: CSR), "sweep_onewg_max" (a sweep of more than 8 levels, each at most 1024 / bs
 * rows, runs in one workgroup when n * bs is at most this; default 4194304, 0: never),
 * "sweep_ahead" (1: such a sweep over an ELL copy, blocks up to 4 x 4, loads each level's
 * operands that the sweep does not write before the previous level's barrier / 0: after it),
 * "sweep_lds" (1: such a sweep keeps the solution vector in LDS when it fits, at most 64 KiB /
 * 0: reads it from global memory), "mgs_onewg" (1: the modified Gram-Schmidt of an Arnoldi step
 * on at most 8192 entries and 64 basis vectors runs in one single-workgroup kernel / 0: one
 * launch per step; bitwise the same),
 * "batch_ell" (1 / 0: ELL or CSR sweeps for batches created afterwards), "bf16_ell" (1: the
 * BF16 MFMA SpMM reads an ELL copy of the BF16 values when no row holds more than 8 blocks /
 * 0: a copy in CSR order; the same results).
 * Unknown names and out-of-range values return BSP_E_INVALID. */
int bsp_ctx_set_option(bsp_ctx* ctx, const char* name, int64_t value);

int bsp_malloc(bsp_ctx* ctx, size_t bytes, void** out_dev);
int bsp_free(bsp_ctx* ctx, void* ptr_dev);
int bsp_memcpy_h2d(bsp_ctx* ctx, void* dst_dev, const void* src_host, size_t bytes);
int bsp_memcpy_d2h(bsp_ctx* ctx, void* dst_host, const void* src_dev, size_t bytes);
int bsp_memcpy_d2d(bsp_ctx* ctx, void* dst_dev, const void* src_dev, size_t bytes);
int bsp_memset_zero(bsp_ctx* ctx, void* dst_dev, size_t bytes);

/* ---- timing on the context stream (HIP events) ---- */
int bsp_event_record(bsp_ctx* ctx, int slot);                  /* slot in [0, 16) */
int bsp_event_elapsed_ms(bsp_ctx* ctx, int slot_a, int slot_b, float* ms);

/* ---- matrices: BlockSp<T,B*D> = Sp::CSR<TinyMatrix> (blocksp_containers.hpp:52-53, 92-496) ---- */

int bsp_bsr_create(bsp_ctx* ctx, int n_rows, int n_cols, int bs, const int32_t* rowptr_host,
                   const int32_t* colidx_host, const double* vals_host, bsp_bsr** out);
int bsp_bsr_destroy(bsp_bsr* A);
int bsp_bsr_info(const bsp_bsr* A, int* n_rows, int* n_cols, int* bs, int64_t* nnzb);
/* Copy structure and values back to host buffers sized by bsp_bsr_info. */
int bsp_bsr_download(bsp_ctx* ctx, const bsp_bsr* A, int32_t* rowptr_host, int32_t* colidx_host,
                     double* vals_host);
/* Replace the values (same structure) from a device buffer. */
int bsp_bsr_set_values_dev(bsp_ctx* ctx, bsp_bsr* A, const double* vals_dev);

/* ---- complex matrices: BlockSp<std::complex<double>, B> (blocksp_containers.hpp:59-83) ----
 * vals_host: nnzb * bs * bs complex numbers as (re, im) pairs -- the memory of
 * std::complex<double> -- in row-major blocks, stored order.  The device holds each complex
 * entry a + ib as the real 2 x 2 block [[a, -b], [b, a]], so the matrix is a real BSR of block
 * size 2 bs (bsp_bsr_info reports 2 bs) and a complex vector is its interleaved (re, im)
 * memory, 2 n_rows bs doubles.  On such a matrix: bsp_bsmv / bsp_bsmv_dim give the complex
 * product; bsp_bsmv_transpose and bsp_bsmm_trans_a the conjugate-transposed ones (the
 * reference's transposed products for complex T, dense_multiply.hpp:80-89, 155-175); bsp_bsmm
 * of two complex matrices is complex; the NoPre, Diag and SOR (LU block inverses) and ILU0
 * (FP64, any elimination order) preconditioners and user host preconditioners (on the
 * interleaved vector) apply to complex vectors; bsp_pgmres runs the complex pGMRES (complex
 * inner products, least squares and harmonic-Ritz restart; solvers.hpp:109-295 with T
 * complex) and bsp_pgmres_ritz reports its Ritz values.  Cholesky block inverses, IChol0,
 * pCG (real only in the reference's solvers, blocksp_solvers.hpp:411, 485), the FP32/BF16
 * factors and the solvers and products with no reference counterpart return BSP_E_UNSUPPORTED.
 * bsp_bsr_save records the complex block size in the cache header's fourth word. */
int bsp_bsr_create_complex(bsp_ctx* ctx, int n_rows, int n_cols, int bs, const int32_t* rowptr_host,
                           const int32_t* colidx_host, const double* vals_host, bsp_bsr** out);
/* The complex block size of A (0 for a real matrix). */
int bsp_bsr_complex_bs(const bsp_bsr* A, int* cbs);
/* As bsp_bsr_download, the values as complex blocks (vals_host: nnzb * cbs * cbs (re, im) pairs). */
int bsp_bsr_download_complex(bsp_ctx* ctx, const bsp_bsr* A, int32_t* rowptr_host, int32_t* colidx_host,
                             double* vals_host);

/* ---- products ---- */

/* y (+)= A x.  Replaces bsmv(A,x) / bsmv(b,A,x) for BlockSp<T,B*D> (blocksp_multiply.hpp:365-390,
 * implementation 170-325).  accumulate=0 overwrites y (returning form), 1 adds (void form). */
int bsp_bsmv(bsp_ctx* ctx, const bsp_bsr* A, const double* x_dev, double* y_dev, int accumulate);
/* Per-dimension product for BlockSp<T,B> times solArray<T,B,D> (blocksp_multiply.hpp:170-193,
 * 252-285); A has bs = B, vectors have n*B*D entries; dim < 0 multiplies all dimensions. */
int bsp_bsmv_dim(bsp_ctx* ctx, const bsp_bsr* A, int D, int dim, const double* x_dev, double* y_dev,
                 int accumulate);
/* y (+)= A^T x (transpose=true in blocksp_multiply.hpp:196-204, 236-245); A^T is built once
 * per matrix (rows of A^T in ascending source row) and reused. */
int bsp_bsmv_transpose(bsp_ctx* ctx, const bsp_bsr* A, const double* x_dev, double* y_dev,
                       int accumulate);
/* Per-dimension transposed product: bsmv(A, x, dim, transpose=true) for BlockSp<T,B> and
 * solArray<T,B,D> (blocksp_multiply.hpp:365-380). */
int bsp_bsmv_dim_transpose(bsp_ctx* ctx, const bsp_bsr* A, int D, int dim, const double* x_dev, double* y_dev,
                           int accumulate);
/* BSR binary cache files (no reference counterpart; the reference reads .mtx text): magic
 * "BSPBSR1\0", int32 n_rows, n_cols, bs, 0, int64 nnzb, int32 rowptr[n_rows+1], int32
 * colidx[nnzb] (stored order), float64 values[nnzb*bs*bs] (row-major blocks), little-endian. */
int bsp_bsr_save(bsp_ctx* ctx, const bsp_bsr* A, const char* path);
int bsp_bsr_load(bsp_ctx* ctx, const char* path, bsp_bsr** out);
/* Multi-RHS product Y = A X with nrhs right-hand sides stored RHS-minor
 * (X[(row*bs + r)*nrhs + k]); no single reference call, it is nrhs bsmv calls. */
int bsp_bsmm_rhs(bsp_ctx* ctx, const bsp_bsr* A, int nrhs, const double* X_dev, double* Y_dev);
/* C = A * B (bsmm, blocksp_multiply.hpp:51-136): structure in the reference's first-touch
 * order (bit-exact), both passes on the device (8 x 8 blocks: numeric pass on the FP64
 * matrix cores).  A and B must outlive C for bsp_bsmm_numeric. */
int bsp_bsmm(bsp_ctx* ctx, const bsp_bsr* A, const bsp_bsr* B, bsp_bsr** C);
/* C = A^T * B (bsmm with Atran = true, blocksp_multiply.hpp:99-134): row j of C in the
 * order the reference inserts it (rows i of A holding column j ascending, then row i of B in
 * its stored order); A^T is built once from A's values and rebuilt when they change. */
int bsp_bsmm_trans_a(bsp_ctx* ctx, const bsp_bsr* A, const bsp_bsr* B, bsp_bsr** C);
/* Recomputes C = A * B (or A^T * B) on the device for new values of A and B (same structures). */
int bsp_bsmm_numeric(bsp_ctx* ctx, const bsp_bsr* A, const bsp_bsr* B, bsp_bsr* C);
/* BF16 block SpMM on the matrix cores (no reference counterpart): Yb = A Xb for nrhs
 * right-hand sides (a multiple of 16), 8 x 8 blocks only, BF16 values (rounded to nearest
 * even from A's FP64 values on first use) and BF16 right-hand sides, FP32 accumulation
 * and results.  Block-major layouts: Xb[(col * nrhs + k) * 8 + c] (uint16 BF16 bits),
 * Yb[(row * nrhs + k) * 8 + r] (float). */
int bsp_bsmm_rhs_bf16(bsp_ctx* ctx, const bsp_bsr* A, int nrhs, const uint16_t* Xb_dev, float* Yb_dev);
/* RHS-minor FP64 X[(row * bs + c) * nrhs + k] (the bsp_bsmm_rhs layout) -> block-major BF16. */
int bsp_rhs_pack_bf16(bsp_ctx* ctx, int64_t rows, int bs, int nrhs, const double* X_dev, uint16_t* Xb_dev);
/* Block-major FP32 -> RHS-minor FP64. */
int bsp_rhs_unpack_f32(bsp_ctx* ctx, int64_t rows, int bs, int nrhs, const float* Yb_dev, double* Y_dev);

/* ---- BLAS-1 on solArrays (blocksp_containers.hpp:858-864, 917-1012) ---- */
int bsp_dot(bsp_ctx* ctx, int64_t len, const double* x_dev, const double* y_dev, double* out_host);
int bsp_nrm2(bsp_ctx* ctx, int64_t len, const double* x_dev, double* out_host);
int bsp_infnorm(bsp_ctx* ctx, int64_t len, const double* x_dev, double* out_host);
int bsp_axpy(bsp_ctx* ctx, int64_t len, double a, const double* x_dev, double* y_dev);
int bsp_scal(bsp_ctx* ctx, int64_t len, double a, double* x_dev);

/* ---- preconditioners (blocksp_preconditioners.hpp) ---- */
/* The PreCond concept of the reference is `void operator()(solArray&) const`,
 * overwriting r with M^{-1} r (preconditioners.hpp:70, 106, 227-234, 352-355, 492-495). */

#define BSP_PRE_NONE 0
#define BSP_PRE_DIAG 1
#define BSP_PRE_SOR 2
#define BSP_PRE_ILU0 3
#define BSP_PRE_ICHOL0 4
#define BSP_PRE_HOST 5

/* NoPre (preconditioners.hpp:66-71). */
int bsp_precond_create_none(bsp_ctx* ctx, const bsp_bsr* A, bsp_precond** out);
/* Diag<Inv> (preconditioners.hpp:75-111); inv_kind 0 = LUFactor, 1 = CholeskyFactor. */
int bsp_precond_create_diag(bsp_ctx* ctx, const bsp_bsr* A, int inv_kind, bsp_precond** out);
/* SOR<Inv>(A, w, nu, type) (preconditioners.hpp:116-240).  A must outlive the preconditioner
 * (the reference keeps a non-owning pointer too, preconditioners.hpp:120). */
int bsp_precond_create_sor(bsp_ctx* ctx, const bsp_bsr* A, double w, int nu, int type, int inv_kind,
                           bsp_precond** out);
/* Multicolour block SOR (no reference counterpart): the SOR above with the rows swept colour by
 * colour (natural order within a colour; `colour`: one non-negative colour per block row), i.e.
 * the reference's SOR of P A P^T applied in A's numbering; sweeps of (number of colours) levels.
 * NULL = natural order (bsp_precond_create_sor). */
int bsp_precond_create_sor_ex(bsp_ctx* ctx, const bsp_bsr* A, double w, int nu, int type, int inv_kind,
                              const int32_t* colour, bsp_precond** out);
/* ILU0(A) (preconditioners.hpp:246-357): factor + diagonal-block inverses. */
int bsp_precond_create_ilu0(bsp_ctx* ctx, const bsp_bsr* A, bsp_precond** out);
/* ILU0 with a chosen elimination order and storage precision (no reference
 * counterpart: the reference eliminates in natural order in the matrix's type).
 * colour (nullable, n_rows entries >= 0): rows are eliminated colour by colour, in
 * natural order within a colour, i.e. ILU0 of P A P^T for that permutation; with a
 * colouring in which no two rows of one colour are coupled, each triangular sweep
 * has at most (number of colours) levels.  NULL = natural order (the reference's).
 * precision: BSP_PREC_F64 or BSP_PREC_F32 (factor, diagonal inverses and applies), or
 * BSP_PREC_BF16 (factor and diagonal inverses stored in BF16, rounded to nearest even from the
 * FP64 factor; applies in FP32 -- an FP32 preconditioner for bsp_fgmres_mixed). */
#define BSP_PREC_F64 0
#define BSP_PREC_F32 1
#define BSP_PREC_BF16 2
int bsp_precond_create_ilu0_ex(bsp_ctx* ctx, const bsp_bsr* A, const int32_t* colour, int precision,
                               bsp_precond** out);
/* IChol0(A) (preconditioners.hpp:364-538), A symmetric positive definite. */
int bsp_precond_create_ichol0(bsp_ctx* ctx, const bsp_bsr* A, bsp_precond** out);
/* IChol0 of A in the elimination order given by `colour` (rows colour by colour, natural order
 * within a colour; no reference counterpart): the IChol0 factor of P A P^T applied to vectors in
 * A's numbering, with sweeps of at most (number of colours) levels.  NULL = natural order
 * (bsp_precond_create_ichol0). */
int bsp_precond_create_ichol0_ex(bsp_ctx* ctx, const bsp_bsr* A, const int32_t* colour, bsp_precond** out);
/* A user preconditioner (the reference's PreCond concept: any object with
 * `void operator()(solArray&) const`, preconditioners.hpp:542-556, lambdas in
 * blocksp_tests.hpp:113).  The library copies r to a host buffer, calls
 * fn(r_host, len, user) which must overwrite it with M^{-1} r, and copies it back.
 * This keeps arbitrary host preconditioners usable by the device solvers. */
typedef void (*bsp_host_precond_fn)(double* r_host, int64_t len, void* user);
int bsp_precond_create_host(bsp_ctx* ctx, const bsp_bsr* A, bsp_host_precond_fn fn, void* user,
                            bsp_precond** out);
int bsp_precond_destroy(bsp_precond* M);
/* r <- M^{-1} r in place (the operator() of every reference preconditioner). */
int bsp_precond_apply(bsp_ctx* ctx, const bsp_precond* M, double* r_dev);
/* Level sets of the forward (lower=1) or backward (lower=0) sweep: level_host[n]. */
int bsp_precond_levels(const bsp_precond* M, int lower, int32_t* level_host, int* nlevels);
/* Widest level of the forward (lower=1) or backward (lower=0) sweep's slot-ordered ELL copy
 * (1-4 entries per slot; each level as wide as its widest row), or 0 when the sweep reads the
 * CSR factor (8 levels or fewer, blocks above 8 x 8, rows with more than 4 blocks, or padding
 * above a quarter of the blocks).  SOR sweeps count their off-diagonal blocks and have the copy
 * for blocks up to 4 x 4; it holds columns and block positions only, so values set on the matrix
 * after the preconditioner was made still take effect.  No reference counterpart: the
 * copy shortens each level's chain of dependent loads and gives bitwise the same results;
 * the context option "sweep_ell" = 0 makes applies read CSR. */
int bsp_precond_ell_width(const bsp_precond* M, int lower);
/* Factor values in the structure of A (ILU0: LU in A's structure; IChol0: L in lowTri(A)). */
int bsp_precond_factor_info(const bsp_precond* M, int64_t* nnzb);
int bsp_precond_download_factor(bsp_ctx* ctx, const bsp_precond* M, int32_t* rowptr_host,
                                 int32_t* colidx_host, double* vals_host);

/* ---- Krylov solvers (blocksp_solvers.hpp) ---- */

typedef struct {
    int iters;          /* Krylov iterations (the reference's totIters / iter counter) */
    int restarts;       /* deflated restarts */
    int converged;      /* 1 when the stopping rule was met */
    int precond_applies;
    double res0;        /* ||M^{-1} b||_2 */
    double res_est;     /* last residual estimate used by the stopping rule */
    double true_relres; /* ||b - A x||_2 / ||b||_2 of the returned x */
    double setup_ms;    /* not filled by the solvers (preconditioners are built separately) */
    double solve_ms;    /* wall time of the solve on the context stream */
    int fallback;       /* bsp_fgmres_fallback: 1 when the fallback preconditioner took over */
    int reorth;         /* bsp_fgmres_*: steps whose Gram-Schmidt needed a second pass */
} bsp_stats;

/* Stopping rule: 0 = the reference's (pGMRES: least-squares residual estimate of the
 * preconditioned system <= tol*||M^{-1}b||, solvers.hpp:177-183; pCG: ||r|| <= tol*||M^{-1}b||,
 * solvers.hpp:64-65), 1 = true relative residual ||b-Ax||/||b|| <= tol (checked on the device,
 * restarting from the current iterate when the estimate converged first). */
#define BSP_TOL_REFERENCE 0
#define BSP_TOL_TRUE_RESIDUAL 1

/* pGMRES(A, b, tol, M, maxIters, restart, k) with deflated restart (solvers.hpp:156-295,
 * blocksp_deflated_restart.hpp:38-269).  x_dev is output (initial guess 0, as in the reference). */
int bsp_pgmres(bsp_ctx* ctx, const bsp_bsr* A, const bsp_precond* M, const double* b_dev, double* x_dev,
               double tol, int maxIters, int restart, int k, int tol_mode, bsp_stats* stats);
/* pCG(A, b, tol, M, maxIters) (solvers.hpp:51-105). */
int bsp_pcg(bsp_ctx* ctx, const bsp_bsr* A, const bsp_precond* M, const double* b_dev, double* x_dev,
            double tol, int maxIters, int tol_mode, bsp_stats* stats);

/* Orthogonalisation step of the mixed-precision FGMRES (no reference counterpart; the
 * reference's one_arnoldi uses modified Gram-Schmidt, blocksp_solvers.hpp:127-135): w against
 * the m (<= 64) orthonormal device vectors Q_dev_ptrs[0..m-1] (a host array of device
 * pointers) by classical Gram-Schmidt, repeated once when ||w|| drops below 0.01 of its
 * value (cancellation).  h_host[0..m-1] = coefficients, h_host[m] = ||w||; qnew = w / ||w||;
 * w_dev is overwritten with the orthogonalised w. */
int bsp_orthogonalize(bsp_ctx* ctx, int64_t len, const double* const* Q_dev_ptrs, int m, double* w_dev,
                      double* h_host, double* qnew_dev);
/* Allocation-free variant of bsp_orthogonalize.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops).  The same kernels run on a file-scope static device scratch array of 265 KiB per GPU
 * (set up by bsp_ctx_create, kept until the process ends) instead of per-call allocations; the
 * call makes no hipMalloc / hipFree and does not synchronise.  h_dev: m + 2 device doubles
 * (h_dev[0..m-1] coefficients, h_dev[m] = ||w||, h_dev[m + 1] = 1 when the second pass ran).
 * m above 64 is rejected with BSP_E_INVALID.  Every call on one GPU shares that GPU's scratch:
 * calls that overlap on one GPU (other streams or host threads) race. */
int bsp_orthogonalize_noalloc(bsp_ctx* ctx, int64_t len, const double* const* Q_dev_ptrs, int m, double* w_dev,
                              double* h_dev, double* qnew_dev);

/* Ritz-value diagnostics (printEigInfo, blocksp_solvers.hpp:272-288): eigenvalues of the
 * square part of the last bsp_pgmres run's Hessenberg on this context -- harmonic Ritz values
 * (*harmonic = 1) when that run restarted.  Up to cap values into re/im; *count = all. */
int bsp_pgmres_ritz(bsp_ctx* ctx, double* re, double* im, int cap, int* count, int* harmonic);
/* arnoldi_eigenvalues(A, iters, preCond, ..., harmonic_Ritz) (blocksp_solvers.hpp:298-333):
 * `iters` Arnoldi steps of M^{-1} A (M may be NULL) from q_0 = e_0 on the device, then the
 * (harmonic) Ritz values of the Hessenberg.  Fewer values when an invariant subspace is found. */
int bsp_arnoldi_eigenvalues(bsp_ctx* ctx, const bsp_bsr* A, const bsp_precond* M, int iters, int harmonic,
                            double* re, double* im, int cap, int* count);

/* Mixed-precision flexible GMRES(restart) with deflated restart k and FP64 iterative
 * refinement (no reference counterpart; it solves the same systems as pGMRES).  The
 * preconditioner M is applied in its own precision (an FP32 ILU0 from
 * bsp_precond_create_ilu0_ex applies in FP32); the Krylov SpMV uses FP32 values with
 * FP32 accumulation when spmv_precision == BSP_PREC_F32; Gram-Schmidt, least squares,
 * deflation, solution update and the true residual b - Ax are FP64.  Each refinement
 * pass reduces the current true residual by inner_reduction (<= 0: 1e-6); the solve
 * stops when ||b - Ax|| <= tol ||b||.  1 <= restart <= 62 and 0 <= k < restart (there is
 * no unrestarted form).  x_dev is output (initial guess 0). */
int bsp_fgmres_mixed(bsp_ctx* ctx, const bsp_bsr* A, const bsp_precond* M, const double* b_dev, double* x_dev,
                     double tol, int maxIters, int restart, int k, int spmv_precision, double inner_reduction,
                     bsp_stats* stats);

/* As bsp_fgmres_mixed, with a fallback preconditioner (no reference counterpart).  When a
 * refinement pass with M stagnates (8 restarts without halving the residual estimate),
 * produces a non-finite value, or leaves the true residual larger than before, the pass's
 * update is discarded if it was not finite and the remaining passes use M_fallback
 * (stats->fallback = 1).  M_fallback may be NULL (then this is bsp_fgmres_mixed). */
int bsp_fgmres_fallback(bsp_ctx* ctx, const bsp_bsr* A, const bsp_precond* M, const bsp_precond* M_fallback,
                        const double* b_dev, double* x_dev, double tol, int maxIters, int restart, int k,
                        int spmv_precision, double inner_reduction, bsp_stats* stats);

/* ---- Learned ILU0 correction (no reference counterpart) ----
 * A 3-layer message-passing network (hidden width 64) over the block graph of A predicts,
 * for each block row i, a correction D_i (bs x bs) of the diagonal factor block and, for
 * each off-diagonal block (i, j), a scaling g_ij of the factor block:
 *     F_ii = LU_ii (I + 0.05 tanh(D_i)),     F_ij = LU_ij (1 + 0.25 tanh g_ij),
 * where LU is the block ILU(0) factor (tanh elementwise; the bound keeps every F_ii
 * invertible with ||F_ii^-1|| <= ||LU_ii^-1|| / 0.6).  The corrected factor is applied by the same
 * level-scheduled sweeps as ILU0, so a learned apply costs what an ILU0 apply costs.
 * Training minimises mean_k ||L'U' z_k - A z_k||^2 / ||A z_k||^2 over sample vectors z_k
 * (recorded Krylov basis vectors of FP64 ILU0-pGMRES solves) with Adam; forward, backward
 * and optimiser run as HIP kernels in FP32.  Requires block size <= 8 and a structurally
 * symmetric block pattern with every diagonal block present.  Parameters are FP32, in the
 * order Wn[64][bs*bs] bn[64] We[64][bs*bs+1] be[64] {Wm[64][128] bm[64] Wu[64][128] bu[64]}x3
 * Wd[bs*bs][64] bd[bs*bs] wg[128] bg[1] (row-major, y = W x + b). */
typedef struct bsp_mpnn bsp_mpnn;
typedef struct bsp_mpnn_graph bsp_mpnn_graph;

/* Network for block size bs; trunk weights uniform(+-1/sqrt(fan_in)) from `seed`, heads zero
 * (so an untrained network reproduces plain ILU0). */
int bsp_mpnn_create(bsp_ctx* ctx, int bs, uint64_t seed, bsp_mpnn** out);
int bsp_mpnn_destroy(bsp_mpnn* net);
int bsp_mpnn_num_params(const bsp_mpnn* net, int64_t* count);
int bsp_mpnn_get_params(bsp_ctx* ctx, const bsp_mpnn* net, float* host);
/* Sets the parameters and resets the Adam state. */
int bsp_mpnn_set_params(bsp_ctx* ctx, bsp_mpnn* net, const float* host);
/* Gradient of the last bsp_mpnn_loss_grad call. */
int bsp_mpnn_get_grads(bsp_ctx* ctx, const bsp_mpnn* net, float* host);
/* Graph of A with the ILU0 factor of M (an ILU0 preconditioner of A, any elimination order)
 * as the factor to correct.  A must outlive the graph. */
int bsp_mpnn_graph_create(bsp_ctx* ctx, const bsp_bsr* A, const bsp_precond* M, bsp_mpnn_graph** out);
int bsp_mpnn_graph_destroy(bsp_mpnn_graph* g);
/* Network outputs on a graph: dD (n x bs*bs) and g (one per off-diagonal block, in CSR order). */
int bsp_mpnn_infer(bsp_ctx* ctx, const bsp_mpnn* net, bsp_mpnn_graph* g, float* dD_host, float* g_host);
/* Loss over K sample vectors Z_dev (K x n*bs, FP64, contiguous) and its gradient (kept in
 * the network for bsp_mpnn_adam_step / bsp_mpnn_get_grads). */
int bsp_mpnn_loss_grad(bsp_ctx* ctx, bsp_mpnn* net, bsp_mpnn_graph* g, const double* Z_dev, int K, double* loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected) with the last gradient. */
int bsp_mpnn_adam_step(bsp_ctx* ctx, bsp_mpnn* net, float lr);
/* `epochs` passes over the graphs; each pass takes minibatches of `batch` vectors from each
 * graph's sample set in turn (one Adam step per minibatch).  epoch_loss[e] (may be NULL)
 * receives the mean minibatch loss of pass e. */
int bsp_mpnn_train(bsp_ctx* ctx, bsp_mpnn* net, int ngraphs, bsp_mpnn_graph* const* graphs,
                   const double* const* Z_dev, const int* K, int epochs, int batch, float lr, double* epoch_loss);
/* ILU0 of A in the elimination order given by `colour` (NULL: natural), corrected by the
 * network's inference on A's graph, as a preconditioner of the given precision. */
int bsp_precond_create_learned(bsp_ctx* ctx, const bsp_bsr* A, const bsp_mpnn* net, const int32_t* colour,
                               int precision, bsp_precond** out);

/* Records the Krylov basis vectors of the following bsp_pgmres solves of length `len`
 * (= n * bs; solves of another length are not recorded) on this context into buf_dev
 * (capacity `cap` vectors; every `every`-th basis vector is kept).  cap = 0 stops recording.
 * bsp_ctx_recorded gives the number recorded so far. */
int bsp_ctx_record_basis(bsp_ctx* ctx, double* buf_dev, int64_t len, int cap, int every);
int bsp_ctx_recorded(const bsp_ctx* ctx, int* count);

/* ---- Batched IChol0-pCG over many independent systems (no reference counterpart) ----
 * The systems are the diagonal blocks of one block-diagonal BSR matrix A: system s owns
 * block rows [sys_rowptr[s], sys_rowptr[s+1]) and may not couple to other systems.  Each
 * (system, right-hand side) is solved by the reference's pCG (blocksp_solvers.hpp:51-105)
 * with the IChol0 preconditioner (blocksp_preconditioners.hpp:364-538) of that system.
 * All systems run in one persistent kernel launch (work items largest system first, one
 * workgroup per item, level-set sweeps separated by workgroup barriers), replayed from a
 * captured HIP graph.  Block sizes 1, 2, 4, 8. */
typedef struct bsp_batch bsp_batch;
/* Builds the IChol0 factor and the per-system level sets (A must outlive the batch and keep
 * its values: the factor and its ELL copies are taken at creation). */
int bsp_batch_create(bsp_ctx* ctx, const bsp_bsr* A, int nsys, const int32_t* sys_rowptr, int nrhs,
                     bsp_batch** out);
/* As bsp_batch_create with each system's IChol0 eliminated colour by colour (`colour`: one
 * non-negative colour per block row of A, e.g. coordinate parity of the grid cells; NULL =
 * natural order; no reference counterpart).  A few levels per sweep instead of ~2 sqrt(n):
 * more pCG iterations, far fewer dependent steps per iteration. */
int bsp_batch_create_ex(bsp_ctx* ctx, const bsp_bsr* A, int nsys, const int32_t* sys_rowptr, int nrhs,
                        const int32_t* colour, bsp_batch** out);
int bsp_batch_destroy(bsp_batch* b);
/* Tiles: workgroup size (multiple of 64, <= 1024) and right-hand sides per work item
 * (1, 2 or 4, dividing nrhs). */
int bsp_batch_set_tiles(bsp_ctx* ctx, bsp_batch* b, int wg_size, int rhs_per_item);
/* Times every tile choice on probe_iters iterations of the batch (B_dev/X_dev as for
 * bsp_batch_pcg; X_dev is overwritten), keeps the fastest of those whose per-(system, RHS)
 * residual norms and ||X|| reproduce the first candidate's (up to rounding), and caches the
 * choice per (device, block size, nrhs, nsys, rows) for later batches. */
int bsp_batch_autotune(bsp_ctx* ctx, bsp_batch* b, const double* B_dev, double* X_dev, int probe_iters,
                       int* wg_size, int* rhs_per_item);
/* As bsp_batch_autotune's probes (not cached), then one complete solve (tol, maxIters,
 * tol_mode as for bsp_batch_pcg; X_dev is overwritten) with each RHS-per-item's fastest
 * workgroup size; keeps the fastest.  Short probes favour throughput, but a batch whose
 * largest systems run far longer than the rest finishes with those items, which complete
 * solves show. */
int bsp_batch_autotune_solve(bsp_ctx* ctx, bsp_batch* b, const double* B_dev, double* X_dev, double tol, int maxIters,
                             int tol_mode, int probe_iters, int* wg_size, int* rhs_per_item);
/* Solves A_s X_{s,k} = B_{s,k} for every system s and right-hand side k.  B_dev and X_dev
 * hold nrhs vectors of length n*bs one after another (vector k: every system's rows in
 * order).  Stopping rule per (s, k): tol_mode 0 ||r|| <= tol ||M^{-1} b|| (the reference's),
 * 1 ||r|| <= tol ||b|| (r the recurrence residual).  iters_host / relres_host (may be NULL):
 * nsys * nrhs entries, (s, k) at s * nrhs + k; relres is ||r|| over the rule's reference
 * norm.  solve_ms: device time of the launch.  Returns BSP_NOT_CONVERGED if any pair did
 * not meet the rule within maxIters. */
int bsp_batch_pcg(bsp_ctx* ctx, bsp_batch* b, const double* B_dev, double* X_dev, double tol, int maxIters,
                  int tol_mode, int32_t* iters_host, double* relres_host, double* solve_ms);
int bsp_batch_info(const bsp_batch* b, int* wg_size, int* rhs_per_item, int* nitems);
/* Entries per slot of the slot-ordered ELL copies of the two factor sweeps (2, 3 or 4), or 0
 * when the sweeps read the CSR factor (rows with more than 4 strict blocks, blocks larger
 * than 4 x 4, or the context option "batch_ell" = 0 when the batch was created).  Both forms give bitwise the same
 * results; the ELL form shortens each level's chain of dependent loads. */
int bsp_batch_ell_width(const bsp_batch* b);

#ifdef __cplusplus
}
#endif
#endif /* BLOCKSP_C_H */
