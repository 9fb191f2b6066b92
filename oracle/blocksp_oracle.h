/*
 * blocksp_oracle.h -- CPU restatement of BlockSp's block-sparse solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation in blocksp_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * Every routine follows the reference's loop and summation order (file:line
 * citations are relative to the BlockSp v1.0 checkout).  LAPACK semantics
 * (getrf/getrs/potrf/pptrf/gels/gees/trsen/geqrf/orgqr) are restated from the
 * textbook algorithms, never from a vendor library.
 *
 * Storage convention (BSR, "block CSR"):
 *   rowptr[n+1], colidx[nnzb] in the reference's per-row insertion order
 *   (blocksp_containers.hpp:147-164), vals[nnzb*bs*bs] row-major per block.
 * Vectors are the solArray flat layout (ie, dim, ib) -> ie*B*D + dim*B + ib
 * (blocksp_containers.hpp:572-736, blocksp_utility.hpp:242-265).
 *
 * Parity status of each function is stated next to it ("pinned by ...").
 */
#ifndef BLOCKSP_ORACLE_H
#define BLOCKSP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* y (+)= A x.  blocksp_multiply.hpp:170-325 (non-transpose, BlockSp<T,B*D> form). */
void orc_bsmv(int n, int bs, const int* rowptr, const int* colidx, const double* vals,
              const double* x, double* y, int accumulate);

/* Per-dimension form: BlockSp<T,B> times solArray<T,B,D> (multiply.hpp:170-193, 252-285).
 * dim < 0 multiplies every dimension. */
void orc_bsmv_dim(int n, int B, int D, int dim, const int* rowptr, const int* colidx,
                  const double* vals, const double* x, double* y, int accumulate);

/* Transposed product y (+)= A^T x for the BlockSp<T,B*D> form (multiply.hpp:236-245, 310-322). */
void orc_bsmv_transpose(int nrows, int ncols, int bs, const int* rowptr, const int* colidx,
                        const double* vals, const double* x, double* y, int accumulate);

/* Dense block factor: LUFactor<T,N> (dense_solvers.hpp:47-110).  For bs<=3 the
 * block is stored unfactored and solved by the explicit formulas (small_solve).
 * lu[bs*bs], piv[bs] (0-based row interchanges, LAPACK getrf convention). */
void orc_lu_factor(int bs, const double* a, double* lu, int* piv);
void orc_lu_solve(int bs, const double* lu, const int* piv, double* x);        /* x <- A^{-1} x */
void orc_lu_rsolve(int bs, const double* lu, const int* piv, double* X);       /* X <- X A^{-1} */

/* CholeskyFactor<T,N> (dense_solvers.hpp:123-190): packed lower, explicit for bs<=3. */
void orc_chol_factor(int bs, const double* a, double* packed);
void orc_chol_solve(int bs, const double* packed, double* x);

/* Block ILU(0): preconditioners.hpp:255-291 (factor) and makeADInv (22-63).
 * lu_vals has the same structure as vals; dlu/dpiv hold LUFactor of each U_ii. */
void orc_ilu0_factor(int n, int bs, const int* rowptr, const int* colidx, const double* vals,
                     double* lu_vals, double* dlu, int* dpiv);
/* ILU0::invmul (preconditioners.hpp:294-349). */
void orc_ilu0_apply(int n, int bs, const int* rowptr, const int* colidx, const double* lu_vals,
                    const double* dlu, const int* dpiv, double* x);

/* Block IC(0): preconditioners.hpp:374-429.  The factor lives in the structure of
 * lowTri(A) (containers.hpp:277-285): l_rowptr/l_colidx are outputs sized n+1 and
 * <= nnzb.  Returns nnz of L. */
int orc_ichol0_factor(int n, int bs, const int* rowptr, const int* colidx, const double* vals,
                      int* l_rowptr, int* l_colidx, double* l_vals);
/* IChol0::invmul (preconditioners.hpp:432-489).  Backward solve visits the rows of
 * column im in ascending order (the reference visits them in hash-set order). */
void orc_ichol0_apply(int n, int bs, const int* l_rowptr, const int* l_colidx,
                      const double* l_vals, double* x);

/* SOR::operator() (preconditioners.hpp:116-234) with LUFactor (inv_kind 0) or
 * CholeskyFactor (inv_kind 1) diagonal inverses.  x <- M^{-1} x, from x0 = 0. */
void orc_sor_apply(int n, int bs, const int* rowptr, const int* colidx, const double* vals,
                   double w, int nu, int type, int inv_kind, double* x);
/* Diag::invert (preconditioners.hpp:75-111). */
void orc_diag_apply(int n, int bs, const int* rowptr, const int* colidx, const double* vals,
                    int inv_kind, double* x);

/* Level sets of the strictly lower (lower=1) or strictly upper (lower=0) block
 * pattern: level[i] = 1 + max level of the rows it reads.  Returns #levels. */
int orc_levels(int n, const int* rowptr, const int* colidx, int lower, int* level);

/* Preconditioner selector for the Krylov oracles. */
enum { ORC_PRE_NONE = 0, ORC_PRE_DIAG = 1, ORC_PRE_SOR = 2, ORC_PRE_ILU0 = 3, ORC_PRE_ICHOL0 = 4 };

typedef struct {
    int kind;        /* ORC_PRE_* */
    int inv_kind;    /* 0 LUFactor, 1 CholeskyFactor (Diag/SOR only) */
    double w;        /* SOR damping */
    int nu;          /* SOR sweeps per direction */
    int type;        /* SOR: <0 backward, 0 symmetric, >0 forward */
} orc_precond;

typedef struct {
    int iters;             /* total Krylov iterations (reference counter) */
    int restarts;          /* deflated restarts performed */
    double res0;           /* ||M^{-1} b|| */
    double res_est;        /* last least-squares / recurrence residual estimate */
    double true_relres;    /* ||b - A x|| / ||b|| of the returned x */
} orc_stats;

/* pCG (solvers.hpp:51-105). */
int orc_pcg(int n, int bs, const int* rowptr, const int* colidx, const double* vals,
            const orc_precond* pre, const double* b, double* x, double tol, int maxIters,
            orc_stats* st);

/* pGMRES with deflated restart (solvers.hpp:109-295, deflated_restart.hpp:38-269). */
int orc_pgmres(int n, int bs, const int* rowptr, const int* colidx, const double* vals,
               const orc_precond* pre, const double* b, double* x, double tol, int maxIters,
               int restart, int k, orc_stats* st);

/* Harmonic-Ritz selection + invariant subspace used by deflated_restart
 * (deflated_restart.hpp:61-193) on an (n+1) x n Hessenberg H (row-major).
 * Writes P (n x kout, row-major, orthonormal) and returns kout (= sdim). */
int orc_harmonic_subspace(int n, const double* H, int k, double* P, double* eig_re,
                          double* eig_im);

/* Real eigenvalues of a general n x n matrix (row-major) via Hessenberg + complex
 * shifted QR; used to pin the Schur code against numpy. */
void orc_eigvals(int n, const double* A, double* re, double* im);

/* Least squares min ||y - H x|| for m x n H (m >= n), Householder QR (gels 'N').
 * y has m entries; on return y[0..n-1] is the solution. */
void orc_gels(int m, int n, const double* H, double* y);

/* bsmm: C = A * B with the reference's first-touch column order
 * (multiply.hpp:51-97).  Returns nnzb of C, or -(needed) if cap is too small. */
int orc_bsmm(int nA, int bs, const int* a_rowptr, const int* a_colidx, const double* a_vals,
             const int* b_rowptr, const int* b_colidx, const double* b_vals,
             int* c_rowptr, int* c_colidx, double* c_vals, int cap);

/* bsmm with Atran: C = A^T * B for A (nA x mA blocks) and B (nA rows), C(j, c) inserted on
 * first touch while walking A's rows in order (multiply.hpp:99-134).  Returns nnzb of C,
 * or -(needed) if cap is too small. */
int orc_bsmm_transA(int nA, int mA, int bs, const int* a_rowptr, const int* a_colidx, const double* a_vals,
                    const int* b_rowptr, const int* b_colidx, const double* b_vals,
                    int* c_rowptr, int* c_colidx, double* c_vals, int cap);

#ifdef __cplusplus
}
#endif
#endif
