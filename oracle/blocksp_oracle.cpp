// blocksp_oracle.cpp -- CPU restatement of BlockSp's block-sparse solve path.
//
// TEST INFRASTRUCTURE ONLY: the parity checker for the HIP product.  Built with
// -O2 -ffp-contract=off so that every expression rounds in the order written
// (the reference's own loop order).  See blocksp_oracle.h for the contract.
//
// Parity anchors (what pins each part; details in tests/test_oracle_*.py):
//  * bsmv, BLAS-1, structure     : dense numpy products, exact integer equality.
//  * LU / Cholesky / small_solve : A*A^{-1} = I, explicit 2x2/3x3 formulas.
//  * ILU0 / IChol0               : (LU)_ij = A_ij on the pattern; exact on
//                                  block-tridiagonal patterns; LDG fixture solve.
//  * SOR                         : dense Gauss-Seidel brute force.
//  * pCG / pGMRES                : Krylov optimality (numpy lstsq), LDG fixture
//                                  known answer ||x - x_sol||_inf ~ 2.9e-4
//                                  (blocksp_tests.hpp:312-342).
//  * deflated restart            : span/Arnoldi-relation invariants
//                                  (deflated_restart.hpp:275-362), numpy eig.
#include "blocksp_oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

namespace {

using cd = std::complex<double>;
constexpr double kEps = std::numeric_limits<double>::epsilon();

// ---------------------------------------------------------------------------
// Dense block kernels (blocksp_dense_multiply.hpp)
// ---------------------------------------------------------------------------

// matVec_nnl (dense_multiply.hpp:29-48): explicit 1x1 / 2x2 forms, else
// val = 0; val += a(r,c) * x(c) in column order.
inline void matvec(int bs, const double* a, const double* x, double* out) {
    if (bs == 1) { out[0] = a[0] * x[0]; return; }
    if (bs == 2) {
        double o0 = a[0] * x[0] + a[1] * x[1];
        double o1 = a[2] * x[0] + a[3] * x[1];
        out[0] = o0; out[1] = o1;
        return;
    }
    for (int r = 0; r < bs; ++r) {
        double val = 0.0;
        for (int c = 0; c < bs; ++c) val += a[r * bs + c] * x[c];
        out[r] = val;
    }
}

// matVec_nnl_transpose (dense_multiply.hpp:96-112): mv = 0; mv(in) += a(im,in) x(im).
inline void matvec_t(int bs, const double* a, const double* x, double* out) {
    if (bs == 1) { out[0] = a[0] * x[0]; return; }
    for (int c = 0; c < bs; ++c) out[c] = 0.0;
    for (int r = 0; r < bs; ++r)
        for (int c = 0; c < bs; ++c) out[c] += a[r * bs + c] * x[r];
}

// matMat (dense_multiply.hpp:129-152): i-k-j loops from zero.
inline void matmat(int bs, const double* a, const double* b, double* out) {
    if (bs == 1) { out[0] = a[0] * b[0]; return; }
    for (int i = 0; i < bs * bs; ++i) out[i] = 0.0;
    for (int i = 0; i < bs; ++i)
        for (int k = 0; k < bs; ++k)
            for (int j = 0; j < bs; ++j) out[i * bs + j] += a[i * bs + k] * b[k * bs + j];
}

inline void transpose(int bs, double* a) {
    for (int i = 0; i < bs; ++i)
        for (int j = i + 1; j < bs; ++j) std::swap(a[i * bs + j], a[j * bs + i]);
}

// explicit_small_Inv (blocksp_utility.hpp:841-890) for N = 2, 3.
void small_inv3(const double* A, double* I) {
    auto a = [&](int r, int c) { return A[r * 3 + c]; };
    I[0 * 3 + 0] = a(1, 1) * a(2, 2) - a(2, 1) * a(1, 2);
    I[1 * 3 + 0] = -(a(1, 0) * a(2, 2) - a(2, 0) * a(1, 2));
    I[2 * 3 + 0] = a(1, 0) * a(2, 1) - a(2, 0) * a(1, 1);
    I[0 * 3 + 1] = -(a(0, 1) * a(2, 2) - a(2, 1) * a(0, 2));
    I[1 * 3 + 1] = a(0, 0) * a(2, 2) - a(2, 0) * a(0, 2);
    I[2 * 3 + 1] = -(a(0, 0) * a(2, 1) - a(2, 0) * a(0, 1));
    I[0 * 3 + 2] = a(0, 1) * a(1, 2) - a(1, 1) * a(0, 2);
    I[1 * 3 + 2] = -(a(0, 0) * a(1, 2) - a(1, 0) * a(0, 2));
    I[2 * 3 + 2] = a(0, 0) * a(1, 1) - a(1, 0) * a(0, 1);
    double det = a(0, 0) * (a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1)) -
                 a(0, 1) * (a(1, 0) * a(2, 2) - a(1, 2) * a(2, 0)) +
                 a(0, 2) * (a(1, 0) * a(2, 1) - a(1, 1) * a(2, 0));
    for (int i = 0; i < 9; ++i) I[i] /= det;
}

// Packed-lower symmetric 3x3 inverse (blocksp_utility.hpp:897-944).
void small_inv3_packed(const double* A, double* I) {
    I[0] = A[2] * A[5] - A[4] * A[4];
    I[3] = -(A[1] * A[5] - A[3] * A[4]);
    I[6] = A[1] * A[4] - A[3] * A[2];
    I[4] = A[0] * A[5] - A[3] * A[3];
    I[7] = -(A[0] * A[4] - A[3] * A[1]);
    I[8] = A[0] * A[2] - A[1] * A[1];
    I[1] = I[3]; I[2] = I[6]; I[5] = I[7];
    double det = A[0] * (A[2] * A[5] - A[4] * A[4]) - A[1] * (A[1] * A[5] - A[4] * A[3]) +
                 A[3] * (A[1] * A[4] - A[2] * A[3]);
    for (int i = 0; i < 9; ++i) I[i] /= det;
}

// getrf: LU with partial pivoting (first max |a| in the column), row-major,
// unit-lower L below the diagonal, U on and above.  Golub & Van Loan alg. 3.4.1.
void getrf(int n, double* a, int* piv) {
    for (int j = 0; j < n; ++j) {
        int p = j;
        double amax = std::fabs(a[j * n + j]);
        for (int i = j + 1; i < n; ++i)
            if (std::fabs(a[i * n + j]) > amax) { amax = std::fabs(a[i * n + j]); p = i; }
        piv[j] = p;
        if (p != j)
            for (int c = 0; c < n; ++c) std::swap(a[j * n + c], a[p * n + c]);
        double d = a[j * n + j];
        if (d != 0.0)
            for (int i = j + 1; i < n; ++i) a[i * n + j] /= d;
        for (int i = j + 1; i < n; ++i) {
            double l = a[i * n + j];
            for (int c = j + 1; c < n; ++c) a[i * n + c] -= l * a[j * n + c];
        }
    }
}

// getrs 'N': x <- A^{-1} x.
void getrs_n(int n, const double* lu, const int* piv, double* x) {
    for (int j = 0; j < n; ++j)
        if (piv[j] != j) std::swap(x[j], x[piv[j]]);
    for (int i = 1; i < n; ++i) {
        double s = x[i];
        for (int c = 0; c < i; ++c) s -= lu[i * n + c] * x[c];
        x[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = x[i];
        for (int c = i + 1; c < n; ++c) s -= lu[i * n + c] * x[c];
        x[i] = s / lu[i * n + i];
    }
}

// getrs 'T': x <- A^{-T} x  (A = P L U, A^T = U^T L^T P^T).
void getrs_t(int n, const double* lu, const int* piv, double* x) {
    for (int i = 0; i < n; ++i) {          // U^T z = x (forward)
        double s = x[i];
        for (int r = 0; r < i; ++r) s -= lu[r * n + i] * x[r];
        x[i] = s / lu[i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {     // L^T w = z (backward, unit)
        double s = x[i];
        for (int r = i + 1; r < n; ++r) s -= lu[r * n + i] * x[r];
        x[i] = s;
    }
    for (int j = n - 1; j >= 0; --j)
        if (piv[j] != j) std::swap(x[j], x[piv[j]]);
}

// LUFactor::invmul on a vector (dense_solvers.hpp:80-92, 1029-1057).
void lu_solve(int bs, const double* lu, const int* piv, double* x) {
    if (bs == 1) { x[0] = x[0] / lu[0]; return; }
    if (bs == 2) {
        const double* A = lu;
        double det = A[0] * A[3] - A[1] * A[2];
        double y0 = (A[3] * x[0] - A[1] * x[1]) / det;
        double y1 = (-A[2] * x[0] + A[0] * x[1]) / det;
        x[0] = y0; x[1] = y1;
        return;
    }
    if (bs == 3) {
        double inv[9], y[3];
        small_inv3(lu, inv);
        matvec(3, inv, x, y);
        x[0] = y[0]; x[1] = y[1]; x[2] = y[2];
        return;
    }
    getrs_n(bs, lu, piv, x);
}

// LUFactor::r_invmul (dense_solvers.hpp:97-108, 1063-1095): X <- X A^{-1}.
void lu_rsolve(int bs, const double* lu, const int* piv, double* X) {
    if (bs == 1) { X[0] = X[0] / lu[0]; return; }
    if (bs == 2) {
        const double* A = lu;
        const double* B = X;
        double det = A[0] * A[3] - A[1] * A[2];
        double x00 = B[0] * A[3] - B[1] * A[2];
        double x01 = -B[0] * A[1] + B[1] * A[0];
        double x10 = B[2] * A[3] - B[3] * A[2];
        double x11 = -B[2] * A[1] + B[3] * A[0];
        X[0] = x00 / det; X[1] = x01 / det; X[2] = x10 / det; X[3] = x11 / det;
        return;
    }
    if (bs == 3) {
        double inv[9], out[9];
        small_inv3(lu, inv);
        matmat(3, X, inv, out);
        std::memcpy(X, out, sizeof(out));
        return;
    }
    // each row r of X: solve A^T y = X(r,:)^T
    for (int r = 0; r < bs; ++r) getrs_t(bs, lu, piv, X + r * bs);
}

void lu_factor(int bs, const double* a, double* lu, int* piv) {
    std::memcpy(lu, a, sizeof(double) * bs * bs);
    for (int i = 0; i < bs; ++i) piv[i] = i;
    if (bs > 3) getrf(bs, lu, piv);
}

// CholeskyFactor (dense_solvers.hpp:123-190): packed lower, pptrf for N > 3.
void chol_factor(int bs, const double* a, double* L) {
    int il = 0;
    for (int i = 0; i < bs; ++i)
        for (int j = 0; j <= i; ++j) L[il++] = a[i * bs + j];
    if (bs <= 3) return;
    auto idx = [](int i, int j) { return i * (i + 1) / 2 + j; };
    for (int j = 0; j < bs; ++j) {
        double s = L[idx(j, j)];
        for (int p = 0; p < j; ++p) s -= L[idx(j, p)] * L[idx(j, p)];
        double d = std::sqrt(s);
        L[idx(j, j)] = d;
        for (int i = j + 1; i < bs; ++i) {
            double t = L[idx(i, j)];
            for (int p = 0; p < j; ++p) t -= L[idx(i, p)] * L[idx(j, p)];
            L[idx(i, j)] = t / d;
        }
    }
}

void chol_solve(int bs, const double* L, double* x) {
    if (bs == 1) { x[0] = x[0] / L[0]; return; }
    if (bs == 2) {
        double det = L[0] * L[2] - L[1] * L[1];
        double y0 = (L[2] * x[0] - L[1] * x[1]) / det;
        double y1 = (-L[1] * x[0] + L[0] * x[1]) / det;
        x[0] = y0; x[1] = y1;
        return;
    }
    if (bs == 3) {
        double inv[9], y[3];
        small_inv3_packed(L, inv);
        matvec(3, inv, x, y);
        x[0] = y[0]; x[1] = y[1]; x[2] = y[2];
        return;
    }
    auto idx = [](int i, int j) { return i * (i + 1) / 2 + j; };
    for (int i = 0; i < bs; ++i) {
        double s = x[i];
        for (int p = 0; p < i; ++p) s -= L[idx(i, p)] * x[p];
        x[i] = s / L[idx(i, i)];
    }
    for (int i = bs - 1; i >= 0; --i) {
        double s = x[i];
        for (int p = i + 1; p < bs; ++p) s -= L[idx(p, i)] * x[p];
        x[i] = s / L[idx(i, i)];
    }
}

// ---------------------------------------------------------------------------
// BSR helpers
// ---------------------------------------------------------------------------

struct Bsr {
    int n, bs;
    const int* rp;
    const int* ci;
    const double* v;
    const double* blk(int p) const { return v + (size_t)p * bs * bs; }
};

// position of column j in row i, or -1 (containers.hpp:206-212 scan/at)
inline int find_block(const int* rp, const int* ci, int i, int j) {
    for (int p = rp[i]; p < rp[i + 1]; ++p)
        if (ci[p] == j) return p;
    return -1;
}

// dot (containers.hpp:974-996): running sum of per-block blitz::dot.
double vdot(int len, int bs, const double* a, const double* b) {
    double d = 0.0;
    for (int ie = 0; ie < len; ++ie) {
        const double* x = a + (size_t)ie * bs;
        const double* y = b + (size_t)ie * bs;
        double bd = x[0] * y[0];
        for (int i = 1; i < bs; ++i) bd += x[i] * y[i];
        d += bd;
    }
    return d;
}

double vnorm(int len, int bs, const double* a) { return std::sqrt(std::fabs(vdot(len, bs, a, a))); }

void bsmv_impl(const Bsr& A, const double* x, double* y, bool acc) {
    const int bs = A.bs;
    std::vector<double> ph(bs);
    for (int i = 0; i < A.n; ++i) {
        double* yi = y + (size_t)i * bs;
        if (!acc)
            for (int r = 0; r < bs; ++r) yi[r] = 0.0;
        for (int p = A.rp[i]; p < A.rp[i + 1]; ++p) {
            matvec(bs, A.blk(p), x + (size_t)A.ci[p] * bs, ph.data());
            for (int r = 0; r < bs; ++r) yi[r] += ph[r];
        }
    }
}

// ---------------------------------------------------------------------------
// Preconditioners (blocksp_preconditioners.hpp)
// ---------------------------------------------------------------------------

struct DiagInv {
    int bs = 0, kind = 0;
    std::vector<double> f;   // LU (bs*bs) or packed Cholesky
    std::vector<int> piv;
    void build(const Bsr& A, int inv_kind) {
        bs = A.bs; kind = inv_kind;
        int fs = kind == 0 ? bs * bs : bs * (bs + 1) / 2;
        f.assign((size_t)A.n * fs, 0.0);
        piv.assign((size_t)A.n * bs, 0);
        for (int i = 0; i < A.n; ++i) {
            int p = find_block(A.rp, A.ci, i, i);
            std::vector<double> z(bs * bs, 0.0);
            const double* a = p >= 0 ? A.blk(p) : z.data();
            if (kind == 0) lu_factor(bs, a, &f[(size_t)i * fs], &piv[(size_t)i * bs]);
            else chol_factor(bs, a, &f[(size_t)i * fs]);
        }
    }
    void solve(int i, double* x) const {
        if (kind == 0) lu_solve(bs, &f[(size_t)i * bs * bs], &piv[(size_t)i * bs], x);
        else chol_solve(bs, &f[(size_t)i * bs * (bs + 1) / 2], x);
    }
};

// One SOR sweep (preconditioners.hpp:135-202).
void sor_sweep(const Bsr& A, const DiagInv& Di, const double* b, double* x, double w, bool fwd) {
    const int bs = A.bs;
    std::vector<double> ph(bs), t(bs);
    for (int s = 0; s < A.n; ++s) {
        int i = fwd ? s : A.n - 1 - s;
        for (int r = 0; r < bs; ++r) ph[r] = 0.0;
        for (int p = A.rp[i]; p < A.rp[i + 1]; ++p) {
            int c = A.ci[p];
            if (c == i) continue;
            matvec(bs, A.blk(p), x + (size_t)c * bs, t.data());
            for (int r = 0; r < bs; ++r) ph[r] += t[r];
        }
        for (int r = 0; r < bs; ++r) ph[r] = b[(size_t)i * bs + r] - ph[r];
        Di.solve(i, ph.data());
        double* xi = x + (size_t)i * bs;
        for (int r = 0; r < bs; ++r) xi[r] += -w * xi[r] + w * ph[r];
    }
}

struct Ilu0 {
    int n, bs;
    const int* rp;
    const int* ci;
    std::vector<double> lu, dlu;
    std::vector<int> dpiv;
};

void ilu0_factor_impl(const Bsr& A, double* LU, double* dlu, int* dpiv) {
    const int n = A.n, bs = A.bs, b2 = bs * bs;
    std::memcpy(LU, A.v, sizeof(double) * (size_t)A.rp[n] * b2);
    // rowIndicesByCol (containers.hpp:410-417): rows with a block in column c.
    std::vector<std::vector<int>> rows_of_col(n);
    for (int i = 0; i < n; ++i)
        for (int p = A.rp[i]; p < A.rp[i + 1]; ++p) rows_of_col[A.ci[p]].push_back(i);
    std::vector<double> f(b2), mm(b2);
    std::vector<int> piv(bs);
    for (int i = 0; i < n; ++i) {
        int pd = find_block(A.rp, A.ci, i, i);
        lu_factor(bs, LU + (size_t)pd * b2, f.data(), piv.data());
        for (int k : rows_of_col[i])
            if (k > i) lu_rsolve(bs, f.data(), piv.data(), LU + (size_t)find_block(A.rp, A.ci, k, i) * b2);
        for (int k : rows_of_col[i]) {
            if (k <= i) continue;
            const double* lki = LU + (size_t)find_block(A.rp, A.ci, k, i) * b2;
            for (int p = A.rp[k]; p < A.rp[k + 1]; ++p) {
                int j = A.ci[p];
                if (j <= i) continue;
                int pij = find_block(A.rp, A.ci, i, j);   // rowInd[j].contains(i)
                if (pij < 0) continue;
                matmat(bs, lki, LU + (size_t)pij * b2, mm.data());
                double* akj = LU + (size_t)p * b2;
                for (int e = 0; e < b2; ++e) akj[e] -= mm[e];
            }
        }
    }
    for (int i = 0; i < n; ++i) {
        int pd = find_block(A.rp, A.ci, i, i);
        lu_factor(bs, LU + (size_t)pd * b2, dlu + (size_t)i * b2, dpiv + (size_t)i * bs);
    }
}

void ilu0_apply_impl(int n, int bs, const int* rp, const int* ci, const double* LU,
                     const double* dlu, const int* dpiv, double* x) {
    const int b2 = bs * bs;
    std::vector<double> t(bs);
    for (int i = 1; i < n; ++i) {
        double* xi = x + (size_t)i * bs;
        for (int p = rp[i]; p < rp[i + 1]; ++p) {
            int c = ci[p];
            if (c >= i) continue;
            matvec(bs, LU + (size_t)p * b2, x + (size_t)c * bs, t.data());
            for (int r = 0; r < bs; ++r) xi[r] -= t[r];
        }
    }
    for (int i = n - 1; i >= 0; --i) {
        double* xi = x + (size_t)i * bs;
        for (int p = rp[i]; p < rp[i + 1]; ++p) {
            int c = ci[p];
            if (c <= i) continue;
            matvec(bs, LU + (size_t)p * b2, x + (size_t)c * bs, t.data());
            for (int r = 0; r < bs; ++r) xi[r] -= t[r];
        }
        lu_solve(bs, dlu + (size_t)i * b2, dpiv + (size_t)i * bs, xi);
    }
}

// In-block triangular substitutions (preconditioners.hpp:499-536).
void lii_fwd_vec(int bs, const double* L, double* x) {
    for (int i = 0; i < bs; ++i) {
        for (int j = 0; j < i; ++j) x[i] -= L[i * bs + j] * x[j];
        x[i] /= L[i * bs + i];
    }
}
void liiT_bwd_vec(int bs, const double* L, double* x) {
    for (int i = bs - 1; i >= 0; --i) {
        for (int j = i + 1; j < bs; ++j) x[i] -= L[j * bs + i] * x[j];
        x[i] /= L[i * bs + i];
    }
}

int ichol0_factor_impl(const Bsr& A, int* lrp, int* lci, double* lv) {
    const int n = A.n, bs = A.bs, b2 = bs * bs;
    // lowTri (containers.hpp:277-285): keep col <= row, in row order.
    int nz = 0;
    lrp[0] = 0;
    for (int i = 0; i < n; ++i) {
        for (int p = A.rp[i]; p < A.rp[i + 1]; ++p)
            if (A.ci[p] <= i) {
                lci[nz] = A.ci[p];
                std::memcpy(lv + (size_t)nz * b2, A.blk(p), sizeof(double) * b2);
                ++nz;
            }
        lrp[i + 1] = nz;
    }
    std::vector<std::vector<int>> rows_of_col(n);   // of the full A pattern
    for (int i = 0; i < n; ++i)
        for (int p = A.rp[i]; p < A.rp[i + 1]; ++p) rows_of_col[A.ci[p]].push_back(i);
    std::vector<double> mm(b2), ljT(b2);
    for (int i = 0; i < n; ++i) {
        double* lii = lv + (size_t)find_block(lrp, lci, i, i) * b2;
        if (bs == 1) {
            lii[0] = std::sqrt(lii[0]);
        } else {
            // potrf 'L' (row-major) then zero the strict upper part.
            for (int j = 0; j < bs; ++j) {
                double s = lii[j * bs + j];
                for (int p = 0; p < j; ++p) s -= lii[j * bs + p] * lii[j * bs + p];
                double d = std::sqrt(s);
                lii[j * bs + j] = d;
                for (int r = j + 1; r < bs; ++r) {
                    double t = lii[r * bs + j];
                    for (int p = 0; p < j; ++p) t -= lii[r * bs + p] * lii[j * bs + p];
                    lii[r * bs + j] = t / d;
                }
            }
            for (int r = 0; r < bs; ++r)
                for (int c = r + 1; c < bs; ++c) lii[r * bs + c] = 0.0;
        }
        for (int k : rows_of_col[i]) {
            if (k <= i) continue;
            double* lki = lv + (size_t)find_block(lrp, lci, k, i) * b2;
            if (bs == 1) { lki[0] /= lii[0]; continue; }
            // X <- X L_ii^{-T}: transpose, forward substitute each column, transpose.
            transpose(bs, lki);
            for (int r = 0; r < bs; ++r)
                for (int c = 0; c < bs; ++c) {
                    for (int j = 0; j < r; ++j) lki[r * bs + c] -= lii[r * bs + j] * lki[j * bs + c];
                    lki[r * bs + c] /= lii[r * bs + r];
                }
            transpose(bs, lki);
        }
        for (int k : rows_of_col[i]) {
            if (k <= i) continue;
            const double* lki = lv + (size_t)find_block(lrp, lci, k, i) * b2;
            for (int p = lrp[k]; p < lrp[k + 1]; ++p) {
                int j = lci[p];
                if (j <= i || j > k) continue;
                if (find_block(A.rp, A.ci, i, j) < 0) continue;   // rowInd[j].contains(i)
                int pji = find_block(lrp, lci, j, i);
                if (pji < 0) continue;
                std::memcpy(ljT.data(), lv + (size_t)pji * b2, sizeof(double) * b2);
                transpose(bs, ljT.data());
                matmat(bs, lki, ljT.data(), mm.data());
                double* lkj = lv + (size_t)p * b2;
                for (int e = 0; e < b2; ++e) lkj[e] -= mm[e];
            }
        }
    }
    return nz;
}

void ichol0_apply_impl(int n, int bs, const int* lrp, const int* lci, const double* lv, double* x) {
    const int b2 = bs * bs;
    std::vector<double> t(bs);
    // column lists of L (rows > col), ascending row order
    std::vector<std::vector<std::pair<int, int>>> colrows(n);
    for (int i = 0; i < n; ++i)
        for (int p = lrp[i]; p < lrp[i + 1]; ++p)
            if (lci[p] < i) colrows[lci[p]].push_back({i, p});
    for (int i = 0; i < n; ++i) {
        double* xi = x + (size_t)i * bs;
        int pd = -1;
        for (int p = lrp[i]; p < lrp[i + 1]; ++p) {
            int c = lci[p];
            if (c == i) pd = p;
            if (c >= i) continue;
            matvec(bs, lv + (size_t)p * b2, x + (size_t)c * bs, t.data());
            for (int r = 0; r < bs; ++r) xi[r] -= t[r];
        }
        lii_fwd_vec(bs, lv + (size_t)pd * b2, xi);
    }
    for (int i = n - 1; i >= 0; --i) {
        double* xi = x + (size_t)i * bs;
        for (auto& rp : colrows[i]) {
            matvec_t(bs, lv + (size_t)rp.second * b2, x + (size_t)rp.first * bs, t.data());
            for (int r = 0; r < bs; ++r) xi[r] -= t[r];
        }
        int pd = find_block(lrp, lci, i, i);
        liiT_bwd_vec(bs, lv + (size_t)pd * b2, xi);
    }
}

// Preconditioner object used by the Krylov oracles.
struct Precond {
    orc_precond cfg{};
    Bsr A{};
    DiagInv di;
    std::vector<double> lu, dlu;
    std::vector<int> dpiv, lrp, lci;
    std::vector<double> lv;
    void build(const Bsr& a, const orc_precond* c) {
        A = a;
        if (c) cfg = *c;
        const int b2 = A.bs * A.bs;
        if (cfg.kind == ORC_PRE_DIAG || cfg.kind == ORC_PRE_SOR) di.build(A, cfg.inv_kind);
        if (cfg.kind == ORC_PRE_ILU0) {
            lu.resize((size_t)A.rp[A.n] * b2);
            dlu.resize((size_t)A.n * b2);
            dpiv.resize((size_t)A.n * A.bs);
            ilu0_factor_impl(A, lu.data(), dlu.data(), dpiv.data());
        }
        if (cfg.kind == ORC_PRE_ICHOL0) {
            lrp.resize(A.n + 1);
            lci.resize(A.rp[A.n]);
            lv.resize((size_t)A.rp[A.n] * b2);
            ichol0_factor_impl(A, lrp.data(), lci.data(), lv.data());
        }
    }
    void apply(double* r) const {
        const size_t len = (size_t)A.n * A.bs;
        switch (cfg.kind) {
            case ORC_PRE_NONE: return;
            case ORC_PRE_DIAG:
                for (int i = 0; i < A.n; ++i) di.solve(i, r + (size_t)i * A.bs);
                return;
            case ORC_PRE_SOR: {
                std::vector<double> x(len, 0.0);
                if (cfg.type < 0) {
                    for (int s = 0; s < cfg.nu; ++s) sor_sweep(A, di, r, x.data(), cfg.w, false);
                } else if (cfg.type == 0) {
                    for (int s = 0; s < cfg.nu; ++s) sor_sweep(A, di, r, x.data(), cfg.w, true);
                    for (int s = 0; s < cfg.nu; ++s) sor_sweep(A, di, r, x.data(), cfg.w, false);
                } else {
                    for (int s = 0; s < cfg.nu; ++s) sor_sweep(A, di, r, x.data(), cfg.w, true);
                }
                std::memcpy(r, x.data(), sizeof(double) * len);
                return;
            }
            case ORC_PRE_ILU0:
                ilu0_apply_impl(A.n, A.bs, A.rp, A.ci, lu.data(), dlu.data(), dpiv.data(), r);
                return;
            case ORC_PRE_ICHOL0:
                ichol0_apply_impl(A.n, A.bs, lrp.data(), lci.data(), lv.data(), r);
                return;
        }
    }
};

// ---------------------------------------------------------------------------
// Small dense algebra for GMRES (gels, Schur, reordering)
// ---------------------------------------------------------------------------

// Householder least squares, m >= n, H row-major m x n (gels 'N').
void gels_impl(int m, int n, const double* Hin, double* y) {
    std::vector<double> H(Hin, Hin + (size_t)m * n);
    std::vector<double> v(m);
    for (int j = 0; j < n; ++j) {
        double nrm = 0.0;
        for (int i = j; i < m; ++i) nrm += H[i * n + j] * H[i * n + j];
        nrm = std::sqrt(nrm);
        if (nrm == 0.0) continue;
        double alpha = H[j * n + j] > 0 ? -nrm : nrm;
        for (int i = 0; i < m; ++i) v[i] = 0.0;
        v[j] = H[j * n + j] - alpha;
        for (int i = j + 1; i < m; ++i) v[i] = H[i * n + j];
        double vv = 0.0;
        for (int i = j; i < m; ++i) vv += v[i] * v[i];
        if (vv == 0.0) continue;
        for (int c = j; c < n; ++c) {
            double s = 0.0;
            for (int i = j; i < m; ++i) s += v[i] * H[i * n + c];
            s = 2.0 * s / vv;
            for (int i = j; i < m; ++i) H[i * n + c] -= s * v[i];
        }
        double s = 0.0;
        for (int i = j; i < m; ++i) s += v[i] * y[i];
        s = 2.0 * s / vv;
        for (int i = j; i < m; ++i) y[i] -= s * v[i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = y[i];
        for (int c = i + 1; c < n; ++c) s -= H[i * n + c] * y[c];
        y[i] = s / H[i * n + i];
    }
}

// Complex Givens (zlartg semantics): [c s; -conj(s) c] [f; g] = [r; 0], c real.
void zlartg(cd f, cd g, double& c, cd& s) {
    if (g == cd(0)) { c = 1; s = 0; return; }
    if (f == cd(0)) { c = 0; s = std::conj(g) / std::abs(g); return; }
    double af = std::abs(f), ag = std::abs(g);
    double nrm = std::hypot(af, ag);
    c = af / nrm;
    s = (f / af) * std::conj(g) / nrm;
}

// Complex Schur form T = Z^H A Z of a real n x n matrix (row-major):
// Householder Hessenberg reduction + shifted QR with Wilkinson shifts
// (Golub & Van Loan 7.4-7.5, complex single-shift variant).
void complex_schur(int n, const double* A, std::vector<cd>& T, std::vector<cd>& Z) {
    T.assign((size_t)n * n, cd(0));
    Z.assign((size_t)n * n, cd(0));
    for (int i = 0; i < n * n; ++i) T[i] = A[i];
    for (int i = 0; i < n; ++i) Z[i * n + i] = 1.0;
    // Hessenberg reduction (real Householder, applied in complex storage)
    std::vector<cd> v(n);
    for (int k = 0; k < n - 2; ++k) {
        double nrm = 0.0;
        for (int i = k + 1; i < n; ++i) nrm += std::norm(T[i * n + k]);
        nrm = std::sqrt(nrm);
        if (nrm == 0.0) continue;
        cd x0 = T[(k + 1) * n + k];
        double ax0 = std::abs(x0);
        cd alpha = ax0 == 0.0 ? cd(-nrm) : -(x0 / ax0) * nrm;
        for (int i = 0; i < n; ++i) v[i] = 0;
        v[k + 1] = x0 - alpha;
        for (int i = k + 2; i < n; ++i) v[i] = T[i * n + k];
        double vv = 0.0;
        for (int i = k + 1; i < n; ++i) vv += std::norm(v[i]);
        if (vv == 0.0) continue;
        for (int c = 0; c < n; ++c) {      // T <- (I - 2vv^H/vv) T
            cd s = 0;
            for (int i = k + 1; i < n; ++i) s += std::conj(v[i]) * T[i * n + c];
            s *= 2.0 / vv;
            for (int i = k + 1; i < n; ++i) T[i * n + c] -= s * v[i];
        }
        for (int r = 0; r < n; ++r) {      // T <- T (I - 2vv^H/vv), Z likewise
            cd s = 0, sz = 0;
            for (int i = k + 1; i < n; ++i) { s += T[r * n + i] * v[i]; sz += Z[r * n + i] * v[i]; }
            s *= 2.0 / vv; sz *= 2.0 / vv;
            for (int i = k + 1; i < n; ++i) {
                T[r * n + i] -= s * std::conj(v[i]);
                Z[r * n + i] -= sz * std::conj(v[i]);
            }
        }
        for (int i = k + 2; i < n; ++i) T[i * n + k] = 0;
    }
    // Shifted QR on the active window [lo, hi]
    int hi = n - 1, iter = 0;
    std::vector<double> cs(n);
    std::vector<cd> sn(n);
    while (hi > 0) {
        int lo = hi;
        while (lo > 0) {
            double s = std::abs(T[(lo - 1) * n + lo - 1]) + std::abs(T[lo * n + lo]);
            if (s == 0.0) s = 1.0;
            if (std::abs(T[lo * n + lo - 1]) <= kEps * s) { T[lo * n + lo - 1] = 0; break; }
            --lo;
        }
        if (lo == hi) { --hi; iter = 0; continue; }
        ++iter;
        cd mu;
        if (iter % 11 == 10) {
            mu = T[hi * n + hi] + 0.75 * std::abs(T[hi * n + hi - 1]);   // exceptional shift
        } else {
            cd a = T[(hi - 1) * n + hi - 1], b = T[(hi - 1) * n + hi];
            cd c = T[hi * n + hi - 1], d = T[hi * n + hi];
            cd tr = a + d, det = a * d - b * c;
            cd disc = std::sqrt(tr * tr / 4.0 - det);
            cd l1 = tr / 2.0 + disc, l2 = tr / 2.0 - disc;
            mu = std::abs(l1 - d) < std::abs(l2 - d) ? l1 : l2;
        }
        for (int i = lo; i <= hi; ++i) T[i * n + i] -= mu;
        for (int i = lo; i < hi; ++i) {
            zlartg(T[i * n + i], T[(i + 1) * n + i], cs[i], sn[i]);
            for (int c = i; c < n; ++c) {
                cd x = T[i * n + c], y = T[(i + 1) * n + c];
                T[i * n + c] = cs[i] * x + sn[i] * y;
                T[(i + 1) * n + c] = cs[i] * y - std::conj(sn[i]) * x;
            }
            T[(i + 1) * n + i] = 0;
        }
        for (int i = lo; i < hi; ++i) {
            int rmax = std::min(i + 2, hi);
            for (int r = 0; r <= rmax; ++r) {
                cd x = T[r * n + i], y = T[r * n + i + 1];
                T[r * n + i] = cs[i] * x + std::conj(sn[i]) * y;
                T[r * n + i + 1] = cs[i] * y - sn[i] * x;
            }
            for (int r = 0; r < n; ++r) {
                cd x = Z[r * n + i], y = Z[r * n + i + 1];
                Z[r * n + i] = cs[i] * x + std::conj(sn[i]) * y;
                Z[r * n + i + 1] = cs[i] * y - sn[i] * x;
            }
        }
        for (int i = lo; i <= hi; ++i) T[i * n + i] += mu;
        if (iter > 100 * n) break;   // no convergence: keep what we have
    }
}

// Swap diagonal entries k and k+1 of the upper-triangular T (ztrexc step).
void schur_swap(int n, std::vector<cd>& T, std::vector<cd>& Z, int k) {
    cd t11 = T[k * n + k], t22 = T[(k + 1) * n + k + 1];
    double c;
    cd s;
    zlartg(T[k * n + k + 1], t22 - t11, c, s);
    for (int col = k + 2; col < n; ++col) {
        cd x = T[k * n + col], y = T[(k + 1) * n + col];
        T[k * n + col] = c * x + s * y;
        T[(k + 1) * n + col] = c * y - std::conj(s) * x;
    }
    for (int r = 0; r < k; ++r) {
        cd x = T[r * n + k], y = T[r * n + k + 1];
        T[r * n + k] = c * x + std::conj(s) * y;
        T[r * n + k + 1] = c * y - s * x;
    }
    T[k * n + k] = t22;
    T[(k + 1) * n + k + 1] = t11;
    for (int r = 0; r < n; ++r) {
        cd x = Z[r * n + k], y = Z[r * n + k + 1];
        Z[r * n + k] = c * x + std::conj(s) * y;
        Z[r * n + k + 1] = c * y - s * x;
    }
}

// Selection of the k smallest harmonic Ritz values with conjugate pairs kept
// together (deflated_restart.hpp:124-165).  Eigenvalues are ordered by modulus
// with a stable sort.
std::vector<int> select_smallest(int n, const std::vector<double>& wr, const std::vector<double>& wi,
                                 int k) {
    double zero_thresh = 10.0 * n * n * kEps;
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int i, int j) {
        return std::sqrt(wr[i] * wr[i] + wi[i] * wi[i]) < std::sqrt(wr[j] * wr[j] + wi[j] * wi[j]);
    });
    std::vector<int> sel(n, 0);
    int cnt = 0;
    for (int i = 0; i < n && cnt < k; ++i) {
        int ip = perm[i];
        if (sel[ip]) continue;
        sel[ip] = 1;
        ++cnt;
        if (std::fabs(wi[ip]) > zero_thresh) {
            for (int j = 0; j < n; ++j)
                if (j != ip && std::fabs(wr[j] - wr[ip]) < zero_thresh &&
                    std::fabs(wi[j] + wi[ip]) < zero_thresh) {
                    if (!sel[j]) { sel[j] = 1; ++cnt; }
                    break;
                }
        }
    }
    return sel;
}

// Orthonormal real basis P (n x kout) of the invariant subspace of Ht belonging
// to the selected eigenvalues: reorder the complex Schur form so they lead,
// then extract a real basis from span{Re z_j, Im z_j} (pivoted Gram-Schmidt).
// Same subspace as gees + trsen + geqrf/orgqr (deflated_restart.hpp:97-193).
int harmonic_subspace(int n, const double* Ht, int k, std::vector<double>& P,
                      std::vector<double>& wr, std::vector<double>& wi) {
    std::vector<cd> T, Z;
    complex_schur(n, Ht, T, Z);
    wr.resize(n); wi.resize(n);
    for (int i = 0; i < n; ++i) { wr[i] = T[i * n + i].real(); wi[i] = T[i * n + i].imag(); }
    std::vector<int> sel = select_smallest(n, wr, wi, k);
    int sdim = 0;
    for (int i = 0; i < n; ++i) sdim += sel[i];
    // bubble selected entries to the front, preserving their relative order
    std::vector<int> tag(sel);
    for (int dst = 0; dst < sdim; ++dst) {
        int src = dst;
        while (!tag[src]) ++src;
        for (int j = src - 1; j >= dst; --j) {
            schur_swap(n, T, Z, j);
            std::swap(tag[j], tag[j + 1]);
        }
    }
    // candidate real vectors
    std::vector<std::vector<double>> cand;
    for (int j = 0; j < sdim; ++j) {
        std::vector<double> re(n), im(n);
        for (int r = 0; r < n; ++r) { re[r] = Z[r * n + j].real(); im[r] = Z[r * n + j].imag(); }
        cand.push_back(re);
        cand.push_back(im);
    }
    P.assign((size_t)n * sdim, 0.0);
    std::vector<char> used(cand.size(), 0);
    for (int col = 0; col < sdim; ++col) {
        int best = -1;
        double bn = -1.0;
        for (size_t c = 0; c < cand.size(); ++c) {
            if (used[c]) continue;
            double s = 0.0;
            for (int r = 0; r < n; ++r) s += cand[c][r] * cand[c][r];
            if (s > bn) { bn = s; best = (int)c; }
        }
        used[best] = 1;
        std::vector<double> q = cand[best];
        double nq = std::sqrt(bn);
        for (int r = 0; r < n; ++r) q[r] /= nq;
        for (int r = 0; r < n; ++r) P[(size_t)r * sdim + col] = q[r];
        for (size_t c = 0; c < cand.size(); ++c) {
            if (used[c]) continue;
            for (int pass = 0; pass < 2; ++pass) {
                double d = 0.0;
                for (int r = 0; r < n; ++r) d += q[r] * cand[c][r];
                for (int r = 0; r < n; ++r) cand[c][r] -= d * q[r];
            }
        }
    }
    return sdim;
}

// deflated_restart (deflated_restart.hpp:38-269) on the GMRES state.
// H: m x n row-major (m = n+1); Qn: m vectors.  On return H is (k+1) x k,
// Qn has k+1 vectors, c_r has k+1 entries.  Returns the new k.
int deflated_restart_impl(std::vector<double>& H, int& m, int& n, std::vector<std::vector<double>>& Qn,
                          const std::vector<double>& res_vec, std::vector<double>& c_r, int k) {
    if (k > n) k = n;
    if (k == 0) {
        H.clear(); m = n = 0;
        Qn.clear();
        c_r.assign(1, 0.0);
        return 0;
    }
    // (1) H_tilde = H_n + h_mn^2 H_n^{-T} e_n e_n^T
    std::vector<double> Hn((size_t)n * n), HnT((size_t)n * n), w(n, 0.0);
    std::vector<int> piv(n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) { Hn[i * n + j] = H[i * n + j]; HnT[j * n + i] = H[i * n + j]; }
    double h_mn = H[(size_t)n * n + n - 1];
    w[n - 1] = 1.0;
    getrf(n, HnT.data(), piv.data());
    getrs_n(n, HnT.data(), piv.data(), w.data());
    double scale = h_mn * h_mn;
    for (int i = 0; i < n; ++i) Hn[i * n + n - 1] += scale * w[i];
    // (2)-(3) harmonic Ritz subspace
    std::vector<double> Pk, wr, wi;
    k = harmonic_subspace(n, Hn.data(), k, Pk, wr, wi);
    // (4) P_{k+1}: [P_k; 0] extended with res_vec, Gram-Schmidt
    const int kp = k + 1;
    std::vector<double> P((size_t)(n + 1) * kp, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < k; ++j) P[i * kp + j] = Pk[i * k + j];
    for (int i = 0; i < n + 1; ++i) P[i * kp + k] = res_vec[i];
    for (int i = 0; i < k; ++i) {
        double d = 0.0;
        for (int j = 0; j < n + 1; ++j) d += P[j * kp + i] * P[j * kp + k];
        for (int j = 0; j < n + 1; ++j) P[j * kp + k] -= d * P[j * kp + i];
    }
    double nrm = 0.0;
    for (int j = 0; j < n + 1; ++j) nrm += P[j * kp + k] * P[j * kp + k];
    nrm = std::sqrt(nrm);
    for (int j = 0; j < n + 1; ++j) P[j * kp + k] /= nrm;
    // (5) c_r = P_{k+1}^T res_vec
    c_r.assign(kp, 0.0);
    for (int i = 0; i < kp; ++i)
        for (int j = 0; j < n + 1; ++j) c_r[i] += P[j * kp + i] * res_vec[j];
    // (6) H_new = P_{k+1}^T (H P_k)
    std::vector<double> HP((size_t)(n + 1) * k, 0.0), Hnew((size_t)kp * k, 0.0);
    for (int i = 0; i < n + 1; ++i)
        for (int l = 0; l < n; ++l)
            for (int j = 0; j < k; ++j) HP[i * k + j] += H[i * n + l] * Pk[l * k + j];
    for (int l = 0; l < n + 1; ++l)
        for (int i = 0; i < kp; ++i)
            for (int j = 0; j < k; ++j) Hnew[i * k + j] += P[l * kp + i] * HP[l * k + j];
    // (7) Qn_new = Qn P_{k+1}
    const size_t len = Qn[0].size();
    std::vector<std::vector<double>> Qnew(kp, std::vector<double>(len, 0.0));
    for (int im = 0; im < n + 1; ++im)
        for (int ik = 0; ik < kp; ++ik) {
            double pv = P[im * kp + ik];
            const double* q = Qn[im].data();
            double* o = Qnew[ik].data();
            for (size_t il = 0; il < len; ++il) o[il] += q[il] * pv;
        }
    H = Hnew; m = kp; n = k;
    Qn.swap(Qnew);
    return k;
}

}  // namespace

// ===========================================================================
// extern "C" surface
// ===========================================================================

extern "C" {

void orc_bsmv(int n, int bs, const int* rowptr, const int* colidx, const double* vals,
              const double* x, double* y, int accumulate) {
    Bsr A{n, bs, rowptr, colidx, vals};
    bsmv_impl(A, x, y, accumulate != 0);
}

void orc_bsmv_dim(int n, int B, int D, int dim, const int* rowptr, const int* colidx,
                  const double* vals, const double* x, double* y, int accumulate) {
    std::vector<double> ph(B);
    const int stride = B * D;
    for (int i = 0; i < n; ++i) {
        for (int d = 0; d < D; ++d) {
            if (dim >= 0 && d != dim) continue;
            double* yi = y + (size_t)i * stride + d * B;
            if (!accumulate)
                for (int r = 0; r < B; ++r) yi[r] = 0.0;
        }
        for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            for (int d = 0; d < D; ++d) {
                if (dim >= 0 && d != dim) continue;
                matvec(B, vals + (size_t)p * B * B, x + (size_t)colidx[p] * stride + d * B, ph.data());
                double* yi = y + (size_t)i * stride + d * B;
                for (int r = 0; r < B; ++r) yi[r] += ph[r];
            }
        }
    }
}

void orc_bsmv_transpose(int nrows, int ncols, int bs, const int* rowptr, const int* colidx,
                        const double* vals, const double* x, double* y, int accumulate) {
    if (!accumulate)
        for (size_t i = 0; i < (size_t)ncols * bs; ++i) y[i] = 0.0;
    std::vector<double> ph(bs);
    for (int i = 0; i < nrows; ++i)
        for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
            matvec_t(bs, vals + (size_t)p * bs * bs, x + (size_t)i * bs, ph.data());
            double* yc = y + (size_t)colidx[p] * bs;
            for (int r = 0; r < bs; ++r) yc[r] += ph[r];
        }
}

void orc_lu_factor(int bs, const double* a, double* lu, int* piv) { lu_factor(bs, a, lu, piv); }
void orc_lu_solve(int bs, const double* lu, const int* piv, double* x) { lu_solve(bs, lu, piv, x); }
void orc_lu_rsolve(int bs, const double* lu, const int* piv, double* X) { lu_rsolve(bs, lu, piv, X); }
void orc_chol_factor(int bs, const double* a, double* packed) { chol_factor(bs, a, packed); }
void orc_chol_solve(int bs, const double* packed, double* x) { chol_solve(bs, packed, x); }

void orc_ilu0_factor(int n, int bs, const int* rowptr, const int* colidx, const double* vals,
                     double* lu_vals, double* dlu, int* dpiv) {
    Bsr A{n, bs, rowptr, colidx, vals};
    ilu0_factor_impl(A, lu_vals, dlu, dpiv);
}

void orc_ilu0_apply(int n, int bs, const int* rowptr, const int* colidx, const double* lu_vals,
                    const double* dlu, const int* dpiv, double* x) {
    ilu0_apply_impl(n, bs, rowptr, colidx, lu_vals, dlu, dpiv, x);
}

int orc_ichol0_factor(int n, int bs, const int* rowptr, const int* colidx, const double* vals,
                      int* l_rowptr, int* l_colidx, double* l_vals) {
    Bsr A{n, bs, rowptr, colidx, vals};
    return ichol0_factor_impl(A, l_rowptr, l_colidx, l_vals);
}

void orc_ichol0_apply(int n, int bs, const int* l_rowptr, const int* l_colidx,
                      const double* l_vals, double* x) {
    ichol0_apply_impl(n, bs, l_rowptr, l_colidx, l_vals, x);
}

void orc_sor_apply(int n, int bs, const int* rowptr, const int* colidx, const double* vals,
                   double w, int nu, int type, int inv_kind, double* x) {
    Bsr A{n, bs, rowptr, colidx, vals};
    orc_precond c{ORC_PRE_SOR, inv_kind, w, nu, type};
    Precond P;
    P.build(A, &c);
    P.apply(x);
}

void orc_diag_apply(int n, int bs, const int* rowptr, const int* colidx, const double* vals,
                    int inv_kind, double* x) {
    Bsr A{n, bs, rowptr, colidx, vals};
    orc_precond c{ORC_PRE_DIAG, inv_kind, 0.0, 0, 0};
    Precond P;
    P.build(A, &c);
    P.apply(x);
}

int orc_levels(int n, const int* rowptr, const int* colidx, int lower, int* level) {
    int nlev = 0;
    if (lower) {
        for (int i = 0; i < n; ++i) {
            int l = 0;
            for (int p = rowptr[i]; p < rowptr[i + 1]; ++p)
                if (colidx[p] < i) l = std::max(l, level[colidx[p]] + 1);
            level[i] = l;
            nlev = std::max(nlev, l + 1);
        }
    } else {
        for (int i = n - 1; i >= 0; --i) {
            int l = 0;
            for (int p = rowptr[i]; p < rowptr[i + 1]; ++p)
                if (colidx[p] > i) l = std::max(l, level[colidx[p]] + 1);
            level[i] = l;
            nlev = std::max(nlev, l + 1);
        }
    }
    return nlev;
}

int orc_pcg(int n, int bs, const int* rowptr, const int* colidx, const double* vals,
            const orc_precond* pre, const double* b, double* x, double tol, int maxIters,
            orc_stats* st) {
    Bsr A{n, bs, rowptr, colidx, vals};
    Precond M;
    M.build(A, pre);
    const size_t len = (size_t)n * bs;
    std::vector<double> r(b, b + len), p(r), z, Ap(len);
    M.apply(p.data());
    z = p;
    for (size_t i = 0; i < len; ++i) x[i] = 0.0;
    double res = std::sqrt(vdot(n, bs, p.data(), p.data()));
    const double res0 = res;
    const double tolerance = res * tol;
    int iter = 0;
    while (res > tolerance) {
        bsmv_impl(A, p.data(), Ap.data(), false);
        double rz = vdot(n, bs, r.data(), z.data());
        double alpha = rz / vdot(n, bs, p.data(), Ap.data());
        for (size_t i = 0; i < len; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * Ap[i];
        }
        res = std::sqrt(vdot(n, bs, r.data(), r.data()));
        if (res > tolerance) {
            z = r;
            M.apply(z.data());
            double beta = vdot(n, bs, r.data(), z.data()) / rz;
            for (size_t i = 0; i < len; ++i) p[i] = z[i] + beta * p[i];
            ++iter;
            if (iter > maxIters) break;
        }
        if (!std::isfinite(res)) break;
    }
    if (st) {
        st->iters = iter;
        st->restarts = 0;
        st->res0 = res0;
        st->res_est = res;
        std::vector<double> ax(len);
        bsmv_impl(A, x, ax.data(), false);
        for (size_t i = 0; i < len; ++i) ax[i] = b[i] - ax[i];
        st->true_relres = vnorm(n, bs, ax.data()) / vnorm(n, bs, b);
    }
    return iter > maxIters ? 1 : 0;
}

int orc_pgmres(int n, int bs, const int* rowptr, const int* colidx, const double* vals,
               const orc_precond* pre, const double* b, double* x, double tol, int maxIters,
               int restart, int k, orc_stats* st) {
    Bsr A{n, bs, rowptr, colidx, vals};
    Precond M;
    M.build(A, pre);
    const size_t len = (size_t)n * bs;
    std::vector<double> x0(len, 0.0), r0(b, b + len), v(len);
    M.apply(r0.data());
    std::vector<double> res_mag;
    res_mag.push_back(vnorm(n, bs, r0.data()));
    std::vector<std::vector<double>> Qn;
    {
        std::vector<double> q0(r0);
        for (auto& e : q0) e /= res_mag.back();
        Qn.push_back(q0);
    }
    const double zero_thresh = std::min(res_mag.back() * kEps, kEps);
    const double tolerance = res_mag.back() * tol;
    const double res0 = res_mag.back();
    std::vector<double> H;   // m x n row-major
    int Hm = 0, Hn = 0;
    int iter = 0, totIters = 0, restarts = 0;
    std::vector<double> y, c_r, res_vec;
    bool restarted = false, breakdown = false;
    while (res_mag.back() > tolerance) {
        // one_arnoldi (solvers.hpp:109-136)
        {
            int m = iter + 2, nn = iter + 1;
            std::vector<double> Hnew((size_t)m * nn, 0.0);
            for (int i = 0; i < std::min(Hm, m); ++i)
                for (int j = 0; j < std::min(Hn, nn); ++j) Hnew[i * nn + j] = H[i * Hn + j];
            H.swap(Hnew); Hm = m; Hn = nn;
            bsmv_impl(A, Qn[iter].data(), v.data(), false);
            M.apply(v.data());
            for (int j = 0; j <= iter; ++j) {
                double h = vdot(n, bs, Qn[j].data(), v.data());
                H[j * Hn + iter] = h;
                const double* q = Qn[j].data();
                for (size_t e = 0; e < len; ++e) v[e] += -h * q[e];
            }
            double h = vnorm(n, bs, v.data());
            H[(iter + 1) * Hn + iter] = h;
            if (std::fabs(h) < zero_thresh) {
                breakdown = true;
            } else {
                std::vector<double> q(v);
                for (auto& e : q) e /= h;
                Qn.push_back(q);
            }
        }
        // least squares (solvers.hpp:192-232)
        int m = iter + 2, nn = iter + 1;
        res_vec.assign(m, 0.0);
        std::vector<double> c(m, 0.0);
        if (!restarted) c[0] = res_mag[0];
        else
            for (int i = 0; i < (int)c_r.size() && i < m; ++i) c[i] = c_r[i];
        y = c;
        gels_impl(m, nn, H.data(), y.data());
        y.resize(nn);
        for (int i = 0; i < m; ++i) {
            double s = 0.0;
            for (int j = 0; j < nn; ++j) s += -1.0 * H[i * nn + j] * y[j];
            res_vec[i] = s + c[i];
        }
        double r = 0.0;
        for (double val : res_vec) r += std::fabs(val * val);
        res_mag.push_back(std::sqrt(r));
        ++iter; ++totIters;
        if (totIters > maxIters || breakdown || !std::isfinite(res_mag.back())) break;
        if (iter == restart && res_mag.back() > tolerance) {
            for (size_t iy = 0; iy < y.size(); ++iy) {
                const double* q = Qn[iy].data();
                for (size_t e = 0; e < len; ++e) x0[e] += y[iy] * q[e];
            }
            // r_0 = M^{-1}(b - A x_0) (solvers.hpp:248); only used by a pure restart
            std::vector<double> rr(len);
            bsmv_impl(A, x0.data(), rr.data(), false);
            for (size_t e = 0; e < len; ++e) rr[e] = b[e] - rr[e];
            M.apply(rr.data());
            int knew = deflated_restart_impl(H, Hm, Hn, Qn, res_vec, c_r, k);
            {
                double last = res_mag.back();
                res_mag.assign(1, last);
            }
            ++restarts;
            if (k == 0) {
                res_mag.clear();
                res_mag.push_back(vnorm(n, bs, rr.data()));
                c_r.assign(1, res_mag.back());
                for (auto& e : rr) e /= res_mag.back();
                Qn.push_back(rr);
                iter = 0;
                restarted = false;
            } else {
                iter = knew;
                restarted = true;
            }
            y.clear();
        }
    }
    for (size_t e = 0; e < len; ++e) x[e] = x0[e];
    for (size_t iy = 0; iy < y.size(); ++iy) {
        const double* q = Qn[iy].data();
        for (size_t e = 0; e < len; ++e) x[e] += y[iy] * q[e];
    }
    if (st) {
        st->iters = totIters;
        st->restarts = restarts;
        st->res0 = res0;
        st->res_est = res_mag.back();
        std::vector<double> ax(len);
        bsmv_impl(A, x, ax.data(), false);
        for (size_t i = 0; i < len; ++i) ax[i] = b[i] - ax[i];
        st->true_relres = vnorm(n, bs, ax.data()) / vnorm(n, bs, b);
    }
    return totIters > maxIters ? 1 : 0;
}

int orc_harmonic_subspace(int n, const double* H, int k, double* P, double* eig_re,
                          double* eig_im) {
    // H is (n+1) x n; form H_tilde exactly as deflated_restart step (1).
    std::vector<double> Hn((size_t)n * n), HnT((size_t)n * n), w(n, 0.0);
    std::vector<int> piv(n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) { Hn[i * n + j] = H[i * n + j]; HnT[j * n + i] = H[i * n + j]; }
    double h_mn = H[(size_t)n * n + n - 1];
    w[n - 1] = 1.0;
    getrf(n, HnT.data(), piv.data());
    getrs_n(n, HnT.data(), piv.data(), w.data());
    for (int i = 0; i < n; ++i) Hn[i * n + n - 1] += h_mn * h_mn * w[i];
    std::vector<double> Pk, wr, wi;
    int kout = harmonic_subspace(n, Hn.data(), k, Pk, wr, wi);
    std::memcpy(P, Pk.data(), sizeof(double) * Pk.size());
    for (int i = 0; i < n; ++i) { eig_re[i] = wr[i]; eig_im[i] = wi[i]; }
    return kout;
}

void orc_eigvals(int n, const double* A, double* re, double* im) {
    std::vector<cd> T, Z;
    complex_schur(n, A, T, Z);
    for (int i = 0; i < n; ++i) { re[i] = T[i * n + i].real(); im[i] = T[i * n + i].imag(); }
}

void orc_gels(int m, int n, const double* H, double* y) { gels_impl(m, n, H, y); }

int orc_bsmm(int nA, int bs, const int* a_rowptr, const int* a_colidx, const double* a_vals,
             const int* b_rowptr, const int* b_colidx, const double* b_vals,
             int* c_rowptr, int* c_colidx, double* c_vals, int cap) {
    const int b2 = bs * bs;
    std::vector<int> cols;
    std::vector<double> acc, mm(b2);
    int nz = 0;
    // First pass: structure (first-touch order per row, multiply.hpp:85-97).
    std::vector<std::vector<int>> rowcols(nA);
    for (int i = 0; i < nA; ++i) {
        auto& rc = rowcols[i];
        for (int pa = a_rowptr[i]; pa < a_rowptr[i + 1]; ++pa) {
            int ka = a_colidx[pa];
            for (int pb = b_rowptr[ka]; pb < b_rowptr[ka + 1]; ++pb) {
                int j = b_colidx[pb];
                if (std::find(rc.begin(), rc.end(), j) == rc.end()) rc.push_back(j);
            }
        }
        nz += (int)rc.size();
    }
    if (nz > cap) return -nz;
    int pos = 0;
    c_rowptr[0] = 0;
    for (int i = 0; i < nA; ++i) {
        auto& rc = rowcols[i];
        int base = pos;
        for (size_t t = 0; t < rc.size(); ++t) {
            c_colidx[base + (int)t] = rc[t];
            for (int e = 0; e < b2; ++e) c_vals[(size_t)(base + t) * b2 + e] = 0.0;
        }
        for (int pa = a_rowptr[i]; pa < a_rowptr[i + 1]; ++pa) {
            int ka = a_colidx[pa];
            for (int pb = b_rowptr[ka]; pb < b_rowptr[ka + 1]; ++pb) {
                int j = b_colidx[pb];
                int t = (int)(std::find(rc.begin(), rc.end(), j) - rc.begin());
                matmat(bs, a_vals + (size_t)pa * b2, b_vals + (size_t)pb * b2, mm.data());
                double* cv = c_vals + (size_t)(base + t) * b2;
                for (int e = 0; e < b2; ++e) cv[e] += mm[e];
            }
        }
        pos += (int)rc.size();
        c_rowptr[i + 1] = pos;
    }
    return nz;
}

int orc_bsmm_transA(int nA, int mA, int bs, const int* a_rowptr, const int* a_colidx, const double* a_vals,
                    const int* b_rowptr, const int* b_colidx, const double* b_vals,
                    int* c_rowptr, int* c_colidx, double* c_vals, int cap) {
    // multiply.hpp:99-134: for each row i of A, each block A(i, j) and each block B(i, c),
    // C(j, c) += A(i, j)^T B(i, c); C(j, c) is inserted on first touch.
    const int b2 = bs * bs;
    std::vector<std::vector<int>> cols(mA);
    std::vector<std::vector<double>> vals(mA);
    std::vector<double> at(b2), mm(b2);
    for (int i = 0; i < nA; ++i)
        for (int pa = a_rowptr[i]; pa < a_rowptr[i + 1]; ++pa) {
            const int j = a_colidx[pa];
            for (int r = 0; r < bs; ++r)
                for (int c = 0; c < bs; ++c) at[r * bs + c] = a_vals[(size_t)pa * b2 + c * bs + r];
            for (int pb = b_rowptr[i]; pb < b_rowptr[i + 1]; ++pb) {
                const int c = b_colidx[pb];
                auto& rc = cols[j];
                size_t t = (size_t)(std::find(rc.begin(), rc.end(), c) - rc.begin());
                if (t == rc.size()) {
                    rc.push_back(c);
                    vals[j].resize(vals[j].size() + b2, 0.0);
                }
                matmat(bs, at.data(), b_vals + (size_t)pb * b2, mm.data());
                double* cv = vals[j].data() + t * b2;
                for (int e = 0; e < b2; ++e) cv[e] += mm[e];
            }
        }
    int nz = 0;
    for (int j = 0; j < mA; ++j) nz += (int)cols[j].size();
    if (nz > cap) return -nz;
    int pos = 0;
    c_rowptr[0] = 0;
    for (int j = 0; j < mA; ++j) {
        for (size_t t = 0; t < cols[j].size(); ++t, ++pos) {
            c_colidx[pos] = cols[j][t];
            std::memcpy(c_vals + (size_t)pos * b2, vals[j].data() + t * b2, sizeof(double) * b2);
        }
        c_rowptr[j + 1] = pos;
    }
    return nz;
}

}  // extern "C"
