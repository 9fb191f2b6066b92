// bsp_zdense.hpp -- small dense complex algebra of the complex pGMRES (host side): the
// least squares of each step, the LU solve and the deflated restart's harmonic-Ritz
// subspace.  The reference gets these from LAPACK's z-routines (zgels, zgetrf/zgetrs, zgees,
// ztrsen, zgeqrf/zungqr; blocksp_solvers.hpp:192-232, blocksp_deflated_restart.hpp:66-245);
// here they are the textbook algorithms (Golub & Van Loan 5.2 Householder QR, 3.4 partial
// pivoting, 7.5 shifted QR through the complex Schur routine of bsp_dense.hpp).
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <numeric>
#include <vector>

#include "bsp_dense.hpp"

namespace bsp::dense {

// |re| + |im|, the pivot measure of LAPACK's izamax
inline double cabs1(cplx a) { return std::fabs(a.real()) + std::fabs(a.imag()); }

// x <- A^{-1} x for an n x n row-major complex A (LU with partial pivoting); false if singular.
inline bool zsolve(int n, std::vector<cplx> A, std::vector<cplx>& x) {
    for (int j = 0; j < n; ++j) {
        int p = j;
        double best = cabs1(A[(size_t)j * n + j]);
        for (int i = j + 1; i < n; ++i)
            if (cabs1(A[(size_t)i * n + j]) > best) {
                best = cabs1(A[(size_t)i * n + j]);
                p = i;
            }
        if (best == 0.0) return false;
        if (p != j) {
            for (int c = 0; c < n; ++c) std::swap(A[(size_t)j * n + c], A[(size_t)p * n + c]);
            std::swap(x[j], x[p]);
        }
        const cplx d = A[(size_t)j * n + j];
        for (int i = j + 1; i < n; ++i) {
            const cplx l = A[(size_t)i * n + j] / d;
            if (l == cplx(0.0)) continue;
            for (int c = j + 1; c < n; ++c) A[(size_t)i * n + c] -= l * A[(size_t)j * n + c];
            x[i] -= l * x[j];
        }
    }
    for (int i = n - 1; i >= 0; --i) {
        cplx s = x[i];
        for (int c = i + 1; c < n; ++c) s -= A[(size_t)i * n + c] * x[c];
        x[i] = s / A[(size_t)i * n + i];
    }
    return true;
}

// min || c - H y ||_2 for an m x n (m >= n) row-major complex H by Householder QR (the
// problem zgels 'N' solves); y gets n entries.  Returns false on a zero pivot of R.
inline bool zlstsq(int m, int n, const cplx* Hin, const cplx* cin, std::vector<cplx>& y) {
    std::vector<cplx> H(Hin, Hin + (size_t)m * n), c(cin, cin + m), v(m);
    for (int j = 0; j < n; ++j) {
        double nrm2 = 0.0;
        for (int i = j; i < m; ++i) nrm2 += std::norm(H[(size_t)i * n + j]);
        const double nrm = std::sqrt(nrm2);
        if (nrm == 0.0) continue;
        const cplx x0 = H[(size_t)j * n + j];
        const double ax0 = std::abs(x0);
        const cplx alpha = ax0 == 0.0 ? cplx(-nrm) : -(x0 / ax0) * nrm;   // H_jj -> alpha
        std::fill(v.begin(), v.end(), cplx(0.0));
        v[j] = x0 - alpha;
        for (int i = j + 1; i < m; ++i) v[i] = H[(size_t)i * n + j];
        double vv = 0.0;
        for (int i = j; i < m; ++i) vv += std::norm(v[i]);
        if (vv == 0.0) continue;
        const double beta = 2.0 / vv;
        for (int col = j; col < n; ++col) {   // H <- (I - beta v v^H) H
            cplx s = 0.0;
            for (int i = j; i < m; ++i) s += std::conj(v[i]) * H[(size_t)i * n + col];
            s *= beta;
            for (int i = j; i < m; ++i) H[(size_t)i * n + col] -= s * v[i];
        }
        cplx s = 0.0;
        for (int i = j; i < m; ++i) s += std::conj(v[i]) * c[i];
        s *= beta;
        for (int i = j; i < m; ++i) c[i] -= s * v[i];
    }
    y.assign(n, 0.0);
    for (int i = n - 1; i >= 0; --i) {
        cplx s = c[i];
        for (int col = i + 1; col < n; ++col) s -= H[(size_t)i * n + col] * y[col];
        const cplx d = H[(size_t)i * n + i];
        if (d == cplx(0.0)) return false;
        y[i] = s / d;
    }
    return true;
}

// Complex deflated restart (deflated_restart.hpp:38-269, complex branch): H is (n+1) x n
// row-major, res_vec the n+1 residual coefficients of the least squares.  H~ = H_n +
// |h_{n+1,n}|^2 H_n^{-H} e_n e_n^T; its complex Schur form is reordered so that the k
// eigenvalues of smallest modulus (stable order by modulus) lead in their Schur order; P_k =
// those k Schur vectors (re-orthonormalised by two passes of modified Gram-Schmidt, the
// subspace geqrf + ungqr returns); P_{k+1} = [P_k; 0] extended by res_vec made orthonormal to
// it; c_r = P_{k+1}^H res_vec, H_new = P_{k+1}^H H P_k.  Outputs: P ((n+1) x (k+1), row-major,
// the basis rotation Q_new = Q P), Hnew ((k+1) x k), c_r (k+1).  Returns k, or -1 when H_n is
// singular or the Schur iteration fails.
inline int zdeflate(int n, const std::vector<cplx>& H, const std::vector<cplx>& res_vec, int k,
                    std::vector<cplx>& P, std::vector<cplx>& Hnew, std::vector<cplx>& c_r) {
    k = std::min(k, n);
    if (k <= 0) return 0;
    std::vector<cplx> Hn((size_t)n * n), HnH((size_t)n * n), w(n, cplx(0.0));
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            Hn[(size_t)i * n + j] = H[(size_t)i * n + j];
            HnH[(size_t)j * n + i] = std::conj(H[(size_t)i * n + j]);
        }
    const cplx h_mn = H[(size_t)n * n + n - 1];
    w[n - 1] = 1.0;
    if (!zsolve(n, HnH, w)) return -1;
    const cplx scale = std::conj(h_mn) * h_mn;
    for (int i = 0; i < n; ++i) Hn[(size_t)i * n + n - 1] += scale * w[i];
    Schur S;
    if (!S.compute(n, Hn.data())) return -1;
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return std::abs(S.t(a, a)) < std::abs(S.t(b, b)); });
    std::vector<int> tag(n, 0);
    for (int q = 0; q < k; ++q) tag[idx[q]] = 1;
    for (int dst = 0; dst < k; ++dst) {   // selected entries to the front, keeping their order
        int src = dst;
        while (!tag[src]) ++src;
        for (int j = src - 1; j >= dst; --j) {
            S.swap_adjacent(j);
            std::swap(tag[j], tag[j + 1]);
        }
    }
    const int kp = k + 1;
    P.assign((size_t)(n + 1) * kp, cplx(0.0));
    for (int j = 0; j < k; ++j) {
        for (int i = 0; i < n; ++i) P[(size_t)i * kp + j] = S.z(i, j);
        for (int pass = 0; pass < 2; ++pass)
            for (int q = 0; q < j; ++q) {
                cplx d = 0.0;
                for (int i = 0; i < n; ++i) d += std::conj(P[(size_t)i * kp + q]) * P[(size_t)i * kp + j];
                for (int i = 0; i < n; ++i) P[(size_t)i * kp + j] -= d * P[(size_t)i * kp + q];
            }
        double nr = 0.0;
        for (int i = 0; i < n; ++i) nr += std::norm(P[(size_t)i * kp + j]);
        nr = std::sqrt(nr);
        if (!(nr > 0.0)) return -1;
        for (int i = 0; i < n; ++i) P[(size_t)i * kp + j] /= nr;
    }
    for (int i = 0; i <= n; ++i) P[(size_t)i * kp + k] = res_vec[i];
    for (int q = 0; q < k; ++q) {
        cplx d = 0.0;
        for (int i = 0; i <= n; ++i) d += std::conj(P[(size_t)i * kp + q]) * P[(size_t)i * kp + k];
        for (int i = 0; i <= n; ++i) P[(size_t)i * kp + k] -= d * P[(size_t)i * kp + q];
    }
    double nr = 0.0;
    for (int i = 0; i <= n; ++i) nr += std::norm(P[(size_t)i * kp + k]);
    nr = std::sqrt(nr);
    if (!(nr > 0.0)) return -1;
    for (int i = 0; i <= n; ++i) P[(size_t)i * kp + k] /= nr;
    c_r.assign(kp, cplx(0.0));
    for (int q = 0; q < kp; ++q)
        for (int i = 0; i <= n; ++i) c_r[q] += std::conj(P[(size_t)i * kp + q]) * res_vec[i];
    std::vector<cplx> HP((size_t)(n + 1) * k, cplx(0.0));
    for (int i = 0; i <= n; ++i)
        for (int l = 0; l < n; ++l) {
            const cplx h = H[(size_t)i * n + l];
            if (h == cplx(0.0)) continue;
            for (int j = 0; j < k; ++j) HP[(size_t)i * k + j] += h * P[(size_t)l * kp + j];
        }
    Hnew.assign((size_t)kp * k, cplx(0.0));
    for (int l = 0; l <= n; ++l)
        for (int q = 0; q < kp; ++q) {
            const cplx pc = std::conj(P[(size_t)l * kp + q]);
            for (int j = 0; j < k; ++j) Hnew[(size_t)q * k + j] += pc * HP[(size_t)l * k + j];
        }
    return k;
}

// Ritz values of a complex Hessenberg H (m x n, row-major): the eigenvalues of its square part,
// or (harmonic) of H_n + |h_{n+1,n}|^2 H_n^{-H} e_n e_n^T when m > n (printEigInfo,
// blocksp_solvers.hpp:272-288, with the deflated restart's H~).
inline bool zritz_values(int m, int n, const std::vector<cplx>& H, bool harmonic, std::vector<cplx>& ev) {
    std::vector<cplx> Ht((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Ht[(size_t)i * n + j] = H[(size_t)i * n + j];
    if (harmonic && m > n) {
        std::vector<cplx> HnH((size_t)n * n), w(n, cplx(0.0));
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) HnH[(size_t)j * n + i] = std::conj(Ht[(size_t)i * n + j]);
        w[n - 1] = 1.0;
        if (!zsolve(n, HnH, w)) return false;
        const double s = std::norm(H[(size_t)n * n + n - 1]);
        for (int i = 0; i < n; ++i) Ht[(size_t)i * n + n - 1] += s * w[i];
    }
    Schur S;
    if (!S.compute(n, Ht.data())) return false;
    ev.resize(n);
    for (int i = 0; i < n; ++i) ev[i] = S.t(i, i);
    return true;
}

}  // namespace bsp::dense
