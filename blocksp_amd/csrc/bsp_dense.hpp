// bsp_dense.hpp -- small dense kernels of the solve path (host side of the library).
//
// These operate on the tiny matrices of the Krylov solvers (Hessenberg <= 65 x 64)
// and on single diagonal blocks during preconditioner setup.  Methods are the
// textbook ones: Gauss-Jordan with partial pivoting, Cholesky, Householder least
// squares (Golub & Van Loan 5.3), Hessenberg reduction + shifted complex QR with
// Wilkinson shifts (GVL 7.4-7.5) and adjacent-eigenvalue swaps by Givens
// rotations to order a Schur form (GVL 7.6.2).
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <limits>
#include <numeric>
#include <vector>

namespace bsp::dense {

using cplx = std::complex<double>;

// Inverse of a bs x bs row-major block by Gauss-Jordan with partial pivoting.
inline bool invert(int bs, const double* a, double* inv) {
    std::vector<double> m(a, a + bs * bs);
    for (int i = 0; i < bs * bs; ++i) inv[i] = 0.0;
    for (int i = 0; i < bs; ++i) inv[i * bs + i] = 1.0;
    for (int col = 0; col < bs; ++col) {
        int piv = col;
        double best = std::fabs(m[col * bs + col]);
        for (int r = col + 1; r < bs; ++r) {
            double v = std::fabs(m[r * bs + col]);
            if (v > best) { best = v; piv = r; }
        }
        if (best == 0.0 || !std::isfinite(best)) return false;
        if (piv != col)
            for (int c = 0; c < bs; ++c) {
                std::swap(m[col * bs + c], m[piv * bs + c]);
                std::swap(inv[col * bs + c], inv[piv * bs + c]);
            }
        const double d = m[col * bs + col];
        for (int c = 0; c < bs; ++c) { m[col * bs + c] /= d; inv[col * bs + c] /= d; }
        for (int r = 0; r < bs; ++r) {
            if (r == col) continue;
            const double f = m[r * bs + col];
            if (f == 0.0) continue;
            for (int c = 0; c < bs; ++c) {
                m[r * bs + c] -= f * m[col * bs + c];
                inv[r * bs + c] -= f * inv[col * bs + c];
            }
        }
    }
    return true;
}

// In-place lower Cholesky (upper part zeroed); false if not positive definite.
inline bool cholesky_lower(int bs, double* a) {
    for (int j = 0; j < bs; ++j) {
        double s = a[j * bs + j];
        for (int p = 0; p < j; ++p) s -= a[j * bs + p] * a[j * bs + p];
        if (!(s > 0.0)) return false;
        const double d = std::sqrt(s);
        a[j * bs + j] = d;
        for (int r = j + 1; r < bs; ++r) {
            double t = a[r * bs + j];
            for (int p = 0; p < j; ++p) t -= a[r * bs + p] * a[j * bs + p];
            a[r * bs + j] = t / d;
        }
    }
    for (int r = 0; r < bs; ++r)
        for (int c = r + 1; c < bs; ++c) a[r * bs + c] = 0.0;
    return true;
}

inline void matmul(int bs, const double* a, const double* b, double* out) {
    for (int i = 0; i < bs; ++i)
        for (int j = 0; j < bs; ++j) {
            double s = 0.0;
            for (int k = 0; k < bs; ++k) s += a[i * bs + k] * b[k * bs + j];
            out[i * bs + j] = s;
        }
}

inline void transpose_to(int bs, const double* a, double* out) {
    for (int i = 0; i < bs; ++i)
        for (int j = 0; j < bs; ++j) out[j * bs + i] = a[i * bs + j];
}

// min_y || c - H y ||_2 for the m x n (m >= n) row-major H; returns the residual
// norm and writes y[0..n-1].  Householder QR of H applied to c.
inline double lstsq(int m, int n, const double* Hin, const double* c, double* y) {
    std::vector<double> H(Hin, Hin + (size_t)m * n), b(c, c + m), v(m);
    for (int j = 0; j < n; ++j) {
        double nrm = 0.0;
        for (int i = j; i < m; ++i) nrm += H[i * n + j] * H[i * n + j];
        nrm = std::sqrt(nrm);
        if (nrm == 0.0) continue;
        const double alpha = H[j * n + j] > 0 ? -nrm : nrm;
        v[j] = H[j * n + j] - alpha;
        for (int i = j + 1; i < m; ++i) v[i] = H[i * n + j];
        double vv = 0.0;
        for (int i = j; i < m; ++i) vv += v[i] * v[i];
        if (vv == 0.0) continue;
        for (int col = j; col < n; ++col) {
            double s = 0.0;
            for (int i = j; i < m; ++i) s += v[i] * H[i * n + col];
            s *= 2.0 / vv;
            for (int i = j; i < m; ++i) H[i * n + col] -= s * v[i];
        }
        double s = 0.0;
        for (int i = j; i < m; ++i) s += v[i] * b[i];
        s *= 2.0 / vv;
        for (int i = j; i < m; ++i) b[i] -= s * v[i];
    }
    double res = 0.0;
    for (int i = n; i < m; ++i) res += b[i] * b[i];
    for (int i = n - 1; i >= 0; --i) {
        double s = b[i];
        for (int col = i + 1; col < n; ++col) s -= H[i * n + col] * y[col];
        y[i] = s / H[i * n + i];
    }
    return std::sqrt(res);
}

// Solve A^T w = e for a small n x n A (row-major) with partial pivoting.
inline bool solve_transposed(int n, const double* A, const double* e, double* w) {
    std::vector<double> M((size_t)n * n), rhs(e, e + n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) M[i * n + j] = A[j * n + i];
    for (int col = 0; col < n; ++col) {
        int piv = col;
        double best = std::fabs(M[col * n + col]);
        for (int r = col + 1; r < n; ++r)
            if (std::fabs(M[r * n + col]) > best) { best = std::fabs(M[r * n + col]); piv = r; }
        if (best == 0.0) return false;
        if (piv != col) {
            for (int c = 0; c < n; ++c) std::swap(M[col * n + c], M[piv * n + c]);
            std::swap(rhs[col], rhs[piv]);
        }
        for (int r = col + 1; r < n; ++r) {
            const double f = M[r * n + col] / M[col * n + col];
            if (f == 0.0) continue;
            for (int c = col; c < n; ++c) M[r * n + c] -= f * M[col * n + c];
            rhs[r] -= f * rhs[col];
        }
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = rhs[i];
        for (int c = i + 1; c < n; ++c) s -= M[i * n + c] * w[c];
        w[i] = s / M[i * n + i];
    }
    return true;
}

// Givens pair (c real, s complex) with [c s; -conj(s) c] [f; g] = [r; 0].
inline void givens(cplx f, cplx g, double& c, cplx& s) {
    const double ag = std::abs(g);
    if (ag == 0.0) { c = 1.0; s = 0.0; return; }
    const double af = std::abs(f);
    if (af == 0.0) { c = 0.0; s = std::conj(g) / ag; return; }
    const double r = std::hypot(af, ag);
    c = af / r;
    s = (f / af) * std::conj(g) / r;
}

// Complex Schur decomposition A = Z T Z^H of a real n x n matrix.
struct Schur {
    int n = 0;
    std::vector<cplx> T, Z;

    cplx& t(int i, int j) { return T[(size_t)i * n + j]; }
    cplx& z(int i, int j) { return Z[(size_t)i * n + j]; }

    void rot_rows(int i, double c, cplx s, int c0) {   // rows i, i+1, columns >= c0
        for (int col = c0; col < n; ++col) {
            const cplx a = t(i, col), b = t(i + 1, col);
            t(i, col) = c * a + s * b;
            t(i + 1, col) = c * b - std::conj(s) * a;
        }
    }
    void rot_cols(int i, double c, cplx s, int r1) {   // columns i, i+1, rows <= r1 (+ Z)
        for (int row = 0; row <= r1; ++row) {
            const cplx a = t(row, i), b = t(row, i + 1);
            t(row, i) = c * a + std::conj(s) * b;
            t(row, i + 1) = c * b - s * a;
        }
        for (int row = 0; row < n; ++row) {
            const cplx a = z(row, i), b = z(row, i + 1);
            z(row, i) = c * a + std::conj(s) * b;
            z(row, i + 1) = c * b - s * a;
        }
    }

    bool compute(int nn, const double* A) {
        n = nn;
        T.assign((size_t)n * n, 0.0);
        Z.assign((size_t)n * n, 0.0);
        for (int i = 0; i < n * n; ++i) T[i] = A[i];
        for (int i = 0; i < n; ++i) z(i, i) = 1.0;
        // Hessenberg reduction with Householder reflectors
        std::vector<cplx> v(n);
        for (int k = 0; k + 2 < n; ++k) {
            double nrm2 = 0.0;
            for (int i = k + 1; i < n; ++i) nrm2 += std::norm(t(i, k));
            const double nrm = std::sqrt(nrm2);
            if (nrm == 0.0) continue;
            const cplx x0 = t(k + 1, k);
            const double ax0 = std::abs(x0);
            const cplx alpha = ax0 == 0.0 ? cplx(-nrm) : -(x0 / ax0) * nrm;
            std::fill(v.begin(), v.end(), cplx(0.0));
            v[k + 1] = x0 - alpha;
            for (int i = k + 2; i < n; ++i) v[i] = t(i, k);
            double vv = 0.0;
            for (int i = k + 1; i < n; ++i) vv += std::norm(v[i]);
            if (vv == 0.0) continue;
            const double beta = 2.0 / vv;
            for (int col = 0; col < n; ++col) {
                cplx s = 0.0;
                for (int i = k + 1; i < n; ++i) s += std::conj(v[i]) * t(i, col);
                s *= beta;
                for (int i = k + 1; i < n; ++i) t(i, col) -= s * v[i];
            }
            for (int row = 0; row < n; ++row) {
                cplx s = 0.0, sz = 0.0;
                for (int i = k + 1; i < n; ++i) { s += t(row, i) * v[i]; sz += z(row, i) * v[i]; }
                s *= beta;
                sz *= beta;
                for (int i = k + 1; i < n; ++i) {
                    t(row, i) -= s * std::conj(v[i]);
                    z(row, i) -= sz * std::conj(v[i]);
                }
            }
            for (int i = k + 2; i < n; ++i) t(i, k) = 0.0;
        }
        // shifted QR sweeps on the unreduced trailing window
        const double eps = std::numeric_limits<double>::epsilon();
        int hi = n - 1, its = 0, total = 0;
        std::vector<double> cs(n);
        std::vector<cplx> sn(n);
        while (hi > 0) {
            int lo = hi;
            for (; lo > 0; --lo) {
                double scale = std::abs(t(lo - 1, lo - 1)) + std::abs(t(lo, lo));
                if (scale == 0.0) scale = 1.0;
                if (std::abs(t(lo, lo - 1)) <= eps * scale) { t(lo, lo - 1) = 0.0; break; }
            }
            if (lo == hi) { --hi; its = 0; continue; }
            if (++total > 200 * n) return false;
            ++its;
            cplx mu;
            if (its % 10 == 0) {
                mu = t(hi, hi) + cplx(std::abs(t(hi, hi - 1)), 0.0);
            } else {
                const cplx a = t(hi - 1, hi - 1), b = t(hi - 1, hi), c = t(hi, hi - 1), d = t(hi, hi);
                const cplx half_tr = 0.5 * (a + d);
                const cplx disc = std::sqrt(half_tr * half_tr - (a * d - b * c));
                const cplx l1 = half_tr + disc, l2 = half_tr - disc;
                mu = std::abs(l1 - d) <= std::abs(l2 - d) ? l1 : l2;
            }
            for (int i = lo; i <= hi; ++i) t(i, i) -= mu;
            for (int i = lo; i < hi; ++i) {
                givens(t(i, i), t(i + 1, i), cs[i], sn[i]);
                rot_rows(i, cs[i], sn[i], i);
                t(i + 1, i) = 0.0;
            }
            for (int i = lo; i < hi; ++i) rot_cols(i, cs[i], sn[i], std::min(i + 2, hi));
            for (int i = lo; i <= hi; ++i) t(i, i) += mu;
        }
        return true;
    }

    // exchange diagonal entries k, k+1
    void swap_adjacent(int k) {
        const cplx t11 = t(k, k), t22 = t(k + 1, k + 1);
        double c;
        cplx s;
        givens(t(k, k + 1), t22 - t11, c, s);
        rot_rows(k, c, s, k + 2);
        // column rotation on rows < k and on Z
        for (int row = 0; row < k; ++row) {
            const cplx a = t(row, k), b = t(row, k + 1);
            t(row, k) = c * a + std::conj(s) * b;
            t(row, k + 1) = c * b - s * a;
        }
        for (int row = 0; row < n; ++row) {
            const cplx a = z(row, k), b = z(row, k + 1);
            z(row, k) = c * a + std::conj(s) * b;
            z(row, k + 1) = c * b - s * a;
        }
        t(k, k) = t22;
        t(k + 1, k + 1) = t11;
    }
};

// Choose the k harmonic Ritz values of smallest modulus, keeping complex
// conjugate pairs together (deflated_restart.hpp:124-165); returns a 0/1 mask.
inline std::vector<int> select_smallest(const std::vector<cplx>& ev, int k) {
    const int n = (int)ev.size();
    const double zero_thresh = 10.0 * n * n * std::numeric_limits<double>::epsilon();
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return std::abs(ev[a]) < std::abs(ev[b]); });
    std::vector<int> sel(n, 0);
    int cnt = 0;
    for (int q = 0; q < n && cnt < k; ++q) {
        const int i = idx[q];
        if (sel[i]) continue;
        sel[i] = 1;
        ++cnt;
        if (std::fabs(ev[i].imag()) > zero_thresh) {
            for (int j = 0; j < n; ++j) {
                if (j != i && std::fabs(ev[j].real() - ev[i].real()) < zero_thresh &&
                    std::fabs(ev[j].imag() + ev[i].imag()) < zero_thresh) {
                    if (!sel[j]) { sel[j] = 1; ++cnt; }
                    break;
                }
            }
        }
    }
    return sel;
}

// Orthonormal real basis (n x kout, row-major, into P) of the invariant subspace of
// the real matrix A for the selected eigenvalues.  Returns kout, or -1 on failure.
inline int invariant_subspace(int n, const double* A, int k, std::vector<double>& P, std::vector<cplx>* evals) {
    Schur S;
    if (!S.compute(n, A)) return -1;
    std::vector<cplx> ev(n);
    for (int i = 0; i < n; ++i) ev[i] = S.t(i, i);
    if (evals) *evals = ev;
    std::vector<int> sel = select_smallest(ev, k);
    const int kout = std::accumulate(sel.begin(), sel.end(), 0);
    // move selected entries to the leading positions (stable)
    int dst = 0;
    for (int src = 0; src < n; ++src) {
        if (!sel[src]) continue;
        for (int j = src - 1; j >= dst; --j) {
            S.swap_adjacent(j);
            std::swap(sel[j], sel[j + 1]);
        }
        ++dst;
    }
    // real basis from real / imaginary parts of the leading Schur vectors
    std::vector<std::vector<double>> cand;
    for (int j = 0; j < kout; ++j) {
        std::vector<double> re(n), im(n);
        for (int r = 0; r < n; ++r) {
            re[r] = S.z(r, j).real();
            im[r] = S.z(r, j).imag();
        }
        cand.push_back(std::move(re));
        cand.push_back(std::move(im));
    }
    P.assign((size_t)n * kout, 0.0);
    std::vector<char> used(cand.size(), 0);
    for (int col = 0; col < kout; ++col) {
        int best = -1;
        double bn = -1.0;
        for (size_t c = 0; c < cand.size(); ++c) {
            if (used[c]) continue;
            double s = 0.0;
            for (double e : cand[c]) s += e * e;
            if (s > bn) { bn = s; best = (int)c; }
        }
        if (best < 0 || !(bn > 0.0)) return -1;
        used[best] = 1;
        const double inv = 1.0 / std::sqrt(bn);
        std::vector<double> q(cand[best]);
        for (double& e : q) e *= inv;
        for (int r = 0; r < n; ++r) P[(size_t)r * kout + col] = q[r];
        for (size_t c = 0; c < cand.size(); ++c) {
            if (used[c]) continue;
            for (int pass = 0; pass < 2; ++pass) {
                double d = 0.0;
                for (int r = 0; r < n; ++r) d += q[r] * cand[c][r];
                for (int r = 0; r < n; ++r) cand[c][r] -= d * q[r];
            }
        }
    }
    return kout;
}

// Ritz values of an Arnoldi Hessenberg H (m x n row-major, m = n + 1 or m = n): the
// eigenvalues of its square part H_n, or with `harmonic` the harmonic Ritz values, the
// eigenvalues of H_n + h_{n+1,n}^2 H_n^{-T} e_n e_n^T (dense_solvers.hpp:919-965).
inline bool ritz_values(int m, int n, const double* H, bool harmonic, std::vector<cplx>& ev) {
    std::vector<double> Ht((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Ht[(size_t)i * n + j] = H[(size_t)i * n + j];
    if (harmonic && m > n) {
        const double h_mn = H[(size_t)n * n + n - 1];
        std::vector<double> w(n), e(n, 0.0);
        e[n - 1] = 1.0;
        if (!solve_transposed(n, Ht.data(), e.data(), w.data())) return false;
        for (int i = 0; i < n; ++i) Ht[(size_t)i * n + n - 1] += h_mn * h_mn * w[i];
    }
    Schur S;
    if (!S.compute(n, Ht.data())) return false;
    ev.resize(n);
    for (int i = 0; i < n; ++i) ev[i] = S.t(i, i);
    return true;
}

// Deflated restart of GMRES (Morgan, GMRES-DR; deflated_restart.hpp:38-269).
// In: H (m x n, m = n+1, row-major), res_vec (m), k.  Out: P (m x kp, row-major,
// kp = kout+1) such that Q_new = Q P, Hnew (kp x kout), c_r (kp).  Returns kout or -1.
inline int deflate(int n, const double* H, const double* res_vec, int k, std::vector<double>& P,
                   std::vector<double>& Hnew, std::vector<double>& c_r) {
    const int m = n + 1;
    if (k > n) k = n;
    std::vector<double> Ht((size_t)n * n), w(n), e(n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Ht[i * n + j] = H[i * n + j];
    const double h_mn = H[(size_t)n * n + n - 1];
    e[n - 1] = 1.0;
    if (!solve_transposed(n, Ht.data(), e.data(), w.data())) return -1;
    for (int i = 0; i < n; ++i) Ht[i * n + n - 1] += h_mn * h_mn * w[i];
    std::vector<double> Pk;
    const int kout = invariant_subspace(n, Ht.data(), k, Pk, nullptr);
    if (kout < 0) return -1;
    const int kp = kout + 1;
    P.assign((size_t)m * kp, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < kout; ++j) P[i * kp + j] = Pk[(size_t)i * kout + j];
    for (int i = 0; i < m; ++i) P[i * kp + kout] = res_vec[i];
    for (int j = 0; j < kout; ++j) {   // Gram-Schmidt of the residual column
        double d = 0.0;
        for (int i = 0; i < m; ++i) d += P[i * kp + j] * P[i * kp + kout];
        for (int i = 0; i < m; ++i) P[i * kp + kout] -= d * P[i * kp + j];
    }
    double nrm = 0.0;
    for (int i = 0; i < m; ++i) nrm += P[i * kp + kout] * P[i * kp + kout];
    nrm = std::sqrt(nrm);
    if (!(nrm > 0.0)) return -1;
    for (int i = 0; i < m; ++i) P[i * kp + kout] /= nrm;
    c_r.assign(kp, 0.0);
    for (int j = 0; j < kp; ++j)
        for (int i = 0; i < m; ++i) c_r[j] += P[i * kp + j] * res_vec[i];
    std::vector<double> HP((size_t)m * kout, 0.0);
    for (int i = 0; i < m; ++i)
        for (int l = 0; l < n; ++l) {
            const double h = H[i * n + l];
            if (h == 0.0) continue;
            for (int j = 0; j < kout; ++j) HP[i * kout + j] += h * Pk[(size_t)l * kout + j];
        }
    Hnew.assign((size_t)kp * kout, 0.0);
    for (int i = 0; i < kp; ++i)
        for (int l = 0; l < m; ++l) {
            const double p = P[l * kp + i];
            for (int j = 0; j < kout; ++j) Hnew[i * kout + j] += p * HP[l * kout + j];
        }
    return kout;
}

}  // namespace bsp::dense
