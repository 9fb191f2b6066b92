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

// Complex Schur decomposition A = Z T Z^H of a real or complex n x n matrix.
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

    // A: n x n row-major, real (double) or complex (cplx)
    template <class S>
    bool compute(int nn, const S* A) {
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

// H = Q^T A Q (upper Hessenberg), Q orthogonal; row-major n x n.
inline void hessenberg_q(int n, std::vector<double>& H, std::vector<double>& Q) {
    Q.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) Q[(size_t)i * n + i] = 1.0;
    std::vector<double> v(n);
    for (int k = 0; k + 2 < n; ++k) {
        double tail = 0.0;   // column k below the subdiagonal: nothing to reduce when zero
        for (int i = k + 2; i < n; ++i) tail += H[(size_t)i * n + k] * H[(size_t)i * n + k];
        if (tail == 0.0) continue;
        const double nrm2 = tail + H[(size_t)(k + 1) * n + k] * H[(size_t)(k + 1) * n + k];
        const double x0 = H[(size_t)(k + 1) * n + k];
        const double alpha = x0 > 0 ? -std::sqrt(nrm2) : std::sqrt(nrm2);
        for (int i = 0; i < n; ++i) v[i] = 0.0;
        v[k + 1] = x0 - alpha;
        for (int i = k + 2; i < n; ++i) v[i] = H[(size_t)i * n + k];
        double vv = 0.0;
        for (int i = k + 1; i < n; ++i) vv += v[i] * v[i];
        if (vv == 0.0) continue;
        const double beta = 2.0 / vv;
        for (int c = 0; c < n; ++c) {   // H <- (I - beta v v^T) H
            double s = 0.0;
            for (int i = k + 1; i < n; ++i) s += v[i] * H[(size_t)i * n + c];
            s *= beta;
            for (int i = k + 1; i < n; ++i) H[(size_t)i * n + c] -= s * v[i];
        }
        for (int r = 0; r < n; ++r) {   // H <- H (I - beta v v^T), Q <- Q (I - beta v v^T)
            double s = 0.0, sq = 0.0;
            for (int i = k + 1; i < n; ++i) {
                s += H[(size_t)r * n + i] * v[i];
                sq += Q[(size_t)r * n + i] * v[i];
            }
            s *= beta;
            sq *= beta;
            for (int i = k + 1; i < n; ++i) {
                H[(size_t)r * n + i] -= s * v[i];
                Q[(size_t)r * n + i] -= sq * v[i];
            }
        }
        for (int i = k + 2; i < n; ++i) H[(size_t)i * n + k] = 0.0;
    }
}

// Eigenvalues of an upper Hessenberg H (destroyed) by Francis double-shift QR steps in real
// arithmetic on the active window (Golub & Van Loan, Matrix Computations, Alg. 7.5.1), with
// negligible-subdiagonal splitting and an exceptional shift every 11th step without one.
inline bool hqr_eigs(int n, std::vector<double>& H, std::vector<cplx>& ev) {
    auto h = [&](int i, int j) -> double& { return H[(size_t)i * n + j]; };
    ev.assign(n, cplx(0.0));
    const double eps = std::numeric_limits<double>::epsilon();
    int hi = n - 1, its = 0, total = 0;
    while (hi >= 0) {
        int lo = hi;
        for (; lo > 0; --lo) {
            double s = std::fabs(h(lo - 1, lo - 1)) + std::fabs(h(lo, lo));
            if (s == 0.0) s = 1.0;
            if (std::fabs(h(lo, lo - 1)) <= eps * s) {
                h(lo, lo - 1) = 0.0;
                break;
            }
        }
        if (lo == hi) {   // 1 x 1
            ev[hi] = h(hi, hi);
            --hi;
            its = 0;
            continue;
        }
        if (lo == hi - 1) {   // 2 x 2
            const double a = h(hi - 1, hi - 1), b = h(hi - 1, hi), c = h(hi, hi - 1), d = h(hi, hi);
            const double tr = 0.5 * (a + d), det = a * d - b * c, disc = tr * tr - det;
            if (disc >= 0.0) {
                const double sq = std::sqrt(disc);
                const double l1 = tr + (tr >= 0 ? sq : -sq);
                const double l2 = l1 != 0.0 ? det / l1 : tr - (tr >= 0 ? sq : -sq);
                ev[hi - 1] = l1;
                ev[hi] = l2;
            } else {
                const double sq = std::sqrt(-disc);
                ev[hi - 1] = cplx(tr, sq);
                ev[hi] = cplx(tr, -sq);
            }
            hi -= 2;
            its = 0;
            continue;
        }
        if (++total > 60 * n) return false;
        ++its;
        double s, t;
        if (its % 11 == 0) {   // exceptional shift
            const double e = std::fabs(h(hi, hi - 1)) + std::fabs(h(hi - 1, hi - 2));
            s = 1.5 * e;
            t = e * e;
        } else {
            s = h(hi - 1, hi - 1) + h(hi, hi);
            t = h(hi - 1, hi - 1) * h(hi, hi) - h(hi - 1, hi) * h(hi, hi - 1);
        }
        double x = h(lo, lo) * h(lo, lo) + h(lo, lo + 1) * h(lo + 1, lo) - s * h(lo, lo) + t;
        double y = h(lo + 1, lo) * (h(lo, lo) + h(lo + 1, lo + 1) - s);
        double z = h(lo + 1, lo) * h(lo + 2, lo + 1);
        for (int k = lo - 1; k <= hi - 3; ++k) {
            // reflector P = I - 2 u u^T / u^T u mapping (x, y, z) to a multiple of e1
            const double nrm = std::sqrt(x * x + y * y + z * z);
            if (nrm == 0.0) break;
            const double alpha = x > 0 ? -nrm : nrm;
            const double u0 = x - alpha, u1 = y, u2 = z;
            const double uu = u0 * u0 + u1 * u1 + u2 * u2;
            if (uu != 0.0) {
                const double beta = 2.0 / uu;
                const int c0 = std::max(k, lo);
                for (int c = c0; c <= hi; ++c) {
                    const double sv = beta * (u0 * h(k + 1, c) + u1 * h(k + 2, c) + u2 * h(k + 3, c));
                    h(k + 1, c) -= sv * u0;
                    h(k + 2, c) -= sv * u1;
                    h(k + 3, c) -= sv * u2;
                }
                const int r1 = std::min(k + 4, hi);
                for (int r = lo; r <= r1; ++r) {
                    const double sv = beta * (h(r, k + 1) * u0 + h(r, k + 2) * u1 + h(r, k + 3) * u2);
                    h(r, k + 1) -= sv * u0;
                    h(r, k + 2) -= sv * u1;
                    h(r, k + 3) -= sv * u2;
                }
            }
            x = h(k + 2, k + 1);
            y = h(k + 3, k + 1);
            if (k < hi - 3) z = h(k + 4, k + 1);
        }
        // final 2-vector Givens on rows / columns hi-1, hi
        {
            const double r = std::hypot(x, y);
            if (r != 0.0) {
                const double c = x / r, sn = y / r;
                const int k = hi - 2;
                for (int col = std::max(k, lo); col <= hi; ++col) {
                    const double a = h(hi - 1, col), b = h(hi, col);
                    h(hi - 1, col) = c * a + sn * b;
                    h(hi, col) = -sn * a + c * b;
                }
                for (int row = lo; row <= hi; ++row) {
                    const double a = h(row, hi - 1), b = h(row, hi);
                    h(row, hi - 1) = c * a + sn * b;
                    h(row, hi) = -sn * a + c * b;
                }
            }
        }
    }
    return true;
}

// Solve (H - theta I) x = b for upper Hessenberg H (complex shift) by Gaussian elimination
// with partial pivoting between adjacent rows; tiny pivots are replaced by eps * scale.  The
// complex arithmetic is written out on real and imaginary parts (one real division per
// pivot, no library calls), and `M` is the caller's workspace of n * n entries.
inline void hess_shift_solve(int n, const std::vector<double>& H, cplx theta, std::vector<cplx>& x, double scale,
                             std::vector<cplx>& M) {
    M.resize((size_t)n * n);
    auto m = [&](int i, int j) -> cplx& { return M[(size_t)i * n + j]; };
    // only the Hessenberg part is read: row i from column i - 1 on
    for (int i = 0; i < n; ++i)
        for (int j = std::max(i - 1, 0); j < n; ++j)
            m(i, j) = cplx(H[(size_t)i * n + j] - (i == j ? theta.real() : 0.0), i == j ? -theta.imag() : 0.0);
    const double tiny = std::numeric_limits<double>::epsilon() * scale;
    const double tiny2 = tiny * tiny;
    auto nrm2 = [](cplx z) { return z.real() * z.real() + z.imag() * z.imag(); };
    auto mul = [](cplx a, cplx b) {
        return cplx(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
    };
    auto inv = [&](cplx z) {
        const double d = nrm2(z);
        return cplx(z.real() / d, -z.imag() / d);
    };
    for (int k = 0; k + 1 < n; ++k) {
        if (nrm2(m(k + 1, k)) > nrm2(m(k, k))) {
            for (int j = k; j < n; ++j) std::swap(m(k, j), m(k + 1, j));
            std::swap(x[k], x[k + 1]);
        }
        if (nrm2(m(k, k)) < tiny2) m(k, k) = tiny;
        const cplx f = mul(m(k + 1, k), inv(m(k, k)));
        if (f != cplx(0.0)) {
            for (int j = k; j < n; ++j) m(k + 1, j) -= mul(f, m(k, j));
            x[k + 1] -= mul(f, x[k]);
        }
    }
    if (nrm2(m(n - 1, n - 1)) < tiny2) m(n - 1, n - 1) = tiny;
    for (int i = n - 1; i >= 0; --i) {
        cplx s = x[i];
        for (int j = i + 1; j < n; ++j) s -= mul(m(i, j), x[j]);
        x[i] = mul(s, inv(m(i, i)));
    }
}

// Orthonormal real basis (n x kout, row-major) of the invariant subspace of the real A for
// its k eigenvalues of smallest modulus (conjugate pairs kept together), without a Schur
// form: Hessenberg reduction H = Q^T A Q, eigenvalues by Francis steps, two steps of
// inverse iteration on H - theta I per selected eigenvalue (one per conjugate pair), the
// real and imaginary parts mapped back by Q and orthonormalised.  About 5x fewer flops
// than the complex Schur route for n = 30 and the same subspace on non-defective spectra.
// Returns kout, or -1 when the QR steps do not converge or the vectors are close to
// dependent (nearly defective eigenvalues), where the caller uses invariant_subspace.
inline int invariant_subspace_fast(int n, const double* A, int k, std::vector<double>& P) {
    std::vector<double> H(A, A + (size_t)n * n), Q;
    hessenberg_q(n, H, Q);
    std::vector<double> Hw(H);
    std::vector<cplx> ev;
    if (!hqr_eigs(n, Hw, ev)) return -1;
    std::vector<int> sel = select_smallest(ev, k);
    const int kout = std::accumulate(sel.begin(), sel.end(), 0);
    double scale = 0.0;
    for (double v : H) scale = std::max(scale, std::fabs(v));
    if (scale == 0.0) scale = 1.0;
    std::vector<std::vector<double>> cand;
    std::vector<char> done(n, 0);
    std::vector<cplx> work;
    const double zt = 10.0 * n * n * std::numeric_limits<double>::epsilon();
    for (int i = 0; i < n; ++i) {
        if (!sel[i] || done[i]) continue;
        done[i] = 1;
        const bool cpx = std::fabs(ev[i].imag()) > zt;
        if (cpx)   // mark the partner
            for (int j = 0; j < n; ++j)
                if (j != i && sel[j] && !done[j] && std::fabs(ev[j].real() - ev[i].real()) < zt &&
                    std::fabs(ev[j].imag() + ev[i].imag()) < zt) {
                    done[j] = 1;
                    break;
                }
        std::vector<cplx> x(n, cplx(1.0));
        for (int it = 0; it < 2; ++it) {
            hess_shift_solve(n, H, ev[i], x, scale, work);
            double nr = 0.0;
            for (auto& e : x) nr += std::norm(e);
            nr = std::sqrt(nr);
            if (!(nr > 0.0) || !std::isfinite(nr)) return -1;
            for (auto& e : x) e /= nr;
        }
        std::vector<double> re(n, 0.0), im(n, 0.0);
        for (int r = 0; r < n; ++r)
            for (int c = 0; c < n; ++c) {
                re[r] += Q[(size_t)r * n + c] * x[c].real();
                im[r] += Q[(size_t)r * n + c] * x[c].imag();
            }
        cand.push_back(re);
        if (cpx) cand.push_back(im);
    }
    if ((int)cand.size() != kout) return -1;
    P.assign((size_t)n * kout, 0.0);
    for (int col = 0; col < kout; ++col) {
        std::vector<double>& q = cand[col];
        double n0 = 0.0;
        for (double e : q) n0 += e * e;
        n0 = std::sqrt(n0);
        for (int pass = 0; pass < 2; ++pass)
            for (int j = 0; j < col; ++j) {
                double d = 0.0;
                for (int r = 0; r < n; ++r) d += P[(size_t)r * kout + j] * q[r];
                for (int r = 0; r < n; ++r) q[r] -= d * P[(size_t)r * kout + j];
            }
        double n1 = 0.0;
        for (double e : q) n1 += e * e;
        n1 = std::sqrt(n1);
        if (!(n1 > 1e-6 * n0)) return -1;
        for (int r = 0; r < n; ++r) P[(size_t)r * kout + col] = q[r] / n1;
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
// fast: the harmonic Ritz subspace from invariant_subspace_fast (Schur route as fallback);
// otherwise the Schur route, as the reference computes it.
inline int deflate(int n, const double* H, const double* res_vec, int k, std::vector<double>& P,
                   std::vector<double>& Hnew, std::vector<double>& c_r, bool fast = false) {
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
    int kout = fast ? invariant_subspace_fast(n, Ht.data(), k, Pk) : -1;
    if (kout < 0) kout = invariant_subspace(n, Ht.data(), k, Pk, nullptr);
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
