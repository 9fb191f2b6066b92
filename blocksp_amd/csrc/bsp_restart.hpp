// bsp_restart.hpp -- the small dense work of GMRES kept on the device: the Givens least-squares
// update of each Arnoldi step, and the deflated restart (Morgan's GMRES-DR; the reference's
// deflated_restart.hpp:38-269).
//
// Every routine is written for a team of threads sharing its workspace: on the device one
// workgroup with the workspace in LDS, on the host one thread (the unit tests compare that
// instantiation with the host routines of bsp_dense.hpp).  Loops over rows or columns are split
// across the team; the scalar control of each algorithm (pivots, shifts, convergence tests) is
// computed by every thread from the same shared values after a barrier, so all threads take
// the same branches.  Methods (textbook, as the host versions in bsp_dense.hpp):
//   * A^T w = e_n by Gaussian elimination with partial pivoting;
//   * Hessenberg reduction H = Q^T G Q by Householder reflections with Q accumulated;
//   * eigenvalues of H by Francis double-shift QR steps (Golub & Van Loan, Alg. 7.5.1);
//   * the k of smallest modulus, conjugate pairs kept together (deflated_restart.hpp:124-165);
//   * their eigenvectors by two steps of complex inverse iteration on H - theta I, mapped back
//     by Q, real and imaginary parts orthonormalised by two-pass Gram-Schmidt;
//   * P_{k+1} = [P_k, residual orthogonalised], c_r = P_{k+1}^T res, H_new = P_{k+1}^T H P_k;
//   * the least-squares problem min ||c - H y|| kept as R = Q0^T G^T H (Householder QR of H_new
//     at a restart, one Givens rotation per later column).
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

namespace bsp::rs {

// Team of one workgroup on the device.
struct DevTeam {
    __device__ int tid() const { return (int)threadIdx.x; }
    __device__ int nt() const { return (int)blockDim.x; }
    __device__ void sync() const { __syncthreads(); }
};

// One host thread (unit tests).
struct HostTeam {
    __host__ int tid() const { return 0; }
    __host__ int nt() const { return 1; }
    __host__ void sync() const {}
};

struct cd {
    double re, im;
};
__host__ __device__ inline double cabs_(cd a) { return hypot(a.re, a.im); }
__host__ __device__ inline cd cmul(cd a, cd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__host__ __device__ inline cd csub(cd a, cd b) { return {a.re - b.re, a.im - b.im}; }
// a / b (Smith's algorithm)
__host__ __device__ inline cd cdiv(cd a, cd b) {
    if (fabs(b.re) >= fabs(b.im)) {
        const double r = b.im / b.re, d = b.re + b.im * r;
        return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
    }
    const double r = b.re / b.im, d = b.re * r + b.im;
    return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}

// Largest Krylov dimension of the device restart (the harmonic matrix is n x n, n <= kMaxN).
constexpr int kMaxN = 64;

// Workspace (doubles) of deflate_team for dimension n: G, Q, X (n*n each), W (2 n*n), nine
// vectors of n and 16 shared scalars.
__host__ __device__ inline int deflate_ws_doubles(int n) { return 5 * n * n + 9 * n + 16; }

// Solve G^T w = e_{n-1} (G: n x n row-major).  M: n*n work, f: n work.  False if singular.
template <class T>
__host__ __device__ bool solve_transposed_team(const T& tm, int n, const double* G, double* M, double* rhs, double* f,
                                               double* w) {
    const int tid = tm.tid(), nt = tm.nt();
    for (int idx = tid; idx < n * n; idx += nt) M[idx] = G[(idx % n) * n + idx / n];
    for (int i = tid; i < n; i += nt) rhs[i] = i == n - 1 ? 1.0 : 0.0;
    tm.sync();
    for (int col = 0; col < n; ++col) {
        int piv = col;
        double best = fabs(M[col * n + col]);
        for (int r = col + 1; r < n; ++r)
            if (fabs(M[r * n + col]) > best) {
                best = fabs(M[r * n + col]);
                piv = r;
            }
        if (best == 0.0) return false;
        if (piv != col) {
            tm.sync();
            for (int c = tid; c < n; c += nt) {
                const double a = M[col * n + c];
                M[col * n + c] = M[piv * n + c];
                M[piv * n + c] = a;
            }
            if (tid == 0) {
                const double a = rhs[col];
                rhs[col] = rhs[piv];
                rhs[piv] = a;
            }
            tm.sync();
        }
        for (int r = col + 1 + tid; r < n; r += nt) f[r] = M[r * n + col] / M[col * n + col];
        tm.sync();
        const int w_ = n - col;
        for (int idx = tid; idx < (n - col - 1) * w_; idx += nt) {
            const int r = col + 1 + idx / w_, c = col + idx % w_;
            if (f[r] != 0.0) M[r * n + c] -= f[r] * M[col * n + c];
        }
        for (int r = col + 1 + tid; r < n; r += nt)
            if (f[r] != 0.0) rhs[r] -= f[r] * rhs[col];
        tm.sync();
    }
    if (tid == 0)
        for (int i = n - 1; i >= 0; --i) {
            double s = rhs[i];
            for (int c = i + 1; c < n; ++c) s -= M[i * n + c] * w[c];
            w[i] = s / M[i * n + i];
        }
    tm.sync();
    return true;
}

// H <- Q^T H Q upper Hessenberg (in place), Q orthogonal (out); v: n work.
template <class T>
__host__ __device__ void hessenberg_team(const T& tm, int n, double* H, double* Q, double* v) {
    const int tid = tm.tid(), nt = tm.nt();
    for (int idx = tid; idx < n * n; idx += nt) Q[idx] = (idx / n == idx % n) ? 1.0 : 0.0;
    tm.sync();
    for (int k = 0; k + 2 < n; ++k) {
        double nrm2 = 0.0;
        for (int i = k + 1; i < n; ++i) nrm2 += H[i * n + k] * H[i * n + k];
        if (nrm2 == 0.0) continue;
        const double x0 = H[(k + 1) * n + k];
        const double alpha = x0 > 0 ? -sqrt(nrm2) : sqrt(nrm2);
        for (int i = tid; i < n; i += nt) v[i] = i < k + 1 ? 0.0 : (i == k + 1 ? x0 - alpha : H[i * n + k]);
        tm.sync();
        double vv = 0.0;
        for (int i = k + 1; i < n; ++i) vv += v[i] * v[i];
        if (vv == 0.0) continue;
        const double beta = 2.0 / vv;
        for (int c = tid; c < n; c += nt) {   // H <- (I - beta v v^T) H
            double s = 0.0;
            for (int i = k + 1; i < n; ++i) s += v[i] * H[i * n + c];
            s *= beta;
            for (int i = k + 1; i < n; ++i) H[i * n + c] -= s * v[i];
        }
        tm.sync();
        for (int r = tid; r < n; r += nt) {   // H <- H (I - beta v v^T), Q <- Q (I - beta v v^T)
            double s = 0.0, sq = 0.0;
            for (int i = k + 1; i < n; ++i) {
                s += H[r * n + i] * v[i];
                sq += Q[r * n + i] * v[i];
            }
            s *= beta;
            sq *= beta;
            for (int i = k + 1; i < n; ++i) {
                H[r * n + i] -= s * v[i];
                Q[r * n + i] -= sq * v[i];
            }
        }
        tm.sync();
        for (int i = k + 2 + tid; i < n; i += nt) H[i * n + k] = 0.0;
        tm.sync();
    }
}

// Eigenvalues of the upper Hessenberg H (destroyed) into (er, ei) by Francis double-shift QR
// steps on the active window with negligible-subdiagonal splitting and an exceptional shift
// every 11th step without one.  False when 60 n steps do not suffice.
template <class T>
__host__ __device__ bool hqr_team(const T& tm, int n, double* H, double* er, double* ei) {
    const int tid = tm.tid(), nt = tm.nt();
    const double eps = DBL_EPSILON;
#define HH(i, j) H[(i) * n + (j)]
    int hi = n - 1, its = 0, total = 0;
    while (hi >= 0) {
        int lo = hi;
        for (; lo > 0; --lo) {
            double s = fabs(HH(lo - 1, lo - 1)) + fabs(HH(lo, lo));
            if (s == 0.0) s = 1.0;
            if (fabs(HH(lo, lo - 1)) <= eps * s) break;
        }
        tm.sync();
        if (lo > 0 && tid == 0) HH(lo, lo - 1) = 0.0;
        if (lo == hi) {
            if (tid == 0) {
                er[hi] = HH(hi, hi);
                ei[hi] = 0.0;
            }
            tm.sync();
            --hi;
            its = 0;
            continue;
        }
        if (lo == hi - 1) {
            const double a = HH(hi - 1, hi - 1), b = HH(hi - 1, hi), c = HH(hi, hi - 1), d = HH(hi, hi);
            const double tr = 0.5 * (a + d), det = a * d - b * c, disc = tr * tr - det;
            if (tid == 0) {
                if (disc >= 0.0) {
                    const double sq = sqrt(disc);
                    const double l1 = tr + (tr >= 0 ? sq : -sq);
                    const double l2 = l1 != 0.0 ? det / l1 : tr - (tr >= 0 ? sq : -sq);
                    er[hi - 1] = l1;
                    ei[hi - 1] = 0.0;
                    er[hi] = l2;
                    ei[hi] = 0.0;
                } else {
                    const double sq = sqrt(-disc);
                    er[hi - 1] = tr;
                    ei[hi - 1] = sq;
                    er[hi] = tr;
                    ei[hi] = -sq;
                }
            }
            tm.sync();
            hi -= 2;
            its = 0;
            continue;
        }
        if (++total > 60 * n) return false;
        ++its;
        tm.sync();
        double s, t;
        if (its % 11 == 0) {
            const double e = fabs(HH(hi, hi - 1)) + fabs(HH(hi - 1, hi - 2));
            s = 1.5 * e;
            t = e * e;
        } else {
            s = HH(hi - 1, hi - 1) + HH(hi, hi);
            t = HH(hi - 1, hi - 1) * HH(hi, hi) - HH(hi - 1, hi) * HH(hi, hi - 1);
        }
        double x = HH(lo, lo) * HH(lo, lo) + HH(lo, lo + 1) * HH(lo + 1, lo) - s * HH(lo, lo) + t;
        double y = HH(lo + 1, lo) * (HH(lo, lo) + HH(lo + 1, lo + 1) - s);
        double z = HH(lo + 1, lo) * HH(lo + 2, lo + 1);
        for (int k = lo - 1; k <= hi - 3; ++k) {
            const double nrm = sqrt(x * x + y * y + z * z);
            if (nrm == 0.0) break;
            const double alpha = x > 0 ? -nrm : nrm;
            const double u0 = x - alpha, u1 = y, u2 = z;
            const double uu = u0 * u0 + u1 * u1 + u2 * u2;
            if (uu != 0.0) {
                const double beta = 2.0 / uu;
                const int c0 = k > lo ? k : lo;
                tm.sync();
                for (int c = c0 + tid; c <= hi; c += nt) {
                    const double sv = beta * (u0 * HH(k + 1, c) + u1 * HH(k + 2, c) + u2 * HH(k + 3, c));
                    HH(k + 1, c) -= sv * u0;
                    HH(k + 2, c) -= sv * u1;
                    HH(k + 3, c) -= sv * u2;
                }
                tm.sync();
                const int r1 = k + 4 < hi ? k + 4 : hi;
                for (int r = lo + tid; r <= r1; r += nt) {
                    const double sv = beta * (HH(r, k + 1) * u0 + HH(r, k + 2) * u1 + HH(r, k + 3) * u2);
                    HH(r, k + 1) -= sv * u0;
                    HH(r, k + 2) -= sv * u1;
                    HH(r, k + 3) -= sv * u2;
                }
                tm.sync();
            }
            x = HH(k + 2, k + 1);
            y = HH(k + 3, k + 1);
            if (k < hi - 3) z = HH(k + 4, k + 1);
        }
        const double r = hypot(x, y);
        if (r != 0.0) {
            const double c = x / r, sn = y / r;
            const int k = hi - 2;
            tm.sync();
            for (int col = (k > lo ? k : lo) + tid; col <= hi; col += nt) {
                const double a = HH(hi - 1, col), b = HH(hi, col);
                HH(hi - 1, col) = c * a + sn * b;
                HH(hi, col) = -sn * a + c * b;
            }
            tm.sync();
            for (int row = lo + tid; row <= hi; row += nt) {
                const double a = HH(row, hi - 1), b = HH(row, hi);
                HH(row, hi - 1) = c * a + sn * b;
                HH(row, hi) = -sn * a + c * b;
            }
        }
        tm.sync();
    }
#undef HH
    return true;
}

// (H - theta I) x = x_in for upper Hessenberg H, complex theta, by Gaussian elimination with
// pivoting between adjacent rows; tiny pivots replaced by eps * scale.  M: 2 n*n work (complex
// interleaved), x: n complex in (xr, xi), overwritten.
template <class T>
__host__ __device__ void hess_shift_solve_team(const T& tm, int n, const double* H, cd theta, double* xr, double* xi,
                                               double scale, double* M) {
    const int tid = tm.tid(), nt = tm.nt();
    cd* Mc = reinterpret_cast<cd*>(M);
    for (int idx = tid; idx < n * n; idx += nt) {
        const int i = idx / n, j = idx % n;
        Mc[idx] = {H[idx] - (i == j ? theta.re : 0.0), i == j ? -theta.im : 0.0};
    }
    tm.sync();
    const double tiny = DBL_EPSILON * scale;
    for (int k = 0; k + 1 < n; ++k) {
        if (cabs_(Mc[(k + 1) * n + k]) > cabs_(Mc[k * n + k])) {
            tm.sync();
            for (int j = k + tid; j < n; j += nt) {
                const cd a = Mc[k * n + j];
                Mc[k * n + j] = Mc[(k + 1) * n + j];
                Mc[(k + 1) * n + j] = a;
            }
            if (tid == 0) {
                double a = xr[k];
                xr[k] = xr[k + 1];
                xr[k + 1] = a;
                a = xi[k];
                xi[k] = xi[k + 1];
                xi[k + 1] = a;
            }
            tm.sync();
        }
        if (cabs_(Mc[k * n + k]) < tiny) {
            tm.sync();
            if (tid == 0) Mc[k * n + k] = {tiny, 0.0};
            tm.sync();
        }
        const cd f = cdiv(Mc[(k + 1) * n + k], Mc[k * n + k]);
        if (f.re != 0.0 || f.im != 0.0) {
            tm.sync();
            for (int j = k + tid; j < n; j += nt) Mc[(k + 1) * n + j] = csub(Mc[(k + 1) * n + j], cmul(f, Mc[k * n + j]));
            if (tid == 0) {
                const cd t = csub({xr[k + 1], xi[k + 1]}, cmul(f, {xr[k], xi[k]}));
                xr[k + 1] = t.re;
                xi[k + 1] = t.im;
            }
            tm.sync();
        }
    }
    if (tid == 0) {
        if (cabs_(Mc[(n - 1) * n + n - 1]) < tiny) Mc[(n - 1) * n + n - 1] = {tiny, 0.0};
        for (int i = n - 1; i >= 0; --i) {
            cd s = {xr[i], xi[i]};
            for (int j = i + 1; j < n; ++j) s = csub(s, cmul(Mc[i * n + j], {xr[j], xi[j]}));
            const cd q = cdiv(s, Mc[i * n + i]);
            xr[i] = q.re;
            xi[i] = q.im;
        }
    }
    tm.sync();
}

// Deflated restart (host counterpart: bsp_dense.hpp deflate with the inverse-iteration route).
// H: (n + 1) x n column-major with leading dimension ldh; res: the n + 1 entries of the least-squares
// residual c - H y.  Out: P ((n + 1) x kp, row-major, kp = kout + 1), Pk (n x kout), Hnew
// (kp x kout), cr (kp).  ws: deflate_ws_doubles(n) doubles shared by the team.  Returns kout,
// or -1 when the harmonic matrix is singular, the QR steps do not converge or the eigenvectors
// are close to dependent (the host route, a complex Schur form, handles those).
template <class T>
__host__ __device__ int deflate_team(const T& tm, int n, const double* H, int ldh, const double* res, int k, double* ws,
                                     double* P, double* Pk, double* Hnew, double* cr) {
    const int tid = tm.tid(), nt = tm.nt();
    const int m = n + 1;
    if (k > n) k = n;
    double* G = ws;              // n*n  harmonic matrix, then its Hessenberg form
    double* Q = G + n * n;       // n*n
    double* W = Q + n * n;       // 2 n*n  work
    double* X = W + 2 * n * n;   // n*n  map-back / candidates
    double* v = X + n * n;       // n
    double* f = v + n;           // n
    double* rhs = f + n;         // n
    double* er = rhs + n;        // n
    double* ei = er + n;         // n
    double* xr = ei + n;         // n
    double* xi = xr + n;         // n
    double* sel = xi + n;        // n  (0 / 1)
    double* done = sel + n;      // n  (0 / 1)
    double* sh = done + n;       // 16 shared scalars
    // harmonic Ritz matrix: G = H_n + h_{n+1,n}^2 H_n^{-T} e_n e_n^T (deflated_restart.hpp:66-99)
    for (int idx = tid; idx < n * n; idx += nt) G[idx] = H[(size_t)(idx % n) * ldh + idx / n];
    tm.sync();
    if (!solve_transposed_team(tm, n, G, W, rhs, f, v)) return -1;
    const double h_mn = H[(size_t)(n - 1) * ldh + n];
    for (int i = tid; i < n; i += nt) G[i * n + n - 1] += h_mn * h_mn * v[i];
    tm.sync();
    // Hessenberg form and eigenvalues
    hessenberg_team(tm, n, G, Q, v);
    for (int idx = tid; idx < n * n; idx += nt) W[idx] = G[idx];
    tm.sync();
    if (!hqr_team(tm, n, W, er, ei)) return -1;
    // selection: the k of smallest modulus, conjugate pairs together (stable order by modulus)
    const double zt = 10.0 * n * n * DBL_EPSILON;
    if (tid == 0) {
        int* idx = reinterpret_cast<int*>(f);   // n ints fit in n doubles
        for (int i = 0; i < n; ++i) {
            idx[i] = i;
            sel[i] = 0.0;
            done[i] = 0.0;
        }
        for (int i = 1; i < n; ++i) {   // insertion sort: stable
            const int a = idx[i];
            const double ma = hypot(er[a], ei[a]);
            int j = i - 1;
            while (j >= 0 && hypot(er[idx[j]], ei[idx[j]]) > ma) {
                idx[j + 1] = idx[j];
                --j;
            }
            idx[j + 1] = a;
        }
        int cnt = 0;
        for (int q = 0; q < n && cnt < k; ++q) {
            const int i = idx[q];
            if (sel[i] != 0.0) continue;
            sel[i] = 1.0;
            ++cnt;
            if (fabs(ei[i]) > zt)
                for (int j = 0; j < n; ++j)
                    if (j != i && fabs(er[j] - er[i]) < zt && fabs(ei[j] + ei[i]) < zt) {
                        if (sel[j] == 0.0) {
                            sel[j] = 1.0;
                            ++cnt;
                        }
                        break;
                    }
        }
        double scale = 0.0;
        for (int i = 0; i < n * n; ++i) scale = fmax(scale, fabs(G[i]));
        sh[0] = scale == 0.0 ? 1.0 : scale;
        sh[1] = cnt;
    }
    tm.sync();
    const int kout = (int)sh[1];
    const double scale = sh[0];
    // inverse iteration per selected eigenvalue (one per conjugate pair); the real (and
    // imaginary) parts mapped back by Q become candidate columns of X
    int ncand = 0;
    for (int i = 0; i < n; ++i) {
        // selected and not yet covered: one vector per eigenvalue, its conjugate partner (the
        // first selected one not yet covered) marked covered
        if (tid == 0) {
            double go = 0.0;
            if (sel[i] != 0.0 && done[i] == 0.0) {
                go = 1.0;
                done[i] = 1.0;
                if (fabs(ei[i]) > zt)
                    for (int j = 0; j < n; ++j)
                        if (j != i && sel[j] != 0.0 && done[j] == 0.0 && fabs(er[j] - er[i]) < zt &&
                            fabs(ei[j] + ei[i]) < zt) {
                            done[j] = 1.0;
                            break;
                        }
            }
            sh[2] = go;
        }
        tm.sync();
        const bool go = sh[2] != 0.0;
        tm.sync();
        if (!go) continue;
        const bool cpx = fabs(ei[i]) > zt;
        for (int r = tid; r < n; r += nt) {
            xr[r] = 1.0;
            xi[r] = 0.0;
        }
        tm.sync();
        for (int it = 0; it < 2; ++it) {
            hess_shift_solve_team(tm, n, G, {er[i], ei[i]}, xr, xi, scale, W);
            double nr = 0.0;
            for (int r = 0; r < n; ++r) nr += xr[r] * xr[r] + xi[r] * xi[r];
            nr = sqrt(nr);
            if (!(nr > 0.0 && nr <= DBL_MAX)) return -1;   // zero, infinite or NaN
            tm.sync();
            for (int r = tid; r < n; r += nt) {
                xr[r] /= nr;
                xi[r] /= nr;
            }
            tm.sync();
        }
        if (ncand + (cpx ? 2 : 1) > kout) return -1;
        for (int r = tid; r < n; r += nt) {
            double a = 0.0, b = 0.0;
            for (int c = 0; c < n; ++c) {
                a += Q[r * n + c] * xr[c];
                b += Q[r * n + c] * xi[c];
            }
            X[ncand * n + r] = a;
            if (cpx) X[(ncand + 1) * n + r] = b;
        }
        tm.sync();
        ncand += cpx ? 2 : 1;
    }
    if (ncand != kout) return -1;
    // orthonormalise the candidates (two Gram-Schmidt passes) into Pk (n x kout)
    for (int col = 0; col < kout; ++col) {
        double* q = X + col * n;
        double n0 = 0.0;
        for (int r = 0; r < n; ++r) n0 += q[r] * q[r];
        n0 = sqrt(n0);
        for (int pass = 0; pass < 2; ++pass)
            for (int j = 0; j < col; ++j) {
                double d = 0.0;
                for (int r = 0; r < n; ++r) d += Pk[r * kout + j] * q[r];
                tm.sync();
                for (int r = tid; r < n; r += nt) q[r] -= d * Pk[r * kout + j];
                tm.sync();
            }
        double n1 = 0.0;
        for (int r = 0; r < n; ++r) n1 += q[r] * q[r];
        n1 = sqrt(n1);
        if (!(n1 > 1e-6 * n0)) return -1;
        for (int r = tid; r < n; r += nt) Pk[r * kout + col] = q[r] / n1;
        tm.sync();
    }
    // P_{k+1} = [P_k; 0 | res orthogonalised], c_r, H_new (deflated_restart.hpp:167-245)
    const int kp = kout + 1;
    for (int idx = tid; idx < m * kp; idx += nt) {
        const int i = idx / kp, j = idx % kp;
        P[idx] = j < kout ? (i < n ? Pk[i * kout + j] : 0.0) : res[i];
    }
    tm.sync();
    for (int j = 0; j < kout; ++j) {
        double d = 0.0;
        for (int i = 0; i < m; ++i) d += P[i * kp + j] * P[i * kp + kout];
        tm.sync();
        for (int i = tid; i < m; i += nt) P[i * kp + kout] -= d * P[i * kp + j];
        tm.sync();
    }
    double nrm = 0.0;
    for (int i = 0; i < m; ++i) nrm += P[i * kp + kout] * P[i * kp + kout];
    nrm = sqrt(nrm);
    if (!(nrm > 0.0)) return -1;
    tm.sync();
    for (int i = tid; i < m; i += nt) P[i * kp + kout] /= nrm;
    tm.sync();
    for (int j = tid; j < kp; j += nt) {
        double s = 0.0;
        for (int i = 0; i < m; ++i) s += P[i * kp + j] * res[i];
        cr[j] = s;
    }
    // HP = H P_k (m x kout) into W, then Hnew = P^T HP
    for (int idx = tid; idx < m * kout; idx += nt) {
        const int i = idx / kout, j = idx % kout;
        double s = 0.0;
        for (int l = 0; l < n; ++l) {
            const double h = H[(size_t)l * ldh + i];
            if (h != 0.0) s += h * Pk[l * kout + j];
        }
        W[idx] = s;
    }
    tm.sync();
    for (int idx = tid; idx < kp * kout; idx += nt) {
        const int i = idx / kout, j = idx % kout;
        double s = 0.0;
        for (int l = 0; l < m; ++l) s += P[l * kp + i] * W[l * kout + j];
        Hnew[idx] = s;
    }
    tm.sync();
    return kout;
}

// ---------------------------------------------------------------------------------------------
// Least-squares state of one GMRES cycle, on the device.  Columns of the Hessenberg H (as the
// Arnoldi steps produce them) and of R are stored column-major with leading dimension ld =
// maxm + 1.  A cycle starts either fresh (c = beta e_1, nq0 = 0) or after a deflated restart:
// then its first kout columns are H_new (kp x kout), Q0t (kp x kp) = the orthogonal factor of
// H_new's QR transposed, R's first kout columns its triangle, and g = Q0t c_r.  Column j >= kout
// (or >= 0 fresh): t = H(:, j); t[0..kp) <- Q0t t[0..kp); the rotations of columns
// kout .. j-1; a new rotation zeroes t[j + 1]; R(:, j) = t[0..j]; |g[j + 1]| is the least-squares
// residual after j + 1 columns.
struct LSState {
    double* H;        // ld x maxm unrotated Hessenberg (column-major)
    double* R;        // ld x maxm
    double* cs;       // maxm rotation cosines
    double* sn;       // maxm rotation sines
    double* g;        // ld rotated right-hand side
    double* Q0t;      // kMaxN+1 x kMaxN+1 (row-major, leading dimension kp)
    double* y;        // maxm solution of R y = g
    double* res;      // maxm + 1: least-squares residual after each column (by column index)
    int* meta;        // [0] i0 = first column of the cycle (kout after a restart, else 0), [1] nq0 = kp or 0
    int ld;           // maxm + 1
    int maxm;         // columns of H and R
};

// Column j of the cycle from the Arnoldi step's column hcol[0..j+1] (copied into st.H(:, j));
// one team, t: ls_column_ws(j) shared doubles (the column, then the rotations staged next to it
// so that the sequential part reads shared memory only).  Writes res[j + 1]; out (optional,
// host-visible): out[0] = residual, out[1] = |h_{j+1,j}|.
__host__ __device__ inline int ls_column_ws(int j) { return 3 * (j + 2); }
template <class T>
__host__ __device__ void ls_column_team(const T& tm, const LSState& st, int j, const double* hcol, double* t,
                                        double* out) {
    const int tid = tm.tid(), nt = tm.nt();
    const int ld = st.ld;
    const int i0 = st.meta[0], nq0 = st.meta[1];
    double* hc = st.H + (size_t)j * ld;
    double* tc = t + (j + 2);
    double* ts = tc + (j + 2);
    for (int i = tid; i < j + 2; i += nt) {
        hc[i] = hcol[i];
        t[i] = hcol[i];
    }
    for (int i = i0 + tid; i < j; i += nt) {
        tc[i] = st.cs[i];
        ts[i] = st.sn[i];
    }
    if (tid == 0) tc[j] = st.g[j];
    tm.sync();
    if (nq0 > 0) {
        for (int i = tid; i < nq0; i += nt) {
            double s = 0.0;
            for (int l = 0; l < nq0; ++l) s += st.Q0t[i * nq0 + l] * hcol[l];
            t[i] = s;
        }
        tm.sync();
    }
    if (tid == 0) {
        for (int i = i0; i < j; ++i) {
            const double a = t[i], b = t[i + 1];
            t[i] = tc[i] * a + ts[i] * b;
            t[i + 1] = -ts[i] * a + tc[i] * b;
        }
        const double a = t[j], b = t[j + 1];
        const double r = hypot(a, b);
        double c = 1.0, s = 0.0;
        if (r != 0.0) {
            c = a / r;
            s = b / r;
        }
        st.cs[j] = c;
        st.sn[j] = s;
        t[j] = r;
        t[j + 1] = 0.0;
        const double gj = tc[j];
        st.g[j] = c * gj;
        st.g[j + 1] = -s * gj;
        st.res[j + 1] = fabs(st.g[j + 1]);
        if (out) {
            out[0] = st.res[j + 1];
            out[1] = fabs(hcol[j + 1]);
        }
    }
    tm.sync();
    for (int i = tid; i <= j; i += nt) st.R[(size_t)j * ld + i] = t[i];
    tm.sync();
}

// y = R(0..nc-1, 0..nc-1)^{-1} g(0..nc-1) and the residual vector resv = c - H y (m = nc + 1
// entries) with c = beta e_1 (fresh cycle, cr == nullptr) or c_r (kp entries, zero-padded).
// w (optional): ls_solve_ws(nc) shared doubles -- R, g, H and y staged there and the back
// substitution done column by column (one division, then the team updates the rest); without
// it one thread substitutes row by row on st's memory.
__host__ __device__ inline int ls_solve_ws(int nc) { return 2 * nc * nc + 3 * nc + 1; }
template <class T>
__host__ __device__ void ls_solve_team(const T& tm, const LSState& st, int nc, double beta, const double* cr, int kp,
                                       double* resv, double* w = nullptr) {
    const int tid = tm.tid(), nt = tm.nt();
    const int ld = st.ld;
    if (w) {
        double* Rs = w;                  // nc x nc column-major
        double* Hs = Rs + nc * nc;       // (nc + 1) x nc column-major
        double* gs = Hs + (nc + 1) * nc; // nc
        double* ys = gs + nc;            // nc
        for (int idx = tid; idx < nc * nc; idx += nt) Rs[idx] = st.R[(size_t)(idx / nc) * ld + idx % nc];
        if (resv)
            for (int idx = tid; idx < (nc + 1) * nc; idx += nt)
                Hs[idx] = st.H[(size_t)(idx / (nc + 1)) * ld + idx % (nc + 1)];
        for (int i = tid; i < nc; i += nt) gs[i] = st.g[i];
        tm.sync();
        for (int i = nc - 1; i >= 0; --i) {
            if (tid == 0) ys[i] = gs[i] / Rs[i * nc + i];
            tm.sync();
            for (int r = tid; r < i; r += nt) gs[r] -= Rs[i * nc + r] * ys[i];
            tm.sync();
        }
        for (int i = tid; i < nc; i += nt) st.y[i] = ys[i];
        if (resv)
            for (int i = tid; i < nc + 1; i += nt) {
                double s = 0.0;
                for (int c = 0; c < nc; ++c) s += -1.0 * Hs[c * (nc + 1) + i] * ys[c];
                const double ci = cr ? (i < kp ? cr[i] : 0.0) : (i == 0 ? beta : 0.0);
                resv[i] = s + ci;
            }
        tm.sync();
        return;
    }
    if (tid == 0)
        for (int i = nc - 1; i >= 0; --i) {
            double s = st.g[i];
            for (int c = i + 1; c < nc; ++c) s -= st.R[(size_t)c * ld + i] * st.y[c];
            st.y[i] = s / st.R[(size_t)i * ld + i];
        }
    tm.sync();
    if (resv) {
        for (int i = tid; i < nc + 1; i += nt) {
            double s = 0.0;
            for (int c = 0; c < nc; ++c) s += -1.0 * st.H[(size_t)c * ld + i] * st.y[c];
            const double ci = cr ? (i < kp ? cr[i] : 0.0) : (i == 0 ? beta : 0.0);
            resv[i] = s + ci;
        }
        tm.sync();
    }
}

// Start of a cycle after a deflated restart: H = R = 0, then H(0..kp, 0..kout) = Hnew (kp x kout), its
// Householder QR -> R's first kout columns, Q0t (kp x kp), g = Q0t c_r; meta = {kout, kp};
// res[kout] = |g[kout]|.  w: ls_restart_ws(kout) shared doubles.
__host__ __device__ inline int ls_restart_ws(int kout) { return (kout + 1) * (2 * kout + 3); }
template <class T>
__host__ __device__ void ls_restart_team(const T& tm, const LSState& st, int kout, const double* Hnew,
                                         const double* cr, double* w) {
    const int tid = tm.tid(), nt = tm.nt();
    const int ld = st.ld, kp = kout + 1;
    double* A = w;                 // kp x kout, row-major
    double* v = A + kp * kout;     // kp
    double* Qs = v + kp;           // kp x kp
    double* gs = Qs + kp * kp;     // kp
    for (int idx = tid; idx < kp * kout; idx += nt) A[idx] = Hnew[idx];
    for (int idx = tid; idx < kp * kp; idx += nt) Qs[idx] = (idx / kp == idx % kp) ? 1.0 : 0.0;
    for (int i = tid; i < kp; i += nt) gs[i] = cr[i];
    for (int i = tid; i < ld; i += nt) st.g[i] = 0.0;
    // the cycle's H and R start from zero; the unrotated leading block of H is H_new (for the
    // residual vector and the next restart)
    for (int idx = tid; idx < ld * st.maxm; idx += nt) {
        st.H[idx] = 0.0;
        st.R[idx] = 0.0;
    }
    tm.sync();
    for (int idx = tid; idx < kp * kout; idx += nt) st.H[(size_t)(idx % kout) * ld + idx / kout] = Hnew[idx];
    for (int j = 0; j < kout; ++j) {
        double nrm = 0.0;
        for (int i = j; i < kp; ++i) nrm += A[i * kout + j] * A[i * kout + j];
        nrm = sqrt(nrm);
        if (nrm == 0.0) continue;
        const double alpha = A[j * kout + j] > 0 ? -nrm : nrm;
        tm.sync();
        for (int i = tid; i < kp; i += nt) v[i] = i < j ? 0.0 : (i == j ? A[j * kout + j] - alpha : A[i * kout + j]);
        tm.sync();
        double vv = 0.0;
        for (int i = j; i < kp; ++i) vv += v[i] * v[i];
        if (vv == 0.0) continue;
        const double beta = 2.0 / vv;
        // A <- (I - beta v v^T) A, Q0t <- (I - beta v v^T) Q0t, g likewise
        for (int col = tid; col < kout + kp + 1; col += nt) {
            double s = 0.0;
            if (col < kout) {
                for (int i = j; i < kp; ++i) s += v[i] * A[i * kout + col];
                s *= beta;
                for (int i = j; i < kp; ++i) A[i * kout + col] -= s * v[i];
            } else if (col < kout + kp) {
                const int c = col - kout;
                for (int i = j; i < kp; ++i) s += v[i] * Qs[i * kp + c];
                s *= beta;
                for (int i = j; i < kp; ++i) Qs[i * kp + c] -= s * v[i];
            } else {
                for (int i = j; i < kp; ++i) s += v[i] * gs[i];
                s *= beta;
                for (int i = j; i < kp; ++i) gs[i] -= s * v[i];
            }
        }
        tm.sync();
    }
    for (int idx = tid; idx < kout * kout; idx += nt) {
        const int i = idx / kout, j = idx % kout;
        st.R[(size_t)j * ld + i] = i <= j ? A[i * kout + j] : 0.0;
    }
    for (int idx = tid; idx < kp * kp; idx += nt) st.Q0t[idx] = Qs[idx];
    for (int i = tid; i < kp; i += nt) st.g[i] = gs[i];
    if (tid == 0) {
        st.meta[0] = kout;
        st.meta[1] = kp;
        st.res[kout] = fabs(gs[kout]);
    }
    tm.sync();
}

// Start of a fresh cycle: g = beta e_1, H = R = 0, meta = {0, 0}, res[0] = beta.
template <class T>
__host__ __device__ void ls_fresh_team(const T& tm, const LSState& st, double beta) {
    const int tid = tm.tid(), nt = tm.nt();
    for (int i = tid; i < st.ld; i += nt) st.g[i] = i == 0 ? beta : 0.0;
    for (int idx = tid; idx < st.ld * st.maxm; idx += nt) {
        st.H[idx] = 0.0;
        st.R[idx] = 0.0;
    }
    if (tid == 0) {
        st.meta[0] = 0;
        st.meta[1] = 0;
        st.res[0] = beta;
    }
    tm.sync();
}

}  // namespace bsp::rs
