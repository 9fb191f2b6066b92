// Host-side checks of the complex dense routines of the complex pGMRES
// (blocksp_amd/csrc/bsp_zdense.hpp): the pivoted complex solve against the residual, the
// Householder least squares against its normal equations, the complex deflated restart's
// identities (orthonormal P_{k+1}, invariant subspace of H~ for the k eigenvalues of smallest
// modulus, c_r = P^H res_vec, H_new = P^H H P_k) and the harmonic Ritz values.  Prints one
// line per check; exits non-zero on a failure.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "../../blocksp_amd/csrc/bsp_zdense.hpp"

using bsp::dense::cplx;

static int failures = 0;
static void report(const char* name, double value, double limit) {
    const bool ok = value <= limit && value == value;
    std::printf("%s %s %.3e (limit %.1e)\n", ok ? "OK" : "FAIL", name, value, limit);
    if (!ok) ++failures;
}

static std::mt19937_64 rng(12345);
static cplx crand() {
    std::normal_distribution<double> nd;
    return cplx(nd(rng), nd(rng));
}

int main() {
    // zsolve: || A x - b || / (||A|| ||x||)
    for (int n : {1, 2, 5, 17}) {
        std::vector<cplx> A((size_t)n * n), b(n), x;
        for (auto& a : A) a = crand();
        for (auto& v : b) v = crand();
        x = b;
        if (!bsp::dense::zsolve(n, A, x)) { report("zsolve_singular", 1, 0); continue; }
        double r = 0, an = 0, xn = 0;
        for (int i = 0; i < n; ++i) {
            cplx s = -b[i];
            for (int j = 0; j < n; ++j) s += A[(size_t)i * n + j] * x[j];
            r = std::max(r, std::abs(s));
            xn = std::max(xn, std::abs(x[i]));
            for (int j = 0; j < n; ++j) an = std::max(an, std::abs(A[(size_t)i * n + j]));
        }
        report("zsolve_residual", r / (an * xn * n), 1e-14);
    }
    // zlstsq: the residual c - H y is orthogonal to H's columns (normal equations)
    for (int n : {1, 4, 12, 30}) {
        const int m = n + 1;
        std::vector<cplx> H((size_t)m * n, cplx(0)), c(m), y;
        for (int i = 0; i < m; ++i)
            for (int j = std::max(0, i - 1); j < n; ++j) H[(size_t)i * n + j] = crand();   // Hessenberg
        for (auto& v : c) v = crand();
        bsp::dense::zlstsq(m, n, H.data(), c.data(), y);
        std::vector<cplx> r(c);
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < n; ++j) r[i] -= H[(size_t)i * n + j] * y[j];
        double worst = 0, scale = 0;
        for (int j = 0; j < n; ++j) {
            cplx s = 0;
            for (int i = 0; i < m; ++i) s += std::conj(H[(size_t)i * n + j]) * r[i];
            worst = std::max(worst, std::abs(s));
        }
        for (auto& v : H) scale = std::max(scale, std::abs(v));
        report("zlstsq_normal_equations", worst / (scale * scale * m), 1e-13);
    }
    // zdeflate
    for (int k : {1, 3, 6}) {
        const int n = 14, kp = k + 1;
        std::vector<cplx> H((size_t)(n + 1) * n, cplx(0)), res(n + 1), P, Hnew, c_r;
        for (int i = 0; i <= n; ++i)
            for (int j = std::max(0, i - 1); j < n; ++j) H[(size_t)i * n + j] = crand();
        for (auto& v : res) v = crand();
        const int kk = bsp::dense::zdeflate(n, H, res, k, P, Hnew, c_r);
        report("zdeflate_k", std::abs(kk - k), 0);
        double orth = 0;
        for (int a = 0; a < kp; ++a)
            for (int b = 0; b < kp; ++b) {
                cplx s = 0;
                for (int i = 0; i <= n; ++i) s += std::conj(P[(size_t)i * kp + a]) * P[(size_t)i * kp + b];
                orth = std::max(orth, std::abs(s - (a == b ? cplx(1) : cplx(0))));
            }
        report("zdeflate_orthonormal", orth, 1e-13);
        // H~ = H_n + |h|^2 H_n^{-H} e_n e_n^T from the solve routine
        std::vector<cplx> Ht((size_t)n * n), HnH((size_t)n * n), w(n, cplx(0));
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                Ht[(size_t)i * n + j] = H[(size_t)i * n + j];
                HnH[(size_t)j * n + i] = std::conj(H[(size_t)i * n + j]);
            }
        w[n - 1] = 1.0;
        bsp::dense::zsolve(n, HnH, w);
        for (int i = 0; i < n; ++i) Ht[(size_t)i * n + n - 1] += std::norm(H[(size_t)n * n + n - 1]) * w[i];
        // invariance: || H~ P_k - P_k (P_k^H H~ P_k) ||
        std::vector<cplx> HP((size_t)n * k, cplx(0)), R((size_t)k * k, cplx(0));
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < k; ++j)
                for (int l = 0; l < n; ++l) HP[(size_t)i * k + j] += Ht[(size_t)i * n + l] * P[(size_t)l * kp + j];
        for (int a = 0; a < k; ++a)
            for (int j = 0; j < k; ++j)
                for (int i = 0; i < n; ++i) R[(size_t)a * k + j] += std::conj(P[(size_t)i * kp + a]) * HP[(size_t)i * k + j];
        double inv = 0, hs = 0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < k; ++j) {
                cplx s = HP[(size_t)i * k + j];
                for (int a = 0; a < k; ++a) s -= P[(size_t)i * kp + a] * R[(size_t)a * k + j];
                inv = std::max(inv, std::abs(s));
            }
        for (auto& v : Ht) hs = std::max(hs, std::abs(v));
        report("zdeflate_invariant_subspace", inv / hs, 1e-10);
        // the k selected eigenvalues are the k of smallest modulus
        std::vector<cplx> evall, evsel;
        bsp::dense::zritz_values(n + 1, n, H, true, evall);
        bsp::dense::Schur S;
        S.compute(k, R.data());
        std::vector<double> a1, a2;
        for (auto& e : evall) a1.push_back(std::abs(e));
        for (int i = 0; i < k; ++i) a2.push_back(std::abs(S.t(i, i)));
        std::sort(a1.begin(), a1.end());
        std::sort(a2.begin(), a2.end());
        double de = 0;
        for (int i = 0; i < k; ++i) de = std::max(de, std::abs(a1[i] - a2[i]) / a1[i]);
        report("zdeflate_smallest_harmonic_ritz", de, 1e-9);
        // c_r = P^H res, H_new = P^H H P_k
        double ec = 0, eh = 0;
        for (int a = 0; a < kp; ++a) {
            cplx s = 0;
            for (int i = 0; i <= n; ++i) s += std::conj(P[(size_t)i * kp + a]) * res[i];
            ec = std::max(ec, std::abs(s - c_r[a]));
            for (int j = 0; j < k; ++j) {
                cplx t = 0;
                for (int i = 0; i <= n; ++i) {
                    cplx hp = 0;
                    for (int l = 0; l < n; ++l) hp += H[(size_t)i * n + l] * P[(size_t)l * kp + j];
                    t += std::conj(P[(size_t)i * kp + a]) * hp;
                }
                eh = std::max(eh, std::abs(t - Hnew[(size_t)a * k + j]));
            }
        }
        report("zdeflate_c_r", ec, 1e-12);
        report("zdeflate_H_new", eh, 1e-11);
    }
    std::printf("%s\n", failures ? "ZDENSE FAILED" : "ZDENSE PASSED");
    return failures ? 1 : 0;
}
