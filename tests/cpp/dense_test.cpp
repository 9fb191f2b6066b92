// Host-side checks of the small dense routines of the solvers (blocksp_amd/csrc/bsp_dense.hpp):
// eigenvalues by Francis double-shift steps against closed forms, and the inverse-iteration
// invariant subspace of the deflated restart against the complex-Schur route.  Prints one line
// per check; exits non-zero on a failure.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "../../blocksp_amd/csrc/bsp_dense.hpp"

using bsp::dense::cplx;

static int failures = 0;
static void report(const char* name, double value, double limit) {
    const bool ok = value <= limit && value == value;
    std::printf("%s %s %.3e (limit %.1e)\n", ok ? "OK" : "FAIL", name, value, limit);
    if (!ok) ++failures;
}

// distance between the spans of two n x k orthonormal bases: max_j ||(I - P2 P2^T) P1 e_j||
static double span_distance(int n, int k, const std::vector<double>& P1, const std::vector<double>& P2) {
    double worst = 0.0;
    for (int j = 0; j < k; ++j) {
        std::vector<double> v(n);
        for (int r = 0; r < n; ++r) v[r] = P1[(size_t)r * k + j];
        for (int i = 0; i < k; ++i) {
            double d = 0.0;
            for (int r = 0; r < n; ++r) d += P2[(size_t)r * k + i] * v[r];
            for (int r = 0; r < n; ++r) v[r] -= d * P2[(size_t)r * k + i];
        }
        double s = 0.0;
        for (double e : v) s += e * e;
        worst = std::max(worst, std::sqrt(s));
    }
    return worst;
}

// ||A P - P (P^T A P)||_max for an n x k basis P
static double invariance_residual(int n, int k, const std::vector<double>& A, const std::vector<double>& P) {
    std::vector<double> AP((size_t)n * k, 0.0), B((size_t)k * k, 0.0);
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c)
            for (int j = 0; j < k; ++j) AP[(size_t)r * k + j] += A[(size_t)r * n + c] * P[(size_t)c * k + j];
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j)
            for (int r = 0; r < n; ++r) B[(size_t)i * k + j] += P[(size_t)r * k + i] * AP[(size_t)r * k + j];
    double worst = 0.0;
    for (int r = 0; r < n; ++r)
        for (int j = 0; j < k; ++j) {
            double s = AP[(size_t)r * k + j];
            for (int i = 0; i < k; ++i) s -= P[(size_t)r * k + i] * B[(size_t)i * k + j];
            worst = std::max(worst, std::fabs(s));
        }
    return worst;
}

int main() {
    std::mt19937_64 g(2024);
    std::normal_distribution<double> N(0.0, 1.0);
    // 1. eigenvalues of Q T Q^T with T upper triangular (known diagonal) and of a block matrix
    //    with 2 x 2 rotation-scaling blocks [a b; -b a] (eigenvalues a +- i b)
    {
        const int n = 12;
        std::vector<double> T((size_t)n * n, 0.0), want(n);
        for (int i = 0; i < n; ++i) {
            want[i] = 0.5 + i;
            T[(size_t)i * n + i] = want[i];
            for (int j = i + 1; j < n; ++j) T[(size_t)i * n + j] = N(g);
        }
        // A = Q T Q^T with a random orthogonal Q (from the Hessenberg reduction of a random matrix)
        std::vector<double> R((size_t)n * n);
        for (auto& e : R) e = N(g);
        std::vector<double> Rh(R), Q;
        bsp::dense::hessenberg_q(n, Rh, Q);
        std::vector<double> QT((size_t)n * n, 0.0), A((size_t)n * n, 0.0);
        for (int r = 0; r < n; ++r)
            for (int c = 0; c < n; ++c)
                for (int l = 0; l < n; ++l) QT[(size_t)r * n + c] += Q[(size_t)r * n + l] * T[(size_t)l * n + c];
        for (int r = 0; r < n; ++r)
            for (int c = 0; c < n; ++c)
                for (int l = 0; l < n; ++l) A[(size_t)r * n + c] += QT[(size_t)r * n + l] * Q[(size_t)c * n + l];
        std::vector<double> H(A), Q2;
        bsp::dense::hessenberg_q(n, H, Q2);
        std::vector<cplx> ev;
        const bool ok = bsp::dense::hqr_eigs(n, H, ev);
        std::vector<double> got;
        for (auto& e : ev) got.push_back(e.real());
        std::sort(got.begin(), got.end());
        double err = ok ? 0.0 : 1.0;
        for (int i = 0; i < n; ++i) err = std::max(err, std::fabs(got[i] - want[i]));
        report("hqr_triangular_eigenvalues", err, 1e-10);
    }
    {
        const int n = 10;   // 5 blocks a_j +- i b_j, then a random orthogonal similarity
        std::vector<double> B((size_t)n * n, 0.0);
        std::vector<cplx> want;
        for (int j = 0; j < n / 2; ++j) {
            const double a = 1.0 + j, b = 0.25 * (j + 1);
            B[(size_t)(2 * j) * n + 2 * j] = a;
            B[(size_t)(2 * j) * n + 2 * j + 1] = b;
            B[(size_t)(2 * j + 1) * n + 2 * j] = -b;
            B[(size_t)(2 * j + 1) * n + 2 * j + 1] = a;
            want.push_back(cplx(a, b));
            want.push_back(cplx(a, -b));
        }
        std::vector<double> R((size_t)n * n);
        for (auto& e : R) e = N(g);
        std::vector<double> Rh(R), Q;
        bsp::dense::hessenberg_q(n, Rh, Q);   // Q: a random orthogonal matrix
        std::vector<double> A((size_t)n * n, 0.0), QB((size_t)n * n, 0.0);
        for (int r = 0; r < n; ++r)
            for (int c = 0; c < n; ++c)
                for (int l = 0; l < n; ++l) QB[(size_t)r * n + c] += Q[(size_t)r * n + l] * B[(size_t)l * n + c];
        for (int r = 0; r < n; ++r)
            for (int c = 0; c < n; ++c)
                for (int l = 0; l < n; ++l) A[(size_t)r * n + c] += QB[(size_t)r * n + l] * Q[(size_t)c * n + l];
        std::vector<double> H(A), Q2;
        bsp::dense::hessenberg_q(n, H, Q2);
        std::vector<cplx> ev;
        const bool ok = bsp::dense::hqr_eigs(n, H, ev);
        double err = ok ? 0.0 : 1.0;
        for (const cplx& w : want) {
            double best = 1e300;
            for (const cplx& e : ev) best = std::min(best, std::abs(e - w));
            err = std::max(err, best);
        }
        report("hqr_complex_pairs", err, 1e-10);
    }
    // 2. fast invariant subspace = Schur-route subspace on seeded random matrices (general,
    //    Hessenberg with small eigenvalues, dense leading block as after a deflated restart)
    {
        const int n = 30, k = 10;
        double worst_span = 0.0, worst_inv = 0.0;
        int fallbacks = 0, mismatches = 0;
        for (int t = 0; t < 60; ++t) {
            std::vector<double> A((size_t)n * n, 0.0);
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) {
                    const int kind = t % 3;
                    if (kind == 0) A[(size_t)i * n + j] = N(g) + (i == j ? 2.0 + 0.1 * i : 0.0);
                    else if (kind == 1 && i <= j + 1) A[(size_t)i * n + j] = 0.3 * N(g) + (i == j ? 0.05 * (i + 1) : 0.0);
                    else if (kind == 2 && (i <= j + 1 || (i < 11 && j < 10))) A[(size_t)i * n + j] = N(g) + (i == j ? 3.0 : 0.0);
                }
            std::vector<double> Pf, Ps;
            const int kf = bsp::dense::invariant_subspace_fast(n, A.data(), k, Pf);
            const int ks = bsp::dense::invariant_subspace(n, A.data(), k, Ps, nullptr);
            if (kf < 0) { ++fallbacks; continue; }
            if (kf != ks) { ++mismatches; continue; }
            worst_span = std::max(worst_span, span_distance(n, kf, Pf, Ps));
            double anorm = 0.0;
            for (double e : A) anorm = std::max(anorm, std::fabs(e));
            worst_inv = std::max(worst_inv, invariance_residual(n, kf, A, Pf) / anorm);
        }
        report("fast_subspace_fallbacks", fallbacks, 0);
        report("fast_subspace_size_mismatches", mismatches, 0);
        report("fast_subspace_equals_schur_subspace", worst_span, 1e-9);
        report("fast_subspace_is_invariant", worst_inv, 1e-10);
    }
    // 3. a defective matrix (one 6 x 6 Jordan block among simple eigenvalues): the fast route
    //    either declines (-1, the solver then takes the Schur route) or returns an invariant basis
    {
        const int n = 16, k = 8;
        std::vector<double> A((size_t)n * n, 0.0);
        for (int i = 0; i < 6; ++i) {
            A[(size_t)i * n + i] = 0.1;
            if (i + 1 < 6) A[(size_t)i * n + i + 1] = 1.0;
        }
        for (int i = 6; i < n; ++i) A[(size_t)i * n + i] = 1.0 + i;
        std::vector<double> P;
        const int kf = bsp::dense::invariant_subspace_fast(n, A.data(), k, P);
        const double inv = kf < 0 ? 0.0 : invariance_residual(n, kf, A, P);
        report("defective_declined_or_invariant", inv, 1e-6);
    }
    // 4. deflate with and without the fast route: the same Q_new span and c_r norm
    {
        const int n = 30, m = 31, k = 10;
        std::vector<double> H((size_t)m * n, 0.0), r(m);
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < n; ++j)
                if (i <= j + 1) H[(size_t)i * n + j] = N(g) + (i == j ? 2.0 : 0.0);
        for (double& e : r) e = N(g);
        std::vector<double> P1, H1, c1, P2, H2, c2;
        const int k1 = bsp::dense::deflate(n, H.data(), r.data(), k, P1, H1, c1, true);
        const int k2 = bsp::dense::deflate(n, H.data(), r.data(), k, P2, H2, c2, false);
        report("deflate_same_size", k1 == k2 && k1 > 0 ? 0.0 : 1.0, 0.0);
        if (k1 == k2 && k1 > 0) {
            report("deflate_same_basis_span", span_distance(m, k1 + 1, P1, P2), 1e-9);
            double n1 = 0.0, n2 = 0.0;
            for (double e : c1) n1 += e * e;
            for (double e : c2) n2 += e * e;
            report("deflate_same_residual_norm", std::fabs(std::sqrt(n1) - std::sqrt(n2)), 1e-10);
        }
    }
    std::printf("%s\n", failures ? "DENSE FAILED" : "DENSE PASSED");
    return failures ? 1 : 0;
}
