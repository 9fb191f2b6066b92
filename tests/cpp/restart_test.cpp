// Unit test of the team-parallel least squares and deflated restart (bsp_restart.hpp), run on
// the host with a one-thread team, against the host routines of bsp_dense.hpp:
//  * the Givens least-squares residual and solution of fresh cycles against dense::lstsq;
//  * deflate_team against dense::deflate with the inverse-iteration route (same method):
//    same kout, the same subspace (P_k P_k^T: the basis inside it depends on the complex phase
//    of each inverse-iteration vector, which rounding sets), the same residual direction and
//    ||c_r||, and the GMRES-DR identities H P_k = P_{k+1} H_new and P_{k+1}^T P_{k+1} = I;
//  * a cycle after a restart: the Householder QR start plus Givens columns give the same
//    least-squares residual and y as dense::lstsq on the assembled Hessenberg.
// Matrices: Arnoldi Hessenbergs of random nonsymmetric matrices (m = n + 1, n = 8 .. 40).
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "../../blocksp_amd/csrc/bsp_dense.hpp"
#include "../../blocksp_amd/csrc/bsp_restart.hpp"

using namespace bsp;

static int failures = 0;
#define CHECK(cond, ...)                          \
    do {                                          \
        if (!(cond)) {                            \
            std::printf("FAIL: " __VA_ARGS__);    \
            std::printf("\n");                    \
            ++failures;                           \
        }                                         \
    } while (0)

// Arnoldi on a random N x N matrix (shifted to make it well conditioned): H is (n+1) x n
// row-major.
static std::vector<double> arnoldi_h(int N, int n, unsigned seed, double shift) {
    std::mt19937 rng(seed);
    std::normal_distribution<double> nd;
    std::vector<double> A((size_t)N * N);
    for (auto& a : A) a = nd(rng) / std::sqrt((double)N);
    for (int i = 0; i < N; ++i) A[(size_t)i * N + i] += shift;
    std::vector<std::vector<double>> V(1, std::vector<double>(N));
    double nr = 0.0;
    for (auto& v : V[0]) {
        v = nd(rng);
        nr += v * v;
    }
    for (auto& v : V[0]) v /= std::sqrt(nr);
    std::vector<double> H((size_t)(n + 1) * n, 0.0);
    for (int j = 0; j < n; ++j) {
        std::vector<double> w(N, 0.0);
        for (int r = 0; r < N; ++r)
            for (int c = 0; c < N; ++c) w[r] += A[(size_t)r * N + c] * V[j][c];
        for (int i = 0; i <= j; ++i) {
            double d = 0.0;
            for (int r = 0; r < N; ++r) d += V[i][r] * w[r];
            H[(size_t)i * n + j] = d;
            for (int r = 0; r < N; ++r) w[r] -= d * V[i][r];
        }
        double h = 0.0;
        for (double x : w) h += x * x;
        h = std::sqrt(h);
        H[(size_t)(j + 1) * n + j] = h;
        for (auto& x : w) x /= h;
        V.push_back(w);
    }
    return H;
}

struct LS {
    int maxm, ld;
    std::vector<double> H, R, cs, sn, g, Q0t, y, res;
    std::vector<int> meta;
    explicit LS(int m) : maxm(m), ld(m + 1) {
        H.assign((size_t)ld * maxm, 0.0);
        R = H;
        cs.assign(maxm, 0.0);
        sn = cs;
        y = cs;
        g.assign(ld, 0.0);
        res.assign(maxm + 1, 0.0);
        Q0t.assign((rs::kMaxN + 1) * (rs::kMaxN + 1), 0.0);
        meta.assign(2, 0);
    }
    rs::LSState st() {
        return {H.data(), R.data(), cs.data(), sn.data(), g.data(), Q0t.data(), y.data(), res.data(), meta.data(),
                ld, maxm};
    }
};

static double maxabs_diff(const std::vector<double>& a, const std::vector<double>& b) {
    double d = 0.0;
    for (size_t i = 0; i < a.size(); ++i) d = std::max(d, std::fabs(a[i] - b[i]));
    return d;
}

int main() {
    const rs::HostTeam tm;
    int cases = 0;
    for (int n : {8, 12, 20, 30, 40}) {
        for (unsigned seed = 1; seed <= 4; ++seed) {
            const int m = n + 1;
            const std::vector<double> H = arnoldi_h(3 * n + 7, n, seed * 31 + n, seed % 2 ? 0.3 : 1.5);
            // ---- fresh cycle: Givens columns vs Householder least squares on each prefix ----
            LS ls(n);
            const double beta = 1.7;
            rs::ls_fresh_team(tm, ls.st(), beta);
            std::vector<double> t(rs::ls_column_ws(n)), hcol(m + 1);
            for (int j = 0; j < n; ++j) {
                for (int i = 0; i < j + 2; ++i) hcol[i] = H[(size_t)i * n + j];
                rs::ls_column_team(tm, ls.st(), j, hcol.data(), t.data(), nullptr);
                // reference on the (j+2) x (j+1) prefix
                std::vector<double> Hp((size_t)(j + 2) * (j + 1)), c(j + 2, 0.0), y(j + 1);
                for (int r = 0; r < j + 2; ++r)
                    for (int cc = 0; cc <= j; ++cc) Hp[(size_t)r * (j + 1) + cc] = H[(size_t)r * n + cc];
                c[0] = beta;
                const double rref = dense::lstsq(j + 2, j + 1, Hp.data(), c.data(), y.data());
                CHECK(std::fabs(ls.res[j + 1] - rref) <= 1e-12 * beta + 1e-9 * rref, "n %d seed %u col %d res %.3e vs %.3e",
                      n, seed, j, ls.res[j + 1], rref);
                if (j == n - 1) {
                    std::vector<double> resv(m), resv2(m), wsol(rs::ls_solve_ws(n));
                    rs::ls_solve_team(tm, ls.st(), n, beta, nullptr, 0, resv2.data());
                    const std::vector<double> y_rows(ls.y.begin(), ls.y.begin() + n);
                    rs::ls_solve_team(tm, ls.st(), n, beta, nullptr, 0, resv.data(), wsol.data());
                    {
                        std::vector<double> yd2(ls.y.begin(), ls.y.begin() + n);
                        double ym2 = 0.0;
                        for (double v : y_rows) ym2 = std::max(ym2, std::fabs(v));
                        CHECK(maxabs_diff(yd2, y_rows) <= 1e-10 * ym2, "staged vs row solve differ by %.3e",
                              maxabs_diff(yd2, y_rows));
                        CHECK(maxabs_diff(resv, resv2) <= 1e-10 * beta, "staged vs row residual vector differ by %.3e",
                              maxabs_diff(resv, resv2));
                    }
                    double ym = 0.0;
                    for (double v : y) ym = std::max(ym, std::fabs(v));
                    std::vector<double> yd(ls.y.begin(), ls.y.begin() + n);
                    CHECK(maxabs_diff(yd, y) <= 1e-9 * ym, "n %d seed %u y differs by %.3e", n, seed,
                          maxabs_diff(yd, y));
                    double rn = 0.0;
                    for (double v : resv) rn += v * v;
                    CHECK(std::fabs(std::sqrt(rn) - rref) <= 1e-10 * beta, "n %d seed %u ||res vec|| %.3e vs %.3e", n,
                          seed, std::sqrt(rn), rref);
                    // ---- deflation: team route vs host route ----
                    for (int k : {1, 4, 10}) {
                        if (k >= n) continue;
                        std::vector<double> P, Hnew, cr;
                        const int kref = dense::deflate(n, H.data(), resv.data(), k, P, Hnew, cr, true);
                        std::vector<double> ws(rs::deflate_ws_doubles(n)), Pd((size_t)m * (n + 1)), Pk((size_t)n * n),
                            Hn((size_t)(n + 1) * n), crd(n + 1);
                        const int kd = rs::deflate_team(tm, n, ls.H.data(), ls.ld, resv.data(), k, ws.data(), Pd.data(),
                                                        Pk.data(), Hn.data(), crd.data());
                        ++cases;
                        CHECK(kd == kref, "n %d seed %u k %d: kout %d vs %d", n, seed, k, kd, kref);
                        if (kd != kref || kd < 0) continue;
                        const int kp = kd + 1;
                        // the same subspace: P_k P_k^T (basis-independent), the same residual
                        // direction (last column, up to sign) and the same ||c_r||
                        std::vector<double> Pkr((size_t)n * kd);
                        for (int i = 0; i < n; ++i)
                            for (int j2 = 0; j2 < kd; ++j2) Pkr[(size_t)i * kd + j2] = P[(size_t)i * kp + j2];
                        double dproj = 0.0;
                        for (int a = 0; a < n; ++a)
                            for (int b = 0; b < n; ++b) {
                                double x = 0.0, y = 0.0;
                                for (int l = 0; l < kd; ++l) {
                                    x += Pkr[(size_t)a * kd + l] * Pkr[(size_t)b * kd + l];
                                    y += Pk[(size_t)a * kd + l] * Pk[(size_t)b * kd + l];
                                }
                                dproj = std::max(dproj, std::fabs(x - y));
                            }
                        CHECK(dproj <= 1e-8, "n %d seed %u k %d: subspaces differ by %.3e", n, seed, k, dproj);
                        double sgn = 0.0, dl = 0.0, nc1 = 0.0, nc2 = 0.0;
                        for (int i = 0; i < m && sgn == 0.0; ++i)
                            if (std::fabs(P[(size_t)i * kp + kd]) > 1e-3)
                                sgn = (P[(size_t)i * kp + kd] > 0) == (Pd[(size_t)i * kp + kd] > 0) ? 1.0 : -1.0;
                        for (int i = 0; i < m; ++i)
                            dl = std::max(dl, std::fabs(P[(size_t)i * kp + kd] - sgn * Pd[(size_t)i * kp + kd]));
                        for (int j2 = 0; j2 < kp; ++j2) {
                            nc1 += cr[j2] * cr[j2];
                            nc2 += crd[j2] * crd[j2];
                        }
                        CHECK(dl <= 1e-8, "n %d seed %u k %d: residual direction differs by %.3e", n, seed, k, dl);
                        CHECK(std::fabs(std::sqrt(nc1) - std::sqrt(nc2)) <= 1e-10 * (1.0 + std::sqrt(nc1)),
                              "n %d seed %u k %d: ||c_r|| %.6e vs %.6e", n, seed, k, std::sqrt(nc2), std::sqrt(nc1));
                        // identities: H P_k = P_{k+1} Hnew, P^T P = I (the host route's residual for scale)
                        auto ident = [&](const std::vector<double>& PP, const std::vector<double>& PK,
                                         const std::vector<double>& HN) {
                            double e = 0.0;
                            for (int i = 0; i < m; ++i)
                                for (int j2 = 0; j2 < kd; ++j2) {
                                    double a = 0.0, b = 0.0;
                                    for (int l = 0; l < n; ++l) a += H[(size_t)i * n + l] * PK[(size_t)l * kd + j2];
                                    for (int l = 0; l < kp; ++l) b += PP[(size_t)i * kp + l] * HN[(size_t)l * kd + j2];
                                    e = std::max(e, std::fabs(a - b));
                                }
                            return e;
                        };
                        const double e1 = ident(Pd, Pk, Hn), e1ref = ident(P, Pkr, Hnew);
                        double e2 = 0.0;
                        for (int a = 0; a < kp; ++a)
                            for (int b = 0; b < kp; ++b) {
                                double s = 0.0;
                                for (int i = 0; i < m; ++i) s += Pd[(size_t)i * kp + a] * Pd[(size_t)i * kp + b];
                                e2 = std::max(e2, std::fabs(s - (a == b ? 1.0 : 0.0)));
                            }
                        CHECK(e1 <= 1e-10 || e1 <= 10.0 * e1ref, "n %d seed %u k %d: H P_k - P Hnew = %.3e (host %.3e)", n,
                              seed, k, e1, e1ref);
                        CHECK(e2 <= 1e-12, "n %d seed %u k %d: P^T P - I = %.3e", n, seed, k, e2);
                        // ---- the next cycle: Householder start + Givens columns vs lstsq ----
                        LS l2(n);
                        std::vector<double> w2(rs::ls_restart_ws(kd));
                        rs::ls_restart_team(tm, l2.st(), kd, Hn.data(), crd.data(), w2.data());
                        {
                            std::vector<double> Hp((size_t)kp * kd), y2(kd);
                            for (int i = 0; i < kp * kd; ++i) Hp[i] = Hn[i];
                            const double r0 = dense::lstsq(kp, kd, Hp.data(), crd.data(), y2.data());
                            CHECK(std::fabs(l2.res[kd] - r0) <= 1e-10 * (1.0 + r0), "restart residual %.6e vs %.6e",
                                  l2.res[kd], r0);
                        }
                        // append columns: H2(:, j) for j >= kd from a fresh Arnoldi of another matrix
                        const std::vector<double> Hx = arnoldi_h(3 * n + 7, n, seed + 100, 1.0);
                        std::vector<double> Hfull((size_t)m * n, 0.0);
                        for (int i = 0; i < kp; ++i)
                            for (int j2 = 0; j2 < kd; ++j2) Hfull[(size_t)i * n + j2] = Hn[(size_t)i * kd + j2];
                        for (int j2 = kd; j2 < n; ++j2) {
                            for (int i = 0; i < j2 + 2; ++i) {
                                hcol[i] = Hx[(size_t)i * n + j2];
                                Hfull[(size_t)i * n + j2] = hcol[i];
                            }
                            rs::ls_column_team(tm, l2.st(), j2, hcol.data(), t.data(), nullptr);
                            std::vector<double> Hp((size_t)(j2 + 2) * (j2 + 1)), c(j2 + 2, 0.0), y2(j2 + 1);
                            for (int r = 0; r < j2 + 2; ++r)
                                for (int cc = 0; cc <= j2; ++cc) Hp[(size_t)r * (j2 + 1) + cc] = Hfull[(size_t)r * n + cc];
                            for (int i = 0; i < kp; ++i) c[i] = crd[i];
                            const double rr = dense::lstsq(j2 + 2, j2 + 1, Hp.data(), c.data(), y2.data());
                            CHECK(std::fabs(l2.res[j2 + 1] - rr) <= 1e-10 * (1.0 + rr),
                                  "n %d seed %u k %d col %d after restart: %.6e vs %.6e", n, seed, k, j2,
                                  l2.res[j2 + 1], rr);
                            if (j2 == n - 1) {
                                std::vector<double> rv(m);
                                rs::ls_solve_team(tm, l2.st(), n, 0.0, crd.data(), kp, rv.data());
                                double ym = 0.0;
                                for (double v : y2) ym = std::max(ym, std::fabs(v));
                                std::vector<double> yd(l2.y.begin(), l2.y.begin() + n);
                                CHECK(maxabs_diff(yd, y2) <= 1e-8 * (1.0 + ym), "y after restart differs by %.3e",
                                      maxabs_diff(yd, y2));
                            }
                        }
                    }
                }
            }
        }
    }
    std::printf("%d deflation cases, %d failures\n", cases, failures);
    if (failures == 0) std::printf("RESTART PASSED\n");
    return failures == 0 ? 0 : 1;
}
