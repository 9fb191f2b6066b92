// Drop-in test: reference-style user code against include/blocksp_include.hpp.
// Builds matrices through the inserting operator(), then calls bsmv, the solver
// structs, ez:: one-call solvers and a user lambda preconditioner, checking the
// true residual of every solve on the host.  Prints one line per check and
// exits non-zero on the first failure.
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "blocksp_include.hpp"

// user code names the matrix type BlockSp::BlockSp<T, B*D> (README.md:48)
template <typename T, int M>
using BSM = BlockSp::BlockSp<T, M>;
using BlockSp::solArray;
using BlockSp::TinyMatrix;
using BlockSp::TinyVector;
using BlockSp::bsmv;
using BlockSp::pCG;
using BlockSp::pGMRES;
using BlockSp::ILU0_pGMRES;
using BlockSp::IChol0_pCG;
namespace ez = BlockSp::ez;
namespace preConditioners = BlockSp::preConditioners;
namespace dense = BlockSp::dense;

static unsigned long long g_state = 88172645463325252ull;
static double urand() {   // xorshift64 in [-1, 1)
    g_state ^= g_state << 13;
    g_state ^= g_state >> 7;
    g_state ^= g_state << 17;
    return (double)(g_state >> 11) / (double)(1ull << 52) - 1.0;
}

static int failures = 0;
static void report(const char* name, double value, double limit) {
    const bool ok = value <= limit && value == value;
    std::printf("%s %s %.3e (limit %.1e)\n", ok ? "OK" : "FAIL", name, value, limit);
    if (!ok) ++failures;
}

// 2D 5-point block stencil, nx x ny, blocks BD x BD; spd selects a symmetric matrix.
template <int BD>
BSM<double, BD> stencil(int nx, int ny, bool spd) {
    BSM<double, BD> A(nx * ny);
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            const int i = iy * nx + ix;
            TinyMatrix<double, BD, BD> d(0.0);
            for (int r = 0; r < BD; ++r) d(r, r) = 4.5;
            A(i, i) = d;
            const int nb[4] = {ix > 0 ? i - 1 : -1, ix < nx - 1 ? i + 1 : -1, iy > 0 ? i - nx : -1,
                               iy < ny - 1 ? i + nx : -1};
            for (int j : nb) {
                if (j < 0) continue;
                TinyMatrix<double, BD, BD> o(0.0);
                for (int r = 0; r < BD; ++r) {
                    o(r, r) = -1.0;
                    if (!spd && r + 1 < BD) o(r, r + 1) = 0.1 * urand();
                }
                A(i, j) = o;
            }
        }
    A.shrink_to_fit();
    return A;
}

template <int B, int D>
double relres(const BSM<double, B * D>& A, const solArray<double, B, D>& x, const solArray<double, B, D>& b) {
    // host residual b - A x, independent of the device
    std::vector<double> r(b.data(), b.data() + b.size_flat());
    const int bs = B * D;
    for (int i = 0; i < A.m(); ++i)
        for (size_t k = 0; k < A.colSize(i); ++k) {
            const int j = A.colInd(i, (int)k);
            const auto& blk = A.get_data(i)[k];
            for (int rr = 0; rr < bs; ++rr)
                for (int c = 0; c < bs; ++c) r[i * bs + rr] -= blk(rr, c) * x.data()[j * bs + c];
        }
    double nr = 0, nb = 0;
    for (size_t i = 0; i < r.size(); ++i) {
        nr += r[i] * r[i];
        nb += b.data()[i] * b.data()[i];
    }
    return std::sqrt(nr / nb);
}

using Z = std::complex<double>;

// complex 2D 5-point block stencil: shifted, non-Hermitian blocks (a Helmholtz-like operator)
template <int CB>
BSM<Z, CB> zstencil(int nx, int ny) {
    BSM<Z, CB> A(nx * ny);
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            const int i = iy * nx + ix;
            TinyMatrix<Z, CB, CB> d(Z(0.0));
            for (int r = 0; r < CB; ++r) {
                d(r, r) = Z(4.5, 0.8);
                if (r + 1 < CB) d(r, r + 1) = Z(0.2 * urand(), 0.2 * urand());
            }
            A(i, i) = d;
            const int nb[4] = {ix > 0 ? i - 1 : -1, ix < nx - 1 ? i + 1 : -1, iy > 0 ? i - nx : -1,
                               iy < ny - 1 ? i + nx : -1};
            for (int j : nb) {
                if (j < 0) continue;
                TinyMatrix<Z, CB, CB> o(Z(0.0));
                for (int r = 0; r < CB; ++r) o(r, r) = Z(-1.0, 0.1 * urand());
                A(i, j) = o;
            }
        }
    A.shrink_to_fit();
    return A;
}

template <int CB>
double zrelres(const BSM<Z, CB>& A, const solArray<Z, CB>& x, const solArray<Z, CB>& b) {
    std::vector<Z> r(b.data(), b.data() + b.size_flat());
    for (int i = 0; i < A.m(); ++i)
        for (size_t k = 0; k < A.colSize(i); ++k) {
            const int j = A.colInd(i, (int)k);
            const auto& blk = A.get_data(i)[k];
            for (int rr = 0; rr < CB; ++rr)
                for (int c = 0; c < CB; ++c) r[i * CB + rr] -= blk(rr, c) * x.data()[j * CB + c];
        }
    double nr = 0, nb = 0;
    for (size_t i = 0; i < r.size(); ++i) {
        nr += std::norm(r[i]);
        nb += std::norm(b.data()[i]);
    }
    return std::sqrt(nr / nb);
}

int main() {
    // ---- B=2, D=1 nonsymmetric
    auto A = stencil<2>(20, 16, false);
    solArray<double, 2> b(A.m());
    for (auto& e : b) e(0) = TinyVector<double, 2>(urand(), urand());
    {
        auto y = bsmv(A, b);
        double err = 0.0;
        for (int i = 0; i < A.m(); ++i) {
            TinyVector<double, 2> acc(0.0);
            for (size_t k = 0; k < A.colSize(i); ++k) {
                const auto& blk = A.get_data(i)[k];
                const auto& xv = b(A.colInd(i, (int)k));
                acc(0) += blk(0, 0) * xv(0) + blk(0, 1) * xv(1);
                acc(1) += blk(1, 0) * xv(0) + blk(1, 1) * xv(1);
            }
            err = std::max(err, std::max(std::abs(acc(0) - y(i)(0)), std::abs(acc(1) - y(i)(1))));
        }
        report("bsmv_b2", err, 1e-13);
        solArray<double, 2> acc = b;
        bsmv(acc, A, b);   // accumulating form: acc = b + A b
        double e2 = 0.0;
        for (size_t i = 0; i < acc.size_flat(); ++i) e2 = std::max(e2, std::abs(acc.data()[i] - b.data()[i] - y.data()[i]));
        report("bsmv_b2_accumulate", e2, 1e-12);
    }
    report("ilu0_pgmres_dr", relres<2, 1>(A, ILU0_pGMRES<double, 2>(A, 1e-10, 5000, 30, 10).solve(A, b), b), 1e-8);
    report("ilu0_pgmres", relres<2, 1>(A, ez::ilu0_pGMRES<double, 2>(A, b), b), 1e-6);
    report("sor_pgmres_dr", relres<2, 1>(A, ez::sor_pGMRES_dr<double, 2>(A, b), b), 1e-6);
    report("sor_pgmres_robust", relres<2, 1>(A, ez::sor_pGMRES_robust<double, 2>(A, b), b), 1e-6);
    {
        preConditioners::SOR<dense::LUFactor<double, 2>, double, 2> sor(A, 1.2, 3, 0);
        auto x = pGMRES<double, 2, 1>(A, b, 1e-10, sor, 5000, 30, 10);
        report("pgmres_sor_struct", relres<2, 1>(A, x, b), 1e-8);
    }
    {
        // user preconditioner: a lambda doing block-Jacobi on the host (the PreCond concept)
        auto jac = [&](solArray<double, 2>& r) {
            for (int i = 0; i < A.m(); ++i) {
                const auto& d = A(i, i);
                const double det = d(0, 0) * d(1, 1) - d(0, 1) * d(1, 0);
                const double r0 = r(i)(0), r1 = r(i)(1);
                r(i)(0) = (d(1, 1) * r0 - d(0, 1) * r1) / det;
                r(i)(1) = (-d(1, 0) * r0 + d(0, 0) * r1) / det;
            }
        };
        auto x = pGMRES<double, 2, 1>(A, b, 1e-10, jac, 5000, 30, 10);
        report("pgmres_lambda_precond", relres<2, 1>(A, x, b), 1e-8);
        auto x2 = pGMRES<double, 2, 1>(A, b, 1e-10, preConditioners::NoPre<double, 2>(), 5000, 30, 10);
        report("pgmres_nopre", relres<2, 1>(A, x2, b), 1e-8);
    }
    // ---- SPD, B=4 D=1: pCG
    auto S = stencil<4>(16, 12, true);
    solArray<double, 4> bs4(S.m());
    for (auto& e : bs4) e(0) = TinyVector<double, 4>(urand(), urand(), urand(), urand());
    report("ichol0_pcg", relres<4, 1>(S, IChol0_pCG<double, 4>(S, 1e-10).solve(S, bs4), bs4), 1e-8);
    report("ez_ichol0_pcg", relres<4, 1>(S, ez::ichol0_pCG<double, 4>(S, bs4), bs4), 1e-6);
    report("ez_sor_pcg", relres<4, 1>(S, ez::sor_pCG<double, 4>(S, bs4), bs4), 1e-6);
    {
        preConditioners::Diag<dense::CholeskyFactor<double, 4>, double, 4> dg(S);
        report("pcg_diag", relres<4, 1>(S, pCG<double, 4, 1>(S, bs4, 1e-10, dg, 5000), bs4), 1e-8);
    }
    // ---- D=2: BlockSp<double, 4> with solArray<double, 2, 2>
    auto A4 = stencil<4>(12, 10, false);
    solArray<double, 2, 2> b22(A4.m());
    for (int i = 0; i < b22.len(); ++i)
        for (int d = 0; d < 2; ++d) b22(i, d) = TinyVector<double, 2>(urand(), urand());
    {
        auto y = bsmv(A4, b22);
        solArray<double, 2, 2> y2(A4.m());
        bsmv(y2, A4, b22);
        double e = 0.0;
        for (size_t i = 0; i < y.size_flat(); ++i) e = std::max(e, std::abs(y.data()[i] - y2.data()[i]));
        report("bsmv_d2_forms_agree", e, 0.0);
    }
    report("ilu0_pgmres_dr_d2", relres<2, 2>(A4, ez::ilu0_pGMRES_dr<double, 2, 2>(A4, b22), b22), 1e-6);
    // ---- float matrices (widened to FP64 on the device)
    BSM<float, 2> Af(A.m());
    for (int i = 0; i < A.m(); ++i)
        for (size_t k = 0; k < A.colSize(i); ++k) {
            TinyMatrix<float, 2, 2> m;
            for (int e = 0; e < 4; ++e) m.data()[e] = (float)A.get_data(i)[k].data()[e];
            Af(i, A.colInd(i, (int)k)) = m;
        }
    solArray<float, 2> bf(A.m());
    for (size_t i = 0; i < bf.size_flat(); ++i) bf.data()[i] = (float)b.data()[i];
    auto xf = ez::ilu0_pGMRES_dr<float, 2>(Af, bf);
    solArray<double, 2> xd(A.m());
    for (size_t i = 0; i < xd.size_flat(); ++i) xd.data()[i] = xf.data()[i];
    report("float_ilu0_pgmres_dr", relres<2, 1>(A, xd, b), 1e-4);
    // ---- modifying a matrix invalidates its device copy
    {
        auto B2 = A;
        auto y0 = bsmv(B2, b);
        B2(0, 0)(0, 0) += 1.0;
        auto y1 = bsmv(B2, b);
        report("cache_invalidation", std::abs((y1(0)(0) - y0(0)(0)) - b(0)(0)), 1e-14);
    }
    // ---- Arnoldi eigenvalues (blocksp_solvers.hpp:298-333): the SPD stencil's components
    // decouple and its eigenvalues are 4.5 - 2 cos(pi j / 9) - 2 cos(pi k / 9).  From e_0 (a
    // corner cell) the Krylov space is the x<->y-symmetric part of component 0 (36 dimensions),
    // so 20 steps stay inside it: the Ritz values lie in [lmin, lmax] and the largest converges.
    {
        auto S = stencil<2>(8, 8, true);
        auto eig = arnoldi_eigenvalues<double, 2>(S, 20);
        const double lmax = 4.5 + 4.0 * std::cos(M_PI / 9.0), lmin = 4.5 - 4.0 * std::cos(M_PI / 9.0);
        report("arnoldi_ritz_count", std::abs((double)eig.real().size() - 20.0), 0.0);
        report("arnoldi_max_eig", std::abs(eig.max_real() - lmax), 1e-5);   // 5.9e-7 after 20 steps
        report("arnoldi_in_spectrum", std::max(0.0, std::max(lmin - eig.min_real(), eig.max_real() - lmax)), 1e-12);
        report("arnoldi_imag", std::max(std::abs(eig.min_imag()), std::abs(eig.max_imag())), 1e-10);
        preConditioners::Diag<dense::LUFactor<double, 2>, double, 2> Md(S);
        auto eigp = arnoldi_eigenvalues<double, 2>(S, 20, &Md, true, false, true);
        report("arnoldi_diag_harmonic_positive", eigp.min_real() > 0.0 ? 0.0 : 1.0, 0.0);
        preConditioners::ILU0<double, 2> Mi(S);
        solArray<double, 2> bs(S.m());
        for (size_t i = 0; i < bs.size_flat(); ++i) bs.data()[i] = urand();
        auto xs = pGMRES<double, 2, 1>(S, bs, 1e-10, Mi, 500, 10, 4, false, true);   // restarted: harmonic info
        report("pgmres_with_eig_info", relres<2, 1>(S, xs, bs), 1e-8);
    }
    // ---- bsmm / bs_id / bs_tr and matrix arithmetic, as in the reference's pseudo-heat test
    // (blocksp_tests.hpp:270-285): ASym = I - delta A^T A is SPD, then solved with pCG.
    {
        constexpr int BD = 4;
        const int l = 40;
        BSM<double, BD> R(l, l);
        for (int i = 0; i < l; ++i)
            for (int k = 0; k < 3; ++k) {
                const int j = (int)((urand() + 1.0) * 0.5 * l) % l;
                auto& blk = R(i, j);
                for (int e = 0; e < BD * BD; ++e) blk.data()[e] += urand();
            }
        auto dense = [&](const BSM<double, BD>& M) {
            std::vector<double> d((size_t)M.m() * BD * M.n() * BD, 0.0);
            const size_t ld = (size_t)M.n() * BD;
            for (int i = 0; i < M.m(); ++i)
                for (size_t k = 0; k < M.colSize(i); ++k) {
                    const int j = M.colInd(i, (int)k);
                    for (int r = 0; r < BD; ++r)
                        for (int c = 0; c < BD; ++c) d[(i * BD + r) * ld + j * BD + c] += M.get_data(i)[k](r, c);
                }
            return d;
        };
        const auto Rd = dense(R);
        const int N = l * BD;
        auto RtR = BlockSp::bsmm(R, R, true);
        auto RR = BlockSp::bsmm(R, R);
        auto RtR2 = BlockSp::bsmm(BlockSp::bs_tr(R), R);   // same product through an explicit transpose
        const auto RtRd = dense(RtR), RRd = dense(RR), RtR2d = dense(RtR2);
        double e1 = 0.0, e2 = 0.0, e3 = 0.0, scale = 0.0;
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j) {
                double t = 0.0, q = 0.0;
                for (int k = 0; k < N; ++k) {
                    t += Rd[(size_t)k * N + i] * Rd[(size_t)k * N + j];
                    q += Rd[(size_t)i * N + k] * Rd[(size_t)k * N + j];
                }
                scale = std::max(scale, std::abs(t));
                e1 = std::max(e1, std::abs(RtRd[(size_t)i * N + j] - t));
                e2 = std::max(e2, std::abs(RRd[(size_t)i * N + j] - q));
                e3 = std::max(e3, std::abs(RtR2d[(size_t)i * N + j] - t));
            }
        report("bsmm_AtA_dense", e1 / scale, 1e-13);
        report("bsmm_AA_dense", e2 / scale, 1e-13);
        report("bsmm_trA_A_dense", e3 / scale, 1e-13);
        bool same_order = RtR.m() == RtR2.m();
        for (int i = 0; same_order && i < RtR.m(); ++i) same_order = RtR.get_colIndex(i) == RtR2.get_colIndex(i);
        report("bsmm_AtA_structure_equals_bs_tr_path", same_order ? 0.0 : 1.0, 0.0);
        report("frobenius_norm_is_sqrt_trace_AtA",
               std::abs(BlockSp::util::frobeniusNorm(R) - std::sqrt(BlockSp::util::trace(RtR))) /
                   BlockSp::util::frobeniusNorm(R), 1e-14);
        const double delta = 1.0 / (2.0 * BlockSp::util::trace(RtR));   // keeps I - delta A^T A SPD
        auto ASym{BlockSp::bs_id<double, BD>(l) - delta * BlockSp::bsmm(R, R, true)};
        ASym.shrink_to_fit();
        double esym = 0.0;
        const auto Sd = dense(ASym);
        for (int i = 0; i < N; ++i)
            for (int j = 0; j < N; ++j)
                esym = std::max(esym, std::abs(Sd[(size_t)i * N + j] - ((i == j ? 1.0 : 0.0) - delta * RtRd[(size_t)i * N + j])));
        report("pseudo_heat_assembly", esym, 1e-15);
        solArray<double, BD> bh(l);
        for (size_t i = 0; i < bh.size_flat(); ++i) bh.data()[i] = urand();
        preConditioners::ILU0<double, BD> Mh(ASym);
        auto xh = pCG<double, BD, 1>(ASym, bh, 1e-12, Mh, 500);
        report("pseudo_heat_pcg", relres<BD, 1>(ASym, xh, bh), 1e-10);
        auto Up = ASym.upTri(), Lo = ASym.lowTri(-1);
        auto Sum = Lo + Up;   // strict lower + upper = the whole matrix
        double esplit = 0.0;
        const auto Sumd = dense(Sum);
        for (size_t i = 0; i < Sd.size(); ++i) esplit = std::max(esplit, std::abs(Sumd[i] - Sd[i]));
        report("upTri_plus_strict_lowTri", esplit, 0.0);
    }
    // ---- complex T: BlockSp<std::complex<double>, 3> (the qc_test path, blocksp_tests.hpp:386-401)
    {
        constexpr int CB = 3;
        auto Az = zstencil<CB>(14, 12);
        solArray<Z, CB> bz(Az.m());
        for (size_t i = 0; i < bz.size_flat(); ++i) bz.data()[i] = Z(urand(), urand());
        auto yz = bsmv(Az, bz);
        double err = 0.0, scale = 0.0;
        for (int i = 0; i < Az.m(); ++i)
            for (int r = 0; r < CB; ++r) {
                Z acc(0.0);
                for (size_t k = 0; k < Az.colSize(i); ++k)
                    for (int c = 0; c < CB; ++c)
                        acc += Az.get_data(i)[k](r, c) * bz.data()[Az.colInd(i, (int)k) * CB + c];
                err = std::max(err, std::abs(acc - yz.data()[i * CB + r]));
                scale = std::max(scale, std::abs(acc));
            }
        report("complex_bsmv", err / scale, 1e-14);
        report("complex_ilu0_pgmres_dr", zrelres<CB>(Az, ez::ilu0_pGMRES_dr<Z, CB>(Az, bz), bz), 1e-6);
        report("complex_ilu0_pgmres", zrelres<CB>(Az, ILU0_pGMRES<Z, CB>(Az, 1e-10, 5000, -1, -1).solve(Az, bz), bz),
               1e-8);
        report("complex_sor_pgmres_robust", zrelres<CB>(Az, ez::sor_pGMRES_robust<Z, CB>(Az, bz), bz), 1e-6);
        // a host lambda preconditioner on complex vectors (block Jacobi with the diagonal's
        // diagonal) through the host callback
        auto zjac = [&](solArray<Z, CB>& r) {
            for (int i = 0; i < Az.m(); ++i)
                for (int c = 0; c < CB; ++c) r.data()[i * CB + c] /= Az(i, i)(c, c);
        };
        report("complex_pgmres_lambda_precond", zrelres<CB>(Az, pGMRES<Z, CB, 1>(Az, bz, 1e-10, zjac, 5000, 20, 5), bz),
               1e-8);
        // A^H A on the device against the host conjugate transpose
        auto AhA = BlockSp::bsmm(Az, Az, true);
        auto Ah = BlockSp::bs_tr(Az);
        auto AhA2 = BlockSp::bsmm(Ah, Az);
        double eh = 0.0, sh = 0.0;
        for (int i = 0; i < AhA.m(); ++i)
            for (size_t k = 0; k < AhA.colSize(i); ++k) {
                const int j = AhA.colInd(i, (int)k);
                for (int r = 0; r < CB; ++r)
                    for (int c = 0; c < CB; ++c) {
                        eh = std::max(eh, std::abs(AhA.get_data(i)[k](r, c) - AhA2(i, j)(r, c)));
                        sh = std::max(sh, std::abs(AhA2(i, j)(r, c)));
                    }
            }
        report("complex_bsmm_AhA_equals_bs_tr_path", eh / sh, 1e-14);
    }
    std::printf("%s\n", failures ? "DROPIN FAILED" : "DROPIN PASSED");
    return failures ? 1 : 0;
}
