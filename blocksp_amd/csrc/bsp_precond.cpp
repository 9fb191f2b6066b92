// bsp_precond.cpp -- preconditioner setup (host) and application (device sweeps).
//
// Setup follows the reference algorithms (blocksp_preconditioners.hpp):
//  * ILU0 (255-291): right-looking block elimination restricted to A's pattern,
//    L_ki = A_ki U_ii^{-1}, A_kj -= L_ki A_ij when (i,j) is in the pattern.
//  * IChol0 (374-429): on lowTri(A), L_ii = chol(A_ii), L_ki = A_ki L_ii^{-T},
//    L_kj -= L_ki L_ji^T for i < j <= k when A(i,j) is in the pattern.
//  * SOR (116-240) / Diag (75-111): inverses of the diagonal blocks.
// The diagonal-block inverses are formed explicitly so that the device applies
// them as one small matrix-vector product per block row.
// Application runs entirely on the device: forward / backward block sweeps
// over level schedules (bsp_kernels.hip).
#include <algorithm>
#include <cstring>
#include <numeric>

#include "bsp_dense.hpp"
#include "bsp_factor.hpp"
#include "bsp_internal.hpp"

namespace bsp {

namespace {

constexpr int kWaveRows = 64;

int find_col(const std::vector<int32_t>& rp, const std::vector<int32_t>& ci, int i, int j) {
    for (int p = rp[i]; p < rp[i + 1]; ++p)
        if (ci[p] == j) return p;
    return -1;
}

// Level schedule of a sweep.  Rows are visited in `topo` (a topological order of
// the dependencies); row i waits on the columns c of its row with dep(i, c):
// level(i) = 1 + max level(c).  Each level is padded to a multiple of the rows
// one wavefront handles; rows within a level are in ascending index order.
template <class Dep>
void build_schedule_topo(Sweep& S, int n, int bs, const std::vector<int32_t>& rp, const std::vector<int32_t>& ci,
                         const std::vector<int32_t>& topo, Dep dep) {
    S.n = n;
    S.bs = bs;
    S.level_of.assign(n, 0);
    int nlev = 0;
    for (int t = 0; t < n; ++t) {
        const int i = topo[t];
        int l = 0;
        for (int p = rp[i]; p < rp[i + 1]; ++p)
            if (dep(i, ci[p])) l = std::max(l, S.level_of[ci[p]] + 1);
        S.level_of[i] = l;
        nlev = std::max(nlev, l + 1);
    }
    S.nlevels = nlev;
    const int rpw = std::max(1, kWaveRows / bs);
    std::vector<int> cnt(nlev, 0);
    for (int i = 0; i < n; ++i) cnt[S.level_of[i]]++;
    S.max_level_rows = nlev ? *std::max_element(cnt.begin(), cnt.end()) : 0;
    S.level_ptr.assign(nlev + 1, 0);
    for (int l = 0; l < nlev; ++l) S.level_ptr[l + 1] = S.level_ptr[l] + (cnt[l] + rpw - 1) / rpw * rpw;
    S.max_level_slots = 0;
    for (int l = 0; l < nlev; ++l) S.max_level_slots = std::max(S.max_level_slots, S.level_ptr[l + 1] - S.level_ptr[l]);
    S.nslots = S.level_ptr[nlev];
    std::vector<int> order(S.nslots, -1), fill(S.level_ptr.begin(), S.level_ptr.end() - 1);
    for (int i = 0; i < n; ++i) order[fill[S.level_of[i]]++] = i;
    (void)S.order.alloc(sizeof(int) * std::max(S.nslots, 1));
    (void)hipMemcpy(S.order.p, order.data(), sizeof(int) * S.nslots, hipMemcpyHostToDevice);
    (void)S.dlevel_ptr.alloc(sizeof(int) * S.level_ptr.size());
    (void)hipMemcpy(S.dlevel_ptr.p, S.level_ptr.data(), sizeof(int) * S.level_ptr.size(), hipMemcpyHostToDevice);
    (void)S.flags.alloc(sizeof(uint32_t) * std::max(n, 1));
    (void)hipMemset(S.flags.p, 0, sizeof(uint32_t) * std::max(n, 1));
    (void)S.ticket.alloc(sizeof(unsigned long long));
    (void)hipMemset(S.ticket.p, 0, sizeof(unsigned long long));
}

// Natural-order schedule: lower waits on columns < i, upper on columns > i.
void build_schedule(Sweep& S, int n, int bs, const std::vector<int32_t>& rp, const std::vector<int32_t>& ci,
                    bool lower) {
    std::vector<int32_t> topo(n);
    for (int i = 0; i < n; ++i) topo[i] = lower ? i : n - 1 - i;
    if (lower) build_schedule_topo(S, n, bs, rp, ci, topo, [](int i, int c) { return c < i; });
    else build_schedule_topo(S, n, bs, rp, ci, topo, [](int i, int c) { return c > i; });
}

hipError_t upload_f32(DevMem& m, const std::vector<double>& v) {
    std::vector<float> f(v.begin(), v.end());
    hipError_t e = m.alloc(sizeof(float) * std::max<size_t>(f.size(), 1));
    if (e != hipSuccess || f.empty()) return e;
    return hipMemcpy(m.p, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice);
}

// FP64 -> BF16 bits: to FP32 (round to nearest), then the upper 16 bits rounded to nearest
// even; NaN stays a quiet NaN.
uint16_t to_bf16(double v) {
    const float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, sizeof(u));
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

hipError_t upload_bf16(DevMem& m, const std::vector<double>& v) {
    std::vector<uint16_t> h(v.size());
    for (size_t i = 0; i < v.size(); ++i) h[i] = to_bf16(v[i]);
    hipError_t e = m.alloc(sizeof(uint16_t) * std::max<size_t>(h.size(), 1));
    if (e != hipSuccess || h.empty()) return e;
    return hipMemcpy(m.p, h.data(), sizeof(uint16_t) * h.size(), hipMemcpyHostToDevice);
}

hipError_t upload_structure(Sweep& S, const std::vector<int32_t>& rp, const std::vector<int32_t>& ci,
                            const std::vector<double>& val, int bs, int precision = BSP_PREC_F64) {
    const bool f32 = precision != BSP_PREC_F64;
    const size_t nnz = ci.size();
    S.nnzb = (int64_t)nnz;
    hipError_t e;
    if ((e = S.rp.alloc(sizeof(int) * rp.size())) != hipSuccess) return e;
    if ((e = S.ci.alloc(sizeof(int) * std::max<size_t>(nnz, 1))) != hipSuccess) return e;
    if ((e = hipMemcpy(S.rp.p, rp.data(), sizeof(int) * rp.size(), hipMemcpyHostToDevice)) != hipSuccess) return e;
    if (nnz && (e = hipMemcpy(S.ci.p, ci.data(), sizeof(int) * nnz, hipMemcpyHostToDevice)) != hipSuccess) return e;
    if (precision == BSP_PREC_BF16) {
        if ((e = upload_bf16(S.val, val)) != hipSuccess) return e;
    } else if (f32) {
        std::vector<float> fv(val.begin(), val.end());
        if ((e = S.val.alloc(sizeof(float) * std::max<size_t>(fv.size(), 1))) != hipSuccess) return e;
        if (nnz && (e = hipMemcpy(S.val.p, fv.data(), sizeof(float) * fv.size(), hipMemcpyHostToDevice)) != hipSuccess)
            return e;
    } else {
        if ((e = S.val.alloc(sizeof(double) * std::max<size_t>(val.size(), 1))) != hipSuccess) return e;
        if (nnz && (e = hipMemcpy(S.val.p, val.data(), sizeof(double) * val.size(), hipMemcpyHostToDevice)) != hipSuccess)
            return e;
    }
    S.f32 = f32;
    S.bf16 = precision == BSP_PREC_BF16;
    S.drp = S.rp.as<int>();
    S.dci = S.ci.as<int>();
    S.dval = S.val.p;
    return hipSuccess;
}

// Structure of a sweep on the device with room for its values in `precision` (filled on the
// device afterwards).
hipError_t upload_pattern(Sweep& S, const std::vector<int32_t>& rp, const std::vector<int32_t>& ci, int bs,
                          int precision) {
    const size_t nnz = ci.size();
    const size_t esz = precision == BSP_PREC_BF16 ? 2 : precision == BSP_PREC_F32 ? 4 : 8;
    S.nnzb = (int64_t)nnz;
    hipError_t e;
    if ((e = S.rp.alloc(sizeof(int) * rp.size())) != hipSuccess) return e;
    if ((e = S.ci.alloc(sizeof(int) * std::max<size_t>(nnz, 1))) != hipSuccess) return e;
    if ((e = hipMemcpy(S.rp.p, rp.data(), sizeof(int) * rp.size(), hipMemcpyHostToDevice)) != hipSuccess) return e;
    if (nnz && (e = hipMemcpy(S.ci.p, ci.data(), sizeof(int) * nnz, hipMemcpyHostToDevice)) != hipSuccess) return e;
    if ((e = S.val.alloc(esz * std::max<size_t>(nnz * bs * bs, 1))) != hipSuccess) return e;
    S.f32 = precision != BSP_PREC_F64;
    S.bf16 = precision == BSP_PREC_BF16;
    S.drp = S.rp.as<int>();
    S.dci = S.ci.as<int>();
    S.dval = S.val.p;
    return hipSuccess;
}

hipError_t upload_ints(DevMem& m, const std::vector<int32_t>& v) {
    hipError_t e = m.alloc(sizeof(int32_t) * std::max<size_t>(v.size(), 1));
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpy(m.p, v.data(), sizeof(int32_t) * v.size(), hipMemcpyHostToDevice);
}

// Split a square block matrix (host) into its strictly lower / strictly upper parts,
// preserving each row's stored order.
void split_strict(int n, int bs, const std::vector<int32_t>& rp, const std::vector<int32_t>& ci,
                  const std::vector<double>& v, bool lower, std::vector<int32_t>& orp, std::vector<int32_t>& oci,
                  std::vector<double>& ov) {
    const size_t b2 = (size_t)bs * bs;
    orp.assign(n + 1, 0);
    oci.clear();
    ov.clear();
    for (int i = 0; i < n; ++i) {
        for (int p = rp[i]; p < rp[i + 1]; ++p) {
            const int c = ci[p];
            if (lower ? (c < i) : (c > i)) {
                oci.push_back(c);
                ov.insert(ov.end(), v.begin() + p * b2, v.begin() + (p + 1) * b2);
            }
        }
        orp[i + 1] = (int32_t)oci.size();
    }
}

int diag_inverses(bsp_ctx* ctx, const bsp_bsr* A, int inv_kind, std::vector<double>& dinv) {
    if (int rc = sync_host_values(ctx, A)) return rc;
    const int n = A->n_rows, bs = A->bs;
    const size_t b2 = (size_t)bs * bs;
    dinv.assign(n * b2, 0.0);
    std::vector<double> blk(b2), linv(b2);
    for (int i = 0; i < n; ++i) {
        const int p = find_col(A->h_rp, A->h_ci, i, i);
        BSP_CHECK(p >= 0, BSP_E_NUMERIC, "missing diagonal block in row " + std::to_string(i));
        const double* a = A->h_val.data() + p * b2;
        if (inv_kind == 1) {
            std::memcpy(blk.data(), a, sizeof(double) * b2);
            BSP_CHECK(dense::cholesky_lower(bs, blk.data()), BSP_E_NUMERIC,
                      "diagonal block " + std::to_string(i) + " is not positive definite");
            BSP_CHECK(dense::invert(bs, blk.data(), linv.data()), BSP_E_NUMERIC, "singular Cholesky factor");
            // A^{-1} = L^{-T} L^{-1}
            for (int r = 0; r < bs; ++r)
                for (int c = 0; c < bs; ++c) {
                    double s = 0.0;
                    for (int k = 0; k < bs; ++k) s += linv[k * bs + r] * linv[k * bs + c];
                    dinv[i * b2 + r * bs + c] = s;
                }
        } else {
            BSP_CHECK(dense::invert(bs, a, dinv.data() + i * b2), BSP_E_NUMERIC,
                      "singular diagonal block in row " + std::to_string(i));
        }
    }
    return BSP_OK;
}

hipError_t upload(DevMem& m, const std::vector<double>& v) {
    hipError_t e = m.alloc(sizeof(double) * std::max<size_t>(v.size(), 1));
    if (e != hipSuccess) return e;
    if (v.empty()) return hipSuccess;
    return hipMemcpy(m.p, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice);
}

}  // namespace

}  // namespace bsp

using namespace bsp;

extern "C" {

int bsp_precond_create_none(bsp_ctx* ctx, const bsp_bsr* A, bsp_precond** out) {
    if (!ctx || !A || !out) return BSP_E_INVALID;
    auto M = std::make_unique<bsp_precond>();
    M->ctx = ctx;
    M->kind = BSP_PRE_NONE;
    M->n = A->n_rows;
    M->bs = A->bs;
    M->A = A;
    *out = M.release();
    return BSP_OK;
}

int bsp_precond_create_diag(bsp_ctx* ctx, const bsp_bsr* A, int inv_kind, bsp_precond** out) {
    if (!ctx || !A || !out) return BSP_E_INVALID;
    *out = nullptr;
    BSP_CHECK(A->n_rows == A->n_cols, BSP_E_INVALID, "Diag needs a square matrix");
    // CholeskyFactor is real only (dense_solvers.hpp:123, 196)
    BSP_CHECK(A->cbs == 0 || inv_kind == 0, BSP_E_UNSUPPORTED, "complex matrices take LU block inverses only");
    BSP_HIP(hipSetDevice(ctx->device));
    auto M = std::make_unique<bsp_precond>();
    M->ctx = ctx;
    M->kind = BSP_PRE_DIAG;
    M->n = A->n_rows;
    M->bs = A->bs;
    M->A = A;
    M->inv_kind = inv_kind;
    std::vector<double> dinv;
    int rc = diag_inverses(ctx, A, inv_kind, dinv);
    if (rc) return rc;
    BSP_HIP(upload(M->dinv_upper, dinv));
    *out = M.release();
    return BSP_OK;
}

int bsp_precond_create_sor(bsp_ctx* ctx, const bsp_bsr* A, double w, int nu, int type, int inv_kind,
                           bsp_precond** out) {
    return bsp_precond_create_sor_ex(ctx, A, w, nu, type, inv_kind, nullptr, out);
}

int bsp_precond_create_sor_ex(bsp_ctx* ctx, const bsp_bsr* A, double w, int nu, int type, int inv_kind,
                              const int32_t* colour, bsp_precond** out) {
    if (!ctx || !A || !out) return BSP_E_INVALID;
    *out = nullptr;
    BSP_CHECK(A->n_rows == A->n_cols, BSP_E_INVALID, "SOR needs a square matrix");
    BSP_CHECK(A->cbs == 0 || inv_kind == 0, BSP_E_UNSUPPORTED, "complex matrices take LU block inverses only");
    BSP_CHECK(nu >= 0, BSP_E_INVALID, "nu must be >= 0");
    BSP_HIP(hipSetDevice(ctx->device));
    auto M = std::make_unique<bsp_precond>();
    M->ctx = ctx;
    M->kind = BSP_PRE_SOR;
    M->n = A->n_rows;
    M->bs = A->bs;
    M->A = A;
    M->w = w;
    M->nu = nu;
    M->type = type;
    M->inv_kind = inv_kind;
    std::vector<double> dinv;
    int rc = diag_inverses(ctx, A, inv_kind, dinv);
    if (rc) return rc;
    BSP_HIP(upload(M->dinv_upper, dinv));
    if (!colour) {
        build_schedule(M->lower, M->n, M->bs, A->h_rp, A->h_ci, true);
        build_schedule(M->upper, M->n, M->bs, A->h_rp, A->h_ci, false);
    } else {
        // multicolour SOR: the sweeps visit the rows colour by colour (natural order within a
        // colour), i.e. the reference's SOR on P A P^T; a row uses the new values of the rows
        // before it in that order
        const int n = M->n;
        for (int i = 0; i < n; ++i) BSP_CHECK(colour[i] >= 0, BSP_E_INVALID, "colours must be non-negative");
        std::vector<int32_t> elim(n), pos(n);
        std::iota(elim.begin(), elim.end(), 0);
        std::stable_sort(elim.begin(), elim.end(), [&](int a, int b) { return colour[a] < colour[b]; });
        for (int t = 0; t < n; ++t) pos[elim[t]] = t;
        std::vector<int32_t> rev(elim.rbegin(), elim.rend());
        build_schedule_topo(M->lower, n, M->bs, A->h_rp, A->h_ci, elim,
                            [&pos](int i, int c) { return pos[c] < pos[i]; });
        build_schedule_topo(M->upper, n, M->bs, A->h_rp, A->h_ci, rev,
                            [&pos](int i, int c) { return pos[c] > pos[i]; });
        BSP_HIP(M->pos.alloc(sizeof(int32_t) * std::max(n, 1)));
        BSP_HIP(hipMemcpy(M->pos.p, pos.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
        M->elim_order = elim;
    }
    for (Sweep* S : {&M->lower, &M->upper}) {
        S->drp = A->rp.as<int>();
        S->dci = A->ci.as<int>();
        S->dval = A->val.as<double>();
        S->nnzb = A->nnzb;
        S->dpos = colour ? M->pos.as<int>() : nullptr;
    }
    rc = build_sor_ell(ctx, M->lower, A->h_rp, A->h_ci, 1);
    if (!rc) rc = build_sor_ell(ctx, M->upper, A->h_rp, A->h_ci, 0);
    if (rc) return rc;
    const size_t len = (size_t)M->n * M->bs;
    BSP_HIP(M->work0.alloc(sizeof(double) * std::max<size_t>(len, 1)));
    BSP_HIP(M->work1.alloc(sizeof(double) * std::max<size_t>(len, 1)));
    *out = M.release();
    return BSP_OK;
}

}  // extern "C"

namespace bsp {

// Elimination order: colour by colour, natural order within a colour (natural order when
// colour == nullptr); pos is its inverse.
int elimination_order(bsp_ctx* ctx, int n, const int32_t* colour, std::vector<int32_t>& elim,
                      std::vector<int32_t>& pos) {
    elim.assign(n, 0);
    pos.assign(n, 0);
    for (int i = 0; i < n; ++i) elim[i] = i;
    if (colour) {
        for (int i = 0; i < n; ++i) BSP_CHECK(colour[i] >= 0, BSP_E_INVALID, "colours must be non-negative");
        std::stable_sort(elim.begin(), elim.end(), [&](int a, int b) { return colour[a] < colour[b]; });
    }
    for (int t = 0; t < n; ++t) pos[elim[t]] = t;
    return BSP_OK;
}

namespace {

// Block ILU(0) on the host in a given elimination order (pos[i] = step at which row i is
// eliminated; natural order pos[i] = i is the reference's, preconditioners.hpp:274-291), for
// block sizes the device factorisation does not take.  The factor is returned in A's block
// layout (LU[p] for A's block p).
int ilu0_factor_rows_host(bsp_ctx* ctx, const bsp_bsr* A, const std::vector<int32_t>& elim,
                          const std::vector<int32_t>& pos, std::vector<double>& LU) {
    if (int rc = sync_host_values(ctx, A)) return rc;
    const int n = A->n_rows, bs = A->bs;
    const size_t b2 = (size_t)bs * bs;
    const auto& rp = A->h_rp;
    const auto& ci = A->h_ci;
    LU = A->h_val;
    // rows holding a block in column c, from A's pattern
    std::vector<int32_t> cptr(n + 1, 0), crow, cpos;
    for (int p = 0; p < (int)ci.size(); ++p) cptr[ci[p] + 1]++;
    for (int c = 0; c < n; ++c) cptr[c + 1] += cptr[c];
    crow.resize(ci.size());
    cpos.resize(ci.size());
    {
        std::vector<int32_t> fill(cptr.begin(), cptr.end() - 1);
        for (int i = 0; i < n; ++i)
            for (int p = rp[i]; p < rp[i + 1]; ++p) {
                crow[fill[ci[p]]] = i;
                cpos[fill[ci[p]]++] = p;
            }
    }
    std::vector<double> inv(b2), tmp(b2), prod(b2);
    for (int t = 0; t < n; ++t) {
        const int i = elim[t];
        const int pd = find_col(rp, ci, i, i);
        BSP_CHECK(pd >= 0, BSP_E_NUMERIC, "ILU0: missing diagonal block in row " + std::to_string(i));
        BSP_CHECK(dense::invert(bs, LU.data() + pd * b2, inv.data()), BSP_E_NUMERIC,
                  "ILU0: singular pivot block in row " + std::to_string(i));
        for (int q = cptr[i]; q < cptr[i + 1]; ++q) {
            if (pos[crow[q]] <= t) continue;
            double* lki = LU.data() + (size_t)cpos[q] * b2;
            std::memcpy(tmp.data(), lki, sizeof(double) * b2);
            dense::matmul(bs, tmp.data(), inv.data(), lki);
        }
        for (int q = cptr[i]; q < cptr[i + 1]; ++q) {
            const int k = crow[q];
            if (pos[k] <= t) continue;
            const double* lki = LU.data() + (size_t)cpos[q] * b2;
            for (int p = rp[k]; p < rp[k + 1]; ++p) {
                const int j = ci[p];
                if (pos[j] <= t) continue;
                const int pij = find_col(rp, ci, i, j);
                if (pij < 0) continue;
                dense::matmul(bs, lki, LU.data() + (size_t)pij * b2, prod.data());
                double* akj = LU.data() + (size_t)p * b2;
                for (size_t e = 0; e < b2; ++e) akj[e] -= prod[e];
            }
        }
    }
    return BSP_OK;
}

}  // namespace

// Block ILU(0) in the elimination order of `colour`, returned on the host in A's block layout
// (device factorisation for blocks up to 8 x 8, downloaded; the host loop otherwise).
int ilu0_factor_host(bsp_ctx* ctx, const bsp_bsr* A, const int32_t* colour, std::vector<int32_t>& elim,
                     std::vector<int32_t>& pos, std::vector<double>& LU) {
    BSP_CHECK(A->n_rows == A->n_cols, BSP_E_INVALID, "ILU0 needs a square matrix");
    if (int rc = elimination_order(ctx, A->n_rows, colour, elim, pos)) return rc;
    if (!factor_device_supported(A->bs) || !ctx->factor_device) return ilu0_factor_rows_host(ctx, A, elim, pos, LU);
    const int n = A->n_rows;
    std::vector<int32_t> lrp(n + 1, 0), lci;
    for (int i = 0; i < n; ++i) {
        for (int p = A->h_rp[i]; p < A->h_rp[i + 1]; ++p)
            if (pos[A->h_ci[p]] < pos[i]) lci.push_back(A->h_ci[p]);
        lrp[i + 1] = (int32_t)lci.size();
    }
    Sweep lower;
    build_schedule_topo(lower, n, A->bs, lrp, lci, elim, [](int, int) { return true; });
    DevMem dpos, F, dinv, diag;
    BSP_HIP(upload_ints(dpos, pos));
    int rc = factor_device(ctx, A, dpos.as<int>(), lower, false, F, dinv, diag);
    if (rc) return rc;
    LU.resize((size_t)A->nnzb * A->bs * A->bs);
    BSP_HIP(hipMemcpy(LU.data(), F.p, sizeof(double) * LU.size(), hipMemcpyDeviceToHost));
    return BSP_OK;
}

// Device preconditioner from an ILU(0) factor in A's block layout: L holds the blocks
// (i, j) with pos[j] < pos[i], U those with pos[j] > pos[i] and the diagonal.
int ilu0_build(bsp_ctx* ctx, const bsp_bsr* A, const std::vector<int32_t>& elim, const std::vector<int32_t>& pos,
               std::vector<double>& LU, int precision, bsp_precond** out) {
    BSP_CHECK(precision == BSP_PREC_F64 || precision == BSP_PREC_F32 || precision == BSP_PREC_BF16, BSP_E_INVALID,
              "ILU0 precision must be F64, F32 or BF16");
    const bool f32 = precision != BSP_PREC_F64;
    BSP_HIP(hipSetDevice(ctx->device));
    const int n = A->n_rows, bs = A->bs;
    const size_t b2 = (size_t)bs * bs;
    const auto& rp = A->h_rp;
    const auto& ci = A->h_ci;
    auto M = std::make_unique<bsp_precond>();
    M->ctx = ctx;
    M->kind = BSP_PRE_ILU0;
    M->n = n;
    M->bs = bs;
    M->f32 = f32;
    M->bf16 = precision == BSP_PREC_BF16;
    std::vector<double> dinv(n * b2);
    for (int i = 0; i < n; ++i) {
        const int pd = find_col(rp, ci, i, i);
        BSP_CHECK(dense::invert(bs, LU.data() + pd * b2, dinv.data() + i * b2), BSP_E_NUMERIC,
                  "ILU0: singular U block in row " + std::to_string(i));
    }
    // strict parts by elimination position, each row in its stored order
    std::vector<int32_t> lrp(n + 1, 0), lci, urp(n + 1, 0), uci;
    std::vector<double> lv, uv;
    for (int i = 0; i < n; ++i) {
        for (int p = rp[i]; p < rp[i + 1]; ++p) {
            const int c = ci[p];
            if (pos[c] < pos[i]) {
                lci.push_back(c);
                lv.insert(lv.end(), LU.begin() + p * b2, LU.begin() + (p + 1) * b2);
            } else if (pos[c] > pos[i]) {
                uci.push_back(c);
                uv.insert(uv.end(), LU.begin() + p * b2, LU.begin() + (p + 1) * b2);
            }
        }
        lrp[i + 1] = (int32_t)lci.size();
        urp[i + 1] = (int32_t)uci.size();
    }
    BSP_HIP(upload_structure(M->lower, lrp, lci, lv, bs, precision));
    BSP_HIP(upload_structure(M->upper, urp, uci, uv, bs, precision));
    std::vector<int32_t> rev(elim.rbegin(), elim.rend());
    build_schedule_topo(M->lower, n, bs, lrp, lci, elim, [](int, int) { return true; });
    build_schedule_topo(M->upper, n, bs, urp, uci, rev, [](int, int) { return true; });
    int rc = build_sweep_ell(ctx, M->lower, lrp);
    if (!rc) rc = build_sweep_ell(ctx, M->upper, urp);
    if (rc) return rc;
    if (f32) {
        if (M->bf16) BSP_HIP(upload_bf16(M->dinv_upper, dinv));
        else BSP_HIP(upload_f32(M->dinv_upper, dinv));
        BSP_HIP(M->work32.alloc(sizeof(float) * std::max<size_t>((size_t)n * bs, 1)));
    } else {
        BSP_HIP(upload(M->dinv_upper, dinv));
    }
    M->elim_order = elim;
    M->f_rp = rp;
    M->f_ci = ci;
    M->f_val.swap(LU);
    *out = M.release();
    return BSP_OK;
}

// Device ILU(0) preconditioner for blocks up to 8 x 8, factored on the device in the
// elimination order elim / pos: sweep structures and schedules from A's structure (host), the
// factor level by level over the forward sweep's schedule, then the strict parts gathered on
// the device in the storage precision and the inverses of U's diagonal blocks computed there.
int ilu0_create_device(bsp_ctx* ctx, const bsp_bsr* A, const std::vector<int32_t>& elim,
                       const std::vector<int32_t>& pos, int precision, bsp_precond** out) {
    BSP_CHECK(precision == BSP_PREC_F64 || precision == BSP_PREC_F32 || precision == BSP_PREC_BF16, BSP_E_INVALID,
              "ILU0 precision must be F64, F32 or BF16");
    BSP_HIP(hipSetDevice(ctx->device));
    PhaseTimer tm(ctx, "ilu0");
    const int n = A->n_rows, bs = A->bs;
    const size_t b2 = (size_t)bs * bs;
    const auto& rp = A->h_rp;
    const auto& ci = A->h_ci;
    auto M = std::make_unique<bsp_precond>();
    M->ctx = ctx;
    M->kind = BSP_PRE_ILU0;
    M->n = n;
    M->bs = bs;
    M->f32 = precision != BSP_PREC_F64;
    M->bf16 = precision == BSP_PREC_BF16;
    std::vector<int32_t> lrp(n + 1, 0), lci, lmap, urp(n + 1, 0), uci, umap;
    lci.reserve(A->nnzb);
    uci.reserve(A->nnzb);
    lmap.reserve(A->nnzb);
    umap.reserve(A->nnzb);
    for (int i = 0; i < n; ++i) {
        for (int p = rp[i]; p < rp[i + 1]; ++p) {
            const int c = ci[p];
            if (pos[c] < pos[i]) {
                lci.push_back(c);
                lmap.push_back(p);
            } else if (pos[c] > pos[i]) {
                uci.push_back(c);
                umap.push_back(p);
            }
        }
        lrp[i + 1] = (int32_t)lci.size();
        urp[i + 1] = (int32_t)uci.size();
    }
    BSP_HIP(upload_pattern(M->lower, lrp, lci, bs, precision));
    BSP_HIP(upload_pattern(M->upper, urp, uci, bs, precision));
    DevMem dl, du, dpos;
    BSP_HIP(upload_ints(dl, lmap));
    BSP_HIP(upload_ints(du, umap));
    BSP_HIP(upload_ints(dpos, pos));
    tm.mark("sweep structures");
    std::vector<int32_t> rev(elim.rbegin(), elim.rend());
    build_schedule_topo(M->lower, n, bs, lrp, lci, elim, [](int, int) { return true; });
    build_schedule_topo(M->upper, n, bs, urp, uci, rev, [](int, int) { return true; });
    tm.mark("schedules");
    DevMem F, dinv, diag;
    int rc = factor_device(ctx, A, dpos.as<int>(), M->lower, false, F, dinv, diag);
    if (rc) return rc;
    tm.mark("device factorisation");
    rc = gather_blocks(ctx, (int64_t)lmap.size(), bs, dl.as<int>(), F.as<double>(), M->lower.val.p, precision, false);
    if (!rc) rc = gather_blocks(ctx, (int64_t)umap.size(), bs, du.as<int>(), F.as<double>(), M->upper.val.p, precision, false);
    if (rc) return rc;
    const size_t esz = precision == BSP_PREC_BF16 ? 2 : precision == BSP_PREC_F32 ? 4 : 8;
    BSP_HIP(M->dinv_upper.alloc(esz * std::max<size_t>(n * b2, 1)));
    if ((rc = invert_diag_device(ctx, n, bs, diag.as<int>(), F.as<double>(), M->dinv_upper.p, nullptr, precision)))
        return fail(ctx, BSP_E_NUMERIC, "ILU0: " + ctx->last_error);
    if (M->f32) BSP_HIP(M->work32.alloc(sizeof(float) * std::max<size_t>((size_t)n * bs, 1)));
    tm.mark("gathers + inverses");
    rc = build_sweep_ell(ctx, M->lower, lrp);
    if (!rc) rc = build_sweep_ell(ctx, M->upper, urp);
    if (rc) return rc;
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    tm.mark("ELL copies");
    M->elim_order = elim;
    M->f_rp = rp;
    M->f_ci = ci;
    M->fdev.p = F.p;
    M->fdev.bytes = F.bytes;
    F.p = nullptr;
    F.bytes = 0;
    *out = M.release();
    return BSP_OK;
}

}  // namespace bsp

namespace {

int create_ilu0_impl(bsp_ctx* ctx, const bsp_bsr* A, const int32_t* colour, int precision, bsp_precond** out) {
    if (!ctx || !A || !out) return BSP_E_INVALID;
    *out = nullptr;
    BSP_CHECK(A->n_rows == A->n_cols, BSP_E_INVALID, "ILU0 needs a square matrix");
    BSP_HIP(hipSetDevice(ctx->device));
    std::vector<int32_t> elim, pos;
    if (factor_device_supported(A->bs) && ctx->factor_device) {
        if (int rc = elimination_order(ctx, A->n_rows, colour, elim, pos)) return rc;
        return ilu0_create_device(ctx, A, elim, pos, precision, out);
    }
    std::vector<double> LU;
    int rc = ilu0_factor_host(ctx, A, colour, elim, pos, LU);
    if (rc) return rc;
    return ilu0_build(ctx, A, elim, pos, LU, precision, out);
}

}  // namespace

namespace {

// IChol0 (blocksp_preconditioners.hpp:374-429) of the block matrix (rp, ci, val) on the host, in
// its own row order: the factor L on lowTri, each row's blocks in stored order.
int ichol0_factor_rows(bsp_ctx* ctx, int n, int bs, const std::vector<int32_t>& rp, const std::vector<int32_t>& ci,
                       const std::vector<double>& val, std::vector<int32_t>& Lrp, std::vector<int32_t>& Lci,
                       std::vector<double>& Lv) {
    const size_t b2 = (size_t)bs * bs;
    // lowTri (blocksp_containers.hpp:277-285): blocks with col <= row, in row order
    Lrp.assign(n + 1, 0);
    Lci.clear();
    Lv.clear();
    for (int i = 0; i < n; ++i) {
        for (int p = rp[i]; p < rp[i + 1]; ++p)
            if (ci[p] <= i) {
                Lci.push_back(ci[p]);
                Lv.insert(Lv.end(), val.begin() + p * b2, val.begin() + (p + 1) * b2);
            }
        Lrp[i + 1] = (int32_t)Lci.size();
    }
    // rows k > c holding L(k, c)
    std::vector<int32_t> cptr(n + 1, 0), crow, cpos;
    for (int i = 0; i < n; ++i)
        for (int p = Lrp[i]; p < Lrp[i + 1]; ++p)
            if (Lci[p] < i) cptr[Lci[p] + 1]++;
    for (int c = 0; c < n; ++c) cptr[c + 1] += cptr[c];
    crow.resize(cptr[n]);
    cpos.resize(cptr[n]);
    {
        std::vector<int32_t> fill(cptr.begin(), cptr.end() - 1);
        for (int i = 0; i < n; ++i)
            for (int p = Lrp[i]; p < Lrp[i + 1]; ++p)
                if (Lci[p] < i) {
                    crow[fill[Lci[p]]] = i;
                    cpos[fill[Lci[p]]++] = p;
                }
    }
    std::vector<double> linv(b2), tmp(b2), prod(b2), ljt(b2);
    for (int i = 0; i < n; ++i) {
        const int pd = find_col(Lrp, Lci, i, i);
        BSP_CHECK(pd >= 0, BSP_E_NUMERIC, "IChol0: missing diagonal block in row " + std::to_string(i));
        double* lii = Lv.data() + (size_t)pd * b2;
        BSP_CHECK(dense::cholesky_lower(bs, lii), BSP_E_NUMERIC,
                  "IChol0: pivot block " + std::to_string(i) + " is not positive definite");
        BSP_CHECK(dense::invert(bs, lii, linv.data()), BSP_E_NUMERIC, "IChol0: singular pivot");
        // L_ki <- L_ki L_ii^{-T}
        for (int q = cptr[i]; q < cptr[i + 1]; ++q) {
            double* lki = Lv.data() + (size_t)cpos[q] * b2;
            std::memcpy(tmp.data(), lki, sizeof(double) * b2);
            for (int r = 0; r < bs; ++r)
                for (int c = 0; c < bs; ++c) {
                    double s = 0.0;
                    for (int t = 0; t < bs; ++t) s += tmp[r * bs + t] * linv[c * bs + t];
                    lki[r * bs + c] = s;
                }
        }
        for (int q = cptr[i]; q < cptr[i + 1]; ++q) {
            const int k = crow[q];
            const double* lki = Lv.data() + (size_t)cpos[q] * b2;
            for (int p = Lrp[k]; p < Lrp[k + 1]; ++p) {
                const int j = Lci[p];
                if (j <= i || j > k) continue;
                if (find_col(rp, ci, i, j) < 0) continue;
                const int pji = find_col(Lrp, Lci, j, i);
                if (pji < 0) continue;
                dense::transpose_to(bs, Lv.data() + (size_t)pji * b2, ljt.data());
                dense::matmul(bs, lki, ljt.data(), prod.data());
                double* lkj = Lv.data() + (size_t)p * b2;
                for (size_t e = 0; e < b2; ++e) lkj[e] -= prod[e];
            }
        }
    }
    return BSP_OK;
}

// Device IChol0 from a factor whose row i holds the blocks (i, j) with pos[j] <= pos[i] (diagonal
// included); elim lists the rows in elimination order and pos is its inverse.  The forward sweep
// runs over the strict blocks of L in elimination order, the backward sweep over L^T (row i:
// (L_ki)^T for the rows k eliminated after i, in elimination order) in reverse.
int ichol0_build(bsp_ctx* ctx, int n, int bs, std::vector<int32_t>& Lrp, std::vector<int32_t>& Lci,
                 std::vector<double>& Lv, const std::vector<int32_t>& elim, bsp_precond** out) {
    const size_t b2 = (size_t)bs * bs;
    auto M = std::make_unique<bsp_precond>();
    M->ctx = ctx;
    M->kind = BSP_PRE_ICHOL0;
    M->n = n;
    M->bs = bs;
    std::vector<double> dlo(n * b2), dup(n * b2);
    for (int i = 0; i < n; ++i) {
        const int pd = find_col(Lrp, Lci, i, i);
        BSP_CHECK(pd >= 0, BSP_E_NUMERIC, "IChol0: missing diagonal block in row " + std::to_string(i));
        BSP_CHECK(dense::invert(bs, Lv.data() + (size_t)pd * b2, dlo.data() + i * b2), BSP_E_NUMERIC,
                  "IChol0: singular factor block");
        dense::transpose_to(bs, dlo.data() + i * b2, dup.data() + i * b2);
    }
    std::vector<int32_t> lrp(n + 1, 0), lci;
    std::vector<double> lv;
    std::vector<int32_t> ucnt(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        for (int p = Lrp[i]; p < Lrp[i + 1]; ++p) {
            const int c = Lci[p];
            if (c == i) continue;
            lci.push_back(c);
            lv.insert(lv.end(), Lv.begin() + (size_t)p * b2, Lv.begin() + (size_t)(p + 1) * b2);
            ucnt[c + 1]++;
        }
        lrp[i + 1] = (int32_t)lci.size();
    }
    // L^T strictly upper: row c holds (L_kc)^T for the rows k eliminated after c, in elimination order
    std::vector<int32_t> urp(n + 1, 0);
    for (int c = 0; c < n; ++c) urp[c + 1] = urp[c] + ucnt[c + 1];
    std::vector<int32_t> uci(urp[n]), fill(urp.begin(), urp.end() - 1);
    std::vector<double> uv((size_t)urp[n] * b2);
    for (int t = 0; t < n; ++t) {
        const int k = elim[t];
        for (int p = Lrp[k]; p < Lrp[k + 1]; ++p) {
            const int c = Lci[p];
            if (c == k) continue;
            const int q = fill[c]++;
            uci[q] = k;
            dense::transpose_to(bs, Lv.data() + (size_t)p * b2, uv.data() + (size_t)q * b2);
        }
    }
    BSP_HIP(upload_structure(M->lower, lrp, lci, lv, bs));
    BSP_HIP(upload_structure(M->upper, urp, uci, uv, bs));
    std::vector<int32_t> rev(elim.rbegin(), elim.rend());
    build_schedule_topo(M->lower, n, bs, lrp, lci, elim, [](int, int) { return true; });
    build_schedule_topo(M->upper, n, bs, urp, uci, rev, [](int, int) { return true; });
    int rc = build_sweep_ell(ctx, M->lower, lrp);
    if (!rc) rc = build_sweep_ell(ctx, M->upper, urp);
    if (rc) return rc;
    BSP_HIP(upload(M->dinv_lower, dlo));
    BSP_HIP(upload(M->dinv_upper, dup));
    M->elim_order = elim;
    M->f_rp.swap(Lrp);
    M->f_ci.swap(Lci);
    M->f_val.swap(Lv);
    *out = M.release();
    return BSP_OK;
}

// IChol0 on the device in the elimination order of `colour` (natural order when null): the
// factor L on the blocks (k, j) with pos[j] <= pos[k] in A's row order (the reference's
// lowTri(A) for the natural order, IChol0 of P A P^T mapped back otherwise), the forward sweep
// over its strict blocks in elimination order and the backward sweep over L^T (row c: (L_kc)^T
// for the rows k eliminated after c, in elimination order), as ichol0_build.
int ichol0_create_device(bsp_ctx* ctx, const bsp_bsr* A, const int32_t* colour, bsp_precond** out) {
    const int n = A->n_rows, bs = A->bs;
    const size_t b2 = (size_t)bs * bs;
    const auto& rp = A->h_rp;
    const auto& ci = A->h_ci;
    PhaseTimer tm(ctx, "ichol0");
    std::vector<int32_t> elim, pos;
    if (int rc = elimination_order(ctx, n, colour, elim, pos)) return rc;
    auto M = std::make_unique<bsp_precond>();
    M->ctx = ctx;
    M->kind = BSP_PRE_ICHOL0;
    M->n = n;
    M->bs = bs;
    std::vector<int32_t> Lrp(n + 1, 0), Lci, fmap, lrp(n + 1, 0), lci, lmap, ucnt(n + 1, 0);
    Lci.reserve(A->nnzb);
    fmap.reserve(A->nnzb);
    lci.reserve(A->nnzb);
    lmap.reserve(A->nnzb);
    for (int k = 0; k < n; ++k) {
        for (int p = rp[k]; p < rp[k + 1]; ++p) {
            const int c = ci[p];
            if (pos[c] > pos[k]) continue;
            Lci.push_back(c);
            fmap.push_back(p);
            if (c == k) continue;
            lci.push_back(c);
            lmap.push_back(p);
            ucnt[c + 1]++;
        }
        Lrp[k + 1] = (int32_t)Lci.size();
        lrp[k + 1] = (int32_t)lci.size();
    }
    std::vector<int32_t> urp(n + 1, 0);
    for (int c = 0; c < n; ++c) urp[c + 1] = urp[c] + ucnt[c + 1];
    std::vector<int32_t> uci(urp[n]), umap(urp[n]), fill(urp.begin(), urp.end() - 1);
    for (int t = 0; t < n; ++t) {
        const int k = elim[t];
        for (int p = rp[k]; p < rp[k + 1]; ++p) {
            const int c = ci[p];
            if (c == k || pos[c] > pos[k]) continue;
            const int q = fill[c]++;
            uci[q] = k;
            umap[q] = p;
        }
    }
    BSP_HIP(upload_pattern(M->lower, lrp, lci, bs, BSP_PREC_F64));
    BSP_HIP(upload_pattern(M->upper, urp, uci, bs, BSP_PREC_F64));
    DevMem dl, du, df, dpos;
    BSP_HIP(upload_ints(dl, lmap));
    BSP_HIP(upload_ints(du, umap));
    BSP_HIP(upload_ints(df, fmap));
    BSP_HIP(upload_ints(dpos, pos));
    tm.mark("sweep structures");
    std::vector<int32_t> rev(elim.rbegin(), elim.rend());
    build_schedule_topo(M->lower, n, bs, lrp, lci, elim, [](int, int) { return true; });
    build_schedule_topo(M->upper, n, bs, urp, uci, rev, [](int, int) { return true; });
    tm.mark("schedules");
    DevMem F, dinv, diag;
    int rc = factor_device(ctx, A, dpos.as<int>(), M->lower, true, F, dinv, diag);
    if (rc) return rc;
    tm.mark("device factorisation");
    BSP_HIP(M->fdev.alloc(sizeof(double) * std::max<size_t>(fmap.size() * b2, 1)));
    rc = gather_blocks(ctx, (int64_t)lmap.size(), bs, dl.as<int>(), F.as<double>(), M->lower.val.p, BSP_PREC_F64, false);
    if (!rc) rc = gather_blocks(ctx, (int64_t)umap.size(), bs, du.as<int>(), F.as<double>(), M->upper.val.p, BSP_PREC_F64, true);
    if (!rc) rc = gather_blocks(ctx, (int64_t)fmap.size(), bs, df.as<int>(), F.as<double>(), M->fdev.p, BSP_PREC_F64, false);
    if (rc) return rc;
    BSP_HIP(M->dinv_lower.alloc(sizeof(double) * std::max<size_t>(n * b2, 1)));
    BSP_HIP(M->dinv_upper.alloc(sizeof(double) * std::max<size_t>(n * b2, 1)));
    if ((rc = invert_diag_device(ctx, n, bs, diag.as<int>(), F.as<double>(), M->dinv_lower.p, M->dinv_upper.p,
                                 BSP_PREC_F64)))
        return fail(ctx, BSP_E_NUMERIC, "IChol0: singular factor block (" + ctx->last_error + ")");
    tm.mark("gathers + inverses");
    rc = build_sweep_ell(ctx, M->lower, lrp);
    if (!rc) rc = build_sweep_ell(ctx, M->upper, urp);
    if (rc) return rc;
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    tm.mark("ELL copies");
    M->elim_order = elim;
    M->f_rp.swap(Lrp);
    M->f_ci.swap(Lci);
    *out = M.release();
    return BSP_OK;
}

}  // namespace

extern "C" {

int bsp_precond_create_ilu0(bsp_ctx* ctx, const bsp_bsr* A, bsp_precond** out) {
    return create_ilu0_impl(ctx, A, nullptr, BSP_PREC_F64, out);
}

int bsp_precond_create_ilu0_ex(bsp_ctx* ctx, const bsp_bsr* A, const int32_t* colour, int precision,
                               bsp_precond** out) {
    if (!ctx) return BSP_E_INVALID;
    BSP_CHECK(precision == BSP_PREC_F64 || precision == BSP_PREC_F32 || precision == BSP_PREC_BF16, BSP_E_INVALID,
              "precision must be F64, F32 or BF16");
    BSP_CHECK(!A || A->cbs == 0 || precision == BSP_PREC_F64, BSP_E_UNSUPPORTED,
              "complex matrices take an FP64 ILU0 only");
    return create_ilu0_impl(ctx, A, colour, precision, out);
}

int bsp_precond_create_ichol0(bsp_ctx* ctx, const bsp_bsr* A, bsp_precond** out) {
    return bsp_precond_create_ichol0_ex(ctx, A, nullptr, out);
}

int bsp_precond_create_ichol0_ex(bsp_ctx* ctx, const bsp_bsr* A, const int32_t* colour, bsp_precond** out) {
    if (!ctx || !A || !out) return BSP_E_INVALID;
    *out = nullptr;
    BSP_CHECK(A->n_rows == A->n_cols, BSP_E_INVALID, "IChol0 needs a square matrix");
    BSP_CHECK(A->cbs == 0, BSP_E_UNSUPPORTED, "IChol0 is real only (IChol0_pCG, blocksp_solvers.hpp:485)");
    BSP_HIP(hipSetDevice(ctx->device));
    if (factor_device_supported(A->bs) && ctx->factor_device) return ichol0_create_device(ctx, A, colour, out);
    if (int rc = sync_host_values(ctx, A)) return rc;
    const int n = A->n_rows, bs = A->bs;
    const size_t b2 = (size_t)bs * bs;
    std::vector<int32_t> elim(n), pos(n);
    std::iota(elim.begin(), elim.end(), 0);
    std::vector<int32_t> Lrp, Lci;
    std::vector<double> Lv;
    if (!colour) {
        int rc = ichol0_factor_rows(ctx, n, bs, A->h_rp, A->h_ci, A->h_val, Lrp, Lci, Lv);
        if (rc) return rc;
        return ichol0_build(ctx, n, bs, Lrp, Lci, Lv, elim, out);
    }
    // colour by colour, natural order within a colour: IChol0 of P A P^T, mapped back
    for (int i = 0; i < n; ++i) BSP_CHECK(colour[i] >= 0, BSP_E_INVALID, "colours must be non-negative");
    std::stable_sort(elim.begin(), elim.end(), [&](int a, int b) { return colour[a] < colour[b]; });
    for (int t = 0; t < n; ++t) pos[elim[t]] = t;
    std::vector<int32_t> prp(n + 1, 0), pci;
    std::vector<double> pv;
    pci.reserve(A->h_ci.size());
    pv.reserve(A->h_val.size());
    for (int t = 0; t < n; ++t) {
        const int i = elim[t];
        for (int p = A->h_rp[i]; p < A->h_rp[i + 1]; ++p) {
            pci.push_back(pos[A->h_ci[p]]);
            pv.insert(pv.end(), A->h_val.begin() + (size_t)p * b2, A->h_val.begin() + (size_t)(p + 1) * b2);
        }
        prp[t + 1] = (int32_t)pci.size();
    }
    std::vector<int32_t> Prp, Pci;
    std::vector<double> Pv;
    int rc = ichol0_factor_rows(ctx, n, bs, prp, pci, pv, Prp, Pci, Pv);
    if (rc) return rc;
    Lrp.assign(n + 1, 0);
    for (int t = 0; t < n; ++t) Lrp[elim[t] + 1] = Prp[t + 1] - Prp[t];
    for (int i = 0; i < n; ++i) Lrp[i + 1] += Lrp[i];
    Lci.resize(Pci.size());
    Lv.resize(Pv.size());
    for (int t = 0; t < n; ++t) {
        const int i = elim[t];
        int q = Lrp[i];
        for (int p = Prp[t]; p < Prp[t + 1]; ++p, ++q) {
            Lci[q] = elim[Pci[p]];
            std::copy(Pv.begin() + (size_t)p * b2, Pv.begin() + (size_t)(p + 1) * b2, Lv.begin() + (size_t)q * b2);
        }
    }
    return ichol0_build(ctx, n, bs, Lrp, Lci, Lv, elim, out);
}

int bsp_precond_create_host(bsp_ctx* ctx, const bsp_bsr* A, bsp_host_precond_fn fn, void* user, bsp_precond** out) {
    if (!ctx || !A || !out || !fn) return BSP_E_INVALID;
    auto M = std::make_unique<bsp_precond>();
    M->ctx = ctx;
    M->kind = BSP_PRE_HOST;
    M->n = A->n_rows;
    M->bs = A->bs;
    M->A = A;
    M->host_fn = fn;
    M->host_user = user;
    *out = M.release();
    return BSP_OK;
}

int bsp_precond_destroy(bsp_precond* M) {
    delete M;
    return BSP_OK;
}

int bsp_precond_apply(bsp_ctx* ctx, const bsp_precond* Mc, double* r) {
    if (!ctx || !Mc) return BSP_E_INVALID;
    auto* M = const_cast<bsp_precond*>(Mc);   // epochs/tickets advance; the operator stays const
    const int64_t len = (int64_t)M->n * M->bs;
    switch (M->kind) {
        case BSP_PRE_NONE: return BSP_OK;
        case BSP_PRE_DIAG: return launch_block_diag(ctx, M->n, M->bs, M->dinv_upper.as<double>(), r);
        case BSP_PRE_HOST: {
            std::vector<double> h((size_t)len);
            BSP_HIP(hipMemcpyAsync(h.data(), r, sizeof(double) * len, hipMemcpyDeviceToHost, ctx->stream));
            BSP_HIP(hipStreamSynchronize(ctx->stream));
            M->host_fn(h.data(), len, M->host_user);
            BSP_HIP(hipMemcpyAsync(r, h.data(), sizeof(double) * len, hipMemcpyHostToDevice, ctx->stream));
            BSP_HIP(hipStreamSynchronize(ctx->stream));
            return BSP_OK;
        }
        case BSP_PRE_ILU0: {
            if (M->f32) {
                float* w = M->work32.as<float>();
                int rc = launch_d2f(ctx, len, r, w);
                if (!rc) rc = precond_apply_f32(ctx, M, w);
                if (!rc) rc = launch_f2d(ctx, len, w, r);
                return rc;
            }
            return launch_ilu0_sweeps(ctx, M->lower, M->upper, M->dinv_upper.as<double>(), r, M->epoch);
        }
        case BSP_PRE_ICHOL0: {
            int rc = launch_trsv(ctx, M->lower, M->dinv_lower.as<double>(), r, ++M->epoch);
            if (rc) return rc;
            return launch_trsv(ctx, M->upper, M->dinv_upper.as<double>(), r, ++M->epoch);
        }
        case BSP_PRE_SOR: {
            double* buf[2] = {M->work0.as<double>(), M->work1.as<double>()};
            const double* xo = nullptr;   // x0 = 0 (preconditioners.hpp:227-234)
            int cur = 0;
            auto sweep = [&](int fwd) -> int {
                double* xn = buf[cur];
                int rc = launch_sor_sweep(ctx, fwd ? M->lower : M->upper, fwd, M->w, M->dinv_upper.as<double>(),
                                          r, xo, xn, ++M->epoch);
                xo = xn;
                cur ^= 1;
                return rc;
            };
            int rc = BSP_OK;
            if (M->type < 0) {
                for (int s = 0; s < M->nu && !rc; ++s) rc = sweep(0);
            } else if (M->type == 0) {
                for (int s = 0; s < M->nu && !rc; ++s) rc = sweep(1);
                for (int s = 0; s < M->nu && !rc; ++s) rc = sweep(0);
            } else {
                for (int s = 0; s < M->nu && !rc; ++s) rc = sweep(1);
            }
            if (rc) return rc;
            if (xo) return launch_copy(ctx, len, xo, r);
            BSP_HIP(hipMemsetAsync(r, 0, sizeof(double) * len, ctx->stream));
            return BSP_OK;
        }
    }
    return fail(ctx, BSP_E_INVALID, "unknown preconditioner");
}

int bsp_precond_levels(const bsp_precond* M, int lower, int32_t* level, int* nlevels) {
    if (!M) return BSP_E_INVALID;
    const Sweep& S = lower ? M->lower : M->upper;
    if (S.level_of.empty() && M->n > 0) return BSP_E_UNSUPPORTED;
    if (level) std::copy(S.level_of.begin(), S.level_of.end(), level);
    if (nlevels) *nlevels = S.nlevels;
    return BSP_OK;
}

int bsp_precond_ell_width(const bsp_precond* M, int lower) {
    if (!M) return BSP_E_INVALID;
    return lower ? M->lower.ell_k : M->upper.ell_k;
}

int bsp_precond_factor_info(const bsp_precond* M, int64_t* nnzb) {
    if (!M) return BSP_E_INVALID;
    if (nnzb) *nnzb = (int64_t)M->f_ci.size();
    return BSP_OK;
}

int bsp_precond_download_factor(bsp_ctx* ctx, const bsp_precond* M, int32_t* rowptr, int32_t* colidx, double* vals) {
    if (!ctx || !M) return BSP_E_INVALID;
    BSP_CHECK(!M->f_rp.empty(), BSP_E_UNSUPPORTED, "this preconditioner has no factor");
    if (rowptr) std::copy(M->f_rp.begin(), M->f_rp.end(), rowptr);
    if (colidx) std::copy(M->f_ci.begin(), M->f_ci.end(), colidx);
    if (vals) {
        if (M->fdev.p) {
            BSP_HIP(hipStreamSynchronize(ctx->stream));
            BSP_HIP(hipMemcpy(vals, M->fdev.p, sizeof(double) * M->f_ci.size() * M->bs * M->bs,
                              hipMemcpyDeviceToHost));
        } else {
            std::copy(M->f_val.begin(), M->f_val.end(), vals);
        }
    }
    return BSP_OK;
}

}  // extern "C"

namespace bsp {

int precond_apply_f32(bsp_ctx* ctx, const bsp_precond* Mc, float* r) {
    auto* M = const_cast<bsp_precond*>(Mc);
    if (!M->f32) return fail(ctx, BSP_E_INVALID, "FP32 apply needs an FP32 preconditioner");
    switch (M->kind) {
        case BSP_PRE_ILU0: return launch_ilu0_sweeps(ctx, M->lower, M->upper, M->dinv_upper.p, r, M->epoch);
        default: return fail(ctx, BSP_E_UNSUPPORTED, "FP32 apply is implemented for ILU0");
    }
}

}  // namespace bsp
