// bsp_spgemm.hip -- block sparse matrix-matrix products C = A B and C = A^T B (bsmm,
// blocksp_multiply.hpp:51-136) and the transposed products A^T x.
//
// Symbolic pass (host, once per structure): row i of C lists the block columns j in the
// order they are first reached by walking row i of A in its stored order and, for each
// A(i, k), row k of B in its stored order -- the reference's first-touch order, so the
// structure of C is bit-exact.  The same walk records, for every block of C, the (A block,
// B block) pairs that contribute to it, in contribution order.
//
// Numeric pass (device): one thread per entry (C block t, row r, column c) sums over the
// block's pairs s_pair = sum_k A_pair[r][k] B_pair[k][c], adding the pair products in the
// reference's order.  For 8x8 blocks a C block is one wavefront; the A row and B column
// a lane reads are shared through L1/L2 by the lanes of the block.
#include <algorithm>
#include <cstring>

#include "bsp_internal.hpp"

namespace bsp {

namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void k_spgemm_numeric(int64_t nnzc, int bs, const int* __restrict__ pair_ptr,
                                                           const int2* __restrict__ pairs,
                                                           const double* __restrict__ A, const double* __restrict__ B,
                                                           double* __restrict__ C) {
    const int b2 = bs * bs;
    const int64_t g = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (g >= nnzc * b2) return;
    const int64_t t = g / b2;
    const int e = (int)(g - t * b2);
    const int r = e / bs, c = e - r * bs;
    double acc = 0.0;
    for (int q = pair_ptr[t]; q < pair_ptr[t + 1]; ++q) {
        const int2 pr = pairs[q];
        const double* a = A + (size_t)pr.x * b2 + r * bs;
        const double* b = B + (size_t)pr.y * b2 + c;
        double s = 0.0;
        for (int k = 0; k < bs; ++k) s = fma(a[k], b[k * bs], s);
        acc += s;
    }
    C[g] = acc;
}

}  // namespace

int sync_host_values(bsp_ctx* ctx, const bsp_bsr* A) {
    if (!A->h_val_stale) return BSP_OK;
    const size_t cnt = (size_t)A->nnzb * A->bs * A->bs;
    A->h_val.resize(cnt);
    BSP_HIP(hipMemcpyAsync(A->h_val.data(), A->val.p, sizeof(double) * cnt, hipMemcpyDeviceToHost, ctx->stream));
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    A->h_val_stale = false;
    return BSP_OK;
}

// Explicit transpose (built once per matrix): row j of A^T holds (A_ij)^T for the rows i
// that hold a block in column j, in ascending i, so each y_j of A^T x accumulates in the
// order the reference's transposed product visits the rows (blocksp_multiply.hpp:196-204).
int ensure_transpose(bsp_ctx* ctx, const bsp_bsr* A) {
    if (A->transposed) return BSP_OK;
    int rc = sync_host_values(ctx, A);
    if (rc) return rc;
    const int n = A->n_rows, m = A->n_cols, bs = A->bs;
    const size_t b2 = (size_t)bs * bs;
    std::vector<int32_t> rp(m + 1, 0), ci(A->nnzb);
    std::vector<double> v((size_t)A->nnzb * b2);
    for (int64_t p = 0; p < A->nnzb; ++p) rp[A->h_ci[p] + 1]++;
    for (int j = 0; j < m; ++j) rp[j + 1] += rp[j];
    std::vector<int32_t> fill(rp.begin(), rp.end() - 1);
    for (int i = 0; i < n; ++i)
        for (int p = A->h_rp[i]; p < A->h_rp[i + 1]; ++p) {
            const int q = fill[A->h_ci[p]]++;
            ci[q] = i;
            const double* src = A->h_val.data() + (size_t)p * b2;
            double* dst = v.data() + (size_t)q * b2;
            for (int r = 0; r < bs; ++r)
                for (int c = 0; c < bs; ++c) dst[c * bs + r] = src[r * bs + c];
        }
    bsp_bsr* T = nullptr;
    rc = bsp_bsr_create(ctx, m, n, bs, rp.data(), ci.data(), v.data(), &T);
    if (rc) return rc;
    A->transposed.reset(T);
    return BSP_OK;
}

// Symbolic pass and first numeric pass of C = L B, where L = A (C = A B) or L = A^T held as
// A's explicit transpose (C = A^T B).  Row j of A^T lists the rows i of A holding column j in
// ascending i, so the first-touch walk over L visits C's blocks in the order the reference's
// transposed product inserts them (blocksp_multiply.hpp:99-134) and adds the pair products
// in the same i order.
int bsmm_build(bsp_ctx* ctx, const bsp_bsr* L, const bsp_bsr* owner, bool transA, const bsp_bsr* B,
               bsp_bsr** out) {
    const int n = L->n_rows, m = B->n_cols, bs = L->bs;
    // symbolic: first-touch columns per row and the contributing pairs per C block
    std::vector<int32_t> crp(n + 1, 0), cci, slot_of(m, -1), touched;
    std::vector<int32_t> pcount;   // pairs per C block
    for (int i = 0; i < n; ++i) {
        touched.clear();
        for (int pa = L->h_rp[i]; pa < L->h_rp[i + 1]; ++pa) {
            const int k = L->h_ci[pa];
            for (int pb = B->h_rp[k]; pb < B->h_rp[k + 1]; ++pb) {
                const int j = B->h_ci[pb];
                if (slot_of[j] < 0) {
                    slot_of[j] = (int32_t)cci.size();
                    cci.push_back(j);
                    pcount.push_back(0);
                    touched.push_back(j);
                }
                pcount[slot_of[j]]++;
            }
        }
        for (int j : touched) slot_of[j] = -1;
        crp[i + 1] = (int32_t)cci.size();
    }
    const int64_t nnzc = (int64_t)cci.size();
    std::vector<int32_t> pptr(nnzc + 1, 0);
    for (int64_t t = 0; t < nnzc; ++t) pptr[t + 1] = pptr[t] + pcount[t];
    std::vector<int32_t> pairs(2 * (size_t)pptr[nnzc]);
    {
        std::vector<int32_t> fill(pptr.begin(), pptr.end() - 1);
        for (int i = 0; i < n; ++i) {
            for (int t = crp[i]; t < crp[i + 1]; ++t) slot_of[cci[t]] = t;
            for (int pa = L->h_rp[i]; pa < L->h_rp[i + 1]; ++pa) {
                const int k = L->h_ci[pa];
                for (int pb = B->h_rp[k]; pb < B->h_rp[k + 1]; ++pb) {
                    const int q = fill[slot_of[B->h_ci[pb]]]++;
                    pairs[2 * (size_t)q] = pa;
                    pairs[2 * (size_t)q + 1] = pb;
                }
            }
            for (int t = crp[i]; t < crp[i + 1]; ++t) slot_of[cci[t]] = -1;
        }
    }
    // C with zero values, then the numeric pass on the device
    const size_t b2 = (size_t)bs * bs;
    auto C = std::make_unique<bsp_bsr>();
    C->ctx = ctx;
    C->n_rows = n;
    C->n_cols = m;
    C->bs = bs;
    C->nnzb = nnzc;
    C->h_rp = crp;
    C->h_ci = cci;
    BSP_HIP(C->rp.alloc(sizeof(int32_t) * (n + 1)));
    BSP_HIP(C->ci.alloc(sizeof(int32_t) * std::max<int64_t>(nnzc, 1)));
    BSP_HIP(C->val.alloc(sizeof(double) * std::max<size_t>(nnzc * b2, 1)));
    BSP_HIP(hipMemcpy(C->rp.p, crp.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice));
    if (nnzc) BSP_HIP(hipMemcpy(C->ci.p, cci.data(), sizeof(int32_t) * nnzc, hipMemcpyHostToDevice));
    BSP_HIP(C->mm_pair_ptr.alloc(sizeof(int32_t) * (nnzc + 1)));
    BSP_HIP(C->mm_pairs.alloc(sizeof(int32_t) * std::max<size_t>(pairs.size(), 2)));
    BSP_HIP(hipMemcpy(C->mm_pair_ptr.p, pptr.data(), sizeof(int32_t) * (nnzc + 1), hipMemcpyHostToDevice));
    if (!pairs.empty())
        BSP_HIP(hipMemcpy(C->mm_pairs.p, pairs.data(), sizeof(int32_t) * pairs.size(), hipMemcpyHostToDevice));
    C->mm_A = owner;
    C->mm_B = B;
    C->mm_Atran = transA;
    C->h_val_stale = true;
    bsp_bsr* raw = C.release();
    int rc = bsp_bsmm_numeric(ctx, owner, B, raw);
    if (rc) {
        bsp_bsr_destroy(raw);
        return rc;
    }
    *out = raw;
    return BSP_OK;
}

}  // namespace bsp

using namespace bsp;

extern "C" {

int bsp_bsmv_transpose(bsp_ctx* ctx, const bsp_bsr* A, const double* x, double* y, int accumulate) {
    if (!ctx || !A) return BSP_E_INVALID;
    BSP_CHECK(x && y && x != y, BSP_E_INVALID, "bsmv_transpose: bad vectors");
    BSP_HIP(hipSetDevice(ctx->device));
    int rc = ensure_transpose(ctx, A);
    if (rc) return rc;
    const bsp_bsr* T = A->transposed.get();
    return launch_bsmv(ctx, T->n_rows, T->bs, T->rp.as<int>(), T->ci.as<int>(), T->val.as<double>(), x, y,
                       accumulate);
}

int bsp_bsmv_dim_transpose(bsp_ctx* ctx, const bsp_bsr* A, int D, int dim, const double* x, double* y,
                           int accumulate) {
    if (!ctx || !A) return BSP_E_INVALID;
    BSP_CHECK(x && y && x != y, BSP_E_INVALID, "bsmv_dim_transpose: bad vectors");
    BSP_CHECK(D >= 1 && dim >= -1 && dim < D, BSP_E_INVALID, "bad dimension arguments");
    BSP_HIP(hipSetDevice(ctx->device));
    int rc = ensure_transpose(ctx, A);
    if (rc) return rc;
    const bsp_bsr* T = A->transposed.get();
    return launch_bsmv_dim(ctx, T->n_rows, T->bs, D, dim, T->rp.as<int>(), T->ci.as<int>(), T->val.as<double>(), x,
                           y, accumulate);
}

int bsp_bsmm(bsp_ctx* ctx, const bsp_bsr* A, const bsp_bsr* B, bsp_bsr** out) {
    if (!ctx || !A || !B || !out) return BSP_E_INVALID;
    *out = nullptr;
    BSP_CHECK(A->n_cols == B->n_rows, BSP_E_INVALID, "bsmm: inner dimensions differ");
    BSP_CHECK(A->bs == B->bs, BSP_E_INVALID, "bsmm: block sizes differ");
    BSP_HIP(hipSetDevice(ctx->device));
    return bsmm_build(ctx, A, A, false, B, out);
}

int bsp_bsmm_trans_a(bsp_ctx* ctx, const bsp_bsr* A, const bsp_bsr* B, bsp_bsr** out) {
    if (!ctx || !A || !B || !out) return BSP_E_INVALID;
    *out = nullptr;
    BSP_CHECK(A->n_rows == B->n_rows, BSP_E_INVALID, "bsmm_trans_a: A and B have different row counts");
    BSP_CHECK(A->bs == B->bs, BSP_E_INVALID, "bsmm_trans_a: block sizes differ");
    BSP_HIP(hipSetDevice(ctx->device));
    int rc = ensure_transpose(ctx, A);
    if (rc) return rc;
    return bsmm_build(ctx, A->transposed.get(), A, true, B, out);
}

int bsp_bsmm_numeric(bsp_ctx* ctx, const bsp_bsr* A, const bsp_bsr* B, bsp_bsr* C) {
    if (!ctx || !A || !B || !C) return BSP_E_INVALID;
    BSP_CHECK(C->mm_A == A && C->mm_B == B && C->mm_pair_ptr.p, BSP_E_INVALID,
              "bsmm_numeric: C was not created by bsp_bsmm(A, B) or bsp_bsmm_trans_a(A, B)");
    BSP_HIP(hipSetDevice(ctx->device));
    const bsp_bsr* L = A;
    if (C->mm_Atran) {
        // A's transpose is rebuilt from its current values when they changed
        int rc = ensure_transpose(ctx, A);
        if (rc) return rc;
        L = A->transposed.get();
    }
    const int64_t total = C->nnzb * C->bs * C->bs;
    if (total > 0) {
        hipLaunchKernelGGL(k_spgemm_numeric, dim3((unsigned)((total + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           ctx->stream, C->nnzb, C->bs, C->mm_pair_ptr.as<int>(), C->mm_pairs.as<int2>(),
                           L->val.as<double>(), B->val.as<double>(), C->val.as<double>());
        BSP_HIP(hipGetLastError());
    }
    C->h_val_stale = true;
    C->val32.release();
    C->val16.release();
    C->valp.release();
    C->transposed.reset();
    return BSP_OK;
}

}  // extern "C"
