// bsp_spgemm.hip -- block sparse matrix-matrix products C = A B and C = A^T B (bsmm,
// blocksp_multiply.hpp:51-136) and the transposed products A^T x.
//
// Symbolic pass (device): row i of C lists the block columns j in the order they are first
// reached by walking row i of A in its stored order and, for each A(i, k), row k of B in its
// stored order -- the reference's first-touch insertion order, so the structure of C is
// bit-exact.  One thread per row walks its candidates and keeps the distinct columns, in
// order, in its own stretch of a scratch array (sized by the row's candidate count); the
// counts are scanned into C's row pointers and the lists copied into its column array.
//
// Numeric pass, 8 x 8 blocks (device, matrix cores): the symbolic pass also lists each C
// block's products in the walk order and spreads each row's blocks over 8 streams (longest
// list first, least loaded stream); one wavefront runs a row's streams two per
// v_mfma_f64_16x16x4f64 tile, each block's sum held in the accumulator from its first
// product to its last.
//
// Numeric pass, other block sizes (device): one workgroup per row of C (bs*bs threads, at least one wavefront;
// thread = entry (r, c) of a block) walks the same (A block, B block) sequence: A(i, k) and a
// batch of B(k, .) blocks are staged in LDS (the batch's loads issued together), each product
// entry is s = sum_m A[r][m] B[m][c] by FMA in m order, and s is added to the register
// accumulator of the C block it lands in -- the reference's order of products per block.
// Rows with more than kSlots blocks are done in chunks of kSlots blocks.  Workgroups are
// remapped so that each XCD works through one contiguous range of rows (the B rows that
// neighbouring rows of C share stay in that XCD's L2).
#include <algorithm>
#include <cstring>

#include <hipcub/hipcub.hpp>

#include "bsp_internal.hpp"

namespace bsp {

namespace {

constexpr int kBlock = 256;
constexpr int kSlots = 16;     // C blocks accumulated in registers per pass over a row
constexpr int kBatch = 4;      // B blocks staged per batch

// candidates of row i: sum over A(i, k) of the length of B's row k
__global__ void k_sg_candidates(int n, const int* __restrict__ arp, const int* __restrict__ aci,
                                const int* __restrict__ brp, int64_t* __restrict__ cand) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t c = 0;
    for (int p = arp[i]; p < arp[i + 1]; ++p) c += brp[aci[p] + 1] - brp[aci[p]];
    cand[i] = c;
}

// distinct columns of row i in first-touch order into scratch[off[i] ..), their count to cnt[i]
__global__ void k_sg_distinct(int n, const int* __restrict__ arp, const int* __restrict__ aci,
                              const int* __restrict__ brp, const int* __restrict__ bci, const int64_t* __restrict__ off,
                              int* __restrict__ scratch, int* __restrict__ cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int* lst = scratch + off[i];
    int m = 0;
    for (int pa = arp[i]; pa < arp[i + 1]; ++pa) {
        const int k = aci[pa];
        for (int pb = brp[k]; pb < brp[k + 1]; ++pb) {
            const int j = bci[pb];
            bool seen = false;
            for (int t = 0; t < m && !seen; ++t) seen = lst[t] == j;
            if (!seen) lst[m++] = j;
        }
    }
    cnt[i] = m;
}

__global__ void k_sg_copy(int n, const int64_t* __restrict__ off, const int* __restrict__ scratch,
                          const int* __restrict__ crp, int* __restrict__ cci) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int* lst = scratch + off[i];
    for (int t = 0, q = crp[i]; q < crp[i + 1]; ++t, ++q) cci[q] = lst[t];
}

// Products of row i grouped by C block: ps[q] = first entry of C block q's (A block, B block)
// list in pairs, the lists in C's column order and each list in the walk order (the order
// the block's sum adds its products in); rows follow each other, so row i's lists fill
// [off[i], off[i + 1]) and ps[nnzc] = off[n].
__global__ void k_sg_pairs(int n, const int* __restrict__ arp, const int* __restrict__ aci,
                           const int* __restrict__ brp, const int* __restrict__ bci, const int64_t* __restrict__ off,
                           const int* __restrict__ crp, const int* __restrict__ cci, int64_t* __restrict__ ps,
                           int2* __restrict__ pairs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c0 = crp[i], nc = crp[i + 1] - c0;
    if (i == n - 1) ps[crp[n]] = off[n];
    if (nc == 0) return;
    int64_t* cur = ps + c0;
    for (int s = 0; s < nc; ++s) cur[s] = 0;
    auto slot = [&](int j) {
        int sl = 0;
        while (sl < nc - 1 && cci[c0 + sl] != j) ++sl;   // every candidate's column is in the row
        return sl;
    };
    for (int pa = arp[i]; pa < arp[i + 1]; ++pa) {
        const int k = aci[pa];
        for (int pb = brp[k]; pb < brp[k + 1]; ++pb) cur[slot(bci[pb])] += 1;
    }
    int64_t run = off[i];
    for (int s = 0; s < nc; ++s) {
        const int64_t c = cur[s];
        cur[s] = run;
        run += c;
    }
    for (int pa = arp[i]; pa < arp[i + 1]; ++pa) {
        const int k = aci[pa];
        for (int pb = brp[k]; pb < brp[k + 1]; ++pb) pairs[cur[slot(bci[pb])]++] = make_int2(pa, pb);
    }
    // cursors now sit at each list's end: shift back to starts
    for (int s = nc - 1; s > 0; --s) cur[s] = cur[s - 1];
    cur[0] = off[i];
}

// The 8 x 8 matrix-core pass runs each row of C as 8 streams of products (two per 16 x 16
// tile, four tiles): a stream takes whole C blocks, one product per step, and writes a
// block out after its last product.  Blocks go to streams longest list first, each to the
// least loaded stream (ties: the earlier block, the lower stream), so a row takes
// K = max stream load steps; the schedule holds 8 K entries per row, step-major.
constexpr int kStreams = 8;

struct SchedPlan {
    int load[kStreams];
    __device__ void reset() {
        for (int h = 0; h < kStreams; ++h) load[h] = 0;
    }
    __device__ int least() const {
        int h = 0;
        for (int g = 1; g < kStreams; ++g)
            if (load[g] < load[h]) h = g;
        return h;
    }
};

__global__ void k_sg_sched_count(int n, const int* __restrict__ crp, const int64_t* __restrict__ ps,
                                 int64_t* __restrict__ steps) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) {
        steps[n] = 0;
        return;
    }
    const int c0 = crp[i], nc = crp[i + 1] - c0;
    int cmax = 0;
    for (int s = 0; s < nc; ++s) cmax = max(cmax, (int)(ps[c0 + s + 1] - ps[c0 + s]));
    SchedPlan pl;
    pl.reset();
    for (int c = cmax; c >= 1; --c)
        for (int s = 0; s < nc; ++s)
            if (ps[c0 + s + 1] - ps[c0 + s] == c) pl.load[pl.least()] += c;
    int k = 0;
    for (int h = 0; h < kStreams; ++h) k = max(k, pl.load[h]);
    steps[i] = (int64_t)k * kStreams;
}

// entry: x = A block, y = B block (x = -1: idle step), z = C block written after this
// product (-1: none)
__global__ void k_sg_sched_fill(int n, const int* __restrict__ crp, const int64_t* __restrict__ ps,
                                const int2* __restrict__ pairs, const int64_t* __restrict__ soff,
                                int4* __restrict__ sched) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c0 = crp[i], nc = crp[i + 1] - c0;
    const int64_t s0 = soff[i];
    const int K = (int)((soff[i + 1] - s0) / kStreams);
    int cmax = 0;
    for (int s = 0; s < nc; ++s) cmax = max(cmax, (int)(ps[c0 + s + 1] - ps[c0 + s]));
    SchedPlan pl;
    pl.reset();
    for (int c = cmax; c >= 1; --c)
        for (int s = 0; s < nc; ++s) {
            const int64_t b = ps[c0 + s];
            if (ps[c0 + s + 1] - b != c) continue;
            const int h = pl.least();
            for (int j = 0; j < c; ++j) {
                const int2 pr = pairs[b + j];
                sched[s0 + (int64_t)(pl.load[h] + j) * kStreams + h] =
                    make_int4(pr.x, pr.y, j == c - 1 ? c0 + s : -1, 0);
            }
            pl.load[h] += c;
        }
    for (int h = 0; h < kStreams; ++h)
        for (int k = pl.load[h]; k < K; ++k) sched[s0 + (int64_t)k * kStreams + h] = make_int4(-1, 0, -1, 0);
}

template <int BS>
struct RowShape {
    static constexpr int G = BS * BS;
    static constexpr int T = G < 64 ? 64 : G;   // threads per workgroup (one row)
};

template <int BS>
__global__ __launch_bounds__(RowShape<BS>::T) void k_spgemm_rows(int n, int nwg, const int* __restrict__ arp,
                                                                 const int* __restrict__ aci,
                                                                 const double* __restrict__ av,
                                                                 const int* __restrict__ brp,
                                                                 const int* __restrict__ bci,
                                                                 const double* __restrict__ bv,
                                                                 const int* __restrict__ crp,
                                                                 const int* __restrict__ cci, double* __restrict__ cv) {
    constexpr int G = RowShape<BS>::G;
    __shared__ double As[G];
    __shared__ double Bs[kBatch][G];
    __shared__ int cols[kSlots];
    __shared__ int bslot[kBatch];
    // XCD-aware row order: XCD x (workgroups x, x + 8, ...) takes one contiguous range of rows
    const int per = (nwg + 7) / 8;
    const int i = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    if (i >= n) return;
    const int e = threadIdx.x;
    const bool act = e < G;
    const int r = e / BS, c = e % BS;
    const int c0 = crp[i], ncol = crp[i + 1] - c0;
    for (int base = 0; base < ncol; base += kSlots) {
        const int cn = min(kSlots, ncol - base);
        double acc[kSlots];
#pragma unroll
        for (int s = 0; s < kSlots; ++s) acc[s] = 0.0;
        __syncthreads();
        if (e < cn) cols[e] = cci[c0 + base + e];
        for (int pa = arp[i]; pa < arp[i + 1]; ++pa) {
            const int k = aci[pa];
            __syncthreads();
            if (act) As[e] = av[(size_t)pa * G + e];
            const int b0 = brp[k], b1 = brp[k + 1];
            for (int pb = b0; pb < b1; pb += kBatch) {
                const int nb = min(kBatch, b1 - pb);
                double bval[kBatch];
#pragma unroll
                for (int q = 0; q < kBatch; ++q)
                    if (q < nb && act) bval[q] = bv[(size_t)(pb + q) * G + e];
                __syncthreads();   // the previous batch's products are done
#pragma unroll
                for (int q = 0; q < kBatch; ++q)
                    if (q < nb && act) Bs[q][e] = bval[q];
                if (e < nb) {
                    const int j = bci[pb + e];
                    int sl = -1;
                    for (int t = 0; t < cn; ++t)
                        if (cols[t] == j) sl = t;
                    bslot[e] = sl;
                }
                __syncthreads();
                if (act) {
#pragma unroll
                    for (int q = 0; q < kBatch; ++q) {
                        if (q >= nb) break;
                        const int sl = bslot[q];
                        if (sl < 0) continue;
                        double sum = 0.0;
#pragma unroll
                        for (int m = 0; m < BS; ++m) sum = fma(As[r * BS + m], Bs[q][m * BS + c], sum);
#pragma unroll
                        for (int s = 0; s < kSlots; ++s)
                            if (s == sl) acc[s] += sum;
                    }
                }
            }
        }
        if (act) {
#pragma unroll
            for (int s = 0; s < kSlots; ++s)
                if (s < cn) cv[(size_t)(c0 + base + s) * G + e] = acc[s];
        }
    }
}

// 8 x 8 blocks on the matrix cores: one wavefront per row of C, four rows per workgroup.
// v_mfma_f64_16x16x4f64 takes, per lane l, A[l % 16][l / 16] and B[l / 16][l % 16] of a
// 16 x 4 by 4 x 16 product and accumulates into D[l / 16 + 4 t][l % 16], t = 0..3 (checked on
// the device with power-of-two operands).  A tile carries two streams: A rows 0-7 and B
// columns 0-7 hold one stream's product, A rows 8-15 and B columns 8-15 the other's, so the
// diagonal quadrants of D accumulate the two streams' current C blocks (the off-diagonal
// ones are never read).  The lanes of a stream (m16 < 8 or m16 >= 8) supply every operand of
// its quadrant and hold all of its results, so a stream writes and clears its quadrant when
// its block ends without touching the other.  A block's sum stays in the accumulator from
// its first product to its last, in the order the schedule lists them (the walk order).
typedef double d4v __attribute__((ext_vector_type(4)));
typedef int i4v __attribute__((ext_vector_type(4)));

// schedule entries and C blocks are touched once: streamed past the caches, which keep the
// A and B blocks that neighbouring rows read again
__device__ inline int4 load_entry(const int4* p) {
    const i4v v = __builtin_nontemporal_load(reinterpret_cast<const i4v*>(p));
    return make_int4(v.x, v.y, v.z, v.w);
}

constexpr int KT = 2;   // tiles per wavefront: each pass runs 2 KT of the row's streams

__global__ __launch_bounds__(256) void k_spgemm_sched_mfma8(int n, int nwg, const int64_t* __restrict__ soff,
                                                           const int4* __restrict__ sched,
                                                           const double* __restrict__ av,
                                                           const double* __restrict__ bv, double* __restrict__ cv) {
    const int per = (nwg + 7) / 8;
    const int wg = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    const int i = wg * 4 + (int)(threadIdx.x >> 6);
    if (wg >= nwg || i >= n) return;   // whole wavefronts leave; no workgroup barrier follows
    const int l = threadIdx.x & 63;
    const int m16 = l & 15, kq = l >> 4, r8 = m16 & 7, half = m16 >= 8 ? 1 : 0;
    const int64_t s0 = soff[i];
    const int K = (int)((soff[i + 1] - s0) / kStreams);
    for (int pass = 0; pass < kStreams / (2 * KT); ++pass) {
        const int4* E = sched + s0 + pass * 2 * KT + half;
        d4v d[KT];
#pragma unroll
        for (int t = 0; t < KT; ++t) d[t] = (d4v){0.0, 0.0, 0.0, 0.0};
        // step k + 1's entries load while step k's operands are in flight
        int4 e[KT];
#pragma unroll
        for (int t = 0; t < KT; ++t) {
            if (K > 0)
                e[t] = load_entry(E + 2 * t);
            else
                e[t] = make_int4(-1, 0, -1, 0);
        }
        for (int k = 0; k < K; ++k) {
            int4 en[KT];
            double xa0[KT], xa1[KT], xb0[KT], xb1[KT];
#pragma unroll
            for (int t = 0; t < KT; ++t) {
                const int4* q = E + (size_t)(k + 1) * kStreams + 2 * t;
                if (k + 1 < K)
                    en[t] = load_entry(q);
                else
                    en[t] = make_int4(-1, 0, -1, 0);
            }
#pragma unroll
            for (int t = 0; t < KT; ++t) {
                xa0[t] = xa1[t] = xb0[t] = xb1[t] = 0.0;
                if (e[t].x >= 0) {
                    const double* A = av + (size_t)e[t].x * 64;
                    const double* Bp = bv + (size_t)e[t].y * 64;
                    xa0[t] = A[r8 * 8 + kq];
                    xa1[t] = A[r8 * 8 + 4 + kq];
                    xb0[t] = Bp[kq * 8 + r8];
                    xb1[t] = Bp[(kq + 4) * 8 + r8];
                }
            }
#pragma unroll
            for (int t = 0; t < KT; ++t) {
                d[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa0[t], xb0[t], d[t], 0, 0, 0);
                d[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa1[t], xb1[t], d[t], 0, 0, 0);
            }
            // lane (kq, m16) holds D[kq + 4 v][m16]: its stream's rows kq, kq + 4 at v = 0, 1
            // (m16 < 8) or v = 2, 3 (m16 >= 8); the other two values are off-diagonal
#pragma unroll
            for (int t = 0; t < KT; ++t) {
                if (e[t].z >= 0) {
                    double* Cb = cv + (size_t)e[t].z * 64;
                    const double v0 = half ? d[t][2] : d[t][0], v1 = half ? d[t][3] : d[t][1];
                    __builtin_nontemporal_store(v0, Cb + kq * 8 + r8);
                    __builtin_nontemporal_store(v1, Cb + (kq + 4) * 8 + r8);
                    d[t] = (d4v){0.0, 0.0, 0.0, 0.0};
                }
            }
#pragma unroll
            for (int t = 0; t < KT; ++t) e[t] = en[t];
        }
    }
}

template <int BS>
void launch_rows(hipStream_t st, int n, const bsp_bsr* L, const bsp_bsr* B, const bsp_bsr* C) {
    const int nwg = n;
    const int grid = (nwg + 7) / 8 * 8;
    hipLaunchKernelGGL(k_spgemm_rows<BS>, dim3(grid), dim3(RowShape<BS>::T), 0, st, n, nwg, L->rp.as<int>(),
                       L->ci.as<int>(), L->val.as<double>(), B->rp.as<int>(), B->ci.as<int>(), B->val.as<double>(),
                       C->rp.as<int>(), C->ci.as<int>(), C->val.as<double>());
}

bool launch_rows_bs(hipStream_t st, int bs, int n, const bsp_bsr* L, const bsp_bsr* B, const bsp_bsr* C,
                    bool mfma) {
    if (bs == 8 && mfma && C->mm_sched.p) {
        const int nwg = (n + 3) / 4, grid = (nwg + 7) / 8 * 8;
        hipLaunchKernelGGL(k_spgemm_sched_mfma8, dim3(grid), dim3(256), 0, st, n, nwg, C->mm_soff.as<int64_t>(),
                           C->mm_sched.as<int4>(), L->val.as<double>(), B->val.as<double>(), C->val.as<double>());
        return true;
    }
    switch (bs) {
        case 1: launch_rows<1>(st, n, L, B, C); return true;
        case 2: launch_rows<2>(st, n, L, B, C); return true;
        case 3: launch_rows<3>(st, n, L, B, C); return true;
        case 4: launch_rows<4>(st, n, L, B, C); return true;
        case 5: launch_rows<5>(st, n, L, B, C); return true;
        case 6: launch_rows<6>(st, n, L, B, C); return true;
        case 7: launch_rows<7>(st, n, L, B, C); return true;
        case 8: launch_rows<8>(st, n, L, B, C); return true;
        case 16: launch_rows<16>(st, n, L, B, C); return true;
        case 32: launch_rows<32>(st, n, L, B, C); return true;
        default: return false;
    }
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
// transposed product inserts them (blocksp_multiply.hpp:99-134) and adds the pair products in
// the same i order.
int bsmm_build(bsp_ctx* ctx, const bsp_bsr* L, const bsp_bsr* owner, bool transA, const bsp_bsr* B,
               bsp_bsr** out) {
    const int n = L->n_rows, m = B->n_cols, bs = L->bs;
    const int tpb = 256, grid = std::max(1, (n + tpb - 1) / tpb);
    DevMem cand, off, cnt, scratch, tmp;
    int64_t total = 0;
    BSP_HIP(cand.alloc(sizeof(int64_t) * (n + 1)));
    BSP_HIP(off.alloc(sizeof(int64_t) * (n + 1)));
    BSP_HIP(cnt.alloc(sizeof(int) * (n + 1)));
    auto C = std::make_unique<bsp_bsr>();
    C->ctx = ctx;
    C->n_rows = n;
    C->n_cols = m;
    C->bs = bs;
    BSP_HIP(C->rp.alloc(sizeof(int32_t) * (n + 1)));
    if (n > 0) {
        hipLaunchKernelGGL(k_sg_candidates, dim3(grid), dim3(tpb), 0, ctx->stream, n, L->rp.as<int>(), L->ci.as<int>(),
                           B->rp.as<int>(), cand.as<int64_t>());
        BSP_HIP(hipGetLastError());
        // candidate offsets (64-bit: the candidate total can exceed 2^31 where C cannot)
        BSP_HIP(hipMemsetAsync(cand.as<int64_t>() + n, 0, sizeof(int64_t), ctx->stream));
        size_t tb = 0;
        BSP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cand.as<int64_t>(), off.as<int64_t>(), n + 1, ctx->stream));
        BSP_HIP(tmp.alloc(std::max<size_t>(tb, 16)));
        BSP_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, cand.as<int64_t>(), off.as<int64_t>(), n + 1, ctx->stream));
        BSP_HIP(hipMemcpyAsync(&total, off.as<int64_t>() + n, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
        BSP_HIP(hipStreamSynchronize(ctx->stream));
        BSP_HIP(scratch.alloc(sizeof(int) * std::max<int64_t>(total, 1)));
        hipLaunchKernelGGL(k_sg_distinct, dim3(grid), dim3(tpb), 0, ctx->stream, n, L->rp.as<int>(), L->ci.as<int>(),
                           B->rp.as<int>(), B->ci.as<int>(), off.as<int64_t>(), scratch.as<int>(), cnt.as<int>());
        BSP_HIP(hipGetLastError());
        BSP_HIP(hipMemsetAsync(cnt.as<int>() + n, 0, sizeof(int), ctx->stream));
        size_t tb2 = 0;
        BSP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, cnt.as<int>(), C->rp.as<int>(), n + 1, ctx->stream));
        if (tb2 > tmp.bytes) BSP_HIP(tmp.alloc(tb2));
        BSP_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb2, cnt.as<int>(), C->rp.as<int>(), n + 1, ctx->stream));
    } else {
        BSP_HIP(hipMemsetAsync(C->rp.p, 0, sizeof(int32_t), ctx->stream));
    }
    C->h_rp.resize(n + 1);
    BSP_HIP(hipMemcpyAsync(C->h_rp.data(), C->rp.p, sizeof(int32_t) * (n + 1), hipMemcpyDeviceToHost, ctx->stream));
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    const int64_t nnzc = C->h_rp[n];
    C->nnzb = nnzc;
    BSP_HIP(C->ci.alloc(sizeof(int32_t) * std::max<int64_t>(nnzc, 1)));
    BSP_HIP(C->val.alloc(sizeof(double) * std::max<int64_t>(nnzc * bs * bs, 1)));
    if (n > 0 && nnzc > 0) {
        hipLaunchKernelGGL(k_sg_copy, dim3(grid), dim3(tpb), 0, ctx->stream, n, off.as<int64_t>(), scratch.as<int>(),
                           C->rp.as<int>(), C->ci.as<int>());
        BSP_HIP(hipGetLastError());
    }
    // the 8 x 8 matrix-core numeric pass's schedule: each C block's product list, then the
    // lists spread over each row's streams
    if (n > 0 && nnzc > 0 && bs == 8) {
        DevMem ps, pairs;
        BSP_HIP(ps.alloc(sizeof(int64_t) * (nnzc + 1)));
        BSP_HIP(pairs.alloc(sizeof(int2) * std::max<int64_t>(total, 1)));
        hipLaunchKernelGGL(k_sg_pairs, dim3(grid), dim3(tpb), 0, ctx->stream, n, L->rp.as<int>(), L->ci.as<int>(),
                           B->rp.as<int>(), B->ci.as<int>(), off.as<int64_t>(), C->rp.as<int>(), C->ci.as<int>(),
                           ps.as<int64_t>(), pairs.as<int2>());
        BSP_HIP(hipGetLastError());
        const int grid1 = (n + 1 + tpb - 1) / tpb;
        hipLaunchKernelGGL(k_sg_sched_count, dim3(grid1), dim3(tpb), 0, ctx->stream, n, C->rp.as<int>(),
                           ps.as<int64_t>(), cand.as<int64_t>());
        BSP_HIP(hipGetLastError());
        BSP_HIP(C->mm_soff.alloc(sizeof(int64_t) * (n + 1)));
        size_t tb3 = 0;
        BSP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb3, cand.as<int64_t>(), C->mm_soff.as<int64_t>(), n + 1,
                                                 ctx->stream));
        if (tb3 > tmp.bytes) BSP_HIP(tmp.alloc(tb3));
        BSP_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb3, cand.as<int64_t>(), C->mm_soff.as<int64_t>(), n + 1,
                                                 ctx->stream));
        int64_t nsched = 0;
        BSP_HIP(hipMemcpyAsync(&nsched, C->mm_soff.as<int64_t>() + n, sizeof(int64_t), hipMemcpyDeviceToHost,
                               ctx->stream));
        BSP_HIP(hipStreamSynchronize(ctx->stream));
        BSP_HIP(C->mm_sched.alloc(sizeof(int4) * std::max<int64_t>(nsched, 1)));
        hipLaunchKernelGGL(k_sg_sched_fill, dim3(grid), dim3(tpb), 0, ctx->stream, n, C->rp.as<int>(),
                           ps.as<int64_t>(), pairs.as<int2>(), C->mm_soff.as<int64_t>(), C->mm_sched.as<int4>());
        BSP_HIP(hipGetLastError());
        BSP_HIP(hipStreamSynchronize(ctx->stream));
    }
    C->h_ci.resize(nnzc);
    if (nnzc)
        BSP_HIP(hipMemcpyAsync(C->h_ci.data(), C->ci.p, sizeof(int32_t) * nnzc, hipMemcpyDeviceToHost, ctx->stream));
    BSP_HIP(hipStreamSynchronize(ctx->stream));
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
    BSP_CHECK(A->cbs == B->cbs, BSP_E_INVALID, "bsmm: one factor is complex, the other real");
    BSP_HIP(hipSetDevice(ctx->device));
    int rc = bsmm_build(ctx, A, A, false, B, out);
    // real-equivalent blocks multiply as the complex ones do: C is complex with A and B
    if (rc == BSP_OK) (*out)->cbs = A->cbs;
    return rc;
}

int bsp_bsmm_trans_a(bsp_ctx* ctx, const bsp_bsr* A, const bsp_bsr* B, bsp_bsr** out) {
    if (!ctx || !A || !B || !out) return BSP_E_INVALID;
    *out = nullptr;
    BSP_CHECK(A->n_rows == B->n_rows, BSP_E_INVALID, "bsmm_trans_a: A and B have different row counts");
    BSP_CHECK(A->bs == B->bs, BSP_E_INVALID, "bsmm_trans_a: block sizes differ");
    BSP_CHECK(A->cbs == B->cbs, BSP_E_INVALID, "bsmm_trans_a: one factor is complex, the other real");
    BSP_HIP(hipSetDevice(ctx->device));
    int rc = ensure_transpose(ctx, A);
    if (rc) return rc;
    // a complex A's real-equivalent transpose is that of A^H: C = A^H B, the reference's
    // conjugate-transposed product for complex T (dense_multiply.hpp:155-175)
    rc = bsmm_build(ctx, A->transposed.get(), A, true, B, out);
    if (rc == BSP_OK) (*out)->cbs = A->cbs;
    return rc;
}

int bsp_bsmm_numeric(bsp_ctx* ctx, const bsp_bsr* A, const bsp_bsr* B, bsp_bsr* C) {
    if (!ctx || !A || !B || !C) return BSP_E_INVALID;
    BSP_CHECK(C->mm_A == A && C->mm_B == B, BSP_E_INVALID,
              "bsmm_numeric: C was not created by bsp_bsmm(A, B) or bsp_bsmm_trans_a(A, B)");
    BSP_HIP(hipSetDevice(ctx->device));
    const bsp_bsr* L = A;
    if (C->mm_Atran) {
        // A's transpose is rebuilt from its current values when they changed
        int rc = ensure_transpose(ctx, A);
        if (rc) return rc;
        L = A->transposed.get();
    }
    if (C->n_rows > 0 && C->nnzb > 0) {
        BSP_CHECK(launch_rows_bs(ctx->stream, C->bs, C->n_rows, L, B, C, ctx->bsmm_mfma != 0), BSP_E_UNSUPPORTED,
                  "bsmm: block size " + std::to_string(C->bs) + " is not supported (1-8, 16, 32)");
        BSP_HIP(hipGetLastError());
    }
    C->h_val_stale = true;
    C->val32.release();
    C->val16.release();
    C->valp.release();
    C->valpe.release();
    C->transposed.reset();
    return BSP_OK;
}

}  // extern "C"
