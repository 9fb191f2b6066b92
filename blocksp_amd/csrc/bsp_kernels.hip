// bsp_kernels.hip -- CDNA4 (gfx950) kernels of the block-sparse solve path.
//
// Mapping used throughout: a block row of a BSR matrix with block size BS is
// handled by BS consecutive lanes of a 64-wide wavefront (lane r owns row r of
// every block of that block row), so one wavefront covers 64/BS block rows.
// Lane r reads the BS contiguous doubles of its block row with 16-byte loads;
// the BS lanes of a group therefore stream the whole BS*BS block contiguously.
// Within a block row, lane r accumulates sum_blocks (sum_c a_rc x_c) in the
// reference's order: blocks in stored order, columns c = 0..BS-1
// (blocksp_multiply.hpp:289-325, blocksp_dense_multiply.hpp:29-48), with FMA.
#include "bsp_internal.hpp"

#include <utility>

#include <algorithm>

namespace bsp {

namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;

template <int BS>
__device__ __forceinline__ void load_vec(const double* __restrict__ src, double (&a)[BS]) {
    if constexpr (BS % 2 == 0) {
        const double2* s2 = reinterpret_cast<const double2*>(src);
#pragma unroll
        for (int c = 0; c < BS / 2; ++c) {
            double2 t = s2[c];
            a[2 * c] = t.x;
            a[2 * c + 1] = t.y;
        }
    } else {
#pragma unroll
        for (int c = 0; c < BS; ++c) a[c] = src[c];
    }
}

template <int BS>
__device__ __forceinline__ double row_dot(const double (&a)[BS], const double (&x)[BS]) {
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < BS; ++c) s = fma(a[c], x[c], s);
    return s;
}

// Apply a BS x BS block (row r of it held by lane r of the group) to the
// group's vector s (element c held by lane base+c): returns sum_c D[r][c] s_c.
template <int BS>
__device__ __forceinline__ double group_matvec(const double* __restrict__ D, int r, int base, double s) {
    double d[BS];
    load_vec<BS>(D + r * BS, d);
    double out = 0.0;
#pragma unroll
    for (int c = 0; c < BS; ++c) out = fma(d[c], __shfl(s, base + c, kWave), out);
    return out;
}

// ---------------------------------------------------------------------------
// bsmv
// ---------------------------------------------------------------------------
template <int BS>
__global__ __launch_bounds__(kBlock) void k_bsmv(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                 const double* __restrict__ val, const double* __restrict__ x,
                                                 double* __restrict__ y, int accumulate) {
    constexpr int RPW = kWave / BS;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
    const int g = lane / BS;
    const int r = lane - g * BS;
    if (g >= RPW) return;
    const int i = wave * RPW + g;
    if (i >= n) return;
    double acc = accumulate ? y[(size_t)i * BS + r] : 0.0;
    const int p0 = rp[i], p1 = rp[i + 1];
    int p = p0;
    for (; p + 1 < p1; p += 2) {
        const int c0 = ci[p], c1 = ci[p + 1];
        double a0[BS], a1[BS], x0[BS], x1[BS];
        load_vec<BS>(val + ((size_t)p * BS + r) * BS, a0);
        load_vec<BS>(val + ((size_t)(p + 1) * BS + r) * BS, a1);
        load_vec<BS>(x + (size_t)c0 * BS, x0);
        load_vec<BS>(x + (size_t)c1 * BS, x1);
        acc += row_dot<BS>(a0, x0);
        acc += row_dot<BS>(a1, x1);
    }
    if (p < p1) {
        const int c0 = ci[p];
        double a0[BS], x0[BS];
        load_vec<BS>(val + ((size_t)p * BS + r) * BS, a0);
        load_vec<BS>(x + (size_t)c0 * BS, x0);
        acc += row_dot<BS>(a0, x0);
    }
    y[(size_t)i * BS + r] = acc;
}

typedef double d2v_p __attribute__((ext_vector_type(2)));

// Same mapping and arithmetic order, reading a copy of the values whose blocks are
// stored chunk-major: entry (r, c) of a BS x BS block sits at ((c/2) * BS + r) * 2 + c%2,
// so the k-th 16-byte load of the BS lanes of a block row covers one contiguous
// BS*16-byte run (in the row-major layout the same load touches BS separate rows).
// Workgroups are remapped so each of the 8 XCDs walks one contiguous range of block rows
// (neighbouring x blocks stay in that XCD's L2), and the values -- read once -- are streamed
// with non-temporal loads so that they do not evict x.
template <int BS>
__global__ __launch_bounds__(kBlock) void k_bsmv_perm(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                      const double* __restrict__ valp, const double* __restrict__ x,
                                                      double* __restrict__ y, int accumulate, int nwg) {
    static_assert(BS % 2 == 0, "chunk-major layout needs an even block size");
    constexpr int RPW = kWave / BS;
    const int lane = threadIdx.x & (kWave - 1);
    const int per = (nwg + 7) / 8;
    const int wg = (blockIdx.x % 8) * per + blockIdx.x / 8;
    if (wg >= nwg) return;
    const int wave = (wg * kBlock + threadIdx.x) / kWave;
    const int g = lane / BS;
    const int r = lane - g * BS;
    if (g >= RPW) return;
    const int i = wave * RPW + g;
    if (i >= n) return;
    double acc = accumulate ? y[(size_t)i * BS + r] : 0.0;
    const int p0 = rp[i], p1 = rp[i + 1];
    auto load_row = [&](int p, double (&a)[BS]) {
        const double2* s2 = reinterpret_cast<const double2*>(valp + (size_t)p * BS * BS) + r;
#pragma unroll
        for (int k = 0; k < BS / 2; ++k) {
            const d2v_p u = __builtin_nontemporal_load(reinterpret_cast<const d2v_p*>(s2 + k * BS));
            a[2 * k] = u.x;
            a[2 * k + 1] = u.y;
        }
    };
    int p = p0;
    for (; p + 1 < p1; p += 2) {
        const int c0 = ci[p], c1 = ci[p + 1];
        double a0[BS], a1[BS], x0[BS], x1[BS];
        load_row(p, a0);
        load_row(p + 1, a1);
        load_vec<BS>(x + (size_t)c0 * BS, x0);
        load_vec<BS>(x + (size_t)c1 * BS, x1);
        acc += row_dot<BS>(a0, x0);
        acc += row_dot<BS>(a1, x1);
    }
    if (p < p1) {
        const int c0 = ci[p];
        double a0[BS], x0[BS];
        load_row(p, a0);
        load_vec<BS>(x + (size_t)c0 * BS, x0);
        acc += row_dot<BS>(a0, x0);
    }
    y[(size_t)i * BS + r] = acc;
}

// The same product over an ELL copy of the chunk-major values (W <= 8 slots per row, column -1 and
// zero blocks on padding): a row's column indices load from its row number, all at once, so the
// blocks and the x entries wait on one load instead of the row pointers and then the columns.
// Real blocks are added in the row's stored order and padding is skipped: the sums are bitwise
// those of k_bsmv_perm.
template <int BS>
__global__ __launch_bounds__(kBlock) void k_bsmv_ell(int n, int W, const int* __restrict__ cie,
                                                     const double* __restrict__ vale, const double* __restrict__ x,
                                                     double* __restrict__ y, int accumulate, int nwg) {
    static_assert(BS % 2 == 0, "chunk-major layout needs an even block size");
    constexpr int RPW = kWave / BS;
    const int lane = threadIdx.x & (kWave - 1);
    const int per = (nwg + 7) / 8;
    const int wg = (blockIdx.x % 8) * per + blockIdx.x / 8;
    if (wg >= nwg) return;
    const int wave = (wg * kBlock + threadIdx.x) / kWave;
    const int g = lane / BS;
    const int r = lane - g * BS;
    if (g >= RPW) return;
    const int i = wave * RPW + g;
    if (i >= n) return;
    double acc = accumulate ? y[(size_t)i * BS + r] : 0.0;
    int col[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) col[s] = s < W ? cie[(size_t)i * W + s] : -1;
    auto load_row = [&](int q, double (&a)[BS]) {
        const double2* s2 = reinterpret_cast<const double2*>(vale + (size_t)q * BS * BS) + r;
#pragma unroll
        for (int k = 0; k < BS / 2; ++k) {
            const d2v_p u = __builtin_nontemporal_load(reinterpret_cast<const d2v_p*>(s2 + k * BS));
            a[2 * k] = u.x;
            a[2 * k + 1] = u.y;
        }
    };
    const size_t q0 = (size_t)i * W;
#pragma unroll
    for (int s = 0; s < 8; s += 2) {
        if (col[s] < 0) break;   // padding is at the end of a row
        double a0[BS], x0[BS];
        load_row(q0 + s, a0);
        load_vec<BS>(x + (size_t)col[s] * BS, x0);
        if (col[s + 1] >= 0) {
            double a1[BS], x1[BS];
            load_row(q0 + s + 1, a1);
            load_vec<BS>(x + (size_t)col[s + 1] * BS, x1);
            acc += row_dot<BS>(a0, x0);
            acc += row_dot<BS>(a1, x1);
        } else {
            acc += row_dot<BS>(a0, x0);
            break;
        }
    }
    y[(size_t)i * BS + r] = acc;
}

// Column indices of the ELL structure: slot (row, s) holds the row's s-th block column, -1 past
// the row's end.
__global__ void k_ell_columns(int64_t slots, int W, const int* __restrict__ rp, const int* __restrict__ ci,
                              int* __restrict__ cie) {
    const int64_t qs = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (qs >= slots) return;
    const int64_t row = qs / W, s = qs - row * W;
    const int p0 = rp[row], len = rp[row + 1] - p0;
    cie[qs] = s < len ? ci[p0 + s] : -1;
}

// ELL copy of the values in the chunk-major order of k_bsmv_perm (FP64, chunks of 2) or of
// k_bsmv_lp_perm (FP32, chunks of 4): thread t writes entry t of the copy; zero on padding.
template <typename T, int CH>
__global__ void k_ell_values(int64_t total, int bs, int W, const int* __restrict__ rp, const double* __restrict__ val,
                             T* __restrict__ vale) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int b2 = bs * bs;
    const int64_t qs = t / b2;
    const int e = (int)(t - qs * b2);
    const int64_t row = qs / W, s = qs - row * W;
    const int p0 = rp[row], len = rp[row + 1] - p0;
    const int chunk = e / (CH * bs), rem = e - chunk * CH * bs;
    const int r = rem / CH, c = chunk * CH + rem % CH;
    vale[t] = s < len ? (T)val[(size_t)(p0 + s) * b2 + r * bs + c] : (T)0;
}

__global__ void k_permute_blocks(int64_t total, int bs, const double* __restrict__ val, double* __restrict__ valp) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int b2 = bs * bs;
    const int64_t p = t / b2;
    const int e = (int)(t - p * b2);
    const int r = e / bs, c = e - r * bs;
    valp[p * b2 + ((c / 2) * bs + r) * 2 + (c & 1)] = val[t];
}

// Runtime block size (bs <= 64): same mapping, scalar loads.
__global__ __launch_bounds__(kBlock) void k_bsmv_generic(int n, int bs, const int* __restrict__ rp,
                                                         const int* __restrict__ ci,
                                                         const double* __restrict__ val,
                                                         const double* __restrict__ x, double* __restrict__ y,
                                                         int accumulate) {
    const int rpw = kWave / bs;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
    const int g = lane / bs;
    const int r = lane - g * bs;
    if (g >= rpw) return;
    const int i = wave * rpw + g;
    if (i >= n) return;
    double acc = accumulate ? y[(size_t)i * bs + r] : 0.0;
    for (int p = rp[i]; p < rp[i + 1]; ++p) {
        const double* a = val + ((size_t)p * bs + r) * bs;
        const double* xc = x + (size_t)ci[p] * bs;
        double s = 0.0;
        for (int c = 0; c < bs; ++c) s = fma(a[c], xc[c], s);
        acc += s;
    }
    y[(size_t)i * bs + r] = acc;
}

// Per-dimension product: A has blocks B x B; x, y have n*B*D entries; each of the
// D dimension slices of a block row is a separate lane group.
__global__ __launch_bounds__(kBlock) void k_bsmv_dim(int n, int B, int D, int dim, const int* __restrict__ rp,
                                                     const int* __restrict__ ci, const double* __restrict__ val,
                                                     const double* __restrict__ x, double* __restrict__ y,
                                                     int accumulate) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t total = (int64_t)n * D * B;
    if (t >= total) return;
    const int r = (int)(t % B);
    const int d = (int)((t / B) % D);
    const int i = (int)(t / ((int64_t)B * D));
    if (dim >= 0 && d != dim) return;
    const int stride = B * D;
    double acc = accumulate ? y[(size_t)i * stride + d * B + r] : 0.0;
    for (int p = rp[i]; p < rp[i + 1]; ++p) {
        const double* a = val + ((size_t)p * B + r) * B;
        const double* xc = x + (size_t)ci[p] * stride + d * B;
        double s = 0.0;
        for (int c = 0; c < B; ++c) s = fma(a[c], xc[c], s);
        acc += s;
    }
    y[(size_t)i * stride + d * B + r] = acc;
}

// Multi-RHS: Y[(i*bs + r)*nrhs + k] = sum_blocks sum_c a_rc X[(col*bs + c)*nrhs + k].
// One thread per (row entry, rhs); for the block sizes k_bsmm_rhs_tile has no instance of.
__global__ __launch_bounds__(kBlock) void k_bsmm_rhs(int n, int bs, int nrhs, const int* __restrict__ rp,
                                                     const int* __restrict__ ci, const double* __restrict__ val,
                                                     const double* __restrict__ X, double* __restrict__ Y) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t total = (int64_t)n * bs * nrhs;
    if (t >= total) return;
    const int k = (int)(t % nrhs);
    const int64_t er = t / nrhs;
    const int r = (int)(er % bs);
    const int i = (int)(er / bs);
    double acc = 0.0;
    for (int p = rp[i]; p < rp[i + 1]; ++p) {
        const double* a = val + ((size_t)p * bs + r) * bs;
        const double* xc = X + (size_t)ci[p] * bs * nrhs + k;
        double s = 0.0;
        for (int c = 0; c < bs; ++c) s = fma(a[c], xc[(size_t)c * nrhs], s);
        acc += s;
    }
    Y[t] = acc;
}

// Multi-RHS product on the RHS tile mapping: one wavefront per (block row i, tile of KT
// right-hand sides), RG = 64 / KT lane groups.  Batches of NB blocks of the row are staged:
// their values go through LDS with one coalesced streaming load per block (read once, so they
// bypass the caches), their X rows load straight into registers (coalesced over k; X is
// cached -- neighbouring rows read the same blocks); the row's column indices come in one
// vector load, 64 at a time.  Two lane-group layouts:
//  * k_bsmm_rhs_splitc (BS a multiple of RG): lane (g, k) sums the columns c of group g's
//    share for every row r of the block -- each X entry is loaded by one lane only -- and the
//    RG partial sums of an output are added across the groups at the end;
//  * k_bsmm_rhs_tile (other BS): lane (g, k) owns rows g R .. g R + R - 1 and all columns;
//    each output adds s = sum_c a_rc x_c (FMA in c order) per block, bsmv's order.
template <int BS>
struct RhsShape {
    static constexpr int NB = BS <= 8 ? 2 : 1;
};

template <int BS, int KT, int NB>
__global__ __launch_bounds__(kBlock) void k_bsmm_rhs_tile(int64_t nwaves, int nrhs, int ntiles,
                                                          const int* __restrict__ rp, const int* __restrict__ ci,
                                                          const double* __restrict__ val,
                                                          const double* __restrict__ X, double* __restrict__ Y) {
    constexpr int RG = kWave / KT, R = (BS + RG - 1) / RG, B2 = BS * BS;
    constexpr int WPB = kBlock / kWave;
    __shared__ double As[WPB][NB][B2];
    const int wv = threadIdx.x / kWave, l = threadIdx.x % kWave;
    // XCD-aware order: XCD x (workgroups x, x + 8, ...) takes one contiguous range of rows
    const int64_t nwg = (nwaves + WPB - 1) / WPB, per = (nwg + 7) / 8;
    const int64_t wg = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    const int64_t gw = wg * WPB + wv;
    if (wg >= nwg || gw >= nwaves) return;   // whole wavefronts leave; no workgroup barrier follows
    const int i = (int)(gw / ntiles), tile = (int)(gw % ntiles);
    const int kk = l % KT, g = l / KT;
    const int k = tile * KT + kk;
    const bool kv = k < nrhs;
    double acc[R];
#pragma unroll
    for (int rr = 0; rr < R; ++rr) acc[rr] = 0.0;
    const int p0 = rp[i], p1 = rp[i + 1];
    // the row's column indices, 64 at a time, one per lane
    int cl = 0;
    for (int pb = p0; pb < p1; pb += NB) {
        const int nb = min(NB, p1 - pb);
        if (((pb - p0) & (kWave - 1)) == 0) cl = pb + l < p1 ? ci[pb + l] : 0;
#pragma unroll
        for (int j = 0; j < NB; ++j)
            if (j < nb)
                for (int e = l; e < B2; e += kWave)
                    As[wv][j][e] = __builtin_nontemporal_load(val + (size_t)(pb + j) * B2 + e);
        double xv[NB][BS];
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const int col = __shfl(cl, (pb - p0 + (j < nb ? j : 0)) & (kWave - 1), kWave);
            const double* xc = X + (size_t)col * BS * nrhs + k;
#pragma unroll
            for (int c = 0; c < BS; ++c) xv[j][c] = (j < nb && kv) ? xc[(size_t)c * nrhs] : 0.0;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            if (j >= nb) break;
#pragma unroll
            for (int rr = 0; rr < R; ++rr) {
                const int r = g * R + rr;
                if (R * RG == BS || r < BS) {
                    double sum = 0.0;
#pragma unroll
                    for (int c = 0; c < BS; ++c) sum = fma(As[wv][j][r * BS + c], xv[j][c], sum);
                    acc[rr] += sum;
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (kv) {
#pragma unroll
        for (int rr = 0; rr < R; ++rr) {
            const int r = g * R + rr;
            if (R * RG == BS || r < BS) __builtin_nontemporal_store(acc[rr], Y + ((size_t)i * BS + r) * nrhs + k);
        }
    }
}

template <int BS, int KT, int NB>
__global__ __launch_bounds__(kBlock) void k_bsmm_rhs_splitc(int64_t nwaves, int nrhs, int ntiles,
                                                            const int* __restrict__ rp, const int* __restrict__ ci,
                                                            const double* __restrict__ val,
                                                            const double* __restrict__ X, double* __restrict__ Y) {
    constexpr int RG = kWave / KT, CG = BS / RG, B2 = BS * BS;
    constexpr int WPB = kBlock / kWave;
    __shared__ double As[WPB][NB][B2];
    const int wv = threadIdx.x / kWave, l = threadIdx.x % kWave;
    const int64_t nwg = (nwaves + WPB - 1) / WPB, per = (nwg + 7) / 8;
    const int64_t wg = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
    const int64_t gw = wg * WPB + wv;
    if (wg >= nwg || gw >= nwaves) return;   // whole wavefronts leave; no workgroup barrier follows
    const int i = (int)(gw / ntiles), tile = (int)(gw % ntiles);
    const int kk = l % KT, g = l / KT;
    const int k = tile * KT + kk;
    const bool kv = k < nrhs;
    double acc[BS];
#pragma unroll
    for (int r = 0; r < BS; ++r) acc[r] = 0.0;
    const int p0 = rp[i], p1 = rp[i + 1];
    int cl = 0;
    for (int pb = p0; pb < p1; pb += NB) {
        const int nb = min(NB, p1 - pb);
        if (((pb - p0) & (kWave - 1)) == 0) cl = pb + l < p1 ? ci[pb + l] : 0;
#pragma unroll
        for (int j = 0; j < NB; ++j)
            if (j < nb)
                for (int e = l; e < B2; e += kWave)
                    As[wv][j][e] = __builtin_nontemporal_load(val + (size_t)(pb + j) * B2 + e);
        double xv[NB][CG];
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const int col = __shfl(cl, (pb - p0 + (j < nb ? j : 0)) & (kWave - 1), kWave);
            const double* xc = X + ((size_t)col * BS + g * CG) * nrhs + k;
#pragma unroll
            for (int c = 0; c < CG; ++c) xv[j][c] = (j < nb && kv) ? xc[(size_t)c * nrhs] : 0.0;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            if (j >= nb) break;
#pragma unroll
            for (int r = 0; r < BS; ++r) {
                double sum = 0.0;
#pragma unroll
                for (int c = 0; c < CG; ++c) sum = fma(As[wv][j][r * BS + g * CG + c], xv[j][c], sum);
                acc[r] += sum;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
#pragma unroll
    for (int r = 0; r < BS; ++r) {
        double v = acc[r];
#pragma unroll
        for (int d = KT; d < kWave; d <<= 1) v += __shfl_xor(v, d, kWave);
        acc[r] = v;
    }
    if (kv && g == 0) {
#pragma unroll
        for (int r = 0; r < BS; ++r) __builtin_nontemporal_store(acc[r], Y + ((size_t)i * BS + r) * nrhs + k);
    }
}

// RHS tile width: the smallest of 8, 16, 32, 64 lanes that covers nrhs (64 for more, in
// tiles), and at least 64 / BS so that every lane has a row
template <int BS>
struct BsmmRhsL {
    using Kern = void (*)(int64_t, int, int, const int*, const int*, const double*, const double*, double*);
    template <int KT, int NB>
    static Kern pick() {
        if constexpr (BS % (kWave / KT) == 0)
            return k_bsmm_rhs_splitc<BS, KT, NB>;
        else
            return k_bsmm_rhs_tile<BS, KT, NB>;
    }
    static void run(hipStream_t st, int n, int nrhs, const int* rp, const int* ci, const double* val,
                    const double* X, double* Y) {
        constexpr int kmin = BS >= 8 ? 8 : (BS >= 4 ? 16 : (BS >= 2 ? 32 : 64));
        int kt = kmin;
        while (kt < 64 && kt < nrhs) kt *= 2;
        const int ntiles = (nrhs + kt - 1) / kt;
        const int64_t nwaves = (int64_t)n * ntiles;
        const int64_t nwg = (nwaves + kBlock / kWave - 1) / (kBlock / kWave);
        const dim3 grid((unsigned)((nwg + 7) / 8 * 8));
        auto go = [&](auto kern) {
            hipLaunchKernelGGL(kern, grid, dim3(kBlock), 0, st, nwaves, nrhs, ntiles, rp, ci, val, X, Y);
        };
        constexpr int NB = RhsShape<BS>::NB;
        switch (kt) {
            case 8:
                if constexpr (kmin <= 8) go(pick<8, NB>());
                break;
            case 16:
                if constexpr (kmin <= 16) go(pick<16, NB>());
                break;
            case 32:
                if constexpr (kmin <= 32) go(pick<32, NB>());
                break;
            default: go(pick<64, NB>()); break;
        }
    }
};

// x_i <- D_i x_i for every block row (Diag preconditioner).
template <int BS>
__global__ __launch_bounds__(kBlock) void k_block_diag(int n, const double* __restrict__ dinv, double* __restrict__ x) {
    constexpr int RPW = kWave / BS;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
    const int g = lane / BS;
    const int r = lane - g * BS;
    const int i = wave * RPW + g;
    const bool active = g < RPW && i < n;
    double s = active ? x[(size_t)i * BS + r] : 0.0;
    double out = 0.0;
    if (active) out = 0.0;
    // all lanes take part in the shuffles
    const int base = g * BS;
    {
        double d[BS];
        if (active) load_vec<BS>(dinv + ((size_t)i * BS + r) * BS, d);
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            double sc = __shfl(s, (base + c) & (kWave - 1), kWave);
            if (active) out = fma(d[c], sc, out);
        }
    }
    if (active) x[(size_t)i * BS + r] = out;
}

__global__ __launch_bounds__(kBlock) void k_block_diag_generic(int n, int bs, const double* __restrict__ dinv,
                                                               const double* __restrict__ xin,
                                                               double* __restrict__ xout) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= (int64_t)n * bs) return;
    const int r = (int)(t % bs);
    const int64_t i = t / bs;
    const double* d = dinv + ((size_t)i * bs + r) * bs;
    const double* s = xin + (size_t)i * bs;
    double out = 0.0;
    for (int c = 0; c < bs; ++c) out = fma(d[c], s[c], out);
    xout[t] = out;
}

// ---------------------------------------------------------------------------
// Reductions (deterministic: fixed grid, fixed partial order, last block sums)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}

template <int KIND>
__device__ __forceinline__ double block_reduce(double v, double* sh) {
    v = KIND == 2 ? wave_max(v) : wave_sum(v);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        r = sh[0];
        for (int i = 1; i < kBlock / kWave; ++i) r = KIND == 2 ? fmax(r, sh[i]) : r + sh[i];
    }
    __syncthreads();
    return r;
}

// Finish a grid reduction: every block stores its partial; the last block to
// arrive (counted with a device-scope atomic) combines the partials in block
// order and writes *out.
template <int KIND>
__device__ __forceinline__ void grid_finish(double part, double* partials, unsigned* counter, double* out,
                                            double* sh) {
    __shared__ bool last;
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = part;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        unsigned prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    double v = KIND == 2 ? 0.0 : 0.0;
    // fixed order: thread t sums partials t, t+256, ...; then block tree
    for (int b = threadIdx.x; b < (int)gridDim.x; b += kBlock) {
        double pv = __hip_atomic_load(&partials[b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        v = KIND == 2 ? fmax(v, pv) : v + pv;
    }
    v = block_reduce<KIND>(v, sh);
    if (threadIdx.x == 0) {
        if (KIND == 1) v = sqrt(fabs(v));
        *out = v;
        *counter = 0u;
    }
}

// Streaming reductions: a block handles tiles of kTile consecutive elements, each
// thread two 16-byte loads per operand issued before use (vector path when every
// operand is 16-byte aligned; scalar path otherwise and for the tail tile).
constexpr int kTile = kBlock * 4;

template <int KIND>
__device__ __forceinline__ double red_op(double s, double a, double b) {
    if constexpr (KIND == 0) return fma(a, b, s);
    else if constexpr (KIND == 1) return fma(a, a, s);
    else return fmax(s, fabs(a));
}

template <int KIND, bool VEC>
__global__ __launch_bounds__(kBlock) void k_reduce(int64_t len, const double* __restrict__ x,
                                                   const double* __restrict__ y, double* partials,
                                                   unsigned* counter, double* out) {
    __shared__ double sh[kBlock / kWave];
    double s0 = 0.0, s1 = 0.0;
    for (int64_t base = (int64_t)blockIdx.x * kTile; base < len; base += (int64_t)gridDim.x * kTile) {
        if (VEC && base + kTile <= len) {
            const double2* x2 = reinterpret_cast<const double2*>(x + base);
            const double2 a0 = x2[threadIdx.x], a1 = x2[threadIdx.x + kBlock];
            double2 b0 = a0, b1 = a1;
            if constexpr (KIND == 0) {
                const double2* y2 = reinterpret_cast<const double2*>(y + base);
                b0 = y2[threadIdx.x];
                b1 = y2[threadIdx.x + kBlock];
            }
            s0 = red_op<KIND>(s0, a0.x, b0.x);
            s1 = red_op<KIND>(s1, a0.y, b0.y);
            s0 = red_op<KIND>(s0, a1.x, b1.x);
            s1 = red_op<KIND>(s1, a1.y, b1.y);
        } else {
            for (int64_t i = base + threadIdx.x; i < len && i < base + kTile; i += kBlock)
                s0 = red_op<KIND>(s0, x[i], KIND == 0 ? y[i] : x[i]);
        }
    }
    double s = KIND == 2 ? fmax(s0, s1) : s0 + s1;
    s = block_reduce<KIND>(s, sh);
    grid_finish<KIND>(s, partials, counter, out, sh);
}

// v -= h_prev * qprev (if qprev), then out = dot(qnext, v) or ||v||.
template <bool VEC>
__global__ __launch_bounds__(kBlock) void k_mgs_step(int64_t len, double* __restrict__ v,
                                                     const double* __restrict__ qprev,
                                                     const double* __restrict__ h_prev,
                                                     const double* __restrict__ qnext, double* partials,
                                                     unsigned* counter, double* out) {
    __shared__ double sh[kBlock / kWave];
    const double mh = qprev ? -(*h_prev) : 0.0;
    double s0 = 0.0, s1 = 0.0;
    for (int64_t base = (int64_t)blockIdx.x * kTile; base < len; base += (int64_t)gridDim.x * kTile) {
        if (VEC && base + kTile <= len) {
            double2* v2 = reinterpret_cast<double2*>(v + base);
            double2 a0 = v2[threadIdx.x], a1 = v2[threadIdx.x + kBlock];
            double2 p0{0.0, 0.0}, p1{0.0, 0.0}, q0{0.0, 0.0}, q1{0.0, 0.0};
            if (qprev) {
                const double2* p2 = reinterpret_cast<const double2*>(qprev + base);
                p0 = p2[threadIdx.x];
                p1 = p2[threadIdx.x + kBlock];
            }
            if (qnext) {
                const double2* q2 = reinterpret_cast<const double2*>(qnext + base);
                q0 = q2[threadIdx.x];
                q1 = q2[threadIdx.x + kBlock];
            }
            if (qprev) {
                a0.x = a0.x + mh * p0.x;
                a0.y = a0.y + mh * p0.y;
                a1.x = a1.x + mh * p1.x;
                a1.y = a1.y + mh * p1.y;
                v2[threadIdx.x] = a0;
                v2[threadIdx.x + kBlock] = a1;
            }
            if (qnext) {
                s0 = fma(q0.x, a0.x, s0);
                s1 = fma(q0.y, a0.y, s1);
                s0 = fma(q1.x, a1.x, s0);
                s1 = fma(q1.y, a1.y, s1);
            } else {
                s0 = fma(a0.x, a0.x, s0);
                s1 = fma(a0.y, a0.y, s1);
                s0 = fma(a1.x, a1.x, s0);
                s1 = fma(a1.y, a1.y, s1);
            }
        } else {
            for (int64_t i = base + threadIdx.x; i < len && i < base + kTile; i += kBlock) {
                double vi = v[i];
                if (qprev) {
                    vi = vi + mh * qprev[i];
                    v[i] = vi;
                }
                s0 = qnext ? fma(qnext[i], vi, s0) : fma(vi, vi, s0);
            }
        }
    }
    double s = block_reduce<0>(s0 + s1, sh);
    if (qnext) grid_finish<0>(s, partials, counter, out, sh);
    else grid_finish<1>(s, partials, counter, out, sh);
}

inline bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

__global__ __launch_bounds__(kBlock) void k_axpy(int64_t len, double a, const double* __restrict__ x,
                                                 double* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < len) y[i] += a * x[i];
}

__global__ __launch_bounds__(kBlock) void k_scal(int64_t len, double a, double* __restrict__ x) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < len) x[i] *= a;
}

template <typename QT>
__global__ __launch_bounds__(kBlock) void k_scale_div(int64_t len, const double* __restrict__ v,
                                                      const double* __restrict__ h, QT* __restrict__ q) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < len) q[i] = (QT)(v[i] / *h);
}

__global__ __launch_bounds__(kBlock) void k_copy(int64_t len, const double* __restrict__ x, double* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < len) y[i] = x[i];
}

__global__ __launch_bounds__(kBlock) void k_fill_u32(int64_t len, uint32_t v, uint32_t* __restrict__ p) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < len) p[i] = v;
}

constexpr int kMaxBasis = 64;

struct PtrTable {
    const double* q[kMaxBasis];
};
struct OutTable {
    double* q[kMaxBasis];
};

// x += sum_i c_i Q_i, summed in index order (solvers.hpp:290-294).
__global__ __launch_bounds__(kBlock) void k_combine_add(int64_t len, int m, PtrTable Q, const double* __restrict__ c,
                                                        double* __restrict__ x) {
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= len) return;
    double acc = x[e];
    for (int i = 0; i < m; ++i) acc += c[i] * Q.q[i][e];
    x[e] = acc;
}

// Qout_j = sum_i Q_i P_ij (deflated_restart.hpp:250-258), m <= 64, kp <= 64.  Two consecutive
// entries per thread (16-byte loads and stores), the basis read once per 16 output vectors,
// four input vectors' loads issued before their FMAs.
// Qout_j = sum_i Q_i P_ij for kp = JN <= 16 output vectors, two entries per thread, eight input
// vectors' loads issued before their FMAs, P read with uniform (scalar) loads: the same sums in
// the same order as k_basis_rotate, without its LDS staging and per-output masks.  len is even
// and every vector 16-byte aligned.
template <int JN>
__global__ __launch_bounds__(kBlock) void k_basis_rotate_jn(int64_t len, int m, PtrTable Q,
                                                            const double* __restrict__ P, OutTable Out) {
    const int64_t e = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * 2;
    if (e >= len) return;
    double a0[JN], a1[JN];
#pragma unroll
    for (int j = 0; j < JN; ++j) a0[j] = a1[j] = 0.0;
    int i = 0;
    for (; i + 8 <= m; i += 8) {
        double2 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = *reinterpret_cast<const double2*>(Q.q[i + u] + e);
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int j = 0; j < JN; ++j) {
                const double pv = P[(i + u) * JN + j];
                a0[j] = fma(v[u].x, pv, a0[j]);
                a1[j] = fma(v[u].y, pv, a1[j]);
            }
    }
    for (; i < m; ++i) {
        const double2 v = *reinterpret_cast<const double2*>(Q.q[i] + e);
#pragma unroll
        for (int j = 0; j < JN; ++j) {
            const double pv = P[i * JN + j];
            a0[j] = fma(v.x, pv, a0[j]);
            a1[j] = fma(v.y, pv, a1[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < JN; ++j) *reinterpret_cast<double2*>(Out.q[j] + e) = make_double2(a0[j], a1[j]);
}

__global__ __launch_bounds__(kBlock) void k_basis_rotate(int64_t len, int m, int kp, PtrTable Q,
                                                         const double* __restrict__ P, OutTable Out, int vec) {
    __shared__ double sP[kMaxBasis * 16];
    const int64_t e = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * 2;
    for (int j0 = 0; j0 < kp; j0 += 16) {
        const int jn = min(16, kp - j0);
        __syncthreads();
        for (int t = threadIdx.x; t < m * jn; t += kBlock) sP[t] = P[(t / jn) * kp + j0 + (t % jn)];
        __syncthreads();
        if (vec && e + 1 < len) {
            double a0[16], a1[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) a0[j] = a1[j] = 0.0;
            int i = 0;
            for (; i + 4 <= m; i += 4) {
                double2 v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) v[u] = *reinterpret_cast<const double2*>(Q.q[i + u] + e);
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int j = 0; j < 16; ++j)
                        if (j < jn) {
                            a0[j] = fma(v[u].x, sP[(i + u) * jn + j], a0[j]);
                            a1[j] = fma(v[u].y, sP[(i + u) * jn + j], a1[j]);
                        }
            }
            for (; i < m; ++i) {
                const double2 v = *reinterpret_cast<const double2*>(Q.q[i] + e);
#pragma unroll
                for (int j = 0; j < 16; ++j)
                    if (j < jn) {
                        a0[j] = fma(v.x, sP[i * jn + j], a0[j]);
                        a1[j] = fma(v.y, sP[i * jn + j], a1[j]);
                    }
            }
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (j < jn) *reinterpret_cast<double2*>(Out.q[j0 + j] + e) = make_double2(a0[j], a1[j]);
        } else {
            for (int64_t f = e; f < e + 2 && f < len; ++f) {
                double acc[16];
#pragma unroll
                for (int j = 0; j < 16; ++j) acc[j] = 0.0;
                for (int i = 0; i < m; ++i) {
                    const double qv = Q.q[i][f];
#pragma unroll
                    for (int j = 0; j < 16; ++j)
                        if (j < jn) acc[j] = fma(qv, sP[i * jn + j], acc[j]);
                }
#pragma unroll
                for (int j = 0; j < 16; ++j)
                    if (j < jn) Out.q[j0 + j][f] = acc[j];
            }
        }
    }
}

// r = b - A x
template <int BS>
__global__ __launch_bounds__(kBlock) void k_residual(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                     const double* __restrict__ val, const double* __restrict__ b,
                                                     const double* __restrict__ x, double* __restrict__ r_out) {
    constexpr int RPW = kWave / BS;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
    const int g = lane / BS;
    const int r = lane - g * BS;
    if (g >= RPW) return;
    const int i = wave * RPW + g;
    if (i >= n) return;
    double acc = 0.0;
    for (int p = rp[i]; p < rp[i + 1]; ++p) {
        double a[BS], xv[BS];
        load_vec<BS>(val + ((size_t)p * BS + r) * BS, a);
        load_vec<BS>(x + (size_t)ci[p] * BS, xv);
        acc += row_dot<BS>(a, xv);
    }
    r_out[(size_t)i * BS + r] = b[(size_t)i * BS + r] - acc;
}

__global__ __launch_bounds__(kBlock) void k_cg_update_xr(int64_t len, const double* __restrict__ num,
                                                         const double* __restrict__ den, const double* __restrict__ p,
                                                         const double* __restrict__ Ap, double* __restrict__ x,
                                                         double* __restrict__ r) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= len) return;
    const double alpha = *num / *den;
    x[i] += alpha * p[i];
    r[i] -= alpha * Ap[i];
}

__global__ __launch_bounds__(kBlock) void k_cg_update_p(int64_t len, const double* __restrict__ num,
                                                        const double* __restrict__ den, const double* __restrict__ z,
                                                        double* __restrict__ p) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= len) return;
    const double beta = *num / *den;
    p[i] = z[i] + beta * p[i];
}

inline unsigned grid_for(int64_t threads) { return (unsigned)((threads + kBlock - 1) / kBlock); }

int reduce_blocks(int64_t len) {
    int64_t b = (len + kTile - 1) / kTile;
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, bsp_ctx::kMaxRedBlocks));
}

#define BSP_LAUNCH_CHECK()                                                                               \
    do {                                                                                                 \
        hipError_t e_ = hipGetLastError();                                                               \
        if (e_ != hipSuccess) return fail(ctx, BSP_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e_)); \
    } while (0)

template <template <int> class F, class... Args>
bool dispatch_bs(int bs, Args&&... args) {
    switch (bs) {
        case 1: F<1>::run(args...); return true;
        case 2: F<2>::run(args...); return true;
        case 3: F<3>::run(args...); return true;
        case 4: F<4>::run(args...); return true;
        case 5: F<5>::run(args...); return true;
        case 6: F<6>::run(args...); return true;
        case 7: F<7>::run(args...); return true;
        case 8: F<8>::run(args...); return true;
        case 16: F<16>::run(args...); return true;
        case 32: F<32>::run(args...); return true;
        default: return false;
    }
}

template <int BS>
struct BsmvL {
    static void run(hipStream_t st, int n, const int* rp, const int* ci, const double* val, const double* x, double* y,
                    int acc) {
        constexpr int RPW = kWave / BS;
        const int64_t waves = (n + RPW - 1) / RPW;
        hipLaunchKernelGGL(k_bsmv<BS>, dim3(grid_for(waves * kWave)), dim3(kBlock), 0, st, n, rp, ci, val, x, y, acc);
    }
};

template <int BS>
struct ResidL {
    static void run(hipStream_t st, int n, const int* rp, const int* ci, const double* val, const double* b,
                    const double* x, double* r) {
        constexpr int RPW = kWave / BS;
        const int64_t waves = (n + RPW - 1) / RPW;
        hipLaunchKernelGGL(k_residual<BS>, dim3(grid_for(waves * kWave)), dim3(kBlock), 0, st, n, rp, ci, val, b, x, r);
    }
};

template <int BS>
struct DiagL {
    static void run(hipStream_t st, int n, const double* dinv, double* x) {
        constexpr int RPW = kWave / BS;
        const int64_t waves = (n + RPW - 1) / RPW;
        hipLaunchKernelGGL(k_block_diag<BS>, dim3(grid_for(waves * kWave)), dim3(kBlock), 0, st, n, dinv, x);
    }
};

}  // namespace

int launch_bsmv(const bsp_ctx* ctx, int n, int bs, const int* rp, const int* ci, const double* val,
                const double* x, double* y, int accumulate) {
    if (n <= 0) return BSP_OK;
    if (!dispatch_bs<BsmvL>(bs, ctx->stream, n, rp, ci, val, x, y, accumulate)) {
        BSP_CHECK(bs <= kWave, BSP_E_UNSUPPORTED, "block size > 64 is not supported");
        const int rpw = kWave / bs;
        const int64_t waves = (n + rpw - 1) / rpw;
        hipLaunchKernelGGL(k_bsmv_generic, dim3(grid_for(waves * kWave)), dim3(kBlock), 0, ctx->stream, n, bs, rp, ci,
                           val, x, y, accumulate);
    }
    BSP_LAUNCH_CHECK();
    return BSP_OK;
}

bool bsmv_perm_supported(int bs) { return bs == 2 || bs == 4 || bs == 6 || bs == 8; }

int launch_permute_blocks(const bsp_ctx* ctx, int64_t nnzb, int bs, const double* val, double* valp) {
    const int64_t total = nnzb * bs * bs;
    if (total <= 0) return BSP_OK;
    hipLaunchKernelGGL(k_permute_blocks, dim3(grid_for(total)), dim3(kBlock), 0, ctx->stream, total, bs, val, valp);
    BSP_LAUNCH_CHECK();
    return BSP_OK;
}

template <int BS>
void launch_perm_bs(hipStream_t st, int n, const int* rp, const int* ci, const double* valp, const double* x,
                    double* y, int accumulate) {
    constexpr int rpw = kWave / BS;
    const int64_t waves = (n + rpw - 1) / rpw;
    const int nwg = (int)grid_for(waves * kWave);
    const int grid8 = (nwg + 7) / 8 * 8;
    hipLaunchKernelGGL(k_bsmv_perm<BS>, dim3(grid8), dim3(kBlock), 0, st, n, rp, ci, valp, x, y, accumulate, nwg);
}

template <int BS>
void launch_ell_bs(hipStream_t st, int n, int W, const int* cie, const double* vale, const double* x, double* y,
                   int accumulate) {
    constexpr int rpw = kWave / BS;
    const int64_t waves = (n + rpw - 1) / rpw;
    const int nwg = (int)grid_for(waves * kWave);
    const int grid8 = (nwg + 7) / 8 * 8;
    hipLaunchKernelGGL(k_bsmv_ell<BS>, dim3(grid8), dim3(kBlock), 0, st, n, W, cie, vale, x, y, accumulate, nwg);
}

int launch_bsmv_ell(const bsp_ctx* ctx, int n, int bs, int W, const int* cie, const double* vale, const double* x,
                    double* y, int accumulate) {
    if (n <= 0) return BSP_OK;
    switch (bs) {
        case 2: launch_ell_bs<2>(ctx->stream, n, W, cie, vale, x, y, accumulate); break;
        case 4: launch_ell_bs<4>(ctx->stream, n, W, cie, vale, x, y, accumulate); break;
        case 6: launch_ell_bs<6>(ctx->stream, n, W, cie, vale, x, y, accumulate); break;
        case 8: launch_ell_bs<8>(ctx->stream, n, W, cie, vale, x, y, accumulate); break;
        default: return fail(ctx, BSP_E_UNSUPPORTED, "ELL bsmv needs block size 2, 4, 6 or 8");
    }
    BSP_LAUNCH_CHECK();
    return BSP_OK;
}

int launch_ell_columns(const bsp_ctx* ctx, int n, int W, const int* rp, const int* ci, int* cie) {
    const int64_t slots = (int64_t)n * W;
    if (slots <= 0) return BSP_OK;
    hipLaunchKernelGGL(k_ell_columns, dim3(grid_for(slots)), dim3(kBlock), 0, ctx->stream, slots, W, rp, ci, cie);
    BSP_LAUNCH_CHECK();
    return BSP_OK;
}

int launch_ell_values_f64(const bsp_ctx* ctx, int n, int bs, int W, const int* rp, const double* val, double* vale) {
    const int64_t total = (int64_t)n * W * bs * bs;
    if (total <= 0) return BSP_OK;
    hipLaunchKernelGGL((k_ell_values<double, 2>), dim3(grid_for(total)), dim3(kBlock), 0, ctx->stream, total, bs, W,
                       rp, val, vale);
    BSP_LAUNCH_CHECK();
    return BSP_OK;
}

int launch_ell_values_f32(const bsp_ctx* ctx, int n, int bs, int W, const int* rp, const double* val, float* vale) {
    const int64_t total = (int64_t)n * W * bs * bs;
    if (total <= 0) return BSP_OK;
    hipLaunchKernelGGL((k_ell_values<float, 4>), dim3(grid_for(total)), dim3(kBlock), 0, ctx->stream, total, bs, W,
                       rp, val, vale);
    BSP_LAUNCH_CHECK();
    return BSP_OK;
}

int launch_bsmv_perm(const bsp_ctx* ctx, int n, int bs, const int* rp, const int* ci, const double* valp,
                     const double* x, double* y, int accumulate) {
    if (n <= 0) return BSP_OK;
    switch (bs) {
        case 2: launch_perm_bs<2>(ctx->stream, n, rp, ci, valp, x, y, accumulate); break;
        case 4: launch_perm_bs<4>(ctx->stream, n, rp, ci, valp, x, y, accumulate); break;
        case 6: launch_perm_bs<6>(ctx->stream, n, rp, ci, valp, x, y, accumulate); break;
        case 8: launch_perm_bs<8>(ctx->stream, n, rp, ci, valp, x, y, accumulate); break;
        default: return fail(ctx, BSP_E_UNSUPPORTED, "chunk-major bsmv needs block size 2, 4, 6 or 8");
    }
    BSP_LAUNCH_CHECK();
    return BSP_OK;
}

int launch_bsmv_dim(const bsp_ctx* ctx, int n, int B, int D, int dim, const int* rp, const int* ci,
                    const double* val, const double* x, double* y, int accumulate) {
    if (n <= 0) return BSP_OK;
    hipLaunchKernelGGL(k_bsmv_dim, dim3(grid_for((int64_t)n * B * D)), dim3(kBlock), 0, ctx->stream, n, B, D, dim, rp,
                       ci, val, x, y, accumulate);
    BSP_LAUNCH_CHECK();
    return BSP_OK;
}

int launch_bsmm_rhs(const bsp_ctx* ctx, int n, int bs, int nrhs, const int* rp, const int* ci, const double* val,
                    const double* X, double* Y) {
    if (n <= 0) return BSP_OK;
    if (nrhs <= 0) return BSP_OK;
    if (!dispatch_bs<BsmmRhsL>(bs, ctx->stream, n, nrhs, rp, ci, val, X, Y)) {
        hipLaunchKernelGGL(k_bsmm_rhs, dim3(grid_for((int64_t)n * bs * nrhs)), dim3(kBlock), 0, ctx->stream, n, bs,
                           nrhs, rp, ci, val, X, Y);
    }
    BSP_LAUNCH_CHECK();
    return BSP_OK;
}

int launch_block_diag(const bsp_ctx* ctx, int n, int bs, const double* dinv, double* x) {
    if (n <= 0) return BSP_OK;
    if (!dispatch_bs<DiagL>(bs, ctx->stream, n, dinv, x)) {
        return fail(ctx, BSP_E_UNSUPPORTED, "block diagonal apply: unsupported block size");
    }
    BSP_LAUNCH_CHECK();
    return BSP_OK;
}

int reduce(const bsp_ctx* ctx, int kind, int64_t len, const double* x, const double* y, int slot, double* host_out) {
    double* out = ctx->scalars.as<double>() + slot;
    const int nb = reduce_blocks(len);
    double* part = ctx->red_partials.as<double>();
    unsigned* cnt = ctx->red_counter.as<unsigned>();
    const bool vec = aligned16(x) && (kind != 0 || aligned16(y));
#define BSP_RED(K, V) hipLaunchKernelGGL((k_reduce<K, V>), dim3(nb), dim3(kBlock), 0, ctx->stream, len, x, y, part, cnt, out)
    switch (kind) {
        case 0: if (vec) BSP_RED(0, true); else BSP_RED(0, false); break;
        case 1: if (vec) BSP_RED(1, true); else BSP_RED(1, false); break;
        default: if (vec) BSP_RED(2, true); else BSP_RED(2, false); break;
    }
#undef BSP_RED
    BSP_LAUNCH_CHECK();
    if (host_out) {
        BSP_HIP(hipMemcpyAsync(ctx->pinned, out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        BSP_HIP(hipStreamSynchronize(ctx->stream));
        *host_out = ctx->pinned[0];
    }
    return BSP_OK;
}

int launch_axpy(const bsp_ctx* ctx, int64_t len, double a, const double* x, double* y) {
    if (len <= 0) return BSP_OK;
    hipLaunchKernelGGL(k_axpy, dim3(grid_for(len)), dim3(kBlock), 0, ctx->stream, len, a, x, y);
    BSP_LAUNCH_CHECK();
    return BSP_OK;
}

int launch_scal(const bsp_ctx* ctx, int64_t len, double a, double* x) {
    if (len <= 0) return BSP_OK;
    hipLaunchKernelGGL(k_scal, dim3(grid_for(len)), dim3(kBlock), 0, ctx->stream, len, a, x);
    BSP_LAUNCH_CHECK();
    return BSP_OK;
}

int launch_mgs_step(const bsp_ctx* ctx, int64_t len, double* v, const double* qprev, const double* h_prev,
                    const double* qnext, double* h_out) {
    const int nb = reduce_blocks(len);
    const bool vec = aligned16(v) && (!qprev || aligned16(qprev)) && (!qnext || aligned16(qnext));
    if (vec)
        hipLaunchKernelGGL(k_mgs_step<true>, dim3(nb), dim3(kBlock), 0, ctx->stream, len, v, qprev, h_prev, qnext,
                           ctx->red_partials.as<double>(), ctx->red_counter.as<unsigned>(), h_out);
    else
        hipLaunchKernelGGL(k_mgs_step<false>, dim3(nb), dim3(kBlock), 0, ctx->stream, len, v, qprev, h_prev, qnext,
                           ctx->red_partials.as<double>(), ctx->red_counter.as<unsigned>(), h_out);
    BSP_LAUNCH_CHECK();
    return BSP_OK;
}

int launch_scale_div(const bsp_ctx* ctx, int64_t len, const double* v, const double* h, double* q) {
    hipLaunchKernelGGL(k_scale_div<double>, dim3(grid_for(len)), dim3(kBlock), 0, ctx->stream, len, v, h, q);
    BSP_LAUNCH_CHECK();
    return BSP_OK;
}

int launch_scale_div(const bsp_ctx* ctx, int64_t len, const double* v, const double* h, float* q) {
    hipLaunchKernelGGL(k_scale_div<float>, dim3(grid_for(len)), dim3(kBlock), 0, ctx->stream, len, v, h, q);
    BSP_LAUNCH_CHECK();
    return BSP_OK;
}

int launch_combine_add(const bsp_ctx* ctx, int64_t len, int m, const double* const* Q, const double* c, double* x) {
    // chunks of 64 vectors, in index order
    for (int i0 = 0; i0 < m; i0 += kMaxBasis) {
        const int mm = std::min(kMaxBasis, m - i0);
        PtrTable t{};
        for (int i = 0; i < mm; ++i) t.q[i] = Q[i0 + i];
        hipLaunchKernelGGL(k_combine_add, dim3(grid_for(len)), dim3(kBlock), 0, ctx->stream, len, mm, t, c + i0, x);
        BSP_LAUNCH_CHECK();
    }
    return BSP_OK;
}

int launch_basis_rotate(const bsp_ctx* ctx, int64_t len, int m, int kp, const double* const* Q, const double* P,
                        double* const* Qout) {
    BSP_CHECK(m <= kMaxBasis && kp <= kMaxBasis, BSP_E_UNSUPPORTED, "Krylov basis larger than 64 vectors");
    PtrTable t{};
    OutTable o{};
    bool vec = true;
    for (int i = 0; i < m; ++i) {
        t.q[i] = Q[i];
        vec = vec && (uintptr_t)Q[i] % 16 == 0;
    }
    for (int j = 0; j < kp; ++j) {
        o.q[j] = Qout[j];
        vec = vec && (uintptr_t)Qout[j] % 16 == 0;
    }
    if (vec && len % 2 == 0 && kp >= 1 && kp <= 16) {
        bool done = false;
        auto go = [&](auto jn) {
            constexpr int JN = decltype(jn)::value;
            if (JN != kp) return;
            hipLaunchKernelGGL(k_basis_rotate_jn<JN>, dim3(grid_for(len / 2)), dim3(kBlock), 0, ctx->stream, len, m,
                               t, P, o);
            done = true;
        };
        [&]<int... J>(std::integer_sequence<int, J...>) { (go(std::integral_constant<int, J + 1>{}), ...); }(
            std::make_integer_sequence<int, 16>{});
        if (done) {
            BSP_LAUNCH_CHECK();
            return BSP_OK;
        }
    }
    hipLaunchKernelGGL(k_basis_rotate, dim3(grid_for((len + 1) / 2)), dim3(kBlock), 0, ctx->stream, len, m, kp, t, P, o,
                       vec ? 1 : 0);
    BSP_LAUNCH_CHECK();
    return BSP_OK;
}

int launch_residual(const bsp_ctx* ctx, int n, int bs, const int* rp, const int* ci, const double* val,
                    const double* b, const double* x, double* r) {
    if (n <= 0) return BSP_OK;
    if (!dispatch_bs<ResidL>(bs, ctx->stream, n, rp, ci, val, b, x, r)) {
        // generic: r = A x then r = b - r
        int rc = launch_bsmv(ctx, n, bs, rp, ci, val, x, r, 0);
        if (rc) return rc;
        rc = launch_scal(ctx, (int64_t)n * bs, -1.0, r);
        if (rc) return rc;
        return launch_axpy(ctx, (int64_t)n * bs, 1.0, b, r);
    }
    BSP_LAUNCH_CHECK();
    return BSP_OK;
}

int launch_copy(const bsp_ctx* ctx, int64_t len, const double* x, double* y) {
    if (len <= 0) return BSP_OK;
    hipLaunchKernelGGL(k_copy, dim3(grid_for(len)), dim3(kBlock), 0, ctx->stream, len, x, y);
    BSP_LAUNCH_CHECK();
    return BSP_OK;
}

int launch_fill_u32(const bsp_ctx* ctx, int64_t len, uint32_t v, uint32_t* p) {
    if (len <= 0) return BSP_OK;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(len)), dim3(kBlock), 0, ctx->stream, len, v, p);
    BSP_LAUNCH_CHECK();
    return BSP_OK;
}

int launch_cg_update_xr(const bsp_ctx* ctx, int64_t len, const double* num, const double* den, const double* p,
                        const double* Ap, double* x, double* r) {
    hipLaunchKernelGGL(k_cg_update_xr, dim3(grid_for(len)), dim3(kBlock), 0, ctx->stream, len, num, den, p, Ap, x, r);
    BSP_LAUNCH_CHECK();
    return BSP_OK;
}

int launch_cg_update_p(const bsp_ctx* ctx, int64_t len, const double* num, const double* den, const double* z,
                       double* p) {
    hipLaunchKernelGGL(k_cg_update_p, dim3(grid_for(len)), dim3(kBlock), 0, ctx->stream, len, num, den, z, p);
    BSP_LAUNCH_CHECK();
    return BSP_OK;
}

}  // namespace bsp
