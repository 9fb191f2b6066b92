// bsp_sweeps.hip -- level-scheduled block triangular solves and SOR sweeps (gfx950).
//
// A sweep visits the block rows of a schedule: rows sorted by level, every
// level padded to whole wavefronts (one wavefront never holds two levels).  A
// block row is owned by BS lanes (lane r owns row r of each BS x BS block), as
// in the bsmv kernel.
//
//  * trsv: x_i <- Dinv_i (x_i - sum_j T_ij x_j) over the rows of T (the strict
//    part of an ILU0 / IChol0 factor); Dinv may be absent (unit diagonal).
//  * SOR:  xn_i = xo_i + (-w xo_i + w Dinv_i (b_i - sum_{j != i} A_ij x_j)), with x_j
//    from xn on the dependency side (forward: j < i) and from xo elsewhere
//    (blocksp_preconditioners.hpp:135-202).
//
// Launch modes:
//  * level launches: one kernel per level (kernel boundaries order the levels), replayed
//    from a captured HIP graph when there are many levels;
//  * one workgroup: a sweep of many narrow levels (each within one pass of a 1024-thread
//    workgroup) in a single workgroup, with a barrier between levels; over an ELL copy of blocks
//    up to 4 x 4 the operands of later levels are loaded ahead and the vector may sit in LDS;
//  * single launch ("sync-free"): a persistent grid takes chunks of the
//    schedule in order from a device ticket counter, so every row a chunk waits
//    on belongs to a chunk that is already held by a running workgroup.  A row
//    polls its dependencies' completion flags with relaxed coherent loads, then
//    issues one acquire fence; a wavefront publishes its rows with one release
//    fence followed by relaxed flag stores (flag = epoch of this launch).  Every
//    wait is bounded (spin_limit polls, or immediately once any wait has failed
//    and set *err), so every wavefront always exits.
// Triangular solves of many levels read a slot-ordered ELL copy of their blocks when it is
// built (columns and blocks indexed by schedule slot, each level as wide as its widest row), so
// a lane's loads do not wait for the row number; the copy keeps each row's CSR order.  SOR
// sweeps of many levels read one of columns and block positions of the off-diagonal blocks
// (the matrix values stay where they are, so a change of values takes effect at once).
#include "bsp_internal.hpp"
#include "bsp_lanes.hpp"

#include <algorithm>
#include <cstring>

namespace bsp {

namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kDirectLevels = 8;     // up to this many levels: direct launches
constexpr size_t kMaxGraphs = 8;     // captured sweep graphs kept per schedule

template <int BS, typename VT>
__device__ __forceinline__ void load_row(const VT* __restrict__ src, VT (&a)[BS]) {
    if constexpr (std::is_same_v<VT, double> && BS % 2 == 0) {
        const double2* s2 = reinterpret_cast<const double2*>(src);
#pragma unroll
        for (int c = 0; c < BS / 2; ++c) {
            const double2 t = s2[c];
            a[2 * c] = t.x;
            a[2 * c + 1] = t.y;
        }
    } else if constexpr (std::is_same_v<VT, float> && BS % 4 == 0) {
        const float4* s4 = reinterpret_cast<const float4*>(src);
#pragma unroll
        for (int c = 0; c < BS / 4; ++c) {
            const float4 t = s4[c];
            a[4 * c] = t.x;
            a[4 * c + 1] = t.y;
            a[4 * c + 2] = t.z;
            a[4 * c + 3] = t.w;
        }
    } else {
#pragma unroll
        for (int c = 0; c < BS; ++c) a[c] = src[c];
    }
}

// BF16 matrix storage (bit patterns in uint16), widened to FP32: a BF16 value is the upper
// half of the FP32 with the same bits.
struct bf16_t {
    uint16_t bits;
};

__device__ __forceinline__ float bf16_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }

// One matrix row of BS values of storage type MT into arithmetic type VT.
template <int BS, typename MT, typename VT>
__device__ __forceinline__ void load_mrow(const MT* __restrict__ src, VT (&a)[BS]) {
    if constexpr (std::is_same_v<MT, VT>) {
        load_row<BS, VT>(src, a);
    } else {
        static_assert(std::is_same_v<MT, bf16_t> && std::is_same_v<VT, float>, "BF16 rows widen to FP32");
        if constexpr (BS % 8 == 0) {
            // factor rows are read once per sweep: streamed, so the vector keeps the caches
            typedef unsigned int u4v __attribute__((ext_vector_type(4)));
            const u4v* s8 = reinterpret_cast<const u4v*>(src);
#pragma unroll
            for (int c = 0; c < BS / 8; ++c) {
                const u4v t = __builtin_nontemporal_load(s8 + c);
                a[8 * c + 0] = bf16_lo(t.x);
                a[8 * c + 1] = bf16_hi(t.x);
                a[8 * c + 2] = bf16_lo(t.y);
                a[8 * c + 3] = bf16_hi(t.y);
                a[8 * c + 4] = bf16_lo(t.z);
                a[8 * c + 5] = bf16_hi(t.z);
                a[8 * c + 6] = bf16_lo(t.w);
                a[8 * c + 7] = bf16_hi(t.w);
            }
        } else {
#pragma unroll
            for (int c = 0; c < BS; ++c) a[c] = __uint_as_float((uint32_t)src[c].bits << 16);
        }
    }
}

template <int BS, typename VT>
__device__ __forceinline__ VT row_dot(const VT (&a)[BS], const VT (&x)[BS]) {
    VT s = VT(0);
#pragma unroll
    for (int c = 0; c < BS; ++c) s = fma(a[c], x[c], s);
    return s;
}

__device__ __forceinline__ void wait_relaxed(const uint32_t* flag, uint32_t epoch, int64_t spin_limit, uint32_t* err) {
    int64_t it = 0;
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch) {
        ++it;
        if (it > spin_limit ||
            ((it & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)) {
            __hip_atomic_fetch_or(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

struct SweepArgs {
    int nchunks;
    const int* order;
    const int* rp;
    const int* ci;
    const void* val;
    const void* dinv;
    void* x;            // trsv: in/out; SOR: xn (out)
    const void* b;      // SOR right-hand side
    const void* xo;     // SOR previous iterate (nullptr: zero)
    double w;
    int forward;
    uint32_t* flags;
    uint32_t epoch;
    unsigned long long* ticket;
    unsigned long long base;
    int64_t spin_limit;
    uint32_t* err;
    const int* pos;     // SOR in an elimination order: step of each row (nullptr: natural order)
    const int* ecol;    // slot-ordered ELL copy (nullptr: CSR); slot s's entry k at
    const void* eval;   // ell_base + s * ell_k + k (per level launch, or uniform)
    const int* epos;    // SOR: the entry's block position in the matrix (its values stay live)
    int ell_k;
    int64_t ell_base;
};

constexpr int kEllMax = 4;     // entries per slot of the ELL copies
constexpr int kEllMaxBs = 8;   // block sizes with an ELL form (a slot's blocks stay in registers)
constexpr int kSorEllMaxBs = 4;   // the SOR form also holds the old iterate's rows: larger
                                  // blocks spill in the 1024-thread one-workgroup sweeps

// One chunk of slots: [slot_begin, slot_end) with kWavesPerBlock * RPW slots per block.
// Factor values and diagonal inverses are stored as MT, vectors and arithmetic are VT.
template <int BS, typename VT, typename MT, bool SF, bool SOR>
__device__ __forceinline__ void sweep_chunk(const SweepArgs& a, int slot_begin, int slot_end) {
    constexpr int RPW = kWave / BS;
    const MT* __restrict__ val = static_cast<const MT*>(a.val);
    const MT* __restrict__ dinv = static_cast<const MT*>(a.dinv);
    VT* x = static_cast<VT*>(a.x);
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = threadIdx.x / kWave;
    const int g = lane / BS;
    const int r = lane - g * BS;
    const int slot = slot_begin + wv * RPW + g;
    const bool have = g < RPW && slot < slot_end;
    const int i = have ? a.order[slot] : -1;
    const bool active = i >= 0;
    const int p0 = active ? a.rp[i] : 0, p1 = active ? a.rp[i + 1] : 0;
    if constexpr (SF) {
        if (active) {
            for (int p = p0; p < p1; ++p) {
                const int c = a.ci[p];
                bool dep = true;
                if constexpr (SOR) dep = a.pos ? (a.forward ? a.pos[c] < a.pos[i] : a.pos[c] > a.pos[i])
                                               : (a.forward ? c < i : c > i);
                if (dep) wait_relaxed(&a.flags[c], a.epoch, a.spin_limit, a.err);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    VT s = VT(0), xold = VT(0);
    if constexpr (SOR && BS <= kSorEllMaxBs) {
        if (a.ecol) {
            // ELL copy of the off-diagonal blocks, each slot's in its row's CSR order, indexed
            // by slot so they load without waiting for the row number: ecol = column on the
            // dependency side, -2 - column on the other (read from the previous iterate), -1
            // padding; epos = the block's position in the matrix, whose values are read live
            const VT* xo = static_cast<const VT*>(a.xo);
            const VT* b = static_cast<const VT*>(a.b);
            const int K = a.ell_k;
            int cq[kEllMax], pq[kEllMax];
#pragma unroll
            for (int k = 0; k < kEllMax; ++k) {
                cq[k] = -1;
                pq[k] = 0;
                if (k < K && have) {
                    const int64_t e = a.ell_base + (int64_t)slot * K + k;
                    cq[k] = a.ecol[e];
                    pq[k] = a.epos[e];
                }
            }
            VT av[kEllMax][BS], xv[kEllMax][BS];
#pragma unroll
            for (int k = 0; k < kEllMax; ++k) {
                if (cq[k] == -1) continue;
                load_mrow<BS, MT, VT>(val + ((size_t)pq[k] * BS + r) * BS, av[k]);
                if (cq[k] >= 0) {
                    load_row<BS, VT>(x + (size_t)cq[k] * BS, xv[k]);
                } else if (xo) {
                    load_row<BS, VT>(xo + (size_t)(-2 - cq[k]) * BS, xv[k]);
                } else {
#pragma unroll
                    for (int c = 0; c < BS; ++c) xv[k][c] = VT(0);
                }
            }
            VT ph = VT(0);
#pragma unroll
            for (int k = 0; k < kEllMax; ++k)
                if (cq[k] != -1) ph += row_dot<BS, VT>(av[k], xv[k]);
            if (active) {
                xold = xo ? xo[(size_t)i * BS + r] : VT(0);
                s = b[(size_t)i * BS + r] - ph;
            }
        }
    }
    if (active) {
        if constexpr (SOR) {
            if (!a.ecol) {
                const VT* xo = static_cast<const VT*>(a.xo);
                const VT* b = static_cast<const VT*>(a.b);
                xold = xo ? xo[(size_t)i * BS + r] : VT(0);
                VT ph = VT(0);
                for (int p = p0; p < p1; ++p) {
                    const int c = a.ci[p];
                    if (c == i) continue;
                    const bool dep = a.pos ? (a.forward ? a.pos[c] < a.pos[i] : a.pos[c] > a.pos[i])
                                           : (a.forward ? c < i : c > i);
                    VT av[BS], xv[BS];
                    load_mrow<BS, MT, VT>(val + ((size_t)p * BS + r) * BS, av);
                    if (dep) {
                        load_row<BS, VT>(x + (size_t)c * BS, xv);
                    } else if (xo) {
                        load_row<BS, VT>(xo + (size_t)c * BS, xv);
                    } else {
#pragma unroll
                        for (int k = 0; k < BS; ++k) xv[k] = VT(0);
                    }
                    ph += row_dot<BS, VT>(av, xv);
                }
                s = b[(size_t)i * BS + r] - ph;
            }
        } else if (!a.ecol) {
            s = x[(size_t)i * BS + r];
            // NB entries at a time: their columns, blocks and vector rows load together (one
            // load latency per NB entries instead of one per entry), then the products are
            // subtracted in the row's order.  FP64 only: the natural-order FP64 ILU0 of the 512^2
            // DG matrix measured 4% faster, the BF16 sweeps of its coloured factor (rows of one or
            // two entries) 25% slower (13.8 -> 17.3 us a level, the registers cost occupancy).
            constexpr int W = 2 * BS * (int)sizeof(VT) / 4;   // registers an entry holds
            constexpr int NB = !std::is_same_v<VT, double> || W >= 64 ? 1 : (64 / W >= 4 ? 4 : 64 / W);
            for (int p = p0; p < p1; p += NB) {
                int c[NB];
#pragma unroll
                for (int u = 0; u < NB; ++u) c[u] = p + u < p1 ? a.ci[p + u] : -1;
                VT av[NB][BS], xv[NB][BS];
#pragma unroll
                for (int u = 0; u < NB; ++u)
                    if (c[u] >= 0) {
                        load_mrow<BS, MT, VT>(val + ((size_t)(p + u) * BS + r) * BS, av[u]);
                        load_row<BS, VT>(x + (size_t)c[u] * BS, xv[u]);
                    }
#pragma unroll
                for (int u = 0; u < NB; ++u)
                    if (c[u] >= 0) s -= row_dot<BS, VT>(av[u], xv[u]);
            }
        }
    }
    if constexpr (!SOR && BS <= kEllMaxBs) {
        if (a.ecol) {
            // ELL copy: columns and blocks are indexed by slot, so they load without waiting
            // for the row number; padding (-1) ends a slot's entries, which keep the CSR order
            const MT* __restrict__ ev = static_cast<const MT*>(a.eval);
            const int K = a.ell_k;
            int cq[kEllMax];
            VT av[kEllMax][BS];
#pragma unroll
            for (int k = 0; k < kEllMax; ++k) {
                cq[k] = -1;
                if (k < K && have) {
                    const int64_t e = a.ell_base + (int64_t)slot * K + k;
                    cq[k] = a.ecol[e];
                    load_mrow<BS, MT, VT>(ev + ((size_t)e * BS + r) * BS, av[k]);
                }
            }
            if (active) s = x[(size_t)i * BS + r];
            VT xv[kEllMax][BS];
#pragma unroll
            for (int k = 0; k < kEllMax; ++k)
                if (cq[k] >= 0) load_row<BS, VT>(x + (size_t)cq[k] * BS, xv[k]);
#pragma unroll
            for (int k = 0; k < kEllMax; ++k)
                if (cq[k] >= 0) s -= row_dot<BS, VT>(av[k], xv[k]);
        }
    }
    VT out = s;
    if (dinv) {
        const int base = (g < RPW ? g : 0) * BS;
        VT d[BS];
        if (active) load_mrow<BS, MT, VT>(dinv + ((size_t)i * BS + r) * BS, d);
        out = VT(0);
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            const VT sc = __shfl(s, base + c, kWave);
            if (active) out = fma(d[c], sc, out);
        }
    }
    if constexpr (SOR) {
        const VT w = (VT)a.w;
        out = xold + (-w * xold + w * out);
    }
    if (active) x[(size_t)i * BS + r] = out;
    if constexpr (SF) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (active && r == 0) __hip_atomic_store(&a.flags[i], a.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <int BS, typename VT, typename MT, bool SOR>
__global__ __launch_bounds__(kBlock) void k_sweep_level(SweepArgs a, int slot0, int count) {
    constexpr int SPB = kWavesPerBlock * (kWave / BS);
    sweep_chunk<BS, VT, MT, false, SOR>(a, slot0 + blockIdx.x * SPB, slot0 + count);
}

template <int BS, typename VT, typename MT, bool SOR>
__global__ __launch_bounds__(kBlock) void k_sweep_syncfree(SweepArgs a, int nslots) {
    constexpr int SPB = kWavesPerBlock * (kWave / BS);
    __shared__ long long s_chunk;
    for (;;) {
        if (threadIdx.x == 0)
            s_chunk = (long long)(__hip_atomic_fetch_add(a.ticket, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) -
                                  a.base);
        __syncthreads();
        const long long chunk = s_chunk;
        __syncthreads();
        if (chunk >= a.nchunks) break;
        sweep_chunk<BS, VT, MT, true, SOR>(a, (int)chunk * SPB, nslots);
    }
}

// A whole sweep (levels [l0, l1)) in one workgroup of kOneBlock threads, a barrier between levels:
// for small systems with many narrow levels a level costs a few dependent loads instead of a
// kernel launch.
constexpr int kOneBlock = 1024;

template <int BS, typename VT, typename MT, bool SOR>
__global__ __launch_bounds__(kOneBlock) void k_sweep_onewg(SweepArgs a, const int* __restrict__ level_ptr,
                                                           const int* __restrict__ ell_kl,
                                                           const int64_t* __restrict__ ell_off, int l0, int l1) {
    constexpr int SPP = (kOneBlock / kWave) * (kWave / BS);   // slots per pass of the workgroup
    for (int l = l0; l < l1; ++l) {
        const int s0 = level_ptr[l], s1 = level_ptr[l + 1];
        SweepArgs b = a;
        if (a.ecol) {   // the level's width and the offset of its entries in the ELL copy
            b.ell_k = ell_kl[l];
            b.ell_base = ell_off[l] - (int64_t)s0 * b.ell_k;
        }
        for (int sb = s0; sb < s1; sb += SPP) sweep_chunk<BS, VT, MT, false, SOR>(b, sb, s1);
        __syncthreads();
    }
}

// The one-workgroup sweep over an ELL copy when every level fits in one pass of the workgroup
// (one slot per lane group; launched with as many wavefronts as the widest level needs, at most
// 1024 threads), with the loads of what no row of the sweep writes pipelined ahead
// of the levels: each slot's row number and entries (columns, and block positions or ELL entry
// indices) two levels ahead, and from them -- right after a level's stores, so they overlap the
// stores and the barrier -- the next level's blocks, diagonal inverse, and the row's own entry
// (triangular solves) or the right-hand side and the previous iterate's rows (SOR).  After the
// barrier only the dependency columns' new values are read.  The arithmetic is sweep_chunk's,
// in the same order, so results are bitwise the same.
//
// A level here is one wavefront's instruction stream between two barriers, and its time is set
// by the number of instructions the wavefront issues (loading further ahead does not shorten
// it), so the per-slot code is kept short: no branches (padding entries and missing rows read
// row 0 and are masked by selects), KM entries per slot (2 when no level is wider), operand
// addresses as 32-bit byte offsets from kernel-argument bases, and the diagonal block's row
// values moved between the lanes of a block row by DPP where the block size divides a quad.
// With the vector in LDS, its rows go to memory once, after the last level.
struct EllIdx {
    bool have;         // the lane group holds a slot of the level
    int i;
    int cq[kEllMax];
    int pq[kEllMax];   // SOR: block positions in the matrix; triangular solves: ELL entry indices
};

template <int BS, typename VT>
struct EllRow {
    int i;                // the slot's row; -1: none (its operands are then row 0's, unused)
    int cq[kEllMax];
    VT av[kEllMax][BS];
    VT xv[kEllMax][BS];   // SOR: rows of the previous iterate (columns on the other side)
    VT d[BS];
    VT s;                 // triangular solve: x_i; SOR: b_i
    VT xold;
};

// An element of `base` at a 32-bit byte offset (the host keeps every operand array of these
// sweeps below 4 GiB): the loads take the base from scalar registers and need no 64-bit
// address arithmetic.
template <typename T>
__device__ __forceinline__ const T* at32(const void* base, uint32_t byte_off) {
    return reinterpret_cast<const T*>(static_cast<const char*>(base) + byte_off);
}

// The indices of a level of width K whose first slot is `first` and whose entries start at
// `ebase`.  Lanes without a slot read the level's first slot (every level holds a row there)
// and are masked by `have` when the operands are loaded.
template <bool SOR, int KM>
__device__ __forceinline__ void ell_idx_load(const SweepArgs& a, int slot, int first, bool have, int K, int ebase,
                                             EllIdx& t) {
    const int sl = have ? slot : first;
    t.have = have;
    t.i = *at32<int>(a.order, (uint32_t)sl * 4u);
    const uint32_t e0 = (uint32_t)(ebase + (sl - first) * K);
#pragma unroll
    for (int k = 0; k < kEllMax; ++k) {
        if (k < KM && k < K) {
            const uint32_t e = e0 + (uint32_t)k;
            t.cq[k] = *at32<int>(a.ecol, e * 4u);
            t.pq[k] = SOR ? *at32<int>(a.epos, e * 4u) : (int)e;
        } else {
            t.cq[k] = -1;
            t.pq[k] = 0;
        }
    }
}

template <int BS, typename VT, typename MT, bool SOR, int KM>
__device__ __forceinline__ void ell_row_load(const SweepArgs& a, const EllIdx& t, EllRow<BS, VT>& q) {
    const uint32_t r = (uint32_t)((threadIdx.x & (kWave - 1)) % BS);
    // The indices were loaded a level ago; the empty asm statements keep the compiler from moving
    // the address arithmetic on them back to their loads (which would wait there for each of them).
    EllIdx u = t;
    asm volatile("" : "+v"(u.i));
#pragma unroll
    for (int k = 0; k < kEllMax; ++k) asm volatile("" : "+v"(u.cq[k]), "+v"(u.pq[k]));
    q.i = u.have ? u.i : -1;
    const uint32_t ir = (uint32_t)max(q.i, 0);
    const void* val = SOR ? a.val : a.eval;
    // without a previous iterate (zero) the right-hand side stands in for it; the products with
    // it are replaced by zeros in ell_row_finish
    const void* xo = a.xo ? a.xo : a.b;
#pragma unroll
    for (int k = 0; k < kEllMax; ++k) {
        q.cq[k] = u.have ? u.cq[k] : -1;
        if (k >= KM) continue;
        const uint32_t p = q.cq[k] == -1 ? 0u : (uint32_t)u.pq[k];
        load_mrow<BS, MT, VT>(at32<MT>(val, (p * BS + r) * BS * (uint32_t)sizeof(MT)), q.av[k]);
        if constexpr (SOR) {
            const uint32_t c = q.cq[k] < -1 ? (uint32_t)(-2 - q.cq[k]) : 0u;
            load_row<BS, VT>(at32<VT>(xo, c * BS * (uint32_t)sizeof(VT)), q.xv[k]);
        }
    }
    if (a.dinv) {
        load_mrow<BS, MT, VT>(at32<MT>(a.dinv, (ir * BS + r) * BS * (uint32_t)sizeof(MT)), q.d);
    } else {
#pragma unroll
        for (int c = 0; c < BS; ++c) q.d[c] = VT(0);
    }
    if constexpr (SOR) {
        q.s = *at32<VT>(a.b, (ir * BS + r) * (uint32_t)sizeof(VT));
        const VT xi = *at32<VT>(xo, (ir * BS + r) * (uint32_t)sizeof(VT));
        q.xold = a.xo ? xi : VT(0);
    } else {
        q.s = *at32<VT>(a.x, (ir * BS + r) * (uint32_t)sizeof(VT));
        q.xold = VT(0);
    }
}

template <int BS, typename VT, bool SOR, bool LX, int KM>
__device__ __forceinline__ void ell_row_finish(const SweepArgs& a, const EllRow<BS, VT>& q, bool have_dinv,
                                               bool have_xo, VT* xs) {
    constexpr int RPW = kWave / BS;
    const int lane = threadIdx.x & (kWave - 1);
    const int g = lane / BS;
    const int r = lane - g * BS;
    const bool active = q.i >= 0;
    // the dependency columns' new values (from LDS: row 0 for the other entries, not used; from
    // memory only where the entry is a dependency -- loads of row 0 there measured 8% slower)
    VT xn[kEllMax][BS];
#pragma unroll
    for (int k = 0; k < KM; ++k) {
        const uint32_t c = (uint32_t)max(q.cq[k], 0);
        if constexpr (LX) {
            load_row<BS, VT>(xs + c * BS, xn[k]);
        } else if (q.cq[k] >= 0) {
            load_row<BS, VT>(at32<VT>(a.x, c * BS * (uint32_t)sizeof(VT)), xn[k]);
        } else {
#pragma unroll
            for (int e = 0; e < BS; ++e) xn[k][e] = VT(0);
        }
    }
    VT s;
    if constexpr (SOR) {
        // each entry's product with the new (dependency) or previous (other side) row, added in
        // the row's order; without a previous iterate the other side's products are zero
        VT ph = VT(0);
#pragma unroll
        for (int k = 0; k < KM; ++k) {
            const bool dep = q.cq[k] >= 0;
            VT xk[BS];
#pragma unroll
            for (int c = 0; c < BS; ++c) xk[c] = dep ? xn[k][c] : q.xv[k][c];
            VT pk = row_dot<BS, VT>(q.av[k], xk);
            pk = (dep || have_xo) ? pk : VT(0);
            ph = q.cq[k] != -1 ? ph + pk : ph;
        }
        s = active ? q.s - ph : VT(0);
    } else {
        s = q.s;
#pragma unroll
        for (int k = 0; k < KM; ++k) {
            const VT pk = row_dot<BS, VT>(q.av[k], xn[k]);
            s = q.cq[k] >= 0 ? s - pk : s;
        }
    }
    VT out = s;
    if (have_dinv) {
        const int base = (g < RPW ? g : 0) * BS;
        out = VT(0);
#pragma unroll
        for (int c = 0; c < BS; ++c) out = fma(q.d[c], block_lane<BS>(s, c, base + c), out);
    }
    if constexpr (SOR) {
        const VT w = (VT)a.w;
        out = q.xold + (-w * q.xold + w * out);
    }
    if (active) {
        // with the vector in LDS the rows go to memory after the last level
        if constexpr (LX)
            xs[(size_t)q.i * BS + r] = out;
        else
            static_cast<VT*>(a.x)[(size_t)q.i * BS + r] = out;
    }
}

template <int BS, typename VT, typename MT, bool SOR, bool LX, int KM>
__global__ __launch_bounds__(kOneBlock) void k_sweep_onewg_ell(SweepArgs a, const int4* __restrict__ level_info,
                                                               int l0, int l1, int nx, int whole) {
    // LX: the solution vector's rows written by this sweep are kept in LDS (nx entries) and
    // dependency columns are read from there; rows of levels before l0 are copied in first, and
    // the rows of levels [l0, l1) are copied out after the last level (all nx entries when
    // `whole`: the sweep covers every level)
    extern __shared__ __align__(16) unsigned char sweep_lds[];
    VT* xs = reinterpret_cast<VT*>(sweep_lds);
    if constexpr (LX) {
        if (l0 > 0) {
            const VT* xg = static_cast<const VT*>(a.x);
            for (int t = threadIdx.x; t < nx; t += blockDim.x) xs[t] = xg[t];
            __syncthreads();
        }
    }
    constexpr int RPW = kWave / BS;
    const int lane = threadIdx.x & (kWave - 1);
    const int g = lane / BS;
    const int off = (int)(threadIdx.x / kWave) * RPW + g;   // the lane group's slot within a level
    const bool grp = g < RPW;
    const bool have_dinv = a.dinv != nullptr;
    const bool have_xo = a.xo != nullptr;
    if (l0 >= l1) return;
    // the indices of level L (clamped to the last level; lanes of levels past it hold no slot)
    auto idx = [&](int L, EllIdx& t) {
        const int4 li = level_info[min(L, l1 - 1)];   // first slot, end slot, width, first entry
        ell_idx_load<SOR, KM>(a, li.x + off, li.x, grp && L < l1 && li.x + off < li.y, li.z, li.w, t);
    };
    EllIdx t;
    EllRow<BS, VT> q;
    idx(l0, t);
    ell_row_load<BS, VT, MT, SOR, KM>(a, t, q);
    idx(l0 + 1, t);
    for (int l = l0;; ++l) {
        ell_row_finish<BS, VT, SOR, LX, KM>(a, q, have_dinv, have_xo, xs);
        if (l + 1 >= l1) break;
        ell_row_load<BS, VT, MT, SOR, KM>(a, t, q);
        idx(l + 2, t);
        __syncthreads();
    }
    if constexpr (LX) {
        __syncthreads();
        VT* xg = static_cast<VT*>(a.x);
        if (whole) {
            for (int e = threadIdx.x; e < nx; e += blockDim.x) xg[e] = xs[e];
        } else {
            const int sb = level_info[l0].x;
            const int ne = (level_info[l1 - 1].y - sb) * BS;
            for (int e = threadIdx.x; e < ne; e += blockDim.x) {
                const int i = a.order[sb + e / BS];   // -1: a padding slot of the level
                if (i < 0) continue;
                const size_t j = (size_t)i * BS + e % BS;
                xg[j] = xs[j];
            }
        }
    }
}

// Levels [l0, l1) of the sweep (l1 < 0: to the last); a partial range is only asked for on
// the direct level-launch path (few levels, level-launch mode) or for a one-workgroup sweep.
template <int BS, typename VT, typename MT, bool SOR>
int run_sweep(bsp_ctx* ctx, Sweep& S, SweepArgs a, int l0, int l1) {
    constexpr int SPB = kWavesPerBlock * (kWave / BS);
    const int nchunks = (S.nslots + SPB - 1) / SPB;
    a.nchunks = nchunks;
    // many levels, each within one pass of the workgroup (wider levels keep their parallel
    // level launches: measured 1.6x slower on the 256^2 x bs 8 DG matrix in one workgroup)
    if (ctx->trsv_mode == 0 && S.nlevels > kDirectLevels && S.max_level_rows <= kOneBlock / BS &&
        (int64_t)S.n * S.bs <= ctx->sweep_onewg_max && S.dlevel_ptr.p) {
        if (!S.dell_kl.p) a.ecol = nullptr;
        constexpr int SPP = (kOneBlock / kWave) * (kWave / BS);
        bool ahead = false;
        if constexpr (BS <= kSorEllMaxBs) {
            // (ELL entry indices are held as int; operand byte offsets as 32-bit)
            // (ell_off is empty when the sweep has no ELL copy)
            const int64_t entries = S.ell_off.empty() ? 0 : S.ell_off.back();
            const int64_t blocks = std::max({S.nnzb, entries, (int64_t)S.n, (int64_t)S.nslots});
            if (a.ecol && S.dlevel_info.p && S.max_level_slots <= SPP && ctx->sweep_ahead &&
                blocks * BS * BS * (int64_t)sizeof(double) < (int64_t(1) << 32)) {
                ahead = true;
                // as many wavefronts as the widest level needs (idle ones would only add loads)
                const int waves = std::max(1, (S.max_level_slots + kWave / BS - 1) / (kWave / BS));
                const int nx = S.n * BS;
                const int64_t lds = (int64_t)nx * sizeof(VT);
                const int lend = l1 < 0 ? S.nlevels : l1;
                const int whole = l0 == 0 && lend == S.nlevels;
                const bool lx = ctx->sweep_lds && lds <= std::min<int64_t>(ctx->lds_per_block, 65536);
                // levels at most 2 entries wide (a triangular factor of a 2D stencil) take the
                // kernel that holds 2 entries per slot
                auto go = [&](auto lxc, auto kmc) {
                    constexpr bool LXc = decltype(lxc)::value;
                    constexpr int KMc = decltype(kmc)::value;
                    hipLaunchKernelGGL((k_sweep_onewg_ell<BS, VT, MT, SOR, LXc, KMc>), dim3(1), dim3(waves * kWave),
                                       LXc ? (size_t)lds : 0, ctx->stream, a, S.dlevel_info.as<int4>(), l0, lend,
                                       nx, whole);
                };
                using T2 = std::integral_constant<int, 2>;
                using T4 = std::integral_constant<int, kEllMax>;
                if (lx) {
                    if (S.ell_k <= 2) go(std::true_type{}, T2{}); else go(std::true_type{}, T4{});
                } else {
                    if (S.ell_k <= 2) go(std::false_type{}, T2{}); else go(std::false_type{}, T4{});
                }
            }
        }
        if (!ahead)
            hipLaunchKernelGGL((k_sweep_onewg<BS, VT, MT, SOR>), dim3(1), dim3(kOneBlock), 0, ctx->stream, a,
                               S.dlevel_ptr.as<int>(), S.dell_kl.as<int>(), S.dell_off.as<int64_t>(), l0,
                               l1 < 0 ? S.nlevels : l1);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail(ctx, BSP_E_HIP, std::string("sweep launch: ") + hipGetErrorString(e));
        return BSP_OK;
    }
    if (ctx->trsv_mode == 1) {
        if (a.ecol && !S.ell_uniform) a.ecol = nullptr;   // levels of different widths: CSR
        const int grid = std::max(1, std::min(nchunks, ctx->num_cus * ctx->sweep_wg_per_cu));
        a.base = S.tickets_used;
        hipLaunchKernelGGL((k_sweep_syncfree<BS, VT, MT, SOR>), dim3(grid), dim3(kBlock), 0, ctx->stream, a, S.nslots);
        S.tickets_used += (unsigned long long)(nchunks + grid);
    } else {
        const bool partial = l0 != 0 || (l1 >= 0 && l1 != S.nlevels);
        if (partial && (S.nlevels > kDirectLevels && ctx->sweep_graphs))
            return fail(ctx, BSP_E_INVALID, "sweep: a level range needs the direct launch path");
        const int lend = l1 < 0 ? S.nlevels : l1;
        auto launch_levels = [&](hipStream_t st) {
            for (int l = l0; l < lend; ++l) {
                const int s0 = S.level_ptr[l], cnt = S.level_ptr[l + 1] - S.level_ptr[l];
                if (cnt <= 0) continue;
                if (a.ecol) {
                    a.ell_k = S.ell_kl[l];
                    a.ell_base = S.ell_off[l] - (int64_t)s0 * S.ell_kl[l];
                }
                hipLaunchKernelGGL((k_sweep_level<BS, VT, MT, SOR>), dim3((cnt + SPB - 1) / SPB), dim3(kBlock), 0, st,
                                   a, s0, cnt);
            }
        };
        if (S.nlevels <= kDirectLevels || !ctx->sweep_graphs) {
            launch_levels(ctx->stream);
        } else {
            // many levels: replay one captured graph instead of nlevels host launches
            uint64_t key[8] = {(uint64_t)(uintptr_t)a.x, (uint64_t)(uintptr_t)a.dinv, (uint64_t)(uintptr_t)a.b,
                               (uint64_t)(uintptr_t)a.xo, 0, (uint64_t)a.forward, (uint64_t)(uintptr_t)a.val,
                               a.ecol ? 2u : 1u};
            std::memcpy(&key[4], &a.w, sizeof(double));
            Sweep::Graph* hit = nullptr;
            for (auto& g : S.graphs)
                if (std::equal(key, key + 8, g.key)) hit = &g;
            if (!hit) {
                hipGraph_t graph = nullptr;
                hipGraphExec_t exec = nullptr;
                hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
                if (e == hipSuccess) {
                    launch_levels(ctx->stream);
                    e = hipStreamEndCapture(ctx->stream, &graph);
                }
                if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                if (graph) (void)hipGraphDestroy(graph);
                if (e != hipSuccess) return fail(ctx, BSP_E_HIP, std::string("sweep graph: ") + hipGetErrorString(e));
                if (S.graphs.size() >= kMaxGraphs) {
                    (void)hipGraphExecDestroy(S.graphs.front().exec);
                    S.graphs.erase(S.graphs.begin());
                }
                S.graphs.push_back({});
                std::copy(key, key + 8, S.graphs.back().key);
                S.graphs.back().exec = exec;
                hit = &S.graphs.back();
            }
            hipError_t e = hipGraphLaunch(hit->exec, ctx->stream);
            if (e != hipSuccess) return fail(ctx, BSP_E_HIP, std::string("sweep graph launch: ") + hipGetErrorString(e));
        }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ctx, BSP_E_HIP, std::string("sweep launch: ") + hipGetErrorString(e));
    return BSP_OK;
}

template <typename VT, typename MT, bool SOR>
int dispatch(bsp_ctx* ctx, Sweep& S, const SweepArgs& a, int l0 = 0, int l1 = -1) {
    switch (S.bs) {
        case 1: return run_sweep<1, VT, MT, SOR>(ctx, S, a, l0, l1);
        case 2: return run_sweep<2, VT, MT, SOR>(ctx, S, a, l0, l1);
        case 3: return run_sweep<3, VT, MT, SOR>(ctx, S, a, l0, l1);
        case 4: return run_sweep<4, VT, MT, SOR>(ctx, S, a, l0, l1);
        case 5: return run_sweep<5, VT, MT, SOR>(ctx, S, a, l0, l1);
        case 6: return run_sweep<6, VT, MT, SOR>(ctx, S, a, l0, l1);
        case 7: return run_sweep<7, VT, MT, SOR>(ctx, S, a, l0, l1);
        case 8: return run_sweep<8, VT, MT, SOR>(ctx, S, a, l0, l1);
        case 16: return run_sweep<16, VT, MT, SOR>(ctx, S, a, l0, l1);
        case 32: return run_sweep<32, VT, MT, SOR>(ctx, S, a, l0, l1);
        default: return fail(ctx, BSP_E_UNSUPPORTED, "sweep: unsupported block size " + std::to_string(S.bs));
    }
}

SweepArgs make_args(bsp_ctx* ctx, Sweep& S, uint32_t epoch) {
    SweepArgs a{};
    a.order = S.order.as<int>();
    a.rp = S.drp;
    a.ci = S.dci;
    a.val = S.dval;
    a.flags = S.flags.as<uint32_t>();
    a.epoch = epoch;
    a.ticket = S.ticket.as<unsigned long long>();
    a.spin_limit = ctx->spin_limit;
    a.err = ctx->err_flag.as<uint32_t>();
    a.pos = S.dpos;
    a.ecol = S.ell_k > 0 && ctx->sweep_ell ? S.ecol.as<int>() : nullptr;
    a.eval = S.eval.p;
    a.epos = S.epos.as<int>();
    a.ell_k = S.ell_k;   // uniform width (single launch); level launches set their level's
    a.ell_base = 0;
    return a;
}

// ELL copy of the levels' slots [s0, s0 + nslots) of width K, entries from e0: thread t
// covers element e of entry k of slot s0 + q.
template <typename MT>
__global__ void k_sweep_ell_gather(int64_t nslots, int K, int b2, int s0, int64_t e0, const int* __restrict__ order,
                                   const int* __restrict__ rp, const int* __restrict__ ci,
                                   const MT* __restrict__ val, int* __restrict__ ecol, MT* __restrict__ eval) {
    const int64_t per = (int64_t)K * b2;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nslots * per) return;
    const int64_t q = t / per;
    const int r = (int)(t - q * per);
    const int k = r / b2, e = r - k * b2;
    const int row = order[s0 + q];
    const int p = row >= 0 ? rp[row] + k : 0;
    const bool have = row >= 0 && p < rp[row + 1];
    const int64_t d = e0 + q * K + k;
    eval[d * b2 + e] = have ? val[(size_t)p * b2 + e] : MT{};
    if (e == 0) ecol[d] = have ? ci[p] : -1;
}

template <typename MT>
hipError_t gather_levels(bsp_ctx* ctx, const Sweep& S, int* ecol, MT* eval) {
    for (int l = 0; l < S.nlevels; ++l) {
        const int s0 = S.level_ptr[l], cnt = S.level_ptr[l + 1] - s0, K = S.ell_kl[l];
        const int64_t total = (int64_t)cnt * K * S.bs * S.bs;
        if (total <= 0) continue;
        hipLaunchKernelGGL(k_sweep_ell_gather<MT>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream,
                           (int64_t)cnt, K, S.bs * S.bs, s0, S.ell_off[l], S.order.as<int>(), S.drp, S.dci,
                           static_cast<const MT*>(S.dval), ecol, eval);
    }
    return hipGetLastError();
}

// SOR copy of the levels' slots [s0, s0 + nslots) of width K, entries from e0: thread t covers
// entry k of slot s0 + q, the k-th off-diagonal block of the slot's row in CSR order.
__global__ void k_sor_ell_gather(int64_t nslots, int K, int s0, int64_t e0, const int* __restrict__ order,
                                 const int* __restrict__ rp, const int* __restrict__ ci, const int* __restrict__ pos,
                                 int forward, int* __restrict__ ecol, int* __restrict__ epos) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nslots * K) return;
    const int64_t q = t / K;
    const int k = (int)(t - q * K);
    const int row = order[s0 + q];
    int col = -1, at = 0;
    if (row >= 0) {
        int seen = 0;
        for (int p = rp[row]; p < rp[row + 1]; ++p) {
            const int c = ci[p];
            if (c == row) continue;
            if (seen++ == k) {
                const bool dep = pos ? (forward ? pos[c] < pos[row] : pos[c] > pos[row])
                                     : (forward ? c < row : c > row);
                col = dep ? c : -2 - c;
                at = p;
                break;
            }
        }
    }
    ecol[e0 + q * K + k] = col;
    epos[e0 + q * K + k] = at;
}

// Level widths kl -> entry offsets, the device copies of both, and the entry count.
hipError_t set_ell_levels(Sweep& S, const std::vector<int>& kl, int64_t& entries) {
    S.ell_kl = kl;
    S.ell_off.assign(S.nlevels + 1, 0);
    for (int l = 0; l < S.nlevels; ++l)
        S.ell_off[l + 1] = S.ell_off[l] + (int64_t)(S.level_ptr[l + 1] - S.level_ptr[l]) * kl[l];
    entries = std::max<int64_t>(S.ell_off[S.nlevels], 1);
    hipError_t e = S.dell_kl.alloc(sizeof(int) * kl.size());
    if (e == hipSuccess) e = S.dell_off.alloc(sizeof(int64_t) * S.ell_off.size());
    if (e == hipSuccess) e = hipMemcpy(S.dell_kl.p, kl.data(), sizeof(int) * kl.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(S.dell_off.p, S.ell_off.data(), sizeof(int64_t) * S.ell_off.size(), hipMemcpyHostToDevice);
    S.dlevel_info.release();
    if (e == hipSuccess && S.ell_off.back() < (int64_t(1) << 31)) {
        std::vector<int4> info(S.nlevels);
        for (int l = 0; l < S.nlevels; ++l)
            info[l] = make_int4(S.level_ptr[l], S.level_ptr[l + 1], kl[l], (int)S.ell_off[l]);
        e = S.dlevel_info.alloc(sizeof(int4) * std::max(S.nlevels, 1));
        if (e == hipSuccess && S.nlevels > 0)
            e = hipMemcpy(S.dlevel_info.p, info.data(), sizeof(int4) * info.size(), hipMemcpyHostToDevice);
    }
    return e;
}

}  // namespace

int build_sor_ell(bsp_ctx* ctx, Sweep& S, const std::vector<int32_t>& rp, const std::vector<int32_t>& ci,
                  int forward) {
    S.ell_k = 0;
    S.ell_uniform = false;
    S.sor_forward = -1;
    S.ell_kl.clear();
    S.ell_off.clear();
    // the triangular solves' rules: many levels, little padding
    if (S.n <= 0 || S.nslots <= 0 || S.bs > kSorEllMaxBs || S.nlevels <= kDirectLevels ||
        (int)rp.size() != S.n + 1 || (int)S.level_of.size() != S.n)
        return BSP_OK;
    std::vector<int> kl(S.nlevels, 0), cnt(S.n, 0);
    int64_t real = 0, padded = 0;
    for (int i = 0; i < S.n; ++i) {
        for (int p = rp[i]; p < rp[i + 1]; ++p) cnt[i] += ci[p] != i;
        kl[S.level_of[i]] = std::max(kl[S.level_of[i]], cnt[i]);
        real += cnt[i];
    }
    for (int i = 0; i < S.n; ++i) padded += kl[S.level_of[i]];
    const int K = *std::max_element(kl.begin(), kl.end());
    if (K == 0 || K > kEllMax || padded * 4 > real * 5) return BSP_OK;
    int64_t entries = 0;
    BSP_HIP(set_ell_levels(S, kl, entries));
    BSP_HIP(S.ecol.alloc(sizeof(int) * (size_t)entries));
    BSP_HIP(S.epos.alloc(sizeof(int) * (size_t)entries));
    for (int l = 0; l < S.nlevels; ++l) {
        const int s0 = S.level_ptr[l], n = S.level_ptr[l + 1] - s0;
        const int64_t total = (int64_t)n * kl[l];
        if (total <= 0) continue;
        hipLaunchKernelGGL(k_sor_ell_gather, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream,
                           (int64_t)n, kl[l], s0, S.ell_off[l], S.order.as<int>(), S.drp, S.dci, S.dpos, forward,
                           S.ecol.as<int>(), S.epos.as<int>());
    }
    BSP_HIP(hipGetLastError());
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    S.ell_k = K;
    S.ell_uniform = std::all_of(kl.begin(), kl.end(), [&](int k) { return k == K; });
    S.sor_forward = forward;
    return BSP_OK;
}

int build_sweep_ell(bsp_ctx* ctx, Sweep& S, const std::vector<int32_t>& rp) {
    S.ell_k = 0;
    S.ell_uniform = false;
    S.ell_kl.clear();
    S.ell_off.clear();
    // a sweep of a few wide levels is bound by bandwidth, not by its load chains: the copy
    // measured no gain there (parity-coloured BF16 ILU0 of the 512^2 DG matrix) and costs memory
    if (S.n <= 0 || S.nslots <= 0 || S.bs > kEllMaxBs || S.nlevels <= kDirectLevels ||
        (int)rp.size() != S.n + 1 || (int)S.level_of.size() != S.n)
        return BSP_OK;
    // each level as wide as its widest row; padding counted over the real rows
    std::vector<int> kl(S.nlevels, 0);
    for (int i = 0; i < S.n; ++i) kl[S.level_of[i]] = std::max(kl[S.level_of[i]], rp[i + 1] - rp[i]);
    int K = 0;
    int64_t padded = 0;
    for (int i = 0; i < S.n; ++i) padded += kl[S.level_of[i]];
    for (int l = 0; l < S.nlevels; ++l) K = std::max(K, kl[l]);
    const int64_t nnz = rp[S.n];
    if (K == 0 || K > kEllMax || padded * 4 > nnz * 5) return BSP_OK;   // padding <= 1/4
    int64_t entries = 0;
    BSP_HIP(set_ell_levels(S, kl, entries));
    const size_t esz = S.bf16 ? sizeof(uint16_t) : S.f32 ? sizeof(float) : sizeof(double);
    BSP_HIP(S.ecol.alloc(sizeof(int) * (size_t)entries));
    BSP_HIP(S.eval.alloc(esz * (size_t)entries * S.bs * S.bs));
    if (S.bf16) BSP_HIP(gather_levels<uint16_t>(ctx, S, S.ecol.as<int>(), S.eval.as<uint16_t>()));
    else if (S.f32) BSP_HIP(gather_levels<float>(ctx, S, S.ecol.as<int>(), S.eval.as<float>()));
    else BSP_HIP(gather_levels<double>(ctx, S, S.ecol.as<int>(), S.eval.as<double>()));
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    S.ell_k = K;
    S.ell_uniform = std::all_of(kl.begin(), kl.end(), [&](int k) { return k == K; });
    return BSP_OK;
}

namespace {

int trsv_levels(bsp_ctx* ctx, Sweep& S, const void* dinv, void* x, uint32_t epoch, int l0, int l1) {
    if (S.n <= 0 || S.nslots <= 0) return BSP_OK;
    SweepArgs a = make_args(ctx, S, epoch);
    a.dinv = dinv;
    a.x = x;
    if (S.bf16) return dispatch<float, bf16_t, false>(ctx, S, a, l0, l1);
    return S.f32 ? dispatch<float, float, false>(ctx, S, a, l0, l1)
                 : dispatch<double, double, false>(ctx, S, a, l0, l1);
}

// The last level of `lower` and the first level of `upper` hold the same rows (rows without
// strict-upper blocks are exactly upper's first level).
bool shared_boundary_level(const Sweep& lower, const Sweep& upper) {
    if (lower.n != upper.n || lower.nlevels < 2 || upper.nlevels < 2) return false;
    if ((int)lower.level_of.size() != lower.n || (int)upper.level_of.size() != upper.n) return false;
    for (int i = 0; i < lower.n; ++i)
        if ((lower.level_of[i] == lower.nlevels - 1) != (upper.level_of[i] == 0)) return false;
    return true;
}

}  // namespace

int launch_trsv(bsp_ctx* ctx, Sweep& S, const void* dinv, void* x, uint32_t epoch) {
    return trsv_levels(ctx, S, dinv, x, epoch, 0, -1);
}

int launch_ilu0_sweeps(bsp_ctx* ctx, Sweep& lower, Sweep& upper, const void* dinv, void* x, uint32_t& epoch) {
    if (lower.fuse_boundary < 0) lower.fuse_boundary = shared_boundary_level(lower, upper) ? 1 : 0;
    const bool direct = ctx->trsv_mode == 0 && (!ctx->sweep_graphs || (lower.nlevels <= kDirectLevels &&
                                                                        upper.nlevels <= kDirectLevels));
    if (lower.fuse_boundary == 1 && direct) {
        // The forward sweep's first level holds the rows without strict-lower blocks, where
        // y = b: nothing to do.  Rows of the shared level: x = D^-1 (b - L y) in the forward
        // sweep; the backward sweep starts at its second level.
        const uint32_t e = ++epoch;
        int rc = trsv_levels(ctx, lower, nullptr, x, e, 1, lower.nlevels - 1);
        if (!rc) rc = trsv_levels(ctx, lower, dinv, x, e, lower.nlevels - 1, lower.nlevels);
        if (!rc) rc = trsv_levels(ctx, upper, dinv, x, ++epoch, 1, upper.nlevels);
        return rc;
    }
    int rc = launch_trsv(ctx, lower, nullptr, x, ++epoch);
    if (rc) return rc;
    return launch_trsv(ctx, upper, dinv, x, ++epoch);
}

int launch_sor_sweep(bsp_ctx* ctx, Sweep& S, int forward, double w, const void* dinv, const void* b, const void* xo,
                     void* xn, uint32_t epoch) {
    if (S.n <= 0 || S.nslots <= 0) return BSP_OK;
    SweepArgs a = make_args(ctx, S, epoch);
    a.dinv = dinv;
    a.x = xn;
    a.b = b;
    a.xo = xo;
    a.w = w;
    a.forward = forward;
    if (!a.epos || forward != S.sor_forward) a.ecol = nullptr;   // the copy encodes one direction
    if (S.bf16) return fail(ctx, BSP_E_UNSUPPORTED, "SOR sweeps have no BF16 form");
    return S.f32 ? dispatch<float, float, true>(ctx, S, a) : dispatch<double, double, true>(ctx, S, a);
}

}  // namespace bsp
