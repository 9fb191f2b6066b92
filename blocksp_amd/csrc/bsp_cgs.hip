// bsp_cgs.hip -- classical Gram-Schmidt with conditional reorthogonalisation (FP64
// arithmetic) for the flexible GMRES of the mixed-precision solver.  The basis is stored in
// FP64 or in FP32 (half the bytes per pass; each q_m is rounded to FP32 once, after it is
// orthogonalised and normalised in FP64).
//
// For w against the orthonormal q_0..q_{m-1} (device pointers, passed by value):
//   A : h1 = Q^T w and ||w||^2                     (reads the m vectors and w once)
//   B : w <- w - Q h1 and ||w||^2                  (reads the m vectors and w, writes w)
//   if ||w_new|| < 0.01 ||w_old|| (cancellation; the Daniel-Gragg-Kaufman-Stewart test with
//   a threshold chosen for FP64) the pass is repeated once (A2, B2) and h = h1 + h2;
//   q_m = w / ||w|| (and, on request, an FP32 copy of q_m for the next preconditioner apply).
// The decision is taken on the device (a flag written after B; A2 and B2 are launched and
// return at once when it is clear), or by the caller, which reads the flag with the H column
// and launches A2 and B2 only when it is set.
// Each pass is followed by a small kernel that adds its per-workgroup partials in a fixed
// order; a step is nine launches and reads 2m + 4 vectors (modified Gram-Schmidt: m + 2
// launches, 4m + 4 vectors).  (Folding those sums into the last workgroup of each pass, with a
// device-scope fence per workgroup, was measured 20% slower per step.)
#include "bsp_internal.hpp"

#include <algorithm>
#include <mutex>
#include <type_traits>

namespace bsp {

namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kGroup = 8;          // dot products per thread in the multi-dot kernel
constexpr int kMaxParts = 512;     // element blocks of a pass
constexpr int kMaxBasis = 64;      // basis vectors (FGMRES restart <= 62)
// Reorthogonalise when ||w_new|| < 0.01 ||w_old||: one classical pass leaves components of
// relative size ~eps ||w_old|| / ||w_new|| along the basis, so below this ratio orthogonality
// would degrade past ~1e-14; above it the second pass only costs bandwidth.
constexpr double kReorth2 = 1e-4;

template <typename QT>
struct QPtrs {
    const QT* q[kMaxBasis];
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Block sums of N values; thread 0 of the block ends with the totals in v.
template <int N>
__device__ __forceinline__ void block_sums(double (&v)[N], double (*sh)[N]) {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < N; ++k) sh[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < N; ++k) v[k] = sh[0][k] + sh[1][k] + sh[2][k] + sh[3][k];
}

// Element groups: a thread handles N consecutive elements per step with 16-byte loads of the
// basis (N = 2 FP64 or 4 FP32 values) and N/2 16-byte loads of w.  Vectors come from hipMalloc
// and the library's pools (256-byte aligned); the len % N tail is done by one thread.
template <typename QT> struct QVec;
template <> struct QVec<double> { static constexpr int N = 2; };
template <> struct QVec<float> { static constexpr int N = 4; };

template <int N>
__device__ __forceinline__ void load_w(const double* w, int64_t i, double (&o)[N]) {
    const double2* p = reinterpret_cast<const double2*>(w) + i * (N / 2);
#pragma unroll
    for (int k = 0; k < N / 2; ++k) {
        const double2 v = p[k];
        o[2 * k] = v.x;
        o[2 * k + 1] = v.y;
    }
}
template <int N>
__device__ __forceinline__ void store_w(double* w, int64_t i, const double (&o)[N]) {
    double2* p = reinterpret_cast<double2*>(w) + i * (N / 2);
#pragma unroll
    for (int k = 0; k < N / 2; ++k) p[k] = make_double2(o[2 * k], o[2 * k + 1]);
}
__device__ __forceinline__ void load_q(const double* q, int64_t i, double (&o)[2]) {
    const double2 v = reinterpret_cast<const double2*>(q)[i];
    o[0] = v.x;
    o[1] = v.y;
}
__device__ __forceinline__ void load_q(const float* q, int64_t i, double (&o)[4]) {
    const float4 v = reinterpret_cast<const float4*>(q)[i];
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
    o[3] = v.w;
}

// Scratch of one step: part[(m + 1) * kMaxParts], part_norm[kMaxParts], st[2m + 2]
// (h1, h2, ||w_old||^2, ||w_new||).
struct Scratch {
    double* part;
    double* part_norm;
    double* st;
};

// Pass A or A2 (gated): part[s][b] = sum over slice b's elements of q_s w (s = 8 g + k for
// vector group g = blockIdx.y) and, for g == 0, part[m][b] = sum of w^2.  (Ordering the
// workgroups so that a slice's groups run back to back on one XCD, sharing w in its L2, was
// measured 5% slower.)
template <typename QT>
__global__ __launch_bounds__(kBlock) void k_cgs_dots(int64_t len, const double* __restrict__ w, QPtrs<QT> Q, int m,
                                                     Scratch sc, int nparts, const int* __restrict__ gate) {
    constexpr int N = QVec<QT>::N;
    if (gate && *gate == 0) return;
    __shared__ double sh[kBlock / kWave][kGroup + 1];
    const int slice = blockIdx.x, g = blockIdx.y;
    const int s0 = g * kGroup;
    const int ns = min(kGroup, m - s0);
    double acc[kGroup + 1];
#pragma unroll
    for (int k = 0; k <= kGroup; ++k) acc[k] = 0.0;
    const QT* q[kGroup];
#pragma unroll
    for (int k = 0; k < kGroup; ++k) q[k] = k < ns ? Q.q[s0 + k] : nullptr;
    const int64_t nvec = len / N;
    const int64_t stride = (int64_t)nparts * kBlock;
    int64_t i = (int64_t)slice * kBlock + threadIdx.x;
    // two element groups per step, all loads first (each sum keeps the one-group order)
    for (; i + stride < nvec; i += 2 * stride) {
        double wa[N], wb[N], qa[kGroup][N], qb[kGroup][N];
        load_w<N>(w, i, wa);
        load_w<N>(w, i + stride, wb);
#pragma unroll
        for (int k = 0; k < kGroup; ++k)
            if (k < ns) {
                load_q(q[k], i, qa[k]);
                load_q(q[k], i + stride, qb[k]);
            }
#pragma unroll
        for (int k = 0; k < kGroup; ++k)
            if (k < ns) {
#pragma unroll
                for (int e = 0; e < N; ++e) acc[k] = fma(qa[k][e], wa[e], acc[k]);
#pragma unroll
                for (int e = 0; e < N; ++e) acc[k] = fma(qb[k][e], wb[e], acc[k]);
            }
        if (g == 0) {
#pragma unroll
            for (int e = 0; e < N; ++e) acc[kGroup] = fma(wa[e], wa[e], acc[kGroup]);
#pragma unroll
            for (int e = 0; e < N; ++e) acc[kGroup] = fma(wb[e], wb[e], acc[kGroup]);
        }
    }
    if (i < nvec) {
        double wi[N];
        load_w<N>(w, i, wi);
#pragma unroll
        for (int k = 0; k < kGroup; ++k)
            if (k < ns) {
                double qi[N];
                load_q(q[k], i, qi);
#pragma unroll
                for (int e = 0; e < N; ++e) acc[k] = fma(qi[e], wi[e], acc[k]);
            }
        if (g == 0)
#pragma unroll
            for (int e = 0; e < N; ++e) acc[kGroup] = fma(wi[e], wi[e], acc[kGroup]);
    }
    if (slice == 0 && threadIdx.x == 0)
        for (int64_t i = nvec * N; i < len; ++i) {
            const double wi = w[i];
#pragma unroll
            for (int k = 0; k < kGroup; ++k)
                if (k < ns) acc[k] = fma((double)q[k][i], wi, acc[k]);
            if (g == 0) acc[kGroup] = fma(wi, wi, acc[kGroup]);
        }
    block_sums<kGroup + 1>(acc, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < kGroup; ++k)
            if (k < ns) sc.part[(size_t)(s0 + k) * nparts + slice] = acc[k];
        if (g == 0) sc.part[(size_t)m * nparts + slice] = acc[kGroup];
    }
}

// Pass B or B2 (gated): w <- w - sum_s h[s] q_s; part_norm[b] = block sum of the new w^2.
template <typename QT>
__global__ __launch_bounds__(kBlock) void k_cgs_update(int64_t len, double* __restrict__ w, QPtrs<QT> Q, int m,
                                                       const double* __restrict__ h, double* __restrict__ part_norm,
                                                       const int* __restrict__ gate) {
    constexpr int N = QVec<QT>::N;
    if (gate && *gate == 0) return;
    __shared__ double sh[kBlock / kWave][1];
    __shared__ const QT* qs[kMaxBasis];
    __shared__ double hs[kMaxBasis];
    for (int s = threadIdx.x; s < m; s += kBlock) {
        qs[s] = Q.q[s];
        hs[s] = h[s];
    }
    __syncthreads();
    double acc[1] = {0.0};
    const int64_t nvec = len / N;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    // two element groups per step (each keeps the one-group order of its updates)
    for (; i + stride < nvec; i += 2 * stride) {
        double wa[N], wb[N];
        load_w<N>(w, i, wa);
        load_w<N>(w, i + stride, wb);
#pragma unroll 4
        for (int s = 0; s < m; ++s) {
            double qa[N], qb[N];
            load_q(qs[s], i, qa);
            load_q(qs[s], i + stride, qb);
#pragma unroll
            for (int e = 0; e < N; ++e) wa[e] = fma(-hs[s], qa[e], wa[e]);
#pragma unroll
            for (int e = 0; e < N; ++e) wb[e] = fma(-hs[s], qb[e], wb[e]);
        }
        store_w<N>(w, i, wa);
        store_w<N>(w, i + stride, wb);
#pragma unroll
        for (int e = 0; e < N; ++e) acc[0] = fma(wa[e], wa[e], acc[0]);
#pragma unroll
        for (int e = 0; e < N; ++e) acc[0] = fma(wb[e], wb[e], acc[0]);
    }
    if (i < nvec) {
        double wi[N];
        load_w<N>(w, i, wi);
        for (int s = 0; s < m; ++s) {
            double qi[N];
            load_q(qs[s], i, qi);
#pragma unroll
            for (int e = 0; e < N; ++e) wi[e] = fma(-hs[s], qi[e], wi[e]);
        }
        store_w<N>(w, i, wi);
#pragma unroll
        for (int e = 0; e < N; ++e) acc[0] = fma(wi[e], wi[e], acc[0]);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (int64_t i = nvec * N; i < len; ++i) {
            double wi = w[i];
            for (int s = 0; s < m; ++s) wi = fma(-hs[s], (double)qs[s][i], wi);
            w[i] = wi;
            acc[0] = fma(wi, wi, acc[0]);
        }
    block_sums<1>(acc, sh);
    if (threadIdx.x == 0) part_norm[blockIdx.x] = acc[0];
}

// Pass B of a first pass whose caller reads the H column on the host, with the normalisation folded
// in: the new norm comes from the dots pass by Pythagoras, ||w - Q h||^2 = ||w||^2 - ||h||^2 (Q
// orthonormal, h = Q^T w), which every workgroup computes from the same values in the same order;
// each element is then updated, stored (the vector a second pass would start from) and scaled:
// q = w / ||w||, z (optional) = FP32 copy of q.  Workgroup 0 writes h, the norm and the
// reorthogonalisation flag to hcol, to the host copy hcol_host and to the scratch.  The formula
// loses relative accuracy ~eps ||w_old||^2 / ||w_new||^2 <= 1e4 eps on steps without a second
// pass (below that ratio the flag asks for the second pass, which recomputes the norm as a sum).
template <typename QT>
__global__ __launch_bounds__(kBlock) void k_cgs_update_normalize(int64_t len, double* __restrict__ w, QPtrs<QT> Q,
                                                                 int m, Scratch sc, QT* __restrict__ q,
                                                                 float* __restrict__ z, double* __restrict__ hcol,
                                                                 double* hcol_host, int* __restrict__ flag_out) {
    constexpr int N = QVec<QT>::N;
    __shared__ const QT* qs[kMaxBasis];
    __shared__ double hs[kMaxBasis];
    __shared__ double s_norm;
    for (int s = threadIdx.x; s < m; s += kBlock) {
        qs[s] = Q.q[s];
        hs[s] = sc.st[s];
    }
    if (threadIdx.x == 0) {
        const double wn2 = sc.st[2 * m];
        double hh = 0.0;
        for (int s = 0; s < m; ++s) hh = fma(sc.st[s], sc.st[s], hh);
        const double v = wn2 - hh;
        const double nrm = sqrt(fabs(v));
        s_norm = nrm;
        if (blockIdx.x == 0) {
            const bool reorth = v < kReorth2 * wn2;
            *flag_out = reorth ? 1 : 0;
            for (int s = 0; s < m; ++s) {
                hcol[s] = sc.st[s];
                hcol_host[s] = sc.st[s];
            }
            hcol[m] = nrm;
            hcol[m + 1] = reorth ? 1.0 : 0.0;
            hcol_host[m] = nrm;
            hcol_host[m + 1] = reorth ? 1.0 : 0.0;
            sc.st[2 * m + 1] = nrm;
            __threadfence_system();
        }
    }
    __syncthreads();
    const double h = s_norm;
    const int64_t nvec = len / N;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    auto out = [&](int64_t e0, const double* wv) {
#pragma unroll
        for (int e = 0; e < N; ++e) {
            const QT v = (QT)(wv[e] / h);
            q[e0 + e] = v;
            if (z) z[e0 + e] = (float)v;
        }
    };
    // two element groups per step (each keeps the one-group order of its updates)
    for (; i + stride < nvec; i += 2 * stride) {
        double wa[N], wb[N];
        load_w<N>(w, i, wa);
        load_w<N>(w, i + stride, wb);
#pragma unroll 4
        for (int s = 0; s < m; ++s) {
            double qa[N], qb[N];
            load_q(qs[s], i, qa);
            load_q(qs[s], i + stride, qb);
#pragma unroll
            for (int e = 0; e < N; ++e) wa[e] = fma(-hs[s], qa[e], wa[e]);
#pragma unroll
            for (int e = 0; e < N; ++e) wb[e] = fma(-hs[s], qb[e], wb[e]);
        }
        store_w<N>(w, i, wa);
        store_w<N>(w, i + stride, wb);
        out(i * N, wa);
        out((i + stride) * N, wb);
    }
    if (i < nvec) {
        double wi[N];
        load_w<N>(w, i, wi);
        for (int s = 0; s < m; ++s) {
            double qi[N];
            load_q(qs[s], i, qi);
#pragma unroll
            for (int e = 0; e < N; ++e) wi[e] = fma(-hs[s], qi[e], wi[e]);
        }
        store_w<N>(w, i, wi);
        out(i * N, wi);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (int64_t e = nvec * N; e < len; ++e) {
            double wi = w[e];
            for (int s = 0; s < m; ++s) wi = fma(-hs[s], (double)qs[s][e], wi);
            w[e] = wi;
            const QT v = (QT)(wi / h);
            q[e] = v;
            if (z) z[e] = (float)v;
        }
}

// Fixed-order sums of the partials.  hcol has m + 2 entries.
//   stage 0 (after A):  h1[s] = sum part[s];  wn2 = sum part[m]
//   stage 1 (after B):  n2 = sum part_norm; reorth = (n2 < 1e-4 wn2); *flag_out = reorth;
//                       hcol[0..m) = h1, hcol[m] = norm, hcol[m + 1] = reorth (as 0.0 / 1.0)
//   stage 2 (after A2): h2[s] = sum part[s]; hcol[s] += h2[s]
//   stage 3 (after B2): hcol[m] = norm = sqrt(sum part_norm)
// Stages 2 and 3 return at once when gate is set and *gate == 0.
__global__ __launch_bounds__(kBlock) void k_cgs_reduce(int stage, int m, Scratch sc, int nparts,
                                                       const int* __restrict__ gate, int* __restrict__ flag_out,
                                                       double* __restrict__ hcol) {
    // st: [0, m) h1, [m, 2m) h2, 2m wn2, 2m + 1 norm
    if ((stage == 2 || stage == 3) && gate && *gate == 0) return;
    __shared__ double sh[kBlock / kWave][1];
    double* st = sc.st;
    const int t = blockIdx.x;   // one block per output
    double acc[1] = {0.0};
    const double* src = (stage == 0 || stage == 2) ? sc.part + (size_t)t * nparts : sc.part_norm;
    for (int b = threadIdx.x; b < nparts; b += kBlock) acc[0] += src[b];
    block_sums<1>(acc, sh);
    if (threadIdx.x != 0) return;
    const double v = acc[0];
    if (stage == 0) {
        if (t < m) st[t] = v;
        else st[2 * m] = v;
    } else if (stage == 1) {
        const bool reorth = v < kReorth2 * st[2 * m];
        *flag_out = reorth ? 1 : 0;
        const double nrm = sqrt(fabs(v));
        for (int s = 0; s < m; ++s) hcol[s] = st[s];
        hcol[m] = nrm;
        hcol[m + 1] = reorth ? 1.0 : 0.0;
        st[2 * m + 1] = nrm;
    } else if (stage == 2) {
        st[m + t] = v;
        hcol[t] += v;
    } else {
        const double nrm = sqrt(fabs(v));
        hcol[m] = nrm;
        st[2 * m + 1] = nrm;
    }
}

// q = w / ||w|| in the basis precision; z (optional) = FP32 copy of q, the input of the next
// step's FP32 preconditioner apply.
// After a first pass whose caller reads the H column on the host: every workgroup adds the
// norm partials in the order of k_cgs_reduce stage 1 (so all get the same bits), workgroup 0
// writes the stage-1 results to hcol and to the host copy hcol_host (pinned, mapped), and all
// scale: q = w / ||w||, z (optional) = FP32 copy of q.
template <typename QT>
__global__ __launch_bounds__(kBlock) void k_cgs_finish_first(int64_t len, const double* __restrict__ w, Scratch sc,
                                                             int nparts, int m, QT* __restrict__ q,
                                                             float* __restrict__ z, double* __restrict__ hcol,
                                                             double* hcol_host, int* __restrict__ flag_out) {
    __shared__ double sh[kBlock / kWave][1];
    __shared__ double s_norm;
    double acc[1] = {0.0};
    for (int b = threadIdx.x; b < nparts; b += kBlock) acc[0] += sc.part_norm[b];
    block_sums<1>(acc, sh);
    if (threadIdx.x == 0) {
        const double v = acc[0];
        const double nrm = sqrt(fabs(v));
        s_norm = nrm;
        if (blockIdx.x == 0) {
            const bool reorth = v < kReorth2 * sc.st[2 * m];
            *flag_out = reorth ? 1 : 0;
            for (int s = 0; s < m; ++s) {
                hcol[s] = sc.st[s];
                hcol_host[s] = sc.st[s];
            }
            hcol[m] = nrm;
            hcol[m + 1] = reorth ? 1.0 : 0.0;
            hcol_host[m] = nrm;
            hcol_host[m + 1] = reorth ? 1.0 : 0.0;
            sc.st[2 * m + 1] = nrm;
            __threadfence_system();
        }
    }
    __syncthreads();
    const double h = s_norm;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < len; i += (int64_t)gridDim.x * kBlock) {
        const QT v = (QT)(w[i] / h);
        q[i] = v;
        if (z) z[i] = (float)v;
    }
}

template <typename QT>
__global__ __launch_bounds__(kBlock) void k_cgs_scale(int64_t len, const double* __restrict__ w,
                                                      const double* __restrict__ st, int m,
                                                      QT* __restrict__ q, float* __restrict__ z) {
    const double h = st[2 * m + 1];
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < len; i += (int64_t)gridDim.x * kBlock) {
        const QT v = (QT)(w[i] / h);
        q[i] = v;
        if (z) z[i] = (float)v;
    }
}

Scratch carve(double* scratch, int m) {
    Scratch sc;
    sc.part = scratch;
    sc.part_norm = sc.part + (size_t)(m + 1) * kMaxParts;
    sc.st = sc.part_norm + kMaxParts;
    return sc;
}

// First pass (A, B) and, when `second` is 1, the second pass (A2, B2) gated on the device flag,
// or, when `second` is 2, the second pass alone and ungated (the caller saw the flag of an
// earlier first pass on this w); then q = w / ||w||.
template <typename QT>
int arnoldi_cgs(const bsp_ctx* ctx, int64_t len, double* w, const QT* const* Q, int m, double* hcol, QT* qnew,
                float* znext, double* scratch, int* flag, int second, double* hcol_host) {
    if (m < 1 || m > kMaxBasis) return fail(ctx, BSP_E_INVALID, "cgs: basis size out of range");
    QPtrs<QT> qp{};
    for (int s = 0; s < m; ++s) qp.q[s] = Q[s];
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>((len + 8 * kBlock - 1) / (8 * kBlock), kMaxParts));
    const Scratch sc = carve(scratch, m);
    const int ng = (m + kGroup - 1) / kGroup;
    if (second != 2) {
        hipLaunchKernelGGL(k_cgs_dots<QT>, dim3(nb, ng), dim3(kBlock), 0, ctx->stream, len, w, qp, m, sc, nb,
                           (const int*)nullptr);
        hipLaunchKernelGGL(k_cgs_reduce, dim3(m + 1), dim3(kBlock), 0, ctx->stream, 0, m, sc, nb, (const int*)nullptr,
                           flag, hcol);
        // (an FP32-stored basis is orthonormal to ~1e-7 only: the Pythagorean norm would be off by
        // that times ||w_old||^2 / ||w_new||^2, so that basis keeps the summed norm)
        if (std::is_same_v<QT, double> && second == 0 && hcol_host && ctx->cgs_pythagoras) {
            hipLaunchKernelGGL(k_cgs_update_normalize<QT>, dim3(nb), dim3(kBlock), 0, ctx->stream, len, w, qp, m, sc,
                               qnew, znext, hcol, hcol_host, flag);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return fail(ctx, BSP_E_HIP, std::string("cgs launch: ") + hipGetErrorString(e));
            return BSP_OK;
        }
        hipLaunchKernelGGL(k_cgs_update<QT>, dim3(nb), dim3(kBlock), 0, ctx->stream, len, w, qp, m, sc.st,
                           sc.part_norm, (const int*)nullptr);
        if (second == 0 && hcol_host) {
            const int gs = (int)std::min<int64_t>((len + kBlock - 1) / kBlock, 1024);
            hipLaunchKernelGGL(k_cgs_finish_first<QT>, dim3(gs), dim3(kBlock), 0, ctx->stream, len, w, sc, nb, m,
                               qnew, znext, hcol, hcol_host, flag);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return fail(ctx, BSP_E_HIP, std::string("cgs launch: ") + hipGetErrorString(e));
            return BSP_OK;
        }
        hipLaunchKernelGGL(k_cgs_reduce, dim3(1), dim3(kBlock), 0, ctx->stream, 1, m, sc, nb, (const int*)nullptr,
                           flag, hcol);
    }
    if (second != 0) {
        const int* gate = second == 1 ? flag : nullptr;
        hipLaunchKernelGGL(k_cgs_dots<QT>, dim3(nb, ng), dim3(kBlock), 0, ctx->stream, len, w, qp, m, sc, nb, gate);
        hipLaunchKernelGGL(k_cgs_reduce, dim3(m), dim3(kBlock), 0, ctx->stream, 2, m, sc, nb, gate, flag, hcol);
        hipLaunchKernelGGL(k_cgs_update<QT>, dim3(nb), dim3(kBlock), 0, ctx->stream, len, w, qp, m, sc.st + m,
                           sc.part_norm, gate);
        hipLaunchKernelGGL(k_cgs_reduce, dim3(1), dim3(kBlock), 0, ctx->stream, 3, m, sc, nb, gate, flag, hcol);
    }
    const int gs = (int)std::min<int64_t>((len + kBlock - 1) / kBlock, 4096);
    hipLaunchKernelGGL(k_cgs_scale<QT>, dim3(gs), dim3(kBlock), 0, ctx->stream, len, w, sc.st, m, qnew, znext);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ctx, BSP_E_HIP, std::string("cgs launch: ") + hipGetErrorString(e));
    return BSP_OK;
}

}  // namespace

size_t cgs_scratch_doubles() { return (size_t)(kMaxBasis + 1) * kMaxParts + kMaxParts + 2 * kMaxBasis + 2; }

namespace {

// Scratch of the allocation-free entry point (bsp_orthogonalize_noalloc): the per-workgroup
// partials and reductions of one Gram-Schmidt step with up to kMaxBasis vectors, then the
// reorthogonalisation flag in the last slot -- 33,922 doubles (265 KiB) on every device the
// library runs on, held until the process ends.  In host code `scratch` names a host-side
// stand-in: kernels get the device address taken from hipGetSymbolAddress per device.
constexpr size_t kNoAllocDoubles = (size_t)(kMaxBasis + 1) * kMaxParts + kMaxParts + 2 * kMaxBasis + 2 + 1;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};   // device address of `scratch` on each device

}  // namespace

// Set-up of the allocation-free entry point, called by bsp_ctx_create with the context's
// device current: records that device's address of `scratch` once.
int noalloc_scratch_init(const bsp_ctx* ctx) {
    BSP_CHECK(ctx->device >= 0 && ctx->device < kMaxDevices, BSP_E_UNSUPPORTED, "device index above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[ctx->device]) return BSP_OK;
    double* ptr = nullptr;
    BSP_HIP(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
    BSP_CHECK(ptr != nullptr, BSP_E_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_dev[ctx->device] = ptr;
    return BSP_OK;
}

// Q: host array of m device pointers; hcol: m + 2 device doubles; scratch: cgs_scratch_doubles()
// doubles; flag: one int.  second: 1 = conditional second pass decided on the device, 0 = first
// pass only (hcol[m + 1] says whether a second pass is needed), 2 = second pass only.
// hcol_host (second == 0, optional): pinned host memory that receives hcol[0..m+1] from the
// kernel itself, complete when the stream reaches the next recorded event.
int launch_arnoldi_cgs(const bsp_ctx* ctx, int64_t len, double* w, const double* const* Q, int m, double* hcol,
                       double* qnew, float* znext, double* scratch, int* flag, int second, double* hcol_host) {
    return arnoldi_cgs<double>(ctx, len, w, Q, m, hcol, qnew, znext, scratch, flag, second, hcol_host);
}

// The same step with the basis stored in FP32 (FP64 arithmetic; q_m is rounded to FP32 once).
int launch_arnoldi_cgs(const bsp_ctx* ctx, int64_t len, double* w, const float* const* Q, int m, double* hcol,
                       float* qnew, float* znext, double* scratch, int* flag, int second, double* hcol_host) {
    return arnoldi_cgs<float>(ctx, len, w, Q, m, hcol, qnew, znext, scratch, flag, second, hcol_host);
}

}  // namespace bsp

using namespace bsp;

extern "C" {

int bsp_orthogonalize(bsp_ctx* ctx, int64_t len, const double* const* Q_dev_ptrs, int m, double* w_dev,
                      double* h_host, double* qnew_dev) {
    if (!ctx || !Q_dev_ptrs || !w_dev || !qnew_dev || len <= 0) return BSP_E_INVALID;
    BSP_CHECK(m >= 1 && m <= 64, BSP_E_INVALID, "orthogonalize: 1 <= m <= 64 vectors");
    BSP_HIP(hipSetDevice(ctx->device));
    DevMem scratch, flag, hcol;
    BSP_HIP(scratch.alloc(sizeof(double) * cgs_scratch_doubles()));
    BSP_HIP(flag.alloc(sizeof(int)));
    BSP_HIP(hcol.alloc(sizeof(double) * (m + 2)));
    int rc = launch_arnoldi_cgs(ctx, len, w_dev, Q_dev_ptrs, m, hcol.as<double>(), qnew_dev, nullptr,
                                scratch.as<double>(), flag.as<int>(), 1, nullptr);
    if (rc) return rc;
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    if (h_host) BSP_HIP(hipMemcpy(h_host, hcol.p, sizeof(double) * (m + 1), hipMemcpyDeviceToHost));
    return BSP_OK;
}

int bsp_orthogonalize_noalloc(bsp_ctx* ctx, int64_t len, const double* const* Q_dev_ptrs, int m, double* w_dev,
                              double* h_dev, double* qnew_dev) {
    if (!ctx || !Q_dev_ptrs || !w_dev || !h_dev || !qnew_dev || len <= 0) return BSP_E_INVALID;
    BSP_CHECK(m >= 1 && m <= kMaxBasis, BSP_E_INVALID,
              "orthogonalize_noalloc: 1 <= m <= 64 vectors (the static scratch holds the partials of 65 sums)");
    double* ptr = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        ptr = ctx->device >= 0 && ctx->device < kMaxDevices ? g_scratch_dev[ctx->device] : nullptr;
    }
    BSP_CHECK(ptr != nullptr, BSP_E_INVALID, "orthogonalize_noalloc: scratch not set up for this device");
    BSP_HIP(hipSetDevice(ctx->device));
    // partials and sums are written before they are read in every call; the flag is written by
    // the first pass's reduction before the gated second pass reads it: nothing to clear
    int* flag = reinterpret_cast<int*>(ptr + cgs_scratch_doubles());
    return launch_arnoldi_cgs(ctx, len, w_dev, Q_dev_ptrs, m, h_dev, qnew_dev, nullptr, ptr, flag, 1, nullptr);
}

}  // extern "C"
