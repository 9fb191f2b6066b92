// bsp_internal.hpp -- internal types of libblocksp_amd (MI355X / gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <ctime>
#include <memory>
#include <string>
#include <complex>
#include <vector>

#include "../../include/blocksp_c.h"

struct bsp_ctx;

namespace bsp {

// Records `msg` as the context's last error and returns `code`.
int fail(const bsp_ctx* ctx, int code, const std::string& msg);

#define BSP_HIP(call)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return ::bsp::fail(ctx, BSP_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define BSP_CHECK(cond, code, msg)                       \
    do {                                                 \
        if (!(cond)) return ::bsp::fail(ctx, code, msg); \
    } while (0)

// Device allocation owned by a context (freed with hipFree).
struct DevMem {
    void* p = nullptr;
    size_t bytes = 0;
    DevMem() = default;
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    ~DevMem() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t alloc(size_t b) {
        release();
        bytes = b;
        if (b == 0) return hipSuccess;
        return hipMalloc(&p, b);
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// Strictly-triangular (or full) block structure on the device plus the level
// schedule used by the sweep kernels.
struct Sweep {
    int n = 0, bs = 0;
    int64_t nnzb = 0;
    DevMem rp, ci, val;          // owned block rows (original per-row order), if any
    const int* drp = nullptr;    // the structure the kernels read: the owned arrays or
    const int* dci = nullptr;    // (SOR) the matrix's own arrays
    const void* dval = nullptr;
    const int* dpos = nullptr;   // SOR in an elimination order: step of each row (device)
    bool f32 = false;            // values / vectors in FP32 instead of FP64
    bool bf16 = false;           // values (and the diagonal inverses applied with them) stored in
                                 // BF16, vectors and arithmetic in FP32 (f32 is set too)
    unsigned long long tickets_used = 0;   // ticket counter value at the next sync-free launch
    // schedule: rows sorted by level, each level padded with -1 to a multiple of
    // the rows handled by one wavefront, so a wavefront never holds two levels.
    DevMem order;                // int32 [nslots]
    std::vector<int> level_ptr;  // host: slot offsets of each level (padded)
    DevMem dlevel_ptr;           // the same on the device (one-workgroup sweeps)
    std::vector<int> level_of;   // host: level of each row (unpadded)
    int nslots = 0;
    int nlevels = 0;
    int max_level_rows = 0;      // rows of the widest level (unpadded)
    int max_level_slots = 0;     // slots of the widest level (padded)
    int fuse_boundary = -1;      // ILU0 lower sweep: its last level is the upper sweep's first
                                 // (1 yes, 0 no, -1 not yet checked)
    // slot-ordered ELL copy of the blocks (triangular solves; built where padding is small):
    // level l's slots hold ell_kl[l] entries each, slot s's entry k at ell_off[l] +
    // (s - level_ptr[l]) * ell_kl[l] + k: column ecol[.] (-1 padding, and -1 on padded slots),
    // block at eval + (.) * bs * bs in the value type of dval.  ell_k: the widest level (0: no
    // copy); ell_uniform: every level has width ell_k (the single-launch sweeps need that).
    int ell_k = 0;
    bool ell_uniform = false;
    std::vector<int> ell_kl;
    std::vector<int64_t> ell_off;
    DevMem ecol, eval;
    // SOR sweeps: ecol = column on the dependency side, -2 - column on the other (-1 padding),
    // epos = the block's position in the matrix; sor_forward: the direction encoded (-1 none)
    DevMem epos;
    int sor_forward = -1;
    DevMem dell_kl, dell_off;    // ell_kl / ell_off on the device (one-workgroup sweeps)
    DevMem dlevel_info;          // int4 per level: first slot, end slot, ELL width, first entry (when
                                 // the entries fit int; the one-workgroup ELL sweeps, one load a level)
    DevMem flags;                // uint32 [n], epoch-stamped completion flags
    DevMem ticket;               // uint64 workgroup ticket counter
    // level-launch sweeps replayed from captured HIP graphs, keyed by their operands
    struct Graph {
        uint64_t key[8] = {};
        hipGraphExec_t exec = nullptr;
    };
    std::vector<Graph> graphs;
    Sweep() = default;
    Sweep(const Sweep&) = delete;
    Sweep& operator=(const Sweep&) = delete;
    ~Sweep() {
        for (auto& g : graphs)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
    }
};

// Host timing of set-up phases, printed to stderr when the context's verbose option is on.
struct PhaseTimer;

}  // namespace bsp

struct bsp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipStream_t aux_stream = nullptr;   // second stream (default priority) for work off the critical
                                        // path: the flexible GMRES's least-squares updates
    mutable std::string last_error;
    int num_cus = 256;
    int64_t lds_per_block = 0;       // device limit of LDS bytes per workgroup
    // options
    int trsv_mode = 0;               // 0 level launches (default: measured faster), 1 sync-free single launch
    int64_t spin_limit = int64_t(1) << 22;
    int verbose = 0;                 // print solver progress every `verbose` iterations (0 = off)
    int sweep_wg_per_cu = 8;         // persistent grid of the sync-free sweeps, per compute unit
    int sweep_graphs = 1;            // replay many-level sweeps from captured HIP graphs
    int sweep_ell = 1;               // triangular sweeps read their slot-ordered ELL copy where built
    int64_t sweep_onewg_max = int64_t(1) << 22;   // level-launch mode: a sweep of more than 8
                                     // levels whose widest level fits one pass of a 1024-thread
                                     // workgroup (1024 / bs rows) runs in one workgroup when n * bs is
                                     // at most this (0: never)
    int sweep_ahead = 1;             // one-workgroup sweeps over an ELL copy (blocks up to 4 x 4,
                                     // levels within one pass) load each level's static operands
                                     // before the previous level's barrier
    int mgs_onewg = 1;               // Arnoldi steps on vectors of at most 8192 entries: all the
                                     // Gram-Schmidt steps in one single-workgroup kernel
    int sweep_lds = 1;               // such sweeps keep the solution vector in LDS when it fits
                                     // (at most 64 KiB and the device's limit per workgroup)
    int bsmm_mfma = 1;               // bsmm numeric pass for 8 x 8 blocks: 1 on the matrix cores (FP64 MFMA), 0 FMA
    int deflate_device = 0;          // GMRES restarts: 1 solve the harmonic-Ritz eigenproblem in one workgroup on the
                                     // device (~1.6 ms per restart at 30 columns, latency-bound), 0 on the host
                                     // (~0.1 ms); least squares, residual vector and rotations stay on the device
    int factor_device = 1;           // ILU0 / IChol0 factorisation: 1 on the device (blocks up to 8 x 8), 0 the
                                     // host loop (the method of the reference, one row at a time)
    int batch_ell = 1;               // batched pCG: slot-ordered ELL copies of the factor sweeps
                                     // (used when no row holds more than 4 strict blocks)
    int bsmv_layout = 2;             // solver bsmv for even bs <= 8 (and the mixed solver's FP32 product for bs 4,
                                     // 8, 16): 2 an ELL copy of the chunk-major values where rows hold at most 8
                                     // blocks (111-112 us on the 512^2 DG matrix, 6.35 TB/s; else as 1), 1 a
                                     // chunk-major copy read by XCD-aware row ranges with streaming loads (117 us),
                                     // 0 the values as stored (131 us)
    int fgmres_ortho = 0;            // mixed FGMRES orthogonalisation: 0 classical Gram-Schmidt with conditional
                                     // reorthogonalisation, 1 modified Gram-Schmidt (the reference's)
    int cgs_pythagoras = 1;          // classical Gram-Schmidt (FP64 basis, first pass read by the host): 1 the new
                                     // norm as ||w||^2 - ||h||^2, normalised inside the update pass (838-843 ms
                                     // headline), 0 summed after the update in a separate kernel (876-899 ms)
    int bf16_ell = 1;                // BF16 MFMA SpMM: 1 reads an ELL copy of the BF16 values where rows hold at
                                     // most 8 blocks (persistent kernel, 136-140 us on the 64^3 DG matrix with 32
                                     // right-hand sides), 0 a copy in CSR order (178-180 us)
    int fgmres_basis_f32 = 0;        // mixed FGMRES Arnoldi basis storage: 0 FP64, 1 FP32 (FP64 arithmetic,
                                     // classical Gram-Schmidt only)
    // reduction scratch
    bsp::DevMem red_partials;        // double [kMaxRedBlocks]
    bsp::DevMem red_counter;         // uint32 [kRedSlots]
    bsp::DevMem scalars;             // double [kScalarSlots]
    bsp::DevMem err_flag;            // uint32 [1], sync-free wait timeouts
    double* pinned = nullptr;        // host pinned [kScalarSlots]
    // final Hessenberg of the last bsp_pgmres (m x n row-major) for the Ritz-value diagnostics;
    // last_Hc instead of last_H after a complex system
    std::vector<double> last_H;
    std::vector<std::complex<double>> last_Hc;
    int last_Hm = 0, last_Hn = 0;
    bool last_restarted = false;
    // Krylov basis recorder (bsp_ctx_record_basis)
    double* rec_buf = nullptr;
    int rec_cap = 0, rec_every = 1, rec_count = 0;
    int64_t rec_seen = 0, rec_len = -1;
    hipEvent_t events[16] = {};
    static constexpr int kMaxRedBlocks = 4096;
    static constexpr int kScalarSlots = 4096;
};

struct bsp_bsr {
    bsp_ctx* ctx = nullptr;
    int n_rows = 0, n_cols = 0, bs = 0;
    int64_t nnzb = 0;
    bsp::DevMem rp, ci, val;
    int cbs = 0;                       // complex matrix (bsp_bsr_create_complex): its complex block size;
                                       // bs = 2 cbs holds the real-equivalent blocks
    std::vector<int32_t> h_rp, h_ci;   // host copy of the structure
    mutable std::vector<double> h_val; // host copy of the values (factorisations start here)
    mutable bool h_val_stale = false;  // values produced on the device (bsmm): copy back on first host use
    // lazily built transpose (for bsmv_transpose)
    mutable std::unique_ptr<bsp_bsr> transposed;
    // C = A B (or A^T B) from bsp_bsmm: the factors its numeric pass reads, and (8 x 8 blocks)
    // each row's product schedule over the matrix-core streams (mm_soff[i]: row i's first entry)
    const bsp_bsr* mm_A = nullptr;
    bsp::DevMem mm_soff, mm_sched;
    const bsp_bsr* mm_B = nullptr;
    bool mm_Atran = false;             // C = A^T B: the pairs index A's explicit transpose
    mutable bsp::DevMem val32;         // FP32 copy of the values (mixed-precision solves)
    mutable bool val32_chunked = false; // val32 is in the chunk-major order of k_bsmv_lp_perm
    mutable bsp::DevMem val16;         // BF16 copy of the values (MFMA block SpMM)
    mutable bsp::DevMem ci16;          // its column indices when val16 is an ELL copy
    mutable int val16_w = -1;          // ELL width for val16 (0: no ELL copy, -1: not yet decided)
    mutable bool val16_ell = false;
    mutable bsp::DevMem valp;          // chunk-major copy of the values (bsmv variant 3)
    // ELL structure shared by the ELL value copies of the products: W = ell_w slots per row (the
    // widest row, at most 8, padding at most a quarter of the blocks; 0: none; -1: not yet decided),
    // ell_ci[row * W + s] the row's s-th block column or -1
    mutable int ell_w = -1;
    mutable bsp::DevMem ell_ci;
    mutable bsp::DevMem valpe;         // FP64 values in that ELL order, chunk-major blocks (k_bsmv_ell)
    mutable bool val32_ell = false;    // val32 is in that ELL order (chunk-major FP32 blocks)
};

struct bsp_precond {
    bsp_ctx* ctx = nullptr;
    int kind = BSP_PRE_NONE;
    int n = 0, bs = 0;
    const bsp_bsr* A = nullptr;       // SOR / Diag: non-owning (reference keeps a pointer too)
    double w = 1.0;
    int nu = 0, type = 0, inv_kind = 0;
    bsp::Sweep lower, upper;          // ILU0/IChol0 factors (strict parts) or SOR schedules
    bsp::DevMem dinv_lower;           // [n][bs][bs] applied after the forward sweep (IChol0) or none
    bsp::DevMem dinv_upper;           // [n][bs][bs] applied after the backward sweep / SOR / Diag
    bsp::DevMem work0, work1;         // vectors for SOR ping-pong
    uint32_t epoch = 0;
    bool f32 = false;                 // factor, diagonal inverses and applies in FP32
    bool bf16 = false;                // factor and diagonal inverses stored in BF16 (applies in FP32)
    bsp::DevMem work32;               // FP32 staging vector for applies on FP64 vectors
    std::vector<int32_t> elim_order;  // elimination order of the factorisation (ILU0, IChol0)
    bsp::DevMem pos;                  // multicolour SOR: step of each row in the sweep order
    bsp_host_precond_fn host_fn = nullptr;   // BSP_PRE_HOST
    void* host_user = nullptr;
    // the factor in the reference layout (ILU0: A's blocks; IChol0: the blocks (k, j) with j
    // eliminated no later than k, in A's row order): structure on the host, values on the
    // device (fdev, device factorisation) or on the host (f_val, host factorisation)
    std::vector<int32_t> f_rp, f_ci;
    std::vector<double> f_val;
    bsp::DevMem fdev;
};

namespace bsp {

// ---- kernel launchers (bsp_kernels.hip) ----
int launch_bsmv(const bsp_ctx* ctx, int n, int bs, const int* rp, const int* ci, const double* val,
                const double* x, double* y, int accumulate);
int launch_bsmv_dim(const bsp_ctx* ctx, int n, int B, int D, int dim, const int* rp, const int* ci,
                    const double* val, const double* x, double* y, int accumulate);
int launch_bsmm_rhs(const bsp_ctx* ctx, int n, int bs, int nrhs, const int* rp, const int* ci,
                    const double* val, const double* X, double* Y);
int launch_block_diag(const bsp_ctx* ctx, int n, int bs, const double* dinv, double* x);
// y (+)= A x with the context's bsmv variant (the chunk-major copy of A's values is built on
// first use for variants >= 3) -- the product every solver uses (bsp_api.cpp)
int bsmv_matrix(bsp_ctx* ctx, const bsp_bsr* A, const double* x, double* y, int accumulate);
// chunk-major copy of the values for k_bsmv_perm (even block sizes <= 8)
bool bsmv_perm_supported(int bs);
int launch_permute_blocks(const bsp_ctx* ctx, int64_t nnzb, int bs, const double* val, double* valp);
int launch_bsmv_ell(const bsp_ctx* ctx, int n, int bs, int W, const int* cie, const double* vale, const double* x,
                    double* y, int accumulate);
int launch_ell_columns(const bsp_ctx* ctx, int n, int W, const int* rp, const int* ci, int* cie);
int launch_ell_values_f64(const bsp_ctx* ctx, int n, int bs, int W, const int* rp, const double* val, double* vale);
int launch_ell_values_f32(const bsp_ctx* ctx, int n, int bs, int W, const int* rp, const double* val, float* vale);
// ELL width of A (computed once from the host structure) and its device column array
int ell_width(const bsp_bsr* A);
int ensure_ell_columns(bsp_ctx* ctx, const bsp_bsr* A);
int launch_bsmv_f32_ell(const bsp_ctx* ctx, int n, int bs, int W, const int* cie, const float* vale, const float* x,
                        double* y);
int launch_bsmv_perm(const bsp_ctx* ctx, int n, int bs, const int* rp, const int* ci, const double* valp,
                     const double* x, double* y, int accumulate);

// Deterministic reductions; the result lands in device scalar slot `slot`
// (ctx->scalars) and, if host_out != nullptr, is copied back synchronously.
// kind: 0 dot(x,y), 1 sqrt(|dot(x,x)|), 2 max|x|.
int reduce(const bsp_ctx* ctx, int kind, int64_t len, const double* x, const double* y, int slot,
           double* host_out);
int launch_axpy(const bsp_ctx* ctx, int64_t len, double a, const double* x, double* y);
int launch_scal(const bsp_ctx* ctx, int64_t len, double a, double* x);

// Sweeps: trsv computes x_i <- Dinv_i (x_i - sum_j T_ij x_j) over the schedule.
// Vectors and values are double, or float when S.f32 (bsp_sweeps.hip).
int launch_trsv(bsp_ctx* ctx, Sweep& S, const void* dinv, void* x, uint32_t epoch);
// ILU0 apply: forward sweep over `lower`, backward over `upper` with the diagonal inverses;
// the shared boundary level is done once when both sweeps are launched level by level.
int launch_ilu0_sweeps(bsp_ctx* ctx, Sweep& lower, Sweep& upper, const void* dinv, void* x, uint32_t& epoch);
// Slot-ordered ELL copy of a triangular sweep's blocks (after its schedule is built; rp: the
// sweep's host row pointers), each level as wide as its widest row, when the sweep has more
// than 8 levels, blocks are at most 8 x 8, every row has at most 4 blocks and the padding stays
// within a quarter of the blocks; otherwise the sweep keeps reading CSR.
int build_sweep_ell(bsp_ctx* ctx, Sweep& S, const std::vector<int32_t>& rp);
// The SOR form of the copy (after the sweep's schedule and its drp / dci / dpos are set; rp, ci:
// the matrix's host structure): each slot's off-diagonal blocks as columns tagged with the side
// they are read from and positions in the matrix, for sweeps in direction `forward`, under the
// same rules counted over off-diagonal blocks, for blocks up to 4 x 4.
int build_sor_ell(bsp_ctx* ctx, Sweep& S, const std::vector<int32_t>& rp, const std::vector<int32_t>& ci,
                  int forward);
// SOR sweep: xn_i = xo_i + (-w xo_i + w Dinv_i (b_i - sum_{j!=i} A_ij x_j)), x_j from xn on the
// dependency side (forward: j < i) and from xo otherwise.
int launch_sor_sweep(bsp_ctx* ctx, Sweep& S, int forward, double w, const void* dinv, const void* b,
                     const void* xo, void* xn, uint32_t epoch);

// Krylov helpers (all operands device-resident; h values read from device memory).
// mgs step: if qprev: v -= h[hprev_slot] * qprev;  then h[hout_slot] = dot(qnext, v) or
// (qnext == nullptr) sqrt(dot(v, v)).
int launch_mgs_step(const bsp_ctx* ctx, int64_t len, double* v, const double* qprev, const double* h_prev,
                    const double* qnext, double* h_out);
// Whole MGS of one Arnoldi step (bsp_mgs.hip): w against Q[0..j]; writes hcol[0..j+1]
// (device) and qnew = w / hcol[j+1].  parts: 2 x 1024 doubles of scratch.
int launch_arnoldi_mgs(const bsp_ctx* ctx, int64_t len, double* w, const double* const* Q, int j, double* hcol,
                       double* qnew, double* parts);
// Classical Gram-Schmidt with one conditional reorthogonalisation (bsp_cgs.hip): w against
// the m (<= 64) orthonormal device vectors Q[0..m-1] (a host array of device pointers);
// hcol[0..m] (device; m + 2 entries, hcol[m + 1] = 1.0 when the first pass asks for a second),
// qnew = w / hcol[m] and, if znext is not null, znext = FP32 copy of qnew.  second: 1 = the
// second pass is decided on the device, 0 = first pass only, 2 = second pass only (after a first
// pass on the same w whose hcol[m + 1] was 1; hcol still holds that pass's coefficients).
// scratch: cgs_scratch_doubles() device doubles; flag: one device int.  hcol_host (second == 0,
// optional): pinned host memory the kernel writes hcol[0..m+1] to directly.
size_t cgs_scratch_doubles();
// Records this context's device address of the static scratch of bsp_orthogonalize_noalloc
// (bsp_cgs.hip); called once per context at creation.
int noalloc_scratch_init(const bsp_ctx* ctx);
int launch_arnoldi_cgs(const bsp_ctx* ctx, int64_t len, double* w, const double* const* Q, int m, double* hcol,
                       double* qnew, float* znext, double* scratch, int* flag, int second, double* hcol_host);
int launch_arnoldi_cgs(const bsp_ctx* ctx, int64_t len, double* w, const float* const* Q, int m, double* hcol,
                       float* qnew, float* znext, double* scratch, int* flag, int second, double* hcol_host);
// q = v / h[0]
int launch_scale_div(const bsp_ctx* ctx, int64_t len, const double* v, const double* h, double* q);
int launch_scale_div(const bsp_ctx* ctx, int64_t len, const double* v, const double* h, float* q);
// x += sum_i c[i] * Q[i]  (c on device, m vectors)
int launch_combine_add(const bsp_ctx* ctx, int64_t len, int m, const double* const* Q, const double* c,
                       double* x);
// Qout[j] = sum_i Q[i] * P[i*kp + j]  (P on device, m x kp)
int launch_basis_rotate(const bsp_ctx* ctx, int64_t len, int m, int kp, const double* const* Q,
                        const double* P, double* const* Qout);
// r = b - A x  (overwrite)
int launch_residual(const bsp_ctx* ctx, int n, int bs, const int* rp, const int* ci, const double* val,
                    const double* b, const double* x, double* r);
int launch_copy(const bsp_ctx* ctx, int64_t len, const double* x, double* y);
int launch_fill_u32(const bsp_ctx* ctx, int64_t len, uint32_t v, uint32_t* p);
// ---- reduced precision (bsp_mixed.hip) ----
// y (double) = A32 x32 with FP32 accumulation; y (float) variant
int launch_bsmv_f32(const bsp_ctx* ctx, int n, int bs, const int* rp, const int* ci, const float* val, const float* x,
                    double* y);
bool bsmv_f32_chunked_supported(int bs);
int launch_d2f_chunked(const bsp_ctx* ctx, int64_t nnzb, int bs, const double* val, float* valp);
int launch_bsmv_f32_chunked(const bsp_ctx* ctx, int n, int bs, const int* rp, const int* ci, const float* valp,
                            const float* x, double* y);
int launch_bsmv_f32f(const bsp_ctx* ctx, int n, int bs, const int* rp, const int* ci, const float* val, const float* x,
                     float* y);
int launch_d2f(const bsp_ctx* ctx, int64_t n, const double* a, float* b);
int launch_f2d(const bsp_ctx* ctx, int64_t n, const float* a, double* b);
int launch_combine_add_f32(const bsp_ctx* ctx, int64_t len, int m, const float* const* Z, const double* c, double* x);
int launch_basis_rotate_f32(const bsp_ctx* ctx, int64_t len, int m, int kp, const float* const* Z, const double* P,
                            float* const* Zout);
// FP32 copy of a matrix's values (built on first use, on the device)
int ensure_values_f32(bsp_ctx* ctx, const bsp_bsr* A);
// r <- M^{-1} r on an FP32 vector (FP32 preconditioners only)
int precond_apply_f32(bsp_ctx* ctx, const bsp_precond* M, float* r);

// Block ILU(0) factor in A's block layout, eliminating rows colour by colour (natural order
// when colour == nullptr), and the device preconditioner built from such a factor
// (bsp_precond.cpp).  ilu0_build takes LU's storage.
int ilu0_factor_host(bsp_ctx* ctx, const bsp_bsr* A, const int32_t* colour, std::vector<int32_t>& elim,
                     std::vector<int32_t>& pos, std::vector<double>& LU);
// precision: BSP_PREC_F64, BSP_PREC_F32 or BSP_PREC_BF16 (storage of factor and diagonal inverses).
int ilu0_build(bsp_ctx* ctx, const bsp_bsr* A, const std::vector<int32_t>& elim, const std::vector<int32_t>& pos,
               std::vector<double>& LU, int precision, bsp_precond** out);

// Host copy of device-produced values (bsp_bsmm) and the explicit transpose (bsp_spgemm.hip).
int sync_host_values(bsp_ctx* ctx, const bsp_bsr* A);
int ensure_transpose(bsp_ctx* ctx, const bsp_bsr* A);

// ||b - A x|| / bnorm (FP64, on the device) and the sync-free timeout check (bsp_krylov.cpp)
// pGMRES core for complex systems (bsp_zgmres.cpp), the contract of gmres_core
int zgmres_core(bsp_ctx* ctx, const bsp_bsr* A, const bsp_precond* M, const double* b, double* x, double tol,
                bool tol_is_relative, int maxIters, int restart, int k, int& iters_out, int& restarts_out,
                double& res0_out, double& res_out, bool& converged, int& applies, int stag_window,
                bool& stagnated);
int true_relres_ext(bsp_ctx* ctx, const bsp_bsr* A, const double* b, const double* x, double* work, double bnorm,
                    double& out);
int check_sweep_errors_ext(bsp_ctx* ctx);

// pCG updates: x += alpha p; r -= alpha Ap with alpha = num/den from device scalars
int launch_cg_update_xr(const bsp_ctx* ctx, int64_t len, const double* num, const double* den,
                        const double* p, const double* Ap, double* x, double* r);
// p = z + (num/den) p
int launch_cg_update_p(const bsp_ctx* ctx, int64_t len, const double* num, const double* den,
                       const double* z, double* p);

// Least squares and deflated restart of a GMRES solve on the device (bsp_restart.hip): the
// Hessenberg and its Givens-rotated triangle, the rotated right-hand side, the solution y and
// the outputs of a deflated restart (P_{k+1}, P_k, H_new, c_r, kout).
struct DeviceLS {
    int maxm = 0, ld = 0;
    bool restarted = false;     // the current cycle started from a deflated restart
    bool ws_in_lds = false;     // the deflation's workspace fits in one workgroup's LDS
    int host_deflations = 0;    // restarts whose eigenproblem ran on the host
    DevMem H, R, cs, sn, g, Q0t, y, res, meta, resv, tbuf, ws, wsg;
    hipEvent_t ev = nullptr;    // the restart's copies down are complete
    bool defl_device = false;   // the current restart runs the device route
    // restart outputs, one device block: P ((n+1) x (n+1)) | P_k (n x n) | H_new ((n+1) x n) |
    // c_r (n+1) | kout; packed by their actual sizes when the host computes them
    DevMem rbuf;
    double *P = nullptr, *Pk = nullptr, *Hnew = nullptr, *cr = nullptr;
    int* kout = nullptr;
    size_t rbuf_doubles = 0;
    double* hpin = nullptr;     // pinned staging: the cycle's H and residual vector down, the
                                // host route's restart outputs up (rbuf_doubles + (n+1)(n+2))
    DeviceLS() = default;
    DeviceLS(const DeviceLS&) = delete;
    DeviceLS& operator=(const DeviceLS&) = delete;
    ~DeviceLS() {
        if (ev) (void)hipEventDestroy(ev);
        if (hpin) (void)hipHostFree(hpin);
    }
};
// Storage for a Hessenberg of maxm columns (restart > 0: also the restart outputs).
int ls_alloc(bsp_ctx* ctx, DeviceLS& ls, int maxm, int restart);
// A fresh cycle with c = beta e_1 (beta: device scalar).
int ls_fresh(bsp_ctx* ctx, DeviceLS& ls, const double* beta_dev);
// Column j from the Arnoldi column hcol_dev[0..j+1]; out_dev (optional, device address of
// host-visible memory) receives {residual, |h_{j+1,j}|}; skip_dev (optional): no update when
// *skip_dev != 0.  stream: the context's stream when null.
int ls_column(bsp_ctx* ctx, DeviceLS& ls, int j, const double* hcol_dev, double* out_dev, const double* skip_dev,
              hipStream_t stream = nullptr);
// y = R^{-1} g over the first nc columns (ls.y), and with want_resv the residual vector c - H y
// (ls.resv); beta_dev: the fresh cycle's beta (device scalar, read also when restarted).
int ls_solve(bsp_ctx* ctx, DeviceLS& ls, int nc, const double* beta_dev, bool want_resv);
// Deflated restart of the n-column cycle (ls.resv must hold its residual vector): ls.P, ls.Pk,
// ls.Hnew, ls.cr on the device, kout to the host.  The harmonic-Ritz eigenproblem runs on the
// host (default; fast: the inverse-iteration route, else the complex Schur route) or, with the
// context option deflate_device, in one workgroup on the device (the inverse-iteration route;
// the host Schur route where it declines).  _begin enqueues the device part and the copies
// down; the caller may enqueue more work before _end, which waits for the copies only, runs
// the host part and enqueues the copy up.
int ls_deflate_begin(bsp_ctx* ctx, DeviceLS& ls, int n, int k);
int ls_deflate_end(bsp_ctx* ctx, DeviceLS& ls, int n, int k, bool fast, int& kout);
// Start the next cycle from the deflation's outputs.
int ls_restart(bsp_ctx* ctx, DeviceLS& ls, int kout);
// The first nc + 1 rows and nc columns of the Hessenberg, row-major, to the host.
int ls_download_h(bsp_ctx* ctx, const DeviceLS& ls, int nc, std::vector<double>& H);

// Scalar slots used by the solvers.
enum : int {
    SLOT_TMP0 = 0,
    SLOT_TMP1 = 1,
    SLOT_TMP2 = 2,
    SLOT_TMP3 = 3,
    SLOT_BETA = 4,    // beta of the current GMRES cycle (read by the device least squares)
    SLOT_H = 64,      // Hessenberg column scratch (<= 1024 entries)
    SLOT_Y = 1200,    // combination coefficients
    SLOT_P = 1400,    // restart rotation matrix (<= 2048 entries)
};

}  // namespace bsp

namespace bsp {

struct PhaseTimer {
    const bsp_ctx* ctx;
    const char* what;
    double t0;
    static double now() {
        timespec ts;
        clock_gettime(CLOCK_MONOTONIC, &ts);
        return ts.tv_sec + 1e-9 * ts.tv_nsec;
    }
    PhaseTimer(const bsp_ctx* c, const char* w) : ctx(c), what(w), t0(now()) {}
    void mark(const char* phase) {
        if (ctx->verbose <= 0) return;
        const double t = now();
        std::fprintf(stderr, "[setup %s] %s %.2f ms\n", what, phase, (t - t0) * 1e3);
        std::fflush(stderr);
        t0 = t;
    }
};

}  // namespace bsp
