// bsp_api.cpp -- C ABI of libblocksp_amd: contexts, memory, matrices, products, BLAS-1.
#include <cmath>
#include <cstring>

#include "bsp_internal.hpp"

namespace bsp {

int fail(const bsp_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->last_error = msg;
    return code;
}

}  // namespace bsp

using namespace bsp;

namespace bsp {

int ell_width(const bsp_bsr* A) {
    if (A->ell_w < 0) {
        int W = 0;
        for (int i = 0; i < A->n_rows; ++i) W = std::max(W, A->h_rp[i + 1] - A->h_rp[i]);
        A->ell_w = W >= 1 && W <= 8 && (int64_t)W * A->n_rows <= A->nnzb + A->nnzb / 4 ? W : 0;
    }
    return A->ell_w;
}

int ensure_ell_columns(bsp_ctx* ctx, const bsp_bsr* A) {
    if (A->ell_ci.p || ell_width(A) == 0) return BSP_OK;
    BSP_HIP(A->ell_ci.alloc(sizeof(int) * std::max<int64_t>((int64_t)A->n_rows * A->ell_w, 1)));
    return launch_ell_columns(ctx, A->n_rows, A->ell_w, A->rp.as<int>(), A->ci.as<int>(), A->ell_ci.as<int>());
}

int bsmv_matrix(bsp_ctx* ctx, const bsp_bsr* A, const double* x, double* y, int accumulate) {
    if (ctx->bsmv_layout == 2 && bsmv_perm_supported(A->bs) && ell_width(A) > 0) {
        int rc = ensure_ell_columns(ctx, A);
        if (rc) return rc;
        if (!A->valpe.p) {
            BSP_HIP(A->valpe.alloc(sizeof(double) * A->bs * A->bs * std::max<int64_t>((int64_t)A->n_rows * A->ell_w, 1)));
            rc = launch_ell_values_f64(ctx, A->n_rows, A->bs, A->ell_w, A->rp.as<int>(), A->val.as<double>(),
                                       A->valpe.as<double>());
            if (rc) return rc;
        }
        return launch_bsmv_ell(ctx, A->n_rows, A->bs, A->ell_w, A->ell_ci.as<int>(), A->valpe.as<double>(), x, y,
                               accumulate);
    }
    if (ctx->bsmv_layout >= 1 && bsmv_perm_supported(A->bs)) {
        if (!A->valp.p) {
            BSP_HIP(A->valp.alloc(sizeof(double) * std::max<int64_t>(A->nnzb * A->bs * A->bs, 1)));
            int rc = launch_permute_blocks(ctx, A->nnzb, A->bs, A->val.as<double>(), A->valp.as<double>());
            if (rc) return rc;
        }
        return launch_bsmv_perm(ctx, A->n_rows, A->bs, A->rp.as<int>(), A->ci.as<int>(), A->valp.as<double>(), x, y,
                                accumulate);
    }
    return launch_bsmv(ctx, A->n_rows, A->bs, A->rp.as<int>(), A->ci.as<int>(), A->val.as<double>(), x, y,
                       accumulate);
}

}  // namespace bsp

extern "C" {

int bsp_ctx_create(int device, void* stream, bsp_ctx** out) {
    if (!out) return BSP_E_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return BSP_E_HIP;
    if (device < 0 || device >= ndev) return BSP_E_INVALID;
    if (hipSetDevice(device) != hipSuccess) return BSP_E_HIP;
    auto* ctx = new bsp_ctx();
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        ctx->num_cus = prop.multiProcessorCount;
        ctx->lds_per_block = (int64_t)prop.sharedMemPerBlock;
    }
    if (stream) {
        ctx->stream = static_cast<hipStream_t>(stream);
    } else {
        if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
            delete ctx;
            return BSP_E_HIP;
        }
        ctx->own_stream = true;
    }
    bool ok = ctx->red_partials.alloc(sizeof(double) * bsp_ctx::kMaxRedBlocks) == hipSuccess &&
              ctx->red_counter.alloc(sizeof(unsigned) * 64) == hipSuccess &&
              ctx->scalars.alloc(sizeof(double) * bsp_ctx::kScalarSlots) == hipSuccess &&
              ctx->err_flag.alloc(sizeof(uint32_t) * 4) == hipSuccess &&
              hipHostMalloc((void**)&ctx->pinned, sizeof(double) * bsp_ctx::kScalarSlots, hipHostMallocDefault) ==
                  hipSuccess;
    if (ok)
        ok = hipMemsetAsync(ctx->red_counter.p, 0, ctx->red_counter.bytes, ctx->stream) == hipSuccess &&
             hipMemsetAsync(ctx->scalars.p, 0, ctx->scalars.bytes, ctx->stream) == hipSuccess &&
             hipMemsetAsync(ctx->err_flag.p, 0, ctx->err_flag.bytes, ctx->stream) == hipSuccess;
    for (int i = 0; ok && i < 16; ++i) ok = hipEventCreate(&ctx->events[i]) == hipSuccess;
    if (ok) ok = hipStreamSynchronize(ctx->stream) == hipSuccess;
    if (!ok) {
        bsp_ctx_destroy(ctx);
        return BSP_E_HIP;
    }
    if (int rc = bsp::noalloc_scratch_init(ctx)) {
        bsp_ctx_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return BSP_OK;
}

int bsp_ctx_destroy(bsp_ctx* ctx) {
    if (!ctx) return BSP_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->aux_stream) {
        (void)hipStreamSynchronize(ctx->aux_stream);
        (void)hipStreamDestroy(ctx->aux_stream);
        ctx->aux_stream = nullptr;
    }
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& e : ctx->events)
        if (e) (void)hipEventDestroy(e);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    ctx->red_partials.release();
    ctx->red_counter.release();
    ctx->scalars.release();
    ctx->err_flag.release();
    if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return BSP_OK;
}

const char* bsp_last_error(const bsp_ctx* ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

int bsp_ctx_sync(bsp_ctx* ctx) {
    if (!ctx) return BSP_E_INVALID;
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    return BSP_OK;
}

int bsp_ctx_set_option(bsp_ctx* ctx, const char* name, int64_t value) {
    if (!ctx || !name) return BSP_E_INVALID;
    const std::string k(name);
    if (k == "trsv") {
        BSP_CHECK(value == 0 || value == 1, BSP_E_INVALID, "trsv option must be 0 or 1");
        ctx->trsv_mode = (int)value;
    } else if (k == "sweep_wg_per_cu") {
        BSP_CHECK(value >= 1 && value <= 64, BSP_E_INVALID, "sweep_wg_per_cu must be in [1, 64]");
        ctx->sweep_wg_per_cu = (int)value;
    } else if (k == "bsmv_layout") {
        BSP_CHECK(value >= 0 && value <= 2, BSP_E_INVALID, "bsmv_layout must be 0, 1 or 2");
        ctx->bsmv_layout = (int)value;
    } else if (k == "fgmres_ortho") {
        BSP_CHECK(value == 0 || value == 1, BSP_E_INVALID, "fgmres_ortho must be 0 (CGS2) or 1 (MGS)");
        ctx->fgmres_ortho = (int)value;
    } else if (k == "fgmres_basis_f32") {
        BSP_CHECK(value == 0 || value == 1, BSP_E_INVALID, "fgmres_basis_f32 must be 0 or 1");
        ctx->fgmres_basis_f32 = (int)value;

    } else if (k == "cgs_pythagoras") {
        BSP_CHECK(value == 0 || value == 1, BSP_E_INVALID, "cgs_pythagoras must be 0 or 1");
        ctx->cgs_pythagoras = (int)value;
    } else if (k == "bf16_ell") {
        BSP_CHECK(value == 0 || value == 1, BSP_E_INVALID, "bf16_ell must be 0 or 1");
        ctx->bf16_ell = (int)value;
    } else if (k == "sweep_graphs") {
        ctx->sweep_graphs = value != 0;
    } else if (k == "sweep_ell") {
        BSP_CHECK(value == 0 || value == 1, BSP_E_INVALID, "sweep_ell must be 0 or 1");
        ctx->sweep_ell = (int)value;
    } else if (k == "sweep_onewg_max") {
        BSP_CHECK(value >= 0, BSP_E_INVALID, "sweep_onewg_max must be >= 0");
        ctx->sweep_onewg_max = value;
    } else if (k == "mgs_onewg") {
        BSP_CHECK(value == 0 || value == 1, BSP_E_INVALID, "mgs_onewg must be 0 or 1");
        ctx->mgs_onewg = (int)value;
    } else if (k == "sweep_lds") {
        BSP_CHECK(value == 0 || value == 1, BSP_E_INVALID, "sweep_lds must be 0 or 1");
        ctx->sweep_lds = (int)value;
    } else if (k == "sweep_ahead") {
        BSP_CHECK(value == 0 || value == 1, BSP_E_INVALID, "sweep_ahead must be 0 or 1");
        ctx->sweep_ahead = (int)value;
    } else if (k == "batch_ell") {
        BSP_CHECK(value == 0 || value == 1, BSP_E_INVALID, "batch_ell must be 0 or 1");
        ctx->batch_ell = (int)value;
    } else if (k == "bsmm_mfma") {
        BSP_CHECK(value == 0 || value == 1, BSP_E_INVALID, "bsmm_mfma must be 0 or 1");
        ctx->bsmm_mfma = (int)value;
    } else if (k == "deflate_device") {
        BSP_CHECK(value == 0 || value == 1, BSP_E_INVALID, "deflate_device must be 0 or 1");
        ctx->deflate_device = (int)value;
    } else if (k == "factor_device") {
        BSP_CHECK(value == 0 || value == 1, BSP_E_INVALID, "factor_device must be 0 or 1");
        ctx->factor_device = (int)value;
    } else if (k == "verbose") {
        ctx->verbose = (int)std::max<int64_t>(0, value);
    } else if (k == "spin_limit") {
        BSP_CHECK(value > 0, BSP_E_INVALID, "spin_limit must be positive");
        ctx->spin_limit = value;
    } else {
        return fail(ctx, BSP_E_INVALID, "unknown option " + k);
    }
    return BSP_OK;
}

int bsp_malloc(bsp_ctx* ctx, size_t bytes, void** out_dev) {
    if (!ctx || !out_dev) return BSP_E_INVALID;
    *out_dev = nullptr;
    if (bytes == 0) return BSP_OK;
    hipError_t e = hipMalloc(out_dev, bytes);
    if (e != hipSuccess) return fail(ctx, BSP_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return BSP_OK;
}

int bsp_free(bsp_ctx* ctx, void* p) {
    if (!ctx) return BSP_E_INVALID;
    if (p) BSP_HIP(hipFree(p));
    return BSP_OK;
}

int bsp_memcpy_h2d(bsp_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return BSP_E_INVALID;
    if (!bytes) return BSP_OK;
    BSP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    return BSP_OK;
}

int bsp_memcpy_d2h(bsp_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return BSP_E_INVALID;
    if (!bytes) return BSP_OK;
    BSP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    return BSP_OK;
}

int bsp_memcpy_d2d(bsp_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return BSP_E_INVALID;
    if (!bytes) return BSP_OK;
    BSP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return BSP_OK;
}

int bsp_memset_zero(bsp_ctx* ctx, void* dst, size_t bytes) {
    if (!ctx) return BSP_E_INVALID;
    if (!bytes) return BSP_OK;
    BSP_HIP(hipMemsetAsync(dst, 0, bytes, ctx->stream));
    return BSP_OK;
}

int bsp_event_record(bsp_ctx* ctx, int slot) {
    if (!ctx || slot < 0 || slot >= 16) return BSP_E_INVALID;
    BSP_HIP(hipEventRecord(ctx->events[slot], ctx->stream));
    return BSP_OK;
}

int bsp_event_elapsed_ms(bsp_ctx* ctx, int a, int b, float* ms) {
    if (!ctx || a < 0 || a >= 16 || b < 0 || b >= 16 || !ms) return BSP_E_INVALID;
    BSP_HIP(hipEventSynchronize(ctx->events[b]));
    BSP_HIP(hipEventElapsedTime(ms, ctx->events[a], ctx->events[b]));
    return BSP_OK;
}

// ---------------------------------------------------------------------------
// matrices
// ---------------------------------------------------------------------------

int bsp_bsr_create(bsp_ctx* ctx, int n_rows, int n_cols, int bs, const int32_t* rowptr, const int32_t* colidx,
                   const double* vals, bsp_bsr** out) {
    if (!ctx || !out) return BSP_E_INVALID;
    *out = nullptr;
    BSP_CHECK(n_rows >= 0 && n_cols >= 0 && bs >= 1 && bs <= 64, BSP_E_INVALID, "bad matrix dimensions");
    BSP_CHECK(rowptr, BSP_E_INVALID, "rowptr is null");
    BSP_CHECK(rowptr[0] == 0, BSP_E_INVALID, "rowptr[0] must be 0");
    for (int i = 0; i < n_rows; ++i)
        BSP_CHECK(rowptr[i + 1] >= rowptr[i], BSP_E_INVALID, "rowptr must be non-decreasing");
    const int64_t nnzb = rowptr[n_rows];
    BSP_CHECK(nnzb == 0 || (colidx && vals), BSP_E_INVALID, "colidx/vals are null");
    for (int64_t p = 0; p < nnzb; ++p)
        BSP_CHECK(colidx[p] >= 0 && colidx[p] < n_cols, BSP_E_INVALID, "column index out of range");
    // the reference inserts each (row, col) block once (blocksp_containers.hpp:147-164)
    {
        std::vector<int> mark(n_cols, -1);
        for (int i = 0; i < n_rows; ++i)
            for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
                BSP_CHECK(mark[colidx[p]] != i, BSP_E_INVALID, "duplicate block in a row");
                mark[colidx[p]] = i;
            }
    }
    BSP_HIP(hipSetDevice(ctx->device));
    auto A = std::make_unique<bsp_bsr>();
    A->ctx = ctx;
    A->n_rows = n_rows;
    A->n_cols = n_cols;
    A->bs = bs;
    A->nnzb = nnzb;
    const size_t b2 = (size_t)bs * bs;
    A->h_rp.assign(rowptr, rowptr + n_rows + 1);
    A->h_ci.assign(colidx, colidx + nnzb);
    A->h_val.assign(vals, vals + nnzb * b2);
    BSP_HIP(A->rp.alloc(sizeof(int) * (n_rows + 1)));
    BSP_HIP(A->ci.alloc(sizeof(int) * std::max<int64_t>(nnzb, 1)));
    BSP_HIP(A->val.alloc(sizeof(double) * std::max<size_t>(nnzb * b2, 1)));
    BSP_HIP(hipMemcpyAsync(A->rp.p, rowptr, sizeof(int) * (n_rows + 1), hipMemcpyHostToDevice, ctx->stream));
    if (nnzb) {
        BSP_HIP(hipMemcpyAsync(A->ci.p, colidx, sizeof(int) * nnzb, hipMemcpyHostToDevice, ctx->stream));
        BSP_HIP(hipMemcpyAsync(A->val.p, vals, sizeof(double) * nnzb * b2, hipMemcpyHostToDevice, ctx->stream));
    }
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    *out = A.release();
    return BSP_OK;
}

int bsp_bsr_create_complex(bsp_ctx* ctx, int n_rows, int n_cols, int bs, const int32_t* rowptr, const int32_t* colidx,
                           const double* vals, bsp_bsr** out) {
    if (!ctx || !out) return BSP_E_INVALID;
    *out = nullptr;
    BSP_CHECK(n_rows >= 0 && n_cols >= 0 && bs >= 1 && bs <= 32, BSP_E_INVALID,
              "bad complex matrix dimensions (block size 1-32)");
    BSP_CHECK(rowptr && rowptr[0] == 0 && rowptr[n_rows] >= 0, BSP_E_INVALID, "bad rowptr");
    const int64_t nnzb = rowptr[n_rows];
    BSP_CHECK(nnzb == 0 || vals, BSP_E_INVALID, "vals is null");
    // complex entry z_rc = a + ib -> rows 2r, 2r+1 and columns 2c, 2c+1 of [[a, -b], [b, a]]
    const int rb = 2 * bs;
    std::vector<double> re((size_t)nnzb * rb * rb);
    for (int64_t p = 0; p < nnzb; ++p) {
        const double* z = vals + (size_t)p * bs * bs * 2;
        double* R = re.data() + (size_t)p * rb * rb;
        for (int r = 0; r < bs; ++r)
            for (int c = 0; c < bs; ++c) {
                const double a = z[(r * bs + c) * 2], bi = z[(r * bs + c) * 2 + 1];
                R[(2 * r) * rb + 2 * c] = a;
                R[(2 * r) * rb + 2 * c + 1] = -bi;
                R[(2 * r + 1) * rb + 2 * c] = bi;
                R[(2 * r + 1) * rb + 2 * c + 1] = a;
            }
    }
    int rc = bsp_bsr_create(ctx, n_rows, n_cols, rb, rowptr, colidx, re.data(), out);
    if (rc) return rc;
    (*out)->cbs = bs;
    return BSP_OK;
}

int bsp_bsr_complex_bs(const bsp_bsr* A, int* cbs) {
    if (!A || !cbs) return BSP_E_INVALID;
    *cbs = A->cbs;
    return BSP_OK;
}

int bsp_bsr_download_complex(bsp_ctx* ctx, const bsp_bsr* A, int32_t* rowptr, int32_t* colidx, double* vals) {
    if (!ctx || !A) return BSP_E_INVALID;
    BSP_CHECK(A->cbs > 0, BSP_E_INVALID, "bsr_download_complex: not a complex matrix");
    const int bs = A->cbs, rb = A->bs;
    std::vector<double> re(vals ? (size_t)A->nnzb * rb * rb : 0);
    int rc = bsp_bsr_download(ctx, A, rowptr, colidx, vals ? re.data() : nullptr);
    if (rc || !vals) return rc;
    for (int64_t p = 0; p < A->nnzb; ++p) {
        const double* R = re.data() + (size_t)p * rb * rb;
        double* z = vals + (size_t)p * bs * bs * 2;
        for (int r = 0; r < bs; ++r)
            for (int c = 0; c < bs; ++c) {
                z[(r * bs + c) * 2] = R[(2 * r) * rb + 2 * c];
                z[(r * bs + c) * 2 + 1] = R[(2 * r + 1) * rb + 2 * c];
            }
    }
    return BSP_OK;
}

int bsp_bsr_destroy(bsp_bsr* A) {
    delete A;
    return BSP_OK;
}

int bsp_bsr_info(const bsp_bsr* A, int* n_rows, int* n_cols, int* bs, int64_t* nnzb) {
    if (!A) return BSP_E_INVALID;
    if (n_rows) *n_rows = A->n_rows;
    if (n_cols) *n_cols = A->n_cols;
    if (bs) *bs = A->bs;
    if (nnzb) *nnzb = A->nnzb;
    return BSP_OK;
}

int bsp_bsr_download(bsp_ctx* ctx, const bsp_bsr* A, int32_t* rowptr, int32_t* colidx, double* vals) {
    if (!ctx || !A) return BSP_E_INVALID;
    const size_t b2 = (size_t)A->bs * A->bs;
    if (rowptr) BSP_HIP(hipMemcpyAsync(rowptr, A->rp.p, sizeof(int) * (A->n_rows + 1), hipMemcpyDeviceToHost, ctx->stream));
    if (colidx && A->nnzb)
        BSP_HIP(hipMemcpyAsync(colidx, A->ci.p, sizeof(int) * A->nnzb, hipMemcpyDeviceToHost, ctx->stream));
    if (vals && A->nnzb)
        BSP_HIP(hipMemcpyAsync(vals, A->val.p, sizeof(double) * A->nnzb * b2, hipMemcpyDeviceToHost, ctx->stream));
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    return BSP_OK;
}

int bsp_bsr_set_values_dev(bsp_ctx* ctx, bsp_bsr* A, const double* vals_dev) {
    if (!ctx || !A || !vals_dev) return BSP_E_INVALID;
    const size_t bytes = sizeof(double) * A->nnzb * A->bs * A->bs;
    BSP_HIP(hipMemcpyAsync(A->val.p, vals_dev, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    A->h_val.resize(A->nnzb * A->bs * A->bs);
    BSP_HIP(hipMemcpyAsync(A->h_val.data(), vals_dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    A->h_val_stale = false;
    // reduced-precision copies and the transpose are rebuilt from the new values on next use
    A->val32.release();
    A->val16.release();
    A->ci16.release();
    A->valp.release();
    A->valpe.release();
    A->transposed.reset();
    return BSP_OK;
}

// ---------------------------------------------------------------------------
// products
// ---------------------------------------------------------------------------

int bsp_bsmv(bsp_ctx* ctx, const bsp_bsr* A, const double* x, double* y, int accumulate) {
    if (!ctx || !A) return BSP_E_INVALID;
    BSP_CHECK(x && y, BSP_E_INVALID, "null vector");
    BSP_CHECK(x != y, BSP_E_INVALID, "bsmv: x and y must not alias");
    return bsmv_matrix(ctx, A, x, y, accumulate);
}

int bsp_bsmv_dim(bsp_ctx* ctx, const bsp_bsr* A, int D, int dim, const double* x, double* y, int accumulate) {
    if (!ctx || !A) return BSP_E_INVALID;
    BSP_CHECK(x && y && x != y, BSP_E_INVALID, "bad vectors");
    BSP_CHECK(D >= 1 && dim < D, BSP_E_INVALID, "bad dimension arguments");
    return launch_bsmv_dim(ctx, A->n_rows, A->bs, D, dim, A->rp.as<int>(), A->ci.as<int>(), A->val.as<double>(), x,
                           y, accumulate);
}

int bsp_bsmm_rhs(bsp_ctx* ctx, const bsp_bsr* A, int nrhs, const double* X, double* Y) {
    if (!ctx || !A) return BSP_E_INVALID;
    BSP_CHECK(A->cbs == 0, BSP_E_UNSUPPORTED, "the multi-RHS product does not take complex matrices");
    BSP_CHECK(nrhs >= 1 && X && Y && X != Y, BSP_E_INVALID, "bad multi-RHS arguments");
    return launch_bsmm_rhs(ctx, A->n_rows, A->bs, nrhs, A->rp.as<int>(), A->ci.as<int>(), A->val.as<double>(), X, Y);
}

// ---------------------------------------------------------------------------
// BLAS-1
// ---------------------------------------------------------------------------

int bsp_dot(bsp_ctx* ctx, int64_t len, const double* x, const double* y, double* out) {
    if (!ctx || !out) return BSP_E_INVALID;
    return reduce(ctx, 0, len, x, y, SLOT_TMP0, out);
}

int bsp_nrm2(bsp_ctx* ctx, int64_t len, const double* x, double* out) {
    if (!ctx || !out) return BSP_E_INVALID;
    return reduce(ctx, 1, len, x, x, SLOT_TMP0, out);
}

int bsp_infnorm(bsp_ctx* ctx, int64_t len, const double* x, double* out) {
    if (!ctx || !out) return BSP_E_INVALID;
    return reduce(ctx, 2, len, x, x, SLOT_TMP0, out);
}

int bsp_axpy(bsp_ctx* ctx, int64_t len, double a, const double* x, double* y) {
    if (!ctx) return BSP_E_INVALID;
    return launch_axpy(ctx, len, a, x, y);
}

int bsp_scal(bsp_ctx* ctx, int64_t len, double a, double* x) {
    if (!ctx) return BSP_E_INVALID;
    return launch_scal(ctx, len, a, x);
}

}  // extern "C"
