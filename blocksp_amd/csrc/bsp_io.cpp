// bsp_io.cpp -- BSR binary cache files.
//
// Layout (little-endian): 8-byte magic "BSPBSR1\0"; int32 n_rows, n_cols, bs, 0; int64 nnzb;
// int32 rowptr[n_rows + 1]; int32 colidx[nnzb] (the stored per-row order, unchanged);
// float64 values[nnzb * bs * bs] (row-major blocks).  blocksp_amd/bsrio.py reads and writes
// the same layout with numpy.
#include <cstdio>
#include <cstring>
#include <memory>

#include "bsp_internal.hpp"

using namespace bsp;

namespace {

constexpr char kMagic[8] = {'B', 'S', 'P', 'B', 'S', 'R', '1', '\0'};

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

}  // namespace

extern "C" {

int bsp_bsr_save(bsp_ctx* ctx, const bsp_bsr* A, const char* path) {
    if (!ctx || !A || !path) return BSP_E_INVALID;
    int rc = sync_host_values(ctx, A);
    if (rc) return rc;
    File out;
    out.f = std::fopen(path, "wb");
    BSP_CHECK(out.f, BSP_E_INVALID, std::string("cannot open ") + path + " for writing");
    // the fourth header word: 0 for a real matrix, the complex block size for a complex one
    // (whose values are then its real-equivalent blocks)
    const int32_t hdr[4] = {A->n_rows, A->n_cols, A->bs, A->cbs};
    const int64_t nnzb = A->nnzb;
    const size_t nval = (size_t)nnzb * A->bs * A->bs;
    bool ok = std::fwrite(kMagic, 1, 8, out.f) == 8 && std::fwrite(hdr, sizeof(int32_t), 4, out.f) == 4 &&
              std::fwrite(&nnzb, sizeof(int64_t), 1, out.f) == 1 &&
              std::fwrite(A->h_rp.data(), sizeof(int32_t), A->h_rp.size(), out.f) == A->h_rp.size() &&
              std::fwrite(A->h_ci.data(), sizeof(int32_t), A->h_ci.size(), out.f) == A->h_ci.size() &&
              std::fwrite(A->h_val.data(), sizeof(double), nval, out.f) == nval;
    BSP_CHECK(ok, BSP_E_INVALID, std::string("short write to ") + path);
    return BSP_OK;
}

int bsp_bsr_load(bsp_ctx* ctx, const char* path, bsp_bsr** out) {
    if (!ctx || !path || !out) return BSP_E_INVALID;
    *out = nullptr;
    File in;
    in.f = std::fopen(path, "rb");
    BSP_CHECK(in.f, BSP_E_INVALID, std::string("cannot open ") + path);
    char magic[8];
    int32_t hdr[4];
    int64_t nnzb = 0;
    BSP_CHECK(std::fread(magic, 1, 8, in.f) == 8 && std::memcmp(magic, kMagic, 8) == 0, BSP_E_INVALID,
              std::string(path) + " is not a BSR cache file");
    BSP_CHECK(std::fread(hdr, sizeof(int32_t), 4, in.f) == 4 && std::fread(&nnzb, sizeof(int64_t), 1, in.f) == 1,
              BSP_E_INVALID, std::string("truncated header in ") + path);
    const int n = hdr[0], m = hdr[1], bs = hdr[2], cbs = hdr[3];
    BSP_CHECK(n >= 0 && m >= 0 && bs >= 1 && nnzb >= 0 && (cbs == 0 || bs == 2 * cbs), BSP_E_INVALID,
              std::string("bad header in ") + path);
    std::vector<int32_t> rp(n + 1), ci(nnzb);
    std::vector<double> v((size_t)nnzb * bs * bs);
    const bool ok = std::fread(rp.data(), sizeof(int32_t), rp.size(), in.f) == rp.size() &&
                    std::fread(ci.data(), sizeof(int32_t), ci.size(), in.f) == ci.size() &&
                    std::fread(v.data(), sizeof(double), v.size(), in.f) == v.size();
    BSP_CHECK(ok, BSP_E_INVALID, std::string("truncated data in ") + path);
    BSP_CHECK(rp[0] == 0 && rp[n] == nnzb, BSP_E_INVALID, std::string("inconsistent row pointers in ") + path);
    int rc = bsp_bsr_create(ctx, n, m, bs, rp.data(), ci.data(), v.data(), out);
    if (rc == BSP_OK) (*out)->cbs = cbs;
    return rc;
}

}  // extern "C"
