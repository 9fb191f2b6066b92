// bsp_batch.hpp -- launchers of the batched IChol0-pCG (bsp_batch.hip).
#pragma once

#include "bsp_internal.hpp"

namespace bsp {
namespace batch {

// Kernel arguments (all device pointers; vectors stored per group of R right-hand sides,
// interleaved inside the group).
struct Args {
    int nitems = 0;
    const int* item_sys = nullptr;    // [nitems] system of each work item (largest first)
    const int* item_rhs0 = nullptr;   // [nitems] first right-hand side of the item
    const int* sys_row0 = nullptr;    // [nsys + 1] first block row of each system
    const int *rp = nullptr, *ci = nullptr;
    const double* val = nullptr;      // A (block diagonal)
    const int *lrp = nullptr, *lci = nullptr, *urp = nullptr, *uci = nullptr;
    const double *lval = nullptr, *uval = nullptr, *dinvL = nullptr, *dinvU = nullptr;   // IChol0 factor
    const int *frows = nullptr, *flev = nullptr, *flev_base = nullptr;   // forward level sets per system
    const int *brows = nullptr, *blev = nullptr, *blev_base = nullptr;   // backward level sets per system
    // ELL copies of the two sweeps, slot-ordered like frows / brows (ell_k == 0: none, CSR sweeps)
    int ell_k = 0;                    // entries per slot
    const int *fcol = nullptr, *bcol = nullptr;           // [slots * ell_k] block column, -1 padding
    const double *fval = nullptr, *bval = nullptr;        // [slots * ell_k][bs * bs] blocks
    const double *fdinv = nullptr, *bdinv = nullptr;      // [slots][bs * bs] inverse diagonal blocks
    int nrhs = 1;
    int64_t len = 0;                  // rows * bs (length of one right-hand side)
    const double* B = nullptr;
    double *X = nullptr, *Rv = nullptr, *P = nullptr, *Z = nullptr, *AP = nullptr;
    double tol = 0.0;
    int maxIters = 0;
    int mode = 0;                     // 0: ||r|| <= tol ||M^{-1} b||, 1: ||r|| <= tol ||b||
    int* iters = nullptr;             // [nsys * nrhs]
    double* res = nullptr;            // [nsys * nrhs][2]: final ||r||, reference norm
    unsigned* work = nullptr;         // work-item counter (0 at launch)
};

bool supported(int bs, int R);
// ELL width for sweeps of bs x bs blocks whose rows hold at most `max_entries` strict blocks
// (0: keep CSR)
int ell_width(int bs, int max_entries);
int max_resident_blocks(const bsp_ctx* ctx, int bs, int R, int K, int wg);
int launch_pcg(const bsp_ctx* ctx, int bs, int R, int wg, int grid, const Args& a);
// ELL copy of a sweep (slot q = position in `rows`): columns, blocks and inverse diagonal blocks
int launch_ell_gather(const bsp_ctx* ctx, int64_t nslots, int K, int bs, const int* rows, const int* rp, const int* ci,
                      const double* val, const double* dinv, int* ecol, double* eval, double* edinv);
// RHS-major (nrhs vectors of len) <-> groups of R right-hand sides interleaved inside the group
int launch_interleave(const bsp_ctx* ctx, int64_t len, int nrhs, int R, const double* src, double* dst,
                      bool to_interleaved);

}  // namespace batch
}  // namespace bsp
