"""Benchmark of the block-sparse solve path on one MI355X.

Workload (N=1): the 2D LDG-style vector Laplacian -- 512x512 quads, Q1 SWIP DG
x 2 components -> BlockSp<double,8>, 262144 block rows, 5 blocks/row,
coefficient lognormal(0,1) per element (seed 1), RHS N(0,1).  Synthetic data.

One step = one complete solve of A x = b to a true FP64 relative residual
||b - Ax|| / ||b|| <= 1e-10 with A, b and the preconditioner resident in HBM:
  * the headline (`value`): mixed-precision FGMRES(30, deflated restart k=10) --
    ILU0 eliminated by parity colours, stored in BF16 and applied in FP32, FP32 SpMV
    inside, FP64 Gram-Schmidt, least squares, deflation, solution update and true
    residual, FP64 refinement (the same solve with the factor in FP32 is reported too);
  * beside it: the plain FP64 ILU0-pGMRES(30, k=10) drop-in (the reference's
    method: natural-order ILU0, left preconditioning), measured once per run.
`value` is the time to solution of one step in seconds (lower is better).
Preconditioner setup is reported separately (`setup_s`).

Also measured in the same run:
  * roofline of the bsmv kernel: algorithmic bytes per launch / average launch time
    (HIP events on the library stream) against 8 TB/s; `traffic` is the measured HBM
    bytes per launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary under profiles/;
  * cpu_baseline: the CPU oracle (this project's restatement of the reference
    algorithm, 1 thread) on a bounded sample, scaled to the same unit;
  * `workloads`: the secondary configurations -- the 3D 64^3 DG Laplacian (BF16 MFMA
    block SpMM with 32 RHS, mixed vs plain FP64 solve with forward errors), 512 batched
    uneven IChol0-pCG systems in one persistent launch, device IChol0-pCG on one large
    system in natural vs parity-coloured order, the learned ILU0 correction on the
    48^3 random-coefficient problem (trained on the GPU in the same run), and the
    64x64 B=2 SOR-pGMRES case beside the oracle on one core.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--skip-cpu] [--skip-plain]
                       [--skip-extras] [--only bf16|batch|batch_solve|learned|learned_solve|pcg|c1|bsmm]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
METRIC = "time to 1e-10 FP64 rel. residual: mixed/learned vs plain FP64 GMRES; bsmv GB/s"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "bsmv_dg2d_veclap_512x512.json")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def bsmv_algorithmic_bytes(n, bs, nnzb, elem=8, accumulate=False):
    """A values + indices + x read once + y written (read too when accumulating)."""
    idx = 4 * nnzb + 4 * (n + 1)
    vec = n * bs * elem * (3 if accumulate else 2)
    return nnzb * bs * bs * elem + idx + vec


def measure_bsmv(bsp, ctx, A, p, reps=50):
    n = p.n * p.bs
    x = ctx.to_device(np.random.default_rng(0).standard_normal(n))
    y = ctx.zeros(n)
    for _ in range(5):
        bsp.bsmv(ctx, A, x, y)
    ctx.sync()
    ctx.record(0)
    for _ in range(reps):
        bsp.bsmv(ctx, A, x, y)
    ctx.record(1)
    ms = ctx.elapsed_ms(0, 1) / reps
    return ms, bsmv_algorithmic_bytes(p.n, p.bs, p.nnzb)


def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count(), len(os.sched_getaffinity(0))


def cpu_baseline(p, gpu_iters, sample_iters=60):
    """Oracle ILU0-pGMRES(30,k=10) built with the reference's release flags (-O3 -funroll-loops
    -fstrict-aliasing -fomit-frame-pointer -ffast-math), pinned to one core: setup + a bounded
    number of iterations on the full matrix, scaled to the plain GPU run's iteration count;
    bsmv is the median of 20 runs after 3 warm-ups."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    model, ncpu, nallowed = cpu_info()
    before = os.sched_getaffinity(0)
    core = min(before)
    os.sched_setaffinity(0, {core})
    po.use_fast_build(True)
    try:
        t0 = time.perf_counter()
        po.ilu0_factor(p)
        t_setup = time.perf_counter() - t0
        t0 = time.perf_counter()
        _, st = po.pgmres(p, p.b, 1e-30, po.PRE_ILU0, maxIters=sample_iters - 1, restart=30, k=10)
        t_run = time.perf_counter() - t0
        per_iter = max(t_run - t_setup, 1e-9) / max(st.iters, 1)   # pgmres rebuilds the factor
        x = np.random.default_rng(0).standard_normal(p.n * p.bs)
        for _ in range(3):
            po.bsmv(p, x)
        ts = []
        for _ in range(20):
            t0 = time.perf_counter()
            po.bsmv(p, x)
            ts.append(time.perf_counter() - t0)
        t_bsmv = float(np.median(ts))
    finally:
        po.use_fast_build(False)
        os.sched_setaffinity(0, before)
    return {
        "value": round(t_setup + per_iter * gpu_iters, 3),
        "unit": "s",
        "cores": 1,
        "kind": "port",
        "sample": f"oracle (reference release flags, pinned to core {core}) ILU0 factor + {st.iters} "
                  f"ILU0-pGMRES(30,k=10) iterations on the full matrix ({t_setup + t_run:.1f} s of CPU work); "
                  f"time/iteration x {gpu_iters} (the plain GPU solve's iterations) + setup",
        "cpu_model": model,
        "nproc": ncpu,
        "cores_allowed": nallowed,
        "setup_s": round(t_setup, 3),
        "s_per_iter": round(per_iter, 4),
        "bsmv_ms": round(t_bsmv * 1e3, 3),
        "bsmv_GBps": round(bsmv_algorithmic_bytes(p.n, p.bs, p.nnzb) / t_bsmv / 1e9, 2),
    }


def timed_ms(ctx, fn, reps):
    fn()
    ctx.sync()
    ctx.record(0)
    for _ in range(reps):
        fn()
    ctx.record(1)
    ctx.sync()
    return ctx.elapsed_ms(0, 1) / reps


def tight_reference(bsp, ctx, A, b, p, colour):
    """FP64 reference solution for forward errors: all-FP64 flexible GMRES (FP64 parity-coloured
    ILU0, FP64 SpMV and bases) to true relres 1e-13, its residual recomputed by the CPU oracle's
    bsmv (parity build)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    M = bsp.ILU0Ex(ctx, A, colour=colour, precision=bsp.PREC_F64)
    xr, st, rc = bsp.FGMRES_mixed(ctx, A, b, 1e-13, M, maxIters=40000, spmv_precision=bsp.PREC_F64)
    xref = xr.numpy()
    bh = b.numpy()
    rel = float(np.linalg.norm(bh - po.bsmv(p, xref)) / np.linalg.norm(bh))
    del M
    return xref, f"all-FP64 coloured FGMRES to true relres {st.true_relres:.1e} (oracle bsmv: {rel:.1e})"


def bench_bsmv_configs(bsp, pr, ctx):
    """The FP64 bsmv on the other named matrices: the 64 x 64 block stencil (B=2), the 64^3 DG
    Laplacian and the 48^3 random-field DG matrix (bs 8), each the median of 3 timings of 20
    launches; algorithmic bytes as for the headline matrix."""
    out = {}
    for name, p in (("stencil2d_b2_64x64", pr.stencil2d_b2(64, 64, seed=0)),
                    ("dg3d_lap_64cubed", pr.dg3d_lap(64, 64, 64, seed=4)),
                    ("dg3d_grf_48cubed", pr.dg3d_grf(48, 48, 48, seed=3))):
        A = bsp.BSR.from_problem(ctx, p)
        ms = sorted(measure_bsmv(bsp, ctx, A, p, reps=20)[0] for _ in range(3))[1]
        nbytes = bsmv_algorithmic_bytes(p.n, p.bs, p.nnzb)
        out[name] = {"n_block_rows": p.n, "bs": p.bs, "nnzb": p.nnzb, "us": round(ms * 1e3, 2),
                     "algorithmic_bytes": nbytes, "GBps": round(nbytes / ms / 1e6, 1),
                     "frac_hbm": round(nbytes / ms / 1e6 / HBM_PEAK_GBPS, 4)}
        log(f"bsmv {name}: {ms * 1e3:.1f} us, {nbytes / ms / 1e6:.0f} GB/s")
        del A
    return out


def bench_bsmm(bsp, ctx, A, p, mfma=1, reps=5):
    """C = A A: both passes on the device; the numeric pass timed alone.  Algorithmic bytes:
    A's values read twice (as left and right factor), C's written once, and the three block
    structures (row pointers, column indices)."""
    ctx.set_option("bsmm_mfma", mfma)
    t0 = time.perf_counter()
    A2 = bsp.bsmm(ctx, A, A)
    ctx.sync()
    t_first = time.perf_counter() - t0
    t_num = timed_ms(ctx, lambda: bsp.bsmm_numeric(ctx, A, A, A2), reps)
    b2 = p.bs * p.bs
    products = int(np.diff(np.asarray(p.rowptr, dtype=np.int64))[np.asarray(p.colidx)].sum())
    alg = (2 * p.nnzb * b2 * 8 + A2.nnzb * b2 * 8 + 2 * (4 * p.nnzb + 4 * (p.n + 1))
           + 4 * A2.nnzb + 4 * (p.n + 1))
    gbps = alg / t_num / 1e6
    out = {"matrix": p.name, "bs": p.bs, "nnzb_A": p.nnzb, "nnzb_C": A2.nnzb, "block_products": products,
           "first_call_s": round(t_first, 3), "matrix_cores": bool(mfma), "numeric_us": round(t_num * 1e3, 1),
           "algorithmic_bytes": alg, "numeric_GBps": round(gbps, 1), "hbm_frac": round(gbps / HBM_PEAK_GBPS, 4),
           "numeric_GFLOPs": round(2.0 * products * p.bs ** 3 / t_num / 1e6, 1)}
    log(f"bsmm(A,A) {p.name}: first call {t_first:.2f}s (both passes on the device), numeric "
        f"{t_num * 1e3:.0f} us, {gbps:.0f} GB/s ({gbps / HBM_PEAK_GBPS:.3f} of HBM)")
    del A2
    ctx.set_option("bsmm_mfma", 1)
    return out


def bench_bsmm_all(bsp, pr, ctx):
    """C = A A on the 2D (512^2 cells) and 3D (64^3 cells) DG matrices, 8 x 8 blocks, on the
    matrix cores and on the FMA row kernel."""
    res = []
    for p in (pr.dg2d_veclap(512, 512, seed=1), pr.dg3d_lap(64, 64, 64, seed=4)):
        A = bsp.BSR.from_problem(ctx, p)
        for mf in (1, 0):
            res.append(bench_bsmm(bsp, ctx, A, p, mfma=mf))
        del A
    return {"bsmm": res}


def bench_bf16(bsp, pr, ctx, solve=True):
    """3D DG Laplacian 64^3 (262144 block rows, bs 8, 7 blocks/row): BF16 MFMA block SpMM
    with 32 N(0,1) right-hand sides against FP64, and the mixed-precision solve against
    the plain FP64 one with forward errors against a tight FP64 solve."""
    p = pr.dg3d_lap(64, 64, 64, seed=4)
    A = bsp.BSR.from_problem(ctx, p)
    nrhs = 32
    X = np.random.default_rng(0).standard_normal((p.n * 8, nrhs))
    Xd = ctx.to_device(X)
    Xb = ctx.empty((p.n * nrhs * 8,), np.uint16)
    bsp.rhs_pack_bf16(ctx, p.n, 8, nrhs, Xd, Xb)
    Yb = ctx.empty((p.n * nrhs * 8,), np.float32)
    Y64 = ctx.empty(X.shape)
    t_bf = timed_ms(ctx, lambda: bsp.bsmm_rhs_bf16(ctx, A, nrhs, Xb, Yb), 20)
    t_64 = timed_ms(ctx, lambda: bsp.bsmm_rhs(ctx, A, nrhs, Xd, Y64), 20)
    Y = ctx.empty(X.shape)
    bsp.rhs_unpack_f32(ctx, p.n, 8, nrhs, Yb, Y)
    a, b64 = Y.numpy(), Y64.numpy()
    nbytes = p.nnzb * 64 * 2 + p.n * 8 * nrhs * (2 + 4) + p.nnzb * 4 + (p.n + 1) * 4
    flops = 2.0 * p.nnzb * 64 * nrhs
    # the same product in FP64 (values, X and Y in FP64, the same index arrays)
    nbytes64 = p.nnzb * 64 * 8 + p.n * 8 * nrhs * (8 + 8) + p.nnzb * 4 + (p.n + 1) * 4
    out = {"workload": "dg3d_lap 64^3 Q1-SWIP-DG (BlockSp<double,8>, 262144 block rows, 7 blocks/row), 32 RHS N(0,1)",
           "bf16_spmm_us": round(t_bf * 1e3, 2), "bf16_spmm_GBps": round(nbytes / t_bf / 1e6, 1),
           "bf16_spmm_frac_hbm": round(nbytes / t_bf / 1e6 / HBM_PEAK_GBPS, 4),
           "bf16_spmm_TFLOPs": round(flops / t_bf / 1e9, 2), "algorithmic_bytes": nbytes,
           "fp64_bsmm_rhs_us": round(t_64 * 1e3, 2), "fp64_bsmm_rhs_GBps": round(nbytes64 / t_64 / 1e6, 1),
           "fp64_bsmm_rhs_frac_hbm": round(nbytes64 / t_64 / 1e6 / HBM_PEAK_GBPS, 4),
           "fp64_algorithmic_bytes": nbytes64, "speedup_vs_fp64_bsmm": round(t_64 / t_bf, 2),
           "bf16_vs_fp64_rel_diff": float(np.linalg.norm(a - b64) / np.linalg.norm(b64))}
    log(f"bf16 MFMA SpMM {t_bf * 1e3:.1f} us ({out['bf16_spmm_GBps']:.0f} GB/s), FP64 {t_64 * 1e3:.1f} us "
        f"({out['fp64_bsmm_rhs_GBps']:.0f} GB/s)")
    if not solve:
        return out
    b = ctx.to_device(p.b)
    col3 = pr.parity_colours((64, 64, 64))
    xref, ref_desc = tight_reference(bsp, ctx, A, b, p, col3)
    fwd = lambda x: float(np.linalg.norm(x.numpy() - xref) / np.linalg.norm(xref))
    M0 = bsp.ILU0(ctx, A)
    x0, st0, _ = bsp.pGMRES(ctx, A, b, 1e-10, M0, maxIters=20000, restart=30, k=10, tol_mode=bsp.TOL_TRUE_RESIDUAL)
    del M0
    M64 = bsp.ILU0Ex(ctx, A, colour=col3, precision=bsp.PREC_F64)
    x3, st3, _ = bsp.FGMRES_mixed(ctx, A, b, 1e-10, M64, spmv_precision=bsp.PREC_F64)
    del M64
    Mc = bsp.ILU0Ex(ctx, A, colour=col3, precision=bsp.PREC_F32)
    x1, st1, _ = bsp.FGMRES_mixed(ctx, A, b, 1e-10, Mc)
    Mh = bsp.ILU0Ex(ctx, A, colour=col3, precision=bsp.PREC_BF16)
    x2, st2, _ = bsp.FGMRES_mixed(ctx, A, b, 1e-10, Mh)
    out.update({"plain_fp64_ms": round(st0.solve_ms, 2), "plain_fp64_iters": st0.iters,
                "plain_fp64_relres": st0.true_relres, "plain_fp64_fwd_err": fwd(x0),
                "fp64_fgmres_coloured_ms": round(st3.solve_ms, 2), "fp64_fgmres_coloured_iters": st3.iters,
                "fp64_fgmres_coloured_relres": st3.true_relres, "fp64_fgmres_coloured_fwd_err": fwd(x3),
                "mixed_f32_ms": round(st1.solve_ms, 2), "mixed_f32_iters": st1.iters,
                "mixed_f32_relres": st1.true_relres, "mixed_f32_fwd_err": fwd(x1),
                "mixed_bf16_factor_ms": round(st2.solve_ms, 2), "mixed_bf16_factor_iters": st2.iters,
                "mixed_bf16_factor_relres": st2.true_relres, "mixed_bf16_factor_fwd_err": fwd(x2),
                "mixed_speedup": round(st0.solve_ms / min(st1.solve_ms, st2.solve_ms), 2),
                "precision_gain_same_ordering": round(st3.solve_ms / min(st1.solve_ms, st2.solve_ms), 2),
                "fwd_err_reference": ref_desc})
    log(f"3D mixed {st1.solve_ms:.0f} ms (FP32 factor), {st2.solve_ms:.0f} ms (BF16 factor) vs all-FP64 coloured "
        f"FGMRES {st3.solve_ms:.0f} ms vs plain FP64 {st0.solve_ms:.0f} ms")
    return out


def bench_batch(bsp, pr, ctx, baseline_every=16):
    """512 independent SPD block Laplacians (B=4, 1e3..5e4 block rows, 4 RHS): IChol0-pCG to
    ||r|| <= 1e-10 ||b|| in one persistent graph-replayed launch with autotuned tiles, and
    the same systems of a subset through one device pCG launch sequence per system."""
    probs = pr.batch_blocklap(nsys=512, lo=1e3, hi=5e4, seed=2, nrhs=4)
    rp, ci, vals, sysp = bsp.block_diagonal(probs)
    n = int(sysp[-1])
    Bh = np.stack([np.concatenate([q.meta["rhs"][k] for q in probs]) for k in range(4)])
    A = bsp.BSR(ctx, n, n, 4, rp, ci, vals)
    del rp, ci, vals
    t0 = time.perf_counter()
    bat = bsp.Batch(ctx, A, sysp, 4)
    setup = time.perf_counter() - t0
    B = ctx.to_device(Bh)
    X = ctx.zeros(Bh.shape)
    wg, r = bat.autotune_solve(B, X, 1e-10, maxIters=5000, tol_mode=bsp.TOL_TRUE_RESIDUAL, probe_iters=8)
    bat.pcg(B, X, 1e-10, maxIters=5000, tol_mode=bsp.TOL_TRUE_RESIDUAL)
    its, rel, ms, rc = bat.pcg(B, X, 1e-10, maxIters=5000, tol_mode=bsp.TOL_TRUE_RESIDUAL)
    Xh = X.numpy()
    # the same solve with the sweeps reading the CSR factor instead of its ELL copies
    ctx.set_option("batch_ell", 0)
    try:
        batl = bsp.Batch(ctx, A, sysp, 4)
    finally:
        ctx.set_option("batch_ell", 1)
    batl.set_tiles(wg, r)
    Xl = ctx.zeros(Bh.shape)
    batl.pcg(B, Xl, 1e-10, maxIters=5000, tol_mode=bsp.TOL_TRUE_RESIDUAL)
    itsl, _, msl, _ = batl.pcg(B, Xl, 1e-10, maxIters=5000, tol_mode=bsp.TOL_TRUE_RESIDUAL)
    csr_same = bool((itsl == its).all() and np.array_equal(Xl.numpy(), Xh))
    log(f"batched 512 systems: ELL sweeps (width {bat.ell_width()}) {ms:.0f} ms, CSR sweeps {msl:.0f} ms, "
        f"bitwise equal: {csr_same}")
    del batl, Xl
    worst = 0.0
    for k in range(4):
        xk = ctx.to_device(Xh[k])
        yk = ctx.zeros(Xh[k].shape)
        bsp.bsmv(ctx, A, xk, yk)
        rk = (Bh[k] - yk.numpy()).reshape(-1, 4)
        bk = Bh[k].reshape(-1, 4)
        rn = np.sqrt(np.add.reduceat((rk ** 2).sum(1), sysp[:-1]))
        bn = np.sqrt(np.add.reduceat((bk ** 2).sum(1), sysp[:-1]))
        worst = max(worst, float((rn / bn).max()))
    sel = list(range(0, 512, baseline_every))
    per_sys = 0.0
    for s in sel:
        q = probs[s]
        Ai = bsp.BSR.from_problem(ctx, q)
        M = bsp.IChol0(ctx, Ai)
        for k in range(4):
            bk = ctx.to_device(q.meta["rhs"][k])
            ctx.sync()
            t0 = time.perf_counter()
            bsp.pCG(ctx, Ai, bk, 1e-10, M, maxIters=5000, tol_mode=bsp.TOL_TRUE_RESIDUAL)
            ctx.sync()
            per_sys += time.perf_counter() - t0
    sub = [probs[s] for s in sel]
    rp2, ci2, v2, sys2 = bsp.block_diagonal(sub)
    A2 = bsp.BSR(ctx, int(sys2[-1]), int(sys2[-1]), 4, rp2, ci2, v2)
    bat2 = bsp.Batch(ctx, A2, sys2, 4)
    bat2.set_tiles(wg, r)
    B2 = ctx.to_device(np.stack([np.concatenate([q.meta["rhs"][k] for q in sub]) for k in range(4)]))
    X2 = ctx.zeros(B2.shape)
    bat2.pcg(B2, X2, 1e-10, maxIters=5000, tol_mode=bsp.TOL_TRUE_RESIDUAL)
    _, _, ms2, _ = bat2.pcg(B2, X2, 1e-10, maxIters=5000, tol_mode=bsp.TOL_TRUE_RESIDUAL)
    out = {"workload": "512 SPD block Laplacians B=4, 1e3..5e4 block rows (log-uniform), 4 RHS, IChol0-pCG to 1e-10",
           "block_rows": n, "batched_s": round(ms / 1e3, 4), "tiles": {"wg": wg, "rhs_per_item": r},
           "sweep_layout": f"ELL width {bat.ell_width()}" if bat.ell_width() else "CSR",
           "csr_sweeps": {"batched_s": round(msl / 1e3, 4), "bitwise_equal": csr_same},
           "setup_s": round(setup, 2), "converged": rc == 0, "iters_max": int(its.max()),
           "iters_mean": round(float(its.mean()), 1), "true_relres_max": worst,
           "subset_systems": len(sel), "subset_per_system_launches_s": round(per_sys, 3),
           "subset_batched_s": round(ms2 / 1e3, 4), "subset_speedup": round(per_sys / (ms2 / 1e3), 1)}
    log(f"batched 512 systems {ms:.0f} ms; subset per-system {per_sys:.2f} s vs batched {ms2:.0f} ms")
    # the same batch with each system's IChol0 eliminated by coordinate-parity colours
    del bat, bat2, A2
    col = np.concatenate([pr.parity_colours(q.meta["dims"]) for q in probs]).astype(np.int32)
    t0 = time.perf_counter()
    batc = bsp.Batch(ctx, A, sysp, 4, colour=col)
    setup_c = time.perf_counter() - t0
    Xc = ctx.zeros(Bh.shape)
    wgc, rcc = batc.autotune_solve(B, Xc, 1e-10, maxIters=5000, tol_mode=bsp.TOL_TRUE_RESIDUAL, probe_iters=8)
    batc.pcg(B, Xc, 1e-10, maxIters=5000, tol_mode=bsp.TOL_TRUE_RESIDUAL)
    itsc, relc, msc, rcode = batc.pcg(B, Xc, 1e-10, maxIters=5000, tol_mode=bsp.TOL_TRUE_RESIDUAL)
    out["parity_coloured"] = {"batched_s": round(msc / 1e3, 4), "tiles": {"wg": wgc, "rhs_per_item": rcc},
                              "setup_s": round(setup_c, 2), "converged": rcode == 0, "iters_max": int(itsc.max()),
                              "iters_mean": round(float(itsc.mean()), 1), "relres_max": float(relc.max()),
                              "speedup_vs_natural": round(ms / msc, 2)}
    log(f"batched 512 systems, parity-coloured IChol0: {msc:.0f} ms (iterations mean {itsc.mean():.0f} "
        f"vs {its.mean():.0f} natural)")
    return out


def bench_batch_solve(bsp, pr, ctx, wg=1024, rhs_per_item=1):
    """The 512-system batch of bench_batch with fixed tiles (the tuned 1024 x 1) and two
    complete IChol0-pCG solves (the first captures the batch's HIP graph), nothing else -- the
    run the batch's FETCH_SIZE / WRITE_SIZE counters are read from."""
    probs = pr.batch_blocklap(nsys=512, lo=1e3, hi=5e4, seed=2, nrhs=4)
    rp, ci, vals, sysp = bsp.block_diagonal(probs)
    n = int(sysp[-1])
    Bh = np.stack([np.concatenate([q.meta["rhs"][k] for q in probs]) for k in range(4)])
    A = bsp.BSR(ctx, n, n, 4, rp, ci, vals)
    del rp, ci, vals
    bat = bsp.Batch(ctx, A, sysp, 4)
    bat.set_tiles(wg, rhs_per_item)
    B = ctx.to_device(Bh)
    X = ctx.zeros(Bh.shape)
    bat.pcg(B, X, 1e-10, maxIters=5000, tol_mode=bsp.TOL_TRUE_RESIDUAL)
    its, rel, ms, rc = bat.pcg(B, X, 1e-10, maxIters=5000, tol_mode=bsp.TOL_TRUE_RESIDUAL)
    log(f"batch solve (tiles {wg} x {rhs_per_item}): {ms:.0f} ms, iterations mean {its.mean():.1f} max {its.max()}")
    return {"batch_solve": {"batched_s": round(ms / 1e3, 4), "tiles": {"wg": wg, "rhs_per_item": rhs_per_item},
                            "iters_mean": round(float(its.mean()), 1), "iters_max": int(its.max()),
                            "converged": rc == 0}}


def bench_c1(bsp, pr, ctx):
    """The reference's CPU-sized case: 64 x 64 block stencil (B=2, 4096 block rows),
    SOR(w=1.2, nu=3)-pGMRES(30, k=10) to tol 1e-10 with the reference's stopping rule, on the GPU
    (exact-order SOR sweeps, and multicolour SOR as the option) beside the oracle on 1 core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    p = pr.stencil2d_b2(64, 64, seed=0)
    A = bsp.BSR.from_problem(ctx, p)
    b = ctx.to_device(p.b)
    out = {"workload": "64x64 block stencil B=2 (4096 block rows), SOR(1.2, nu 3)-pGMRES(30, k 10), tol 1e-10"}
    for label, colour in (("exact_order_sor", None), ("multicolour_sor", pr.parity_colours((64, 64)))):
        M = bsp.SOR(ctx, A, 1.2, 3, 0, colour=colour)
        bsp.pGMRES(ctx, A, b, 1e-10, M, maxIters=5000, restart=30, k=10)
        ctx.sync()
        t0 = time.perf_counter()
        _, st, rc = bsp.pGMRES(ctx, A, b, 1e-10, M, maxIters=5000, restart=30, k=10)
        ctx.sync()
        out[label] = {"ms": round((time.perf_counter() - t0) * 1e3, 2), "iters": st.iters,
                      "true_relres": st.true_relres, "converged": rc == 0}
    t0 = time.perf_counter()
    _, sto = po.pgmres(p, p.b, 1e-10, po.PRE_SOR, maxIters=5000, restart=30, k=10, w=1.2, nu=3)
    out["cpu_oracle_1core"] = {"ms": round((time.perf_counter() - t0) * 1e3, 2), "iters": sto.iters,
                               "true_relres": sto.true_relres}
    log(f"C1 SOR-pGMRES: GPU exact-order {out['exact_order_sor']['ms']} ms ({out['exact_order_sor']['iters']} it), "
        f"multicolour {out['multicolour_sor']['ms']} ms ({out['multicolour_sor']['iters']} it); "
        f"oracle 1 core {out['cpu_oracle_1core']['ms']} ms ({sto.iters} it)")
    return out


def bench_pcg(bsp, pr, ctx):
    """Device pCG with IChol0 to true relres 1e-10 on one large system, the reference's natural
    elimination order against the parity-coloured one: the largest 2D block Laplacian of the
    batch (223 x 224 cells, B=4) and the 3D DG Laplacian 48^3 (bs=8)."""
    out = {}
    big = max(pr.batch_blocklap(nsys=512, lo=1e3, hi=5e4, seed=2, nrhs=4), key=lambda q: q.n)
    cases = [("blocklap2d_223x224_B4", big, big.meta["rhs"][0], big.meta["dims"]),
             ("dg3d_lap_48cubed_bs8", pr.dg3d_lap(48, 48, 48, seed=4), None, (48, 48, 48))]
    for name, p, rhs, dims in cases:
        A = bsp.BSR.from_problem(ctx, p)
        b = ctx.to_device(p.b if rhs is None else rhs)
        res = {}
        for label, colour in (("natural", None), ("parity_coloured", pr.parity_colours(dims))):
            M = bsp.IChol0(ctx, A, colour=colour)
            bsp.pCG(ctx, A, b, 1e-10, M, maxIters=5000, tol_mode=bsp.TOL_TRUE_RESIDUAL)
            _, st, rc = bsp.pCG(ctx, A, b, 1e-10, M, maxIters=5000, tol_mode=bsp.TOL_TRUE_RESIDUAL)
            res[label] = {"ms": round(st.solve_ms, 2), "iters": st.iters, "true_relres": st.true_relres,
                          "converged": rc == 0}
        res["speedup"] = round(res["natural"]["ms"] / res["parity_coloured"]["ms"], 2)
        out[name] = res
        log(f"pCG IChol0 {name}: natural {res['natural']['ms']:.1f} ms ({res['natural']['iters']} it), "
            f"parity-coloured {res['parity_coloured']['ms']:.1f} ms ({res['parity_coloured']['iters']} it)")
    return out


def bench_learned(bsp, pr, ctx, train_only=False):
    """3D random-coefficient diffusion 48^3 (seed 3): the learned ILU0 correction trained on
    the GPU from Krylov vectors of FP64 ILU0-pGMRES solves of 16 seeded 24^3 problems
    (2000 pairs), then FGMRES(30) with it (ILU0 fallback armed) against the same FGMRES
    without the network and against plain FP64 ILU0-pGMRES."""
    from blocksp_amd import learned
    tr = [pr.dg3d_grf(24, 24, 24, seed=10 + s) for s in range(16)]
    t0 = time.perf_counter()
    data = learned.training_set(ctx, tr, count=250)
    ctx.sync()
    t_data = time.perf_counter() - t0
    t0 = time.perf_counter()
    net, losses = learned.train(ctx, data, epochs=20, batch=16, lr=1e-3)
    ctx.sync()
    t_train = time.perf_counter() - t0
    out = {"workload": "dg3d_grf 48^3 (110592 block rows, bs 8), lognormal coefficient, correlation 4 cells, seed 3",
           "pairs": int(sum(d[4] for d in data)), "data_s": round(t_data, 2), "train_s": round(t_train, 2),
           "loss_first_epoch": round(float(losses[0]), 5), "loss_last_epoch": round(float(losses[-1]), 5)}
    if train_only:
        return out
    p = pr.dg3d_grf(48, 48, 48, seed=3)
    A = bsp.BSR.from_problem(ctx, p)
    b = ctx.to_device(p.b)
    M0 = bsp.ILU0(ctx, A)
    _, stp, _ = bsp.pGMRES(ctx, A, b, 1e-10, M0, maxIters=20000, restart=30, k=10, tol_mode=bsp.TOL_TRUE_RESIDUAL)
    t0 = time.perf_counter()
    Ml = bsp.LearnedILU0(ctx, A, net, precision=bsp.PREC_F32)
    ctx.sync()
    t_setup = time.perf_counter() - t0
    Mn = bsp.ILU0Ex(ctx, A, precision=bsp.PREC_F32)
    _, stl, _ = bsp.FGMRES_mixed(ctx, A, b, 1e-10, Ml, fallback=Mn)
    _, stn, _ = bsp.FGMRES_mixed(ctx, A, b, 1e-10, Mn)
    c48 = pr.parity_colours((48, 48, 48))
    Mc = bsp.ILU0Ex(ctx, A, colour=c48, precision=bsp.PREC_F32)
    bsp.FGMRES_mixed(ctx, A, b, 1e-10, Mc)                      # warm-up (graph captures, buffers)
    _, stc, _ = bsp.FGMRES_mixed(ctx, A, b, 1e-10, Mc)
    # the correction on the parity-coloured factor, trained on the coloured solver's own Krylov
    # vectors: the elimination order the fast mixed solves use
    t0 = time.perf_counter()
    data_c = learned.training_set(ctx, tr, count=250, colour_fn=lambda q: pr.parity_colours(q.meta["dims"]),
                                  krylov_colour=True)
    net_c, losses_c = learned.train(ctx, data_c, epochs=20, batch=16, lr=1e-3)
    ctx.sync()
    t_train_c = time.perf_counter() - t0
    del data_c
    Mlc = bsp.LearnedILU0(ctx, A, net_c, colour=c48, precision=bsp.PREC_F32)
    bsp.FGMRES_mixed(ctx, A, b, 1e-10, Mlc, fallback=Mc)
    _, stlc, _ = bsp.FGMRES_mixed(ctx, A, b, 1e-10, Mlc, fallback=Mc)
    out.update({"plain_fp64": {"ms": round(stp.solve_ms, 1), "iters": stp.iters, "relres": stp.true_relres},
                "learned_fgmres": {"ms": round(stl.solve_ms, 1), "iters": stl.iters, "relres": stl.true_relres,
                                   "fallback": stl.fallback, "setup_s": round(t_setup, 3)},
                "same_fgmres_without_net": {"ms": round(stn.solve_ms, 1), "iters": stn.iters,
                                            "relres": stn.true_relres},
                "coloured_fgmres_without_net": {"ms": round(stc.solve_ms, 1), "iters": stc.iters,
                                                "relres": stc.true_relres},
                "learned_coloured_fgmres": {"ms": round(stlc.solve_ms, 1), "iters": stlc.iters,
                                            "relres": stlc.true_relres, "fallback": stlc.fallback,
                                            "train_s": round(t_train_c, 2),
                                            "loss_first_epoch": round(float(losses_c[0]), 5),
                                            "loss_last_epoch": round(float(losses_c[-1]), 5),
                                            "iteration_reduction": round(1 - stlc.iters / stc.iters, 3)},
                "iteration_reduction": round(1 - stl.iters / stn.iters, 3),
                "learned_speedup_vs_plain_fp64": round(stp.solve_ms / stl.solve_ms, 2)})
    log(f"learned FGMRES {stl.iters} its {stl.solve_ms:.0f} ms vs without net {stn.iters} its; "
        f"plain FP64 {stp.iters} its {stp.solve_ms:.0f} ms; coloured factor: learned {stlc.iters} its "
        f"{stlc.solve_ms:.0f} ms vs {stc.iters} its {stc.solve_ms:.0f} ms")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--bsmv-only", action="store_true", help="only the bsmv roofline (for profiling)")
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--verbose", type=int, default=0)
    ap.add_argument("--profile-iters", type=int, default=0, help="run N iterations of each solver and exit")
    ap.add_argument("--profile-mixed-only", action="store_true",
                    help="with --profile-iters: only the mixed solver (for counter runs)")
    ap.add_argument("--skip-extras", action="store_true", help="skip the 3D BF16, batched and learned workloads")
    ap.add_argument("--only", choices=["bf16", "batch", "batch_solve", "learned", "learned_solve", "pcg", "c1", "bsmm"], default=None,
                    help="run one secondary workload alone (for profiling) and print its JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the solve path is built and "
              f"measured on one GPU; its multi-GPU sharding is a design note, not a measured configuration.",
              file=sys.stderr)
        sys.exit(2)

    import blocksp_amd as bsp
    from blocksp_amd import problems as pr

    if args.only:
        ctx = bsp.Context(0)
        # profiling runs launch sweeps level by level: rocprofv3's kernel tracing has crashed
        # (host SIGSEGV inside the tool) while a sweep graph was being captured
        ctx.set_option("sweep_graphs", 0)
        if args.verbose:
            ctx.set_option("verbose", args.verbose)
        fn = {"bf16": lambda: bench_bf16(bsp, pr, ctx, solve=False),
              "batch": lambda: bench_batch(bsp, pr, ctx, baseline_every=512),
              "learned": lambda: bench_learned(bsp, pr, ctx, train_only=True),
              "learned_solve": lambda: bench_learned(bsp, pr, ctx),
              "pcg": lambda: bench_pcg(bsp, pr, ctx),
              "c1": lambda: bench_c1(bsp, pr, ctx),
              "bsmm": lambda: bench_bsmm_all(bsp, pr, ctx),
              "batch_solve": lambda: bench_batch_solve(bsp, pr, ctx)}[args.only]
        print(json.dumps(fn()))
        return

    t0 = time.perf_counter()
    p = pr.dg2d_veclap(args.nx, args.nx, seed=1)
    colour = pr.parity_colours((args.nx, args.nx))
    log(f"generated {p.name}: n={p.n} bs={p.bs} nnzb={p.nnzb} in {time.perf_counter() - t0:.1f}s")
    ctx = bsp.Context(0)
    # host handoff of the drop-in path: the matrix from host arrays (structure checks, host copy, upload)
    # and the right-hand side up, then (after the solves) a solution vector down
    ctx.sync()
    t0 = time.perf_counter()
    A = bsp.BSR.from_problem(ctx, p)
    ctx.sync()
    t_bsr = time.perf_counter() - t0
    t0 = time.perf_counter()
    b = ctx.to_device(p.b)
    ctx.sync()
    t_b = time.perf_counter() - t0
    mat_bytes = p.nnzb * p.bs * p.bs * 8 + 4 * p.nnzb + 4 * (p.n + 1)

    bsmv_ms, bsmv_bytes = measure_bsmv(bsp, ctx, A, p)
    bsmv_gbps = bsmv_bytes / (bsmv_ms * 1e-3) / 1e9
    log(f"bsmv {bsmv_ms * 1e3:.1f} us/launch, {bsmv_gbps:.0f} GB/s")
    traffic = None
    if os.path.exists(TRAFFIC_FILE) and args.nx == 512:
        traffic = json.load(open(TRAFFIC_FILE)).get("traffic_bytes_per_launch")
    roofline = {"kernel": "k_bsmv_ell<8> (ELL copy of chunk-major values, XCD-aware rows, streaming loads)", "bound": "hbm", "achieved": round(bsmv_gbps, 1), "peak": HBM_PEAK_GBPS,
                "unit": "GB/s", "frac": round(bsmv_gbps / HBM_PEAK_GBPS, 4), "traffic": traffic,
                "algorithmic_bytes": bsmv_bytes, "avg_us": round(bsmv_ms * 1e3, 2)}
    bsmm = bench_bsmm(bsp, ctx, A, p)
    if args.bsmv_only:
        # calibration kernel of known traffic for the FETCH_SIZE / WRITE_SIZE counters:
        # axpy on 2^21 doubles reads 2 x 16 MiB and writes 16 MiB per launch
        import ctypes
        n = 1 << 21
        xa, ya = ctx.to_device(np.ones(n)), ctx.to_device(np.ones(n))
        for _ in range(20):
            ctx.check(ctx.L.bsp_axpy(ctx.h, n, ctypes.c_double(0.5), xa.ptr, ya.ptr), "axpy")
        ctx.sync()
        print(json.dumps({"bsmv_us": bsmv_ms * 1e3, "bsmv_GBps": bsmv_gbps, "algorithmic_bytes": bsmv_bytes}))
        return
    if args.verbose:
        ctx.set_option("verbose", args.verbose)
    if args.profile_iters:
        # a short run of each solver for kernel traces (not a benchmark step); sweeps are
        # launched level by level (no graph replay) so every kernel is traced directly
        ctx.set_option("sweep_graphs", 0)
        Mh = bsp.ILU0Ex(ctx, A, colour, bsp.PREC_BF16)
        _, st, _ = bsp.FGMRES_mixed(ctx, A, b, 1e-10, Mh, maxIters=args.profile_iters, restart=30, k=10)
        log(f"mixed: {st.iters} iterations in {st.solve_ms:.1f} ms")
        if args.profile_mixed_only:
            return
        M = bsp.ILU0(ctx, A)
        for mode in (0, 1):
            ctx.set_option("trsv", mode)
            _, st, _ = bsp.pGMRES(ctx, A, b, 1e-10, M, maxIters=args.profile_iters, restart=30, k=10)
            log(f"plain trsv={mode}: {st.iters} iterations in {st.solve_ms:.1f} ms")
        return

    # ---- mixed precision (headline) ----
    t0 = time.perf_counter()
    Mh = bsp.ILU0Ex(ctx, A, colour, bsp.PREC_BF16)
    ctx.sync()
    setup_mixed = time.perf_counter() - t0
    log(f"BF16 parity-coloured ILU0 setup {setup_mixed:.2f}s")

    _x_last = [None]

    def mixed_step(M=Mh):
        x, st, rc = bsp.FGMRES_mixed(ctx, A, b, 1e-10, M, maxIters=20000, restart=30, k=10,
                                     spmv_precision=bsp.PREC_F32, inner_reduction=1e-6)
        _x_last[0] = x
        return x, st, rc

    for w in range(args.warmup):
        _, st, rc = mixed_step()
        log(f"mixed warmup {w}: iters={st.iters} relres={st.true_relres:.2e} {st.solve_ms:.1f} ms")
    ctx.sync()
    t0 = time.perf_counter()
    stats = []
    for s in range(args.steps):
        _, st, rc = mixed_step()
        stats.append(st.as_dict())
        log(f"mixed step {s}: iters={st.iters} relres={st.true_relres:.2e} {st.solve_ms:.1f} ms")
    ctx.sync()
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3
    ok = all(s["true_relres"] <= 1e-10 for s in stats)
    x_head = _x_last[0]
    # forward error of the headline solution against a tight FP64 reference solve
    xref, ref_desc = tight_reference(bsp, ctx, A, b, p, colour)
    fwd_err = float(np.linalg.norm(x_head.numpy() - xref) / np.linalg.norm(xref))
    log(f"headline forward error {fwd_err:.2e} ({ref_desc})")
    # the same ordering in all-FP64 (FP64 factor, SpMV and bases): the precision gain alone
    M64 = bsp.ILU0Ex(ctx, A, colour, bsp.PREC_F64)
    x64, st64, _ = bsp.FGMRES_mixed(ctx, A, b, 1e-10, M64, maxIters=20000, restart=30, k=10,
                                    spmv_precision=bsp.PREC_F64)
    fp64_coloured = {"time_s": round(st64.solve_ms / 1e3, 4), "iters": st64.iters, "true_relres": st64.true_relres,
                     "fwd_err": float(np.linalg.norm(x64.numpy() - xref) / np.linalg.norm(xref))}
    log(f"all-FP64 coloured FGMRES: iters={st64.iters} {st64.solve_ms:.1f} ms")
    del M64, x64
    # the same solve with the factor stored in FP32 (one run, for comparison)
    M32 = bsp.ILU0Ex(ctx, A, colour, bsp.PREC_F32)
    _, st32, _ = mixed_step(M32)
    f32_factor = {"time_s": round(st32.solve_ms / 1e3, 4), "iters": st32.iters, "true_relres": st32.true_relres}
    log(f"mixed with FP32 factor: iters={st32.iters} {st32.solve_ms:.1f} ms")
    del M32

    # ---- plain FP64 drop-in (the reference's method) ----
    plain = None
    if not args.skip_plain:
        t0 = time.perf_counter()
        M = bsp.ILU0(ctx, A)
        ctx.sync()
        setup_plain = time.perf_counter() - t0
        apply_ms = {}
        r = ctx.to_device(np.random.default_rng(1).standard_normal(p.n * p.bs))
        for mode in (1, 0):
            ctx.set_option("trsv", mode)
            M(r)
            ctx.sync()
            ctx.record(2)
            for _ in range(3):
                M(r)
            ctx.record(3)
            apply_ms[mode] = ctx.elapsed_ms(2, 3) / 3
            log(f"natural ILU0 apply trsv_mode={mode}: {apply_ms[mode]:.3f} ms")
        ctx.set_option("trsv", 1 if apply_ms[1] <= apply_ms[0] else 0)
        t0 = time.perf_counter()
        _, stp, rcp = bsp.pGMRES(ctx, A, b, 1e-10, M, maxIters=20000, restart=30, k=10,
                                 tol_mode=bsp.TOL_TRUE_RESIDUAL)
        ctx.sync()
        t_plain = time.perf_counter() - t0
        log(f"plain FP64 ILU0-pGMRES: iters={stp.iters} relres={stp.true_relres:.2e} {t_plain:.2f} s")
        plain = {"time_s": round(t_plain, 4), "iters": stp.iters, "restarts": stp.restarts,
                 "true_relres": stp.true_relres, "converged": bool(stp.true_relres <= 1e-10),
                 "setup_s": round(setup_plain, 3), "setup_plus_solve_s": round(setup_plain + t_plain, 4),
                 "ilu0_apply_ms": {"syncfree": round(apply_ms[1], 4), "level_launch": round(apply_ms[0], 4)}}

    cpu = None
    if not args.skip_cpu:
        cpu = cpu_baseline(p, plain["iters"] if plain else stats[-1]["iters"])
        log(f"cpu baseline {cpu['value']:.1f}s (est.)")

    extras = None
    if not args.skip_extras:
        extras = {"dg3d_bf16_mixed": bench_bf16(bsp, pr, ctx), "batched_ichol0_pcg": bench_batch(bsp, pr, ctx),
                  "ichol0_pcg_single": bench_pcg(bsp, pr, ctx), "learned_ilu0": bench_learned(bsp, pr, ctx),
                  "sor_pgmres_c1": bench_c1(bsp, pr, ctx), "bsmv_other_matrices": bench_bsmv_configs(bsp, pr, ctx)}

    value = ms_per_step / 1e3
    t0 = time.perf_counter()
    _x_host = _x_last[0].numpy()
    t_x = time.perf_counter() - t0
    host_handoff = {"bsr_create_s": round(t_bsr, 3), "matrix_bytes": mat_bytes,
                    "bsr_create_GBps": round(mat_bytes / t_bsr / 1e9, 2), "rhs_upload_ms": round(t_b * 1e3, 2),
                    "solution_download_ms": round(t_x * 1e3, 2),
                    "mixed_with_handoff_s": round(value + setup_mixed + t_bsr + t_b + t_x, 4),
                    "plain_with_handoff_s": (round(plain["time_s"] + plain["setup_s"] + t_bsr + t_b + t_x, 4)
                                             if plain else None)}
    out = {
        "metric": METRIC,
        "value": round(value, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 outer / f32 inner (bf16 factor storage)",
        "data": "synthetic",
        "config": {"workload": f"dg2d_veclap {args.nx}x{args.nx} Q1-SWIP-DG x D=2 (BlockSp<double,8>), solve to "
                               f"true FP64 relres 1e-10",
                   "n_block_rows": p.n, "bs": p.bs, "nnzb": p.nnzb,
                   "method": "mixed FGMRES(30,k=10): parity-coloured ILU0 stored in BF16, applied in FP32 + FP32 "
                             "SpMV, FP64 classical Gram-Schmidt/LS/deflation/update/residual, FP64 refinement"},
        "converged": ok,
        "iters": stats[-1]["iters"],
        "restarts": stats[-1]["restarts"],
        "true_relres": stats[-1]["true_relres"],
        "fwd_err": fwd_err,
        "fwd_err_reference": ref_desc,
        "setup_s": round(setup_mixed, 3),
        "setup_plus_solve_s": round(setup_mixed + value, 4),
        "mixed_with_f32_factor": f32_factor,
        "fp64_fgmres_same_coloured_ilu0": fp64_coloured,
        "precision_gain_same_ordering": round(fp64_coloured["time_s"] / value, 3),
        "plain_fp64": plain,
        "speedup_vs_plain_fp64": round(plain["time_s"] / value, 3) if plain else None,
        "speedup_vs_plain_fp64_with_setup": (round((plain["time_s"] + plain["setup_s"]) / (value + setup_mixed), 3)
                                             if plain else None),
        "host_handoff": host_handoff,
        "bsmv_GBps": round(bsmv_gbps, 1),
        "bsmm_AA": bsmm,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "workloads": extras,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
