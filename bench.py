#!/usr/bin/env python3
"""Benchmark: BEM surface-pressure solves/sec on one MI355X.

Headline workload ("large"): seeded perturbed ellipsoid (semi-axes
0.075/0.09/0.11 m, 8 spherical-harmonic bumps up to degree 6, amplitude <= 5 %)
at icosphere subdivision 6 (40962 nodes, 81920 triangles), rigid, 12 kHz
(ka ~ 24), c = 343 m/s, rho = 1.21 kg/m^3, 32 CHIEF points (M = 40994 rows,
a 13.4 GB complex64 matrix), 32 unit point sources 1 cm outside the surface
as right-hand sides.  One step = complex FP32 assembly of A in HBM + the
least-squares solve of all 32 right-hand sides (block conjugate gradients on
the normal equations over the 32 sources, FP32 MFMA products, FP64 refinement
of the normal-equation residual to 1e-10); value = 32 solves / step time.

Also reported on that line: p50/p99 step latency, assembly time and GB/s,
the error against the all-FP64 device solve of the same system, the plain
single-RHS FP64 GPU solve time (the 1x yardstick), the MFMA roofline of the
dominant kernels (the multi-right-hand-side products, timed with HIP events on
the library stream in an instrumented step after the timed ones), and the CPU
baseline: the oracle's complex-FP32 port of the assembly on a 2048-row sample
plus a BLAS product on the same rows, on all the host cores this process may
use, extrapolated to the whole step.

Sections of the same JSON line: precision (icosphere s5, 2 kHz, 16 CHIEF, one
plane wave; GMRES on the normal equations), plain (s3, QR + field operator),
batching (64 frequencies x 16 plane waves), serving (Poisson request stream).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--only SECTION]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 (vector and dense MFMA)
PEAK_FP64_TFLOPS = 78.6    # MI355X FP64 (vector and dense MFMA)


def executed_share():
    """Share of the algorithmic complex-product flops (8 per complex
    multiply-add) that the MFMA pipes execute: the products' three-product
    form takes 6 real MFMA flops per complex multiply-add, the four-product
    form (BEM_GEMM_4M) all 8."""
    return 1.0 if os.environ.get("BEM_GEMM_4M") else 0.75
PEAK_HBM_GBS = 8000.0      # MI355X HBM3E
PEAK_FP16_TFLOPS = 2500.0  # MI355X dense FP16 MFMA
# measured: a read-only kernel (64-row x 640-column wave tiles, 16-byte loads,
# 2 workgroups per CU) streams 13.4 GB at 6.15-6.24 TB/s on MI355X
READ_CEILING_GBS = 6200.0
FLOP_PER_EVAL = 80.0       # nominal FP32 flops per (row, quadrature point) evaluation
# measured VALU issue rate of the FP32 assembly kernel (wave-instructions per
# SIMD-cycle, SQ counters of profiles/precision_s5_asm_valu_counters.txt)
ASM_VALU_ISSUE = 0.237

PRECISION = dict(subdiv=5, radius=0.1, freq=2000.0, nchief=16, seed=102)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed steps of the headline (large) workload")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-batching", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--large-method", choices=["cgls", "gmres", "block"], default="block",
                    help="solver of the headline workload (cgls: CG per right-hand side; block: block CG over all "
                         "32; gmres: BA-GMRES(100) on the normal equations)")
    ap.add_argument("--large-refine-tol", type=float, default=1e-10,
                    help="FP64 refinement target (normal-equation residual) of the headline workload; the error "
                         "against the all-FP64 solve is 5.81e-7 at 1e-10 and at 1e-11 (FP32 assembly bound)")
    ap.add_argument("--precision-steps", type=int, default=10)
    ap.add_argument("--plain-subdiv", type=int, default=3, help="icosphere subdivision of the plain workload")
    ap.add_argument("--plain-reps", type=int, default=20)
    ap.add_argument("--only", choices=["plain", "precision", "batching", "large", "serving"],
                    help="run one workload alone and print its section (not the benchmark line)")
    ap.add_argument("--workers", type=int, default=4, help="concurrent solvers of the batch driver")
    ap.add_argument("--batch-method", choices=["gmres", "block", "cgls"], default="gmres",
                    help="solver of the batching workload")
    ap.add_argument("--batch-tune", action="store_true",
                    help="run the product tile tuner for the batching size before the timed batch")
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--serve-rate", type=float, default=2.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-method", choices=sorted(SERVE_METHODS), default="block",
                    help="solver of the serving requests (cgls: CG per right-hand side; block: block CG)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run one device-to-device copy of MIB MiB (scale for FETCH/WRITE_SIZE counters)")
    return ap.parse_args()


def calibrate_copy(mib):
    """A plain copy kernel of known size, to scale the FETCH_SIZE / WRITE_SIZE
    hardware counters of the library's kernels."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes": n * 4}), flush=True)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(kernel, workload):
    """HBM bytes per launch measured with FETCH_SIZE / WRITE_SIZE counters (scaled
    by a copy kernel of known size) and recorded under profiles/ as lines
    'traffic_bytes_per_launch <kernel> <workload> <bytes>'; None if absent."""
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None
    for name in sorted(os.listdir(pdir)):
        if not name.endswith("_hbm_traffic.txt"):
            continue
        for line in open(os.path.join(pdir, name)):
            f = line.split()
            if len(f) == 4 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel and f[2] == workload:
                return float(f[3])
    return None


def host_cores():
    """Cores this process may run on: its CPU affinity, bounded by a cgroup
    CPU quota when one is set (cpu.max), so the CPU baseline uses every core
    it has and no more."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            n = min(n, max(1, int(float(quota) / float(period))))
    except (OSError, ValueError):
        pass
    return n


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def median_of(fn, n=5):
    """1 warm-up call, then the median of n timed calls (seconds)."""
    fn()
    ts = []
    for _ in range(n):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts


def cpu_baseline_large(mesh, k, rho, c, R, products, threads):
    """The CPU path of the headline step, extrapolated from samples: the
    oracle's complex-FP32 port of the reference assembly (OpenMP over rows) on
    512-row samples (1 warm-up, median of 5 at 5 places, 2048 rows in all),
    scaled to all M rows, plus the solve's `products` passes of the matrix
    against the 32 right-hand sides as complex64 BLAS products (numpy, the
    same threads) on a 2048-row block, scaled to M rows."""
    import oracle_py as O
    import bemcases as C
    from threadpoolctl import threadpool_limits
    O.quad_init(6)
    p, e, bc = C.mesh_arrays(mesh)
    pts_all = np.vstack([p, mesh.chief()]).astype(np.float32)
    M, N = pts_all.shape[0], p.shape[0]
    starts = iter([int(M * f) for f in (0.1, 0.3, 0.5, 0.7, 0.9, 0.05)])

    def asm():
        r0 = next(starts)
        O.assemble_rows(k, pts_all, N, e, bc, r0, r0 + 512, dens=rho, speed=c, dtype=np.float32, nthreads=threads)

    t_asm, ts_asm = median_of(asm)
    rng = np.random.default_rng(7)
    a = (rng.standard_normal((2048, N)) + 1j * rng.standard_normal((2048, N))).astype(np.complex64)
    x = (rng.standard_normal((N, R)) + 1j * rng.standard_normal((N, R))).astype(np.complex64)
    with threadpool_limits(limits=threads):
        t_mv, ts_mv = median_of(lambda: a @ x)
    asm_full = t_asm / 512 * M
    solve_full = products * t_mv / 2048 * M
    total = asm_full + solve_full
    return {
        "value": R / total, "unit": "solves/s", "cores": threads, "kind": "port", "extrapolated": True,
        "cpu_model": cpu_model(),
        "sample": "extrapolated: complex-FP32 oracle assembly, median of 5 samples of 512 rows (%.2f s each) "
                  "-> %.1f s for all %d rows; %d products of the matrix with the %d right-hand sides as complex64 "
                  "BLAS on a 2048-row block (median of 5, %.3f s) -> %.1f s; %.1f s per %d-solve step on %d "
                  "threads" % (t_asm, asm_full, M, products, R, t_mv, solve_full, total, R, threads),
        "assembly_s": asm_full, "solve_s": solve_full,
    }


def bench_batching(bem, C, workers, fp64=True, method=1, tune=False):
    """Multi-frequency, multi-right-hand-side batch: seeded perturbed ellipsoid
    (semi-axes 0.075/0.09/0.11 m, 8 spherical-harmonic bumps up to degree 6,
    amplitude <= 5%) at icosphere subdivision 4 (2562 nodes), rigid, 64
    frequencies log-spaced 100 Hz - 8 kHz, 16 plane-wave directions per
    frequency drawn uniformly on the sphere; FP32 assembly + lockstep GMRES(100)
    (MFMA multi-RHS products) with FP64 refinement to 1e-8.  The system has no
    CHIEF rows, so frequencies near interior resonances are ill-conditioned; a
    restart length of 100 keeps their Krylov spaces long enough."""
    c = 343.0
    mesh = bem.Mesh().ellipsoid(4, (0.075, 0.09, 0.11), nbumps=8, lmax=6, amp=0.05, seed=103)
    mesh.set_bc(0, 1, 0)
    N, E, M = mesh.num_pnts, mesh.num_elems, mesh.num_pnts
    nf, R = 64, 16
    freqs = 100.0 * 80.0 ** (np.arange(nf) / (nf - 1.0))
    ks = (2 * np.pi * freqs / c).astype(np.float32)
    dirs = np.stack([C.sphere_points(1030 + f, R, 1.0) for f in range(nf)]).astype(np.float32)
    kw = dict(method=method, restart=100, tol=1e-6, maxit=6000, refine_tol=1e-8, max_refine=50)
    tuned = None
    if tune:
        # the product tile tuner for this size class (outside the timed region)
        tuned = bem.autotune_products(M, N, R)
        log("batching: product tuner %s" % tuned)
    # the batch context (the workers' resident solvers) is set up outside the
    # timed region, as the mesh upload is; its first run allocates the solve
    # workspaces
    bem.sync()
    ts = time.perf_counter()
    batch = bem.Batch(mesh, workers)
    batch.run(ks[-workers:], bem.RHS_PLANE, dirs[-workers:], **kw)  # warm-up
    bem.sync()
    setup_s = time.perf_counter() - ts
    t0 = time.perf_counter()
    X, stats = batch.run(ks, bem.RHS_PLANE, dirs, **kw)
    wall = time.perf_counter() - t0
    batch.close()
    # the same batch through bem_batch_solve, which creates and destroys the
    # workers' solvers inside the call
    bem.sync()
    t1 = time.perf_counter()
    bem.batch_solve(mesh, ks, bem.RHS_PLANE, dirs, workers=workers, **kw)
    bem.sync()
    wall_oneshot = time.perf_counter() - t1
    conv = all(s["converged"] == 1 for s in stats)
    its = [s["iterations"] for s in stats]
    log("batching: %d x %d solves in %.3f s (%d workers), iterations min/median/max %d/%d/%d, converged %s" %
        (nf, R, wall, workers, min(its), int(np.median(its)), max(its), conv))
    # kernel roofline of the MFMA product, measured with one worker so the
    # kernel has the GPU to itself
    bem.timing_reset()
    bem.timing_enable(True)
    bem.batch_solve(mesh, ks[-2:], bem.RHS_PLANE, dirs[-2:], workers=1, **kw)
    bem.sync()
    bem.timing_enable(False)
    g_ms, g_n = bem.timing_get("gemm_n_f32")
    bem.timing_reset()
    g_avg = g_ms / max(g_n, 1)
    flops = 8.0 * M * N * R
    gbytes = 8.0 * M * N + 8.0 * (M + N) * R
    # accuracy against the all-FP64 path and the 1x yardstick (plain single-RHS
    # FP64 solve) at the lowest, middle and highest frequency
    errs, t64 = [], []
    for f in ((0, nf // 2, nf - 1) if fp64 else ()):
        s64 = bem.Solver(mesh)
        s64.assemble(float(ks[f]), bem.PREC_FP64)
        s64.set_rhs(bem.RHS_PLANE, dirs[f], 1.0)
        s64.solve(bem.SOLVE_GMRES, restart=30, tol=1e-12, maxit=20000, refine_tol=0, max_refine=0)
        x64 = s64.solution()
        errs.append(float(np.max(np.linalg.norm(X[:, :, f] - x64, axis=0) / np.linalg.norm(x64, axis=0))))
        s64.set_rhs(bem.RHS_PLANE, dirs[f][:1], 1.0)
        bem.sync()
        ts = time.perf_counter()
        s64.assemble(float(ks[f]), bem.PREC_FP64)
        s64.solve(bem.SOLVE_GMRES, restart=30, tol=1e-12, maxit=20000, refine_tol=0, max_refine=0)
        bem.sync()
        t64.append(time.perf_counter() - ts)
        s64.close()
    value = nf * R / wall
    fp64_rate = 1.0 / float(np.mean(t64)) if t64 else None
    return {
        "workload": "batching: perturbed ellipsoid s4 (N=%d, E=%d), rigid, 64 freqs 100 Hz-8 kHz x 16 plane waves; "
                    "FP32 assembly + %s with MFMA multi-RHS products, FP64 refinement to 1e-8, "
                    "%d concurrent workers of a batch context set up before the timed run" % (N, E, {bem.SOLVE_GMRES: "lockstep GMRES(100,1e-6)",
                                                      bem.SOLVE_BLOCK_CG: "block CG (normal equations, 1e-6)",
                                                      bem.SOLVE_CGLS: "lockstep CG (normal equations, 1e-6)"}[method],
                                               workers),
        "value": round(value, 2), "unit": "solves/s", "wall_s": round(wall, 4), "solves": nf * R,
        "setup_s": round(setup_s, 4),
        "oneshot": {"wall_s": round(wall_oneshot, 4), "solves_per_s": round(nf * R / wall_oneshot, 2),
                    "note": "bem_batch_solve: the workers' solvers created and destroyed inside the call"},
        "iterations": {"min": int(min(its)), "median": int(np.median(its)), "max": int(max(its))},
        "converged": conv,
        "product_tuner": tuned,
        "rel_err_vs_fp64_max": max(errs) if errs else None,
        "fp64_single_rhs_solves_per_s": fp64_rate and round(fp64_rate, 3),
        "speedup_vs_fp64_single_rhs": fp64_rate and round(value / fp64_rate, 2),
        "roofline_gemm": {"bound": "mfma", "kernel": "gemm_n_f32", "avg_ms": round(g_avg, 5), "launches": g_n,
                          "achieved": round(flops / (g_avg * 1e-3) / 1e12, 3), "peak": PEAK_FP32_TFLOPS,
                          "unit": "TFLOP/s", "frac": round(flops / (g_avg * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4),
                          "hbm_gbs": round(gbytes / (g_avg * 1e-3) / 1e9, 1), "flops_per_launch": flops,
                          "executed_flops_per_launch": flops * executed_share(),
                          "executed_frac": round(flops * executed_share() / (g_avg * 1e-3) / 1e12 / PEAK_FP32_TFLOPS,
                                                 4),
                          "bytes_per_launch": gbytes},
    }


def bench_plain(bem, C, reps=20, subdiv=3):
    """Plain workload (the reference's own small case): rigid sphere r = 0.1 m
    as icosphere s3 (642 nodes), 171 Hz, 8 CHIEF points, unit plane wave along
    +z; FP32 assembly + QR least squares with FP64 refinement + the field at 64
    seeded points on r = 1 m through the field operator.  Errors: surface vs
    the rigid-sphere series, scattered field vs the exterior series.  A larger
    subdivision runs the same path at that size (for profiling the QR); one
    extra solve with kernel timing on gives the QR's panel and trailing-update
    times and the trailing updates' MFMA rate."""
    c, a = 343.0, 0.1
    k = 2 * np.pi * 171.0 / c
    m = bem.Mesh().icosphere(subdiv, a)
    m.set_bc(0, 1, 0)
    m.set_chief(C.ball_points(101, 8, a / 2))
    xs = C.sphere_points(101, 64, 1.0)
    s = bem.Solver(m)
    s.set_rhs(bem.RHS_PLANE, [[0, 0, 1]], 1.0)

    def one():
        s.assemble(k, bem.PREC_FP32)
        st = s.solve(bem.SOLVE_QR, refine_tol=1e-10, max_refine=10)
        s.field_matrix(xs)
        return st, s.field_apply()[:, 0]

    one()
    bem.sync()
    lat = []
    t0 = time.perf_counter()
    for _ in range(reps):
        ts = time.perf_counter()
        st, p = one()
        lat.append((time.perf_counter() - ts) * 1e3)
    wall = time.perf_counter() - t0
    bem.timing_reset()
    bem.timing_enable(True)
    s.solve(bem.SOLVE_QR, refine_tol=1e-10, max_refine=10)
    bem.sync()
    bem.timing_enable(False)
    p_ms, p_n = bem.timing_get("qr_panel_f32")
    t_ms, t_n = bem.timing_get("qr_trailing_f32")
    bem.timing_reset()
    Mq, Nq = m.num_pnts + m.num_chief, m.num_pnts
    qr_flop = 8.0 * Mq * Nq * Nq - 8.0 * Nq ** 3 / 3
    x = s.solution()[:, 0]
    pn = m.pnts().astype(np.float64)
    ser = C.rigid_sphere_surface(a, k, pn[:, 2] / np.linalg.norm(pn, axis=1))
    inc = np.exp(-1j * k * xs[:, 2].astype(np.float64))
    fser = C.rigid_sphere_field(a, k, xs)
    s.close()
    return {
        "workload": "plain: rigid sphere icosphere s%d (N=%d), 171 Hz, 8 CHIEF, plane wave; FP32 assembly + QR with "
                    "FP64 refinement + field at 64 points on r=1 m (field operator)" % (subdiv, Nq),
        "value": round(reps / wall, 2), "unit": "solves/s", "latency_ms": {
            "p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3)},
        "solve": {kk: st[kk] for kk in ("refinements", "converged", "residual")},
        "rel_err_surface_vs_series": float(np.linalg.norm(x - ser) / np.linalg.norm(ser)),
        "rel_err_scattered_field_vs_series": float(np.linalg.norm((p - inc) - (fser - inc)) /
                                                   np.linalg.norm(fser - inc)),
        "qr_factor": {"panels": p_n, "panel_ms": round(p_ms, 3), "trailing_ms": round(t_ms, 3),
                      "trailing_tflops_mfma": round(qr_flop / (t_ms * 1e-3) / 1e12, 2) if t_ms else None,
                      "flop": qr_flop},
    }


def bench_large(bem, C, steps, warmup, fp64=True, cpu=True, method="block", refine_tol=1e-10):
    """Headline: seeded perturbed ellipsoid at icosphere subdivision 6 (40962
    nodes, 81920 triangles, 13.4 GB complex64 matrix), rigid, 12 kHz, 32 CHIEF
    points, 32 point sources 1 cm outside the surface.  One step = FP32
    assembly + the least-squares solve of the 32 right-hand sides (block CG on
    the normal equations over all 32 by default, FP32 MFMA products, FP64
    refinement to a normal-equation residual of refine_tol).  Timed steps run
    without per-launch events; kernel durations for the rooflines come from
    one instrumented step after them."""
    c, rho = 343.0, 1.21
    m, src = C.large_case(bem)
    N, E, Cn = m.num_pnts, m.num_elems, m.num_chief
    M = N + Cn
    R = src.shape[0]
    k = 2 * np.pi * 12000.0 / c
    if method == "gmres":
        kw = dict(method=bem.SOLVE_GMRES, restart=100, tol=1e-6, maxit=6000, refine_tol=refine_tol, max_refine=20)
    elif method == "block":
        kw = dict(method=bem.SOLVE_BLOCK_CG, tol=1e-6, maxit=6000, refine_tol=refine_tol, max_refine=10)
    else:
        kw = dict(method=bem.SOLVE_CGLS, tol=1e-6, maxit=6000, refine_tol=refine_tol, max_refine=10)
    sv = bem.Solver(m)
    sv.set_rhs(bem.RHS_POINT, src, 1.0)

    def step():
        sv.assemble(k, bem.PREC_FP32)
        return sv.solve(**kw)

    st = None
    for i in range(warmup):
        st = step()
        log("large warm-up %d: %s" % (i, st))
    bem.sync()
    lat = []
    t0 = time.perf_counter()
    for i in range(steps):
        ts = time.perf_counter()
        st = step()
        lat.append((time.perf_counter() - ts) * 1e3)
        log("large step %d: %.3f s, %s" % (i, lat[-1] / 1e3, st))
    bem.sync()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / steps
    x32 = sv.solution()
    # one instrumented step: per-launch HIP events on the library stream
    bem.timing_reset()
    bem.timing_enable(True)
    sti = step()
    bem.sync()
    bem.timing_enable(False)
    t = {name: bem.timing_get(name) for name in ("gemm_n_f32", "gemm_h_f32", "gemm_hc_f32", "gemm_n_f16s",
                                                 "gemm_hc_f16s", "adjoint_copy", "asm_tile_f32", "sing_elems",
                                                 "asm_finalize", "ghost_reduce", "mres_f64acc", "mh_f64acc")}
    bem.timing_reset()
    # the same step with block CG's products on the FP32 kernels (BEM_GEMM_NO_F16),
    # reported beside the headline: one warm-up, then the median of 2 steps
    f32p = None
    if method == "block":
        os.environ["BEM_GEMM_NO_F16"] = "1"
        try:
            step()
            bem.sync()
            lat32 = []
            for i in range(2):
                ts = time.perf_counter()
                st32 = step()
                bem.sync()
                lat32.append((time.perf_counter() - ts) * 1e3)
            x32f = sv.solution()
        finally:
            del os.environ["BEM_GEMM_NO_F16"]
        ms32 = float(np.median(lat32))
        f32p = {"note": "block CG products on the FP32 32x32x2 MFMA kernels (BEM_GEMM_NO_F16), same step",
                "ms_per_step": round(ms32, 2), "value": round(R / (ms32 / 1e3), 3),
                "solve": {kk: st32[kk] for kk in ("iterations", "refinements", "converged", "residual")}}
        log("large, FP32 products: %.3f s, %s" % (ms32 / 1e3, st32))
    sv.close()
    avg = {kk: (v[0] / v[1] if v[1] else None) for kk, v in t.items()}
    flops = 8.0 * M * N * R
    gbytes = 8.0 * M * N + 8.0 * (M + N) * R
    # block CG runs its products on the FP16-split kernels ("gemm_n_f16s",
    # "gemm_hc_f16s") unless BEM_GEMM_NO_F16 is set; the adjoint products run on
    # the Y = A X kernel over the conjugate-transposed copy when the solver made
    # one ("gemm_hc_*"), else on the adjoint kernel
    f16s = bool(t["gemm_n_f16s"][1])
    nname = "gemm_n_f16s" if f16s else "gemm_n_f32"
    hname = "gemm_hc_f16s" if f16s else ("gemm_hc_f32" if t["gemm_hc_f32"][1] else "gemm_h_f32")
    g_avg, h_avg = avg[nname], avg[hname]
    asm_ms = sum(v[0] for kk, v in t.items() if kk in ("asm_tile_f32", "sing_elems", "asm_finalize",
                                                        "ghost_reduce")) / max(t["asm_tile_f32"][1], 1)
    out = {
        "workload": "large: perturbed ellipsoid s6 (N=%d, E=%d, M=%d, A %.1f GB), rigid, 12 kHz, 32 CHIEF, "
                    "32 point sources 1 cm outside; FP32 assembly + %s (%s, R=32) with FP64 "
                    "refinement to %.0e" % (N, E, M, 8.0 * M * N / 1e9,
                                            {"gmres": "BA-GMRES(100) on the normal equations",
                                             "block": "block CG on the normal equations"}.get(
                                                method, "CG on the normal equations"),
                                            "FP32 matrix, FP16-split MFMA products with FP32 accumulation" if f16s
                                            else "FP32 MFMA products", refine_tol),
        "N": N, "E": E, "M": M, "rhs": R,
        "value": R / (ms_per_step / 1e3), "ms_per_step": ms_per_step,
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2), "p99": round(float(np.percentile(lat, 99)), 2)},
        "assembly_ms": round(asm_ms, 2),
        "assembly_gbs": round((8.0 * M * N + 8.0 * M * R) / (asm_ms * 1e-3) / 1e9, 1),
        "solve": {kk: st[kk] for kk in ("iterations", "refinements", "converged", "residual")},
        "instrumented_step": {"products_n": t[nname][1], "products_h": t[hname][1],
                              "adjoint_copy_ms": round(t["adjoint_copy"][0], 3) if t["adjoint_copy"][1] else None,
                              "solve_ms": round(sti["solve_ms"], 1)},
        "roofline": {"bound": "mfma", "kernel": "gemm_n_f32", "achieved": round(flops / (g_avg * 1e-3) / 1e12, 2),
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(flops / (g_avg * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4),
                     "traffic": measured_traffic("gemm_n_f32", "large"), "avg_ms": round(g_avg, 3),
                     "launches": t["gemm_n_f32"][1], "flops_per_launch": flops, "bytes_per_launch": gbytes,
                     "executed_flops_per_launch": flops * executed_share(),
                     "executed_frac": round(flops * executed_share() / (g_avg * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4),
                     "hbm_gbs": round(8.0 * M * N / (g_avg * 1e-3) / 1e9, 1),
                     "adjoint": {"kernel": hname, "avg_ms": round(h_avg, 3),
                                 "achieved": round(flops / (h_avg * 1e-3) / 1e12, 2),
                                 "frac": round(flops / (h_avg * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4),
                                 "traffic": measured_traffic(hname, "large")}},
    }
    if f16s:
        # the FP16-split product executes 24 FP16 MFMA flops per complex
        # multiply-add (four real products, each as hi*hi + hi*lo + lo*hi) at
        # 2.5 PFLOP/s: its time floor is the A stream, so its roofline is HBM
        gbs = gbytes / (g_avg * 1e-3) / 1e9
        hgbs = gbytes / (h_avg * 1e-3) / 1e9
        out["roofline"] = {"bound": "hbm", "kernel": nname, "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS,
                           "unit": "GB/s", "frac": round(gbs / PEAK_HBM_GBS, 4),
                           "read_ceiling": {"gbs": READ_CEILING_GBS, "frac": round(gbs / READ_CEILING_GBS, 4),
                                            "note": "a read-only kernel with the product's access pattern over "
                                                    "the same 13.4 GB on MI355X (measured once)"},
                           "traffic": measured_traffic(nname, "large"), "avg_ms": round(g_avg, 3),
                           "launches": t[nname][1], "bytes_per_launch": gbytes, "flops_per_launch": flops,
                           "complex_tflops": round(flops / (g_avg * 1e-3) / 1e12, 2),
                           "mfma_fp16": {"executed_flops_per_launch": 3.0 * flops,
                                         "achieved": round(3.0 * flops / (g_avg * 1e-3) / 1e12, 1),
                                         "peak": PEAK_FP16_TFLOPS, "unit": "TFLOP/s",
                                         "frac": round(3.0 * flops / (g_avg * 1e-3) / 1e12 / PEAK_FP16_TFLOPS, 4)},
                           "adjoint": {"kernel": hname, "avg_ms": round(h_avg, 3), "achieved": round(hgbs, 1),
                                       "frac": round(hgbs / PEAK_HBM_GBS, 4),
                                       "traffic": measured_traffic(hname, "large")}}
    if f32p is not None:
        out["fp32_products"] = f32p
    if avg["mres_f64acc"]:
        out["roofline_fp64_residual"] = {"bound": "mfma", "kernel": "mres_f64acc (FP64 MFMA, FP32 A widened)",
                                         "avg_ms": round(avg["mres_f64acc"], 3),
                                         "achieved": round(flops / (avg["mres_f64acc"] * 1e-3) / 1e12, 2),
                                         "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                                         "frac": round(flops / (avg["mres_f64acc"] * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
                                                       4),
                                         "executed_frac": round(flops * executed_share() /
                                                                (avg["mres_f64acc"] * 1e-3) / 1e12 /
                                                                PEAK_FP64_TFLOPS, 4)}
    if fp64:
        # 1x yardstick: plain single-RHS FP64 solve (FP64 assembly + FP64 CG),
        # also the reference solution for the first two sources
        s64 = bem.Solver(m)
        s64.set_rhs(bem.RHS_POINT, src[:1], 1.0)
        bem.sync()
        ts = time.perf_counter()
        s64.assemble(k, bem.PREC_FP64)
        st64 = s64.solve(bem.SOLVE_CGLS, tol=1e-6, maxit=6000, refine_tol=1e-12, max_refine=10)
        bem.sync()
        t64 = time.perf_counter() - ts
        x64 = s64.solution()
        s64.close()
        err = float(np.linalg.norm(x32[:, 0] - x64[:, 0]) / np.linalg.norm(x64[:, 0]))
        log("large fp64: %.2f s, stats %s, err %.3g" % (t64, st64, err))
        out.update({"rel_err_vs_fp64": err, "fp64_single_rhs_s": round(t64, 3),
                    "speedup_vs_fp64_single_rhs": round(out["value"] * t64, 2)})
        if f32p is not None:
            f32p["rel_err_vs_fp64"] = float(np.linalg.norm(x32f[:, 0] - x64[:, 0]) / np.linalg.norm(x64[:, 0]))
    if cpu:
        products = 2 * st["iterations"] + 2 * (st["refinements"] + 1)
        out["cpu_baseline"] = cpu_baseline_large(m, k, rho, c, R, products, host_cores())
        log("cpu baseline", out["cpu_baseline"])
    return out


def bench_precision(bem, C, steps, warmup=2, fp64=True):
    """Precision: rigid sphere r = 0.1 m, icosphere s5 (10242 nodes, 20480
    triangles), 2 kHz, 16 CHIEF points in r < 0.05 m, one plane wave along +z.
    One step = FP32 assembly + GMRES(30, 1e-6) on the normal equations with
    FP64 refinement to 1e-8.  Reports the step rate, the error against the
    all-FP64 GPU solve (also timed: the 1x yardstick) and the sphere series,
    the assembly kernel's nominal and counter-based VALU figures, the FP64
    yardstick assembly against the FP64 peak, and the one-pass
    normal-equation product against HBM."""
    cfg = PRECISION
    c = 343.0
    a = cfg["radius"]
    k = 2 * np.pi * cfg["freq"] / c
    mesh = bem.Mesh().icosphere(cfg["subdiv"], a)
    mesh.set_bc(0, 1, 0)
    mesh.set_chief(C.ball_points(cfg["seed"], cfg["nchief"], a / 2))
    N, E, Cn = mesh.num_pnts, mesh.num_elems, mesh.num_chief
    M = N + Cn
    solver = bem.Solver(mesh)
    solver.set_rhs(bem.RHS_PLANE, [[0, 0, 1]], 1.0)

    def step():
        solver.assemble(k, bem.PREC_FP32)
        return solver.solve(bem.SOLVE_GMRES, restart=30, tol=1e-6, maxit=3000, refine_tol=1e-8, max_refine=50)

    for _ in range(warmup):
        st = step()
    bem.sync()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        st = step()
        lat.append((time.perf_counter() - ts) * 1e3)
    bem.sync()
    ms_per_step = (time.perf_counter() - t0) * 1e3 / steps
    bem.timing_reset()
    bem.timing_enable(True)
    for _ in range(3):
        step()
    bem.sync()
    bem.timing_enable(False)
    tm = {name: bem.timing_get(name) for name in ("asm_tile_f32", "sing_elems", "asm_finalize", "ghost_reduce",
                                                  "ne_apply_part", "ne_apply_f32")}
    bem.timing_reset()
    x32 = solver.solution()
    evals = 36.0 * (M * E - 3 * E)
    asm_avg_ms = tm["asm_tile_f32"][0] / max(tm["asm_tile_f32"][1], 1)
    asm_total_ms = sum(tm[kk][0] for kk in ("asm_tile_f32", "sing_elems", "asm_finalize", "ghost_reduce")) / max(
        tm["asm_tile_f32"][1], 1)
    nominal = FLOP_PER_EVAL * evals / (asm_avg_ms * 1e-3) / 1e12
    ne_ms, ne_n = tm["ne_apply_part"]
    ne_bytes = 8.0 * M * N + 16.0 * N
    ne_avg_us = ne_ms / ne_n * 1e3 if ne_n else None
    ne_gbs = ne_bytes / (ne_avg_us * 1e-6) / 1e9 if ne_n else None
    out = {
        "workload": "precision: rigid sphere icosphere s5 r=0.10 m (N=%d, E=%d, M=%d), 2000 Hz, 16 CHIEF, plane wave; "
                    "FP32 assembly + GMRES(30,1e-6) on the normal equations with FP64 refinement to 1e-8" % (N, E, M),
        "value": round(1000.0 / ms_per_step, 3), "unit": "solves/s", "ms_per_step": round(ms_per_step, 3),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3)},
        "assembly_ms": round(asm_total_ms, 3),
        "assembly_gbs": round((8.0 * M * N + 8.0 * M) / (asm_total_ms * 1e-3) / 1e9, 2),
        "solve": {kk: st[kk] for kk in ("iterations", "refinements", "converged", "residual")},
        "roofline_asm": {"bound": "valu", "kernel": "asm_tile_f32", "avg_ms": round(asm_avg_ms, 4),
                         "nominal_flops_per_launch": FLOP_PER_EVAL * evals,
                         "nominal_achieved": round(nominal, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                         "nominal_frac": round(nominal / PEAK_FP32_TFLOPS, 4),
                         "valu_issue_frac_counters": ASM_VALU_ISSUE,
                         "traffic": measured_traffic("asm_tile_f32", "precision")},
        "roofline_ne": {"bound": "hbm", "kernel": "ne_apply_part", "achieved": ne_gbs and round(ne_gbs, 1),
                        "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": ne_gbs and round(ne_gbs / PEAK_HBM_GBS, 4),
                        "bytes_per_launch": ne_bytes, "avg_us": ne_avg_us and round(ne_avg_us, 2),
                        "avg_us_with_reduce": tm["ne_apply_f32"][1] and round(
                            tm["ne_apply_f32"][0] / tm["ne_apply_f32"][1] * 1e3, 2),
                        "traffic": measured_traffic("ne_apply_part", "precision"), "launches": ne_n},
    }
    if fp64:
        solver64 = bem.Solver(mesh)
        solver64.set_rhs(bem.RHS_PLANE, [[0, 0, 1]], 1.0)
        t64 = []
        for i in range(2):
            ts = time.perf_counter()
            solver64.assemble(k, bem.PREC_FP64)
            st64 = solver64.solve(bem.SOLVE_GMRES, restart=30, tol=1e-12, maxit=5000, refine_tol=0, max_refine=0)
            bem.sync()
            t64.append((time.perf_counter() - ts) * 1e3)
        x64 = solver64.solution()
        # the yardstick's assembly kernel against the FP64 peak (same nominal count)
        bem.timing_reset()
        bem.timing_enable(True)
        solver64.assemble(k, bem.PREC_FP64)
        bem.sync()
        bem.timing_enable(False)
        a64_ms, a64_n = bem.timing_get("asm_tile_f64")
        bem.timing_reset()
        solver64.close()
        a64 = a64_ms / max(a64_n, 1)
        nominal64 = FLOP_PER_EVAL * evals / (a64 * 1e-3) / 1e12
        out.update({
            "rel_err_vs_fp64": float(np.linalg.norm(x32 - x64) / np.linalg.norm(x64)),
            "fp64_gpu_ms_per_solve": round(min(t64), 3),
            "speedup_vs_fp64_gpu": round(min(t64) / ms_per_step, 3),
            "fp64_yardstick_assembly": {"kernel": "asm_tile_f64 (generic FP64 kernel, halo plan)",
                                        "avg_ms": round(a64, 3), "nominal_achieved": round(nominal64, 3),
                                        "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                                        "nominal_frac": round(nominal64 / PEAK_FP64_TFLOPS, 4)},
        })
    pn = mesh.pnts().astype(np.float64)
    ser = C.rigid_sphere_surface(a, k, pn[:, 2] / np.linalg.norm(pn, axis=1))
    out["rel_err_vs_series"] = float(np.linalg.norm(x32[:, 0] - ser) / np.linalg.norm(ser))
    solver.close()
    return out


def serving_requests(bem, C, n, seed):
    """Seeded request mix: sphere (r = 0.1 m) or perturbed ellipsoid with
    p = 1/2, icosphere subdivision uniform in 3..6, ka uniform in [0.5, 20],
    1..32 right-hand sides, plane waves or point sources 1 cm radially outside
    the surface with p = 1/2; 16 CHIEF points inside each body."""
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        s = int(rng.integers(3, 7))
        if rng.random() < 0.5:
            m = bem.Mesh().icosphere(s, 0.1)
            a = 0.1
            m.set_chief(C.ball_points(seed * 1000 + i, 16, 0.05))
        else:
            m = bem.Mesh().ellipsoid(s, (0.075, 0.09, 0.11), nbumps=8, lmax=6, amp=0.05, seed=seed * 1000 + i)
            a = 0.11
            m.set_chief(C.ball_points(seed * 1000 + i, 16, 0.03))
        m.set_bc(0, 1, 0)
        k = float(rng.uniform(0.5, 20.0) / a)
        R = int(rng.integers(1, 33))
        if rng.random() < 0.5:
            kind, params = bem.RHS_PLANE, C.sphere_points(seed * 7000 + i, R, 1.0)
        else:
            pn = m.pnts()
            v = pn[rng.integers(0, pn.shape[0], R)].astype(np.float64)
            params = v * (1.0 + 0.01 / np.linalg.norm(v, axis=1))[:, None]
            kind = bem.RHS_POINT
        reqs.append(dict(mesh=m, k=k, kind=kind, params=np.asarray(params, np.float32), s=s, R=R))
    return reqs


SERVE_METHODS = {"cgls": lambda bem: bem.SOLVE_CGLS, "block": lambda bem: bem.SOLVE_BLOCK_CG,
                 "gmres": lambda bem: bem.SOLVE_GMRES}


def bench_serving(bem, C, n, rate, slots, seed=104, method=None):
    """Poisson request stream into the request server; latency = completion -
    arrival (queueing included).  The CHIEF systems are solved with block CG
    (or CG per right-hand side) on the normal equations (FP32 MFMA products,
    FP64 refinement to 1e-8)."""
    import threading
    reqs = serving_requests(bem, C, n, seed)
    gaps = np.random.default_rng(seed + 1).exponential(1.0 / rate, n)
    arrivals = np.cumsum(gaps) - gaps[0]
    kw = dict(method=bem.SOLVE_BLOCK_CG if method is None else method, restart=30, tol=1e-6, maxit=6000,
              refine_tol=1e-8, max_refine=50)
    srv = bem.Server(slots=slots)
    # warm-up: one small request
    srv.wait(srv.submit(reqs[0]["mesh"], reqs[0]["k"], reqs[0]["kind"], reqs[0]["params"], **kw))
    ids = [None] * n
    results = [None] * n
    ready = [threading.Event() for _ in range(n)]

    def collect():
        for i in range(n):
            ready[i].wait()
            results[i] = srv.wait(ids[i])

    th = threading.Thread(target=collect)
    th.start()
    t0 = time.perf_counter()
    for i, r in enumerate(reqs):
        dt = arrivals[i] - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        ids[i] = srv.submit(r["mesh"], r["k"], r["kind"], r["params"], **kw)
        ready[i].set()
        if (i + 1) % 8 == 0:
            log("serving: submitted %d/%d" % (i + 1, n))
    th.join()
    wall = time.perf_counter() - t0
    st = srv.status()
    # the same requests all at once: the server's capacity (saturated queue)
    t0 = time.perf_counter()
    bids = [srv.submit(r["mesh"], r["k"], r["kind"], r["params"], **kw) for r in reqs]
    bres = [srv.wait(i) for i in bids]
    bwall = time.perf_counter() - t0
    srv.close()
    blat = np.array([res[1]["end_ms"] - res[1]["submit_ms"] for res in bres])
    lat = np.array([res[1]["end_ms"] - res[1]["submit_ms"] for res in results])
    queue = np.array([res[1]["start_ms"] - res[1]["submit_ms"] for res in results])
    conv = all(res[1]["solve"]["converged"] == 1 for res in results)
    cols = sum(r["R"] for r in reqs)
    # accuracy: the smallest requests against the all-FP64 path
    errs = []
    for i in sorted(range(n), key=lambda j: reqs[j]["mesh"].num_pnts)[:2]:
        r = reqs[i]
        s64 = bem.Solver(r["mesh"])
        s64.assemble(r["k"], bem.PREC_FP64)
        s64.set_rhs(r["kind"], r["params"], 1.0)
        s64.solve(bem.SOLVE_CGLS, tol=1e-6, maxit=20000, refine_tol=1e-12, max_refine=10)
        x64 = s64.solution()
        s64.close()
        errs.append(float(np.max(np.linalg.norm(results[i][0] - x64, axis=0) / np.linalg.norm(x64, axis=0))))
    sizes = {}
    for r in reqs:
        sizes[r["s"]] = sizes.get(r["s"], 0) + 1
    return {
        "workload": "serving: %d Poisson requests (rate %.3g/s, seed %d), sphere/perturbed ellipsoid s3-s6, "
                    "ka U[0.5,20], 1-32 plane-wave or point-source RHS, 16 CHIEF; FP32 assembly + %s on the normal "
                    "equations with FP64 refinement; %d slots, HBM-footprint admission" %
                    (n, rate, seed, "block CG" if kw["method"] == bem.SOLVE_BLOCK_CG else "CG", slots),
        "requests": n, "rhs_columns": cols, "wall_s": round(wall, 3),
        "solves_per_s": round(cols / wall, 2), "requests_per_s": round(n / wall, 3),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2), "p99": round(float(np.percentile(lat, 99)), 2),
                       "max": round(float(lat.max()), 2)},
        "queue_ms": {"p50": round(float(np.percentile(queue, 50)), 2), "p99": round(float(np.percentile(queue, 99)), 2)},
        "burst": {"wall_s": round(bwall, 3), "solves_per_s": round(cols / bwall, 2),
                  "requests_per_s": round(n / bwall, 3),
                  "latency_ms": {"p50": round(float(np.percentile(blat, 50)), 2),
                                 "p99": round(float(np.percentile(blat, 99)), 2)}},
        "converged": conv, "subdiv_counts": {str(k): v for k, v in sorted(sizes.items())},
        "max_running": st["max_running"], "peak_hbm_gb": round(st["peak_in_use_bytes"] / 1e9, 2),
        "budget_gb": round(st["budget_bytes"] / 1e9, 1),
        "rel_err_vs_fp64_smallest2": errs,
    }


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print("bench.py: measures one GPU only (--gpus %d requested); across GPUs the solver would "
              "shard by collocation rows (one large system) or by frequency and request (many small "
              "ones)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import bem_gpu_clean_cpp_amd as bem
    import bemcases as C

    bem.quad_init(6)
    bem.set_medium(1.21, 343.0)
    fp64 = not args.no_fp64
    if args.only:
        # one workload alone (for profiling); prints that section's JSON
        if args.only == "plain":
            res = bench_plain(bem, C, args.plain_reps, args.plain_subdiv)
        elif args.only == "precision":
            res = bench_precision(bem, C, args.precision_steps, fp64=fp64)
        elif args.only == "batching":
            res = bench_batching(bem, C, args.workers, fp64=fp64, method=SERVE_METHODS[args.batch_method](bem),
                                 tune=args.batch_tune)
        elif args.only == "large":
            res = bench_large(bem, C, args.steps, args.warmup, fp64=fp64, cpu=not args.no_cpu_baseline,
                              method=args.large_method, refine_tol=args.large_refine_tol)
        else:
            res = bench_serving(bem, C, args.serve_requests, args.serve_rate, args.serve_slots,
                                method=SERVE_METHODS[args.serve_method](bem))
        print(json.dumps({args.only: res}), flush=True)
        return

    large = bench_large(bem, C, args.steps, args.warmup, fp64=fp64, cpu=not args.no_cpu_baseline,
                        method=args.large_method, refine_tol=args.large_refine_tol)
    log("large", {kk: v for kk, v in large.items() if kk != "cpu_baseline"})
    sections = {}
    if not args.no_precision:
        sections["precision"] = bench_precision(bem, C, args.precision_steps, fp64=fp64)
        log("precision", sections["precision"])
    if not args.no_plain:
        sections["plain"] = bench_plain(bem, C, args.plain_reps, args.plain_subdiv)
        log("plain", sections["plain"])
    if not args.no_batching:
        sections["batching"] = bench_batching(bem, C, args.workers, fp64=fp64,
                                              method=SERVE_METHODS[args.batch_method](bem), tune=args.batch_tune)
        log("batching", sections["batching"])
    if not args.no_serving:
        sections["serving"] = bench_serving(bem, C, args.serve_requests, args.serve_rate, args.serve_slots,
                                            method=SERVE_METHODS[args.serve_method](bem))
        log("serving", sections["serving"])

    head_keys = ("value", "ms_per_step", "latency_ms", "roofline", "cpu_baseline", "workload", "N", "E", "M", "rhs")
    out = {
        "metric": "BEM solves/sec (surface pressure, <=1e-5 rel err vs FP64)",
        "value": round(large["value"], 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(large["ms_per_step"], 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": large["workload"], "N": large["N"], "E": large["E"], "M": large["M"],
                   "rhs": large["rhs"]},
        "latency_ms": large["latency_ms"],
        "roofline": large["roofline"],
        "cpu_baseline": large.get("cpu_baseline") and {kk: large["cpu_baseline"][kk] for kk in (
            "value", "unit", "cores", "kind", "sample", "extrapolated", "cpu_model")},
    }
    out.update({kk: v for kk, v in large.items() if kk not in head_keys})
    out.update(sections)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
