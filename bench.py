#!/usr/bin/env python3
"""Benchmark: BEM surface-pressure solves/sec on one MI355X.

Workload (default, "precision"): rigid sphere r = 0.1 m as an icosphere at
subdivision 5 (10242 nodes, 20480 triangles), 2 kHz, c = 343 m/s,
rho = 1.21 kg/m^3, 16 CHIEF points seeded uniformly in r < 0.05 m (M = 10258
rows), unit plane wave along +z.  One step = complex FP32 assembly of A
(M x N, in HBM) + GMRES(30, tol 1e-6) on the CHIEF least-squares system (the
normal equations) with FP64 residual refinement -> nodal surface pressure.

Also reported: assembly GB/s (8 M N + 8 M R bytes written / assembly time),
p50/p99 step latency, the error of the FP32 result against a plain all-FP64 GPU
solve of the same discretisation and against the rigid-sphere series, the
roofline of the dominant kernel (the FP32 quadrature/assembly kernel, timed with
HIP events on the library stream), and the CPU port (the oracle's complex-FP32
restatement) timed on a row sample on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config NAME]
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
PEAK_HBM_GBS = 8000.0      # MI355X HBM3E
FLOP_PER_EVAL = 80.0       # nominal FP32 flops per (row, quadrature point) evaluation

CONFIGS = {
    # name: (subdiv, radius, freq, n_chief, chief_seed)
    "plain": dict(subdiv=3, radius=0.1, freq=171.0, nchief=8, seed=101, method="qr"),
    "precision": dict(subdiv=5, radius=0.1, freq=2000.0, nchief=16, seed=102, method="gmres"),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="precision", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run one device-to-device copy of MIB MiB (scale for FETCH/WRITE_SIZE counters)")
    return ap.parse_args()


def calibrate_copy(mib):
    """A plain copy kernel of known size, to scale the FETCH_SIZE / WRITE_SIZE
    hardware counters of the assembly kernel."""
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


def measured_traffic(kernel, config):
    """HBM bytes per launch measured with FETCH_SIZE / WRITE_SIZE counters (scaled
    by a copy kernel of known size) and recorded under profiles/; None if absent."""
    path = os.path.join(ROOT, "profiles", "%s_s5_hbm_traffic.txt" % config)
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 4 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel and f[2] == config:
            return float(f[3])
    return None


def cpu_baseline(mesh, k, rho, c, target_s, threads):
    """The oracle's complex-FP32 port of the reference assembly, timed on a row
    sample on the host cores and scaled to a full assembly (solve not counted)."""
    import oracle_py as O
    import bemcases as C
    O.quad_init(6)
    p, e, bc = C.mesh_arrays(mesh)
    bc = np.tile(np.array([0, 1, 0], complex), (len(e), 1))
    pts_all = np.vstack([p, mesh.chief()]).astype(np.float32)
    M, N, E = pts_all.shape[0], p.shape[0], e.shape[0]

    def run(rows):
        r0 = (M - rows) // 2
        t0 = time.perf_counter()
        O.assemble_rows(k, pts_all, N, e, bc, r0, r0 + rows, dens=rho, speed=c, dtype=np.float32,
                        nthreads=threads)
        return time.perf_counter() - t0

    # calibrate on a small sample, then size the timed sample for ~target_s
    t_cal = run(min(M, 2 * threads))
    per_row = t_cal / min(M, 2 * threads)
    rows = int(max(threads, min(M, target_s / max(per_row, 1e-9))))
    t = run(rows)
    per_row = t / rows
    full = per_row * M
    return {
        "value": 1.0 / full, "unit": "solves/s", "cores": threads, "kind": "port",
        "sample": "complex-FP32 oracle assembly of %d of %d rows on %d threads in %.1f s, scaled to the full "
                  "%d x %d system: %.2f s per solve (solve not counted)" % (rows, M, threads, t, M, N, full),
        "full_assembly_s": full,
    }


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (--gpus %d requested); the path shards by "
              "collocation rows or by request, see the design notes" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import bem_gpu_clean_cpp_amd as bem
    import bemcases as C

    cfg = CONFIGS[args.config]
    rho, c = 1.21, 343.0
    a = cfg["radius"]
    k = 2 * np.pi * cfg["freq"] / c
    bem.quad_init(6)
    bem.set_medium(rho, c)
    mesh = bem.Mesh().icosphere(cfg["subdiv"], a)
    mesh.set_bc(0, 1, 0)
    mesh.set_chief(C.ball_points(cfg["seed"], cfg["nchief"], a / 2))
    N, E, Cn = mesh.num_pnts, mesh.num_elems, mesh.num_chief
    M = N + Cn
    R = 1
    log("config %s: N=%d E=%d C=%d M=%d k=%.4f ka=%.3f" % (args.config, N, E, Cn, M, k, k * a))
    method = bem.SOLVE_GMRES if cfg["method"] == "gmres" else bem.SOLVE_QR
    solver = bem.Solver(mesh)
    solver.set_rhs(bem.RHS_PLANE, [[0, 0, 1]], 1.0)

    def step():
        solver.assemble(k, bem.PREC_FP32)
        return solver.solve(method, restart=30, tol=1e-6, maxit=3000, refine_tol=1e-8, max_refine=50)

    st = None
    for _ in range(args.warmup):
        st = step()
    log("warmup stats", st)
    bem.sync()
    bem.timing_reset()
    bem.timing_enable(True)
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        st = step()
        lat.append((time.perf_counter() - ts) * 1e3)
    bem.sync()
    t1 = time.perf_counter()
    bem.timing_enable(False)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    asm_ms, asm_n = bem.timing_get("asm_tile_f32")
    sing_ms, sing_n = bem.timing_get("sing_elems")
    fin_ms, fin_n = bem.timing_get("asm_finalize")
    gn_ms, gn_n = bem.timing_get("gemv_n_f32")
    gh_ms, gh_n = bem.timing_get("gemv_h_f32")
    rn_ms, rn_n = bem.timing_get("resid_n_f32a")
    bem.timing_reset()
    x32 = solver.solution()
    log("timed stats", st)

    # roofline of the dominant kernel: regular quadrature + assembly
    evals = 36.0 * (M * E - 3 * E)
    asm_avg_ms = asm_ms / max(asm_n, 1)
    achieved_tflops = FLOP_PER_EVAL * evals / (asm_avg_ms * 1e-3) / 1e12
    asm_total_ms = (asm_ms + sing_ms + fin_ms) / max(asm_n, 1)
    asm_gbs = (8.0 * M * N + 8.0 * M * R) / (asm_total_ms * 1e-3) / 1e9
    gemv_bytes = 8.0 * M * N + 8.0 * (M + N)
    gemv_gbs = gemv_bytes / (gn_ms / max(gn_n, 1) * 1e-3) / 1e9 if gn_n else None
    gemvh_gbs = gemv_bytes / (gh_ms / max(gh_n, 1) * 1e-3) / 1e9 if gh_n else None

    # accuracy: plain all-FP64 GPU solve of the same discretisation (also timed:
    # the 1x yardstick), and the rigid-sphere series
    err_fp64 = None
    fp64_ms = None
    if not args.no_fp64:
        solver64 = bem.Solver(mesh)
        solver64.set_rhs(bem.RHS_PLANE, [[0, 0, 1]], 1.0)
        t64 = []
        for i in range(2):
            ts = time.perf_counter()
            solver64.assemble(k, bem.PREC_FP64)
            st64 = solver64.solve(bem.SOLVE_GMRES, restart=30, tol=1e-12, maxit=5000, refine_tol=0, max_refine=0)
            bem.sync()
            t64.append((time.perf_counter() - ts) * 1e3)
        fp64_ms = min(t64)
        x64 = solver64.solution()
        err_fp64 = float(np.linalg.norm(x32 - x64) / np.linalg.norm(x64))
        log("fp64 stats", st64, "ms", t64)
        solver64.close()
    pn = mesh.pnts().astype(np.float64)
    ser = C.rigid_sphere_surface(a, k, pn[:, 2] / np.linalg.norm(pn, axis=1))
    err_series = float(np.linalg.norm(x32[:, 0] - ser) / np.linalg.norm(ser))

    cpu = None
    if not args.no_cpu_baseline:
        threads = min(os.cpu_count() or 1, 16)
        cpu = cpu_baseline(mesh, k, rho, c, args.cpu_seconds, threads)
        log("cpu baseline", cpu)

    value = 1000.0 / ms_per_step * R
    out = {
        "metric": "BEM solves/sec (surface pressure, <=1e-5 rel err vs FP64)",
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": "%s: rigid sphere icosphere s%d r=%.2f m, %g Hz, %d CHIEF, plane wave; "
                               "FP32 assembly + %s with FP64 refinement" %
                               (args.config, cfg["subdiv"], a, cfg["freq"], Cn,
                                "GMRES(30,1e-6) on normal equations" if method == bem.SOLVE_GMRES else "QR"),
                   "N": N, "E": E, "M": M, "rhs": R},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3)},
        "assembly_gbs": round(asm_gbs, 2),
        "assembly_ms": round(asm_total_ms, 3),
        "solve": {"iterations": st["iterations"], "refinements": st["refinements"], "converged": st["converged"],
                  "residual": st["residual"]},
        "rel_err_vs_fp64": err_fp64,
        "rel_err_vs_series": err_series,
        "fp64_gpu_ms_per_solve": fp64_ms,
        "speedup_vs_fp64_gpu": (fp64_ms / ms_per_step) if fp64_ms else None,
        "roofline": {"bound": "valu", "kernel": "asm_tile_f32", "achieved": round(achieved_tflops, 3),
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved_tflops / PEAK_FP32_TFLOPS, 4),
                     "traffic": measured_traffic("asm_tile_f32", args.config), "avg_ms": round(asm_avg_ms, 4),
                     "flops_per_launch": FLOP_PER_EVAL * evals},
        "roofline_gemv": {"bound": "hbm", "kernel": "gemv_n_f32", "achieved": gemv_gbs and round(gemv_gbs, 1),
                          "achieved_h": gemvh_gbs and round(gemvh_gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                          "frac": gemv_gbs and round(gemv_gbs / PEAK_HBM_GBS, 4), "bytes_per_launch": gemv_bytes,
                          "traffic": measured_traffic("gemv_n_f32", args.config),
                          "launches": gn_n},
        "cpu_baseline": cpu and {kk: cpu[kk] for kk in ("value", "unit", "cores", "kind", "sample")},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
