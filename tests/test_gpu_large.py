"""Large system: 40962 nodes, 81920 triangles, 32 CHIEF rows (a 13.4 GB
complex64 matrix), 12 kHz (ka ~ 24), 32 point sources 1 cm outside.

Checks (an FP64 oracle of the whole system would take hours on the CPU):
  * 2048 rows (8 seeded blocks of 256) of the FP32 and FP64 device matrices
    against the FP64 oracle: max |dA| <= 2e-5 max |A| (FP32) and
    <= 1e-10 max |A| (FP64);
  * the FP32 solve (block conjugate gradients on the normal equations over
    all 32 sources, FP32 MFMA products, FP64 refinement) against the all-FP64
    device solve (conjugate gradients per source) of the same system for 2 of
    the sources: relative 2-norm error <= 1e-5;
  * the same FP32 solve with conjugate gradients per source: the two agree to
    1e-7 after refinement to 1e-11;
  * the block solve refined to 1e-10 only (the benchmark's setting): against
    the FP64 solve at 1e-5 as well, and within 1e-6 of the 1e-11 solution;
  * every source's refined FP64 normal-equation residual <= 1e-11, and the
    residual of the FP32 solution on the sampled rows evaluated with the
    oracle's rows equals the one with the device rows to 1e-6.
"""
import os

import numpy as np
import pytest

import bem_gpu_clean_cpp_amd as bem
import bemcases as C
import oracle_py as O

pytestmark = pytest.mark.gpu

RHO, SPEED = float(np.float32(1.21)), float(np.float32(343.0))


@pytest.fixture(autouse=True, scope="module")
def _setup():
    bem.quad_init(6)
    O.quad_init(6)
    bem.set_medium(RHO, SPEED)
    yield
    bem.set_medium(1.2041, 343.21)


def test_large_system_rows_and_solution():
    m, src = C.large_case(bem)
    N, Cn = m.num_pnts, m.num_chief
    M = N + Cn
    assert (N, m.num_elems, Cn) == (40962, 81920, 32)
    k = float(np.float32(2 * np.pi * 12000.0 / SPEED))
    s32 = bem.Solver(m)
    s32.assemble(k, bem.PREC_FP32)
    s32.set_rhs(bem.RHS_POINT, src, 1.0)
    st = s32.solve(bem.SOLVE_CGLS, tol=1e-6, maxit=6000, refine_tol=1e-11, max_refine=10)
    print("fp32 lockstep CG stats", st, flush=True)
    assert st["converged"] == 1 and st["residual"] <= 1e-11
    xl = s32.solution()
    st = s32.solve(bem.SOLVE_BLOCK_CG, tol=1e-6, maxit=6000, refine_tol=1e-11, max_refine=10)
    print("fp32 block CG stats", st, flush=True)
    assert st["converged"] == 1 and st["residual"] <= 1e-11
    x32 = s32.solution()
    dbl = (np.linalg.norm(x32 - xl, axis=0) / np.linalg.norm(xl, axis=0)).max()
    print("block vs lockstep CG: %.3g" % dbl, flush=True)
    assert dbl <= 1e-7
    st10 = s32.solve(bem.SOLVE_BLOCK_CG, tol=1e-6, maxit=6000, refine_tol=1e-10, max_refine=10)
    x10 = s32.solution()
    d10 = (np.linalg.norm(x10 - x32, axis=0) / np.linalg.norm(x32, axis=0)).max()
    print("block CG to 1e-10: stats %s, vs the 1e-11 solution %.3g" % (st10, d10), flush=True)
    assert st10["converged"] == 1 and d10 <= 1e-6
    starts = np.sort(np.random.default_rng(1053).choice(M // 256, 8, replace=False)) * 256
    rows = np.concatenate([np.arange(r, min(r + 256, M)) for r in starts])
    a32 = s32.rows(rows)
    s32.close()
    s64 = bem.Solver(m)
    s64.assemble(k, bem.PREC_FP64)
    s64.set_rhs(bem.RHS_POINT, src[:2], 1.0)
    st64 = s64.solve(bem.SOLVE_CGLS, tol=1e-6, maxit=6000, refine_tol=1e-12, max_refine=10)
    print("fp64 stats", st64, flush=True)
    x64 = s64.solution()
    a64 = s64.rows(rows)
    s64.close()
    err = np.linalg.norm(x32[:, :2] - x64, axis=0) / np.linalg.norm(x64, axis=0)
    err10 = np.linalg.norm(x10[:, :2] - x64, axis=0) / np.linalg.norm(x64, axis=0)
    print("solution error vs FP64: refined to 1e-11 %s, to 1e-10 %s" % (err, err10), flush=True)
    assert st64["converged"] == 1
    assert err.max() <= 1e-5 and err10.max() <= 1e-5
    p, e, bc = C.mesh_arrays(m)
    pts = np.vstack([p, m.chief()]).astype(np.float64)
    ao = np.zeros_like(a32)
    for r0 in starts:
        r1 = min(r0 + 256, M)
        sel = (rows >= r0) & (rows < r1)
        ao[sel] = O.assemble_rows(k, pts, N, e, bc, int(r0), int(r1), dens=RHO, speed=SPEED, dtype=np.float64,
                                  nthreads=min(16, os.cpu_count() or 1))
    amax = np.abs(ao).max()
    e32, e64 = np.abs(a32 - ao).max() / amax, np.abs(a64 - ao).max() / amax
    # residual of the FP32 solution on the sampled rows: oracle rows vs device rows
    dist = np.linalg.norm(pts[rows][:, None, :] - src.astype(np.float64)[None, :, :], axis=2)
    b = np.exp(-1j * k * dist) / (4 * np.pi * dist)
    res_o = np.linalg.norm(b - ao @ x32) / np.linalg.norm(b)
    res_d = np.linalg.norm(b - a32 @ x32) / np.linalg.norm(b)
    print("row sample: fp32 %.3g fp64 %.3g (relative to max |A| %.3g); residual oracle rows %.3e device rows %.3e"
          % (e32, e64, amax, res_o, res_d))
    assert e32 <= 2e-5
    assert e64 <= 1e-10
    assert abs(res_o - res_d) <= 1e-6
