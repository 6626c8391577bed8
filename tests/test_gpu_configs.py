"""The batching and serving workloads at their own sizes.

Batching (64 frequencies 100 Hz - 8 kHz on a seeded perturbed ellipsoid at
icosphere subdivision 4, 2562 nodes, no CHIEF rows, 16 plane waves per
frequency; FP32 assembly + lockstep GMRES(100) with MFMA products and FP64
refinement): three of its frequencies -- the lowest (100 Hz), the middle one
(926 Hz) and the highest (8 kHz, which also needs the most GMRES iterations of
the 64, about 300) -- solved through the batch driver and checked against the
FP64 oracle (assembly + LAPACK) at the lowest and the highest and the all-FP64
device path at the middle one.  Bar: relative 2-norm error <= 1e-5 per
right-hand side.

Serving (requests on seeded perturbed spheres and ellipsoids at subdivision
3-6, ka up to 20, 1-32 plane-wave or point-source right-hand sides, 16 CHIEF
rows each, block CG on the normal equations with FP64 refinement): a mix that holds
an s5 request (sphere, 32 plane waves, ka = 20) and an s6 request (ellipsoid,
32 point sources 1 cm outside, ka = 19) besides small ones; every result is
bit-identical alone (one slot) and in the mix (four slots); the s5 and s6
results agree with the all-FP64 device solve to 1e-5, and the s5 result's
residual on 512 sampled rows of the oracle matrix equals that of the FP64
solution to 1e-6.
"""
import os

import numpy as np
import pytest

import bem_gpu_clean_cpp_amd as bem
import bemcases as C
import oracle_py as O

pytestmark = pytest.mark.gpu

RHO, SPEED = float(np.float32(1.21)), float(np.float32(343.0))

# the batch driver's iteration counts grow with the frequency; the highest
# (index 63) needs the most, about 300 GMRES(100) steps over 3 refinements
HARDEST = 63


@pytest.fixture(autouse=True, scope="module")
def _setup():
    bem.quad_init(6)
    O.quad_init(6)
    bem.set_medium(RHO, SPEED)
    yield
    bem.set_medium(1.2041, 343.21)


def _batching_problem():
    mesh = bem.Mesh().ellipsoid(4, (0.075, 0.09, 0.11), nbumps=8, lmax=6, amp=0.05, seed=103)
    mesh.set_bc(0, 1, 0)
    nf = 64
    freqs = 100.0 * 80.0 ** (np.arange(nf) / (nf - 1.0))
    ks = (2 * np.pi * freqs / SPEED).astype(np.float32)
    dirs = np.stack([C.sphere_points(1030 + f, 16, 1.0) for f in range(nf)]).astype(np.float32)
    return mesh, ks, dirs


def test_batching_workload_frequencies():
    mesh, ks, dirs = _batching_problem()
    assert (mesh.num_pnts, mesh.num_elems, mesh.num_chief) == (2562, 5120, 0)
    sel = [0, 32, HARDEST]
    kw = dict(restart=100, tol=1e-6, maxit=6000, refine_tol=1e-8, max_refine=50)
    X, stats = bem.batch_solve(mesh, ks[sel], bem.RHS_PLANE, dirs[sel], workers=3, **kw)
    for i, f in enumerate(sel):
        print("case: batching f[%d] = %.1f Hz (ka %.2f), stats %s" % (f, ks[f] * SPEED / (2 * np.pi),
                                                                       ks[f] * 0.11, stats[i]), flush=True)
        assert stats[i]["converged"] == 1
    p, e, bc = C.mesh_arrays(mesh)
    pts = p.astype(np.float64)
    nt = min(16, os.cpu_count() or 1)
    for i, f in enumerate(sel):
        k = float(ks[f])
        d = dirs[f].astype(np.float64)
        if f != 32:
            binit = np.exp(-1j * k * pts @ d.T)
            A, B = O.assemble(k, pts, p.shape[0], e, bc, binit=binit, dens=RHO, speed=SPEED, nthreads=nt)
            xr = np.linalg.solve(A, B)
            what = "FP64 oracle"
        else:
            s64 = bem.Solver(mesh)
            s64.assemble(k, bem.PREC_FP64)
            s64.set_rhs(bem.RHS_PLANE, dirs[f], 1.0)
            st = s64.solve(bem.SOLVE_GMRES, restart=100, tol=1e-12, maxit=20000, refine_tol=0, max_refine=0)
            xr = s64.solution()
            s64.close()
            what = "FP64 device solve (%d iterations)" % st["iterations"]
        err = np.linalg.norm(X[:, :, i] - xr, axis=0) / np.linalg.norm(xr, axis=0)
        print("case: batching f[%d], error vs %s: max %.3e median %.3e" % (f, what, err.max(), np.median(err)),
              flush=True)
        assert err.max() <= 1e-5


def _serving_requests():
    reqs = []
    # s5 sphere, ka = 20, 32 plane waves
    m5 = bem.Mesh().icosphere(5, 0.1)
    m5.set_chief(C.ball_points(5101, 16, 0.05))
    m5.set_bc(0, 1, 0)
    reqs.append((m5, 20.0 / 0.1, bem.RHS_PLANE, C.sphere_points(5102, 32, 1.0)))
    # s6 perturbed ellipsoid, ka = 19, 32 point sources 1 cm outside
    m6 = bem.Mesh().ellipsoid(6, (0.075, 0.09, 0.11), nbumps=8, lmax=6, amp=0.05, seed=6101)
    m6.set_chief(C.ball_points(6102, 16, 0.03))
    m6.set_bc(0, 1, 0)
    reqs.append((m6, 19.0 / 0.11, bem.RHS_POINT, C.sources_outside(m6, 6103, 32, 0.01)))
    # small ones around them
    rng = np.random.default_rng(104)
    for i in range(4):
        s = 3 + (i % 2)
        if i % 2:
            m = bem.Mesh().ellipsoid(s, (0.075, 0.09, 0.11), nbumps=8, lmax=6, amp=0.05, seed=7000 + i)
            m.set_chief(C.ball_points(7100 + i, 16, 0.03))
            a = 0.11
        else:
            m = bem.Mesh().icosphere(s, 0.1)
            m.set_chief(C.ball_points(7200 + i, 16, 0.05))
            a = 0.1
        m.set_bc(0, 1, 0)
        R = int(rng.integers(1, 33))
        k = float(rng.uniform(0.5, 20.0) / a)
        if i % 2:
            reqs.append((m, k, bem.RHS_POINT, C.sources_outside(m, 7300 + i, R, 0.01)))
        else:
            reqs.append((m, k, bem.RHS_PLANE, C.sphere_points(7400 + i, R, 1.0)))
    return reqs


def test_serving_workload_sizes():
    reqs = _serving_requests()
    kw = dict(method=bem.SOLVE_BLOCK_CG, restart=30, tol=1e-6, maxit=6000, refine_tol=1e-8, max_refine=50)
    alone = []
    srv = bem.Server(slots=1)
    for q, (m, k, kind, params) in enumerate(reqs):
        x, info = srv.wait(srv.submit(m, k, kind, params, **kw))
        print("case: request %d alone: N %d, R %d, ka %.1f, stats %s" % (q, m.num_pnts, len(params),
                                                                          k * (0.1 if q != 1 else 0.11),
                                                                          info["solve"]), flush=True)
        assert info["status"] == 0 and info["solve"]["converged"] == 1, info
        alone.append(x)
    srv.close()
    srv = bem.Server(slots=4)
    ids = [srv.submit(m, k, kind, params, **kw) for m, k, kind, params in reqs]
    mix = [srv.wait(i) for i in ids]
    st = srv.status()
    srv.close()
    print("case: mix of %d requests on 4 slots, status %s" % (len(reqs), st), flush=True)
    for q, ((x, info), xa) in enumerate(zip(mix, alone)):
        assert info["status"] == 0, info
        assert np.array_equal(x, xa), q
    assert st["max_running"] > 1
    # the s5 and s6 results against the all-FP64 device solve (4 and 2 of their right-hand sides)
    for q, nr in ((0, 4), (1, 2)):
        m, k, kind, params = reqs[q]
        s64 = bem.Solver(m)
        s64.assemble(k, bem.PREC_FP64)
        s64.set_rhs(kind, params[:nr], 1.0)
        st64 = s64.solve(bem.SOLVE_CGLS, tol=1e-6, maxit=20000, refine_tol=1e-12, max_refine=10)
        x64 = s64.solution()
        s64.close()
        err = np.linalg.norm(alone[q][:, :nr] - x64, axis=0) / np.linalg.norm(x64, axis=0)
        print("case: request %d (N %d) vs FP64 device solve: errors %s, FP64 stats %s" % (q, m.num_pnts, err, st64),
              flush=True)
        assert st64["converged"] == 1
        assert err.max() <= 1e-5
        if q == 0:
            # residual of the served solution on 512 sampled rows of the oracle matrix
            N, M = m.num_pnts, m.num_pnts + m.num_chief
            p, e, bc = C.mesh_arrays(m)
            pts = np.vstack([p, m.chief()]).astype(np.float64)
            r0 = int(np.random.default_rng(5104).integers(0, M - 512))
            ao = O.assemble_rows(k, pts, N, e, bc, r0, r0 + 512, dens=RHO, speed=SPEED, dtype=np.float64,
                                 nthreads=min(16, os.cpu_count() or 1))
            d = np.asarray(params[:nr], np.float64)
            b = np.exp(-1j * k * pts[r0:r0 + 512] @ d.T)
            res_served = np.linalg.norm(b - ao @ alone[q][:, :nr]) / np.linalg.norm(b)
            res_fp64 = np.linalg.norm(b - ao @ x64) / np.linalg.norm(b)
            print("case: request 0, rows %d-%d of the oracle matrix: residual served %.4e, FP64 %.4e" %
                  (r0, r0 + 512, res_served, res_fp64), flush=True)
            assert abs(res_served - res_fp64) <= 1e-6
