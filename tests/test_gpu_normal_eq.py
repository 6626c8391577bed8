"""The one-pass normal-equation product A^H (A v) over a row-major copy of A,
used by GMRES on the CHIEF normal equations with one right-hand side.

Checks, for a size in each of its launch shapes (N = 642: 1024 threads x 1
column pair; 2562: 1024 x 2; 5124, two spheres: 512 x 8; 10242: 512 x 12):
  * the refined solution against the two-GEMV operator (BEM_NE_TWO_PASS):
    relative 2-norm difference <= 1e-7 (both refined to an FP64 residual of
    1e-10);
  * at N = 642, the solution against the oracle's FP64 least squares: <= 1e-5.
"""
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


def _mesh(case):
    if case == "two_s4":
        m = C.two_spheres(bem, 4, 0.1, gap=3.0)
        m.set_chief(C.ball_points(23, 16, 0.05) - [0.25, 0, 0])
    else:
        m = bem.Mesh().icosphere(int(case[1]), 0.1)
        m.set_bc(0, 1, 0)
        m.set_chief(C.ball_points(23, 16, 0.05))
    return m


def _solve(m, k, two_pass, monkeypatch):
    if two_pass:
        monkeypatch.setenv("BEM_NE_TWO_PASS", "1")
    else:
        monkeypatch.delenv("BEM_NE_TWO_PASS", raising=False)
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP32)
    s.set_rhs(bem.RHS_PLANE, [[0, 0, 1]], 1.0)
    st = s.solve(bem.SOLVE_GMRES, restart=30, tol=1e-6, refine_tol=1e-10, max_refine=50)
    x = s.solution()
    s.close()
    return x, st


@pytest.mark.parametrize("case", ["s3", "s4", "two_s4", "s5"])
def test_one_pass_product_matches_two_gemvs(case, monkeypatch):
    m = _mesh(case)
    k = 2 * np.pi * 1000.0 / SPEED
    print("case %s: N %d, CHIEF %d" % (case, m.num_pnts, m.num_chief), flush=True)
    x1, st1 = _solve(m, k, False, monkeypatch)
    x2, st2 = _solve(m, k, True, monkeypatch)
    d = np.linalg.norm(x1 - x2) / np.linalg.norm(x2)
    print("one pass: %s\ntwo GEMVs: %s\nrelative difference %.3g" % (st1, st2, d))
    assert st1["converged"] == 1 and st2["converged"] == 1
    assert d < 1e-7


def test_one_pass_product_matches_oracle():
    m = _mesh("s3")
    k = 2 * np.pi * 1000.0 / SPEED
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP32)
    s.set_rhs(bem.RHS_PLANE, [[0, 0, 1]], 1.0)
    st = s.solve(bem.SOLVE_GMRES, restart=30, tol=1e-6, refine_tol=1e-10, max_refine=50)
    x = s.solution()
    p, e, bc = C.mesh_arrays(m)
    pts = np.vstack([p, m.chief()]).astype(np.float64)
    binit = np.exp(-1j * k * pts[:, 2:3])
    Ao, Bo = O.assemble(k, pts, p.shape[0], e, bc, binit=binit, dens=RHO, speed=SPEED)
    xo = np.linalg.lstsq(Ao, Bo, rcond=None)[0]
    err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    print("error vs oracle %.3g, %s" % (err, st))
    assert st["converged"] == 1
    assert err < 1e-5


def test_workspace_kept_between_solves_matches_fresh_solver(monkeypatch):
    """A solver keeps its GMRES workspace (basis, row-major copy of A) between
    solves of one shape and replaces it when the shape changes: a repeated
    solve, a solve after reassembly at another wavenumber and a solve with
    another restart length must each equal a fresh solver's, bit for bit."""
    monkeypatch.delenv("BEM_NE_TWO_PASS", raising=False)
    m = _mesh("s3")
    k1 = 2 * np.pi * 1000.0 / SPEED
    k2 = 2 * np.pi * 1500.0 / SPEED

    def fresh(k, restart):
        s = bem.Solver(m)
        s.assemble(k, bem.PREC_FP32)
        s.set_rhs(bem.RHS_PLANE, [[0, 0, 1]], 1.0)
        s.solve(bem.SOLVE_GMRES, restart=restart, tol=1e-6, refine_tol=1e-10, max_refine=50)
        x = s.solution()
        s.close()
        return x

    s = bem.Solver(m)
    s.assemble(k1, bem.PREC_FP32)
    s.set_rhs(bem.RHS_PLANE, [[0, 0, 1]], 1.0)
    kw = dict(tol=1e-6, refine_tol=1e-10, max_refine=50)
    s.solve(bem.SOLVE_GMRES, restart=30, **kw)
    xa = s.solution()
    s.solve(bem.SOLVE_GMRES, restart=30, **kw)
    xb = s.solution()
    s.assemble(k2, bem.PREC_FP32)
    s.solve(bem.SOLVE_GMRES, restart=30, **kw)
    xc = s.solution()
    s.solve(bem.SOLVE_GMRES, restart=20, **kw)
    xd = s.solution()
    s.close()
    np.testing.assert_array_equal(xa, fresh(k1, 30))
    np.testing.assert_array_equal(xb, xa)
    np.testing.assert_array_equal(xc, fresh(k2, 30))
    np.testing.assert_array_equal(xd, fresh(k2, 20))
