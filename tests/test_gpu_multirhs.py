"""Many right-hand sides solved in lockstep (MFMA multi-RHS products).

The batched GMRES applies the FP32 operator to every system's basis vector in
one MFMA product (A X, and A^H T for CHIEF's normal equations) and computes the
refinement residuals for all systems with FP64 accumulation.  Checks:
  * each column against the FP64 oracle's least-squares solution, relative
    2-norm error <= 1e-5 (the same bar as single right-hand sides);
  * each column against the same system solved alone (the GEMV path) after
    refinement to 1e-10: relative difference <= 1e-8;
  * the FP64 matrix path (per-system GEMVs, FP64 multi-RHS residuals) against
    the oracle at 1e-9.
Right-hand-side counts cover one and two 16-wide MFMA groups, ragged groups
(20, 40) and the maximum of 64.
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


def _mesh(s, nchief, r=0.1, seed=7):
    m = bem.Mesh().icosphere(s, r)
    if nchief:
        m.set_chief(C.ball_points(seed, nchief, r / 2))
    return m


def _dirs(n, seed=11):
    return C.sphere_points(seed, n, 1.0)


def _oracle_lstsq(m, k, d):
    p, e, bc = C.mesh_arrays(m)
    pts = np.vstack([p, m.chief()]).astype(np.float64)
    binit = np.exp(-1j * k * pts @ np.asarray(d, np.float64).T)
    A, B = O.assemble(k, pts, p.shape[0], e, bc, binit=binit, dens=RHO, speed=SPEED)
    return np.linalg.lstsq(A, B, rcond=None)[0]


def _relcols(x, y):
    return np.linalg.norm(x - y, axis=0) / np.linalg.norm(y, axis=0)


@pytest.mark.parametrize("nchief", [0, 8])
@pytest.mark.parametrize("R", [16, 20, 40])
def test_multi_rhs_gmres_matches_oracle(nchief, R):
    m = _mesh(3, nchief)
    k = 31.3
    d = _dirs(R)
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP32)
    s.set_rhs(bem.RHS_PLANE, d, 1.0)
    st = s.solve(bem.SOLVE_GMRES, restart=30, tol=1e-6, maxit=3000, refine_tol=1e-10, max_refine=50)
    x = s.solution()
    assert x.shape == (m.num_pnts, R)
    xo = _oracle_lstsq(m, k, d)
    err = _relcols(x, xo)
    print("R %d nchief %d stats %s worst err %.3g" % (R, nchief, st, err.max()))
    assert st["converged"] == 1
    assert err.max() < 1e-5


def test_autotuned_products_keep_parity():
    """The product tile tuner (kernel family and split-K chunks per size
    class) on the s3 CHIEF system with 20 right-hand sides: the solve with the
    tuned shapes in force still matches the oracle at the 1e-5 bar."""
    m = _mesh(3, 8)
    M, N = m.num_pnts + m.num_chief, m.num_pnts
    k = 31.3
    d = _dirs(20)
    try:
        r = bem.autotune_products(M, N, 20)
        print("case: product tuner, M %d N %d R 20: %s" % (M, N, r))
        assert r["rhs_per_wave_n"] in (16, 32) and 1 <= r["chunks_n"] <= 16
        assert r["rhs_per_wave_h"] in (16, 32) and 1 <= r["chunks_h"] <= 16
        assert 0 < r["best_ms_n"] and 0 < r["best_ms_h"]
        s = bem.Solver(m)
        s.assemble(k, bem.PREC_FP32)
        s.set_rhs(bem.RHS_PLANE, d, 1.0)
        st = s.solve(bem.SOLVE_GMRES, restart=30, tol=1e-6, maxit=3000, refine_tol=1e-10, max_refine=50)
        err = _relcols(s.solution(), _oracle_lstsq(m, k, d))
        print("case: tuned solve stats %s worst err %.3g" % (st, err.max()))
        assert st["converged"] == 1 and err.max() < 1e-5
    finally:
        bem.tune_clear()


@pytest.mark.parametrize("nchief", [0, 8])
def test_multi_rhs_equals_single_rhs(nchief):
    m = _mesh(3, nchief)
    k = 31.3
    d = _dirs(20, seed=5)
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP32)
    s.set_rhs(bem.RHS_PLANE, d, 1.0)
    s.solve(bem.SOLVE_GMRES, restart=30, tol=1e-6, maxit=3000, refine_tol=1e-10, max_refine=50)
    xb = s.solution()
    for q in (0, 7, 19):
        s.set_rhs(bem.RHS_PLANE, d[q:q + 1], 1.0)
        st = s.solve(bem.SOLVE_GMRES, restart=30, tol=1e-6, maxit=3000, refine_tol=1e-10, max_refine=50)
        x1 = s.solution()[:, 0]
        rel = np.linalg.norm(xb[:, q] - x1) / np.linalg.norm(x1)
        assert st["converged"] == 1
        assert rel < 1e-8, (q, rel)


def test_multi_rhs_max_batch_point_sources():
    """64 right-hand sides (the largest batch) of point sources inside the body."""
    m = _mesh(2, 4)
    k = 12.0
    src = C.ball_points(3, 64, 0.04)
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP32)
    s.set_rhs(bem.RHS_POINT, src, 1.0)
    st = s.solve(bem.SOLVE_GMRES, restart=40, tol=1e-6, maxit=3000, refine_tol=1e-10, max_refine=50)
    x = s.solution()
    p, e, bc = C.mesh_arrays(m)
    pts = np.vstack([p, m.chief()]).astype(np.float64)
    r = np.linalg.norm(pts[:, None, :] - np.asarray(src, np.float64)[None, :, :], axis=2)
    binit = np.exp(-1j * k * r) / (4 * np.pi * r)
    A, B = O.assemble(k, pts, p.shape[0], e, bc, binit=binit, dens=RHO, speed=SPEED)
    xo = np.linalg.lstsq(A, B, rcond=None)[0]
    err = _relcols(x, xo)
    print("stats %s worst err %.3g" % (st, err.max()))
    assert st["converged"] == 1
    assert err.max() < 1e-5


def test_multi_rhs_fp64_matrix():
    m = _mesh(2, 6)
    k = 20.0
    d = _dirs(9, seed=2)
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP64)
    s.set_rhs(bem.RHS_PLANE, d, 1.0)
    st = s.solve(bem.SOLVE_GMRES, restart=30, tol=1e-12, maxit=3000, refine_tol=1e-12, max_refine=20)
    x = s.solution()
    xo = _oracle_lstsq(m, k, d)
    err = _relcols(x, xo)
    assert st["converged"] == 1, st
    assert err.max() < 1e-9, err.max()


def test_multi_rhs_qr_matches_gmres():
    m = _mesh(3, 8)
    k = 31.3
    d = _dirs(24, seed=9)
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP32)
    s.set_rhs(bem.RHS_PLANE, d, 1.0)
    s.solve(bem.SOLVE_GMRES, restart=30, tol=1e-6, maxit=3000, refine_tol=1e-10, max_refine=50)
    xg = s.solution()
    st = s.solve(bem.SOLVE_QR, refine_tol=1e-10, max_refine=20)
    xq = s.solution()
    assert st["converged"] == 1
    assert _relcols(xq, xg).max() < 1e-7


@pytest.mark.parametrize("R", [1, 20])
@pytest.mark.parametrize("prec", [bem.PREC_FP32, bem.PREC_FP64])
def test_cgls_matches_oracle(R, prec):
    """Conjugate gradients on the CHIEF normal equations (lockstep, FP64
    refinement) against the oracle's least-squares solution."""
    m = _mesh(3, 8)
    k = 31.3
    d = _dirs(R, seed=13)
    s = bem.Solver(m)
    s.assemble(k, prec)
    s.set_rhs(bem.RHS_PLANE, d, 1.0)
    st = s.solve(bem.SOLVE_CGLS, tol=1e-6, maxit=5000, refine_tol=1e-10, max_refine=20)
    x = s.solution()
    xo = _oracle_lstsq(m, k, d)
    err = _relcols(x, xo)
    print("R %d prec %d stats %s worst err %.3g" % (R, prec, st, err.max()))
    assert st["converged"] == 1
    assert err.max() < (1e-5 if prec == bem.PREC_FP32 else 1e-8)


def test_gmres_graph_replays_equal_eager(monkeypatch):
    """GMRES Arnoldi chunks captured as HIP graphs (one per first step, step
    count and Gram-Schmidt mode, on a chunk's third launch; opt-in through
    BEM_GMRES_GRAPH) and replayed in
    later cycles and later solves of the same solver, with the host's Givens rotations, the FP64
    residuals and the solution updates between replays: three solves in a row
    at different wavenumbers return the same bits as eager launches, and the
    oracle's solution."""
    m = _mesh(3, 0)
    s = bem.Solver(m)
    d = _dirs(16, seed=5)
    s.set_rhs(bem.RHS_PLANE, d, 1.0)
    monkeypatch.delenv("BEM_NO_GRAPH", raising=False)
    for k in (20.0, 31.3, 42.0):
        s.assemble(k, bem.PREC_FP32)
        monkeypatch.setenv("BEM_GMRES_GRAPH", "1")
        st = s.solve(bem.SOLVE_GMRES, restart=30, tol=1e-6, maxit=3000, refine_tol=1e-10, max_refine=50)
        xg = s.solution()
        monkeypatch.delenv("BEM_GMRES_GRAPH")
        s.solve(bem.SOLVE_GMRES, restart=30, tol=1e-6, maxit=3000, refine_tol=1e-10, max_refine=50)
        xe = s.solution()
        print("case: GMRES graphs, k %.1f, stats %s" % (k, st))
        assert st["converged"] == 1
        assert np.array_equal(xg, xe), k
        assert _relcols(xg, _oracle_lstsq(m, k, d)).max() < 1e-5


def test_cgls_graph_replays_equal_eager(monkeypatch):
    """The CG steps run as a captured HIP graph replayed every 8 steps, with
    the host's flag reads and the FP64 residuals between replays.  Every
    solve (three in a row, each replaying the graph many times) returns the
    same bits as the same solve launched eagerly, and the oracle's solution."""
    m = _mesh(3, 8)
    k = 25.0
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP32)
    for seed in (1, 2, 3):
        d = _dirs(12, seed=seed)
        s.set_rhs(bem.RHS_PLANE, d, 1.0)
        monkeypatch.delenv("BEM_NO_GRAPH", raising=False)
        st = s.solve(bem.SOLVE_CGLS, tol=1e-6, maxit=5000, refine_tol=1e-10, max_refine=20)
        xg = s.solution()
        monkeypatch.setenv("BEM_NO_GRAPH", "1")
        s.solve(bem.SOLVE_CGLS, tol=1e-6, maxit=5000, refine_tol=1e-10, max_refine=20)
        xe = s.solution()
        assert st["converged"] == 1 and st["iterations"] >= 16
        assert np.array_equal(xg, xe), seed
        xo = _oracle_lstsq(m, k, d)
        assert _relcols(xg, xo).max() < 1e-5


@pytest.mark.parametrize("method", [bem.SOLVE_CGLS, bem.SOLVE_GMRES])
def test_adjoint_copy_matches_adjoint_kernel(monkeypatch, method):
    """Adjoint products through the conjugate-transposed copy of A (taken for
    more than 16 right-hand sides) against the adjoint kernel on A itself
    (BEM_NO_ADJOINT_COPY), on an s3 sphere with 7 CHIEF rows (M = 649: odd, so
    the copy's last row pair and its last tiles are ragged) and 24 right-hand
    sides.  After FP64 refinement to 1e-10 the two solutions agree to 1e-8, and
    the copy's matches the oracle's least-squares solution at 1e-5."""
    m = _mesh(3, 7)
    assert m.num_pnts + m.num_chief == 649
    k = 31.3
    d = _dirs(24, seed=17)
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP32)
    s.set_rhs(bem.RHS_PLANE, d, 1.0)
    kw = dict(restart=30, tol=1e-6, maxit=5000, refine_tol=1e-10, max_refine=50)
    monkeypatch.delenv("BEM_NO_ADJOINT_COPY", raising=False)
    st = s.solve(method, **kw)
    xc = s.solution()
    monkeypatch.setenv("BEM_NO_ADJOINT_COPY", "1")
    st2 = s.solve(method, **kw)
    xk = s.solution()
    rel = _relcols(xc, xk).max()
    err = _relcols(xc, _oracle_lstsq(m, k, d)).max()
    print("case: adjoint copy, method %d: stats %s / %s, copy vs kernel %.3g, vs oracle %.3g" %
          (method, st, st2, rel, err))
    assert st["converged"] == 1 and st2["converged"] == 1
    assert rel < 1e-8
    assert err < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("nrhs,method", [(24, bem.SOLVE_BLOCK_CG), (24, bem.SOLVE_GMRES), (8, bem.SOLVE_GMRES)])
def test_three_and_four_product_forms_agree(monkeypatch, nrhs, method):
    """The complex products in the three-product form (default) against the
    four-product form (BEM_GEMM_4M): 24 right-hand sides take the 32 x 32
    kernels (the Krylov product with paired X loads, the adjoint through the
    conjugate-transposed copy, the FP64 residuals), 8 the 16 x 16 ones.  On an
    s3 sphere with 7 CHIEF rows (M = 649, ragged tiles), after FP64
    refinement to 1e-10 the two solutions agree to 1e-8, and the default one
    matches the oracle's least-squares solution at 1e-5."""
    m = _mesh(3, 7)
    k = 27.1
    d = _dirs(nrhs, seed=23)
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP32)
    s.set_rhs(bem.RHS_PLANE, d, 1.0)
    kw = dict(restart=30, tol=1e-6, maxit=5000, refine_tol=1e-10, max_refine=50)
    monkeypatch.delenv("BEM_GEMM_4M", raising=False)
    if method == bem.SOLVE_BLOCK_CG:
        # block CG on the FP32 product kernels (the FP16-split ones have one form)
        monkeypatch.setenv("BEM_BLOCK_CG_ALWAYS", "1")
        monkeypatch.setenv("BEM_GEMM_NO_F16", "1")
    st3 = s.solve(method, **kw)
    x3 = s.solution()
    monkeypatch.setenv("BEM_GEMM_4M", "1")
    st4 = s.solve(method, **kw)
    x4 = s.solution()
    rel = _relcols(x3, x4).max()
    err = _relcols(x3, _oracle_lstsq(m, k, d)).max()
    print("case: product forms, %d rhs, method %d: stats %s / %s, 3M vs 4M %.3g, vs oracle %.3g" %
          (nrhs, method, st3, st4, rel, err))
    assert st3["converged"] == 1 and st4["converged"] == 1
    assert rel < 1e-8
    assert err < 1e-5


@pytest.mark.gpu
def test_product_row_tiles_same_bits(monkeypatch):
    """The FP32 32-column product (BEM_GEMM_NO_F16) with two 64-row tiles per
    wave (taken for large launches) against one tile per wave
    (BEM_GEMM_ROW_TILES forces either):
    every element's sum runs in the same order, so a block CG solve with 24
    right-hand sides on an s3 sphere with 7 CHIEF rows (ragged row tiles)
    returns the same bits."""
    m = _mesh(3, 7)
    k = 27.1
    d = _dirs(24, seed=29)
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP32)
    s.set_rhs(bem.RHS_PLANE, d, 1.0)
    monkeypatch.setenv("BEM_BLOCK_CG_ALWAYS", "1")
    monkeypatch.setenv("BEM_GEMM_NO_F16", "1")
    kw = dict(tol=1e-6, maxit=5000, refine_tol=1e-10, max_refine=50)
    out = {}
    for tiles in ("1", "2"):
        monkeypatch.setenv("BEM_GEMM_ROW_TILES", tiles)
        st = s.solve(bem.SOLVE_BLOCK_CG, **kw)
        assert st["converged"] == 1, st
        out[tiles] = (s.solution(), st["iterations"])
    print("case: row tiles, iterations %d / %d" % (out["1"][1], out["2"][1]))
    assert out["1"][1] == out["2"][1]
    assert np.array_equal(out["1"][0], out["2"][0])
