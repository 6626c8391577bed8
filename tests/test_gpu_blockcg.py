"""Block conjugate gradients on the CHIEF normal equations (BEM_SOLVE_BLOCK_CG).

All right-hand sides share one block Krylov space (O'Leary 1980, residual
block orthonormalised by CholeskyQR2 as in Dubrulle 2001), with FP64 residual
refinement.  Checks:
  * each column against the FP64 oracle's least-squares solution (relative
    2-norm error <= 1e-5 for the FP32 matrix, <= 1e-8 for the FP64 matrix), for
    2, 20 and 40 plane waves and the maximum block of 64 point sources;
  * the same solution as the lockstep CG after refinement to 1e-10 (1e-8);
  * fewer block steps than lockstep CG steps for 20 right-hand sides;
  * two equal right-hand sides (a rank-deficient block): the block cannot
    start, the solve falls back to the lockstep CG and still meets the bar;
  * below 4e8 of M N nrhs the method takes the lockstep iteration: the result
    equals BEM_SOLVE_CGLS's bit for bit;
  * the second CholeskyQR pass and the explicit CholeskyQR2 of S, which a step
    takes only when the Gram terms leave orthogonality in doubt, forced at
    every step: the same solution to 1e-9.
These systems are small, so every test but the last sets BEM_BLOCK_CG_ALWAYS
to take the block iteration.
"""
import numpy as np
import pytest

import bem_gpu_clean_cpp_amd as bem
import bemcases as C
import oracle_py as O

pytestmark = pytest.mark.gpu

RHO, SPEED = float(np.float32(1.21)), float(np.float32(343.0))
KW = dict(tol=1e-6, maxit=5000, refine_tol=1e-10, max_refine=20)


@pytest.fixture(autouse=True, scope="module")
def _setup():
    bem.quad_init(6)
    O.quad_init(6)
    bem.set_medium(RHO, SPEED)
    yield
    bem.set_medium(1.2041, 343.21)


@pytest.fixture(autouse=True)
def _always_block(request, monkeypatch):
    if "size_policy" not in request.node.name:
        monkeypatch.setenv("BEM_BLOCK_CG_ALWAYS", "1")


def _mesh(s, nchief, r=0.1, seed=7):
    m = bem.Mesh().icosphere(s, r)
    m.set_chief(C.ball_points(seed, nchief, r / 2))
    return m


def _oracle(m, k, kind, params):
    p, e, bc = C.mesh_arrays(m)
    pts = np.vstack([p, m.chief()]).astype(np.float64)
    q = np.asarray(params, np.float64)
    if kind == bem.RHS_PLANE:
        binit = np.exp(-1j * k * pts @ q.T)
    else:
        r = np.linalg.norm(pts[:, None, :] - q[None, :, :], axis=2)
        binit = np.exp(-1j * k * r) / (4 * np.pi * r)
    A, B = O.assemble(k, pts, p.shape[0], e, bc, binit=binit, dens=RHO, speed=SPEED)
    return np.linalg.lstsq(A, B, rcond=None)[0]


def _relcols(x, y):
    return np.linalg.norm(x - y, axis=0) / np.linalg.norm(y, axis=0)


@pytest.mark.parametrize("R", [2, 20, 40])
def test_block_cg_matches_oracle(R):
    m = _mesh(3, 8)
    k = 31.3
    d = C.sphere_points(11, R, 1.0)
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP32)
    s.set_rhs(bem.RHS_PLANE, d, 1.0)
    st = s.solve(bem.SOLVE_BLOCK_CG, **KW)
    err = _relcols(s.solution(), _oracle(m, k, bem.RHS_PLANE, d))
    print("case: block CG, R %d: stats %s, worst err %.3g" % (R, st, err.max()), flush=True)
    assert st["converged"] == 1 and st["method"] == bem.SOLVE_BLOCK_CG
    assert err.max() < 1e-5


def test_block_cg_fp64_matrix():
    m = _mesh(2, 6)
    k = 20.0
    d = C.sphere_points(2, 9, 1.0)
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP64)
    s.set_rhs(bem.RHS_PLANE, d, 1.0)
    st = s.solve(bem.SOLVE_BLOCK_CG, tol=1e-12, maxit=5000, refine_tol=1e-12, max_refine=20)
    err = _relcols(s.solution(), _oracle(m, k, bem.RHS_PLANE, d))
    print("case: block CG FP64, stats %s, worst err %.3g" % (st, err.max()), flush=True)
    assert st["converged"] == 1
    assert err.max() < 1e-8


def test_block_cg_max_block_point_sources():
    """64 point sources inside the body (the largest block)."""
    m = _mesh(2, 4)
    k = 12.0
    src = C.ball_points(3, 64, 0.04)
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP32)
    s.set_rhs(bem.RHS_POINT, src, 1.0)
    st = s.solve(bem.SOLVE_BLOCK_CG, **KW)
    err = _relcols(s.solution(), _oracle(m, k, bem.RHS_POINT, src))
    print("case: block CG, 64 point sources: stats %s, worst err %.3g" % (st, err.max()), flush=True)
    assert st["converged"] == 1
    assert err.max() < 1e-5


def test_block_cg_equals_lockstep_cg():
    m = _mesh(3, 8)
    k = 31.3
    d = C.sphere_points(5, 20, 1.0)
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP32)
    s.set_rhs(bem.RHS_PLANE, d, 1.0)
    sb = s.solve(bem.SOLVE_BLOCK_CG, **KW)
    xb = s.solution()
    sl = s.solve(bem.SOLVE_CGLS, **KW)
    xl = s.solution()
    rel = _relcols(xb, xl).max()
    print("case: block vs lockstep CG: %s / %s, difference %.3g" % (sb, sl, rel), flush=True)
    assert sb["converged"] == 1 and sl["converged"] == 1
    assert rel < 1e-8
    assert sb["iterations"] < sl["iterations"]


def test_block_cg_equal_right_hand_sides_fall_back():
    """Columns 0 and 3 are the same plane wave: the block is rank deficient."""
    m = _mesh(3, 8)
    k = 25.0
    d = C.sphere_points(9, 6, 1.0)
    d[3] = d[0]
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP32)
    s.set_rhs(bem.RHS_PLANE, d, 1.0)
    st = s.solve(bem.SOLVE_BLOCK_CG, **KW)
    x = s.solution()
    err = _relcols(x, _oracle(m, k, bem.RHS_PLANE, d))
    print("case: block CG with equal columns: stats %s, worst err %.3g" % (st, err.max()), flush=True)
    assert st["converged"] == 1
    assert err.max() < 1e-5
    assert np.array_equal(x[:, 0], x[:, 3])


def test_block_cg_size_policy_small_system_takes_lockstep(monkeypatch):
    """s3 with 8 plane waves (M N nrhs = 3.3e6): the block method runs the
    lockstep CG iteration, so the solution is BEM_SOLVE_CGLS's bit for bit."""
    monkeypatch.delenv("BEM_BLOCK_CG_ALWAYS", raising=False)
    m = _mesh(3, 8)
    k = 31.3
    d = C.sphere_points(11, 8, 1.0)
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP32)
    s.set_rhs(bem.RHS_PLANE, d, 1.0)
    sb = s.solve(bem.SOLVE_BLOCK_CG, **KW)
    xb = s.solution()
    sl = s.solve(bem.SOLVE_CGLS, **KW)
    xl = s.solution()
    print("case: block CG size policy at s3: %s / %s" % (sb, sl), flush=True)
    assert sb["iterations"] == sl["iterations"]
    assert np.array_equal(xb, xl)


@pytest.mark.parametrize("R", [8, 24, 36])
def test_block_cg_fp16_split_products(monkeypatch, R):
    """Block CG's products on the FP16-split kernels (the default with an FP32
    matrix) against the FP32 kernels (BEM_GEMM_NO_F16): per-launch timing
    shows which kernels ran (one product pair per block step); point sources
    at 0.12 m to 120 m from the centre of a 0.1 m sphere give right-hand sides
    whose norms differ by about a thousand, which the per-column scales of the
    split absorb; 8 right-hand sides fill a quarter of a launch group (the
    conjugate-transposed copy is made for them too), 36 take two groups, the
    second ragged.  Both solutions match the oracle's least-squares solution
    at 1e-5 and each other at 1e-9 after FP64 refinement."""
    m = _mesh(3, 7)
    k = 27.1
    u = C.sphere_points(41, R, 1.0)
    src = (u * np.logspace(np.log10(0.12), np.log10(120.0), R)[:, None]).astype(np.float32)
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP32)
    s.set_rhs(bem.RHS_POINT, src, 1.0)
    names = ("gemm_n_f16s", "gemm_hc_f16s", "gemm_n_f32", "gemm_hc_f32")
    out = {}
    try:
        for mode in ("f16", "f32"):
            if mode == "f32":
                monkeypatch.setenv("BEM_GEMM_NO_F16", "1")
            bem.timing_reset()
            bem.timing_enable(True)
            st = s.solve(bem.SOLVE_BLOCK_CG, **KW)
            bem.sync()
            bem.timing_enable(False)
            out[mode] = (st, s.solution(), {n: bem.timing_get(n)[1] for n in names})
    finally:
        bem.timing_enable(False)
        bem.timing_reset()
    ref = _oracle(m, k, bem.RHS_POINT, src)
    (s16, x16, c16), (s32, x32, c32) = out["f16"], out["f32"]
    rel = _relcols(x16, x32).max()
    e16, e32 = _relcols(x16, ref).max(), _relcols(x32, ref).max()
    print("case: FP16-split products, R %d: %s %s / FP32 %s %s, difference %.3g, vs oracle %.3g / %.3g" %
          (R, s16, c16, s32, c32, rel, e16, e32), flush=True)
    assert s16["converged"] == 1 and s32["converged"] == 1
    # (a refinement cycle left with one unconverged system runs lockstep CG on
    # the FP32 kernels, so those may appear in the default solve too; the
    # first residual A^H b is one more FP16-split adjoint product)
    assert c16["gemm_n_f16s"] > 0 and c16["gemm_hc_f16s"] == c16["gemm_n_f16s"] + 1
    assert c32["gemm_n_f16s"] == 0 and c32["gemm_hc_f16s"] == 0 and c32["gemm_n_f32"] > 0
    assert rel < 1e-9
    assert e16 < 1e-5 and e32 < 1e-5


def test_block_cg_fp16_split_impedance(monkeypatch):
    """An impedance boundary condition (complex ca, cb, cc: the h - (ca/cb) g
    terms change the size and phase of A's entries and of the right-hand
    sides) with 12 point sources 0.5 m from the centre: block CG on the
    FP16-split products matches the oracle's least-squares solution at 1e-5
    and the FP32 products at 1e-9."""
    m = _mesh(3, 7)
    m.set_bc(0.5 + 0.25j, 1.0 - 0.5j, 0.3 + 0.1j)
    k = 25.0
    src = C.sphere_points(43, 12, 0.5)
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP32)
    s.set_rhs(bem.RHS_POINT, src, 1.0)
    s16 = s.solve(bem.SOLVE_BLOCK_CG, **KW)
    x16 = s.solution()
    monkeypatch.setenv("BEM_GEMM_NO_F16", "1")
    s32 = s.solve(bem.SOLVE_BLOCK_CG, **KW)
    x32 = s.solution()
    ref = _oracle(m, k, bem.RHS_POINT, src)
    rel = _relcols(x16, x32).max()
    err = _relcols(x16, ref).max()
    print("case: FP16-split products, impedance BC: %s / FP32 %s, difference %.3g, vs oracle %.3g" %
          (s16, s32, rel, err), flush=True)
    assert s16["converged"] == 1 and s32["converged"] == 1
    assert rel < 1e-9
    assert err < 1e-5


@pytest.mark.parametrize("switch,nrhs", [("BEM_BLOCK_CG_TWO_PASS", 20), ("BEM_BLOCK_CG_EXPLICIT", 20),
                                         ("BEM_GRAM_VALU", 20), ("BEM_GRAM_VALU", 40),
                                         ("BEM_GEMM_NO_F16", 20), ("BEM_GEMM_NO_F16", 40)])
def test_block_cg_qr_variants_agree(monkeypatch, switch, nrhs):
    """The second CholeskyQR pass at every step, the explicit QR at every step,
    the Gram products on the VALU kernel instead of MFMA (20 right-hand sides:
    one 16 x 16 tile per wave; 40: four), and the block products on the FP32
    kernels instead of the FP16-split ones (20 right-hand sides: one launch
    group; 40: two) give the default solve's solution to 1e-9 after
    refinement."""
    m = _mesh(3, 8)
    k = 31.3
    d = C.sphere_points(23, nrhs, 1.0)
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP32)
    s.set_rhs(bem.RHS_PLANE, d, 1.0)
    sd = s.solve(bem.SOLVE_BLOCK_CG, **KW)
    xd = s.solution()
    monkeypatch.setenv(switch, "1")
    sv = s.solve(bem.SOLVE_BLOCK_CG, **KW)
    xv = s.solution()
    rel = _relcols(xv, xd).max()
    print("case: block CG with %s: %s / default %s, difference %.3g" % (switch, sv, sd, rel), flush=True)
    assert sv["converged"] == 1 and rel < 1e-9
