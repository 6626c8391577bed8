"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle.

Tolerances:
  * FP32 assembly vs FP64 oracle: max |dA| <= 2e-5 * max |A| (complex FP32
    quadrature; the reference computes in FP32 with fast-math);
  * FP64 assembly vs FP64 oracle: max |dA| <= 1e-11 * max |A|;
  * FP32 solves with FP64 refinement vs FP64 oracle solution: relative 2-norm
    error <= 1e-5 (the accuracy a surface-pressure solve must reach against
    the FP64 solution);
  * repeat assemblies are bitwise identical (no atomics).
"""
import os

import numpy as np
import pytest

import bem_gpu_clean_cpp_amd as bem
import bemcases as C
import oracle_py as O

pytestmark = pytest.mark.gpu

# the C-ABI takes float density / speed; the oracle gets the same rounded values
RHO, SPEED = float(np.float32(1.21)), float(np.float32(343.0))


@pytest.fixture(autouse=True, scope="module")
def _setup():
    bem.quad_init(6)
    O.quad_init(6)
    bem.set_medium(RHO, SPEED)
    yield
    bem.set_medium(1.2041, 343.21)


def _oracle_system(mesh, k, binit=None):
    p, e, bc = C.mesh_arrays(mesh)
    pts = np.vstack([p, mesh.chief()]).astype(np.float64)
    return O.assemble(k, pts, p.shape[0], e, bc, binit=binit, dens=RHO, speed=SPEED)


def _rel_max(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def _icosphere(s, r=0.1, nchief=0, seed=101):
    m = bem.Mesh().icosphere(s, r)
    if nchief:
        m.set_chief(C.ball_points(seed, nchief, r / 2))
    return m


@pytest.mark.parametrize("qg", ["1", "1g", "4"])
@pytest.mark.parametrize("k", [3.13, 36.6])
def test_assembly_fp32_matches_oracle(qg, k, monkeypatch):
    """QG=1 runs the separable fast kernel, "1g" the generic one-lane-per-row
    kernel, QG=4 the kernel that splits the 36 points over 4 lanes per row."""
    monkeypatch.setenv("BEM_QG", qg[0])
    if qg == "1g":
        monkeypatch.setenv("BEM_GENERIC_TILE", "1")
    m = _icosphere(2, nchief=4)
    print("case: icosphere s2 + 4 CHIEF, k=%g, QG=%s" % (k, qg), flush=True)
    A, B = bem.bem_system(m, k)
    Ao, Bo = _oracle_system(m, k)
    assert A.shape == Ao.shape
    err = _rel_max(A, Ao)
    print("FP32 assembly max rel err %.3g" % err)
    assert err < 2e-5


@pytest.mark.parametrize("far", [False, True])
def test_assembly_fp32_phase_range(far, monkeypatch):
    """The fast kernel hands k r / (2 pi) unreduced to v_sin_f32 / v_cos_f32
    while the bounding-box diagonal of the points keeps it within their
    256-revolution range, and reduces it first otherwise.  A CHIEF point 500 m
    away (k = 36.6: about 2900 revolutions) selects the reduced form; both
    forms must match the oracle."""
    monkeypatch.setenv("BEM_QG", "1")
    m = _icosphere(2)
    ch = C.ball_points(101, 3, 0.05)
    if far:
        ch = np.vstack([ch, [[500.0, 0.0, 0.0]]])
    m.set_chief(ch)
    k = 36.6
    print("case: icosphere s2 + %d CHIEF (far point: %s), k=%g" % (len(ch), far, k), flush=True)
    A, B = bem.bem_system(m, k)
    Ao, Bo = _oracle_system(m, k)
    err = _rel_max(A, Ao)
    print("FP32 assembly max rel err %.3g" % err)
    assert err < 2e-5


@pytest.mark.parametrize("imp", [False, True])
def test_assembly_fp32_element_once_plan(imp, monkeypatch):
    """The FP32 fast kernel's element-once plan: every element integrated by
    one column tile, its vertices that other tiles own written to side columns
    that ghost_reduce adds in a fixed order.  Forced on icosphere s3 with 12
    tiles (where it is taken: 258 side columns, and 29 % of the elements
    straddle tiles), it matches the oracle and the halo plan (every straddling
    element integrated once per tile); with ca != 0 as well (the h - (ca/cb) g
    path)."""
    monkeypatch.setenv("BEM_QG", "1")
    monkeypatch.setenv("BEM_TILES", "12")
    m = _icosphere(3, nchief=8)
    srcs = None
    if imp:
        m.set_bc(0.5 + 0.25j, 1.0 - 0.5j, 0.3 + 0.1j)
        srcs = [[0.5, 0.2, 0.1]]
    k = 36.6
    print("case: icosphere s3 + 8 CHIEF, 12 tiles, k=%g, impedance %s" % (k, imp), flush=True)
    bem.timing_reset()
    bem.timing_enable(True)
    A, B = bem.bem_system(m, k, srcs=srcs)
    bem.sync()
    bem.timing_enable(False)
    _, n_ghost = bem.timing_get("ghost_reduce")
    bem.timing_reset()
    assert n_ghost >= 1, "the element-once plan was not taken"
    monkeypatch.setenv("BEM_ASM_HALO", "1")
    Ah, Bh = bem.bem_system(m, k, srcs=srcs)
    p, e, bc = C.mesh_arrays(m)
    pts = np.vstack([p, m.chief()]).astype(np.float64)
    binit = np.zeros((pts.shape[0], 1)) if imp else None
    Ao, Bo = O.assemble(k, pts, p.shape[0], e, bc, binit=binit, dens=RHO, speed=SPEED)
    err, dh = _rel_max(A, Ao), _rel_max(A, Ah)
    print("vs oracle %.3g, vs halo plan %.3g" % (err, dh))
    assert err < 2e-5
    assert dh < 1e-6
    if imp:
        assert _rel_max(B, Bo) < 2e-5


def test_assembly_fp32_element_once_plan_at_s5(monkeypatch):
    """At icosphere s5 + 16 CHIEF (10258 x 10242) the default launch takes the
    element-once plan (204 column tiles, 4348 side columns).  Four blocks of
    64 rows, one of them over the CHIEF rows, match the oracle's FP64 rows
    (2e-5 of max |A|) and the halo plan's rows (1e-6)."""
    for v in ("BEM_QG", "BEM_TILES", "BEM_ASM_HALO", "BEM_GENERIC_TILE", "BEM_NO_GHOST_PLAN"):
        monkeypatch.delenv(v, raising=False)
    m = _icosphere(5, nchief=16)
    k = 2 * np.pi * 2000.0 / SPEED
    N, M = m.num_pnts, m.num_pnts + m.num_chief
    print("case: icosphere s5 + 16 CHIEF (M %d, N %d), k=%g" % (M, N, k), flush=True)
    s = bem.Solver(m)
    bem.timing_reset()
    bem.timing_enable(True)
    s.assemble(k, bem.PREC_FP32)
    bem.sync()
    bem.timing_enable(False)
    _, n_ghost = bem.timing_get("ghost_reduce")
    bem.timing_reset()
    assert n_ghost >= 1, "the element-once plan was not taken"
    starts = (0, 3000, 6100, M - 64)
    rows = np.concatenate([np.arange(r, r + 64) for r in starts])
    a = s.rows(rows)
    monkeypatch.setenv("BEM_ASM_HALO", "1")
    s.assemble(k, bem.PREC_FP32)
    ah = s.rows(rows)
    s.close()
    p, e, bc = C.mesh_arrays(m)
    pts = np.vstack([p, m.chief()]).astype(np.float64)
    ao = np.vstack([O.assemble_rows(k, pts, N, e, bc, r, r + 64, dens=RHO, speed=SPEED, dtype=np.float64,
                                    nthreads=min(16, os.cpu_count() or 1)) for r in starts])
    err, dh = _rel_max(a, ao), _rel_max(a, ah)
    print("rows vs oracle %.3g, vs halo plan %.3g" % (err, dh))
    assert err < 2e-5
    assert dh < 1e-6


def test_assembly_fp64_matches_oracle():
    m = _icosphere(2, nchief=4)
    s = bem.Solver(m)
    k = 20.0
    s.assemble(k, bem.PREC_FP64)
    s.set_rhs(bem.RHS_PLANE, [[0, 0, 1]], 1.0)
    A, B = s.system()  # complex64 copy of the FP64 matrix
    Ao, _ = _oracle_system(m, k)
    assert _rel_max(A, Ao) < 1e-6  # limited by the complex64 copy-out
    # the FP64 solve checks the FP64 matrix to FP64 accuracy
    s.solve(bem.SOLVE_QR)
    x = s.solution()
    p = m.pnts().astype(np.float64)
    binit = np.exp(-1j * k * np.vstack([p, m.chief()])[:, 2:3].astype(np.float64))
    Ao, Bo = _oracle_system(m, k, binit=binit)
    xo = np.linalg.lstsq(Ao, Bo, rcond=None)[0]
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-10


def test_assembly_reference_mesh_and_bc():
    """sphere.obj with the reference's default bc (0,1,-1) and 1 CHIEF point."""
    m = bem.Mesh().read_obj(os.path.join(C.GOLDEN, "sphere.obj"))
    m.set_chief([[0.1, -0.2, 0.05]])
    k = float(np.float32(2 * np.pi * 171.5 / 343.21))
    bem.set_medium(1.2041, 343.21)
    try:
        A, B = bem.bem_system(m, k, srcs=[[1e6, 1e6, 1e6]])
        p, e, bc = C.mesh_arrays(m)
        pts = np.vstack([p, m.chief()]).astype(np.float64)
        Ao, Bo = O.assemble(k, pts, p.shape[0], e, bc, binit=np.zeros((pts.shape[0], 1)),
                            dens=float(np.float32(1.2041)), speed=float(np.float32(343.21)))
        assert _rel_max(A, Ao) < 2e-5
        assert _rel_max(B, Bo) < 2e-5
    finally:
        bem.set_medium(RHO, SPEED)


@pytest.mark.parametrize("qg", ["1", "4"])
def test_assembly_impedance_bc(qg, monkeypatch):
    """ca != 0 exercises the h - (ca/cb) g path of the kernels."""
    monkeypatch.setenv("BEM_QG", qg)
    m = _icosphere(2, nchief=2)
    m.set_bc(0.5 + 0.25j, 1.0 - 0.5j, 0.3 + 0.1j)
    k = 12.0
    A, B = bem.bem_system(m, k, srcs=[[0.5, 0.2, 0.1]])
    p, e, bc = C.mesh_arrays(m)
    pts = np.vstack([p, m.chief()]).astype(np.float64)
    Ao, Bo = O.assemble(k, pts, p.shape[0], e, bc, binit=np.zeros((pts.shape[0], 1)), dens=RHO, speed=SPEED)
    assert _rel_max(A, Ao) < 2e-5
    assert _rel_max(B, Bo) < 2e-5


@pytest.mark.parametrize("qg", ["1", "4"])
def test_assembly_bitwise_repeatable(qg, monkeypatch):
    monkeypatch.setenv("BEM_QG", qg)
    m = _icosphere(3, nchief=8)
    A1, _ = bem.bem_system(m, 30.0)
    A2, _ = bem.bem_system(m, 30.0)
    a1 = np.ascontiguousarray(A1).view(np.uint32)
    a2 = np.ascontiguousarray(A2).view(np.uint32)
    assert np.array_equal(a1, a2)


def test_point_source_rhs_strength():
    """B = q G(x, src) - BC term (atomicFuncs.cu:238-242) with q set."""
    m = _icosphere(2)
    k = 10.0
    src = np.array([[0.3, -0.1, 0.2]], np.float32)
    bem.set_source_strength(2.5)
    try:
        A, B = bem.bem_system(m, k, srcs=src)
    finally:
        bem.set_source_strength(0.0)
    p = m.pnts().astype(np.float64)
    r = np.linalg.norm(p - src.astype(np.float64), axis=1)
    inc = 2.5 * np.exp(-1j * k * r) / (4 * np.pi * r)
    _, Bo = _oracle_system(m, k, binit=inc[:, None])
    assert _rel_max(B, Bo) < 2e-5


def test_qr_solve_matches_lapack():
    rng = np.random.default_rng(1)
    m, n = 300, 260
    A = (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))).astype(np.complex64)
    A += 20 * np.eye(m, n, dtype=np.complex64)
    B = (rng.standard_normal((m, 2)) + 1j * rng.standard_normal((m, 2))).astype(np.complex64)
    X = bem.qr_solve(A, B)
    Xo = np.linalg.lstsq(A.astype(np.complex128), B.astype(np.complex128), rcond=None)[0]
    assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) < 1e-5


def test_qr_panel_variants_match_lapack(monkeypatch):
    """1100 x 1040: the first panels (over 1024 rows) run one launch per
    column over row-split workgroups, then the register panels with 2 rows per
    thread (<= 1024 rows), 1 row per thread (<= 512, <= 256), and the last 16
    columns the one-workgroup global-memory panel.  Against LAPACK, and
    against the one-workgroup panel for every panel (BEM_QR_GPANEL)."""
    rng = np.random.default_rng(7)
    m, n = 1100, 1040
    A = (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))).astype(np.complex64)
    A += 30 * np.eye(m, n, dtype=np.complex64)
    B = (rng.standard_normal((m, 3)) + 1j * rng.standard_normal((m, 3))).astype(np.complex64)
    X = bem.qr_solve(A, B)
    Xo = np.linalg.lstsq(A.astype(np.complex128), B.astype(np.complex128), rcond=None)[0]
    assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) < 2e-5
    monkeypatch.setenv("BEM_QR_GPANEL", "1")
    Xg = bem.qr_solve(A, B)
    assert np.linalg.norm(Xg - Xo) / np.linalg.norm(Xo) < 2e-5
    assert np.linalg.norm(X - Xg) / np.linalg.norm(Xo) < 2e-5
    # the one-workgroup solve (Q^H b and back substitution) against the
    # row-split launches
    monkeypatch.setenv("BEM_QR_ONE_WG_SOLVE", "1")
    Xn = bem.qr_solve(A, B)
    assert np.linalg.norm(Xn - Xo) / np.linalg.norm(Xo) < 2e-5
    assert np.linalg.norm(X - Xn) / np.linalg.norm(Xo) < 2e-5


def test_qr_fp64_one_workgroup_solve_matches_row_split(monkeypatch):
    """FP64 QR of a CHIEF system with a ragged last panel (N = 162): the
    one-workgroup solve against the row-split launches, to FP64 accuracy."""
    m = _icosphere(2, nchief=6)
    k = 17.0
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP64)
    s.set_rhs(bem.RHS_PLANE, [[0, 0, 1], [0.6, 0.8, 0], [0, 1, 0]], 1.0)
    s.solve(bem.SOLVE_QR)
    x = s.solution()
    monkeypatch.setenv("BEM_QR_ONE_WG_SOLVE", "1")
    s.solve(bem.SOLVE_QR)
    xn = s.solution()
    assert np.linalg.norm(x - xn) / np.linalg.norm(xn) < 1e-12


def test_qr_launch_per_panel_solve_many_rhs(monkeypatch):
    """More than 64 right-hand sides take the multi-workgroup solve (one launch
    per panel for Q^H B, one per block column for R X = Q^H B): against LAPACK
    and against the per-column kernels (BEM_QR_PER_COLUMN_SOLVE)."""
    rng = np.random.default_rng(11)
    m, n, r = 300, 261, 70
    A = (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))).astype(np.complex64)
    A += 20 * np.eye(m, n, dtype=np.complex64)
    B = (rng.standard_normal((m, r)) + 1j * rng.standard_normal((m, r))).astype(np.complex64)
    X = bem.qr_solve(A, B)
    Xo = np.linalg.lstsq(A.astype(np.complex128), B.astype(np.complex128), rcond=None)[0]
    assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) < 1e-5
    monkeypatch.setenv("BEM_QR_PER_COLUMN_SOLVE", "1")
    Xc = bem.qr_solve(A, B)
    assert np.linalg.norm(X - Xc) / np.linalg.norm(Xo) < 1e-5


def test_qr_launch_per_panel_solve_tall(monkeypatch):
    """6400 x 6300 (past the one-workgroup solve's 6144 rows), 2 right-hand
    sides: the normal-equation residual A^H (b - A x) against the right-hand
    side's A^H b, and the per-column kernels' solution."""
    rng = np.random.default_rng(12)
    m, n = 6400, 6300
    A = (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))).astype(np.complex64)
    A += 80 * np.eye(m, n, dtype=np.complex64)
    B = (rng.standard_normal((m, 2)) + 1j * rng.standard_normal((m, 2))).astype(np.complex64)
    print("case: QR least squares 6400 x 6300, 2 right-hand sides", flush=True)
    X = bem.qr_solve(A, B)
    A64, B64, X64 = A.astype(np.complex128), B.astype(np.complex128), X.astype(np.complex128)
    res = A64.conj().T @ (B64 - A64 @ X64)
    ref = A64.conj().T @ B64
    rel = np.linalg.norm(res) / np.linalg.norm(ref)
    monkeypatch.setenv("BEM_QR_PER_COLUMN_SOLVE", "1")
    Xc = bem.qr_solve(A, B)
    diff = np.linalg.norm(X - Xc) / np.linalg.norm(Xc)
    print("normal-equation residual %.3e, multi-workgroup vs per-column solve %.3e (bar 1e-5)" % (rel, diff))
    assert rel < 1e-5
    assert diff < 1e-5


def test_qr_fp64_multi_workgroup_solve_matches_per_column(monkeypatch):
    """FP64 QR of a two-sphere CHIEF system with M = 5140 rows (past the
    one-workgroup FP64 solve's 3072) and 3 right-hand sides: the
    multi-workgroup solve against the per-column kernels, to FP64 accuracy."""
    m = C.two_spheres(bem, 4, 0.1)
    m.set_chief(C.ball_points(29, 16, 0.05) - [0.125, 0, 0])
    print("case: FP64 QR, two s4 spheres, N %d M %d" % (m.num_pnts, m.num_pnts + m.num_chief), flush=True)
    s = bem.Solver(m)
    s.assemble(9.0, bem.PREC_FP64)
    s.set_rhs(bem.RHS_PLANE, [[0, 0, 1], [0.6, 0.8, 0], [1, 0, 0]], 1.0)
    s.solve(bem.SOLVE_QR)
    x = s.solution()
    monkeypatch.setenv("BEM_QR_PER_COLUMN_SOLVE", "1")
    s.solve(bem.SOLVE_QR)
    xc = s.solution()
    assert np.linalg.norm(x - xc) / np.linalg.norm(xc) < 1e-12


def test_lsq_solve_returns_r_and_q():
    rng = np.random.default_rng(2)
    m, n = 90, 70
    A = (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))).astype(np.complex64)
    B = (rng.standard_normal((m, 1)) + 1j * rng.standard_normal((m, 1))).astype(np.complex64)
    X, R, Q = bem.lsq_solve(A, B, want_q=True)
    assert np.allclose(np.tril(R, -1), 0)
    assert np.linalg.norm(Q.conj().T @ Q - np.eye(m)) < 1e-4
    assert np.linalg.norm(Q[:, :n] @ R[:n] - A) / np.linalg.norm(A) < 1e-5
    Xo = np.linalg.lstsq(A.astype(np.complex128), B.astype(np.complex128), rcond=None)[0]
    assert np.linalg.norm(X - Xo) / np.linalg.norm(Xo) < 1e-5


def test_gmres_square_matches_lapack():
    m = _icosphere(3)
    k = 20.0
    A, _ = bem.bem_system(m, k)
    p = m.pnts().astype(np.float64)
    b = np.exp(-1j * k * p[:, 2]).astype(np.complex64)
    x = bem.gmres(A, b, 200, 1e-6)
    xo = np.linalg.solve(A.astype(np.complex128), b.astype(np.complex128))
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 2e-5


def test_gmres_reference_main_call_pattern():
    """GMRES(A, B, N, N, 0, x) as the reference's main.cpp:86 calls it: the
    square s3 system from bemSystem, maxIter = N, threshold 0, so the Krylov
    space grows to N steps without a restart (GMRES.cu:145-279), against
    LAPACK on the same FP32 matrix."""
    m = _icosphere(3)
    N = m.num_pnts
    k = float(np.float32(2 * np.pi * 171.5 / SPEED * 8))
    bem.set_source_strength(1.0)
    try:
        A, B = bem.bem_system(m, k, [[0.0, 0.0, 0.02]])
    finally:
        bem.set_source_strength(0.0)
    b = B[:, 0]
    x = bem.gmres(A, b, N, 0.0)
    xo = np.linalg.solve(A.astype(np.complex128), b.astype(np.complex128))
    err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    print("case: GMRES(A, B, %d, %d, 0, x), error vs LAPACK %.3e" % (N, N, err))
    assert err < 1e-5


def test_gen_system_with_chief_rows_matches_oracle():
    """atomicGenSystem (atomicFuncs.cu:218-298) on raw arrays: the nodes and
    the CHIEF points in one pnts array (numNods then numCHIEF), two point
    sources of strength 1; A and B against the FP64 oracle, and bit-identical
    to bemSystem on the same mesh."""
    import ctypes
    m = _icosphere(3, nchief=8, seed=77)
    m.set_bc(0, 1, 0)
    N, Cn = m.num_pnts, m.num_chief
    M = N + Cn
    k = float(np.float32(25.0))
    src = np.array([[0.0, 0.02, 0.01], [0.3, -0.2, 0.1]], np.float32)
    pnts = np.ascontiguousarray(np.vstack([m.pnts(), m.chief()]), np.float32)
    el = m.elems()
    A = np.zeros((N, M), np.complex64)  # column-major M x N
    B = np.zeros((2, M), np.complex64)
    bem.set_source_strength(1.0)
    try:
        rc = bem.lib().bem_gen_system(ctypes.c_float(k), bem._ptr(el), el.shape[0], bem._ptr(pnts), N, Cn,
                                      bem._ptr(src), 2, bem._ptr(A), M, bem._ptr(B), M)
        assert rc == 0, bem.last_error()
        As, Bs = bem.bem_system(m, k, src)
    finally:
        bem.set_source_strength(0.0)
    A, B = A.T, B.T
    assert np.array_equal(A, As) and np.array_equal(B, Bs)
    p, e, bc = C.mesh_arrays(m)
    pts = pnts.astype(np.float64)
    d = np.linalg.norm(pts[:, None, :] - src.astype(np.float64)[None, :, :], axis=2)
    binit = np.exp(-1j * k * d) / (4 * np.pi * d)
    Ao, Bo = O.assemble(k, pts, N, e, bc, binit=binit, dens=RHO, speed=SPEED)
    ea, eb = _rel_max(A, Ao), _rel_max(B, Bo)
    print("case: atomicGenSystem, N %d + %d CHIEF, 2 sources: A err %.3g, B err %.3g" % (N, Cn, ea, eb))
    assert ea < 2e-5 and eb < 2e-5


@pytest.mark.parametrize("method", [bem.SOLVE_GMRES, bem.SOLVE_QR])
def test_refined_solve_chief_1e5(method):
    """FP32 assembly + FP32 solve + FP64 refinement vs the FP64 oracle path."""
    m = _icosphere(3, nchief=8)
    k = 2 * np.pi * 171 / SPEED * 10  # ka ~ 3.1
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP32)
    s.set_rhs(bem.RHS_PLANE, [[0, 0, 1], [0.6, 0.0, 0.8]], 1.0)
    st = s.solve(method, restart=30, tol=1e-6, maxit=2000, refine_tol=1e-10, max_refine=50)
    x = s.solution()
    pts = np.vstack([m.pnts(), m.chief()]).astype(np.float64)
    d = np.array([[0, 0, 1], [0.6, 0.0, 0.8]])
    binit = np.exp(-1j * k * pts @ d.T)
    Ao, Bo = _oracle_system(m, k, binit=binit)
    xo = np.linalg.lstsq(Ao, Bo, rcond=None)[0]
    err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    print("method %d stats %s err %.3g" % (method, st, err))
    assert st["converged"] == 1
    assert err < 1e-5


def test_fp64_gmres_plain():
    m = _icosphere(3, nchief=4)
    k = 15.0
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP64)
    s.set_rhs(bem.RHS_POINT, [[0.3, 0.0, 0.0]], 1.0)
    st = s.solve(bem.SOLVE_GMRES, restart=30, tol=1e-12, maxit=3000, refine_tol=0, max_refine=0)
    x = s.solution()
    pts = np.vstack([m.pnts(), m.chief()]).astype(np.float64)
    r = np.linalg.norm(pts - np.array([0.3, 0.0, 0.0]), axis=1)
    Ao, Bo = _oracle_system(m, k, binit=(np.exp(-1j * k * r) / (4 * np.pi * r))[:, None])
    xo = np.linalg.lstsq(Ao, Bo, rcond=None)[0]
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-9, st


def test_field_matches_oracle():
    m = _icosphere(2, nchief=2)
    k = 8.0
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP32)
    dirs = [[0, 0, 1], [1, 0, 0], [0, 0.6, 0.8]]
    s.set_rhs(bem.RHS_PLANE, dirs, 1.0)
    s.solve(bem.SOLVE_QR)
    x = s.solution()
    xs = C.sphere_points(5, 64, 1.0)
    out = s.field(xs)
    p, e, bc = C.mesh_arrays(m)
    inc = np.exp(-1j * k * xs.astype(np.float64) @ np.array(dirs, float).T)
    fo = O.field(k, p.astype(np.float64), p.shape[0], e, bc, xs, x, inc=inc, dens=RHO, speed=SPEED)
    assert np.max(np.abs(out - fo)) / np.max(np.abs(fo)) < 2e-5


def test_ext_pressure_single_point():
    m = _icosphere(2)
    k = 5.0
    rng = np.random.default_rng(3)
    sp = (rng.standard_normal(m.num_pnts) + 1j * rng.standard_normal(m.num_pnts)).astype(np.complex64)
    x = [0.0, 0.0, 0.7]
    got = bem.ext_pressure(k, m, [1e6, 1e6, 1e6], x, sp)
    p, e, bc = C.mesh_arrays(m)
    fo = O.field(k, p.astype(np.float64), p.shape[0], e, bc, np.array([x]), sp, dens=RHO, speed=SPEED)
    assert abs(got - fo[0, 0]) / abs(fo[0, 0]) < 2e-5


def test_pulsating_sphere_reference_golden():
    """The reference's CHIEF golden file (sphere.obj, f = 171.5 Hz, 1 CHIEF point)."""
    bem.set_medium(1.2041, 343.21)
    try:
        m = bem.Mesh().read_obj(os.path.join(C.GOLDEN, "sphere.obj"))
        m.find_bb(1e-4).gen_chief(1, 0.1, seed=1)
        k = float(np.float32(2 * np.pi * 171.5 / 343.21))
        s = bem.Solver(m)
        s.assemble(k, bem.PREC_FP32)
        s.set_rhs(bem.RHS_POINT, [[1e6, 1e6, 1e6]], 0.0)
        s.solve(bem.SOLVE_QR)
        x = s.solution()
        r = np.arange(1.0, 3.01, 0.2)
        xs = np.array([[-(1 + 0.2 * i), 0, 0] for i in range(1, 11)], np.float32)
        p = np.concatenate([[x[0, 0]], s.field(xs)[:, 0]])
        gold = C.read_golden("sphere171Hz_CHIEF")
        ana = C.pulsating_sphere(r, 1.0, k, 1.2041, 343.21)
        e_gold = np.linalg.norm(p - gold) / np.linalg.norm(gold)
        e_ana = np.linalg.norm(p - ana) / np.linalg.norm(ana)
        print("vs golden %.4f, vs analytic %.4f" % (e_gold, e_ana))
        assert e_gold < 0.015
        assert e_ana < 0.025
    finally:
        bem.set_medium(RHO, SPEED)


def test_rigid_sphere_series_s3_171hz():
    """Plain small case: icosphere s3, r = 0.1 m, 171 Hz, 8 CHIEF, plane wave +z."""
    a = 0.1
    m = _icosphere(3, a, nchief=8)
    m.set_bc(0, 1, 0)  # rigid: zero normal velocity
    k = 2 * np.pi * 171 / SPEED
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP32)
    s.set_rhs(bem.RHS_PLANE, [[0, 0, 1]], 1.0)
    s.solve(bem.SOLVE_QR)
    x = s.solution()[:, 0]
    p = m.pnts().astype(np.float64)
    ser = C.rigid_sphere_surface(a, k, p[:, 2] / np.linalg.norm(p, axis=1))
    err = np.linalg.norm(x - ser) / np.linalg.norm(ser)
    print("rigid sphere series error %.4g" % err)
    assert err < 0.005


def test_autotune_assembly_keeps_parity():
    """The launch-shape autotuner picks a shape for the mesh's size class; the
    assembly with it still matches the oracle."""
    m = _icosphere(3, nchief=8)
    k = 20.0
    try:
        r = bem.autotune_assembly(m, k)
        assert r["lanes_per_row"] in (1, 4) and r["waves"] > 0
        assert 0 < r["best_ms"] <= r["default_ms"] * 1.0001 or r["default_ms"] < 0
        s = bem.Solver(m)
        s.assemble(k, bem.PREC_FP32)
        A, _ = s.system()
        Ao, _ = _oracle_system(m, k)
        assert _rel_max(A, Ao) < 2e-5
    finally:
        bem.tune_clear()
