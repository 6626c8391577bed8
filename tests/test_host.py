"""The C-ABI library on the CPU (no GPU compute): it loads, exports every symbol
include/bem.h declares, and its host-side mesh front-end and quadrature tables
behave as the reference's (OBJ reader, bounding box, CHIEF generator) and as the
oracle."""
import os
import re
import tempfile

import numpy as np
import pytest

import bem_gpu_clean_cpp_amd as bem
import bemcases as C
import oracle_py as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    src = open(os.path.join(ROOT, "include", "bem.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bem_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    names = _header_functions()
    assert len(names) >= 40
    L = bem.lib()
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(bem.EXPORTED)
    assert b"gfx950" in L.bem_version()


def test_struct_layout_matches_reference():
    import ctypes
    assert ctypes.sizeof(bem.bem_complex) == 8      # cuFloatComplex
    assert ctypes.sizeof(bem.bem_cart_coord) == 12  # cartCoord
    assert ctypes.sizeof(bem.bem_tri_elem) == 40    # triElem
    assert bem.bem_tri_elem.bc.offset == 16


def test_quadrature_tables():
    bem.quad_init(6)
    x, w = bem.quad_get(6)
    xr, wr = np.polynomial.legendre.leggauss(6)
    assert np.allclose(x, xr, atol=1e-14) and np.allclose(w, wr, atol=1e-14)
    with pytest.raises(bem.BemError):
        bem.quad_init(5)  # kernels are built for the 6-point rule (INTORDER)


@pytest.mark.parametrize("name", ["sphere.obj", "sphere1.obj", "cube.obj", "test.obj"])
def test_read_obj_matches_reference_semantics(name):
    path = os.path.join(C.GOLDEN, name)
    m = bem.Mesh().read_obj(path)
    p, e = O.read_obj(path)
    assert np.array_equal(m.pnts(), p)
    el = m.elems()
    assert np.array_equal(el["nodes"], e)
    assert np.all(el["bc"] == np.array([0, 1, -1], np.complex64))  # mesh.cu:1240-1242


def test_read_obj_robust_lines():
    txt = "# c\nv 0 0 0\nvn 0 0 1\nvt 0.5 0.5\nv 1 0 0\nv 0 1 0\nv 0 0 1\nf 1/1/1 3/2/1 2/3/1\nf -4 -2 -1\nf 1 2 3 4\n"
    with tempfile.NamedTemporaryFile("w", suffix=".obj", delete=False) as f:
        f.write(txt)
    try:
        m = bem.Mesh().read_obj(f.name)
        assert m.num_pnts == 4  # vn / vt are not vertices
        assert m.num_elems == 4  # quad fanned into two triangles
        assert list(m.elems()["nodes"][1]) == [0, 2, 3]
    finally:
        os.unlink(f.name)


def test_mesh_errors_are_reported():
    with pytest.raises(bem.BemError, match="open"):
        bem.Mesh().read_obj("/nonexistent/x.obj")
    with pytest.raises(bem.BemError, match="outside"):
        bem.Mesh().set([[0, 0, 0], [1, 0, 0], [0, 1, 0]], [[0, 1, 5]])
    with pytest.raises(bem.BemError):
        bem.Mesh().find_bb(0.1)  # empty mesh (mesh.cu:1405-1407)
    with pytest.raises(bem.BemError):
        bem.set_medium(-1.0, 343.0)
    m = bem.Mesh().icosphere(1, 1.0)
    with pytest.raises(bem.BemError):
        m.set_bc(0, 0, 1)  # cb = 0 would divide by zero


def test_reference_entry_points_reject_bad_arguments():
    """The reference entry points return EXIT_FAILURE with a message on bad
    arguments (numerical.h:60-122), and do so before any device work, so these
    calls run without a GPU."""
    L = bem.lib()
    m = bem.Mesh().icosphere(1, 1.0)  # N = 42, no CHIEF points, so M = 42
    A = np.zeros((42, 42), np.complex64)
    B = np.zeros(42, np.complex64)
    src = np.zeros(3, np.float32)
    P = bem._ptr

    def fails(rc, pattern):
        assert rc == 1
        assert re.search(pattern, bem.last_error()), bem.last_error()

    # bemSystem (mesh.cu:985-1006)
    empty = bem.Mesh()  # held: a temporary would be destroyed before the call
    fails(L.bem_system(empty.handle, 1.0, None, 0, P(A), 42, None, 42), r"Not enough mesh information! \(0 nodes")
    fails(L.bem_system(m.handle, 1.0, None, 0, P(A), 41, None, 42), r"lda \(41\) must be >= M \(42\)")
    fails(L.bem_system(m.handle, 1.0, P(src), 1, P(A), 42, P(B), 41), "bad B / sources")
    fails(L.bem_system(m.handle, 1.0, P(src), -1, P(A), 42, P(B), 42), "bad B / sources")
    # atomicGenSystem (atomicFuncs.cu:218-298)
    fails(L.bem_gen_system(1.0, None, 0, None, 0, 0, None, 0, P(A), 42, P(B), 42), "bad system arguments")
    # qrSolver (numerical.cu:580-644) and lsqSolver (numerical.cu:506-554) need more rows than columns
    fails(L.bem_qr_solve(P(A), 4, 5, 4, P(B), 1, 4), "Senario not included yet")
    fails(L.bem_qr_solve(P(A), 6, 5, 5, P(B), 1, 6), "bad QR arguments")
    fails(L.bem_lsq_solve(P(A), 5, 5, 5, P(B), 1, 5, None), "lsqSolver needs m > n")
    # GMRES (GMRES.cu:145-279)
    fails(L.bem_gmres(P(A), P(B), 0, 10, 1e-6, P(B)), "bad GMRES arguments")
    fails(L.bem_gmres(None, P(B), 42, 10, 1e-6, P(B)), "bad GMRES arguments")
    # the device-resident extension checks its handles and arguments the same way
    fails(L.bem_solver_assemble(None, 1.0, bem.PREC_FP32), "null solver")
    fails(L.bem_solver_set_rhs(None, bem.RHS_POINT, P(src), 1, 1.0), "null solver")
    fails(L.bem_solver_solve(None, bem.SOLVE_GMRES, 30, 1e-6, 100, 1e-9, 5, None), "null solver")
    fails(L.bem_solver_get_solution(None, P(B), 42), "null argument")
    fails(L.bem_batch_solve(m.handle, None, 0, bem.PREC_FP32, bem.RHS_POINT, None, 1, 0, 1.0, None, 1, None, 42,
                            None), "bad batch arguments")
    fails(L.bem_mesh_get_bb(m.handle, P(np.zeros(6, np.float32))), "bounding box not computed")
    fails(L.bem_mesh_set_chief(m.handle, None, 2), "bad CHIEF arguments")


@pytest.mark.parametrize("s", [0, 1, 2, 3, 4])
def test_icosphere(s):
    a = 0.1
    m = bem.Mesh().icosphere(s, a)
    assert m.num_pnts == 10 * 4 ** s + 2 and m.num_elems == 20 * 4 ** s
    p = m.pnts().astype(np.float64)
    assert np.allclose(np.linalg.norm(p, axis=1), a, rtol=1e-6)
    assert m.signed_volume() < 0  # normals into the body, like the reference meshes
    if s >= 3:
        assert abs(m.surf_area() - 4 * np.pi * a * a) / (4 * np.pi * a * a) < 0.01


def test_perturbed_ellipsoid():
    ax = (0.075, 0.09, 0.11)
    m = bem.Mesh().ellipsoid(4, ax, nbumps=8, lmax=6, amp=0.05, seed=3)
    assert m.num_pnts == 2562
    vol = -m.signed_volume()
    ref = 4 / 3 * np.pi * np.prod(ax)
    assert abs(vol - ref) / ref < 0.1
    m2 = bem.Mesh().ellipsoid(4, ax, nbumps=8, lmax=6, amp=0.05, seed=3)
    assert np.array_equal(m.pnts(), m2.pnts())


def test_find_bb_and_chief_generator():
    a = 0.1
    m = bem.Mesh().icosphere(3, a)
    m.find_bb(1e-4)
    bb = m.bb()
    assert np.allclose(bb, [-a - 1e-4, a + 1e-4] * 3, atol=1e-6)
    m.gen_chief(8, 0.03, seed=5)
    c = m.chief()
    assert c.shape == (8, 3)
    p = m.pnts()
    for x in c:
        assert np.min(np.linalg.norm(p - x, axis=1)) >= 0.03
        assert np.linalg.norm(x) < a  # inside the sphere
    m2 = bem.Mesh().icosphere(3, a).find_bb(1e-4).gen_chief(8, 0.03, seed=5)
    assert np.array_equal(m2.chief(), c)
    with pytest.raises(bem.BemError):
        bem.Mesh().icosphere(2, a).find_bb(1e-4).gen_chief(4, 0.5, seed=1)  # impossible threshold


def test_wrt_cplx_mat_golden_format():
    v = np.array([[1.5 + 2.25j, -3.125 + 0.5j]])
    with tempfile.NamedTemporaryFile("r", suffix=".txt", delete=False) as f:
        name = f.name
    try:
        bem.wrt_cplx_mat(v, name)
        txt = open(name).read()
        assert txt == "(1.500000,2.250000) (-3.125000,0.500000) \n"
        assert np.allclose(C.read_golden(os.path.basename(name)) if False else
                           np.array([complex(float(a), float(b)) for a, b in
                                     re.findall(r"\(([-\d.]+),([-\d.]+)\)", txt)]), v[0])
    finally:
        os.unlink(name)


def test_python_binding_covers_every_declared_symbol():
    assert set(_header_functions()) <= set(bem.EXPORTED)


def test_request_footprint_bounds_the_allocations():
    """The HBM footprint used for admission covers A and grows with the
    right-hand sides and the Krylov basis; QR keeps a factored copy of A, and
    GMRES on the normal equations with one right-hand side a row-major one."""
    M, N, E = 10258, 10242, 20480
    ld = (M + 63) // 64 * 64
    f1 = bem.request_footprint(M, N, E, 1)
    f2 = bem.request_footprint(M, N, E, 2)
    f32 = bem.request_footprint(M, N, E, 32)
    assert f1 >= 16.0 * ld * N  # A and its row-major copy
    assert f2 >= 8.0 * ld * N
    assert f32 > f2 + 16.0 * ld * 30 * 31  # 30 more FP64 bases of 31 vectors
    assert bem.request_footprint(M, N, E, 1, restart=100) > f1
    assert bem.request_footprint(M, N, E, 1, method=bem.SOLVE_CGLS) < f1 - 8.0 * M * N
    assert bem.request_footprint(M, N, E, 1, method=bem.SOLVE_QR) >= 16.0 * ld * N
    assert bem.request_footprint(M, N, E, 1, precision=bem.PREC_FP64) >= 16.0 * ld * N
    assert bem.request_footprint(N - 1, N, E, 1) < 0  # underdetermined: refused


def test_tune_table_roundtrip(tmp_path):
    src = tmp_path / "tune.txt"
    src.write_text("26 29 1 16384\n22 25 4 8192\nbad line\n")
    bem.tune_clear()
    try:
        bem.tune_load(src)
        out = tmp_path / "out.txt"
        bem.tune_save(out)
        rows = sorted(tuple(map(int, l.split())) for l in out.read_text().splitlines())
        assert rows == [(22, 25, 4, 8192), (26, 29, 1, 16384)]
        with pytest.raises(bem.BemError):
            bem.tune_load(tmp_path / "missing.txt")
    finally:
        bem.tune_clear()
