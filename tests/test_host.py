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
    assert not L.bem_batch_create(None, 2) and "null mesh" in bem.last_error()
    assert not L.bem_batch_create(m.handle, 0) and "workers" in bem.last_error()
    fails(L.bem_batch_run(None, None, 0, bem.PREC_FP32, bem.RHS_POINT, None, 1, 0, 1.0, None, None, 42, None),
          "bad batch arguments")
    L.bem_batch_destroy(None)
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


def _winding_number(pnts, elems, xs):
    """Generalised winding number of a closed triangle surface at each x: the
    sum of the signed solid angles of its triangles (Van Oosterom & Strackee,
    IEEE Trans. Biomed. Eng. 30, 1983) over 4 pi; +-1 inside, 0 outside.  An
    inside test that casts no rays."""
    P = np.asarray(pnts, np.float64)[np.asarray(elems)]
    out = np.zeros(len(xs))
    for i, x in enumerate(np.asarray(xs, np.float64)):
        a, b, c = P[:, 0] - x, P[:, 1] - x, P[:, 2] - x
        la, lb, lc = (np.linalg.norm(v, axis=1) for v in (a, b, c))
        num = np.einsum("ij,ij->i", a, np.cross(b, c))
        den = la * lb * lc + (a * b).sum(1) * lc + (b * c).sum(1) * la + (c * a).sum(1) * lb
        out[i] = 2.0 * np.arctan2(num, den).sum() / (4.0 * np.pi)
    return out


@pytest.mark.parametrize("case", ["ellipsoid", "sphere.obj"])
def test_inside_test_matches_winding_number_and_oracle(case):
    """The CHIEF generator's 3-ray majority inside test against the winding
    number (no rays) and the oracle's one-ray inObj (numerical.cu:119-131,
    rayTrnglInt mesh.cu:1184-1189) at 1200 seeded points in the bounding box of
    a perturbed ellipsoid and of the reference's own sphere.obj.  The
    reference's triangle test compares normalised edge cross products within
    EPS = 5e-5 (mesh.cu:1167-1182, numerical.h:124-126), which fails for hits
    near an edge, where one cross product is short and its direction is set by
    rounding: its one-ray answer is wrong at about 3 % of such points, the
    majority of three Moller-Trumbore rays at none."""
    if case == "ellipsoid":
        m = bem.Mesh().ellipsoid(3, (0.075, 0.09, 0.11), nbumps=8, lmax=6, amp=0.05, seed=7)
    else:
        m = bem.Mesh().read_obj(os.path.join(C.GOLDEN, "sphere.obj"))
    p, e, _ = C.mesh_arrays(m)
    lo, hi = p.min(0), p.max(0)
    rng = np.random.default_rng(11)
    xs = (lo + (hi - lo) * rng.random((1200, 3))).astype(np.float32)
    wn = _winding_number(p, e, xs)
    assert np.all(np.minimum(np.abs(wn), np.abs(np.abs(wn) - 1)) < 1e-6)  # closed surface
    truth = np.abs(wn) > 0.5
    got = m.points_inside(xs, seed=3)
    dirs = C.sphere_points(12, len(xs), 1.0)
    orc = np.array([O.ray_inside(x, d, p, e) for x, d in zip(xs, dirs)], bool)
    print("%s: %d of %d inside; library disagrees with the winding number at %d, the one-ray oracle at %d" %
          (case, truth.sum(), len(xs), (got != truth).sum(), (orc != truth).sum()))
    assert 300 < truth.sum() < 900
    assert np.array_equal(got, truth)
    assert (orc != truth).mean() < 0.06


def test_chief_points_are_inside_a_nonconvex_body():
    """CHIEF points generated on a perturbed ellipsoid (not convex where the
    bumps dent it) lie inside by the winding number."""
    m = bem.Mesh().ellipsoid(3, (0.075, 0.09, 0.11), nbumps=8, lmax=6, amp=0.05, seed=7)
    m.find_bb(1e-4).gen_chief(32, 0.01, seed=9)
    p, e, _ = C.mesh_arrays(m)
    assert np.all(np.abs(_winding_number(p, e, m.chief())) > 0.99)


def test_fix_orientation():
    """Randomly flipped windings, and a part wound outward as a whole, are
    repaired to the reference's winding (normals into the body)."""
    a = bem.Mesh().icosphere(2, 0.1)
    p, e0, bc = C.mesh_arrays(a)
    assert a.fix_orientation() == 0  # already consistent and inward
    e = e0.copy()
    flip = np.random.default_rng(4).random(len(e)) < 0.3
    e[flip] = e[flip][:, [0, 2, 1]]
    b = bem.Mesh().set(p, e, (0, 1, 0))
    assert b.fix_orientation() == flip.sum()
    got = b.elems()["nodes"]
    assert np.array_equal(got, e0)  # the flipped ones are swapped back exactly
    assert b.signed_volume() < 0
    # two spheres, the second wound outward as a whole
    p2 = np.vstack([p - [0.3, 0, 0], p + [0.3, 0, 0]])
    e2 = np.vstack([e0, e0[:, [0, 2, 1]] + p.shape[0]])
    t = bem.Mesh().set(p2, e2, (0, 1, 0))
    assert t.fix_orientation() == len(e0)
    assert np.array_equal(t.elems()["nodes"], np.vstack([e0, e0 + p.shape[0]]))
    assert t.signed_volume() < 0


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
    # block CG keeps the conjugate-transposed copy (for its FP16-split
    # products) at any count above one; lockstep CG only above 16
    blk8 = bem.request_footprint(M, N, E, 8, method=bem.SOLVE_BLOCK_CG)
    cg8 = bem.request_footprint(M, N, E, 8, method=bem.SOLVE_CGLS)
    assert blk8 >= cg8 + 8.0 * M * ((N + 63) // 64 * 64)


def test_tune_table_roundtrip(tmp_path):
    src = tmp_path / "tune.txt"
    src.write_text("26 29 1 16384\n22 25 4 8192\nbad line\np 0 31 31 2 1 3\np 1 23 23 1 0 12\n"
                   "p 0 31 31 2 1 99\n")  # last: chunk count past the limit, ignored
    bem.tune_clear()
    try:
        bem.tune_load(src)
        out = tmp_path / "out.txt"
        bem.tune_save(out)
        lines = out.read_text().splitlines()
        rows = sorted(tuple(map(int, l.split())) for l in lines if not l.startswith("p"))
        assert rows == [(22, 25, 4, 8192), (26, 29, 1, 16384)]
        prods = sorted(tuple(map(int, l.split()[1:])) for l in lines if l.startswith("p"))
        assert prods == [(0, 31, 31, 2, 1, 3), (1, 23, 23, 1, 0, 12)]
        with pytest.raises(bem.BemError):
            bem.tune_load(tmp_path / "missing.txt")
    finally:
        bem.tune_clear()
