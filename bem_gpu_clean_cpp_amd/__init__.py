"""MI355X boundary-element Helmholtz solver: Python binding of libbem's C-ABI.

The product is the C-ABI shared library ``libbem.so`` (include/bem.h): host C++
plus hand-written gfx950 HIP kernels.  This module is a thin ctypes layer over
it, mirroring the reference's C++ entry points (mesh::readObj, mesh::findBB,
mesh::genCHIEF, gaussQuad, bemSystem, qrSolver, lsqSolver, GMRES,
genExtPressure, wrtCplxMat) plus the device-resident solver extension.  There
is no fallback: if ``libbem.so`` is missing or the GPU is absent, calls fail
loudly.
"""
import ctypes
import os
import subprocess

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libbem.so")

PREC_FP32 = 32
PREC_FP64 = 64
RHS_POINT = 0
RHS_PLANE = 1
SOLVE_QR = 0
SOLVE_GMRES = 1
SOLVE_CGLS = 2
SOLVE_BLOCK_CG = 3

_lib = None


class BemError(RuntimeError):
    pass


class bem_complex(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float)]


class bem_cart_coord(ctypes.Structure):
    _fields_ = [("coords", ctypes.c_float * 3)]


class bem_tri_elem(ctypes.Structure):
    _fields_ = [("nodes", ctypes.c_int * 3), ("_pad", ctypes.c_int), ("bc", bem_complex * 3)]


class bem_solve_stats(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("iterations", ctypes.c_int), ("refinements", ctypes.c_int),
                ("converged", ctypes.c_int), ("residual", ctypes.c_double), ("assemble_ms", ctypes.c_double),
                ("solve_ms", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class bem_solve_params(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("restart", ctypes.c_int), ("tol", ctypes.c_float),
                ("maxit", ctypes.c_int), ("refine_tol", ctypes.c_double), ("max_refine", ctypes.c_int)]


class bem_request_info(ctypes.Structure):
    _fields_ = [("id", ctypes.c_long), ("status", ctypes.c_int), ("slot", ctypes.c_int), ("N", ctypes.c_int),
                ("M", ctypes.c_int), ("nrhs", ctypes.c_int), ("footprint_bytes", ctypes.c_double),
                ("submit_ms", ctypes.c_double), ("start_ms", ctypes.c_double), ("end_ms", ctypes.c_double),
                ("solve", bem_solve_stats)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_ if f != "solve"}
        d["solve"] = self.solve.as_dict()
        return d


class bem_tune_result(ctypes.Structure):
    _fields_ = [("lanes_per_row", ctypes.c_int), ("waves", ctypes.c_int), ("best_ms", ctypes.c_double),
                ("default_ms", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class bem_product_tune_result(ctypes.Structure):
    _fields_ = [("rhs_per_wave_n", ctypes.c_int), ("chunks_n", ctypes.c_int), ("rhs_per_wave_h", ctypes.c_int),
                ("chunks_h", ctypes.c_int), ("best_ms_n", ctypes.c_double), ("default_ms_n", ctypes.c_double),
                ("best_ms_h", ctypes.c_double), ("default_ms_h", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class bem_server_info(ctypes.Structure):
    _fields_ = [("queued", ctypes.c_int), ("running", ctypes.c_int), ("max_running", ctypes.c_int),
                ("slots", ctypes.c_int), ("budget_bytes", ctypes.c_double), ("in_use_bytes", ctypes.c_double),
                ("peak_in_use_bytes", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


assert ctypes.sizeof(bem_tri_elem) == 40 and ctypes.sizeof(bem_cart_coord) == 12

ELEM_DTYPE = np.dtype([("nodes", np.int32, 3), ("_pad", np.int32), ("bc", np.complex64, 3)])
assert ELEM_DTYPE.itemsize == 40

_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_D = ctypes.c_double

# name: (restype, argtypes)
_SIGS = {
    "bem_mesh_create": (_P, []),
    "bem_mesh_destroy": (None, [_P]),
    "bem_mesh_read_obj": (_I, [_P, ctypes.c_char_p]),
    "bem_mesh_set": (_I, [_P, _P, _I, _P, _I]),
    "bem_mesh_set_bc": (_I, [_P, bem_complex, bem_complex, bem_complex]),
    "bem_mesh_find_bb": (_I, [_P, _F]),
    "bem_mesh_get_bb": (_I, [_P, _P]),
    "bem_mesh_gen_chief": (_I, [_P, _I, _F, ctypes.c_ulonglong]),
    "bem_mesh_set_chief": (_I, [_P, _P, _I]),
    "bem_mesh_num_pnts": (_I, [_P]),
    "bem_mesh_num_elems": (_I, [_P]),
    "bem_mesh_num_chief": (_I, [_P]),
    "bem_mesh_get_pnts": (_I, [_P, _P]),
    "bem_mesh_get_elems": (_I, [_P, _P]),
    "bem_mesh_get_chief": (_I, [_P, _P]),
    "bem_mesh_gen_icosphere": (_I, [_P, _I, _F]),
    "bem_mesh_gen_ellipsoid": (_I, [_P, _I, _F, _F, _F, _I, _I, _F, ctypes.c_ulonglong]),
    "bem_mesh_signed_volume": (_D, [_P]),
    "bem_mesh_fix_orientation": (_I, [_P, ctypes.POINTER(_I)]),
    "bem_mesh_points_inside": (_I, [_P, _P, _I, ctypes.c_ulonglong, _P]),
    "bem_mesh_surf_area": (_F, [_P]),
    "bem_quad_init": (_I, [_I]),
    "bem_quad_get": (_I, [_P, _P]),
    "bem_set_medium": (_I, [_F, _F]),
    "bem_set_source_strength": (_I, [_F]),
    "bem_set_near_field": (_I, [_F, _I]),
    "bem_system": (_I, [_P, _F, _P, _I, _P, _I, _P, _I]),
    "bem_gen_system": (_I, [_F, _P, _I, _P, _I, _I, _P, _I, _P, _I, _P, _I]),
    "bem_qr_solve": (_I, [_P, _I, _I, _I, _P, _I, _I]),
    "bem_lsq_solve": (_I, [_P, _I, _I, _I, _P, _I, _I, _P]),
    "bem_gmres": (_I, [_P, _P, _I, _I, _F, _P]),
    "bem_ext_pressure": (bem_complex, [_F, _P, bem_cart_coord, bem_cart_coord, _P]),
    "bem_ext_pressure_noalloc": (_I, [_F, _P, bem_cart_coord, _P, _I, _P, _P, _P]),
    "bem_wrt_cplx_mat": (_I, [_P, _I, _I, _I, ctypes.c_char_p]),
    "bem_print_complex_matrix": (None, [_P, _I, _I, _I]),
    "bem_solver_create": (_P, [_P]),
    "bem_solver_destroy": (None, [_P]),
    "bem_solver_assemble": (_I, [_P, _F, _I]),
    "bem_solver_set_rhs": (_I, [_P, _I, _P, _I, _F]),
    "bem_solver_solve": (_I, [_P, _I, _I, _F, _I, _D, _I, ctypes.POINTER(bem_solve_stats)]),
    "bem_solver_get_solution": (_I, [_P, _P, _I]),
    "bem_solver_get_solution_z": (_I, [_P, _P, _I]),
    "bem_solver_get_system": (_I, [_P, _P, _I, _P, _I]),
    "bem_solver_field": (_I, [_P, _P, _I, _P, _I]),
    "bem_solver_dims": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "bem_solver_stats": (_I, [_P, ctypes.POINTER(bem_solve_stats)]),
    "bem_solver_get_rows": (_I, [_P, _P, _I, _P]),
    "bem_solver_field_matrix": (_I, [_P, _P, _I]),
    "bem_solver_field_apply": (_I, [_P, _P, _I]),
    "bem_solver_get_field_matrix": (_I, [_P, _P, _I]),
    "bem_solver_set_mesh": (_I, [_P, _P]),
    "bem_solver_release": (_I, [_P]),
    "bem_batch_solve": (_I, [_P, _P, _I, _I, _I, _P, _I, _I, _F, ctypes.POINTER(bem_solve_params), _I, _P,
                             _I, _P]),
    "bem_batch_create": (_P, [_P, _I]),
    "bem_batch_run": (_I, [_P, _P, _I, _I, _I, _P, _I, _I, _F, ctypes.POINTER(bem_solve_params), _P, _I, _P]),
    "bem_batch_destroy": (None, [_P]),
    "bem_server_create": (_P, [_I, _D]),
    "bem_server_submit": (ctypes.c_long, [_P, _P, _F, _I, _I, _P, _I, _F, ctypes.POINTER(bem_solve_params)]),
    "bem_server_wait": (_I, [_P, ctypes.c_long, _P, _I, ctypes.POINTER(bem_request_info)]),
    "bem_server_status": (_I, [_P, ctypes.POINTER(bem_server_info)]),
    "bem_server_destroy": (None, [_P]),
    "bem_request_footprint": (_D, [_I, _I, _I, _I, _I, _I, _I]),
    "bem_autotune_assembly": (_I, [_P, _F, _I, ctypes.POINTER(bem_tune_result)]),
    "bem_autotune_products": (_I, [_I, _I, _I, ctypes.POINTER(bem_product_tune_result)]),
    "bem_tune_save": (_I, [ctypes.c_char_p]),
    "bem_tune_load": (_I, [ctypes.c_char_p]),
    "bem_tune_clear": (_I, []),
    "bem_timing_enable": (_I, [_I]),
    "bem_timing_reset": (_I, []),
    "bem_timing_get": (_I, [ctypes.c_char_p, ctypes.POINTER(_D), ctypes.POINTER(ctypes.c_long)]),
    "bem_sync": (_I, []),
    "bem_last_error": (ctypes.c_char_p, []),
    "bem_version": (ctypes.c_char_p, []),
}

EXPORTED = tuple(_SIGS)


def build(jobs=8):
    """Compile libbem.so for gfx950 (hipcc) in-tree."""
    subprocess.check_call(["make", "-s", "-C", os.path.join(PKG_DIR, "csrc"), "-j%d" % jobs])


def lib():
    """The loaded libbem (raises BemError if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BemError("libbem.so not built (%s); run bem_gpu_clean_cpp_amd.build()" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error():
    return lib().bem_last_error().decode()


def _check(rc, what):
    if rc != 0:
        raise BemError("%s failed: %s" % (what, last_error()))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


def _c64(a):
    return np.ascontiguousarray(a, np.complex64)


# --------------------------------------------------------------- global ---
def quad_init(order=6):
    """gaussQuad(order) + sendToDevice (numerical.cu:152-156, 227-231)."""
    _check(lib().bem_quad_init(order), "bem_quad_init")


def quad_get(order=6):
    x = np.zeros(order)
    w = np.zeros(order)
    _check(lib().bem_quad_get(_ptr(x), _ptr(w)), "bem_quad_get")
    return x, w


def set_medium(density, speed):
    _check(lib().bem_set_medium(density, speed), "bem_set_medium")


def set_source_strength(q):
    _check(lib().bem_set_source_strength(q), "bem_set_source_strength")


def set_near_field(factor, max_level=12):
    """FP64 near-singular integration for (row, element) pairs closer than
    factor x the element's longest edge (0: off, the reference's rule)."""
    _check(lib().bem_set_near_field(factor, max_level), "bem_set_near_field")


# ----------------------------------------------------------------- mesh ---
class Mesh:
    """The reference's `mesh` class (mesh.h:216-259) over bem_mesh."""

    def __init__(self):
        self._h = lib().bem_mesh_create()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.bem_mesh_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def read_obj(self, path):
        _check(lib().bem_mesh_read_obj(self._h, path.encode()), "readObj(%s)" % path)
        return self

    def set(self, pnts, elems, bc=(0, 1, -1)):
        """Set nodes (N, 3) and triangles (E, 3); bc = (ca, cb, cc) or (E, 3) complex."""
        p = np.ascontiguousarray(pnts, np.float32).reshape(-1, 3)
        e = np.zeros(len(elems), ELEM_DTYPE)
        e["nodes"] = np.asarray(elems, np.int32).reshape(-1, 3)
        bcv = np.asarray(bc, np.complex64)
        e["bc"] = bcv if bcv.ndim == 2 else np.broadcast_to(bcv, (len(e), 3))
        _check(lib().bem_mesh_set(self._h, _ptr(p), p.shape[0], _ptr(e), e.shape[0]), "bem_mesh_set")
        return self

    def set_bc(self, ca, cb, cc):
        C = lambda z: bem_complex(complex(z).real, complex(z).imag)
        _check(lib().bem_mesh_set_bc(self._h, C(ca), C(cb), C(cc)), "bem_mesh_set_bc")
        return self

    def find_bb(self, threshold=1e-4):
        _check(lib().bem_mesh_find_bb(self._h, threshold), "findBB")
        return self

    def bb(self):
        b = np.zeros(6, np.float32)
        _check(lib().bem_mesh_get_bb(self._h, _ptr(b)), "bem_mesh_get_bb")
        return b

    def gen_chief(self, num, threshold, seed=0):
        _check(lib().bem_mesh_gen_chief(self._h, num, threshold, seed), "genCHIEF")
        return self

    def set_chief(self, pts):
        p = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
        _check(lib().bem_mesh_set_chief(self._h, _ptr(p), p.shape[0]), "bem_mesh_set_chief")
        return self

    def icosphere(self, subdiv, radius):
        _check(lib().bem_mesh_gen_icosphere(self._h, subdiv, radius), "icosphere")
        return self

    def ellipsoid(self, subdiv, axes, nbumps=8, lmax=6, amp=0.05, seed=0):
        _check(lib().bem_mesh_gen_ellipsoid(self._h, subdiv, axes[0], axes[1], axes[2], nbumps, lmax, amp, seed),
               "ellipsoid")
        return self

    @property
    def num_pnts(self):
        return lib().bem_mesh_num_pnts(self._h)

    @property
    def num_elems(self):
        return lib().bem_mesh_num_elems(self._h)

    @property
    def num_chief(self):
        return lib().bem_mesh_num_chief(self._h)

    def pnts(self):
        p = np.zeros((self.num_pnts, 3), np.float32)
        _check(lib().bem_mesh_get_pnts(self._h, _ptr(p)), "get_pnts")
        return p

    def elems(self):
        e = np.zeros(self.num_elems, ELEM_DTYPE)
        _check(lib().bem_mesh_get_elems(self._h, _ptr(e)), "get_elems")
        return e

    def chief(self):
        c = np.zeros((self.num_chief, 3), np.float32)
        if self.num_chief:
            _check(lib().bem_mesh_get_chief(self._h, _ptr(c)), "get_chief")
        return c

    def signed_volume(self):
        return lib().bem_mesh_signed_volume(self._h)

    def fix_orientation(self):
        """Consistent winding with the normals into the body; returns the
        number of elements whose winding changed."""
        n = _I()
        _check(lib().bem_mesh_fix_orientation(self._h, ctypes.byref(n)), "fix_orientation")
        return n.value

    def points_inside(self, pts, seed=0):
        """The CHIEF generator's inside test (3-ray majority) for (n, 3) points."""
        p = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
        out = np.zeros(p.shape[0], np.int32)
        _check(lib().bem_mesh_points_inside(self._h, _ptr(p), p.shape[0], seed, _ptr(out)), "points_inside")
        return out.astype(bool)

    def surf_area(self):
        return lib().bem_mesh_surf_area(self._h)


# ------------------------------------------------- reference entry points ---
def bem_system(mesh, k, srcs=None):
    """bemSystem (mesh.cu:985-1006): returns (A (M x N), B (M x nsrc)) complex64."""
    M = mesh.num_pnts + mesh.num_chief
    N = mesh.num_pnts
    s = np.zeros((0, 3), np.float32) if srcs is None else np.ascontiguousarray(srcs, np.float32).reshape(-1, 3)
    ns = s.shape[0]
    A = np.zeros((N, M), np.complex64)  # column-major M x N
    B = np.zeros((max(ns, 1), M), np.complex64)
    _check(lib().bem_system(mesh.handle, k, _ptr(s), ns, _ptr(A), M, _ptr(B), M), "bemSystem")
    return A.T, B[:ns].T


def qr_solve(A, B):
    """qrSolver (numerical.cu:580-644): least squares, returns X (N x nrhs) and Q^H B rows."""
    A = np.asarray(A)
    m, n = A.shape
    a = np.array(_c64(A).T, order="C", copy=True)
    Bm = _c64(B)
    one = Bm.ndim == 1
    Bm = Bm.reshape(m, -1)
    b = np.array(Bm.T, order="C", copy=True)  # the solve overwrites it
    _check(lib().bem_qr_solve(_ptr(a), m, n, m, _ptr(b), b.shape[0], m), "qrSolver")
    X = b.T[:n]
    return X[:, 0].copy() if one else X.copy()


def lsq_solve(A, B, want_q=False):
    """lsqSolver (numerical.cu:506-554): returns (X, R, Q or None)."""
    A = np.asarray(A)
    m, n = A.shape
    a = np.array(_c64(A).T, order="C", copy=True)  # overwritten with R
    Bm = _c64(B).reshape(m, -1)
    b = np.array(Bm.T, order="C", copy=True)  # overwritten with the solution
    q = np.zeros((m, m), np.complex64) if want_q else None
    _check(lib().bem_lsq_solve(_ptr(a), m, n, m, _ptr(b), b.shape[0], m, _ptr(q)), "lsqSolver")
    return b.T[:n].copy(), a.T.copy(), (q.T.copy() if want_q else None)


def gmres(A, b, max_iter, threshold, x0=None):
    """GMRES (GMRES.cu:145-279) for a square system; returns x."""
    A = np.asarray(A)
    m = A.shape[0]
    a = np.ascontiguousarray(_c64(A).T)
    bb = _c64(b).reshape(m)
    x = np.zeros(m, np.complex64) if x0 is None else _c64(x0).reshape(m).copy()
    _check(lib().bem_gmres(_ptr(a), _ptr(bb), m, max_iter, threshold, _ptr(x)), "GMRES")
    return x


def ext_pressure(k, mesh, src, x, surf_p):
    """genExtPressure (mesh.cu:1038-1073) at one point."""
    sp = _c64(surf_p).reshape(-1)
    r = lib().bem_ext_pressure(k, mesh.handle, bem_cart_coord((_F * 3)(*src)), bem_cart_coord((_F * 3)(*x)),
                               _ptr(sp))
    if np.isnan(r.x):
        raise BemError("genExtPressure failed: %s" % last_error())
    return complex(r.x, r.y)


def ext_pressure_noalloc(k, mesh, src, xs_ptr, P, surf_ptr, out_ptr, stream=None):
    """bem_ext_pressure_noalloc: the field at P points from device buffers
    (raw device addresses), nothing allocated; queued on `stream`
    (a hipStream_t handle, None: the library's stream), no wait."""
    _check(lib().bem_ext_pressure_noalloc(k, mesh.handle, bem_cart_coord((_F * 3)(*src)), xs_ptr, P, surf_ptr,
                                          out_ptr, stream), "bem_ext_pressure_noalloc")


def wrt_cplx_mat(mat, path):
    """wrtCplxMat (numerical.cu:556-578)."""
    m = _c64(mat)
    if m.ndim == 1:
        m = m[None, :]
    a = np.ascontiguousarray(m.T)
    _check(lib().bem_wrt_cplx_mat(_ptr(a), m.shape[0], m.shape[1], m.shape[0], path.encode()), "wrtCplxMat")


# -------------------------------------------------- device-resident solver ---
class Solver:
    """Device-resident solver: mesh, A, right-hand sides and solutions stay in HBM."""

    def __init__(self, mesh):
        self._h = lib().bem_solver_create(mesh.handle)
        if not self._h:
            raise BemError("bem_solver_create failed: %s" % last_error())

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.bem_solver_destroy(h)
            self._h = None

    def close(self):
        self.__del__()

    def dims(self):
        M, N, R = _I(), _I(), _I()
        _check(lib().bem_solver_dims(self._h, ctypes.byref(M), ctypes.byref(N), ctypes.byref(R)), "dims")
        return M.value, N.value, R.value

    def assemble(self, k, precision=PREC_FP32):
        _check(lib().bem_solver_assemble(self._h, k, precision), "assemble")
        return self

    def set_rhs(self, kind, params, amp=1.0):
        p = np.ascontiguousarray(params, np.float32).reshape(-1, 3)
        _check(lib().bem_solver_set_rhs(self._h, kind, _ptr(p), p.shape[0], amp), "set_rhs")
        return self

    def solve(self, method=SOLVE_GMRES, restart=30, tol=1e-6, maxit=3000, refine_tol=1e-9, max_refine=50):
        st = bem_solve_stats()
        _check(lib().bem_solver_solve(self._h, method, restart, tol, maxit, refine_tol, max_refine,
                                      ctypes.byref(st)), "solve")
        return st.as_dict()

    def solution(self, fp64=True):
        M, N, R = self.dims()
        if fp64:
            x = np.zeros((R, N), np.complex128)
            _check(lib().bem_solver_get_solution_z(self._h, _ptr(x), N), "get_solution_z")
        else:
            x = np.zeros((R, N), np.complex64)
            _check(lib().bem_solver_get_solution(self._h, _ptr(x), N), "get_solution")
        return x.T.copy()

    def system(self):
        M, N, R = self.dims()
        A = np.zeros((N, M), np.complex64)
        B = np.zeros((max(R, 1), M), np.complex64)
        _check(lib().bem_solver_get_system(self._h, _ptr(A), M, _ptr(B) if R else None, M), "get_system")
        return A.T.copy(), B[:R].T.copy()

    def rows(self, rows):
        """Rows of the assembled A (len(rows) x N, complex128)."""
        r = np.ascontiguousarray(rows, np.int32).ravel()
        M, N, R = self.dims()
        out = np.zeros((r.size, N), np.complex128)
        _check(lib().bem_solver_get_rows(self._h, _ptr(r), r.size, _ptr(out)), "get_rows")
        return out

    def field_matrix(self, pts):
        """Assemble the radiation operator of the field points (kept in HBM)."""
        p = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
        _check(lib().bem_solver_field_matrix(self._h, _ptr(p), p.shape[0]), "field_matrix")
        self._fP = p.shape[0]
        return self

    def field_apply(self):
        """Field at the points of field_matrix for every current solution (P x nrhs)."""
        M, N, R = self.dims()
        P = self._fP
        out = np.zeros((R, P), np.complex64)
        _check(lib().bem_solver_field_apply(self._h, _ptr(out), P), "field_apply")
        return out.T.copy()

    def get_field_matrix(self):
        M, N, R = self.dims()
        P = self._fP
        F = np.zeros((N, P), np.complex64)
        _check(lib().bem_solver_get_field_matrix(self._h, _ptr(F), P), "get_field_matrix")
        return F.T.copy()

    def field(self, pts):
        p = np.ascontiguousarray(pts, np.float32).reshape(-1, 3)
        M, N, R = self.dims()
        out = np.zeros((R, p.shape[0]), np.complex64)
        _check(lib().bem_solver_field(self._h, _ptr(p), p.shape[0], _ptr(out), p.shape[0]), "field")
        return out.T.copy()

    def stats(self):
        st = bem_solve_stats()
        _check(lib().bem_solver_stats(self._h, ctypes.byref(st)), "stats")
        return st.as_dict()

    def set_mesh(self, mesh):
        _check(lib().bem_solver_set_mesh(self._h, mesh.handle), "set_mesh")
        return self

    def release(self):
        _check(lib().bem_solver_release(self._h), "release")


# ------------------------------------------------- batches and serving ---
def solve_params(method=SOLVE_GMRES, restart=30, tol=1e-6, maxit=3000, refine_tol=1e-8, max_refine=50):
    return bem_solve_params(method, restart, tol, maxit, refine_tol, max_refine)


def batch_solve(mesh, ks, kind, params, amp=1.0, precision=PREC_FP32, workers=4, **solve_kw):
    """One mesh at many wavenumbers.  params: (nrhs, 3) for the same right-hand
    sides at every frequency, or (nk, nrhs, 3) for one set per frequency.
    Returns (X with shape (N, nrhs, nk) complex128, list of per-frequency stats)."""
    k = np.ascontiguousarray(ks, np.float32).ravel()
    nk = k.size
    p = np.ascontiguousarray(params, np.float32)
    per_freq = 1 if p.ndim == 3 else 0
    if per_freq and p.shape[0] != nk:
        raise BemError("params must be (nrhs, 3) or (nk, nrhs, 3)")
    R = p.shape[-2]
    N = mesh.num_pnts
    X = np.zeros((nk, R, N), np.complex128)
    st = (bem_solve_stats * nk)()
    sp = solve_params(**solve_kw)
    _check(lib().bem_batch_solve(mesh.handle, _ptr(k), nk, precision, kind, _ptr(p), R, per_freq, amp,
                                 ctypes.byref(sp), workers, _ptr(X), N, st), "batch_solve")
    return X.transpose(2, 1, 0).copy(), [x.as_dict() for x in st]


class Batch:
    """Batch context (bem_batch): `workers` resident solvers of one mesh kept
    between runs; run() takes the arguments of batch_solve and gives the same
    results."""

    def __init__(self, mesh, workers=4):
        self._h = lib().bem_batch_create(mesh.handle, workers)
        if not self._h:
            raise BemError("bem_batch_create failed: %s" % last_error())
        self.num_pnts = mesh.num_pnts

    def run(self, ks, kind, params, amp=1.0, precision=PREC_FP32, **solve_kw):
        k = np.ascontiguousarray(ks, np.float32).ravel()
        nk = k.size
        p = np.ascontiguousarray(params, np.float32)
        per_freq = 1 if p.ndim == 3 else 0
        if per_freq and p.shape[0] != nk:
            raise BemError("params must be (nrhs, 3) or (nk, nrhs, 3)")
        R = p.shape[-2]
        N = self.num_pnts
        X = np.zeros((nk, R, N), np.complex128)
        st = (bem_solve_stats * nk)()
        sp = solve_params(**solve_kw)
        _check(lib().bem_batch_run(self._h, _ptr(k), nk, precision, kind, _ptr(p), R, per_freq, amp,
                                   ctypes.byref(sp), _ptr(X), N, st), "batch_run")
        return X.transpose(2, 1, 0).copy(), [x.as_dict() for x in st]

    def close(self):
        if self._h:
            lib().bem_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def request_footprint(M, N, E, nrhs, precision=PREC_FP32, method=SOLVE_GMRES, restart=30):
    return lib().bem_request_footprint(M, N, E, nrhs, precision, method, restart)


class Server:
    """Request server with continuous admission by HBM footprint (bem_server)."""

    def __init__(self, slots=4, hbm_budget_bytes=0.0):
        self._h = lib().bem_server_create(slots, float(hbm_budget_bytes))
        if not self._h:
            raise BemError("bem_server_create failed: %s" % last_error())
        self._n = {}

    def submit(self, mesh, k, kind, params, amp=1.0, precision=PREC_FP32, **solve_kw):
        p = np.ascontiguousarray(params, np.float32).reshape(-1, 3)
        sp = solve_params(**solve_kw)
        rid = lib().bem_server_submit(self._h, mesh.handle, k, precision, kind, _ptr(p), p.shape[0], amp,
                                      ctypes.byref(sp))
        if rid < 0:
            raise BemError("bem_server_submit failed: %s" % last_error())
        self._n[rid] = (mesh.num_pnts, p.shape[0])
        return rid

    def wait(self, rid):
        """(solution N x nrhs complex128, request record dict); raises on failure."""
        N, R = self._n.pop(rid)
        x = np.zeros((R, N), np.complex128)
        info = bem_request_info()
        _check(lib().bem_server_wait(self._h, rid, _ptr(x), N, ctypes.byref(info)), "server_wait")
        return x.T.copy(), info.as_dict()

    def status(self):
        info = bem_server_info()
        _check(lib().bem_server_status(self._h, ctypes.byref(info)), "server_status")
        return info.as_dict()

    def close(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.bem_server_destroy(h)
            self._h = None

    def __del__(self):
        self.close()


# ---------------------------------------------------------------- tuning ---
def autotune_assembly(mesh, k, precision=PREC_FP32):
    """Time candidate assembly launch shapes for this mesh; the fastest is kept
    for meshes of its size class.  Returns the choice and both timings."""
    r = bem_tune_result()
    _check(lib().bem_autotune_assembly(mesh.handle, k, precision, ctypes.byref(r)), "autotune_assembly")
    return r.as_dict()


def autotune_products(M, N, nrhs):
    """Time the multi-right-hand-side product shapes for an M x N matrix and
    nrhs right-hand sides; the fastest is kept for its size class."""
    r = bem_product_tune_result()
    _check(lib().bem_autotune_products(M, N, nrhs, ctypes.byref(r)), "autotune_products")
    return r.as_dict()


def tune_save(path):
    _check(lib().bem_tune_save(str(path).encode()), "tune_save")


def tune_load(path):
    _check(lib().bem_tune_load(str(path).encode()), "tune_load")


def tune_clear():
    _check(lib().bem_tune_clear(), "tune_clear")


# ---------------------------------------------------------------- timing ---
def timing_enable(on=True):
    _check(lib().bem_timing_enable(1 if on else 0), "timing_enable")


def timing_reset():
    _check(lib().bem_timing_reset(), "timing_reset")


def timing_get(kernel):
    ms, cnt = _D(), ctypes.c_long()
    _check(lib().bem_timing_get(kernel.encode(), ctypes.byref(ms), ctypes.byref(cnt)), "timing_get")
    return ms.value, cnt.value


def sync():
    _check(lib().bem_sync(), "sync")
