"""ctypes access to the oracle (oracle/liboracle.so) for the tests and the bench's
cpu_baseline leg.  The oracle is the checker, never the thing measured or shipped.

Builds the oracle with `make -C oracle` when the shared library is missing.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
_lib = None

_d = ctypes.c_double
_f = ctypes.c_float
_i = ctypes.c_int
_l = ctypes.c_long
_p = ctypes.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_gauss_legendre.argtypes = [_i, _p, _p]
        L.orc_quad_init.argtypes = [_i]
        for sfx, R in (("_d", _d), ("_f", _f)):
            getattr(L, "orc_elem_regular" + sfx).argtypes = [R, R, R, _p, _p, _p, _p, _p]
            getattr(L, "orc_elem_singular" + sfx).argtypes = [_i, R, R, R, _p, _p, _p, _p]
            getattr(L, "orc_assemble" + sfx).argtypes = [R, R, R, _p, _i, _i, _p, _p, _i, _p, _i,
                                                       _p, _l, _p, _l, _i]
            getattr(L, "orc_assemble_rows" + sfx).argtypes = [R, R, R, _p, _i, _i, _p, _p, _i, _p, _i,
                                                            _p, _l, _p, _l, _i, _i, _i]
            getattr(L, "orc_field" + sfx).argtypes = [R, R, R, _p, _i, _p, _p, _i, _p, _i, _p, _l,
                                                    _i, _p, _p, _l, _i]
            getattr(L, "orc_lsq_qr" + sfx).argtypes = [_p, _i, _i, _l, _p, _i, _l]
            getattr(L, "orc_gmres" + sfx).argtypes = [_p, _p, _i, _l, _i, R, _p]
        L.orc_read_obj.argtypes = [ctypes.c_char_p, _p, _i, _p, _i, ctypes.POINTER(_i),
                                   ctypes.POINTER(_i)]
        L.orc_ray_trngl_int.argtypes = [_p, _p, _p, _p, _p]
        L.orc_ray_inside.argtypes = [_p, _p, _p, _p, _i]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_p) if a is not None else None


def _prec(dtype):
    """('_d', float64, complex128) or ('_f', float32, complex64)."""
    if np.dtype(dtype) in (np.dtype(np.float32), np.dtype(np.complex64)):
        return "_f", np.float32, np.complex64
    return "_d", np.float64, np.complex128


def gauss_legendre(n):
    x = np.zeros(n)
    w = np.zeros(n)
    assert lib().orc_gauss_legendre(n, _ptr(x), _ptr(w)) == 0
    return x, w


def quad_init(n=6):
    assert lib().orc_quad_init(n) == 0


def elem_regular(k, x, p, dens=1.2041, speed=343.21, dtype=np.float64):
    sfx, R, C = _prec(dtype)
    x = np.ascontiguousarray(x, R)
    p = np.ascontiguousarray(p, R).reshape(9)
    g = np.zeros(6, R)
    h = np.zeros(6, R)
    c = np.zeros(1, R)
    getattr(lib(), "orc_elem_regular" + sfx)(k, dens, speed, _ptr(x), _ptr(p), _ptr(g), _ptr(h), _ptr(c))
    return g.view(C), h.view(C), float(c[0])


def elem_singular(v, k, p, dens=1.2041, speed=343.21, dtype=np.float64):
    sfx, R, C = _prec(dtype)
    p = np.ascontiguousarray(p, R).reshape(9)
    g = np.zeros(6, R)
    h = np.zeros(6, R)
    c = np.zeros(1, R)
    getattr(lib(), "orc_elem_singular" + sfx)(v, k, dens, speed, _ptr(p), _ptr(g), _ptr(h), _ptr(c))
    return g.view(C), h.view(C), float(c[0])


def assemble(k, pnts, n_nod, elems, bc, binit=None, dens=1.2041, speed=343.21,
             dtype=np.float64, nthreads=None):
    """Oracle collocation system: returns (A (M x N), B (M x R)) as complex arrays.

    pnts: (M, 3) nodes then CHIEF points; elems: (E, 3) int; bc: (E, 3) complex
    (ca, cb, cc); binit: (M, R) complex initial B (incident field) or None."""
    sfx, R, C = _prec(dtype)
    pnts = np.ascontiguousarray(pnts, R)
    M = pnts.shape[0]
    elems = np.ascontiguousarray(elems, np.int32)
    bc = np.ascontiguousarray(bc, C)
    nrhs = 0 if binit is None else binit.shape[1]
    A = np.zeros((n_nod, M), C)  # column-major M x N == row-major N x M
    B = np.zeros((max(nrhs, 1), M), C)
    bi = None if binit is None else np.ascontiguousarray(np.asarray(binit, C).T)
    nt = nthreads or os.cpu_count() or 1
    rc = getattr(lib(), "orc_assemble" + sfx)(k, dens, speed, _ptr(pnts), n_nod, M - n_nod, _ptr(elems),
                                            _ptr(bc.view(R)), elems.shape[0], _ptr(bi), nrhs,
                                            _ptr(A), M, _ptr(B), M, nt)
    assert rc == 0
    return A.T.copy(), B[:nrhs].T.copy()


def assemble_rows(k, pnts, n_nod, elems, bc, r0, r1, dens=1.2041, speed=343.21, dtype=np.float32,
                  nthreads=None):
    """Rows [r0, r1) of the oracle A (the cpu_baseline sample); returns A rows (r1-r0, N)."""
    sfx, R, C = _prec(dtype)
    pnts = np.ascontiguousarray(pnts, R)
    M = pnts.shape[0]
    elems = np.ascontiguousarray(elems, np.int32)
    bc = np.ascontiguousarray(bc, C)
    nr = r1 - r0
    A = np.zeros((n_nod, nr), C)
    B = np.zeros((1, nr), C)
    nt = nthreads or os.cpu_count() or 1
    rc = getattr(lib(), "orc_assemble_rows" + sfx)(k, dens, speed, _ptr(pnts), n_nod, M - n_nod, _ptr(elems),
                                                 _ptr(bc.view(R)), elems.shape[0], None, 0, _ptr(A), nr,
                                                 _ptr(B), nr, r0, r1, nt)
    assert rc == 0
    return A.T.copy()


def field(k, pnts, n_nod, elems, bc, xs, surf_p, inc=None, dens=1.2041, speed=343.21,
          dtype=np.float64, nthreads=None):
    """Exterior pressure at points xs (P, 3) for surface solutions surf_p (N, R)."""
    sfx, R, C = _prec(dtype)
    pnts = np.ascontiguousarray(pnts, R)
    xs = np.ascontiguousarray(xs, R)
    elems = np.ascontiguousarray(elems, np.int32)
    bc = np.ascontiguousarray(bc, C)
    sp = np.asarray(surf_p, C)
    if sp.ndim == 1:
        sp = sp[:, None]
    nr = sp.shape[1]
    spT = np.ascontiguousarray(sp.T)
    P = xs.shape[0]
    out = np.zeros((nr, P), C)
    incT = None if inc is None else np.ascontiguousarray(np.asarray(inc, C).reshape(P, nr).T)
    nt = nthreads or os.cpu_count() or 1
    rc = getattr(lib(), "orc_field" + sfx)(k, dens, speed, _ptr(pnts), n_nod, _ptr(elems), _ptr(bc.view(R)),
                                         elems.shape[0], _ptr(xs), P, _ptr(spT), sp.shape[0], nr,
                                         _ptr(incT), _ptr(out), P, nt)
    assert rc == 0
    return out.T.copy()


def lsq_qr(A, B, dtype=np.complex128):
    sfx, R, C = _prec(dtype)
    A = np.asarray(A, C)
    m, n = A.shape
    B = np.asarray(B, C)
    if B.ndim == 1:
        B = B[:, None]
    a = np.ascontiguousarray(A.T)
    b = np.ascontiguousarray(B.T)
    rc = getattr(lib(), "orc_lsq_qr" + sfx)(_ptr(a), m, n, m, _ptr(b), b.shape[0], m)
    assert rc == 0, rc
    return b.T[:n].copy()


def gmres(A, b, max_iter, tol, x0=None, dtype=np.complex128):
    sfx, R, C = _prec(dtype)
    A = np.asarray(A, C)
    m = A.shape[0]
    a = np.ascontiguousarray(A.T)
    b = np.ascontiguousarray(b, C)
    x = np.zeros(m, C) if x0 is None else np.array(x0, C)
    it = getattr(lib(), "orc_gmres" + sfx)(_ptr(a), _ptr(b), m, m, max_iter, tol, _ptr(x))
    assert it >= 0
    return x, it


def read_obj(path):
    np_, ne = _i(0), _i(0)
    assert lib().orc_read_obj(path.encode(), None, 0, None, 0, ctypes.byref(np_), ctypes.byref(ne)) == 0
    p = np.zeros((np_.value, 3), np.float32)
    e = np.zeros((ne.value, 3), np.int32)
    assert lib().orc_read_obj(path.encode(), _ptr(p), np_.value, _ptr(e), ne.value,
                              ctypes.byref(np_), ctypes.byref(ne)) == 0
    return p, e


def ray_inside(sp, d, pnts, elems):
    sp = np.ascontiguousarray(sp, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    pnts = np.ascontiguousarray(pnts, np.float32)
    elems = np.ascontiguousarray(elems, np.int32)
    return lib().orc_ray_inside(_ptr(sp), _ptr(d), _ptr(pnts), _ptr(elems), elems.shape[0])
