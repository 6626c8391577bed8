"""Host-side dense algebra of the Krylov solvers, on the CPU.

`small_dense.hpp` (the R x R algebra of block CG: Cholesky factor, triangular
inverse, conjugate transpose, product) and `hessenberg_ls.hpp` (the GMRES
least-squares problem min |beta e1 - H y| kept as Givens rotations, the
reference's givens_coeffs / apply_givens / Rsolver, GMRES.cu:26-88) are
compiled on their own into a small driver (tests/native/dense_check.cpp) and
compared with numpy / LAPACK on seeded inputs:
  * Cholesky factor of random Hermitian positive definite matrices (n = 1, 5,
    32, 64) against numpy.linalg.cholesky, and refusal of a singular one;
  * triangular inverse, product and conjugate transpose against numpy;
  * the Givens least-squares update: the residual estimate after every column
    against numpy's least-squares residual of the same leading block, the
    final y against numpy.linalg.lstsq, and a column whose rotation is zero
    (a singular Hessenberg) refused without touching the solution.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("native") / "dense_check")
    subprocess.check_call([gxx, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "bem_gpu_clean_cpp_amd", "csrc"),
                           "-o", exe, os.path.join(ROOT, "tests", "native", "dense_check.cpp")])
    return exe


def _fmt(*mats):
    out = []
    for m in mats:
        for v in np.asarray(m).reshape(-1, order="F"):
            out.append("%.17g %.17g" % (v.real, v.imag))
    return "\n".join(out)


def _run(exe, text):
    res = subprocess.run([exe], input=text, capture_output=True, text=True, check=True)
    return res.stdout.split()


def _mat(tok, n):
    v = np.array([float(t) for t in tok[:2 * n * n]])
    return (v[0::2] + 1j * v[1::2]).reshape((n, n), order="F"), tok[2 * n * n:]


def _hpd(rng, n, cond=1e3):
    q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    ev = np.logspace(0, np.log10(cond), n)
    return (q * ev) @ q.conj().T


@pytest.mark.parametrize("n", [1, 5, 32, 64])
def test_cholesky_matches_numpy(driver, n):
    G = _hpd(np.random.default_rng(100 + n), n)
    tok = _run(driver, "chol %d\n%s\n" % (n, _fmt(G)))
    assert tok[0] == "1"
    L, _ = _mat(tok[1:], n)
    Lref = np.linalg.cholesky(G)
    assert np.abs(L - Lref).max() <= 1e-12 * np.abs(Lref).max()
    assert np.allclose(np.triu(L, 1), 0)


def test_cholesky_refuses_singular(driver):
    rng = np.random.default_rng(7)
    B = rng.standard_normal((8, 7)) + 1j * rng.standard_normal((8, 7))
    tok = _run(driver, "chol 8\n%s\n" % _fmt(B @ B.conj().T))
    assert tok[0] == "0"


@pytest.mark.parametrize("n", [1, 6, 40])
def test_inverse_product_transpose(driver, n):
    rng = np.random.default_rng(200 + n)
    L = np.tril(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)), -1) / n
    L += np.diag(1.0 + rng.random(n))
    X, _ = _mat(_run(driver, "linv %d\n%s\n" % (n, _fmt(L))), n)
    assert np.abs(X - np.linalg.inv(L)).max() <= 1e-12 * np.abs(np.linalg.inv(L)).max()
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    B = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    C, _ = _mat(_run(driver, "mul %d\n%s\n" % (n, _fmt(A, B))), n)
    assert np.abs(C - A @ B).max() <= 1e-13 * np.abs(A @ B).max()
    H, _ = _mat(_run(driver, "herm %d\n%s\n" % (n, _fmt(A))), n)
    assert np.array_equal(H, A.conj().T)


def _lsq_input(K, beta, cols):
    lines = ["lsq %d %.17g %d" % (K, beta, len(cols))]
    for hn, h in cols:
        lines.append("%.17g" % hn)
        lines.append(_fmt(h))
    return "\n".join(lines) + "\n"


def test_givens_least_squares_matches_lstsq(driver):
    rng = np.random.default_rng(31)
    K, beta = 12, 2.5
    H = np.zeros((K + 1, K), complex)
    for j in range(K):
        H[:j + 1, j] = rng.standard_normal(j + 1) + 1j * rng.standard_normal(j + 1)
        H[j + 1, j] = 0.2 + rng.random()
    tok = _run(driver, _lsq_input(K, beta, [(H[j + 1, j].real, H[:j + 1, j]) for j in range(K)]))
    e1 = np.zeros(K + 1, complex)
    e1[0] = beta
    for j in range(K):
        ok, res = tok[2 * j], float(tok[2 * j + 1])
        y = np.linalg.lstsq(H[:j + 2, :j + 1], e1[:j + 2], rcond=None)[0]
        ref = np.linalg.norm(e1[:j + 2] - H[:j + 2, :j + 1] @ y)
        assert ok == "1"
        assert abs(res - ref) <= 1e-12 * beta, (j, res, ref)
    rest = tok[2 * K:]
    assert int(rest[0]) == K
    v = np.array([float(t) for t in rest[1:1 + 2 * K]])
    y = v[0::2] + 1j * v[1::2]
    yref = np.linalg.lstsq(H, e1, rcond=None)[0]
    assert np.linalg.norm(y - yref) <= 1e-10 * np.linalg.norm(yref)


def test_givens_refuses_zero_column(driver):
    """A last column whose rotated entries are all zero (h = 0, hn = 0): the
    update refuses it, the residual and the solution stay those of the
    earlier columns (no division by a zero diagonal)."""
    rng = np.random.default_rng(41)
    K, beta, n = 6, 1.0, 4
    H = np.zeros((K + 1, K), complex)
    for j in range(n):
        H[:j + 1, j] = rng.standard_normal(j + 1) + 1j * rng.standard_normal(j + 1)
        H[j + 1, j] = 0.5 + rng.random()
    cols = [(H[j + 1, j].real, H[:j + 1, j]) for j in range(n)] + [(0.0, np.zeros(n + 1, complex))]
    tok = _run(driver, _lsq_input(K, beta, cols))
    assert tok[2 * n] == "0"
    assert float(tok[2 * n + 1]) == float(tok[2 * n - 1])
    rest = tok[2 * (n + 1):]
    assert int(rest[0]) == n
    v = np.array([float(t) for t in rest[1:1 + 2 * n]])
    y = v[0::2] + 1j * v[1::2]
    e1 = np.zeros(n + 1, complex)
    e1[0] = beta
    yref = np.linalg.lstsq(H[:n + 1, :n], e1, rcond=None)[0]
    assert np.all(np.isfinite(y)) and np.linalg.norm(y - yref) <= 1e-10 * np.linalg.norm(yref)
