"""Frequencies fused into one lockstep GMRES in the batch driver (opt-in).

With BEM_BATCH_FUSE=4, square FP32 systems (no CHIEF rows) with several
right-hand sides go through bem_batch_solve in groups of 4 frequencies whose
matrices are resident
side by side; every Arnoldi step is one MFMA product launch over the group
(grid.z = frequency) and one host wait per chunk for all its systems.  On a
seeded perturbed ellipsoid at icosphere subdivision 3 (642 nodes), 7
frequencies 400 Hz - 6 kHz (so the last group is ragged) and 5 plane waves
each, refined to 1e-10:
  * each frequency's solution agrees with the same batch solved one frequency
    at a time (the default) to 1e-8, and with the FP64 oracle's solution
    to 1e-5 at the lowest and the highest frequency;
  * the results do not depend on the worker count (1 or 3 workers: equal bit
    for bit).
"""
import numpy as np
import pytest

import bem_gpu_clean_cpp_amd as bem
import bemcases as C
import oracle_py as O

pytestmark = pytest.mark.gpu

RHO, SPEED = float(np.float32(1.21)), float(np.float32(343.0))
KW = dict(restart=40, tol=1e-6, maxit=6000, refine_tol=1e-10, max_refine=50)


@pytest.fixture(autouse=True, scope="module")
def _setup():
    bem.quad_init(6)
    O.quad_init(6)
    bem.set_medium(RHO, SPEED)
    yield
    bem.set_medium(1.2041, 343.21)


def _problem():
    m = bem.Mesh().ellipsoid(3, (0.075, 0.09, 0.11), nbumps=8, lmax=6, amp=0.05, seed=321)
    m.set_bc(0, 1, 0)
    freqs = 400.0 * 15.0 ** (np.arange(7) / 6.0)
    ks = (2 * np.pi * freqs / SPEED).astype(np.float32)
    d = C.sphere_points(322, 5, 1.0)
    return m, ks, d


def test_fused_batch_matches_one_by_one_and_oracle(monkeypatch):
    m, ks, d = _problem()
    assert m.num_chief == 0
    monkeypatch.setenv("BEM_BATCH_FUSE", "4")
    Xf, sf = bem.batch_solve(m, ks, bem.RHS_PLANE, d, workers=3, **KW)
    X1, s1 = bem.batch_solve(m, ks, bem.RHS_PLANE, d, workers=1, **KW)
    monkeypatch.delenv("BEM_BATCH_FUSE")
    Xs, ss = bem.batch_solve(m, ks, bem.RHS_PLANE, d, workers=3, **KW)
    for f in range(len(ks)):
        rel = np.linalg.norm(Xf[:, :, f] - Xs[:, :, f], axis=0) / np.linalg.norm(Xs[:, :, f], axis=0)
        print("case: fused batch f %d (k %.2f): fused %s, alone %s, difference %.3g" %
              (f, ks[f], sf[f], ss[f], rel.max()), flush=True)
        assert sf[f]["converged"] == 1 and ss[f]["converged"] == 1
        assert rel.max() < 1e-8
    # fused groups: the same 4 iteration counts across each group
    assert len({sf[f]["iterations"] for f in range(3, 7)}) == 1
    assert np.array_equal(Xf, X1)
    p, e, bc = C.mesh_arrays(m)
    pts = p.astype(np.float64)
    for f in (0, len(ks) - 1):
        k = float(ks[f])
        binit = np.exp(-1j * k * pts @ np.asarray(d, np.float64).T)
        A, B = O.assemble(k, pts, p.shape[0], e, bc, binit=binit, dens=RHO, speed=SPEED)
        xo = np.linalg.solve(A, B)
        err = np.linalg.norm(Xf[:, :, f] - xo, axis=0) / np.linalg.norm(xo, axis=0)
        print("case: fused batch f %d vs FP64 oracle: %.3g" % (f, err.max()), flush=True)
        assert err.max() < 1e-5
