"""FP64 near-singular integration (bem_set_near_field) on the GPU.

The element integrals of a point a quarter or half of the element size away
(tests/golden/elem_vectors.json, "near" cases: scipy adaptive quadrature with
the normal derivative as a finite difference, no formula shared with the
library) against single-element systems assembled through the C-ABI: the
triangle's three vertices are the nodes and the near points are extra rows, so
with the boundary condition (0, 1, 0) row x holds H_j(x) and with (1, 1, 0)
it holds H_j(x) - g_j(x), g_j = i rho omega G_j.

  * default (off): the entries equal the oracle's 6x6 rule (the reference's
    g_h_c_l_nsgl, mesh.cu:597-672) and are off by more than 1e-5 at 0.25 h;
  * on: FP64 entries within 1e-7 of the adaptive integrals (relative to the
    largest entry of the row), FP32 entries within 5e-7;
  * on, a whole mesh: CHIEF rows 0.3 element sizes inside a perturbed
    ellipsoid sum to the interior solid-angle integral (1) at k -> 0 to 1e-8,
    where the 6x6 rule misses by more than 1e-4.
"""
import json
import os

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
    bem.set_near_field(0.0)
    bem.set_medium(1.2041, 343.21)


def _near_groups():
    cases = json.load(open(os.path.join(C.GOLDEN, "elem_vectors.json")))["cases"]
    groups = {}
    for c in cases:
        if c["kind"] != "near":
            continue
        key = (json.dumps(c["p"]), c["k"])
        groups.setdefault(key, []).append(c)
    return list(groups.values())


def _single_element_rows(p, xs, k, prec, bc):
    m = bem.Mesh().set(np.asarray(p, np.float32), [[0, 1, 2]], bc)
    m.set_chief(np.asarray(xs, np.float32))
    s = bem.Solver(m)
    s.assemble(k, prec)
    rows = s.rows(np.arange(3, 3 + len(xs)))
    s.close()
    return rows


def _entries(p, xs, k, prec):
    h = _single_element_rows(p, xs, k, prec, (0, 1, 0))
    hg = _single_element_rows(p, xs, k, prec, (1, 1, 0))
    return h - hg, h  # g, H


@pytest.mark.parametrize("prec", [bem.PREC_FP64, bem.PREC_FP32])
def test_near_field_matches_adaptive_integrals(prec):
    worst_off = worst_on = 0.0
    for group in _near_groups():
        p = np.array(group[0]["p"])
        k = float(group[0]["k"])
        xs = np.array([c["x"] for c in group])
        G = np.array([[complex(*z) for z in c["G"]] for c in group])
        H = np.array([[complex(*z) for z in c["H"]] for c in group])
        gref = 1j * RHO * k * SPEED * G
        bem.set_near_field(0.0)
        g0, h0 = _entries(p, xs, k, prec)
        bem.set_near_field(1.0)
        g1, h1 = _entries(p, xs, k, prec)
        bem.set_near_field(0.0)
        for q, c in enumerate(group):
            scale = max(np.abs(H[q]).max(), np.abs(gref[q]).max())
            e_off = max(np.abs(h0[q] - H[q]).max(), np.abs(g0[q] - gref[q]).max()) / scale
            e_on = max(np.abs(h1[q] - H[q]).max(), np.abs(g1[q] - gref[q]).max()) / scale
            print("case: %s at %.2f h, k %.1f, prec %d: rule %.2e, near field %.2e" %
                  (c["where"], c["dist_over_h"], k, prec, e_off, e_on))
            if c["dist_over_h"] == 0.25:
                assert e_off > 1e-5
            worst_off = max(worst_off, e_off)
            worst_on = max(worst_on, e_on)
            if prec == bem.PREC_FP64:
                # default mode is the reference's rule: the oracle's 6x6 values
                go, ho, _ = O.elem_regular(k, c["x"], c["p"], dens=RHO, speed=SPEED)
                assert np.abs(h0[q] - ho).max() <= 1e-10 * scale
                assert np.abs(g0[q] - go).max() <= 1e-10 * scale
    print("worst: rule %.3e, near field %.3e" % (worst_off, worst_on))
    assert worst_on <= (1e-7 if prec == bem.PREC_FP64 else 5e-7)


def test_near_field_on_a_mesh():
    """A perturbed ellipsoid (s3) with 16 CHIEF rows 0.3 element sizes inside
    the surface.  At k -> 0 a CHIEF row sums to the solid-angle integral
    int dG0/dn dS = 1 of an interior point (normals into the body), so the
    quadrature error near the surface shows in the row sums: the 6x6 rule
    misses by more than 1e-4, the near-field integration by less than 1e-8.
    The node rows keep A.1 = 1 to rounding in both modes (their h and c terms
    use the same points, corrected together)."""
    m = bem.Mesh().ellipsoid(3, (0.075, 0.09, 0.11), nbumps=8, lmax=6, amp=0.05, seed=31)
    m.set_bc(0, 1, 0)
    p, e, _ = C.mesh_arrays(m)
    N = m.num_pnts
    edges = np.concatenate([np.linalg.norm(p[e[:, a]] - p[e[:, b]], axis=1) for a, b in ((0, 1), (1, 2), (2, 0))])
    h = float(edges.mean())
    vn = C.outward_vertex_normals(p, e)
    idx = np.random.default_rng(33).choice(N, 16, replace=False)
    chief = (p[idx].astype(np.float64) - 0.3 * h * vn[idx]).astype(np.float32)
    m.set_chief(chief)
    out = {}
    for f in (0.0, 1.0):
        bem.set_near_field(f)
        s = bem.Solver(m)
        s.assemble(1e-4, bem.PREC_FP64)
        out[f] = s.rows(np.arange(N + 16))
        s.close()
    bem.set_near_field(0.0)
    A0, A1 = out[0.0], out[1.0]
    c0 = np.abs(A0[N:].sum(axis=1) - 1).max()
    c1 = np.abs(A1[N:].sum(axis=1) - 1).max()
    n0 = np.abs(A0[:N].sum(axis=1) - 1).max()
    n1 = np.abs(A1[:N].sum(axis=1) - 1).max()
    changed = (np.abs(A1 - A0).max(axis=1) > 0).sum()
    print("case: s3 ellipsoid, k 1e-4, h %.4f: %d of %d rows changed; CHIEF rows |sum - 1| rule %.3e, near field "
          "%.3e; node rows |A.1 - 1| rule %.3e, near field %.3e" % (h, changed, N + 16, c0, c1, n0, n1))
    assert c0 > 1e-4
    assert c1 < 1e-8
    assert n0 < 1e-9 and n1 < 1e-9
    assert changed > N // 2
