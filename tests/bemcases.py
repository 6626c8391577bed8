"""Shared problem set-ups for the tests: meshes, CHIEF points, right-hand sides,
the analytic pulsating- and rigid-sphere solutions, and golden-file parsing."""
import os
import re

import numpy as np
from scipy import special

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def splitmix64(seed):
    s = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        s = (s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        yield (z ^ (z >> 31)) >> 11


def uniforms(seed, n):
    g = splitmix64(seed)
    return np.array([next(g) * (1.0 / 9007199254740992.0) for _ in range(n)])


def ball_points(seed, n, radius):
    """n points uniform in the ball |x| < radius (rejection from the cube)."""
    out = []
    u = iter(uniforms(seed, 300 * n + 300))
    while len(out) < n:
        p = np.array([2 * next(u) - 1, 2 * next(u) - 1, 2 * next(u) - 1]) * radius
        if np.linalg.norm(p) < radius:
            out.append(p)
    return np.array(out, np.float32)


def sphere_points(seed, n, radius):
    """n points uniform on the sphere |x| = radius (normalised Gaussians)."""
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((n, 3))
    return (radius * v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def read_golden(name):
    s = open(name if os.path.isabs(name) else os.path.join(GOLDEN, name)).read()
    return np.array([complex(float(a), float(b)) for a, b in re.findall(r"\(([-\d.eE+]+),([-\d.eE+]+)\)", s)])


def pulsating_sphere(r, a, k, rho, c):
    """sphere.m:11 -- p = i rho omega a^2/(r (1 + i k a)) exp(-i k (r - a))."""
    omega = k * c
    return 1j * rho * omega * a ** 2 / (r * (1 + 1j * k * a)) * np.exp(-1j * k * (r - a))


def rigid_sphere_surface(a, k, cos_theta, nmax=None):
    """Total pressure on a rigid sphere of radius a for p_inc = exp(-i k z)
    (e^{+i omega t}): p = -(i/(ka)^2) sum_n (2n+1) (-i)^n P_n(cos t) / h_n^(2)'(ka)."""
    x = k * a
    if nmax is None:
        nmax = int(x + 4 * x ** (1 / 3) + 12)
    p = np.zeros_like(np.asarray(cos_theta, float), dtype=complex)
    for n in range(nmax + 1):
        dh = special.spherical_jn(n, x, derivative=True) - 1j * special.spherical_yn(n, x, derivative=True)
        p += (2 * n + 1) * (-1j) ** n * special.eval_legendre(n, cos_theta) / dh
    return -1j / x ** 2 * p


def mesh_arrays(mesh):
    """(pnts float32 N x 3, elems int32 E x 3, bc complex E x 3) of a bem Mesh."""
    e = mesh.elems()
    return mesh.pnts(), np.ascontiguousarray(e["nodes"]), np.ascontiguousarray(e["bc"]).astype(np.complex128)


def outward_vertex_normals(pnts, elems):
    """Area-weighted vertex normals pointing out of the body.  Element normals
    (p1 - p3) x (p2 - p3) point into the body for the reference's winding, so
    they are negated."""
    p = np.asarray(pnts, np.float64)
    e = np.asarray(elems)
    fn = np.cross(p[e[:, 0]] - p[e[:, 2]], p[e[:, 1]] - p[e[:, 2]])
    vn = np.zeros_like(p)
    for j in range(3):
        np.add.at(vn, e[:, j], fn)
    vn = -vn
    return vn / np.linalg.norm(vn, axis=1, keepdims=True)


def sources_outside(mesh, seed, n, offset):
    """n point sources `offset` metres outside the surface along the outward
    normal at n seeded random vertices."""
    p, e, _ = mesh_arrays(mesh)
    idx = np.random.default_rng(seed).choice(p.shape[0], n, replace=False)
    vn = outward_vertex_normals(p, e)
    return (p[idx].astype(np.float64) + offset * vn[idx]).astype(np.float32)


def large_case(bem):
    """Large workload: seeded perturbed ellipsoid (semi-axes 0.075/0.09/0.11 m,
    8 spherical-harmonic bumps up to degree 6, amplitude <= 5%) at icosphere
    subdivision 6 (40962 nodes, 81920 triangles), rigid, 12 kHz, 32 CHIEF points
    in r < 0.05 m, 32 unit point sources 1 cm outside the surface."""
    m = bem.Mesh().ellipsoid(6, (0.075, 0.09, 0.11), nbumps=8, lmax=6, amp=0.05, seed=105)
    m.set_bc(0, 1, 0)
    m.set_chief(ball_points(1051, 32, 0.05))
    src = sources_outside(m, 1052, 32, 0.01)
    return m, src


def rigid_sphere_field(a, k, x):
    """Total pressure at exterior points x (P, 3) around a rigid sphere of
    radius a for p_inc = exp(-i k z) (e^{+i omega t}):
    p = sum_n (2n+1) (-i)^n [j_n(kr) - j_n'(ka) / h_n^(2)'(ka) h_n^(2)(kr)] P_n(cos t)."""
    x = np.asarray(x, np.float64)
    r = np.linalg.norm(x, axis=1)
    ct = x[:, 2] / r
    ka = k * a
    nmax = int(max(ka, k * r.max()) + 4 * max(ka, 1.0) ** (1 / 3) + 20)
    p = np.zeros(r.shape, complex)
    for n in range(nmax + 1):
        dh = special.spherical_jn(n, ka, derivative=True) - 1j * special.spherical_yn(n, ka, derivative=True)
        dj = special.spherical_jn(n, ka, derivative=True)
        h = special.spherical_jn(n, k * r) - 1j * special.spherical_yn(n, k * r)
        p += (2 * n + 1) * (-1j) ** n * (special.spherical_jn(n, k * r) - dj / dh * h) * special.eval_legendre(n, ct)
    return p


def two_spheres(bem, subdiv, radius, gap=0.5, bc=(0, 1, 0)):
    """Two icospheres of the given radius centred at x = -/+ (1 + gap/2) radius,
    so their surfaces are gap * radius apart: one mesh of 2 (10 4^s + 2) nodes."""
    a = bem.Mesh().icosphere(subdiv, radius)
    p, e = a.pnts(), a.elems()["nodes"]
    d = (1.0 + gap / 2) * radius
    pts = np.vstack([p - [d, 0, 0], p + [d, 0, 0]]).astype(np.float32)
    els = np.vstack([e, e + p.shape[0]]).astype(np.int32)
    return bem.Mesh().set(pts, els, bc)
