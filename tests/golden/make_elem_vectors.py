"""Writes tests/golden/elem_vectors.json: element integrals computed independently
of the oracle, to pin its regular (6x6 collapsed Gauss) and vertex-singular rules.

Quantities (for element vertices p1, p2, p3, y(xi) = xi1 p1 + xi2 p2 + (1 - xi1 - xi2) p3,
N = (xi1, xi2, 1 - xi1 - xi2), n = unit (p1 - p3) x (p2 - p3), G = exp(-ikr)/(4 pi r)):
    G_j = int_T N_j G dS
    H_j = int_T N_j dG/dn_y dS
    C   = int_T dG0/dn_y dS     (G0 = 1 / (4 pi r))
computed with scipy's adaptive quadrature (dblquad); the normal derivative is a
central finite difference of G along n (no closed-form derivative is used).
Regular cases integrate over the triangle in (xi1, xi2); singular cases (x at a
vertex) integrate in polar coordinates centred on the vertex, where the 1/r
singularity cancels against the area element.

Near-singular cases ("near": x at 0.25 and 0.5 element sizes above the centroid or
an edge midpoint) pin the opt-in FP64 near-field integration, where the 6x6 rule is
off by more than 1e-5.

Run: python tests/golden/make_elem_vectors.py
"""
import json
import os

import numpy as np
from scipy import integrate

HERE = os.path.dirname(os.path.abspath(__file__))


def G(k, x, y):
    r = np.linalg.norm(y - x)
    return np.exp(-1j * k * r) / (4 * np.pi * r)


def dGdn(k, x, y, n, h):
    return (G(k, x, y + h * n) - G(k, x, y - h * n)) / (2 * h)


def dG0dn(x, y, n, h):
    return (G(0.0, x, y + h * n) - G(0.0, x, y - h * n)).real / (2 * h)


def regular_case(k, x, p):
    p1, p2, p3 = p
    cr = np.cross(p1 - p3, p2 - p3)
    J = np.linalg.norm(cr)
    n = cr / J
    h = 1e-5 * np.linalg.norm(p1 - p3)
    out = {}

    def integ(f):
        re = integrate.dblquad(lambda b, a: f(a, b).real, 0, 1, 0, lambda a: 1 - a, epsabs=0, epsrel=1e-11)[0]
        im = integrate.dblquad(lambda b, a: f(a, b).imag, 0, 1, 0, lambda a: 1 - a, epsabs=0, epsrel=1e-11)[0]
        return J * complex(re, im)

    def y(a, b):
        return a * p1 + b * p2 + (1 - a - b) * p3

    N = [lambda a, b: a, lambda a, b: b, lambda a, b: 1 - a - b]
    out["G"] = [integ(lambda a, b, j=j: N[j](a, b) * G(k, x, y(a, b))) for j in range(3)]
    out["H"] = [integ(lambda a, b, j=j: N[j](a, b) * dGdn(k, x, y(a, b), n, h)) for j in range(3)]
    out["C"] = integ(lambda a, b: complex(dG0dn(x, y(a, b), n, h), 0)).real
    return out


def singular_case(k, v, p):
    """x = p[v]; polar coordinates (r, phi) in the element plane around x."""
    p = [np.asarray(q, float) for q in p]
    x = p[v]
    a, b = [p[i] for i in range(3) if i != v]
    cr = np.cross(p[0] - p[2], p[1] - p[2])
    J = np.linalg.norm(cr)
    n = cr / J
    e1 = (a - x) / np.linalg.norm(a - x)
    e2 = np.cross(n, e1)
    # polar angle of b in the (e1, e2) frame
    ang_b = np.arctan2(np.dot(b - x, e2), np.dot(b - x, e1))
    lo, hi = (0.0, ang_b) if ang_b > 0 else (ang_b, 0.0)
    # opposite edge a + t (b - a); distance along direction d(phi) to that edge
    Minv = None

    def rmax(phi):
        d = np.cos(phi) * e1 + np.sin(phi) * e2
        # solve x + s d = a + t (b - a) in the plane
        Mm = np.array([[np.dot(d, e1), -np.dot(b - a, e1)], [np.dot(d, e2), -np.dot(b - a, e2)]])
        rhs = np.array([np.dot(a - x, e1), np.dot(a - x, e2)])
        s, t = np.linalg.solve(Mm, rhs)
        return s

    # barycentric coordinates of a point y in the element
    T = np.column_stack([p[0] - p[2], p[1] - p[2]])

    def bary(y):
        xi = np.linalg.lstsq(T, y - p[2], rcond=None)[0]
        return np.array([xi[0], xi[1], 1 - xi[0] - xi[1]])

    def point(r, phi):
        return x + r * (np.cos(phi) * e1 + np.sin(phi) * e2)

    def integ(f):
        re = integrate.dblquad(lambda r, phi: (f(r, phi) * r).real, lo, hi, 0, rmax, epsabs=0, epsrel=1e-11)[0]
        im = integrate.dblquad(lambda r, phi: (f(r, phi) * r).imag, lo, hi, 0, rmax, epsabs=0, epsrel=1e-11)[0]
        return complex(re, im)

    out = {}
    out["G"] = [integ(lambda r, phi, j=j: bary(point(r, phi))[j] * G(k, x, point(r, phi))) for j in range(3)]
    return out


def main():
    rng = np.random.default_rng(20240611)
    tris = [
        np.array([[0.0, 0.0, 0.0], [0.01, 0.0, 0.0], [0.0, 0.012, 0.0]]),
        np.array([[0.1, 0.02, -0.03], [0.095, 0.035, -0.02], [0.088, 0.018, -0.012]]),
        np.array([[0.0, 0.0, 1.0], [0.2, 0.05, 0.97], [0.05, 0.21, 0.96]]),
    ]
    cases = []
    for ti, p in enumerate(tris):
        cen = p.mean(axis=0)
        hsz = max(np.linalg.norm(p[0] - p[1]), np.linalg.norm(p[1] - p[2]), np.linalg.norm(p[2] - p[0]))
        cr = np.cross(p[0] - p[2], p[1] - p[2])
        n = cr / np.linalg.norm(cr)
        for dist in (1.0, 2.0, 5.0):
            d = rng.standard_normal(3)
            d = d / np.linalg.norm(d) + 0.5 * n
            d /= np.linalg.norm(d)
            x = cen + dist * hsz * d
            for k in (0.5 / hsz, 30.0):
                r = regular_case(k, x, p)
                cases.append({"kind": "regular", "k": k, "x": x.tolist(), "p": p.tolist(), "dist_over_h": dist,
                              "G": [[z.real, z.imag] for z in r["G"]], "H": [[z.real, z.imag] for z in r["H"]],
                              "C": r["C"]})
        for v in range(3):
            for k in (1e-4, 0.5 / hsz):
                r = singular_case(k, v, p)
                cases.append({"kind": "singular", "k": k, "v": v, "p": p.tolist(),
                              "G": [[z.real, z.imag] for z in r["G"]]})
    # near-singular points over the two small triangles (drawn after the cases
    # above so that those stay the same)
    # (coordinates and k rounded to FP32, the library's input precision, so
    # that the device's FP64 near-field integrals can be compared at 1e-7)
    for ti, p in enumerate(tris[:2]):
        p = p.astype(np.float32).astype(np.float64)
        cen = p.mean(axis=0)
        hsz = max(np.linalg.norm(p[0] - p[1]), np.linalg.norm(p[1] - p[2]), np.linalg.norm(p[2] - p[0]))
        cr = np.cross(p[0] - p[2], p[1] - p[2])
        n = cr / np.linalg.norm(cr)
        mid = 0.5 * (p[0] + p[1])
        for where, base in (("centroid", cen), ("edge", mid)):
            for dist in (0.25, 0.5):
                x = (base + dist * hsz * n).astype(np.float32).astype(np.float64)
                for k in (float(np.float32(0.5 / hsz)), 30.0):
                    r = regular_case(k, x, p)
                    cases.append({"kind": "near", "where": where, "k": k, "x": x.tolist(), "p": p.tolist(),
                                  "dist_over_h": dist, "G": [[z.real, z.imag] for z in r["G"]],
                                  "H": [[z.real, z.imag] for z in r["H"]], "C": r["C"]})
    with open(os.path.join(HERE, "elem_vectors.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_elem_vectors.py", "cases": cases}, f, indent=1)
    print("wrote %d cases" % len(cases))


if __name__ == "__main__":
    main()
