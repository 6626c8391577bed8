/*
 * TEST INFRASTRUCTURE ONLY -- the oracle.
 *
 * A CPU restatement, in plain C, of the reference's boundary-element path
 * (ZiqiFrancisFan/BEM_GPU_Clean_CPP): Gauss-Legendre tables, the regular and
 * vertex-singular element rules, collocation assembly with CHIEF rows, exterior
 * field evaluation, a Householder least-squares solve and GMRES.  It is the
 * checker for the HIP library: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it.  The product never links or calls it.
 *
 * Built by oracle/Makefile into oracle/liboracle.so (gcc -O2 -fopenmp).
 *
 * How each part is pinned (tests/test_oracle.py):
 *   orc_gauss_legendre        textbook 6-point nodes/weights (Abramowitz &
 *                             Stegun 25.4) and numpy.polynomial.legendre.leggauss
 *   orc_elem_regular_*        adaptive scipy quadrature of the same integrals
 *                             (golden vectors in tests/golden/), k=0 solid-angle
 *                             identity
 *   orc_elem_singular_*       closed form of the k=0 vertex integral of 1/r,
 *                             h = c = 0 on flat elements, mpmath quadrature
 *   orc_assemble_*            A*1 = 1 at k = 0 on node rows (h reduces to c),
 *                             pulsating sphere: reference golden files
 *                             sphere171Hz_NoCHIEF / sphere171Hz_CHIEF and the
 *                             analytic formula of sphere.m:11
 *   orc_field_*               the same golden files (exterior points r=1.2..3)
 *   orc_lsq_qr_*              numpy.linalg.lstsq (LAPACK)
 *   orc_gmres_*               numpy.linalg.solve (LAPACK)
 *   orc_read_obj              vertex/face counts printed in the reference OBJ
 *                             headers, signed volume of the shipped meshes
 *   orc_ray_inside            convex meshes with known interior points
 */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.1415926535897932 /* numerical.h:36-38 */
#define ORC_MAXORDER 32
#define ORC_MAXRHS 64
#define ORC_EPS 0.00005f /* numerical.h:124-126 */

/* ---------------- FP64 oracle ---------------- */
#define REAL double
#define CPLX double complex
#define SFX _d
#define SQRT sqrt
#define COS cos
#define SIN sin
#define CREAL creal
#define CIMAG cimag
#define CONJ conj
#define CABS cabs
#include "bem_oracle_impl.h"
#undef REAL
#undef CPLX
#undef SFX
#undef SQRT
#undef COS
#undef SIN
#undef CREAL
#undef CIMAG
#undef CONJ
#undef CABS

/* ---------------- complex FP32 reference-mimic ---------------- */
#define REAL float
#define CPLX float complex
#define SFX _f
#define SQRT sqrtf
#define COS cosf
#define SIN sinf
#define CREAL crealf
#define CIMAG cimagf
#define CONJ conjf
#define CABS cabsf
#include "bem_oracle_impl.h"

/* Gauss-Legendre nodes/weights by Golub-Welsch, the method of
 * gaussQuad::genGaussParams (numerical.cu:186-225): symmetric tridiagonal
 * Jacobi matrix with off-diagonal (i+1)/sqrt((2i+2)^2-1), nodes = eigenvalues,
 * weights = 2 * (first eigenvector component)^2.  The reference calls GSL's
 * gsl_eigen_symmv; here a cyclic Jacobi eigen-solver does the same job, and
 * the pairs are sorted by node (GSL leaves them unsorted). */
int orc_gauss_legendre(int n, double *x, double *w) {
    if (n < 1 || n > ORC_MAXORDER) return 1;
    double a[ORC_MAXORDER][ORC_MAXORDER], v[ORC_MAXORDER][ORC_MAXORDER];
    memset(a, 0, sizeof a);
    memset(v, 0, sizeof v);
    for (int i = 0; i < n; i++) v[i][i] = 1;
    for (int i = 0; i < n - 1; i++) {
        double t = sqrt(pow(2.0 * (i + 1), 2) - 1);
        a[i][i + 1] = a[i + 1][i] = (i + 1) / t;
    }
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) off += a[p][q] * a[p][q];
        if (off < 1e-30) break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                if (fabs(a[p][q]) < 1e-300) continue;
                double th = (a[q][q] - a[p][p]) / (2 * a[p][q]);
                double t = (th >= 0 ? 1 : -1) / (fabs(th) + sqrt(th * th + 1));
                double c = 1 / sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < n; k++) { /* A <- A J */
                    double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) { /* A <- J^T A */
                    double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) { /* V <- V J */
                    double vkp = v[k][p], vkq = v[k][q];
                    v[k][p] = c * vkp - s * vkq;
                    v[k][q] = s * vkp + c * vkq;
                }
            }
    }
    int idx[ORC_MAXORDER];
    for (int i = 0; i < n; i++) idx[i] = i;
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++)
            if (a[idx[j]][idx[j]] < a[idx[i]][idx[i]]) {
                int t = idx[i]; idx[i] = idx[j]; idx[j] = t;
            }
    for (int i = 0; i < n; i++) {
        x[i] = a[idx[i]][idx[i]];
        w[i] = 2 * v[0][idx[i]] * v[0][idx[i]];
    }
    return 0;
}

/* Both precisions read the same table (gaussQuad::sendToDevice,
 * numerical.cu:227-231). */
int orc_quad_init(int n) {
    double x[ORC_MAXORDER], w[ORC_MAXORDER];
    if (orc_gauss_legendre(n, x, w)) return 1;
    orc_set_quadrature_d(n, x, w);
    orc_set_quadrature_f(n, x, w);
    return 0;
}

/* OBJ text reader, the semantics of mesh::readObj (mesh.cu:1215-1265): a line
 * whose first character is 'v' is a vertex "%s %f %f %f", a line whose first
 * character is 'f' is a face "%s %d %d %d" with 1-based indices stored 0-based.
 * (The reference reads 48-character chunks; lines here are read whole.)
 * Call with NULL arrays to get the counts. */
int orc_read_obj(const char *path, float *pnts, int maxPnts, int *elems, int maxElems,
                 int *nPnts, int *nElems) {
    FILE *fp = fopen(path, "r");
    if (!fp) return 1;
    char line[4096], type[64];
    int np = 0, ne = 0;
    while (fgets(line, sizeof line, fp)) {
        if (line[0] == 'v') {
            float c[3] = {0, 0, 0};
            sscanf(line, "%63s %f %f %f", type, &c[0], &c[1], &c[2]);
            if (pnts && np < maxPnts) memcpy(pnts + 3 * np, c, sizeof c);
            np++;
        }
        if (line[0] == 'f') {
            int t[3] = {0, 0, 0};
            sscanf(line, "%63s %d %d %d", type, &t[0], &t[1], &t[2]);
            if (elems && ne < maxElems)
                for (int j = 0; j < 3; j++) elems[3 * ne + j] = t[j] - 1;
            ne++;
        }
    }
    fclose(fp);
    *nPnts = np;
    *nElems = ne;
    return 0;
}

/* ---- CHIEF interior test: rayTrnglInt (mesh.cu:1184-1189), rayPlaneInt
 * (mesh.cu:1128-1147), isInsideTrngl (mesh.cu:1167-1182), inObj
 * (numerical.cu:119-131), all in float like the reference. ---- */
static void v3sub(const float *a, const float *b, float *o) { o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2]; }
static void v3cross(const float *a, const float *b, float *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static float v3dot(const float *a, const float *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static int v3nrmlzd(const float *a, float *o) {
    float n = sqrtf(v3dot(a, a));
    if (n == 0) { o[0] = o[1] = o[2] = NAN; return 0; }
    o[0] = a[0] / n; o[1] = a[1] / n; o[2] = a[2] / n;
    return 1;
}
static int v3legal(const float *a) { return a[0] == a[0] && a[1] == a[1] && a[2] == a[2]; }
static int v3equal(const float *a, const float *b) {
    return fabsf(a[0] - b[0]) < ORC_EPS && fabsf(a[1] - b[1]) < ORC_EPS && fabsf(a[2] - b[2]) < ORC_EPS;
}

int orc_ray_trngl_int(const float *sp, const float *dir, const float *p1, const float *p2, const float *p3) {
    float e1[3], e2[3], n[3], ip[3], d[3];
    v3sub(p2, p1, e1);
    v3sub(p3, p1, e2);
    v3cross(e1, e2, n);
    v3sub(sp, p1, d);
    if (v3dot(n, d) == 0) {
        memcpy(ip, sp, sizeof ip);
    } else if (v3dot(n, dir) == 0) {
        ip[0] = ip[1] = ip[2] = NAN;
    } else {
        float t = (v3dot(n, p1) - v3dot(n, sp)) / v3dot(n, dir);
        if (t > 0) {
            for (int i = 0; i < 3; i++) ip[i] = sp[i] + t * dir[i];
        } else {
            ip[0] = ip[1] = ip[2] = NAN;
        }
    }
    if (!v3legal(ip)) return 0;
    float v12[3], v23[3], v31[3], v1p[3], v2p[3], v3p[3], t1[3], t2[3], t3[3], n1[3], n2[3], n3[3];
    v3sub(p2, p1, v12); v3sub(p3, p2, v23); v3sub(p1, p3, v31);
    v3sub(ip, p1, v1p); v3sub(ip, p2, v2p); v3sub(ip, p3, v3p);
    v3cross(v12, v1p, t1); v3cross(v23, v2p, t2); v3cross(v31, v3p, t3);
    v3nrmlzd(t1, n1); v3nrmlzd(t2, n2); v3nrmlzd(t3, n3);
    return v3equal(n1, n2) && v3equal(n2, n3);
}

/* Parity of ray hits from sp along dir: 1 = inside (inObj). */
int orc_ray_inside(const float *sp, const float *dir, const float *pnts, const int *elems, int nElem) {
    int cnt = 0;
    for (int e = 0; e < nElem; e++)
        cnt += orc_ray_trngl_int(sp, dir, pnts + 3 * elems[3 * e], pnts + 3 * elems[3 * e + 1],
                                 pnts + 3 * elems[3 * e + 2]);
    return cnt % 2;
}
