/*
 * TEST INFRASTRUCTURE ONLY -- the oracle.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and only as the checker (or, for
 * the cpu_baseline leg, as the timed CPU port).  The product never calls it.
 *
 * Precision-generic body of the CPU restatement of the reference's BEM path.
 * bem_oracle.c includes this file twice: once with REAL=double (suffix _d, the
 * FP64 oracle) and once with REAL=float (suffix _f, the "reference-mimic" that
 * computes in complex FP32 like the reference does, used as the CPU baseline).
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Summation order is fixed (row-wise, element order), unlike the reference's
 * float atomics, so results are repeatable for any thread count.
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define F(name) CAT(name, SFX)

typedef REAL F(real_t);

/* Quadrature table the integrators read, like the reference's __constant__
 * INTPNTS/INTWGTS (numerical.cu:17-19, filled by gaussQuad::sendToDevice,
 * numerical.cu:227-231). */
static REAL F(g_qp)[ORC_MAXORDER];
static REAL F(g_qw)[ORC_MAXORDER];
static int F(g_qn) = 0;

void F(orc_set_quadrature)(int n, const double *x, const double *w) {
    F(g_qn) = n;
    for (int i = 0; i < n; i++) {
        F(g_qp)[i] = (REAL)x[i];
        F(g_qw)[i] = (REAL)w[i];
    }
}

/* Literal restatement of g_h_c_l_nsgl (mesh.cu:597-672) for a general Duffy
 * map.  `map` selects the local coordinates of the collapsed square:
 *   3: (xi1, xi2) = (rho(1-theta), rho - xi1)      regular rule, mesh.cu:633-634
 *   0: (1-rho, rho(1-theta))                       vertex 1, mesh.cu:272-273
 *   1: (rho*theta, 1-rho)                          vertex 2, mesh.cu:312-313
 *   2: (rho(1-theta), rho*theta)                   vertex 3, mesh.cu:352-353
 * Outputs g[3], h[3] (complex, interleaved) and c (real):
 *   g_j = i*rho*omega * sum w N_j G,  h_j = sum w N_j dG/dn,  c = sum w dG0/dn,
 * with G = exp(-ikr)/(4 pi r) (numerical.cu:70-74), dG/dn = G(-1/r - ik) dr/dn
 * (mesh.cu:130-136) and dG0/dn = -dr/dn / (4 pi r^2) (mesh.cu:125-128). */
static void F(elem_rule)(int map, REAL k, REAL dens, REAL speed, const REAL x[3],
                         const REAL p[9], CPLX g[3], CPLX h[3], REAL *cc) {
    const REAL *p1 = p, *p2 = p + 3, *p3 = p + 6;
    REAL e1[3], e2[3], cr[3], nrm[3];
    for (int i = 0; i < 3; i++) {
        e1[i] = p1[i] - p3[i];
        e2[i] = p2[i] - p3[i];
    }
    /* crossProd = (p1-p3) x (p2-p3), mesh.cu:612 / mesh.cu:64-70 */
    cr[0] = e1[1] * e2[2] - e1[2] * e2[1];
    cr[1] = e1[2] * e2[0] - e1[0] * e2[2];
    cr[2] = e1[0] * e2[1] - e1[1] * e2[0];
    REAL J = SQRT(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    for (int i = 0; i < 3; i++) nrm[i] = cr[i] / J; /* nrmlzd, mesh.cu:1119-1126 */
    const REAL omega = k * speed;
    const REAL prodRhoOmega = dens * omega;
    const REAL fourPI = (REAL)4.0 * (REAL)ORC_PI;
    CPLX gs[3] = {0, 0, 0}, hs[3] = {0, 0, 0};
    REAL cs = 0;
    const int nq = F(g_qn);
    for (int n = 0; n < nq; n++) {
        REAL eta2 = F(g_qp)[n], wn = F(g_qw)[n];
        REAL theta = (REAL)0.5 + (REAL)0.5 * eta2;
        for (int m = 0; m < nq; m++) {
            REAL eta1 = F(g_qp)[m], wm = F(g_qw)[m];
            REAL rho = (REAL)0.5 + (REAL)0.5 * eta1;
            REAL temp = (REAL)0.25 * wn * wm * rho * J;
            REAL xi1, xi2;
            switch (map) {
            case 0: xi1 = 1 - rho; xi2 = rho * (1 - theta); break;
            case 1: xi1 = rho * theta; xi2 = 1 - rho; break;
            case 2: xi1 = rho * (1 - theta); xi2 = rho * theta; break;
            default: xi1 = rho * (1 - theta); xi2 = rho - xi1; break;
            }
            REAL N1 = xi1, N2 = xi2, N3 = (REAL)1 - xi1 - xi2; /* mesh.cu:1656-1666 */
            REAL y[3], rv[3];
            for (int i = 0; i < 3; i++) {
                y[i] = N1 * p1[i] + N2 * p2[i] + N3 * p3[i]; /* xiToElem, mesh.cu:1692-1696 */
                rv[i] = y[i] - x[i];
            }
            REAL radius = SQRT(rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]);
            REAL prpn2 = (rv[0] * nrm[0] + rv[1] * nrm[1] + rv[2] * nrm[2]) / radius;
            REAL pPsiLpn2 = -(REAL)1 / fourPI / (radius * radius) * prpn2;
            CPLX Psi = (COS(-k * radius) + I * SIN(-k * radius)) / (fourPI * radius);
            CPLX pPsipn2 = Psi * ((REAL)(-1) / radius - I * k) * prpn2;
            REAL Nj[3] = {N1, N2, N3};
            for (int j = 0; j < 3; j++) {
                gs[j] += temp * Nj[j] * Psi;
                hs[j] += temp * Nj[j] * pPsipn2;
            }
            cs += temp * pPsiLpn2;
        }
    }
    for (int j = 0; j < 3; j++) {
        g[j] = I * prodRhoOmega * gs[j]; /* mesh.cu:668-670 */
        h[j] = hs[j];
    }
    *cc = cs;
}

/* Regular element, x not a vertex: g_h_c_l_nsgl (mesh.cu:597-672).
 * g, h are 3 complex each as 6 interleaved reals. */
void F(orc_elem_regular)(REAL k, REAL dens, REAL speed, const REAL *x, const REAL *p,
                         REAL *g, REAL *h, REAL *c) {
    CPLX gg[3], hh[3];
    F(elem_rule)(3, k, dens, speed, x, p, gg, hh, c);
    for (int j = 0; j < 3; j++) {
        g[2 * j] = CREAL(gg[j]); g[2 * j + 1] = CIMAG(gg[j]);
        h[2 * j] = CREAL(hh[j]); h[2 * j + 1] = CIMAG(hh[j]);
    }
}

/* Singular element, x = vertex v of the element: g_l_sgl{1,2,3}
 * (mesh.cu:253-371), h_l_sgl{1,2,3} (mesh.cu:373-491), c_l_sgl{1,2,3}
 * (mesh.cu:519-595); the collocation point is the vertex itself, as
 * atomicPntsElems_sgl passes it (atomicFuncs.cu:130-149). */
void F(orc_elem_singular)(int v, REAL k, REAL dens, REAL speed, const REAL *p,
                          REAL *g, REAL *h, REAL *c) {
    CPLX gg[3], hh[3];
    F(elem_rule)(v, k, dens, speed, p + 3 * v, p, gg, hh, c);
    for (int j = 0; j < 3; j++) {
        g[2 * j] = CREAL(gg[j]); g[2 * j + 1] = CIMAG(gg[j]);
        h[2 * j] = CREAL(hh[j]); h[2 * j + 1] = CIMAG(hh[j]);
    }
}

/* Collocation system, the algorithm of atomicGenSystem (atomicFuncs.cu:218-298)
 * with the row order of bemSystem (mesh.cu:993-999): rows 0..nNod-1 are the
 * nodes, rows nNod..nNod+nChief-1 the CHIEF points; columns are nodes.
 *   A starts as I on the node rows (atomicFuncs.cu:227-235), B as Binit
 *   (atomicFuncs.cu:238-242: q*G(x, src), zero when q = 0);
 *   regular pairs (atomicFuncs.cu:69-117): A[x, nodes[j]] += h_j - (ca/cb) g_j,
 *     A[x, x] -= c on node rows, B[x, s] -= (cc/cb) sum_j g_j;
 *   singular pairs (atomicFuncs.cu:119-216): the same with the vertex rules.
 * pnts: (nNod+nChief) x 3, elems: nElem x 3 (0-based), bc: nElem x 3 complex
 * (ca, cb, cc) interleaved, A: column-major complex, lda >= M; B: M x nrhs.
 * Rows are independent, so OpenMP over rows keeps the per-row order fixed. */
/* Rows [r0, r1) only; row i is stored at A[(i - r0) + col * lda]. */
int F(orc_assemble_rows)(REAL k, REAL dens, REAL speed, const REAL *pnts, int nNod, int nChief,
                         const int *elems, const REAL *bc, int nElem, const REAL *Binit, int nrhs,
                         REAL *A, long lda, REAL *B, long ldb, int r0, int r1, int nthreads) {
    const int M = nNod + nChief;
    if (nNod <= 0 || nElem <= 0 || r0 < 0 || r1 > M || r0 > r1 || lda < r1 - r0 ||
        (nrhs > 0 && ldb < r1 - r0))
        return 1;
    CPLX *Ac = (CPLX *)A, *Bc = (CPLX *)B;
    const CPLX *Bi = (const CPLX *)Binit;
    const CPLX *bcc = (const CPLX *)bc;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads > 0 ? nthreads : 1)
    for (int i = r0; i < r1; i++) {
        CPLX *Ar = Ac - r0, *Br = Bc - r0;
        for (int j = 0; j < nNod; j++) Ar[(long)j * lda + i] = (i == j) ? 1 : 0;
        CPLX bterm = 0;
        const REAL *x = pnts + 3 * (long)i;
        for (int e = 0; e < nElem; e++) {
            const int *nd = elems + 3 * (long)e;
            REAL p[9];
            for (int v = 0; v < 3; v++)
                for (int d = 0; d < 3; d++) p[3 * v + d] = pnts[3 * (long)nd[v] + d];
            int sv = -1;
            for (int v = 0; v < 3; v++)
                if (nd[v] == i) sv = v; /* exclusion test, atomicFuncs.cu:75-76 */
            CPLX g[3], h[3];
            REAL c;
            F(elem_rule)(sv < 0 ? 3 : sv, k, dens, speed, x, p, g, h, &c);
            CPLX beta = bcc[3 * e] / bcc[3 * e + 1];
            CPLX gam = bcc[3 * e + 2] / bcc[3 * e + 1];
            for (int j = 0; j < 3; j++) Ar[(long)nd[j] * lda + i] += h[j] - beta * g[j];
            if (i < nNod) Ar[(long)i * lda + i] -= c;
            bterm += gam * (g[0] + g[1] + g[2]);
        }
        for (int s = 0; s < nrhs; s++) Br[(long)s * ldb + i] = (Bi ? Bi[(long)s * M + i] : 0) - bterm;
    }
    return 0;
}

int F(orc_assemble)(REAL k, REAL dens, REAL speed, const REAL *pnts, int nNod, int nChief,
                    const int *elems, const REAL *bc, int nElem, const REAL *Binit, int nrhs,
                    REAL *A, long lda, REAL *B, long ldb, int nthreads) {
    return F(orc_assemble_rows)(k, dens, speed, pnts, nNod, nChief, elems, bc, nElem, Binit, nrhs, A, lda,
                                B, ldb, 0, nNod + nChief, nthreads);
}

/* Exterior field: genExtPressure (mesh.cu:1038-1073) with pntElemOffset
 * (mesh.cu:1008-1027): p(x) = inc(x) - sum_e [ sum_j (h_j - (ca/cb) g_j) p_j
 *                                             + (cc/cb) sum_j g_j ],
 * for P points and R surface solutions (surfP: nNod x R, column-major ldp).
 * inc: P x R incident values (point source q*G in the reference). */
int F(orc_field)(REAL k, REAL dens, REAL speed, const REAL *pnts, int nNod, const int *elems,
                 const REAL *bc, int nElem, const REAL *xs, int P, const REAL *surfP, long ldp,
                 int R, const REAL *inc, REAL *out, long ldo, int nthreads) {
    const CPLX *sp = (const CPLX *)surfP, *bcc = (const CPLX *)bc, *ic = (const CPLX *)inc;
    CPLX *o = (CPLX *)out;
    if (R > ORC_MAXRHS) return 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
    for (int q = 0; q < P; q++) {
        const REAL *x = xs + 3 * (long)q;
        CPLX acc[ORC_MAXRHS];
        for (int r = 0; r < R; r++) acc[r] = 0;
        for (int e = 0; e < nElem; e++) {
            const int *nd = elems + 3 * (long)e;
            REAL p[9];
            for (int v = 0; v < 3; v++)
                for (int d = 0; d < 3; d++) p[3 * v + d] = pnts[3 * (long)nd[v] + d];
            CPLX g[3], h[3];
            REAL c;
            F(elem_rule)(3, k, dens, speed, x, p, g, h, &c);
            CPLX beta = bcc[3 * e] / bcc[3 * e + 1];
            CPLX gam = bcc[3 * e + 2] / bcc[3 * e + 1];
            for (int r = 0; r < R; r++) {
                CPLX contrib = 0;
                for (int j = 0; j < 3; j++) contrib += (h[j] - beta * g[j]) * sp[(long)r * ldp + nd[j]];
                contrib += gam * (g[0] + g[1] + g[2]);
                acc[r] += contrib;
            }
        }
        for (int r = 0; r < R; r++) o[(long)r * ldo + q] = (ic ? ic[(long)r * P + q] : 0) - acc[r];
    }
    (void)nNod;
    return 0;
}

/* Least squares min ||A X - B|| by Householder QR, the semantics of qrSolver
 * (numerical.cu:580-644: geqrf, unmqr with Q^H, trsm; solution in B[0:n, :]).
 * A (m x n, lda) is overwritten with R and the reflectors, like geqrf. */
int F(orc_lsq_qr)(REAL *A, int m, int n, long lda, REAL *B, int nrhs, long ldb) {
    if (m < n || n <= 0) return 1;
    CPLX *a = (CPLX *)A, *b = (CPLX *)B;
    CPLX *v = (CPLX *)malloc(sizeof(CPLX) * (size_t)m);
    if (!v) return 1;
    for (int j = 0; j < n; j++) {
        /* Householder vector for column j, rows j..m-1 */
        REAL nrm2 = 0;
        for (int i = j; i < m; i++) nrm2 += CREAL(a[(long)j * lda + i] * CONJ(a[(long)j * lda + i]));
        REAL nrm = SQRT(nrm2);
        if (nrm == 0) { free(v); return 2; }
        CPLX x0 = a[(long)j * lda + j];
        REAL ax0 = CABS(x0);
        CPLX ph = ax0 > 0 ? x0 / ax0 : 1;
        CPLX alpha = -ph * nrm; /* R_jj */
        for (int i = j; i < m; i++) v[i] = a[(long)j * lda + i];
        v[j] -= alpha;
        REAL vn2 = 0;
        for (int i = j; i < m; i++) vn2 += CREAL(v[i] * CONJ(v[i]));
        /* H = I - 2 v v^H / (v^H v) applied to the trailing columns and to B */
        for (int c = j + 1; c < n; c++) {
            CPLX s = 0;
            for (int i = j; i < m; i++) s += CONJ(v[i]) * a[(long)c * lda + i];
            s = s * (REAL)2 / vn2;
            for (int i = j; i < m; i++) a[(long)c * lda + i] -= s * v[i];
        }
        for (int c = 0; c < nrhs; c++) {
            CPLX s = 0;
            for (int i = j; i < m; i++) s += CONJ(v[i]) * b[(long)c * ldb + i];
            s = s * (REAL)2 / vn2;
            for (int i = j; i < m; i++) b[(long)c * ldb + i] -= s * v[i];
        }
        a[(long)j * lda + j] = alpha;
        for (int i = j + 1; i < m; i++) a[(long)j * lda + i] = v[i];
    }
    /* R x = (Q^H b)[0:n]  (upper, non-unit: numerical.cu:622-623) */
    for (int c = 0; c < nrhs; c++) {
        for (int i = n - 1; i >= 0; i--) {
            CPLX s = b[(long)c * ldb + i];
            for (int t = i + 1; t < n; t++) s -= a[(long)t * lda + i] * b[(long)c * ldb + t];
            b[(long)c * ldb + i] = s / a[(long)i * lda + i];
        }
    }
    free(v);
    return 0;
}

/* GMRES for a square system, the semantics of GMRES (GMRES.cu:145-279):
 * unrestarted, initial guess in x, stop when the Givens residual estimate
 * |err|/||b|| < threshold or after maxIter Arnoldi steps; modified Gram-Schmidt
 * Arnoldi (GMRES.cu:90-143); back substitution of the rotated Hessenberg
 * (GMRES.cu:26-42, 236-263).  Returns the number of iterations (>= 0). */
int F(orc_gmres)(const REAL *A, const REAL *bvec, int m, long lda, int maxIter, REAL threshold,
                 REAL *xvec) {
    const CPLX *a = (const CPLX *)A, *b = (const CPLX *)bvec;
    CPLX *x = (CPLX *)xvec;
    if (maxIter > m) maxIter = m;
    if (maxIter <= 0) return 0;
    const int K = maxIter;
    CPLX *Q = (CPLX *)calloc((size_t)m * (K + 1), sizeof(CPLX));
    CPLX *H = (CPLX *)calloc((size_t)(K + 1) * K, sizeof(CPLX));
    CPLX *cs = (CPLX *)calloc((size_t)K, sizeof(CPLX));
    CPLX *sn = (CPLX *)calloc((size_t)K, sizeof(CPLX));
    CPLX *err = (CPLX *)calloc((size_t)K + 1, sizeof(CPLX));
    CPLX *w = (CPLX *)calloc((size_t)m, sizeof(CPLX));
    if (!Q || !H || !cs || !sn || !err || !w) return -1;
    REAL nb = 0;
    for (int i = 0; i < m; i++) nb += CREAL(b[i] * CONJ(b[i]));
    nb = SQRT(nb);
    for (int i = 0; i < m; i++) {
        CPLX s = b[i];
        for (int j = 0; j < m; j++) s -= a[(long)j * lda + i] * x[j];
        w[i] = s;
    }
    REAL nr = 0;
    for (int i = 0; i < m; i++) nr += CREAL(w[i] * CONJ(w[i]));
    nr = SQRT(nr);
    int it = 0;
    if (nb == 0 || nr / nb < threshold) goto done;
    for (int i = 0; i < m; i++) Q[i] = w[i] / nr;
    err[0] = nr;
    for (it = 1; it <= K; it++) {
        const CPLX *q = Q + (long)(it - 1) * m;
        for (int i = 0; i < m; i++) {
            CPLX s = 0;
            for (int j = 0; j < m; j++) s += a[(long)j * lda + i] * q[j];
            w[i] = s;
        }
        CPLX *hc = H + (long)(it - 1) * (K + 1);
        for (int l = 0; l < it; l++) {
            const CPLX *ql = Q + (long)l * m;
            CPLX d = 0;
            for (int i = 0; i < m; i++) d += CONJ(ql[i]) * w[i];
            hc[l] = d;
            for (int i = 0; i < m; i++) w[i] -= d * ql[i];
        }
        REAL hn = 0;
        for (int i = 0; i < m; i++) hn += CREAL(w[i] * CONJ(w[i]));
        hn = SQRT(hn);
        hc[it] = hn;
        if (hn > 0)
            for (int i = 0; i < m; i++) Q[(long)it * m + i] = w[i] / hn;
        /* previous rotations, then a new one (apply_givens, GMRES.cu:58-88) */
        for (int l = 0; l < it - 1; l++) {
            CPLX t0 = cs[l] * hc[l] + sn[l] * hc[l + 1];
            CPLX t1 = -CONJ(sn[l]) * hc[l] + CONJ(cs[l]) * hc[l + 1];
            hc[l] = t0;
            hc[l + 1] = t1;
        }
        REAL mag = SQRT(CREAL(hc[it - 1] * CONJ(hc[it - 1])) + CREAL(hc[it] * CONJ(hc[it])));
        cs[it - 1] = CONJ(hc[it - 1]) / mag; /* givens_coeffs, GMRES.cu:44-56 */
        sn[it - 1] = CONJ(hc[it]) / mag;
        hc[it - 1] = cs[it - 1] * hc[it - 1] + sn[it - 1] * hc[it];
        hc[it] = 0;
        err[it] = -CONJ(sn[it - 1]) * err[it - 1];
        err[it - 1] = cs[it - 1] * err[it - 1];
        if (CABS(err[it]) / nb < threshold || hn == 0) break;
    }
    if (it > K) it = K;
    {
        /* R y = err (Rsolver, GMRES.cu:26-42), then x += Q y (GMRES.cu:262-263) */
        CPLX *y = (CPLX *)calloc((size_t)it, sizeof(CPLX));
        for (int i = it - 1; i >= 0; i--) {
            CPLX s = err[i];
            for (int j = i + 1; j < it; j++) s -= H[(long)j * (K + 1) + i] * y[j];
            y[i] = s / H[(long)i * (K + 1) + i];
        }
        for (int j = 0; j < it; j++)
            for (int i = 0; i < m; i++) x[i] += Q[(long)j * m + i] * y[j];
        free(y);
    }
done:
    free(Q); free(H); free(cs); free(sn); free(err); free(w);
    return it;
}

#undef F
