// Collocation assembly on gfx950.
//
// The reference (atomicFuncs.cu:69-298) launches one thread per (collocation
// row, element) pair and scatters every result with float atomicAdd into A and
// B, so the summation order changes from run to run.  Here a wave owns a block
// of consecutive rows (one row per lane) and a column tile (a compact patch of
// nodes); it walks the tile's element list in a fixed order and accumulates
// straight into its own rows of A in HBM.  No other wave writes those entries,
// so there are no atomics and the result is bitwise repeatable.  Writes are
// coalesced: the 64 lanes of a wave write 64 consecutive rows of one column.
//
// Kernels:
//   prep_elems      one thread per element: p3, e1, e2, unit normal, J, ca/cb,
//                   cc/cb -- the element-uniform data the quadrature reads with
//                   scalar loads;
//   sing_elems      one thread per (element, vertex): the vertex-collapsed rule
//                   of the reference (mesh.cu:253-595) in FP64;
//   asm_tile        the regular 6x6 rule (mesh.cu:597-672) for every (row,
//                   element) pair of a tile, FP32 (or FP64), with the singular
//                   lanes taking the FP64 values;
//   asm_finalize    per row: 1 - sum c on the node diagonal
//                   (atomicFuncs.cu:227-235, 99-102), and the boundary-condition
//                   vector sum_e (cc/cb) sum_j g_j (atomicFuncs.cu:104-115);
//   rhs_fill        B[:, s] = incident_s - bterm.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>
#include <vector>

#include "bem_internal.hpp"
#include "xcd.hpp"

namespace bem {

// ------------------------------------------------------------ device math ---
// sin(2 pi t), cos(2 pi t) for t >= 0 in revolutions, FP32: the quadrant is
// removed exactly (u = t - j/4 with j = rint(4t) is exact in binary floating
// point), then near-minimax polynomials in u on [-1/8, 1/8] (coefficients fitted
// for this library; max abs error 9e-8 after FP32 rounding).
__device__ __forceinline__ void sincos_rev(float t, float &sn, float &cs) {
    const float j = __builtin_rintf(4.0f * t);
    const float u = __builtin_fmaf(j, -0.25f, t);
    const float u2 = u * u;
    float s = __builtin_fmaf(u2, -75.24006652832031f, 81.58812713623047f);
    s = __builtin_fmaf(u2, s, -41.34162902832031f);
    s = __builtin_fmaf(u2, s, 6.283185005187988f);
    s = u * s;
    float c = __builtin_fmaf(u2, 59.24602508544922f, -85.44374084472656f);
    c = __builtin_fmaf(u2, c, 64.93932342529297f);
    c = __builtin_fmaf(u2, c, -19.739208221435547f);
    c = __builtin_fmaf(u2, c, 1.0f);
    const int q = (int)j;
    const float ss = (q & 1) ? c : s;
    const float cc = (q & 1) ? s : c;
    sn = (q & 2) ? -ss : ss;
    cs = ((q + 1) & 2) ? -cc : cc;
}

__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rsq(double x) { return 1.0 / sqrt(x); }

// ------------------------------------------------------------- arguments ---
template <typename R>
struct QTab {
    R xi1[kNQ], xi2[kNQ], w[kNQ];
};

struct SingTab {
    double xi1[3][kNQ], xi2[3][kNQ], w[kNQ];
};

template <typename R>
struct AsmArgs {
    const float4 *__restrict__ pts;
    const ElemRec<R> *__restrict__ recs;
    const R *__restrict__ sing;  // (e*3+v)*16: p[0..5], c[6], bterm[7..8]
    const int *__restrict__ tile_ptr;
    const int *__restrict__ tile_entries;
    const int *__restrict__ tile_slots;  // element-once plan: side column of each ghost vertex
    float2 *side;
    typename std::conditional<sizeof(R) == 4, float2, double2>::type *A;
    R *cpart;
    typename std::conditional<sizeof(R) == 4, float2, double2>::type *bpart;
    long ldA;
    long Mp;
    int M, N, nrbg, nwg;
    int r0;  // first row assembled (rows r0..M-1 land in A rows 0..M-r0-1)
    R k, kf, rw;  // k, k/(2 pi), rho*omega
    int impedance;  // any element with ca != 0
    QTab<R> q;
};

// (xcd_remap, xcd.hpp: consecutive logical ids, here the row blocks of one
// tile, which share the tile's element records, land on the same XCD)

// --------------------------------------------------------- element prep ---
// 6-point Gauss-Legendre nodes and weights on [-1, 1]
struct GL6 {
    static constexpr double eta[6] = {-0.9324695142031520278123016, -0.6612093864662645136613996,
                                      -0.2386191860831969086305017, 0.2386191860831969086305017,
                                      0.6612093864662645136613996,  0.9324695142031520278123016};
    static constexpr double w[6] = {0.1713244923791703450402961, 0.3607615730481386075698335,
                                    0.4679139345726910473898703, 0.4679139345726910473898703,
                                    0.3607615730481386075698335, 0.1713244923791703450402961};
};

template <typename R>
__global__ void prep_elems(const double *__restrict__ pts64, const int *__restrict__ elems,
                           const float2 *__restrict__ bc, int E, ElemRec<R> *__restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const int n0 = elems[3 * e], n1 = elems[3 * e + 1], n2 = elems[3 * e + 2];
    double p1[3], p2[3], p3[3], e1[3], e2[3];
    for (int d = 0; d < 3; d++) {
        p1[d] = pts64[3 * n0 + d];
        p2[d] = pts64[3 * n1 + d];
        p3[d] = pts64[3 * n2 + d];
        e1[d] = p1[d] - p3[d];
        e2[d] = p2[d] - p3[d];
    }
    // crossProd = (p1-p3) x (p2-p3), normal = crossProd / |crossProd| (mesh.cu:612-617)
    double cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    double J = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    ElemRec<R> r;
    for (int d = 0; d < 3; d++) {
        r.p3[d] = (R)p3[d];
        r.e1[d] = (R)e1[d];
        r.e2[d] = (R)e2[d];
        r.n[d] = (R)(cr[d] / J);
    }
    r.J = (R)J;
    // bc = ca/cb and cc/cb (atomicFuncs.cu:87, 105), complex division
    const float2 ca = bc[3 * e], cb = bc[3 * e + 1], cc = bc[3 * e + 2];
    const double den = (double)cb.x * cb.x + (double)cb.y * cb.y;
    r.beta[0] = (R)(((double)ca.x * cb.x + (double)ca.y * cb.y) / den);
    r.beta[1] = (R)(((double)ca.y * cb.x - (double)ca.x * cb.y) / den);
    r.gamma[0] = (R)(((double)cc.x * cb.x + (double)cc.y * cb.y) / den);
    r.gamma[1] = (R)(((double)cc.y * cb.x - (double)cc.x * cb.y) / den);
    r.nodes[0] = n0;
    r.nodes[1] = n1;
    r.nodes[2] = n2;
    double f[3];
    for (int d = 0; d < 3; d++) {
        f[d] = e2[d] - e1[d];
        r.f[d] = (R)f[d];
    }
    for (int t = 0; t < 6; t++) {
        const double th = 0.5 + 0.5 * GL6::eta[t];
        double vv = 0;
        for (int d = 0; d < 3; d++) vv += (e1[d] + th * f[d]) * (e1[d] + th * f[d]);
        r.vv[t] = (R)vv;
    }
    for (int i = 0; i < (int)(sizeof(r.pad) / sizeof(int)); i++) r.pad[i] = 0;
    out[e] = r;
}

// ------------------------------------------------------ singular, FP64 ---
// One thread per (element, vertex v): the reference's vertex rule for x = p_v,
// g_l_sgl{v+1}/h_l_sgl{v+1}/c_l_sgl{v+1} (mesh.cu:253-595), in FP64.  Stores
// the A contributions h_j - (ca/cb) g_j (j = 0..2), c, and (cc/cb) sum_j g_j.
template <typename Rout>
__global__ void sing_elems(const double *__restrict__ pts64, const int *__restrict__ elems,
                           const float2 *__restrict__ bc, int E, double k, double rw, SingTab tab,
                           Rout *__restrict__ out) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= 3 * E) return;
    const int e = id / 3, v = id % 3;
    double p[3][3];
    for (int j = 0; j < 3; j++)
        for (int d = 0; d < 3; d++) p[j][d] = pts64[3 * elems[3 * e + j] + d];
    double e1[3], e2[3];
    for (int d = 0; d < 3; d++) {
        e1[d] = p[0][d] - p[2][d];
        e2[d] = p[1][d] - p[2][d];
    }
    double cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const double J = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    const double nx = cr[0] / J, ny = cr[1] / J, nz = cr[2] / J;
    const double *x = p[v];
    const double inv4pi = 1.0 / (4.0 * kPi);
    double gr[3] = {0, 0, 0}, gi[3] = {0, 0, 0}, hr[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, c = 0;
    for (int q = 0; q < kNQ; q++) {
        const double xi1 = tab.xi1[v][q], xi2 = tab.xi2[v][q];
        const double N[3] = {xi1, xi2, 1.0 - xi1 - xi2};
        const double temp = tab.w[q] * J;  // 0.25 w_n w_m rho J
        double rv[3];
        for (int d = 0; d < 3; d++) rv[d] = N[0] * p[0][d] + N[1] * p[1][d] + N[2] * p[2][d] - x[d];
        const double r = sqrt(rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]);
        const double prpn2 = (rv[0] * nx + rv[1] * ny + rv[2] * nz) / r;
        double sn, cs;
        sincos(k * r, &sn, &cs);
        const double Gr = cs * inv4pi / r, Gi = -sn * inv4pi / r;  // exp(-ikr)/(4 pi r)
        // dG/dn = G (-1/r - ik) dr/dn
        const double Hr = (Gr * (-1.0 / r) + Gi * k) * prpn2;
        const double Hi = (Gi * (-1.0 / r) - Gr * k) * prpn2;
        const double dG0 = -inv4pi / (r * r) * prpn2;
        for (int j = 0; j < 3; j++) {
            gr[j] += temp * N[j] * Gr;
            gi[j] += temp * N[j] * Gi;
            hr[j] += temp * N[j] * Hr;
            hi[j] += temp * N[j] * Hi;
        }
        c += temp * dG0;
    }
    // g_j *= i rho omega (mesh.cu:668-670)
    double Gjr[3], Gji[3];
    for (int j = 0; j < 3; j++) {
        Gjr[j] = -rw * gi[j];
        Gji[j] = rw * gr[j];
    }
    const float2 ca = bc[3 * e], cb = bc[3 * e + 1], cc = bc[3 * e + 2];
    const double den = (double)cb.x * cb.x + (double)cb.y * cb.y;
    const double br = ((double)ca.x * cb.x + (double)ca.y * cb.y) / den;
    const double bi = ((double)ca.y * cb.x - (double)ca.x * cb.y) / den;
    const double yr = ((double)cc.x * cb.x + (double)cc.y * cb.y) / den;
    const double yi = ((double)cc.y * cb.x - (double)cc.x * cb.y) / den;
    Rout *o = out + (size_t)id * 16;
    double sgr = 0, sgi = 0;
    for (int j = 0; j < 3; j++) {
        o[2 * j] = (Rout)(hr[j] - (br * Gjr[j] - bi * Gji[j]));
        o[2 * j + 1] = (Rout)(hi[j] - (br * Gji[j] + bi * Gjr[j]));
        sgr += Gjr[j];
        sgi += Gji[j];
    }
    o[6] = (Rout)c;
    o[7] = (Rout)(yr * sgr - yi * sgi);
    o[8] = (Rout)(yr * sgi + yi * sgr);
    for (int i = 9; i < 16; i++) o[i] = 0;
}

// ------------------------------------------------- regular quadrature ---
template <typename R>
__device__ __forceinline__ void sincos_phase(R r, R k, R kf, R &sn, R &cs);
template <>
__device__ __forceinline__ void sincos_phase<float>(float r, float, float kf, float &sn, float &cs) {
    sincos_rev(r * kf, sn, cs);
}
template <>
__device__ __forceinline__ void sincos_phase<double>(double r, double k, double, double &sn, double &cs) {
    sincos(k * r, &sn, &cs);
}

template <typename R>
__device__ __forceinline__ R shfl_xor(R v, int m) {
    return __shfl_xor(v, m, 64);
}

// One wave = RPW = 64/QG consecutive rows; QG lanes share a row and split the
// 36 quadrature points (QG = 1: a lane does all of them).  A workgroup is 4
// waves on 4 consecutive row blocks of the same tile.
template <typename R, int QG, bool IMP>
__global__ __launch_bounds__(256) void asm_tile(AsmArgs<R> a) {
    using C2 = typename std::conditional<sizeof(R) == 4, float2, double2>::type;
    constexpr int RPW = 64 / QG;
    constexpr int NQL = kNQ / QG;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int L = xcd_remap(blockIdx.x, a.nwg);
    const int tile = L / a.nrbg;
    const int rbg = L - tile * a.nrbg;
    const int row = a.r0 + (rbg * 4 + wave) * RPW + lane / QG;
    const int qg = (QG == 1) ? 0 : (lane % QG);
    const bool active = row < a.M;
    const float4 xp = a.pts[active ? row : 0];
    const R x0 = xp.x, x1 = xp.y, x2 = xp.z;

    R txi1[NQL], txi2[NQL], tw[NQL];
#pragma unroll
    for (int i = 0; i < NQL; i++) {
        const int q = qg + QG * i;
        txi1[i] = a.q.xi1[q];
        txi2[i] = a.q.xi2[q];
        tw[i] = a.q.w[q];
    }
    const R k = a.k, kf = a.kf;
    const R inv4pi = (R)(1.0 / (4.0 * kPi));
    R c_acc = 0, br_acc = 0, bi_acc = 0;

    const int it0 = a.tile_ptr[tile], it1 = a.tile_ptr[tile + 1];
    for (int it = it0; it < it1; ++it) {
        const int ent = a.tile_entries[it];
        const int e = ent >> 8;
        const ElemRec<R> &rec = a.recs[e];
        const R ax = rec.p3[0] - x0, ay = rec.p3[1] - x1, az = rec.p3[2] - x2;
        const R e10 = rec.e1[0], e11 = rec.e1[1], e12 = rec.e1[2];
        const R e20 = rec.e2[0], e21 = rec.e2[1], e22 = rec.e2[2];
        // (y - x).n is constant over a flat element: n.(p3 - x)
        const R dn = rec.n[0] * ax + rec.n[1] * ay + rec.n[2] * az;
        R h1r = 0, h1i = 0, h2r = 0, h2i = 0, htr = 0, hti = 0, gr = 0, gi = 0, cc = 0;
        R g1r = 0, g1i = 0, g2r = 0, g2i = 0;
#pragma unroll
        for (int i = 0; i < NQL; i++) {
            const R dx = fma(txi1[i], e10, fma(txi2[i], e20, ax));
            const R dy = fma(txi1[i], e11, fma(txi2[i], e21, ay));
            const R dz = fma(txi1[i], e12, fma(txi2[i], e22, az));
            const R r2 = fma(dx, dx, fma(dy, dy, dz * dz));
            const R rinv = rsq(r2);
            const R r = r2 * rinv;
            R sn, cs;
            sincos_phase<R>(r, k, kf, sn, cs);
            // 4 pi G = (cos kr - i sin kr) / r
            const R Gr = cs * rinv, Gi = -sn * rinv;
            const R w2 = tw[i] * rinv;
            // 4 pi r G (1/r + ik), weighted; dG/dn = -(dn/4pi) rinv G~ (rinv + ik)
            const R Hr = (fma(Gr, rinv, -k * Gi)) * w2;
            const R Hi = (fma(Gi, rinv, k * Gr)) * w2;
            htr += Hr;
            hti += Hi;
            h1r = fma(txi1[i], Hr, h1r);
            h1i = fma(txi1[i], Hi, h1i);
            h2r = fma(txi2[i], Hr, h2r);
            h2i = fma(txi2[i], Hi, h2i);
            const R wGr = tw[i] * Gr, wGi = tw[i] * Gi;
            gr += wGr;
            gi += wGi;
            if (IMP) {
                g1r = fma(txi1[i], wGr, g1r);
                g1i = fma(txi1[i], wGi, g1i);
                g2r = fma(txi2[i], wGr, g2r);
                g2i = fma(txi2[i], wGi, g2i);
            }
            cc = fma(w2 * rinv, rinv, cc);
        }
        if (QG > 1) {
#pragma unroll
            for (int m = 1; m < QG; m <<= 1) {
                h1r += shfl_xor(h1r, m); h1i += shfl_xor(h1i, m);
                h2r += shfl_xor(h2r, m); h2i += shfl_xor(h2i, m);
                htr += shfl_xor(htr, m); hti += shfl_xor(hti, m);
                gr += shfl_xor(gr, m); gi += shfl_xor(gi, m);
                cc += shfl_xor(cc, m);
                if (IMP) {
                    g1r += shfl_xor(g1r, m); g1i += shfl_xor(g1i, m);
                    g2r += shfl_xor(g2r, m); g2i += shfl_xor(g2i, m);
                }
            }
        }
        const R sJ = rec.J * inv4pi;
        const R sh = -dn * sJ;
        // h_j (mesh.cu:661-663), N3 = 1 - N1 - N2
        R p0r = sh * h1r, p0i = sh * h1i;
        R p1r = sh * h2r, p1i = sh * h2i;
        R p2r = sh * (htr - h1r - h2r), p2i = sh * (hti - h1i - h2i);
        // sum_j g_j = i rho omega (J/4pi) sum w G~  (mesh.cu:657-659, 668-670)
        const R sg = a.rw * sJ;
        const R gsr = -sg * gi, gsi = sg * gr;
        if (IMP) {
            // p_j = h_j - beta g_j
            const R br = rec.beta[0], bi = rec.beta[1];
            const R G0r = -sg * g1i, G0i = sg * g1r;
            const R G1r = -sg * g2i, G1i = sg * g2r;
            const R G2r = gsr - G0r - G1r, G2i = gsi - G0i - G1i;
            p0r -= br * G0r - bi * G0i; p0i -= br * G0i + bi * G0r;
            p1r -= br * G1r - bi * G1i; p1i -= br * G1i + bi * G1r;
            p2r -= br * G2r - bi * G2i; p2i -= br * G2i + bi * G2r;
        }
        R cval = -dn * sJ * cc;  // c_l_nsgl (mesh.cu:645, 665)
        R btr = rec.gamma[0] * gsr - rec.gamma[1] * gsi;
        R bti = rec.gamma[0] * gsi + rec.gamma[1] * gsr;
        // singular rows: x is a vertex of e (atomicFuncs.cu:75-76): take the FP64 rule
        const int sv = (row == rec.nodes[0]) ? 0 : (row == rec.nodes[1]) ? 1 : (row == rec.nodes[2]) ? 2 : -1;
        if (sv >= 0) {
            const R *s = a.sing + ((size_t)e * 3 + sv) * 16;
            p0r = s[0]; p0i = s[1]; p1r = s[2]; p1i = s[3]; p2r = s[4]; p2i = s[5];
            cval = s[6];
            btr = s[7];
            bti = s[8];
        }
        // scatter into this wave's own rows of A (first touch writes)
        const bool writer = active && (QG == 1 || qg == 0);
        if (writer) {
            const R pr[3] = {p0r, p1r, p2r}, pi[3] = {p0i, p1i, p2i};
#pragma unroll
            for (int j = 0; j < 3; j++) {
                if ((ent >> j) & 1) {
                    C2 *dst = a.A + (size_t)rec.nodes[j] * a.ldA + (row - a.r0);
                    C2 v;
                    if ((ent >> (3 + j)) & 1) {
                        v.x = pr[j];
                        v.y = pi[j];
                    } else {
                        v = *dst;
                        v.x += pr[j];
                        v.y += pi[j];
                    }
                    *dst = v;
                }
            }
        }
        if ((ent >> 6) & 1) {
            c_acc += cval;
            br_acc += btr;
            bi_acc += bti;
        }
    }
    if (active && (QG == 1 || qg == 0)) {
        a.cpart[(size_t)tile * a.Mp + (row - a.r0)] = c_acc;
        C2 b;
        b.x = br_acc;
        b.y = bi_acc;
        a.bpart[(size_t)tile * a.Mp + (row - a.r0)] = b;
    }
}

// ------------------------------------------- FP32 fast path (one lane/row) ---
// The 6x6 collapsed rule is separable: with theta_n, rho_m the mapped
// Gauss-Legendre nodes, xi = rho_m ((1 - theta_n), theta_n) and
// W = 0.25 w_n w_m rho_m, so the point offset is y - x = a + rho_m v_n with
// v_n = e1 + theta_n (e2 - e1) (one vector per theta row), and the
// shape-function sums factor per theta row:
//   h_1 = sum_n (1 - theta_n) A_n,  h_2 = sum_n theta_n A_n,  h_tot = sum_n B_n,
//   A_n = sum_m W rho_m H~,  B_n = sum_m W H~.
// All rule constants are compile-time literals.  sin / cos of the phase come
// from v_sin_f32 / v_cos_f32, which take revolutions u = k r / (2 pi) and are
// specified for |u| <= 256 (measured on gfx950: max abs error 1.2e-7 against
// sin(2 pi u) of the FP32 u over that whole range).  RAW passes u = r k/(2 pi)
// as one rounded product; otherwise u is first reduced to |u| <= 1/2 with the
// round-to-nearest magic constant 1.5 * 2^23 (for geometries that reach past
// 256 revolutions).  The rounding of the product adds an error of the order of
// the one r already carries: at 8 kHz on icosphere s5 the relative Frobenius
// error of A against FP64 goes from 1.060e-7 to 1.063e-7, and the kernel
// runs 5 % faster (two fewer dependent instructions per point).
// The squared distance is expanded per theta row,
//   r^2 = |a|^2 + rho_m (2 a.v_n + rho_m |v_n|^2),
// two FMAs per point instead of six.  Its rounding error grows as
// (|a| / r)^2 where the point is much closer to the element than to the
// vertex p3; the rows where r -> 0 (x a vertex of e) take the FP64 singular
// rule, and on icosphere s5 at 2 and 8 kHz the entries are as close to FP64
// as with the direct form (relative Frobenius error 9.3e-8 and 1.06e-7).
__device__ __forceinline__ constexpr float gl_theta(int n) { return (float)(0.5 + 0.5 * GL6::eta[n]); }
__device__ __forceinline__ constexpr float gl_rho(int m) { return (float)(0.5 + 0.5 * GL6::eta[m]); }
// theta_n - theta_{n+1} for n < 5, theta_5 for n = 5
__device__ __forceinline__ constexpr float gl_dtheta(int n) {
    return n < 5 ? (float)(0.5 * (GL6::eta[n] - GL6::eta[n < 5 ? n + 1 : n])) : (float)(0.5 + 0.5 * GL6::eta[5]);
}
__device__ __forceinline__ constexpr float gl_W(int n, int m) {
    return (float)(0.25 * GL6::w[n] * GL6::w[m] * (0.5 + 0.5 * GL6::eta[m]));
}
__device__ __forceinline__ constexpr float gl_Wrho(int n, int m) {
    return (float)(0.25 * GL6::w[n] * GL6::w[m] * (0.5 + 0.5 * GL6::eta[m]) * (0.5 + 0.5 * GL6::eta[m]));
}

template <bool IMP, bool RAW, bool GH>
__global__ __launch_bounds__(256) void asm_tile_fast(AsmArgs<float> a) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int L = xcd_remap(blockIdx.x, a.nwg);
    const int tile = L / a.nrbg;
    const int rbg = L - tile * a.nrbg;
    const int row = a.r0 + (rbg * 4 + wave) * 64 + lane;
    const bool active = row < a.M;
    const float4 xp = a.pts[active ? row : 0];
    const float k = a.k;
    const float krev = a.kf;  // k / (2 pi)
    const float inv4pi = (float)(1.0 / (4.0 * kPi));
    const float MAGIC = 12582912.0f;  // 1.5 * 2^23
    float c_acc = 0, br_acc = 0, bi_acc = 0;
    const int it0 = a.tile_ptr[tile], it1 = a.tile_ptr[tile + 1];
    for (int it = it0; it < it1; ++it) {
        const int ent = a.tile_entries[it];
        const int e = ent >> 8;
        const ElemRec<float> &rec = a.recs[e];
        const float ax = rec.p3[0] - xp.x, ay = rec.p3[1] - xp.y, az = rec.p3[2] - xp.z;
        const float dn = rec.n[0] * ax + rec.n[1] * ay + rec.n[2] * az;
        const float aa = __builtin_fmaf(ax, ax, __builtin_fmaf(ay, ay, az * az));
        // 2 a.v_n = 2 a.e1 + 2 theta_n a.f
        const float ae1 = __builtin_fmaf(ax, rec.e1[0], __builtin_fmaf(ay, rec.e1[1], az * rec.e1[2]));
        const float af = __builtin_fmaf(ax, rec.f[0], __builtin_fmaf(ay, rec.f[1], az * rec.f[2]));
        const float ae1x2 = ae1 + ae1;
        // S = sum W rho H~ and ht = sum W H~ over all 36 points; the theta
        // moment h2 = sum_n theta_n A_n of the per-row sums A_n comes from the
        // running totals S_n of S after each row,
        //   h2 = sum_{n<5} (theta_n - theta_{n+1}) S_n + theta_5 S_5,
        // and h1 = S - h2 (likewise g1, g2 from SC = sum W rho G~)
        float Sr = 0, Si = 0, h2r = 0, h2i = 0, htr = 0, hti = 0, gr = 0, gi = 0, cc = 0;
        float SCr = 0, SCi = 0, g2r = 0, g2i = 0;
#pragma unroll
        for (int n = 0; n < 6; n++) {
            const float th = gl_theta(n);
            const float tav = __builtin_fmaf(2.0f * th, af, ae1x2);
            const float vv = rec.vv[n];
#pragma unroll
            for (int m = 0; m < 6; m++) {
                const float rho = gl_rho(m), W = gl_W(n, m), Wr = gl_Wrho(n, m);
                const float r2 = __builtin_fmaf(rho, __builtin_fmaf(rho, vv, tav), aa);
                const float rinv = __builtin_amdgcn_rsqf(r2);
                const float r = r2 * rinv;
                float u;
                if (RAW) {
                    u = r * krev;
                } else {
                    const float t2 = __builtin_fmaf(r, krev, MAGIC);
                    const float jf = t2 - MAGIC;
                    u = __builtin_fmaf(r, krev, -jf);
                }
                const float sn = __builtin_amdgcn_sinf(u), cs = __builtin_amdgcn_cosf(u);
                const float rinv2 = rinv * rinv;
                const float q3 = rinv2 * rinv;  // 1/r^3
                const float q2k = k * rinv2;    // k/r^2
                // 4 pi r dG/dn / (-dn) = (cos kr - i sin kr)(1/r^3 + i k/r^2)
                const float Hr = __builtin_fmaf(cs, q3, sn * q2k);
                const float Hi = __builtin_fmaf(cs, q2k, -sn * q3);
                Sr = __builtin_fmaf(Wr, Hr, Sr);
                Si = __builtin_fmaf(Wr, Hi, Si);
                htr = __builtin_fmaf(W, Hr, htr);
                hti = __builtin_fmaf(W, Hi, hti);
                // 4 pi G = (cos kr - i sin kr) / r
                if (IMP) {
                    const float Gr = cs * rinv, Gi = -sn * rinv;
                    gr = __builtin_fmaf(W, Gr, gr);
                    gi = __builtin_fmaf(W, Gi, gi);
                    SCr = __builtin_fmaf(Wr, Gr, SCr);
                    SCi = __builtin_fmaf(Wr, Gi, SCi);
                } else {
                    const float Wg = W * rinv;
                    gr = __builtin_fmaf(Wg, cs, gr);
                    gi = __builtin_fmaf(-Wg, sn, gi);
                }
                cc = __builtin_fmaf(W, q3, cc);
            }
            // theta moment of the rho-weighted sums from their running totals
            const float dth = gl_dtheta(n);
            h2r = __builtin_fmaf(dth, Sr, h2r);
            h2i = __builtin_fmaf(dth, Si, h2i);
            if (IMP) {
                g2r = __builtin_fmaf(dth, SCr, g2r);
                g2i = __builtin_fmaf(dth, SCi, g2i);
            }
        }
        const float h1r = Sr - h2r, h1i = Si - h2i;
        const float g1r = SCr - g2r, g1i = SCi - g2i;
        const float sJ = rec.J * inv4pi;
        const float sh = -dn * sJ;
        float p0r = sh * h1r, p0i = sh * h1i;
        float p1r = sh * h2r, p1i = sh * h2i;
        float p2r = sh * (htr - Sr), p2i = sh * (hti - Si);
        const float sg = a.rw * sJ;
        const float gsr = -sg * gi, gsi = sg * gr;
        if (IMP) {
            const float br = rec.beta[0], bi = rec.beta[1];
            const float G0r = -sg * g1i, G0i = sg * g1r;
            const float G1r = -sg * g2i, G1i = sg * g2r;
            const float G2r = gsr - G0r - G1r, G2i = gsi - G0i - G1i;
            p0r -= br * G0r - bi * G0i; p0i -= br * G0i + bi * G0r;
            p1r -= br * G1r - bi * G1i; p1i -= br * G1i + bi * G1r;
            p2r -= br * G2r - bi * G2i; p2i -= br * G2i + bi * G2r;
        }
        float cval = -dn * sJ * cc;
        float btr = rec.gamma[0] * gsr - rec.gamma[1] * gsi;
        float bti = rec.gamma[0] * gsi + rec.gamma[1] * gsr;
        const int sv = (row == rec.nodes[0]) ? 0 : (row == rec.nodes[1]) ? 1 : (row == rec.nodes[2]) ? 2 : -1;
        if (sv >= 0) {
            const float *sg2 = a.sing + ((size_t)e * 3 + sv) * 16;
            p0r = sg2[0]; p0i = sg2[1]; p1r = sg2[2]; p1i = sg2[3]; p2r = sg2[4]; p2i = sg2[5];
            cval = sg2[6];
            btr = sg2[7];
            bti = sg2[8];
        }
        if (active) {
            const float pr[3] = {p0r, p1r, p2r}, pi[3] = {p0i, p1i, p2i};
#pragma unroll
            for (int j = 0; j < 3; j++) {
                float2 *dst = nullptr;
                if ((ent >> j) & 1) dst = a.A + (size_t)rec.nodes[j] * a.ldA + (row - a.r0);
                else if (GH) dst = a.side + (size_t)a.tile_slots[3 * it + j] * a.Mp + (row - a.r0);
                if (dst) {
                    float2 v;
                    if ((ent >> (3 + j)) & 1) {
                        v.x = pr[j];
                        v.y = pi[j];
                    } else {
                        v = *dst;
                        v.x += pr[j];
                        v.y += pi[j];
                    }
                    *dst = v;
                }
            }
        }
        if ((ent >> 6) & 1) {
            c_acc += cval;
            br_acc += btr;
            bi_acc += bti;
        }
    }
    if (active) {
        a.cpart[(size_t)tile * a.Mp + (row - a.r0)] = c_acc;
        a.bpart[(size_t)tile * a.Mp + (row - a.r0)] = make_float2(br_acc, bi_acc);
    }
}

// Columns of nodes no element references: only the diagonal survives.
template <typename C2>
__global__ void zero_columns(C2 *A, long ldA, int M, const int *__restrict__ cols, int ncols) {
    const int c = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncols || i >= M) return;
    C2 z;
    z.x = 0;
    z.y = 0;
    A[(size_t)cols[c] * ldA + i] = z;
}

// Element-once plan: column of ghost node g += its side columns, in the order
// they were made (the column is set when the node's own tile wrote nothing).
__global__ void ghost_reduce(float2 *A, long ldA, int nrows, const float2 *__restrict__ side, long Mp,
                             const int *__restrict__ gptr, const int *__restrict__ gnode,
                             const int *__restrict__ gown, const int *__restrict__ gslot) {
    // two adjacent rows per thread (16-byte accesses: ldA and Mp are even)
    const int g = blockIdx.y;
    const int i = 2 * (blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= nrows) return;
    float2 *col = A + (size_t)gnode[g] * ldA;
    const int s0 = gptr[g], s1 = gptr[g + 1];
    const bool own = gown[g] != 0;
    if (i + 1 < nrows) {
        float4 acc = own ? *(const float4 *)(col + i) : make_float4(0.f, 0.f, 0.f, 0.f);
        for (int t = s0; t < s1; t++) {
            const float4 v = *(const float4 *)(side + (size_t)gslot[t] * Mp + i);
            acc.x += v.x;
            acc.y += v.y;
            acc.z += v.z;
            acc.w += v.w;
        }
        *(float4 *)(col + i) = acc;
    } else {
        float2 acc = own ? col[i] : make_float2(0.f, 0.f);
        for (int t = s0; t < s1; t++) {
            const float2 v = side[(size_t)gslot[t] * Mp + i];
            acc.x += v.x;
            acc.y += v.y;
        }
        col[i] = acc;
    }
}

// Per row: fixed-order sum of the tile partials in FP64.
// Rows r0..M-1 (local index i = row - r0 in A, the partials and bterm).
template <typename R, typename C2>
__global__ void asm_finalize(C2 *A, long ldA, int M, int N, int P, long Mp, int r0, const R *__restrict__ cpart,
                             const C2 *__restrict__ bpart, double2 *__restrict__ bterm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = r0 + i;
    if (row >= M) return;
    double cs = 0, br = 0, bi = 0;
    // 8 tiles' partials in flight per step, added in tile order
    int t = 0;
    for (; t + 8 <= P; t += 8) {
        R c[8];
        C2 b[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            c[u] = cpart[(size_t)(t + u) * Mp + i];
            b[u] = bpart[(size_t)(t + u) * Mp + i];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            cs += (double)c[u];
            br += (double)b[u].x;
            bi += (double)b[u].y;
        }
    }
    for (; t < P; t++) {
        cs += (double)cpart[(size_t)t * Mp + i];
        const C2 b = bpart[(size_t)t * Mp + i];
        br += (double)b.x;
        bi += (double)b.y;
    }
    if (row < N) {
        C2 *d = A + (size_t)row * ldA + i;
        C2 v = *d;
        v.x = (R)((double)v.x + (1.0 - cs));
        *d = v;
    }
    bterm[i] = make_double2(br, bi);
}

// B[:, s] = inc_s(x) - bterm: point source q G(x, src) (atomicFuncs.cu:238-242)
// or plane wave amp exp(-i k d.x); computed in FP64.
template <typename C2>
__global__ void rhs_fill(const double *__restrict__ pts64, int M, double k, int kind, double amp,
                         const float *__restrict__ params, int nrhs, const double2 *__restrict__ bterm,
                         C2 *B, long ldB) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (i >= M || s >= nrhs) return;
    const double x = pts64[3 * i], y = pts64[3 * i + 1], z = pts64[3 * i + 2];
    const double px = params[3 * s], py = params[3 * s + 1], pz = params[3 * s + 2];
    double ir, ii;
    if (kind == BEM_RHS_PLANE) {
        double sn, cs;
        sincos(k * (px * x + py * y + pz * z), &sn, &cs);
        ir = amp * cs;
        ii = -amp * sn;
    } else {
        const double dx = x - px, dy = y - py, dz = z - pz;
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        double sn, cs;
        sincos(k * r, &sn, &cs);
        const double g = amp / (4.0 * kPi * r);
        ir = g * cs;
        ii = -g * sn;
        if (amp == 0.0) ir = ii = 0.0;
    }
    const double2 bt = bterm[i];
    C2 v;
    v.x = ir - bt.x;
    v.y = ii - bt.y;
    B[(size_t)s * ldB + i] = v;
}

// ------------------------------------------------------------- field ---
// Exterior pressure (genExtPressure / pntElemOffset, mesh.cu:1008-1073) at P
// points for R nodal solutions.  Lane = field point; the wave walks a chunk of
// elements (element data uniform, read with scalar loads) with the regular
// 6x6 rule and accumulates sum_j (h_j - beta g_j) X[node_j] + gamma sum_j g_j;
// chunk partials are summed in a fixed order by field_finalize.
constexpr int kFieldChunk = 64;
constexpr int kFieldR = 16;

struct FieldArgs {
    const float4 *__restrict__ xs;
    const ElemRec<float> *__restrict__ recs;
    const float2 *__restrict__ X;  // N x R, ldx
    long ldx;
    float2 *part;                  // [chunk][r][P]
    int P, E, R, r0;
    float k, kf, rw;
    QTab<float> q;
};

__global__ __launch_bounds__(64) void field_part(FieldArgs a) {
    const int p = blockIdx.x * 64 + threadIdx.x;
    const int chunk = blockIdx.y;
    const bool active = p < a.P;
    const float4 xp = a.xs[active ? p : 0];
    const int e0 = chunk * kFieldChunk, e1 = min(a.E, e0 + kFieldChunk);
    const int nr = min(kFieldR, a.R - a.r0);
    float accr[kFieldR], acci[kFieldR];
#pragma unroll
    for (int r = 0; r < kFieldR; r++) {
        accr[r] = 0;
        acci[r] = 0;
    }
    const float inv4pi = (float)(1.0 / (4.0 * kPi));
    for (int e = e0; e < e1; e++) {
        const ElemRec<float> &rec = a.recs[e];
        const float ax = rec.p3[0] - xp.x, ay = rec.p3[1] - xp.y, az = rec.p3[2] - xp.z;
        const float dn = rec.n[0] * ax + rec.n[1] * ay + rec.n[2] * az;
        float h1r = 0, h1i = 0, h2r = 0, h2i = 0, htr = 0, hti = 0, gr = 0, gi = 0;
        float g1r = 0, g1i = 0, g2r = 0, g2i = 0;
#pragma unroll 6
        for (int i = 0; i < kNQ; i++) {
            const float xi1 = a.q.xi1[i], xi2 = a.q.xi2[i], w = a.q.w[i];
            const float dx = fmaf(xi1, rec.e1[0], fmaf(xi2, rec.e2[0], ax));
            const float dy = fmaf(xi1, rec.e1[1], fmaf(xi2, rec.e2[1], ay));
            const float dz = fmaf(xi1, rec.e1[2], fmaf(xi2, rec.e2[2], az));
            const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
            const float rinv = rsq(r2);
            const float r = r2 * rinv;
            float sn, cs;
            sincos_rev(r * a.kf, sn, cs);
            const float Gr = cs * rinv, Gi = -sn * rinv;
            const float w2 = w * rinv;
            const float Hr = fmaf(Gr, rinv, -a.k * Gi) * w2;
            const float Hi = fmaf(Gi, rinv, a.k * Gr) * w2;
            htr += Hr; hti += Hi;
            h1r = fmaf(xi1, Hr, h1r); h1i = fmaf(xi1, Hi, h1i);
            h2r = fmaf(xi2, Hr, h2r); h2i = fmaf(xi2, Hi, h2i);
            const float wGr = w * Gr, wGi = w * Gi;
            gr += wGr; gi += wGi;
            g1r = fmaf(xi1, wGr, g1r); g1i = fmaf(xi1, wGi, g1i);
            g2r = fmaf(xi2, wGr, g2r); g2i = fmaf(xi2, wGi, g2i);
        }
        const float sJ = rec.J * inv4pi, sh = -dn * sJ, sg = a.rw * sJ;
        float pr[3] = {sh * h1r, sh * h2r, sh * (htr - h1r - h2r)};
        float pi[3] = {sh * h1i, sh * h2i, sh * (hti - h1i - h2i)};
        const float gsr = -sg * gi, gsi = sg * gr;
        const float G0r = -sg * g1i, G0i = sg * g1r, G1r = -sg * g2i, G1i = sg * g2r;
        const float Gjr[3] = {G0r, G1r, gsr - G0r - G1r}, Gji[3] = {G0i, G1i, gsi - G0i - G1i};
        const float br = rec.beta[0], bi = rec.beta[1];
#pragma unroll
        for (int j = 0; j < 3; j++) {
            pr[j] -= br * Gjr[j] - bi * Gji[j];
            pi[j] -= br * Gji[j] + bi * Gjr[j];
        }
        const float btr = rec.gamma[0] * gsr - rec.gamma[1] * gsi;
        const float bti = rec.gamma[0] * gsi + rec.gamma[1] * gsr;
#pragma unroll
        for (int r = 0; r < kFieldR; r++) {
            if (r < nr) {
                const float2 *xc = a.X + (size_t)(a.r0 + r) * a.ldx;
                float sr = btr, si = bti;
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const float2 xv = xc[rec.nodes[j]];
                    sr += pr[j] * xv.x - pi[j] * xv.y;
                    si += pr[j] * xv.y + pi[j] * xv.x;
                }
                accr[r] += sr;
                acci[r] += si;
            }
        }
    }
    if (active) {
#pragma unroll
        for (int r = 0; r < kFieldR; r++)
            if (r < nr) a.part[((size_t)chunk * a.R + a.r0 + r) * a.P + p] = make_float2(accr[r], acci[r]);
    }
}

// out[p, r] = inc_r(x_p) - sum_chunks part
__global__ void field_finalize(const float4 *__restrict__ xs, int P, int R, int nchunks, const float2 *__restrict__ part,
                               double k, int kind, double amp, const float *__restrict__ params, float2 *out,
                               long ldo) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (p >= P || r >= R) return;
    double sr = 0, si = 0;
    for (int c = 0; c < nchunks; c++) {
        const float2 v = part[((size_t)c * R + r) * P + p];
        sr += v.x;
        si += v.y;
    }
    const double x = xs[p].x, y = xs[p].y, z = xs[p].z;
    double ir = 0, ii = 0;
    if (params) {
        const double px = params[3 * r], py = params[3 * r + 1], pz = params[3 * r + 2];
        double sn, cs;
        if (kind == BEM_RHS_PLANE) {
            sincos(k * (px * x + py * y + pz * z), &sn, &cs);
            ir = amp * cs;
            ii = -amp * sn;
        } else if (amp != 0.0) {
            const double dx = x - px, dy = y - py, dz = z - pz;
            const double rr = sqrt(dx * dx + dy * dy + dz * dz);
            sincos(k * rr, &sn, &cs);
            ir = amp * cs / (4.0 * kPi * rr);
            ii = -amp * sn / (4.0 * kPi * rr);
        }
    }
    out[(size_t)r * ldo + p] = make_float2((float)(ir - sr), (float)(ii - si));
}

int field_fp32(DeviceProblem &dp, float k, const float4 *xs, int P, const float2 *X, long ldx, int R,
               const RhsSpec &rhs, float2 *out, long ldo, hipStream_t s) {
    if (P <= 0 || R <= 0) return EXIT_SUCCESS;
    const Config &cfg = config();
    QuadTables qt;
    quad_tables(qt);
    const int nch = (dp.E + kFieldChunk - 1) / kFieldChunk;
    float2 *part = nullptr;
    float *params = nullptr;
    HOST_CALL(salloc(&part, (size_t)nch * R * P, s));
    if (!rhs.params.empty()) {
        HOST_CALL(salloc(&params, rhs.params.size(), s));
        HIP_CALL(hipMemcpyAsync(params, rhs.params.data(), rhs.params.size() * sizeof(float), hipMemcpyHostToDevice, s));
    }
    FieldArgs a;
    a.xs = xs;
    a.recs = dp.rec32;
    a.X = X;
    a.ldx = ldx;
    a.part = part;
    a.P = P;
    a.E = dp.E;
    a.R = R;
    a.k = k;
    a.kf = (float)((double)k / (2.0 * kPi));
    a.rw = (float)((double)cfg.density * (double)(k * cfg.speed));
    for (int q = 0; q < kNQ; q++) {
        a.q.xi1[q] = qt.xi1[q];
        a.q.xi2[q] = qt.xi2[q];
        a.q.w[q] = qt.w[q];
    }
    TimedRegion tr;
    timing_begin(tr, "field_part", s);
    for (int r0 = 0; r0 < R; r0 += kFieldR) {
        a.r0 = r0;
        dim3 grid((P + 63) / 64, nch);
        field_part<<<grid, 64, 0, s>>>(a);
        HIP_CALL(hipGetLastError());
    }
    timing_end(tr, s);
    dim3 g2((P + 255) / 256, R);
    field_finalize<<<g2, 256, 0, s>>>(xs, P, R, nch, part, (double)k, rhs.kind, (double)rhs.amp,
                                      rhs.nrhs >= R ? params : nullptr, out, ldo);
    hipError_t e = hipGetLastError();
    (void)hipStreamSynchronize(s);
    sfree(part, s);
    sfree(params, s);
    HIP_CALL(e);
    return EXIT_SUCCESS;
}

// ------------------------------------------- near-singular, FP64 ---
// The reference integrates every element that does not hold the collocation
// point with the 6x6 collapsed rule (g_h_c_l_nsgl, mesh.cu:597-672; the only
// exclusion is the vertex test, atomicFuncs.cu:75-76).  For a point closer to
// the element than about its size the integrand peaks inside the element and
// the rule's error grows to 1e-3 at a quarter of the element size.  With
// bem_set_near_field on, every (row, element) pair closer than a factor times
// the element's longest edge gets an FP64 correction
//     delta = (integral by adaptive subdivision) - (the 6x6 rule in FP64),
// added to the assembled entries after the row-owned assembly, in a fixed
// order per row.  The subdivision splits a triangle into four at its edge
// midpoints, depth first; a sub-triangle is integrated with the same 6x6
// collapsed rule once its distance from the point (centroid distance minus
// circumradius) is at least kNearEta times its longest edge, or at the depth
// limit.  The path through the quadtree is kept as two bits per level, so the
// walk needs no stack: each sub-triangle is rebuilt from the root.
constexpr double kNearEta = 1.5;
constexpr int kNearOut = 13;  // g (3 complex), h (3 complex), c

struct NearArgs {
    const double *__restrict__ pts64;  // 3 M: nodes then extra rows
    const int *__restrict__ elems;     // 3 E
    const int *__restrict__ row;       // per pair
    const int *__restrict__ elem;      // per pair
    int p0, p1;                        // pairs [p0, p1)
    int levels;
    double k, rw;                      // k, rho omega
    double qu[kNQ], qv[kNQ], qw[kNQ];  // 6x6 collapsed rule on the unit triangle: (u, v) and weight
    double *__restrict__ out;
};

struct NearAcc {
    double gr[3], gi[3], hr[3], hi[3], c;
};

// the collapsed rule over the sub-triangle (a, b, c) of the element's
// (xi1, xi2) plane, scaled by its area ratio `det`; y = p3 + xi1 e1 + xi2 e2
__device__ void near_leaf(const NearArgs &a, const double *x, const double *p3, const double *e1, const double *e2,
                          const double *n, double J, const double *va, const double *vb, const double *vc,
                          double det, NearAcc &acc) {
    const double inv4pi = 1.0 / (4.0 * kPi);
    for (int q = 0; q < kNQ; q++) {
        const double u = a.qu[q], v = a.qv[q];
        const double xi1 = vc[0] + u * (va[0] - vc[0]) + v * (vb[0] - vc[0]);
        const double xi2 = vc[1] + u * (va[1] - vc[1]) + v * (vb[1] - vc[1]);
        const double N[3] = {xi1, xi2, 1.0 - xi1 - xi2};
        double rv[3];
        for (int d = 0; d < 3; d++) rv[d] = p3[d] + xi1 * e1[d] + xi2 * e2[d] - x[d];
        const double r2 = rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2];
        const double r = sqrt(r2);
        const double prpn2 = (rv[0] * n[0] + rv[1] * n[1] + rv[2] * n[2]) / r;  // dr/dn_y
        double sn, cs;
        sincos(a.k * r, &sn, &cs);
        const double w = a.qw[q] * det * J;
        // G = exp(-ikr) / (4 pi r); dG/dn = G (-1/r - ik) dr/dn; dG0/dn = -dr/dn / (4 pi r^2)
        const double Gr = cs * inv4pi / r, Gi = -sn * inv4pi / r;
        const double Hr = (Gr * (-1.0 / r) + Gi * a.k) * prpn2;
        const double Hi = (Gi * (-1.0 / r) - Gr * a.k) * prpn2;
        for (int j = 0; j < 3; j++) {
            acc.gr[j] += w * N[j] * Gr;
            acc.gi[j] += w * N[j] * Gi;
            acc.hr[j] += w * N[j] * Hr;
            acc.hi[j] += w * N[j] * Hi;
        }
        acc.c += -w * prpn2 * inv4pi / r2;
    }
}

// One thread per pair: delta = adaptive - rule, with g scaled by i rho omega
// as the reference scales it (mesh.cu:668-670).
__global__ __launch_bounds__(64) void near_elems(NearArgs a) {
    const int id = a.p0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= a.p1) return;
    const int e = a.elem[id], row = a.row[id];
    double x[3], p[3][3];
    for (int d = 0; d < 3; d++) x[d] = a.pts64[3 * (size_t)row + d];
    for (int j = 0; j < 3; j++)
        for (int d = 0; d < 3; d++) p[j][d] = a.pts64[3 * (size_t)a.elems[3 * e + j] + d];
    double e1[3], e2[3];
    for (int d = 0; d < 3; d++) {
        e1[d] = p[0][d] - p[2][d];
        e2[d] = p[1][d] - p[2][d];
    }
    const double cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const double J = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    const double n[3] = {cr[0] / J, cr[1] / J, cr[2] / J};
    const double A0[2] = {1.0, 0.0}, B0[2] = {0.0, 1.0}, C0[2] = {0.0, 0.0};
    NearAcc rule = {}, ad = {};
    near_leaf(a, x, p[2], e1, e2, n, J, A0, B0, C0, 1.0, rule);
    // depth-first walk of the quadtree; digits of the path, 2 bits per level
    unsigned long long path = 0;
    int level = 0;
    for (;;) {
        double va[2] = {1.0, 0.0}, vb[2] = {0.0, 1.0}, vc[2] = {0.0, 0.0};
        for (int l = 1; l <= level; l++) {
            const int dgt = (int)((path >> (2 * (l - 1))) & 3);
            const double ab[2] = {0.5 * (va[0] + vb[0]), 0.5 * (va[1] + vb[1])};
            const double bc[2] = {0.5 * (vb[0] + vc[0]), 0.5 * (vb[1] + vc[1])};
            const double ca[2] = {0.5 * (vc[0] + va[0]), 0.5 * (vc[1] + va[1])};
            double na[2], nb[2], nc[2];
            if (dgt == 0) { na[0] = va[0]; na[1] = va[1]; nb[0] = ab[0]; nb[1] = ab[1]; nc[0] = ca[0]; nc[1] = ca[1]; }
            else if (dgt == 1) { na[0] = ab[0]; na[1] = ab[1]; nb[0] = vb[0]; nb[1] = vb[1]; nc[0] = bc[0]; nc[1] = bc[1]; }
            else if (dgt == 2) { na[0] = ca[0]; na[1] = ca[1]; nb[0] = bc[0]; nb[1] = bc[1]; nc[0] = vc[0]; nc[1] = vc[1]; }
            else { na[0] = bc[0]; na[1] = bc[1]; nb[0] = ca[0]; nb[1] = ca[1]; nc[0] = ab[0]; nc[1] = ab[1]; }
            va[0] = na[0]; va[1] = na[1]; vb[0] = nb[0]; vb[1] = nb[1]; vc[0] = nc[0]; vc[1] = nc[1];
        }
        // physical vertices, centroid distance, circumradius bound, longest edge
        double Y[3][3], cen[3] = {0, 0, 0};
        const double *vv[3] = {va, vb, vc};
        for (int t = 0; t < 3; t++)
            for (int d = 0; d < 3; d++) {
                Y[t][d] = p[2][d] + vv[t][0] * e1[d] + vv[t][1] * e2[d];
                cen[d] += Y[t][d] / 3.0;
            }
        double rad = 0, diam = 0;
        for (int t = 0; t < 3; t++) {
            double q = 0, s = 0;
            for (int d = 0; d < 3; d++) {
                q += (Y[t][d] - cen[d]) * (Y[t][d] - cen[d]);
                s += (Y[t][d] - Y[(t + 1) % 3][d]) * (Y[t][d] - Y[(t + 1) % 3][d]);
            }
            rad = fmax(rad, sqrt(q));
            diam = fmax(diam, sqrt(s));
        }
        const double dc = sqrt((x[0] - cen[0]) * (x[0] - cen[0]) + (x[1] - cen[1]) * (x[1] - cen[1]) +
                               (x[2] - cen[2]) * (x[2] - cen[2]));
        if (level >= a.levels || dc - rad >= kNearEta * diam) {
            near_leaf(a, x, p[2], e1, e2, n, J, va, vb, vc, ldexp(1.0, -2 * level), ad);
            // next sibling, climbing past last children
            while (level > 0 && ((path >> (2 * (level - 1))) & 3) == 3) {
                path &= ~(3ull << (2 * (level - 1)));
                level--;
            }
            if (level == 0) break;
            path += 1ull << (2 * (level - 1));
        } else {
            level++;  // first child: digit 0
        }
    }
    double *o = a.out + (size_t)id * kNearOut;
    // g_j = i rho omega int N_j G
    for (int j = 0; j < 3; j++) {
        const double dgr = ad.gr[j] - rule.gr[j], dgi = ad.gi[j] - rule.gi[j];
        o[2 * j] = -a.rw * dgi;
        o[2 * j + 1] = a.rw * dgr;
        o[6 + 2 * j] = ad.hr[j] - rule.hr[j];
        o[6 + 2 * j + 1] = ad.hi[j] - rule.hi[j];
    }
    o[12] = ad.c - rule.c;
}

// One thread per row r0 + i: the row's pair corrections in pair order,
// A[i, node_j] += dh_j - (ca/cb) dg_j, A[i, i] -= dc on node rows,
// bterm[i] += (cc/cb) sum_j dg_j (atomicFuncs.cu:86-115 for the corrections)
template <typename R, typename C2>
__global__ void near_apply(C2 *A, long ldA, int M, int N, int r0, const int *__restrict__ ptr,
                           const int *__restrict__ elem, const int *__restrict__ elems, const float2 *__restrict__ bc,
                           const double *__restrict__ d, double2 *__restrict__ bterm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int row = r0 + i;
    if (row >= M) return;
    double btr = 0, bti = 0, dc = 0;
    for (int q = ptr[row]; q < ptr[row + 1]; q++) {
        const int e = elem[q];
        const double *o = d + (size_t)q * kNearOut;
        const float2 ca = bc[3 * e], cb = bc[3 * e + 1], cc = bc[3 * e + 2];
        const double den = (double)cb.x * cb.x + (double)cb.y * cb.y;
        const double br = ((double)ca.x * cb.x + (double)ca.y * cb.y) / den;
        const double bi = ((double)ca.y * cb.x - (double)ca.x * cb.y) / den;
        const double gr = ((double)cc.x * cb.x + (double)cc.y * cb.y) / den;
        const double gi = ((double)cc.y * cb.x - (double)cc.x * cb.y) / den;
        double sgr = 0, sgi = 0;
        for (int j = 0; j < 3; j++) {
            const double dgr = o[2 * j], dgi = o[2 * j + 1];
            const double pr = o[6 + 2 * j] - (br * dgr - bi * dgi);
            const double pi = o[6 + 2 * j + 1] - (br * dgi + bi * dgr);
            C2 *a = A + (size_t)elems[3 * e + j] * ldA + i;
            C2 v = *a;
            v.x = (R)((double)v.x + pr);
            v.y = (R)((double)v.y + pi);
            *a = v;
            sgr += dgr;
            sgi += dgi;
        }
        btr += gr * sgr - gi * sgi;
        bti += gr * sgi + gi * sgr;
        dc += o[12];
    }
    if (row < N && dc != 0.0) {
        C2 *a = A + (size_t)row * ldA + i;
        C2 v = *a;
        v.x = (R)((double)v.x - dc);
        *a = v;
    }
    double2 b = bterm[i];
    b.x += btr;
    b.y += bti;
    bterm[i] = b;
}

namespace {
// distance from x to the triangle (a, b, c): closest point by the regions of
// the plane around the triangle (the vertex, edge and face regions of the
// barycentric test)
double point_triangle_distance(const double *x, const double *a, const double *b, const double *c) {
    auto sub = [](const double *u, const double *v, double *o) {
        for (int d = 0; d < 3; d++) o[d] = u[d] - v[d];
    };
    auto dot = [](const double *u, const double *v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
    double ab[3], ac[3], ap[3], bp[3], cp[3], q[3];
    sub(b, a, ab);
    sub(c, a, ac);
    sub(x, a, ap);
    auto dist_to = [&](const double *y) {
        return std::sqrt((x[0] - y[0]) * (x[0] - y[0]) + (x[1] - y[1]) * (x[1] - y[1]) + (x[2] - y[2]) * (x[2] - y[2]));
    };
    auto on = [&](const double *o, const double *dir, double t) {
        for (int d = 0; d < 3; d++) q[d] = o[d] + t * dir[d];
        return dist_to(q);
    };
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) return dist_to(a);
    sub(x, b, bp);
    const double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) return dist_to(b);
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) return on(a, ab, d1 / (d1 - d3));
    sub(x, c, cp);
    const double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) return dist_to(c);
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) return on(a, ac, d2 / (d2 - d6));
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        double bc[3];
        sub(c, b, bc);
        return on(b, bc, (d4 - d3) / ((d4 - d3) + (d5 - d6)));
    }
    const double den = 1.0 / (va + vb + vc);
    const double v = vb * den, w = vc * den;
    for (int d = 0; d < 3; d++) q[d] = a[d] + ab[d] * v + ac[d] * w;
    return dist_to(q);
}
}  // namespace

// (row, element) pairs with dist(x_row, element) < factor * (longest edge),
// the element not holding the row's node; sorted by row, then element.  A
// uniform grid over the element centroids, cells as wide as the largest
// reach, bounds the search to the 27 cells around each row point.
void near_pairs(const std::vector<double> &pts, int M, int N, const std::vector<int> &elems, double factor,
                std::vector<int> &ptr, std::vector<int> &prow, std::vector<int> &pelem) {
    const int E = (int)(elems.size() / 3);
    std::vector<double> cen(3 * (size_t)E), reach(E), hmax(E);
    double cell = 0, lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int e = 0; e < E; e++) {
        const double *v[3] = {&pts[3 * (size_t)elems[3 * e]], &pts[3 * (size_t)elems[3 * e + 1]],
                              &pts[3 * (size_t)elems[3 * e + 2]]};
        double h = 0, r = 0;
        for (int d = 0; d < 3; d++) cen[3 * e + d] = (v[0][d] + v[1][d] + v[2][d]) / 3.0;
        for (int t = 0; t < 3; t++) {
            double s = 0, q = 0;
            for (int d = 0; d < 3; d++) {
                s += (v[t][d] - v[(t + 1) % 3][d]) * (v[t][d] - v[(t + 1) % 3][d]);
                q += (v[t][d] - cen[3 * e + d]) * (v[t][d] - cen[3 * e + d]);
            }
            h = std::max(h, std::sqrt(s));
            r = std::max(r, std::sqrt(q));
        }
        hmax[e] = h;
        reach[e] = r + factor * h;
        cell = std::max(cell, reach[e]);
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], cen[3 * e + d]);
            hi[d] = std::max(hi[d], cen[3 * e + d]);
        }
    }
    cell = std::max(cell, 1e-30);
    int nc[3];
    for (int d = 0; d < 3; d++) nc[d] = std::max(1, std::min(1024, (int)((hi[d] - lo[d]) / cell) + 1));
    auto cell_of = [&](const double *y, int *c) {
        for (int d = 0; d < 3; d++) c[d] = std::max(0, std::min(nc[d] - 1, (int)((y[d] - lo[d]) / cell)));
    };
    std::vector<std::vector<int>> grid((size_t)nc[0] * nc[1] * nc[2]);
    for (int e = 0; e < E; e++) {
        int c[3];
        cell_of(&cen[3 * e], c);
        grid[((size_t)c[0] * nc[1] + c[1]) * nc[2] + c[2]].push_back(e);
    }
    ptr.assign((size_t)M + 1, 0);
    prow.clear();
    pelem.clear();
    std::vector<int> found;
    for (int i = 0; i < M; i++) {
        const double *x = &pts[3 * (size_t)i];
        int c[3];
        cell_of(x, c);
        found.clear();
        for (int a = std::max(0, c[0] - 1); a <= std::min(nc[0] - 1, c[0] + 1); a++)
            for (int b = std::max(0, c[1] - 1); b <= std::min(nc[1] - 1, c[1] + 1); b++)
                for (int z = std::max(0, c[2] - 1); z <= std::min(nc[2] - 1, c[2] + 1); z++)
                    for (int e : grid[((size_t)a * nc[1] + b) * nc[2] + z]) {
                        const int *n = &elems[3 * e];
                        if (i < N && (n[0] == i || n[1] == i || n[2] == i)) continue;  // singular rule
                        const double dc = std::sqrt((x[0] - cen[3 * e]) * (x[0] - cen[3 * e]) +
                                                    (x[1] - cen[3 * e + 1]) * (x[1] - cen[3 * e + 1]) +
                                                    (x[2] - cen[3 * e + 2]) * (x[2] - cen[3 * e + 2]));
                        if (dc >= reach[e]) continue;
                        const double d = point_triangle_distance(x, &pts[3 * (size_t)n[0]], &pts[3 * (size_t)n[1]],
                                                                 &pts[3 * (size_t)n[2]]);
                        if (d < factor * hmax[e]) found.push_back(e);
                    }
        std::sort(found.begin(), found.end());
        for (int e : found) {
            prow.push_back(i);
            pelem.push_back(e);
        }
        ptr[i + 1] = (int)prow.size();
    }
}

// the pairs of the device problem for the current bem_set_near_field setting
static int near_prepare(DeviceProblem &dp, hipStream_t s) {
    const Config &cfg = config();
    if (cfg.near_factor == dp.near_factor && cfg.near_levels == dp.near_levels) return EXIT_SUCCESS;
    sfree(dp.near_ptr, s);
    sfree(dp.near_row, s);
    sfree(dp.near_elem, s);
    sfree(dp.near_out, s);
    dp.near_npairs = 0;
    dp.hnear_ptr.clear();
    dp.near_factor = cfg.near_factor;
    dp.near_levels = cfg.near_levels;
    if (cfg.near_factor <= 0) return EXIT_SUCCESS;
    std::vector<int> prow, pelem;
    near_pairs(dp.hpts, dp.M, dp.N, dp.helems, cfg.near_factor, dp.hnear_ptr, prow, pelem);
    dp.near_npairs = (int)prow.size();
    HOST_CALL(salloc(&dp.near_ptr, dp.hnear_ptr.size(), s));
    HOST_CALL(salloc(&dp.near_row, std::max<size_t>(1, prow.size()), s));
    HOST_CALL(salloc(&dp.near_elem, std::max<size_t>(1, pelem.size()), s));
    HOST_CALL(salloc(&dp.near_out, std::max<size_t>(1, prow.size()) * kNearOut, s));
    HIP_CALL(hipMemcpyAsync(dp.near_ptr, dp.hnear_ptr.data(), dp.hnear_ptr.size() * sizeof(int), hipMemcpyHostToDevice,
                            s));
    if (!prow.empty()) {
        HIP_CALL(hipMemcpyAsync(dp.near_row, prow.data(), prow.size() * sizeof(int), hipMemcpyHostToDevice, s));
        HIP_CALL(hipMemcpyAsync(dp.near_elem, pelem.data(), pelem.size() * sizeof(int), hipMemcpyHostToDevice, s));
    }
    HIP_CALL(hipStreamSynchronize(s));  // the host vectors go out of scope
    return EXIT_SUCCESS;
}

// corrections of rows r0..M-1 of A (ld ldA, row r0 first) and their bterm
template <typename R, typename C2>
static int near_correct(DeviceProblem &dp, double k, double rw, C2 *A, long ldA, int r0, double2 *bterm,
                        hipStream_t s) {
    HOST_CALL(near_prepare(dp, s));
    if (dp.near_npairs == 0) return EXIT_SUCCESS;
    const int p0 = dp.hnear_ptr[r0], p1 = dp.hnear_ptr[dp.M];
    if (p1 > p0) {
        NearArgs a;
        a.pts64 = dp.pts64;
        a.elems = dp.elems;
        a.row = dp.near_row;
        a.elem = dp.near_elem;
        a.p0 = p0;
        a.p1 = p1;
        a.levels = dp.near_levels;
        a.k = k;
        a.rw = rw;
        const Config &cfg = config();
        for (int n = 0; n < kQOrder; n++) {
            const double theta = 0.5 + 0.5 * cfg.qp[n];
            for (int m = 0; m < kQOrder; m++) {
                const double rho = 0.5 + 0.5 * cfg.qp[m];
                const int q = n * kQOrder + m;
                a.qu[q] = rho * (1 - theta);  // xi1 = rho (1 - theta), xi2 = rho theta (mesh.cu:623-636)
                a.qv[q] = rho * theta;
                a.qw[q] = 0.25 * cfg.qw[n] * cfg.qw[m] * rho;
            }
        }
        a.out = dp.near_out;
        TimedRegion tr;
        timing_begin(tr, "near_elems", s);
        near_elems<<<(p1 - p0 + 63) / 64, 64, 0, s>>>(a);
        HIP_CALL(hipGetLastError());
        timing_end(tr, s);
    }
    const int nrows = dp.M - r0;
    near_apply<R, C2><<<(nrows + 255) / 256, 256, 0, s>>>(A, ldA, dp.M, dp.N, r0, dp.near_ptr, dp.near_elem, dp.elems,
                                                        dp.bc, dp.near_out, bterm);
    HIP_CALL(hipGetLastError());
    return EXIT_SUCCESS;
}

// ------------------------------------------- field without allocations ---
// The exterior field for callers that cannot allocate device memory (code
// under HIP graph capture, serving loops): the partial sums, the points in
// the kernels' float4 layout and the source point live in one file-scope
// device array of fixed size.  Each device holds its own copy of the array;
// host code reaches it only through the address hipGetSymbolAddress returns
// on that device, kept per device below.  Every call on a device shares that
// device's array, so calls that overlap on one device race.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// the scratch array's device address on the current device (looked up once per device)
int scratch_address(void **out) {
    int dev = 0;
    HIP_CALL(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) BEM_FAIL("device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIP_CALL(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) BEM_FAIL("no device address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return EXIT_SUCCESS;
}

size_t round256(size_t b) { return (b + 255) / 256 * 256; }
}  // namespace

// 12-byte points (cartCoord) to the kernels' float4 layout, and the
// incident point source's position next to them
__global__ void k_pack_field_points(const float *__restrict__ xs, int P, float4 *__restrict__ xs4, float sx, float sy,
                                    float sz, float *__restrict__ params) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < P) xs4[p] = make_float4(xs[3 * p], xs[3 * p + 1], xs[3 * p + 2], 0.f);
    if (p == 0) {
        params[0] = sx;
        params[1] = sy;
        params[2] = sz;
    }
}

size_t field_noalloc_bytes(int E, int P) {
    const size_t nch = (size_t)(E + kFieldChunk - 1) / kFieldChunk;
    return round256(sizeof(float4) * (size_t)P) + 256 + sizeof(float2) * nch * (size_t)P;
}

size_t field_noalloc_capacity() { return kNoAllocScratchBytes; }

int field_fp32_noalloc(const DeviceProblem &dp, float k, const float *xs, int P, const float2 *X, float strength,
                       const float *src, float2 *out, hipStream_t s) {
    if (P <= 0) return EXIT_SUCCESS;
    const size_t need = field_noalloc_bytes(dp.E, P);
    if (need > kNoAllocScratchBytes)
        BEM_FAIL("%d field points on %d elements need %zu bytes of scratch, more than its %zu", P, dp.E, need,
                 kNoAllocScratchBytes);
    void *base = nullptr;
    HOST_CALL(scratch_address(&base));
    unsigned char *b = (unsigned char *)base;
    float4 *xs4 = (float4 *)b;
    float *params = (float *)(b + round256(sizeof(float4) * (size_t)P));
    float2 *part = (float2 *)(b + round256(sizeof(float4) * (size_t)P) + 256);
    k_pack_field_points<<<(P + 255) / 256, 256, 0, s>>>(xs, P, xs4, src[0], src[1], src[2], params);
    HIP_CALL(hipGetLastError());
    const Config &cfg = config();
    QuadTables qt;
    quad_tables(qt);
    FieldArgs a;
    a.xs = xs4;
    a.recs = dp.rec32;
    a.X = X;
    a.ldx = dp.N;
    a.part = part;
    a.P = P;
    a.E = dp.E;
    a.R = 1;
    a.r0 = 0;
    a.k = k;
    a.kf = (float)((double)k / (2.0 * kPi));
    a.rw = (float)((double)cfg.density * (double)(k * cfg.speed));
    for (int q = 0; q < kNQ; q++) {
        a.q.xi1[q] = qt.xi1[q];
        a.q.xi2[q] = qt.xi2[q];
        a.q.w[q] = qt.w[q];
    }
    const int nch = (dp.E + kFieldChunk - 1) / kFieldChunk;
    field_part<<<dim3((P + 63) / 64, nch), 64, 0, s>>>(a);
    HIP_CALL(hipGetLastError());
    field_finalize<<<dim3((P + 255) / 256, 1), 256, 0, s>>>(xs4, P, 1, nch, part, (double)k, BEM_RHS_POINT,
                                                            (double)strength, params, out, P);
    HIP_CALL(hipGetLastError());
    return EXIT_SUCCESS;
}

// --------------------------------------------------------------- host ---
// Launch shape of the assembly: qg lanes per row (1 or 4) and P column tiles,
// P x ceil(M / (64 / qg)) waves in all.  More tiles cost more halo elements
// (an element straddling tiles is integrated once per tile) but give more
// waves to balance over the 1024 SIMDs; the default aims at kDefaultWaves.
// bem_autotune_assembly times candidate shapes for a mesh on the device and
// records the fastest for its size class (M and E in half-octave buckets),
// which device problems of that class built afterwards use.
constexpr int kDefaultWaves = 16384;

namespace {
struct TileShape {
    int qg = 0, waves = 0;
};
std::mutex g_tune_mu;
std::map<std::pair<int, int>, TileShape> g_tune;
thread_local const TileShape *t_force = nullptr;  // candidate under test

int half_octave(long v) { return (int)std::lround(2.0 * std::log2((double)std::max(1L, v))); }

bool tuned_shape(int M, int E, TileShape &out) {
    if (t_force) {
        out = *t_force;
        return true;
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find({half_octave(M), half_octave(E)});
    if (it == g_tune.end()) return false;
    out = it->second;
    return true;
}
}  // namespace

// halo share (elements integrated per element) from which the element-once
// plan is taken
constexpr double kGhostMinDup = 1.05;

// side columns the element-once plan may use: about the nodes on the tile
// boundaries (4 sqrt(N P) + P; 4190 for N = 10242 in 102 tiles, which need
// 3110), and never more than N / 2
static long ghost_slot_bound(int N, int P) {
    return std::min((long)N / 2, (long)(4.0 * std::sqrt((double)N * P)) + P);
}

// wave_factor: the element-once plan has no halo to pay for more tiles, and
// runs at twice the waves (icosphere s5: 5.56 ms with 102 tiles, 5.35 with 204)
static int choose_tiles(int M, int N, int E, int rows_per_wave, int wave_factor = 1) {
    const char *env = getenv("BEM_TILES");
    if (env && atoi(env) > 0) return std::min(atoi(env), N);
    TileShape t;
    const int target_waves = (tuned_shape(M, E, t) ? t.waves : kDefaultWaves) * wave_factor;
    const int nrb = (M + rows_per_wave - 1) / rows_per_wave;
    int P = (target_waves + nrb - 1) / nrb;
    P = std::max(1, std::min(P, std::max(1, N / 8)));
    return P;
}

static int choose_qg(int M, int E) {
    const char *env = getenv("BEM_QG");
    if (env) {
        int q = atoi(env);
        if (q == 1 || q == 4) return q;
    }
    TileShape t;
    if (tuned_shape(M, E, t)) return t.qg;
    return ((long)M * E < 4000000L) ? 4 : 1;
}

// Times the FP32 (or FP64) assembly of mesh m at wavenumber k for each
// candidate shape (2 warm + 3 timed launches each, HIP events on stream s)
// and records the fastest for the mesh's size class.
int autotune_assembly(const Mesh &m, double k, int precision, int *best_qg, int *best_waves, double *best_ms,
                      double *default_ms, hipStream_t s) {
    const int M = m.N() + m.C(), E = m.E();
    std::vector<TileShape> cand;
    for (int qg : {1, 4})
        for (int w : {4096, 8192, 16384, 32768}) cand.push_back(TileShape{qg, w});
    TileShape def{((long)M * E < 4000000L) ? 4 : 1, kDefaultWaves};
    double bms = 1e300, dms = -1;
    TileShape best = def;
    hipEvent_t e0, e1;
    HIP_CALL(hipEventCreate(&e0));
    HIP_CALL(hipEventCreate(&e1));
    size_t n = ((size_t)M + 63) / 64 * 64 * (size_t)m.N();
    void *A = nullptr;
    HIP_CALL(hipMallocAsync(&A, n * (precision == BEM_PREC_FP64 ? sizeof(double2) : sizeof(float2)), s));
    int rc = EXIT_SUCCESS;
    for (const TileShape &c : cand) {
        DeviceProblem dp;
        t_force = &c;
        rc = device_problem_build(m, dp, s);
        t_force = nullptr;
        if (rc != EXIT_SUCCESS) break;
        float ms = 0;
        for (int rep = 0; rep < 5 && rc == EXIT_SUCCESS; rep++) {
            if (rep == 2) (void)hipEventRecord(e0, s);
            rc = precision == BEM_PREC_FP64 ? assemble_fp64(dp, k, (double2 *)A, s)
                                            : assemble_fp32(dp, (float)k, (float2 *)A, s);
        }
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        (void)hipEventElapsedTime(&ms, e0, e1);
        device_problem_free(dp);
        if (rc != EXIT_SUCCESS) break;
        const double avg = ms / 3.0;
        if (c.qg == def.qg && c.waves == def.waves) dms = avg;
        if (avg < bms) {
            bms = avg;
            best = c;
        }
    }
    (void)hipFreeAsync(A, s);
    (void)hipStreamSynchronize(s);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HOST_CALL(rc);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune[{half_octave(M), half_octave(E)}] = best;
    }
    if (best_qg) *best_qg = best.qg;
    if (best_waves) *best_waves = best.waves;
    if (best_ms) *best_ms = bms;
    if (default_ms) *default_ms = dms;
    return EXIT_SUCCESS;
}

// The tuned shapes as text lines: "M-bucket E-bucket qg waves" for the
// assembly, then the products' "p ..." lines (gemm.hip), and back.
int tune_save(const char *file) {
    FILE *f = fopen(file, "w");
    if (!f) BEM_FAIL("cannot write %s", file);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        for (const auto &kv : g_tune)
            fprintf(f, "%d %d %d %d\n", kv.first.first, kv.first.second, kv.second.qg, kv.second.waves);
    }
    product_tune_save(f);
    fclose(f);
    return EXIT_SUCCESS;
}

int tune_load(const char *file) {
    FILE *f = fopen(file, "r");
    if (!f) BEM_FAIL("cannot read %s", file);
    char line[256];
    while (fgets(line, sizeof line, f)) {
        int a, b, q, w;
        if (product_tune_parse(line)) continue;
        if (sscanf(line, "%d %d %d %d", &a, &b, &q, &w) == 4 && (q == 1 || q == 4) && w > 0) {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tune[{a, b}] = TileShape{q, w};
        }
    }
    fclose(f);
    return EXIT_SUCCESS;
}

void tune_clear() {
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune.clear();
    }
    product_tune_clear();
}

// Upper bound of the HBM bytes device_problem_build allocates for a mesh of
// M rows (nodes + CHIEF), N nodes and E elements.
size_t device_problem_bytes(int M, int N, int E) {
    const size_t Mp = ((size_t)M + 63) / 64 * 64;
    const int qg = choose_qg(M, E);
    const size_t P = (size_t)choose_tiles(M, N, E, 64 / qg);
    size_t b = 0;
    b += sizeof(float4) * M + sizeof(double) * 3 * M + sizeof(double2) * M;        // points, bterm
    b += (sizeof(int) + sizeof(float2)) * 3 * (size_t)E;                            // elements, bc
    b += (sizeof(ElemRec<float>) + sizeof(ElemRec<double>)) * (size_t)E;            // element records
    b += (sizeof(float) + sizeof(double)) * 48 * (size_t)E;                          // singular terms
    b += sizeof(int) * (P + 1 + 3 * (size_t)E + N);                                  // tile CSR, isolated
    b += (sizeof(float) + sizeof(float2) + sizeof(double) + sizeof(double2)) * P * Mp;  // tile partials
    if (qg == 1) {
        // element-once plan: its side columns and index arrays
        const int gP = choose_tiles(M, N, E, 64, 2);
        const size_t gs = (size_t)ghost_slot_bound(N, gP);
        b += sizeof(float2) * gs * Mp;
        if ((size_t)gP > P) b += (sizeof(float) + sizeof(float2)) * ((size_t)gP - P) * Mp;  // its tile partials
        b += sizeof(int) * (gP + 1 + 4 * (size_t)E + 2 * (size_t)N + 2 * gs);
    }
    return b;
}

int device_problem_build(const Mesh &m, DeviceProblem &dp, hipStream_t s) {
    HOST_CALL(validate_mesh(m));
    HOST_CALL(ensure_quad());
    device_problem_free(dp);
    dp.st = s;
    dp.N = m.N();
    dp.C = m.C();
    dp.M = dp.N + dp.C;
    dp.E = m.E();
    dp.ldA = ((long)dp.M + 63) / 64 * 64;
    dp.Mp = dp.ldA;
    std::vector<float4> p4(dp.M);
    std::vector<double> p64(3 * (size_t)dp.M);
    for (int i = 0; i < dp.M; i++) {
        const float *src = i < dp.N ? &m.pnts[3 * i] : &m.chief[3 * (i - dp.N)];
        p4[i] = make_float4(src[0], src[1], src[2], 0.f);
        for (int d = 0; d < 3; d++) p64[3 * i + d] = src[d];
    }
    {
        double lo[3], hi[3];
        for (int d = 0; d < 3; d++) lo[d] = hi[d] = dp.M ? p64[d] : 0.0;
        for (int i = 0; i < dp.M; i++)
            for (int d = 0; d < 3; d++) {
                lo[d] = std::min(lo[d], p64[3 * i + d]);
                hi[d] = std::max(hi[d], p64[3 * i + d]);
            }
        dp.diam = std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                            (hi[2] - lo[2]) * (hi[2] - lo[2]));
    }
    std::vector<int> el(3 * (size_t)dp.E);
    std::vector<float2> bc(3 * (size_t)dp.E);
    for (int e = 0; e < dp.E; e++)
        for (int j = 0; j < 3; j++) {
            el[3 * e + j] = m.elems[e].nodes[j];
            bc[3 * e + j] = make_float2(m.elems[e].bc[j].x, m.elems[e].bc[j].y);
        }
    TilePlan plan;
    const int qg = choose_qg(dp.M, dp.E);
    build_tile_plan(m, choose_tiles(dp.M, dp.N, dp.E, 64 / qg), plan);
    dp.P = plan.P;
    HOST_CALL(salloc(&dp.pts, dp.M, s));
    HOST_CALL(salloc(&dp.pts64, 3 * (size_t)dp.M, s));
    HOST_CALL(salloc(&dp.elems, 3 * (size_t)dp.E, s));
    HOST_CALL(salloc(&dp.bc, 3 * (size_t)dp.E, s));
    HOST_CALL(salloc(&dp.rec32, dp.E, s));
    HOST_CALL(salloc(&dp.rec64, dp.E, s));
    HOST_CALL(salloc(&dp.sing32, (size_t)dp.E * 48, s));
    HOST_CALL(salloc(&dp.sing64, (size_t)dp.E * 48, s));
    HOST_CALL(salloc(&dp.tile_ptr, plan.ptr.size(), s));
    HOST_CALL(salloc(&dp.tile_entries, plan.entries.size(), s));
    dp.n_isolated = (int)plan.isolated.size();
    HOST_CALL(salloc(&dp.isolated, std::max<size_t>(1, plan.isolated.size()), s));
    HOST_CALL(salloc(&dp.cpart, (size_t)dp.P * dp.Mp, s));
    HOST_CALL(salloc(&dp.bpart, (size_t)dp.P * dp.Mp, s));
    HOST_CALL(salloc(&dp.cpart64, (size_t)dp.P * dp.Mp, s));
    HOST_CALL(salloc(&dp.bpart64, (size_t)dp.P * dp.Mp, s));
    HOST_CALL(salloc(&dp.bterm, (size_t)dp.M, s));
    dp.hpts = p64;
    dp.helems = el;
    HIP_CALL(hipMemcpyAsync(dp.pts, p4.data(), p4.size() * sizeof(float4), hipMemcpyHostToDevice, s));
    HIP_CALL(hipMemcpyAsync(dp.pts64, p64.data(), p64.size() * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_CALL(hipMemcpyAsync(dp.elems, el.data(), el.size() * sizeof(int), hipMemcpyHostToDevice, s));
    HIP_CALL(hipMemcpyAsync(dp.bc, bc.data(), bc.size() * sizeof(float2), hipMemcpyHostToDevice, s));
    HIP_CALL(hipMemcpyAsync(dp.tile_ptr, plan.ptr.data(), plan.ptr.size() * sizeof(int), hipMemcpyHostToDevice, s));
    HIP_CALL(hipMemcpyAsync(dp.tile_entries, plan.entries.data(), plan.entries.size() * sizeof(int),
                            hipMemcpyHostToDevice, s));
    if (!plan.isolated.empty())
        HIP_CALL(hipMemcpyAsync(dp.isolated, plan.isolated.data(), plan.isolated.size() * sizeof(int),
                                hipMemcpyHostToDevice, s));
    dp.qg = qg;
    // element-once plan for the FP32 fast kernel: taken when the halo plan
    // integrates at least 5 % of the elements twice and the side columns stay
    // within ghost_slot_bound (device_problem_bytes counts that bound).  It
    // runs at twice the waves as well, so it pays below 10 %: the 40962-node
    // ellipsoid (52 tiles, 7.9 % twice) assembles in 82.9 ms against 91.0 ms.
    dp.use_ghost = 0;
    // Too many side columns for the bound: fewer tiles (two thirds each try;
    // the 2562-node perturbed ellipsoid starts at 320 tiles of 8 nodes, and at
    // 64 tiles assembles in 0.48 ms against 0.70 ms on the halo plan).  A halo
    // under 5 % is left as it is.
    GhostPlan gp;
    int gP = choose_tiles(dp.M, dp.N, dp.E, 64, 2);
    bool take = false;
    for (int tries = 0; qg == 1 && !getenv("BEM_NO_GHOST_PLAN") && tries < 8 && gP >= 4; tries++) {
        build_ghost_plan(m, gP, gp);
        if (gp.P != gP || gp.dup < kGhostMinDup || gp.nslots == 0) break;
        // (ghost_reduce puts the ghost nodes on grid.y: at most 65535)
        if (gp.nslots <= ghost_slot_bound(dp.N, gP) && gp.gnode.size() <= 65535) {
            take = true;
            break;
        }
        gP = gP * 2 / 3;
    }
    {
        if (take) {
            dp.use_ghost = 1;
            dp.gP = gP;
            if (gP > dp.P) {
                // the FP32 tile partials of c and the BC sums, one row per tile
                sfree(dp.cpart, s);
                sfree(dp.bpart, s);
                HOST_CALL(salloc(&dp.cpart, (size_t)gP * dp.Mp, s));
                HOST_CALL(salloc(&dp.bpart, (size_t)gP * dp.Mp, s));
            }
            dp.nslots = gp.nslots;
            dp.ngnodes = (int)gp.gnode.size();
            HOST_CALL(salloc(&dp.gtile_ptr, gp.ptr.size(), s));
            HOST_CALL(salloc(&dp.gtile_entries, gp.entries.size(), s));
            HOST_CALL(salloc(&dp.gtile_slots, gp.slots.size(), s));
            HOST_CALL(salloc(&dp.gred_ptr, gp.gptr.size(), s));
            HOST_CALL(salloc(&dp.gred_node, std::max<size_t>(1, gp.gnode.size()), s));
            HOST_CALL(salloc(&dp.gred_own, std::max<size_t>(1, gp.gown.size()), s));
            HOST_CALL(salloc(&dp.gred_slot, std::max<size_t>(1, gp.gslot.size()), s));
            HOST_CALL(salloc(&dp.side, (size_t)dp.nslots * dp.Mp, s));
            HIP_CALL(hipMemcpyAsync(dp.gtile_ptr, gp.ptr.data(), gp.ptr.size() * sizeof(int), hipMemcpyHostToDevice, s));
            HIP_CALL(hipMemcpyAsync(dp.gtile_entries, gp.entries.data(), gp.entries.size() * sizeof(int),
                                    hipMemcpyHostToDevice, s));
            HIP_CALL(hipMemcpyAsync(dp.gtile_slots, gp.slots.data(), gp.slots.size() * sizeof(int),
                                    hipMemcpyHostToDevice, s));
            HIP_CALL(hipMemcpyAsync(dp.gred_ptr, gp.gptr.data(), gp.gptr.size() * sizeof(int), hipMemcpyHostToDevice, s));
            if (!gp.gnode.empty()) {
                HIP_CALL(hipMemcpyAsync(dp.gred_node, gp.gnode.data(), gp.gnode.size() * sizeof(int),
                                        hipMemcpyHostToDevice, s));
                HIP_CALL(hipMemcpyAsync(dp.gred_own, gp.gown.data(), gp.gown.size() * sizeof(int),
                                        hipMemcpyHostToDevice, s));
                HIP_CALL(hipMemcpyAsync(dp.gred_slot, gp.gslot.data(), gp.gslot.size() * sizeof(int),
                                        hipMemcpyHostToDevice, s));
            }
            HIP_CALL(hipStreamSynchronize(s));  // gp goes out of scope
        }
    }
    dp.impedance = 0;
    for (const auto &e : m.elems)
        if (e.bc[0].x != 0.f || e.bc[0].y != 0.f) dp.impedance = 1;
    const int bs = 256;
    prep_elems<float><<<(dp.E + bs - 1) / bs, bs, 0, s>>>(dp.pts64, dp.elems, dp.bc, dp.E, dp.rec32);
    HIP_CALL(hipGetLastError());
    prep_elems<double><<<(dp.E + bs - 1) / bs, bs, 0, s>>>(dp.pts64, dp.elems, dp.bc, dp.E, dp.rec64);
    HIP_CALL(hipGetLastError());
    HIP_CALL(hipStreamSynchronize(s));  // host vectors above go out of scope
    dp.mesh_version = m.version;
    return EXIT_SUCCESS;
}

void device_problem_free(DeviceProblem &dp) {
    sfree(dp.pts, dp.st);
    sfree(dp.pts64, dp.st);
    sfree(dp.elems, dp.st);
    sfree(dp.bc, dp.st);
    sfree(dp.rec32, dp.st);
    sfree(dp.rec64, dp.st);
    sfree(dp.sing32, dp.st);
    sfree(dp.sing64, dp.st);
    sfree(dp.tile_ptr, dp.st);
    sfree(dp.tile_entries, dp.st);
    sfree(dp.isolated, dp.st);
    sfree(dp.gtile_ptr, dp.st);
    sfree(dp.gtile_entries, dp.st);
    sfree(dp.gtile_slots, dp.st);
    sfree(dp.gred_ptr, dp.st);
    sfree(dp.gred_node, dp.st);
    sfree(dp.gred_own, dp.st);
    sfree(dp.gred_slot, dp.st);
    sfree(dp.side, dp.st);
    dp.use_ghost = 0;
    dp.nslots = dp.ngnodes = 0;
    sfree(dp.cpart, dp.st);
    sfree(dp.bpart, dp.st);
    sfree(dp.cpart64, dp.st);
    sfree(dp.bpart64, dp.st);
    sfree(dp.bterm, dp.st);
    sfree(dp.near_ptr, dp.st);
    sfree(dp.near_row, dp.st);
    sfree(dp.near_elem, dp.st);
    sfree(dp.near_out, dp.st);
    dp.near_factor = 0;
    dp.near_levels = 0;
    dp.near_npairs = 0;
    dp.hnear_ptr.clear();
    dp.hpts.clear();
    dp.helems.clear();
    dp.N = dp.C = dp.M = dp.E = dp.P = 0;
}

static void sing_tab(SingTab &t) {
    QuadTables q;
    quad_tables(q);
    const Config &c = config();
    for (int n = 0; n < kQOrder; n++) {
        const double theta = 0.5 + 0.5 * c.qp[n];
        for (int m = 0; m < kQOrder; m++) {
            const double rho = 0.5 + 0.5 * c.qp[m];
            const int qi = n * kQOrder + m;
            // vertex 1: (1-rho, rho(1-theta)); vertex 2: (rho theta, 1-rho);
            // vertex 3: (rho(1-theta), rho theta) (mesh.cu:272-273, 312-313, 352-353)
            t.xi1[0][qi] = 1 - rho;
            t.xi2[0][qi] = rho * (1 - theta);
            t.xi1[1][qi] = rho * theta;
            t.xi2[1][qi] = 1 - rho;
            t.xi1[2][qi] = rho * (1 - theta);
            t.xi2[2][qi] = rho * theta;
            t.w[qi] = q.dw[qi];
        }
    }
}

// Rows r0..M-1 of the system into A (ld ldA, row r0 first) and their
// boundary-condition sums into bterm; r0 = 0 is the whole collocation system.
template <typename R, typename C2>
static int assemble_impl(DeviceProblem &dp, R k, C2 *A, long ldA, int r0, double2 *bterm, hipStream_t s) {
    HOST_CALL(ensure_quad());
    if (dp.M == 0) BEM_FAIL("device problem not built");
    const Config &cfg = config();
    const double kd = (double)k;
    const double rw = (double)cfg.density * (double)(k * (R)cfg.speed);  // rho * omega, omega = k c
    QuadTables qt;
    quad_tables(qt);
    SingTab st;
    sing_tab(st);
    R *sing = sizeof(R) == 4 ? (R *)dp.sing32 : (R *)dp.sing64;
    {
        TimedRegion tr;
        timing_begin(tr, "sing_elems", s);
        const int n = 3 * dp.E, bs = 128;
        sing_elems<R><<<(n + bs - 1) / bs, bs, 0, s>>>(dp.pts64, dp.elems, dp.bc, dp.E, kd, rw, st, sing);
        HIP_CALL(hipGetLastError());
        timing_end(tr, s);
    }
    AsmArgs<R> a;
    a.pts = dp.pts;
    a.recs = (const ElemRec<R> *)(sizeof(R) == 4 ? (void *)dp.rec32 : (void *)dp.rec64);
    a.sing = sing;
    a.tile_ptr = dp.tile_ptr;
    a.tile_entries = dp.tile_entries;
    a.A = A;
    a.cpart = sizeof(R) == 4 ? (R *)dp.cpart : (R *)dp.cpart64;
    a.bpart = sizeof(R) == 4 ? (C2 *)dp.bpart : (C2 *)dp.bpart64;
    a.ldA = ldA;
    a.Mp = dp.Mp;
    a.M = dp.M;
    a.N = dp.N;
    a.r0 = r0;
    const int rpw = 64 / dp.qg;
    const int nrb = (dp.M - r0 + rpw - 1) / rpw;
    a.nrbg = (nrb + 3) / 4;
    a.nwg = a.nrbg * dp.P;
    a.k = k;
    a.kf = (R)(kd / (2.0 * kPi));
    a.rw = (R)rw;
    a.impedance = dp.impedance;
    for (int q = 0; q < kNQ; q++) {
        a.q.xi1[q] = sizeof(R) == 4 ? (R)qt.xi1[q] : (R)qt.dxi1[q];
        a.q.xi2[q] = sizeof(R) == 4 ? (R)qt.xi2[q] : (R)qt.dxi2[q];
        a.q.w[q] = sizeof(R) == 4 ? (R)qt.w[q] : (R)qt.dw[q];
    }
    a.tile_slots = nullptr;
    a.side = nullptr;
    bool ghost = false;
    {
        TimedRegion tr;
        timing_begin(tr, sizeof(R) == 4 ? "asm_tile_f32" : "asm_tile_f64", s);
        dim3 grid(a.nwg), block(256);
        if (dp.qg == 4) {
            if (dp.impedance) asm_tile<R, 4, true><<<grid, block, 0, s>>>(a);
            else asm_tile<R, 4, false><<<grid, block, 0, s>>>(a);
        } else if (sizeof(R) == 4 && !getenv("BEM_GENERIC_TILE")) {
            // v_sin_f32 / v_cos_f32 are specified for |x| <= 256 revolutions:
            // the phase k r / (2 pi) goes to them unreduced while every
            // distance (at most the bounding-box diagonal of the points) keeps
            // it below that
            const bool raw = kd / (2.0 * kPi) * dp.diam < 250.0;
            ghost = dp.use_ghost && !getenv("BEM_ASM_HALO");
            AsmArgs<float> &af = *(AsmArgs<float> *)&a;
            if (ghost) {
                af.nwg = af.nrbg * dp.gP;
                grid = dim3(af.nwg);
                af.tile_ptr = dp.gtile_ptr;
                af.tile_entries = dp.gtile_entries;
                af.tile_slots = dp.gtile_slots;
                af.side = dp.side;
            }
#define BEM_FAST(IMP_, RAW_)                                                          \
    do {                                                                              \
        if (ghost) asm_tile_fast<IMP_, RAW_, true><<<grid, block, 0, s>>>(af);       \
        else asm_tile_fast<IMP_, RAW_, false><<<grid, block, 0, s>>>(af);            \
    } while (0)
            if (dp.impedance) {
                if (raw) BEM_FAST(true, true);
                else BEM_FAST(true, false);
            } else {
                if (raw) BEM_FAST(false, true);
                else BEM_FAST(false, false);
            }
#undef BEM_FAST
        } else {
            if (dp.impedance) asm_tile<R, 1, true><<<grid, block, 0, s>>>(a);
            else asm_tile<R, 1, false><<<grid, block, 0, s>>>(a);
        }
        HIP_CALL(hipGetLastError());
        timing_end(tr, s);
    }
    const int nrows = dp.M - r0;
    if (ghost && dp.ngnodes > 0) {
        if (ldA % 2 || dp.Mp % 2) BEM_FAIL("ghost_reduce needs even leading dimensions (ldA %ld, Mp %ld)", ldA, dp.Mp);
        TimedRegion tr;
        timing_begin(tr, "ghost_reduce", s);
        dim3 grid((nrows + 511) / 512, dp.ngnodes);
        ghost_reduce<<<grid, 256, 0, s>>>((float2 *)A, ldA, nrows, dp.side, dp.Mp, dp.gred_ptr, dp.gred_node,
                                          dp.gred_own, dp.gred_slot);
        HIP_CALL(hipGetLastError());
        timing_end(tr, s);
    }
    if (dp.n_isolated > 0) {
        dim3 grid((nrows + 255) / 256, dp.n_isolated);
        zero_columns<C2><<<grid, 256, 0, s>>>(A, ldA, nrows, dp.isolated, dp.n_isolated);
        HIP_CALL(hipGetLastError());
    }
    {
        TimedRegion tr;
        timing_begin(tr, "asm_finalize", s);
        asm_finalize<R, C2><<<(nrows + 255) / 256, 256, 0, s>>>(A, ldA, dp.M, dp.N, ghost ? dp.gP : dp.P, dp.Mp, r0,
                                                             a.cpart,
                                                             a.bpart, bterm);
        HIP_CALL(hipGetLastError());
        timing_end(tr, s);
    }
    // opt-in FP64 near-singular corrections, after every row is complete
    if (cfg.near_factor > 0 || dp.near_factor > 0) HOST_CALL((near_correct<R, C2>(dp, kd, rw, A, ldA, r0, bterm, s)));
    return EXIT_SUCCESS;
}

int assemble_fp32(DeviceProblem &dp, float k, float2 *A, hipStream_t s) {
    return assemble_impl<float, float2>(dp, k, A, dp.ldA, 0, dp.bterm, s);
}
int assemble_fp64(DeviceProblem &dp, double k, double2 *A, hipStream_t s) {
    return assemble_impl<double, double2>(dp, k, A, dp.ldA, 0, dp.bterm, s);
}

template <typename C2>
static int rhs_impl(DeviceProblem &dp, double k, C2 *B, long ldB, const RhsSpec &rhs, hipStream_t s) {
    if (rhs.nrhs <= 0) return EXIT_SUCCESS;
    if (rhs.nrhs > kMaxRhs) BEM_FAIL("at most %d right-hand sides, got %d", kMaxRhs, rhs.nrhs);
    float *params = nullptr;
    HOST_CALL(salloc(&params, 3 * (size_t)rhs.nrhs, s));
    HIP_CALL(hipMemcpyAsync(params, rhs.params.data(), 3 * sizeof(float) * rhs.nrhs, hipMemcpyHostToDevice, s));
    dim3 grid((dp.M + 255) / 256, rhs.nrhs);
    rhs_fill<C2><<<grid, 256, 0, s>>>(dp.pts64, dp.M, k, rhs.kind, (double)rhs.amp, params, rhs.nrhs, dp.bterm, B, ldB);
    hipError_t e = hipGetLastError();
    (void)hipStreamSynchronize(s);
    sfree(params, s);
    HIP_CALL(e);
    return EXIT_SUCCESS;
}

// ------------------------------------------------- field as a GEMM ---
// fdp is the mesh with the P field points as its extra rows (the CHIEF
// position).  Rows N..N+P-1 of the collocation formula are the radiation
// operator: p(x) = inc(x) - bterm(x) - sum_j F[x, j] p_j (genExtPressure,
// mesh.cu:1038-1073, as a matrix).  F is P x N (ldF), fbterm P values.
int field_matrix_fp32(DeviceProblem &fdp, float k, float2 *F, long ldF, double2 *fbterm, hipStream_t s) {
    if (fdp.M <= fdp.N) BEM_FAIL("field problem has no field points");
    if (ldF < fdp.M - fdp.N) BEM_FAIL("ldF must be >= P");
    return assemble_impl<float, float2>(fdp, k, F, ldF, fdp.N, fbterm, s);
}

__global__ void k_sub_c(float2 *out, long ldo, const float2 *__restrict__ Y, long ldy, int P) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (i >= P) return;
    const float2 y = Y[(size_t)r * ldy + i];
    float2 o = out[(size_t)r * ldo + i];
    o.x -= y.x;
    o.y -= y.y;
    out[(size_t)r * ldo + i] = o;
}

// out (P x R, ldo) = inc_r(x_p) - bterm_p - (F X)_p for the FP32 solutions X
// (N x R, ldx): one GEMV (R = 1) or one MFMA product (R > 1) over F.
// ws >= max(ws_gemv_n(P, N, 1), ws_gemm(P, N, R)); Y: P x R scratch (ldF).
int field_apply_fp32(DeviceProblem &fdp, double k, const float2 *F, long ldF, const double2 *fbterm, const float2 *X,
                     long ldx, const RhsSpec &rhs, float2 *Y, float2 *out, long ldo, void *ws, hipStream_t s) {
    const int N = fdp.N, P = fdp.M - fdp.N, R = rhs.nrhs;
    if (R <= 0 || P <= 0) return EXIT_SUCCESS;
    {
        TimedRegion tr;
        timing_begin(tr, "field_gemm", s);
        if (R == 1) HOST_CALL(gemv_n_f32(F, ldF, P, N, X, 0, Y, 0, 1, ws, s));
        else HOST_CALL(gemm_n_f32(F, ldF, P, N, X, ldx, Y, ldF, R, ws, s));
        timing_end(tr, s);
    }
    float *params = nullptr;
    HOST_CALL(salloc(&params, 3 * (size_t)R, s));
    HIP_CALL(hipMemcpyAsync(params, rhs.params.data(), 3 * sizeof(float) * R, hipMemcpyHostToDevice, s));
    dim3 grid((P + 255) / 256, R);
    rhs_fill<float2><<<grid, 256, 0, s>>>(fdp.pts64 + 3 * (size_t)N, P, k, rhs.kind, (double)rhs.amp, params, R,
                                          fbterm, out, ldo);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        k_sub_c<<<grid, 256, 0, s>>>(out, ldo, Y, ldF, P);
        e = hipGetLastError();
    }
    (void)hipStreamSynchronize(s);  // params is a host-side copy source
    sfree(params, s);
    HIP_CALL(e);
    return EXIT_SUCCESS;
}

int rhs_fp32(DeviceProblem &dp, double k, float2 *B, long ldB, const RhsSpec &rhs, hipStream_t s) {
    return rhs_impl<float2>(dp, k, B, ldB, rhs, s);
}
int rhs_fp64(DeviceProblem &dp, double k, double2 *B, long ldB, const RhsSpec &rhs, hipStream_t s) {
    return rhs_impl<double2>(dp, k, B, ldB, rhs, s);
}

}  // namespace bem
