// Restarted GMRES on the device with FP64 Arnoldi orthogonalisation.
//
// The reference GMRES (GMRES.cu:90-279) keeps the m x m Krylov basis on the
// host, re-uploads A on every Arnoldi step and runs modified Gram-Schmidt with
// one host round trip per basis vector.  Here A, the basis and the residuals
// stay in HBM; each Arnoldi step is one operator application (the HBM-bound
// GEMV), a classical Gram-Schmidt pass repeated twice (CGS2) in FP64, and one
// small device-to-host copy of the new Hessenberg column for the Givens
// rotations (GMRES.cu:44-88) done on the host.
//
// Every restart cycle starts from the true residual computed with FP64
// accumulation (b - A x for a square system, A^H (b - A x) for the normal
// equations of an overdetermined CHIEF system, which is GMRES on
// A^H A x = A^H b, i.e. BA-GMRES with B = A^H, Hayami, Yin & Ito 2010).  With an
// FP32 matrix this is iterative refinement: the Krylov correction runs with FP32
// operator applications, the residual and the accumulated solution are FP64.
//
// Many right-hand sides (plane waves, sources) run in lockstep: every system
// keeps its own basis, Hessenberg matrix and rotations, while the operator is
// applied to all current basis vectors at once, so A is streamed from HBM once
// per Arnoldi step for the whole batch (the MFMA products of gemm.hip) instead
// of once per system.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <map>
#include <memory>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <complex>
#include <vector>

#include "bem_internal.hpp"
#include "hessenberg_ls.hpp"

namespace bem {

int gemv_n_f32a(const float2 *A, long lda, long M, long N, const double2 *x, long ldx, double2 *y, long ldy, int R,
                void *ws, hipStream_t s);
int resid_n_f64(const double2 *A, long lda, long M, long N, const double2 *x, long ldx, const double2 *b, long ldb,
                double2 *r, long ldr, int R, void *ws, hipStream_t s);

using zc = std::complex<double>;

// ---------------------------------------------------------------- kernels ---
// CGS2 steps for R systems at once: system r = blockIdx.y, its vectors at
// fixed strides (basis vstride, vectors wstride, Hessenberg columns hstride).
// part[blk * nv + l] of system r (blockIdx.y) = sum over the 1024 rows of
// block blk of conj(V_l) w; one wave per basis vector l (4 per workgroup,
// blockIdx.z selects the group), no workgroup barriers.
__global__ __launch_bounds__(256) void kb_mdot_part(const double2 *__restrict__ V, long ldv, long vstride, int nv,
                                                    const double2 *__restrict__ w, long wstride, int n,
                                                    double2 *__restrict__ part, long pstride) {
    const int r = blockIdx.y;
    const int l = blockIdx.z * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (l >= nv) return;
    const double2 *vl = V + (size_t)r * vstride + (size_t)l * ldv;
    const double2 *wr = w + (size_t)r * wstride;
    const int base = blockIdx.x * 1024;
    double sr = 0, si = 0;
    // loads in flight 8 pairs at a time: clamped rows, v zeroed past n
#pragma unroll
    for (int u0 = 0; u0 < 16; u0 += 8) {
        double2 v[8], x[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int i = base + lane + 64 * (u0 + u), ic = min(i, n - 1);
            v[u] = vl[ic];
            x[u] = wr[ic];
            if (i >= n) v[u] = make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            sr = fma(v[u].x, x[u].x, fma(v[u].y, x[u].y, sr));
            si = fma(v[u].x, x[u].y, fma(-v[u].y, x[u].x, si));
        }
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        sr += __shfl_xor(sr, m, 64);
        si += __shfl_xor(si, m, 64);
    }
    if (lane == 0) part[(size_t)r * pstride + (size_t)blockIdx.x * nv + l] = make_double2(sr, si);
}

// out[r * ostride + l] = sum_blocks part; with acc: acc[r * ostride + l] += the sum
__global__ void kb_mdot_final(const double2 *__restrict__ part, long pstride, int nblk, int nv,
                              double2 *__restrict__ out, long ostride, double2 *__restrict__ acc, long astride) {
    const int r = blockIdx.y, l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nv) return;
    part += (size_t)r * pstride;
    double sr = 0, si = 0;
    int b = 0;
    for (; b + 8 <= nblk; b += 8) {
        double2 p[8];
#pragma unroll
        for (int u = 0; u < 8; u++) p[u] = part[(size_t)(b + u) * nv + l];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            sr += p[u].x;
            si += p[u].y;
        }
    }
    for (; b < nblk; b++) {
        const double2 p = part[(size_t)b * nv + l];
        sr += p.x;
        si += p.y;
    }
    out[(size_t)r * ostride + l] = make_double2(sr, si);
    if (acc) {
        double2 a = acc[(size_t)r * astride + l];
        a.x += sr;
        a.y += si;
        acc[(size_t)r * astride + l] = a;
    }
}

// out[r * ostride] = (|w_r|^2, 0): one workgroup per system, fixed-order tree
__global__ __launch_bounds__(1024) void kb_norm2(const double2 *__restrict__ w, long wstride, int n,
                                                 double2 *__restrict__ out, long ostride) {
    __shared__ double red[1024];
    const double2 *wr = w + (size_t)blockIdx.x * wstride;
    double sum = 0;
    int i = threadIdx.x;
    for (; i + 3 * 1024 < n; i += 4 * 1024) {
        double2 x[4];
#pragma unroll
        for (int u = 0; u < 4; u++) x[u] = wr[i + 1024 * u];
#pragma unroll
        for (int u = 0; u < 4; u++) sum = fma(x[u].x, x[u].x, fma(x[u].y, x[u].y, sum));
    }
    for (; i < n; i += 1024) sum = fma(wr[i].x, wr[i].x, fma(wr[i].y, wr[i].y, sum));
    red[threadIdx.x] = sum;
    __syncthreads();
    for (int m = 512; m >= 1; m >>= 1) {
        if (threadIdx.x < m) red[threadIdx.x] += red[threadIdx.x + m];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[(size_t)blockIdx.x * ostride] = make_double2(red[0], 0.0);
}

__global__ void kb_maxpy(const double2 *__restrict__ V, long ldv, long vstride, int nv, const double2 *__restrict__ h,
                         long hstride, double sign, double2 *__restrict__ w, long wstride, int n) {
    const int r = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    V += (size_t)r * vstride;
    h += (size_t)r * hstride;
    w += (size_t)r * wstride;
    double sr = 0, si = 0;
    // basis vectors 8 at a time with their loads in flight, added in order
    int l = 0;
    for (; l + 8 <= nv; l += 8) {
        double2 v[8], c[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            v[u] = V[(size_t)(l + u) * ldv + i];
            c[u] = h[l + u];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            sr = fma(v[u].x, c[u].x, fma(-v[u].y, c[u].y, sr));
            si = fma(v[u].x, c[u].y, fma(v[u].y, c[u].x, si));
        }
    }
    for (; l < nv; l++) {
        const double2 v = V[(size_t)l * ldv + i];
        const double2 c = h[l];
        sr = fma(v.x, c.x, fma(-v.y, c.y, sr));
        si = fma(v.x, c.y, fma(v.y, c.x, si));
    }
    double2 o = w[i];
    o.x += sign * sr;
    o.y += sign * si;
    w[i] = o;
}

__global__ void kb_normalize(const double2 *__restrict__ w, long wstride, const double2 *__restrict__ nrm2,
                             long nstride, double2 *__restrict__ out, long ostride, float2 *__restrict__ out32,
                             long o32stride, int n) {
    const int r = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double q = nrm2[(size_t)r * nstride].x;
    const double sc = q > 0 ? 1.0 / sqrt(q) : 0.0;
    const double2 x = w[(size_t)r * wstride + i];
    const double2 v = make_double2(x.x * sc, x.y * sc);
    out[(size_t)r * ostride + i] = v;
    if (out32) out32[(size_t)r * o32stride + i] = make_float2((float)v.x, (float)v.y);
}

// One Gram-Schmidt pass after kb_mdot_part, in one launch: every workgroup
// forms h_r[l] = sum_b part_r[b][l] (fixed order) in LDS; workgroup 0 stores
// it to hout (and adds it to acc); then w_r += sign V_r h_r over the
// workgroup's 256 rows, and the workgroup's partial of |w_r|^2 (fixed-order
// tree) goes to npart_r[blockIdx.x].  Dynamic LDS: nv double2.
__global__ __launch_bounds__(256) void kb_final_maxpy(const double2 *__restrict__ part, long pstride, int nblk,
                                                      const double2 *__restrict__ V, long ldv, long vstride, int nv,
                                                      double2 *__restrict__ hout, long hstride,
                                                      double2 *__restrict__ acc, long astride, double sign,
                                                      double2 *__restrict__ w, long wstride, int n,
                                                      double *__restrict__ npart, long nstride) {
    extern __shared__ double2 s_h[];
    __shared__ double s_n[4];
    const int r = blockIdx.y;
    const double2 *pr = part + (size_t)r * pstride;
    for (int l = threadIdx.x; l < nv; l += 256) {
        double sr = 0, si = 0;
        int b = 0;
        for (; b + 8 <= nblk; b += 8) {
            double2 q[8];
#pragma unroll
            for (int u = 0; u < 8; u++) q[u] = pr[(size_t)(b + u) * nv + l];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                sr += q[u].x;
                si += q[u].y;
            }
        }
        for (; b < nblk; b++) {
            const double2 q = pr[(size_t)b * nv + l];
            sr += q.x;
            si += q.y;
        }
        s_h[l] = make_double2(sr, si);
        if (blockIdx.x == 0) {
            hout[(size_t)r * hstride + l] = make_double2(sr, si);
            if (acc) {
                double2 a = acc[(size_t)r * astride + l];
                a.x += sr;
                a.y += si;
                acc[(size_t)r * astride + l] = a;
            }
        }
    }
    __syncthreads();
    const int i = blockIdx.x * 256 + threadIdx.x;
    double nrm = 0;
    if (i < n) {
        const double2 *Vr = V + (size_t)r * vstride;
        double sr = 0, si = 0;
        int l = 0;
        for (; l + 8 <= nv; l += 8) {
            double2 v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) v[u] = Vr[(size_t)(l + u) * ldv + i];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const double2 c = s_h[l + u];
                sr = fma(v[u].x, c.x, fma(-v[u].y, c.y, sr));
                si = fma(v[u].x, c.y, fma(v[u].y, c.x, si));
            }
        }
        for (; l < nv; l++) {
            const double2 v = Vr[(size_t)l * ldv + i];
            const double2 c = s_h[l];
            sr = fma(v.x, c.x, fma(-v.y, c.y, sr));
            si = fma(v.x, c.y, fma(v.y, c.x, si));
        }
        double2 o = w[(size_t)r * wstride + i];
        o.x += sign * sr;
        o.y += sign * si;
        w[(size_t)r * wstride + i] = o;
        nrm = fma(o.x, o.x, o.y * o.y);
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) nrm += __shfl_xor(nrm, m, 64);
    if ((threadIdx.x & 63) == 0) s_n[threadIdx.x >> 6] = nrm;
    __syncthreads();
    if (threadIdx.x == 0) npart[(size_t)r * nstride + blockIdx.x] = (s_n[0] + s_n[1]) + (s_n[2] + s_n[3]);
}

// |w_r|^2 = sum of the nb partials of kb_final_maxpy (one wave, fixed order);
// workgroup 0 stores it to hslot_r; out_r = w_r / |w_r| (and its FP32 copy)
__global__ __launch_bounds__(256) void kb_normalize_np(const double2 *__restrict__ w, long wstride,
                                                       const double *__restrict__ npart, long nstride, int nb,
                                                       double2 *__restrict__ hslot, long hsstride,
                                                       double2 *__restrict__ out, long ostride,
                                                       float2 *__restrict__ out32, long o32stride, int n) {
    __shared__ double s_q;
    const int r = blockIdx.y;
    if (threadIdx.x < 64) {
        const double *np = npart + (size_t)r * nstride;
        double q = 0;
        for (int b = threadIdx.x; b < nb; b += 64) q += np[b];
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) q += __shfl_xor(q, m, 64);
        if (threadIdx.x == 0) {
            s_q = q;
            if (blockIdx.x == 0) hslot[(size_t)r * hsstride] = make_double2(q, 0.0);
        }
    }
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double q = s_q;
    const double sc = q > 0 ? 1.0 / sqrt(q) : 0.0;
    const double2 x = w[(size_t)r * wstride + i];
    const double2 v = make_double2(x.x * sc, x.y * sc);
    out[(size_t)r * ostride + i] = v;
    if (out32) out32[(size_t)r * o32stride + i] = make_float2((float)v.x, (float)v.y);
}

__global__ void kb_cvt(const float2 *__restrict__ in, long istride, double2 *__restrict__ out, long ostride, int n) {
    const int r = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float2 v = in[(size_t)r * istride + i];
        out[(size_t)r * ostride + i] = make_double2(v.x, v.y);
    }
}

// ---------------------------------------------------------------- driver ---
constexpr int kMaxChunk = 8;  // Arnoldi steps between reads of the Hessenberg columns
constexpr size_t kMaxGraphs = 512;  // captured Arnoldi chunks kept per workspace
constexpr int kGraphAfter = 2;       // eager launches of a chunk before it is captured

struct KrylovWS {
    int n = 0, m = 0, R = 0, restart = 0, nblk = 0;
    long ldv = 0, vstride = 0, hstride = 0, pstride = 0;
    double2 *V = nullptr, *w64 = nullptr, *t64 = nullptr, *r = nullptr, *h = nullptr, *part = nullptr;
    float2 *v32 = nullptr, *w32 = nullptr, *t32 = nullptr;
    char *gws = nullptr;
    // row-major copy of A for the one-pass normal-equation product (one
    // right-hand side, FP32 A); null when the two GEMVs are used
    float2 *Ar = nullptr;
    long ldr = 0;
    char *nws = nullptr;
    // conjugate-transposed copy of an FP32 A for the adjoint products of the
    // normal equations (several right-hand sides); refreshed every solve
    float2 *AH = nullptr;
    double2 *hch = nullptr;  // Hessenberg columns of an Arnoldi chunk
    double *npart = nullptr;  // per system: |w|^2 partials of the 256-row workgroups
    long nstride = 0;
    hipStream_t st = nullptr;
    ~KrylovWS() {
        sfree(V, st); sfree(w64, st); sfree(t64, st); sfree(r, st); sfree(h, st); sfree(part, st);
        sfree(v32, st); sfree(w32, st); sfree(t32, st); sfree(gws, st); sfree(Ar, st); sfree(nws, st);
        sfree(AH, st); sfree(hch, st); sfree(npart, st);
    }
    int init(int n_, int m_, int R_, int restart_, hipStream_t s) {
        st = s;
        n = n_;
        m = m_;
        R = R_;
        restart = restart_;
        ldv = ((long)std::max(n, m) + 63) / 64 * 64;
        vstride = ldv * (restart + 1);
        hstride = 2L * (restart + 2);
        nblk = (int)((std::max(n, m) + 1023) / 1024);
        pstride = (long)nblk * (restart + 2);
        HOST_CALL(salloc(&V, (size_t)vstride * R, s));
        HOST_CALL(salloc(&w64, (size_t)ldv * R, s));
        HOST_CALL(salloc(&t64, (size_t)ldv * R, s));
        HOST_CALL(salloc(&r, (size_t)ldv * R, s));
        HOST_CALL(salloc(&h, (size_t)hstride * R, s));
        HOST_CALL(salloc(&part, (size_t)pstride * R, s));
        HOST_CALL(salloc(&v32, (size_t)ldv * R, s));
        HOST_CALL(salloc(&w32, (size_t)ldv * R, s));
        HOST_CALL(salloc(&t32, (size_t)ldv * R, s));
        const size_t g = std::max(std::max(ws_gemv_n(m, n, 1), ws_gemm(m, n, R)), ws_mres(m, n, R));
        HOST_CALL(salloc(&gws, g, s));
        nstride = (n + 255) / 256;
        HOST_CALL(salloc(&npart, (size_t)nstride * R, s));
        return EXIT_SUCCESS;
    }
};

// batched dots <V_r[:, 0..nv), w_r> -> out[r * ostride + l] (and acc += out)
static int bdot(KrylovWS &ws, const double2 *V, long vstride, int nv, const double2 *w, long wstride, int n,
                double2 *out, long ostride, hipStream_t s, double2 *acc = nullptr, long astride = 0) {
    const int nblk = (n + 1023) / 1024;
    dim3 g1(nblk, ws.R, (nv + 3) / 4);
    kb_mdot_part<<<g1, 256, 0, s>>>(V, ws.ldv, vstride, nv, w, wstride, n, ws.part, ws.pstride);
    HIP_CALL(hipGetLastError());
    // one wave per 64 sums, at most 256 threads per workgroup (long restarts
    // take several workgroups per system)
    const int bt = std::min(256, 64 * ((nv + 63) / 64));
    dim3 g2((nv + bt - 1) / bt, ws.R);
    kb_mdot_final<<<g2, bt, 0, s>>>(ws.part, ws.pstride, nblk, nv, out, ostride, acc, astride ? astride : ostride);
    HIP_CALL(hipGetLastError());
    return EXIT_SUCCESS;
}

// the partials of V_r^H w_r alone (kb_final_maxpy sums them)
static int bdot_part(KrylovWS &ws, const double2 *V, long vstride, int nv, const double2 *w, long wstride, int n,
                     hipStream_t s) {
    const int nblk = (n + 1023) / 1024;
    dim3 g1(nblk, ws.R, (nv + 3) / 4);
    kb_mdot_part<<<g1, 256, 0, s>>>(V, ws.ldv, vstride, nv, w, wstride, n, ws.part, ws.pstride);
    HIP_CALL(hipGetLastError());
    return EXIT_SUCCESS;
}

// |w_r|^2 -> out[r * ostride]
static int bnorm2(KrylovWS &ws, const double2 *w, long wstride, int n, double2 *out, long ostride, hipStream_t s) {
    kb_norm2<<<ws.R, 1024, 0, s>>>(w, wstride, n, out, ostride);
    HIP_CALL(hipGetLastError());
    return EXIT_SUCCESS;
}

// The operator of the Krylov iteration: square A, or A^H A for M > N.
struct Operator {
    const float2 *A32 = nullptr;
    const float2 *AH = nullptr;  // conjugate-transposed copy of A32, or null
    const double2 *A64 = nullptr;
    long lda = 0;
    int M = 0, N = 0;
    // nmat > 1 square FP32 matrices astride apart, each for R / nmat systems
    long astride = 0;
    int nmat = 1;
    bool normal() const { return M > N; }
    // w64_r <- op(V_r[:, j]) for every system (FP32 A reads v32)
    int apply(KrylovWS &ws, int j, hipStream_t s) const {
        const int bs = 256;
        const int R = ws.R;
        if (A32) {
            if (R == 1 && normal() && ws.Ar) {
                // A^H (A v) in one pass over the row-major copy, FP64 out
                return ne_apply_f32(ws.Ar, ws.ldr, M, N, ws.v32, ws.w64, ws.nws, s);
            }
            if (R == 1) {
                if (!normal()) {
                    HOST_CALL(gemv_n_f32(A32, lda, M, N, ws.v32, 0, ws.w32, 0, 1, ws.gws, s));
                } else {
                    HOST_CALL(gemv_n_f32(A32, lda, M, N, ws.v32, 0, ws.t32, 0, 1, ws.gws, s));
                    HOST_CALL(gemv_h_f32(A32, lda, M, N, ws.t32, 0, ws.w32, 0, 1, s));
                }
                dim3 g((N + bs - 1) / bs, R);
                kb_cvt<<<g, bs, 0, s>>>(ws.w32, ws.ldv, ws.w64, ws.ldv, N);
                HIP_CALL(hipGetLastError());
            } else {
                // the products' fixed-order reduce writes the FP64 vector directly
                if (nmat > 1) {
                    HOST_CALL(gemm_n_f32z_batched(A32, lda, astride, nmat, M, N, ws.v32, ws.ldv, ws.w64, ws.ldv, R,
                                                  ws.gws, s));
                } else if (!normal()) {
                    HOST_CALL(gemm_n_f32z(A32, lda, M, N, ws.v32, ws.ldv, ws.w64, ws.ldv, R, ws.gws, s));
                } else {
                    HOST_CALL(gemm_n_f32(A32, lda, M, N, ws.v32, ws.ldv, ws.t32, ws.ldv, R, ws.gws, s, true));
                    if (AH) HOST_CALL(gemm_hc_f32z(AH, M, N, ws.t32, ws.ldv, ws.w64, ws.ldv, R, ws.gws, s));
                    else HOST_CALL(gemm_h_f32z(A32, lda, M, N, ws.t32, ws.ldv, ws.w64, ws.ldv, R, ws.gws, s));
                }
            }
        } else {
            for (int q = 0; q < R; q++) {
                const double2 *vj = ws.V + (size_t)q * ws.vstride + (size_t)j * ws.ldv;
                double2 *w = ws.w64 + (size_t)q * ws.ldv, *t = ws.t64 + (size_t)q * ws.ldv;
                if (!normal()) {
                    HOST_CALL(gemv_n_f64(A64, lda, M, N, vj, 0, w, 0, 1, ws.gws, s));
                } else {
                    HOST_CALL(gemv_n_f64(A64, lda, M, N, vj, 0, t, 0, 1, ws.gws, s));
                    HOST_CALL(gemv_h_f64(A64, lda, M, N, t, 0, w, 0, 1, s));
                }
            }
        }
        return EXIT_SUCCESS;
    }
    // r <- the residual at x = 0 (square: b; normal equations: A^H b)
    int rhs(KrylovWS &ws, const double2 *B, long ldb, double2 *r, hipStream_t s) const {
        const int R = ws.R;
        if (!normal()) {
            HIP_CALL(hipMemcpy2DAsync(r, sizeof(double2) * ws.ldv, B, sizeof(double2) * ldb, sizeof(double2) * N, R,
                                      hipMemcpyDeviceToDevice, s));
        } else if (A32) {
            if (R == 1) HOST_CALL(gemv_h_f32a(A32, lda, M, N, B, 0, r, 0, 1, s));
            else if (AH) HOST_CALL(mh_f32a_hc(AH, M, N, B, ldb, r, ws.ldv, R, ws.gws, s));
            else HOST_CALL(mh_f32a(A32, lda, M, N, B, ldb, r, ws.ldv, R, ws.gws, s));
        } else {
            if (R == 1) HOST_CALL(gemv_h_f64(A64, lda, M, N, B, 0, r, 0, 1, s));
            else HOST_CALL(mh_f64(A64, lda, M, N, B, ldb, r, ws.ldv, R, s));
        }
        return EXIT_SUCCESS;
    }
    // r <- rhs - op(x) in FP64 for every system: square b - A x; normal A^H (b - A x)
    int residual(KrylovWS &ws, const double2 *B, long ldb, const double2 *X, long ldx, double2 *r,
                 hipStream_t s) const {
        const int R = ws.R;
        if (A32 && R == 1 && normal() && ws.Ar && !getenv("BEM_NE_RESID_TWO_PASS")) {
            // A^H (b - A x) in one pass over the row-major copy, FP64 accumulation
            return ne_resid_f32a(ws.Ar, ws.ldr, M, N, X, B, r, ws.nws, s);
        }
        if (A32) {
            if (R == 1) {
                if (!normal()) {
                    HOST_CALL(resid_n_f32a(A32, lda, M, N, X, 0, B, 0, r, 0, 1, ws.gws, s));
                } else {
                    HOST_CALL(resid_n_f32a(A32, lda, M, N, X, 0, B, 0, ws.t64, 0, 1, ws.gws, s));
                    HOST_CALL(gemv_h_f32a(A32, lda, M, N, ws.t64, 0, r, 0, 1, s));
                }
            } else if (nmat > 1) {
                HOST_CALL(mres_f32a_batched(A32, lda, astride, nmat, M, N, X, ldx, B, ldb, r, ws.ldv, R, ws.gws, s));
            } else if (!normal()) {
                HOST_CALL(mres_f32a(A32, lda, M, N, X, ldx, B, ldb, r, ws.ldv, R, ws.gws, s));
            } else {
                HOST_CALL(mres_f32a(A32, lda, M, N, X, ldx, B, ldb, ws.t64, ws.ldv, R, ws.gws, s));
                if (AH) HOST_CALL(mh_f32a_hc(AH, M, N, ws.t64, ws.ldv, r, ws.ldv, R, ws.gws, s));
                else HOST_CALL(mh_f32a(A32, lda, M, N, ws.t64, ws.ldv, r, ws.ldv, R, ws.gws, s));
            }
        } else if (R == 1) {
            if (!normal()) {
                HOST_CALL(resid_n_f64(A64, lda, M, N, X, 0, B, 0, r, 0, 1, ws.gws, s));
            } else {
                HOST_CALL(resid_n_f64(A64, lda, M, N, X, 0, B, 0, ws.t64, 0, 1, ws.gws, s));
                HOST_CALL(gemv_h_f64(A64, lda, M, N, ws.t64, 0, r, 0, 1, s));
            }
        } else {
            if (!normal()) {
                HOST_CALL(mres_f64(A64, lda, M, N, X, ldx, B, ldb, r, ws.ldv, R, ws.gws, s));
            } else {
                HOST_CALL(mres_f64(A64, lda, M, N, X, ldx, B, ldb, ws.t64, ws.ldv, R, ws.gws, s));
                HOST_CALL(mh_f64(A64, lda, M, N, ws.t64, ws.ldv, r, ws.ldv, R, s));
            }
        }
        return EXIT_SUCCESS;
    }
};

// Restarted GMRES for R right-hand sides in lockstep: each system has its own
// FP64 Krylov basis, Hessenberg matrix and Givens rotations; the operator is
// applied to all of them at once (one MFMA product per Arnoldi step when
// R > 1).  Each cycle starts from the true FP64 residual and runs up to
// `restart` steps; a system leaves the cycle when its estimated residual fell by
// `tol` relative to the cycle's start (or by only as much as reaching
// 0.1 outer_tol takes, when that is less), and the solve stops when every system's
// FP64 residual is below `outer_tol` relative to ||rhs|| (or maxit / max_cycles).
// B: M x R (ldb), X: N x R (ldx, initial guess on entry).
struct GmresCache {
    KrylovWS ws;
    bool ne = false;
    // pinned host copies of the norms, the Hessenberg columns of a chunk and
    // the update coefficients (device-to-host reads of pageable memory are
    // staged, and each one stalls the host loop)
    double2 *pin = nullptr;
    // Arnoldi chunks captured as HIP graphs, by (first step, steps, second
    // Gram-Schmidt pass): the kernels' arguments depend only on those and on
    // the workspace and operator pointers, so a chunk replays in later cycles
    // and later solves of the same shape.  A chunk is captured on its third
    // launch (a one-off chunk costs more to capture than to launch).  Valid
    // for the operator recorded with them.
    std::map<std::array<int, 3>, hipGraphExec_t> graphs;
    std::map<std::array<int, 3>, int> uses;  // eager launches of each chunk so far
    const void *graph_op = nullptr;
    long graph_lda = 0;
    void clear_graphs() {
        for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
        uses.clear();
    }
    ~GmresCache() {
        clear_graphs();
        if (pin) (void)hipHostFree(pin);
    }
    bool matches(int N, int M, int R, int restart, bool ne_, hipStream_t s) const {
        return ws.n == N && ws.m == M && ws.R == R && ws.restart == restart && ne == ne_ && ws.st == s;
    }
};

void gmres_cache_free(GmresCache *&c) {
    delete c;
    c = nullptr;
}

int gmres_device(const float2 *A32, const double2 *A64, long lda, int M, int N, const double2 *B, long ldb,
                 double2 *X, long ldx, bool x_zero, int R, int restart, double tol, int maxit, double outer_tol,
                 int max_cycles, bem_solve_stats *st, hipStream_t s, GmresCache **cache, long astride, int nmat) {
    if (M < N) BEM_FAIL("GMRES needs M >= N (got %d x %d)", M, N);
    if (R <= 0) return EXIT_SUCCESS;
    if (nmat > 1 && (!A32 || M != N || R % nmat || R / nmat < 2))
        BEM_FAIL("GMRES over %d matrices needs square FP32 matrices and at least 2 right-hand sides each", nmat);
    restart = std::max(1, std::min(restart, N));
    const int K = restart;
    // Hessenberg column slots of a chunk: per system kMaxChunk slots of K + 2
    // entries (the products at 0..j, |w|^2 at K + 1)
    const long hs = K + 2, hcs = hs * kMaxChunk;
    const bool ne = A32 && M > N && R == 1 && ne_fused_fits(M, N) && !getenv("BEM_NE_TWO_PASS");
    std::unique_ptr<GmresCache> local;
    GmresCache *gc = nullptr;
    if (cache && *cache && (*cache)->matches(N, M, R, restart, ne, s)) {
        gc = *cache;
    } else {
        if (cache) gmres_cache_free(*cache);
        local.reset(new GmresCache);
        local->ne = ne;
        KrylovWS &w = local->ws;
        HOST_CALL(w.init(N, M, R, restart, s));
        if (ne) {
            w.ldr = ne_ldr(N);
            HOST_CALL(salloc(&w.Ar, (size_t)M * w.ldr, s));
            HOST_CALL(salloc(&w.nws, std::max({ws_ne_apply(M, N), ws_ne_resid(M, N), ws_ne_rhs(M, N)}), s));
        }
        HOST_CALL(salloc(&w.hch, (size_t)hcs * R, s));
        if (cache)
            HIP_CALL(hipHostMalloc((void **)&local->pin, sizeof(double2) * ((size_t)w.hstride * R * 2 + hcs * R),
                                   hipHostMallocDefault));
        if (cache) {
            *cache = local.release();
            gc = *cache;
        } else {
            gc = local.get();
        }
    }
    KrylovWS &ws = gc->ws;
    // Arnoldi chunks replay as captured graphs when the workspace is kept
    // between solves and BEM_GMRES_GRAPH is set (not while kernel timing
    // records events around launches)
    const void *opp = A32 ? (const void *)A32 : (const void *)A64;
    if (gc->graph_op != opp || gc->graph_lda != lda) {
        gc->clear_graphs();
        gc->graph_op = opp;
        gc->graph_lda = lda;
    }
    // opt-in: with several solvers sharing the GPU (the batch driver's
    // workers) graph launches measured slower than the eager chunks
    const bool use_graphs = cache && !timing_on() && getenv("BEM_GMRES_GRAPH") && !getenv("BEM_NO_GRAPH");
    Operator op;
    op.A32 = A32;
    op.A64 = A64;
    op.lda = lda;
    op.M = M;
    op.N = N;
    op.astride = astride;
    op.nmat = std::max(1, nmat);
    if (A32 && M > N && R > 1) {
        if (ws.AH) HOST_CALL(adjoint_copy_fill(A32, lda, M, N, ws.AH, s));
        else HOST_CALL(adjoint_copy_make(A32, lda, M, N, R, &ws.AH, s));
        op.AH = ws.AH;
    }
    const int bs = 256;
    const long ldv = ws.ldv;
    // ||rhs|| per system: the residual at x = 0.  With the row-major copy, the
    // copy and A^H b come from one pass over A.
    if (ws.Ar) HOST_CALL(transpose_rm_ahb_f32(A32, lda, M, N, ws.Ar, ws.ldr, B, ws.r, ws.nws, s));
    else HOST_CALL(op.rhs(ws, B, ldb, ws.r, s));
    // host buffers: pinned when the workspace is kept, else plain vectors
    const size_t nh = (size_t)ws.hstride * R;
    std::vector<double2> hvec;
    double2 *hbuf, *yh, *hch;
    if (gc->pin) {
        hbuf = gc->pin;
    } else {
        hvec.resize(2 * nh + (size_t)hcs * R);
        hbuf = hvec.data();
    }
    yh = hbuf + nh;
    hch = yh + nh;
    auto norms = [&](const double2 *w, std::vector<double> &out) -> int {
        HOST_CALL(bnorm2(ws, w, ldv, N, ws.h, ws.hstride, s));
        HIP_CALL(hipMemcpyAsync(hbuf, ws.h, sizeof(double2) * ws.hstride * R, hipMemcpyDeviceToHost, s));
        HIP_CALL(hipStreamSynchronize(s));
        for (int q = 0; q < R; q++) out[q] = std::sqrt(std::max(0.0, hbuf[(size_t)q * ws.hstride].x));
        return EXIT_SUCCESS;
    };
    std::vector<double> nb(R), beta(R), relres(R, 0.0);
    // Tolerance of a cycle relative to its start: `tol`, or looser when the
    // cycle starts within a factor 1/tol of outer_tol; the margin 0.1 makes a
    // further cycle unlikely (the FP32 operator's error on the correction is
    // far below it)
    constexpr double kCycleMargin = 0.1;
    std::vector<double> ctol(R, tol);
    const bool fixed_ctol = getenv("BEM_FIXED_CYCLE_TOL") != nullptr;
    const bool trace = getenv("BEM_TRACE_SOLVE") != nullptr;
    HOST_CALL(norms(ws.r, nb));
    std::vector<char> done(R, 0);
    for (int q = 0; q < R; q++)
        if (nb[q] == 0) {
            done[q] = 1;
            HIP_CALL(hipMemsetAsync(X + (size_t)q * ldx, 0, sizeof(double2) * N, s));
        }
    double rate = 0.0;
    // Classical Gram-Schmidt, one pass while it is safe: the host sees every
    // step's products h and the new norm hn, and |w before|^2 = hn^2 + sum |h|^2,
    // so a step whose w lost more than 99 % of its squared norm to the
    // projection (where one pass leaves w measurably non-orthogonal) is seen;
    // from then on every step of the solve takes the second pass (CGS2,
    // "twice is enough").  BEM_CGS2 forces two passes throughout.
    bool cgs2 = getenv("BEM_CGS2") != nullptr;
    const bool full_chunks = getenv("BEM_GMRES_FULL_CHUNKS") != nullptr;
    // fused Gram-Schmidt passes while h fits in LDS
    const bool fused = !getenv("BEM_ARNOLDI_UNFUSED") && (size_t)(K + 1) * sizeof(double2) <= 48 * 1024;
    constexpr double kCancel = 1e-2;
    std::vector<double> g_start(R), g_end(R);
    std::vector<int> steps(R);
    std::vector<HessenbergLS> ls(R);
    std::vector<zc> hcol((size_t)K + 1), ysol((size_t)K);
    std::vector<char> live(R, 0);
    int total = 0, cycles = 0;
    // ws.r and beta hold the residual of the current X (and ws.h[.., 0] its
    // squared norms) while res_current is set
    bool res_current = x_zero;
    if (x_zero) beta = nb;
    for (int cyc = 0; cyc < max_cycles; cyc++) {
        if (!res_current) {
            HOST_CALL(op.residual(ws, B, ldb, X, ldx, ws.r, s));
            HOST_CALL(norms(ws.r, beta));
            res_current = true;
        }
        bool any = false;
        double worst = 0.0;
        for (int q = 0; q < R; q++) {
            if (nb[q] == 0) continue;
            relres[q] = beta[q] / nb[q];
            if (relres[q] <= outer_tol) done[q] = 1;
            any = any || !done[q];
            // a cycle that starts near outer_tol only has to close the gap
            ctol[q] = tol;
            if (outer_tol > 0 && relres[q] > 0 && !fixed_ctol)
                ctol[q] = std::max(tol, std::min(0.5, kCycleMargin * outer_tol / relres[q]));
            worst = std::max(worst, relres[q]);
        }
        if (trace)
            fprintf(stderr, "gmres: cycle %d, worst relative residual %.3e, %d steps so far\n", cycles, worst, total);
        if (!any || total >= maxit) break;
        cycles++;
        // V_r[:, 0] = r_r / beta_r (norms still in ws.h[.., 0])
        {
            dim3 gg((N + bs - 1) / bs, R);
            kb_normalize<<<gg, bs, 0, s>>>(ws.r, ldv, ws.h, ws.hstride, ws.V, ws.vstride, op.A32 ? ws.v32 : nullptr,
                                           ldv, N);
            HIP_CALL(hipGetLastError());
        }
        for (int q = 0; q < R; q++) {
            live[q] = !done[q];
            ls[q].reset(K, beta[q]);
            if (done[q] && op.A32) HIP_CALL(hipMemsetAsync(ws.v32 + (size_t)q * ldv, 0, sizeof(float2) * ldv, s));
        }
        // Arnoldi steps in chunks: the steps of a chunk are launched back to
        // back and write their Hessenberg columns to slots of hch; one copy to
        // the host then gives all of them, and the Givens rotations and the
        // convergence test run over the chunk's steps in order (steps past a
        // system's convergence are ignored).  The chunk length is the number
        // of steps the slowest live system still needs at its observed rate
        // of decrease of |g| (4 for the first chunk of the first cycle), at
        // most kMaxChunk, so few steps run past convergence.
        int j = 0;
        while (j < K && total < maxit) {
            int c = std::min({kMaxChunk, K - j, maxit - total});
            if (rate < 1.0 && rate > 0.0) {
                int need = 1;
                for (int q = 0; q < R; q++) {
                    if (!live[q]) continue;
                    const double gj = ls[q].residual();
                    if (gj <= ctol[q] * beta[q]) continue;
                    need = std::max(need, (int)std::ceil(std::log(ctol[q] * beta[q] / gj) / std::log(rate)));
                }
                // one step short of the estimate: GMRES usually converges faster
                // than its last rate says, and a further chunk costs one host
                // round trip where a step past convergence costs a product
                if (!full_chunks) need = need - 1;
                c = std::max(1, std::min(c, need));
            } else if (j == 0 && cycles == 1) {
                c = std::min(c, 4);
            }
            auto launch_chunk = [&](int j, int c) -> int {
            for (int t = 0; t < c; t++) {
                const int jj = j + t;
                HOST_CALL(op.apply(ws, jj, s));
                double2 *h1 = ws.hch + (size_t)t * hs, *h2 = ws.h + (K + 2);
                dim3 gg((N + bs - 1) / bs, R);
                if (fused) {
                    // per pass: partial dots, then sums + update + |w|^2 partials
                    // in one launch; the normalisation sums the partials
                    const int nblk = (N + 1023) / 1024;
                    const size_t lds = sizeof(double2) * (jj + 1);
                    HOST_CALL(bdot_part(ws, ws.V, ws.vstride, jj + 1, ws.w64, ldv, N, s));
                    kb_final_maxpy<<<gg, bs, lds, s>>>(ws.part, ws.pstride, nblk, ws.V, ldv, ws.vstride, jj + 1, h1,
                                                       hcs, nullptr, 0, -1.0, ws.w64, ldv, N, ws.npart, ws.nstride);
                    HIP_CALL(hipGetLastError());
                    if (cgs2) {
                        HOST_CALL(bdot_part(ws, ws.V, ws.vstride, jj + 1, ws.w64, ldv, N, s));
                        kb_final_maxpy<<<gg, bs, lds, s>>>(ws.part, ws.pstride, nblk, ws.V, ldv, ws.vstride, jj + 1,
                                                           h2, ws.hstride, h1, hcs, -1.0, ws.w64, ldv, N, ws.npart,
                                                           ws.nstride);  // h1 += h2
                        HIP_CALL(hipGetLastError());
                    }
                    kb_normalize_np<<<gg, bs, 0, s>>>(ws.w64, ldv, ws.npart, ws.nstride, (int)gg.x, h1 + (K + 1), hcs,
                                                      ws.V + (size_t)(jj + 1) * ldv, ws.vstride,
                                                      op.A32 ? ws.v32 : nullptr, ldv, N);
                    HIP_CALL(hipGetLastError());
                } else {
                HOST_CALL(bdot(ws, ws.V, ws.vstride, jj + 1, ws.w64, ldv, N, h1, hcs, s));
                kb_maxpy<<<gg, bs, 0, s>>>(ws.V, ldv, ws.vstride, jj + 1, h1, hcs, -1.0, ws.w64, ldv, N);
                HIP_CALL(hipGetLastError());
                if (cgs2) {
                    HOST_CALL(bdot(ws, ws.V, ws.vstride, jj + 1, ws.w64, ldv, N, h2, ws.hstride, s, h1, hcs));  // h1 += h2
                    kb_maxpy<<<gg, bs, 0, s>>>(ws.V, ldv, ws.vstride, jj + 1, h2, ws.hstride, -1.0, ws.w64, ldv, N);
                    HIP_CALL(hipGetLastError());
                }
                HOST_CALL(bnorm2(ws, ws.w64, ldv, N, h1 + (K + 1), hcs, s));
                kb_normalize<<<gg, bs, 0, s>>>(ws.w64, ldv, h1 + (K + 1), hcs, ws.V + (size_t)(jj + 1) * ldv,
                                               ws.vstride, op.A32 ? ws.v32 : nullptr, ldv, N);
                HIP_CALL(hipGetLastError());
                }
            }
            return EXIT_SUCCESS;
            };
            if (use_graphs) {
                const std::array<int, 3> key = {j, c, cgs2 ? 1 : 0};
                auto it = gc->graphs.find(key);
                hipGraphExec_t exec = it == gc->graphs.end() ? nullptr : it->second;
                // a chunk is captured once it has recurred: capturing and
                // instantiating costs more than launching it a few times
                if (!exec && gc->graphs.size() < kMaxGraphs && gc->uses[key]++ >= kGraphAfter) {
                    hipGraph_t graph = nullptr;
                    HIP_CALL(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                    const int rc = launch_chunk(j, c);
                    hipError_t e = hipStreamEndCapture(s, &graph);
                    HOST_CALL(rc);
                    HIP_CALL(e);
                    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(graph);
                    HIP_CALL(e);
                    gc->graphs[key] = exec;
                }
                if (exec) HIP_CALL(hipGraphLaunch(exec, s));
                else HOST_CALL(launch_chunk(j, c));
            } else {
                HOST_CALL(launch_chunk(j, c));
            }
            HIP_CALL(hipMemcpyAsync(hch, ws.hch, sizeof(double2) * hcs * R, hipMemcpyDeviceToHost, s));
            HIP_CALL(hipStreamSynchronize(s));
            for (int q = 0; q < R; q++) {
                g_start[q] = ls[q].residual();
                steps[q] = 0;
            }
            bool anylive = false;
            for (int t = 0; t < c; t++) {
                const int jj = j + t;
                total++;
                anylive = false;
                for (int q = 0; q < R; q++) {
                    if (!live[q]) continue;
                    const double2 *hq = &hch[(size_t)q * hcs + (size_t)t * hs];
                    double proj2 = 0;
                    for (int l = 0; l <= jj; l++) {
                        hcol[l] = zc(hq[l].x, hq[l].y);
                        proj2 += std::norm(hcol[l]);
                    }
                    const double hn2 = std::max(0.0, hq[K + 1].x);
                    const double hn = std::sqrt(hn2);
                    if (!cgs2 && hn2 < kCancel * (hn2 + proj2)) {
                        cgs2 = true;
                        if (trace)
                            fprintf(stderr, "gmres: step %d, system %d kept %.2e of |w|^2: second Gram-Schmidt pass from here\n",
                                    jj, q, hn2 / (hn2 + proj2));
                    }
                    // previous rotations, then a new one (GMRES.cu:58-88 / 44-56);
                    // a zero rotated column (singular H) is left out
                    if (!ls[q].add_column(jj, hcol.data(), hn)) {
                        live[q] = 0;
                        continue;
                    }
                    const double gnew = ls[q].residual();
                    steps[q]++;
                    g_end[q] = gnew;
                    if (gnew <= ctol[q] * beta[q] || hn == 0) live[q] = 0;
                    anylive = anylive || live[q];
                }
                if (!anylive) {
                    // later steps of this chunk ran past convergence
                    j = jj + 1;
                    break;
                }
            }
            // geometric-mean rate over the chunk; the slowest system sets the next length
            double rate_obs = 0.0;
            for (int q = 0; q < R; q++)
                if (steps[q] > 0 && g_start[q] > 0 && g_end[q] > 0)
                    rate_obs = std::max(rate_obs, std::pow(g_end[q] / g_start[q], 1.0 / steps[q]));
            if (rate_obs > 0) rate = rate_obs;
            if (!anylive) break;
            j += c;
        }
        // back substitution per system (Rsolver, GMRES.cu:26-42), x_r += V_r y_r
        int tmax = 0;
        for (int q = 0; q < R; q++)
            if (!done[q]) tmax = std::max(tmax, ls[q].len());
        if (tmax == 0) continue;
        std::fill(yh, yh + nh, make_double2(0, 0));
        for (int q = 0; q < R; q++) {
            const int t = ls[q].len();
            if (done[q] || t == 0) continue;
            ls[q].solve(ysol.data());
            for (int i = 0; i < t; i++)
                yh[(size_t)q * ws.hstride + i] = make_double2(ysol[i].real(), ysol[i].imag());
        }
        HIP_CALL(hipMemcpyAsync(ws.h, yh, sizeof(double2) * ws.hstride * R, hipMemcpyHostToDevice, s));
        dim3 gg((N + bs - 1) / bs, R);
        kb_maxpy<<<gg, bs, 0, s>>>(ws.V, ldv, ws.vstride, tmax, ws.h, ws.hstride, 1.0, X, ldx, N);
        HIP_CALL(hipGetLastError());
        // no wait here: yh is next written after the next chunk's synchronisation
        // (and a copy from pageable memory is staged before the call returns)
        res_current = false;
    }
    // final FP64 residuals (unless the last check already saw the current X)
    if (!res_current) {
        HOST_CALL(op.residual(ws, B, ldb, X, ldx, ws.r, s));
        HOST_CALL(norms(ws.r, beta));
    }
    bool conv = true;
    double worst = 0;
    for (int q = 0; q < R; q++) {
        const double rr = nb[q] > 0 ? beta[q] / nb[q] : 0.0;
        worst = std::max(worst, rr);
        conv = conv && rr <= outer_tol;
    }
    if (st) {
        st->iterations = total;
        st->refinements = cycles;
        st->converged = conv ? 1 : 0;
        st->residual = worst;
    }
    return EXIT_SUCCESS;
}

}  // namespace bem
