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
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "bem_internal.hpp"

namespace bem {

int gemv_n_f32a(const float2 *A, long lda, long M, long N, const double2 *x, long ldx, double2 *y, long ldy, int R,
                void *ws, hipStream_t s);
int resid_n_f64(const double2 *A, long lda, long M, long N, const double2 *x, long ldx, const double2 *b, long ldb,
                double2 *r, long ldr, int R, void *ws, hipStream_t s);

using zc = std::complex<double>;

// ---------------------------------------------------------------- kernels ---
__global__ void k_cvt_c2z(const float2 *__restrict__ in, double2 *__restrict__ out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_double2(in[i].x, in[i].y);
}

__global__ void k_copy_z(const double2 *__restrict__ in, double2 *__restrict__ out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}

// part[blk * nv + l] = sum_{i in blk} conj(V[i, l]) w[i], l < nv; 1024 rows per block
__global__ __launch_bounds__(256) void k_mdot_part(const double2 *__restrict__ V, long ldv, int nv,
                                                   const double2 *__restrict__ w, int n, double2 *__restrict__ part) {
    __shared__ double red[2][8];
    const int base = blockIdx.x * 1024;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    double2 wv[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int i = base + threadIdx.x + 256 * u;
        wv[u] = i < n ? w[i] : make_double2(0, 0);
    }
    for (int l = 0; l < nv; l++) {
        double sr = 0, si = 0;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int i = base + threadIdx.x + 256 * u;
            if (i < n) {
                const double2 v = V[(size_t)l * ldv + i];
                sr = fma(v.x, wv[u].x, fma(v.y, wv[u].y, sr));
                si = fma(v.x, wv[u].y, fma(-v.y, wv[u].x, si));
            }
        }
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            sr += __shfl_xor(sr, m, 64);
            si += __shfl_xor(si, m, 64);
        }
        if (lane == 0) {
            red[0][wid] = sr;
            red[1][wid] = si;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double tr = 0, ti = 0;
            for (int q = 0; q < 4; q++) {
                tr += red[0][q];
                ti += red[1][q];
            }
            part[(size_t)blockIdx.x * nv + l] = make_double2(tr, ti);
        }
        __syncthreads();
    }
}

// out[l] = sum_blk part[blk * nv + l] (fixed order)
__global__ void k_mdot_final(const double2 *__restrict__ part, int nblk, int nv, double2 *__restrict__ out) {
    const int l = threadIdx.x;
    if (l >= nv) return;
    double sr = 0, si = 0;
    for (int b = 0; b < nblk; b++) {
        const double2 p = part[(size_t)b * nv + l];
        sr += p.x;
        si += p.y;
    }
    out[l] = make_double2(sr, si);
}

// w[i] += sign * sum_l V[i, l] h[l]
__global__ void k_maxpy(const double2 *__restrict__ V, long ldv, int nv, const double2 *__restrict__ h, double sign,
                        double2 *__restrict__ w, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double sr = 0, si = 0;
    for (int l = 0; l < nv; l++) {
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

// accumulate h1 += h2 (nv entries)
__global__ void k_hadd(double2 *__restrict__ h1, const double2 *__restrict__ h2, int nv) {
    const int l = threadIdx.x;
    if (l < nv) {
        h1[l].x += h2[l].x;
        h1[l].y += h2[l].y;
    }
}

// out = w / sqrt(nrm2[0].x) in FP64 and (optionally) FP32
__global__ void k_normalize(const double2 *__restrict__ w, const double2 *__restrict__ nrm2, double2 *__restrict__ out,
                            float2 *__restrict__ out32, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double s = nrm2[0].x > 0 ? 1.0 / sqrt(nrm2[0].x) : 0.0;
    const double2 v = make_double2(w[i].x * s, w[i].y * s);
    out[i] = v;
    if (out32) out32[i] = make_float2((float)v.x, (float)v.y);
}

// ---------------------------------------------------------------- driver ---
struct KrylovWS {
    long ldv = 0;
    int n = 0, m = 0, restart = 0;
    double2 *V = nullptr;
    double2 *w64 = nullptr;
    double2 *t64 = nullptr;
    float2 *v32 = nullptr;
    float2 *w32 = nullptr;
    float2 *t32 = nullptr;
    double2 *r = nullptr;
    double2 *h = nullptr;   // 2 * (restart + 2)
    double2 *part = nullptr;
    char *gws = nullptr;
    size_t gws_bytes = 0;
    int nblk = 0;
    hipStream_t st = nullptr;
    ~KrylovWS() {
        sfree(V, st); sfree(w64, st); sfree(t64, st); sfree(v32, st); sfree(w32, st); sfree(t32, st);
        sfree(r, st); sfree(h, st); sfree(part, st); sfree(gws, st);
    }
    int init(int n_, int m_, int restart_, hipStream_t s) {
        st = s;
        n = n_;
        m = m_;
        restart = restart_;
        ldv = ((long)std::max(n, m) + 63) / 64 * 64;
        nblk = (int)((std::max(n, m) + 1023) / 1024);
        HOST_CALL(salloc(&V, (size_t)ldv * (restart + 1), s));
        HOST_CALL(salloc(&w64, ldv, s));
        HOST_CALL(salloc(&t64, ldv, s));
        HOST_CALL(salloc(&v32, ldv, s));
        HOST_CALL(salloc(&w32, ldv, s));
        HOST_CALL(salloc(&t32, ldv, s));
        HOST_CALL(salloc(&r, ldv, s));
        HOST_CALL(salloc(&h, 2 * (size_t)(restart + 2), s));
        HOST_CALL(salloc(&part, (size_t)nblk * (restart + 2), s));
        gws_bytes = ws_gemv_n(m, n, 1);
        HOST_CALL(salloc(&gws, gws_bytes, s));
        return EXIT_SUCCESS;
    }
};

// dots of V[:, 0..nv-1] against w -> out (device)
static int mdot(KrylovWS &ws, const double2 *V, int nv, const double2 *w, int n, double2 *out, hipStream_t s) {
    const int nblk = (n + 1023) / 1024;
    k_mdot_part<<<nblk, 256, 0, s>>>(V, ws.ldv, nv, w, n, ws.part);
    HIP_CALL(hipGetLastError());
    k_mdot_final<<<1, 64 * ((nv + 63) / 64), 0, s>>>(ws.part, nblk, nv, out);
    HIP_CALL(hipGetLastError());
    return EXIT_SUCCESS;
}

static int norm2_host(KrylovWS &ws, const double2 *w, int n, double &nrm, hipStream_t s) {
    HOST_CALL(mdot(ws, w, 1, w, n, ws.h, s));
    double2 hv;
    HIP_CALL(hipMemcpyAsync(&hv, ws.h, sizeof(double2), hipMemcpyDeviceToHost, s));
    HIP_CALL(hipStreamSynchronize(s));
    nrm = std::sqrt(std::max(0.0, hv.x));
    return EXIT_SUCCESS;
}

// The operator of the Krylov iteration: square A, or A^H A for M > N.
struct Operator {
    const float2 *A32 = nullptr;
    const double2 *A64 = nullptr;
    long lda = 0;
    int M = 0, N = 0;
    bool normal() const { return M > N; }
    // w64 <- op(V[:, j]) ; for FP32 A the input is v32
    int apply(KrylovWS &ws, const double2 *vj, hipStream_t s) const {
        const int bs = 256;
        if (A32) {
            if (!normal()) {
                HOST_CALL(gemv_n_f32(A32, lda, M, N, ws.v32, 0, ws.w32, 0, 1, ws.gws, s));
            } else {
                HOST_CALL(gemv_n_f32(A32, lda, M, N, ws.v32, 0, ws.t32, 0, 1, ws.gws, s));
                HOST_CALL(gemv_h_f32(A32, lda, M, N, ws.t32, 0, ws.w32, 0, 1, s));
            }
            k_cvt_c2z<<<(N + bs - 1) / bs, bs, 0, s>>>(ws.w32, ws.w64, N);
            HIP_CALL(hipGetLastError());
        } else {
            if (!normal()) {
                HOST_CALL(gemv_n_f64(A64, lda, M, N, vj, 0, ws.w64, 0, 1, ws.gws, s));
            } else {
                HOST_CALL(gemv_n_f64(A64, lda, M, N, vj, 0, ws.t64, 0, 1, ws.gws, s));
                HOST_CALL(gemv_h_f64(A64, lda, M, N, ws.t64, 0, ws.w64, 0, 1, s));
            }
        }
        return EXIT_SUCCESS;
    }
    // r <- rhs_ne - op(x) in FP64: square: b - A x; normal: A^H (b - A x)
    int residual(KrylovWS &ws, const double2 *b, const double2 *x, double2 *r, hipStream_t s) const {
        if (A32) {
            if (!normal()) {
                HOST_CALL(resid_n_f32a(A32, lda, M, N, x, 0, b, 0, r, 0, 1, ws.gws, s));
            } else {
                HOST_CALL(resid_n_f32a(A32, lda, M, N, x, 0, b, 0, ws.t64, 0, 1, ws.gws, s));
                HOST_CALL(gemv_h_f32a(A32, lda, M, N, ws.t64, 0, r, 0, 1, s));
            }
        } else {
            if (!normal()) {
                HOST_CALL(resid_n_f64(A64, lda, M, N, x, 0, b, 0, r, 0, 1, ws.gws, s));
            } else {
                HOST_CALL(resid_n_f64(A64, lda, M, N, x, 0, b, 0, ws.t64, 0, 1, ws.gws, s));
                HOST_CALL(gemv_h_f64(A64, lda, M, N, ws.t64, 0, r, 0, 1, s));
            }
        }
        return EXIT_SUCCESS;
    }
};

struct GmresResult {
    int iterations = 0;
    int cycles = 0;
    bool converged = false;
    double relres = 0;
};

// Solve op(x) = rhs for one right-hand side.  b: M-vector (FP64, device), x:
// N-vector (FP64, device, initial guess on entry).  Each restart cycle runs up
// to `restart` Arnoldi steps or until the estimated residual fell by `tol`
// relative to the cycle's start; cycles repeat until the FP64 residual is below
// `outer_tol` relative to ||rhs|| (or maxit / max_cycles is reached).
static int gmres_solve(const Operator &op, KrylovWS &ws, const double2 *b, double2 *x, int restart, double tol,
                       int maxit, double outer_tol, int max_cycles, GmresResult &res, hipStream_t s) {
    const int N = op.N;
    const int bs = 256;
    const int gN = (N + bs - 1) / bs;
    // ||rhs||: rhs = b (square) or A^H b (normal equations)
    double nb;
    {
        // rhs norm via the residual at x = 0
        HIP_CALL(hipMemsetAsync(ws.w64, 0, sizeof(double2) * ws.ldv, s));
        HOST_CALL(op.residual(ws, b, ws.w64, ws.r, s));
        HOST_CALL(norm2_host(ws, ws.r, N, nb, s));
    }
    res = GmresResult();
    if (nb == 0) {
        HIP_CALL(hipMemsetAsync(x, 0, sizeof(double2) * N, s));
        res.converged = true;
        return EXIT_SUCCESS;
    }
    std::vector<zc> H((size_t)(restart + 1) * restart), cs(restart), sn(restart), g(restart + 1), y(restart);
    std::vector<double2> hcol(restart + 2);
    int total = 0;
    for (int cyc = 0; cyc < max_cycles; cyc++) {
        HOST_CALL(op.residual(ws, b, x, ws.r, s));
        double beta;
        HOST_CALL(norm2_host(ws, ws.r, N, beta, s));
        res.relres = beta / nb;
        if (beta / nb <= outer_tol || total >= maxit) {
            res.converged = beta / nb <= outer_tol;
            break;
        }
        res.cycles++;
        // V[:, 0] = r / beta
        HOST_CALL(mdot(ws, ws.r, 1, ws.r, N, ws.h, s));
        k_normalize<<<gN, bs, 0, s>>>(ws.r, ws.h, ws.V, op.A32 ? ws.v32 : nullptr, N);
        HIP_CALL(hipGetLastError());
        std::fill(g.begin(), g.end(), zc(0));
        g[0] = beta;
        int j = 0;
        for (; j < restart && total < maxit; j++) {
            total++;
            HOST_CALL(op.apply(ws, ws.V + (size_t)j * ws.ldv, s));
            // CGS2 in FP64
            double2 *h1 = ws.h, *h2 = ws.h + (restart + 2);
            HOST_CALL(mdot(ws, ws.V, j + 1, ws.w64, N, h1, s));
            k_maxpy<<<gN, bs, 0, s>>>(ws.V, ws.ldv, j + 1, h1, -1.0, ws.w64, N);
            HIP_CALL(hipGetLastError());
            HOST_CALL(mdot(ws, ws.V, j + 1, ws.w64, N, h2, s));
            k_maxpy<<<gN, bs, 0, s>>>(ws.V, ws.ldv, j + 1, h2, -1.0, ws.w64, N);
            HIP_CALL(hipGetLastError());
            k_hadd<<<1, 64 * ((j + 1 + 63) / 64), 0, s>>>(h1, h2, j + 1);
            HIP_CALL(hipGetLastError());
            HOST_CALL(mdot(ws, ws.w64, 1, ws.w64, N, h2, s));
            k_normalize<<<gN, bs, 0, s>>>(ws.w64, h2, ws.V + (size_t)(j + 1) * ws.ldv, op.A32 ? ws.v32 : nullptr, N);
            HIP_CALL(hipGetLastError());
            HIP_CALL(hipMemcpyAsync(hcol.data(), h1, sizeof(double2) * (j + 1), hipMemcpyDeviceToHost, s));
            double2 hn2;
            HIP_CALL(hipMemcpyAsync(&hn2, h2, sizeof(double2), hipMemcpyDeviceToHost, s));
            HIP_CALL(hipStreamSynchronize(s));
            zc *hc = &H[(size_t)j * (restart + 1)];
            for (int l = 0; l <= j; l++) hc[l] = zc(hcol[l].x, hcol[l].y);
            const double hn = std::sqrt(std::max(0.0, hn2.x));
            hc[j + 1] = hn;
            // previous rotations, new rotation (GMRES.cu:58-88 / 44-56)
            for (int l = 0; l < j; l++) {
                zc t0 = cs[l] * hc[l] + sn[l] * hc[l + 1];
                zc t1 = -std::conj(sn[l]) * hc[l] + std::conj(cs[l]) * hc[l + 1];
                hc[l] = t0;
                hc[l + 1] = t1;
            }
            const double mag = std::sqrt(std::norm(hc[j]) + std::norm(hc[j + 1]));
            if (mag == 0) {
                j++;
                break;
            }
            cs[j] = std::conj(hc[j]) / mag;
            sn[j] = std::conj(hc[j + 1]) / mag;
            hc[j] = cs[j] * hc[j] + sn[j] * hc[j + 1];
            hc[j + 1] = 0;
            g[j + 1] = -std::conj(sn[j]) * g[j];
            g[j] = cs[j] * g[j];
            if (std::abs(g[j + 1]) <= tol * beta || hn == 0) {
                j++;
                break;
            }
        }
        // back substitution (Rsolver, GMRES.cu:26-42), x += V y
        const int t = j;
        for (int i = t - 1; i >= 0; i--) {
            zc acc = g[i];
            for (int l = i + 1; l < t; l++) acc -= H[(size_t)l * (restart + 1) + i] * y[l];
            y[i] = acc / H[(size_t)i * (restart + 1) + i];
        }
        std::vector<double2> yh(t);
        for (int i = 0; i < t; i++) yh[i] = make_double2(y[i].real(), y[i].imag());
        HIP_CALL(hipMemcpyAsync(ws.h, yh.data(), sizeof(double2) * t, hipMemcpyHostToDevice, s));
        k_maxpy<<<gN, bs, 0, s>>>(ws.V, ws.ldv, t, ws.h, 1.0, x, N);
        HIP_CALL(hipGetLastError());
        HIP_CALL(hipStreamSynchronize(s));
    }
    res.iterations = total;
    if (!res.converged) {
        HOST_CALL(op.residual(ws, b, x, ws.r, s));
        double beta;
        HOST_CALL(norm2_host(ws, ws.r, N, beta, s));
        res.relres = beta / nb;
        res.converged = res.relres <= outer_tol;
    }
    return EXIT_SUCCESS;
}

// Public (internal) entry: solve for R right-hand sides with FP32 or FP64 A.
// B: M x R (ldb) FP64 on the device, X: N x R (ldx) FP64 on the device (in/out).
int gmres_device(const float2 *A32, const double2 *A64, long lda, int M, int N, const double2 *B, long ldb,
                 double2 *X, long ldx, int R, int restart, double tol, int maxit, double outer_tol, int max_cycles,
                 bem_solve_stats *st, hipStream_t s) {
    if (M < N) BEM_FAIL("GMRES needs M >= N (got %d x %d)", M, N);
    restart = std::max(1, std::min(restart, N));
    KrylovWS ws;
    HOST_CALL(ws.init(N, M, restart, s));
    Operator op;
    op.A32 = A32;
    op.A64 = A64;
    op.lda = lda;
    op.M = M;
    op.N = N;
    int iters = 0, cycles = 0;
    bool conv = true;
    double worst = 0;
    for (int r = 0; r < R; r++) {
        GmresResult res;
        HOST_CALL(gmres_solve(op, ws, B + (size_t)r * ldb, X + (size_t)r * ldx, restart, tol, maxit, outer_tol,
                              max_cycles, res, s));
        iters += res.iterations;
        cycles += res.cycles;
        conv = conv && res.converged;
        worst = std::max(worst, res.relres);
    }
    if (st) {
        st->iterations = iters;
        st->refinements = cycles;
        st->converged = conv ? 1 : 0;
        st->residual = worst;
    }
    return EXIT_SUCCESS;
}

}  // namespace bem
