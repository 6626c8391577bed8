// Conjugate gradients on the normal equations of the overdetermined CHIEF
// least-squares systems, many right-hand sides in lockstep.
//
// min ||b - A x|| over x (M > N) is the Hermitian positive definite system
// A^H A x = A^H b.  Conjugate gradients on it (Hestenes & Stiefel 1952,
// section 10: CG applied to the normal equations) with the recurrence
//     q = A p,  alpha = gamma / |q|^2,  x += alpha p,  s -= alpha A^H q,
//     gamma' = |s|^2,  p = s + (gamma' / gamma) p
// costs one product with A and one with A^H per step and keeps four vectors
// per system: no Krylov basis, no restarts, no orthogonalisation.  The
// normal-equation residual s is updated through A^H q rather than recomputed
// as A^H r: at a least-squares solution r = b - A x stays finite while A^H r
// vanishes, so an FP32 product A^H r would be swamped by the rounding of r;
// A^H q only ever sees the correction.
//
// The products are the MFMA multi-RHS kernels for an FP32 A (the GEMV for one
// right-hand side), the adjoint one through a conjugate-transposed copy of A
// made at the start of the solve when the HBM allows; the recurrence vectors and scalars are FP64 and stay on
// the device (alpha and beta are formed by the update kernels from device
// scalars), so an iteration needs no host round trip; the host reads the
// per-system activity flags every 8 steps, and those 8 steps are one captured
// HIP graph, replayed for every chunk of the solve.
//
// Accuracy comes from FP64 residual refinement, as for GMRES: each cycle
// computes r = b - A x and s = A^H r with FP64 accumulation, runs CG on the
// correction until |s| fell by `tol` (or by only as much as reaching
// 0.1 outer_tol takes, when that is less), and adds it to x; the solve stops
// when every system's |A^H (b - A x)| <= outer_tol |A^H b|.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "bem_internal.hpp"

namespace bem {

int gemv_n_f32a(const float2 *A, long lda, long M, long N, const double2 *x, long ldx, double2 *y, long ldy, int R,
                void *ws, hipStream_t s);
int resid_n_f64(const double2 *A, long lda, long M, long N, const double2 *x, long ldx, const double2 *b, long ldb,
                double2 *r, long ldr, int R, void *ws, hipStream_t s);

// out[r] = sum_i |v_r[i]|^2 (FP64, fixed-order tree); with act: the system
// stays active while out[r] > thr[r]
template <typename C2>
__global__ __launch_bounds__(1024) void kc_norm2(const C2 *__restrict__ v, long ld, int n, double *__restrict__ out,
                                                 const double *__restrict__ thr, int *__restrict__ act) {
    __shared__ double red[1024];
    const int r = blockIdx.x;
    const C2 *vr = v + (size_t)r * ld;
    double sum = 0;
    for (int i = threadIdx.x; i < n; i += 1024) {
        const double x = (double)vr[i].x, y = (double)vr[i].y;
        sum = fma(x, x, fma(y, y, sum));
    }
    red[threadIdx.x] = sum;
    __syncthreads();
    for (int w = 512; w >= 1; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[r] = red[0];
        if (act) act[r] = act[r] && red[0] > thr[r];
    }
}

// alpha = gamma / delta: x += alpha p, s -= alpha w (w = A^H q)
template <typename W2>
__global__ void kc_upd_xs(int N, const double *__restrict__ gam, const double *__restrict__ del,
                          const int *__restrict__ act, const double2 *__restrict__ p, double2 *__restrict__ x,
                          const W2 *__restrict__ w, double2 *__restrict__ sv, long ld) {
    const int r = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double d = del[r];
    const double a = (act[r] && d > 0) ? gam[r] / d : 0.0;
    const size_t o = (size_t)r * ld + i;
    const double2 pv = p[o];
    double2 xv = x[o];
    xv.x = fma(a, pv.x, xv.x);
    xv.y = fma(a, pv.y, xv.y);
    x[o] = xv;
    const W2 wv = w[o];
    double2 ss = sv[o];
    ss.x = fma(-a, (double)wv.x, ss.x);
    ss.y = fma(-a, (double)wv.y, ss.y);
    sv[o] = ss;
}

// beta = gamma' / gamma: p = s + beta p, p32 = p
template <typename S2>
__global__ void kc_upd_p(int N, const double *__restrict__ gnew, const double *__restrict__ gold,
                         const int *__restrict__ act, const S2 *__restrict__ sv, double2 *__restrict__ p,
                         float2 *__restrict__ p32, long ld) {
    const int r = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double go = gold[r];
    const double b = (act[r] && go > 0) ? gnew[r] / go : 0.0;
    const size_t o = (size_t)r * ld + i;
    const S2 s = sv[o];
    double2 pv = p[o];
    pv.x = fma(b, pv.x, (double)s.x);
    pv.y = fma(b, pv.y, (double)s.y);
    p[o] = pv;
    if (p32) p32[o] = make_float2((float)pv.x, (float)pv.y);
}

// p = s, p32 = s, d = 0
__global__ void kc_start(int N, const double2 *__restrict__ s64, double2 *__restrict__ p, float2 *__restrict__ p32,
                         double2 *__restrict__ d, long ld) {
    const int r = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const size_t o = (size_t)r * ld + i;
    const double2 sv = s64[o];
    p[o] = sv;
    if (p32) p32[o] = make_float2((float)sv.x, (float)sv.y);
    d[o] = make_double2(0, 0);
}

__global__ void kc_add(int N, const double2 *__restrict__ d, long ldd, double2 *__restrict__ X, long ldx) {
    const int r = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double2 v = d[(size_t)r * ldd + i];
    double2 x = X[(size_t)r * ldx + i];
    x.x += v.x;
    x.y += v.y;
    X[(size_t)r * ldx + i] = x;
}

__global__ void kc_act_init(int R, const double *__restrict__ g0, const int *__restrict__ done, int *__restrict__ act) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R) act[r] = !done[r] && g0[r] > 0;
}

namespace {

constexpr int kChunk = 8;  // CG steps between reads of the activity flags

struct CglsWS {
    long ld = 0;
    hipStream_t st = nullptr;
    double2 *rr = nullptr, *s64 = nullptr, *d = nullptr, *p = nullptr, *q64 = nullptr, *w64 = nullptr;
    float2 *p32 = nullptr, *q32 = nullptr, *w32 = nullptr;
    float2 *AH = nullptr;  // conjugate-transposed copy of an FP32 A (null: adjoint products read A)
    double *gam = nullptr, *del = nullptr, *g0 = nullptr, *thr = nullptr;
    int *act = nullptr, *done = nullptr;
    char *gws = nullptr;
    ~CglsWS() {
        sfree(rr, st); sfree(s64, st); sfree(d, st); sfree(p, st); sfree(q64, st); sfree(w64, st);
        sfree(p32, st); sfree(q32, st); sfree(w32, st); sfree(AH, st);
        sfree(gam, st); sfree(del, st); sfree(g0, st); sfree(thr, st); sfree(act, st); sfree(done, st); sfree(gws, st);
    }
};

}  // namespace

int cgls_device(const float2 *A32, const double2 *A64, long lda, int M, int N, const double2 *B, long ldb,
                double2 *X, long ldx, int R, double tol, int maxit, double outer_tol, int max_cycles,
                bem_solve_stats *st, hipStream_t s) {
    if (M < N) BEM_FAIL("least squares needs M >= N (got %d x %d)", M, N);
    if (R <= 0) return EXIT_SUCCESS;
    const bool f32 = A32 != nullptr;
    CglsWS w;
    w.st = s;
    const long ld = ((long)std::max(M, N) + 63) / 64 * 64;
    w.ld = ld;
    const size_t nv = (size_t)ld * R;
    HOST_CALL(salloc(&w.rr, nv, s));
    HOST_CALL(salloc(&w.s64, nv, s));
    HOST_CALL(salloc(&w.d, nv, s));
    HOST_CALL(salloc(&w.p, nv, s));
    if (f32) {
        HOST_CALL(salloc(&w.p32, nv, s));
        HOST_CALL(salloc(&w.q32, nv, s));
        HOST_CALL(salloc(&w.w32, nv, s));
    } else {
        HOST_CALL(salloc(&w.q64, nv, s));
        HOST_CALL(salloc(&w.w64, nv, s));
    }
    HOST_CALL(salloc(&w.gam, 2 * (size_t)R, s));
    HOST_CALL(salloc(&w.del, R, s));
    HOST_CALL(salloc(&w.g0, R, s));
    HOST_CALL(salloc(&w.thr, R, s));
    HOST_CALL(salloc(&w.act, R, s));
    HOST_CALL(salloc(&w.done, R, s));
    HOST_CALL(salloc(&w.gws, std::max(std::max(ws_gemv_n(M, N, 1), ws_gemm(M, N, R)), ws_mres(M, N, R)), s));
    if (f32 && R > 1) HOST_CALL(adjoint_copy_make(A32, lda, M, N, R, &w.AH, s));
    const int bs = 256;
    const dim3 gN((N + bs - 1) / bs, R);

    // FP64-accumulated residual rr = b - A X and s64 = A^H rr
    auto outer_residual = [&](bool x_zero) -> int {
        if (x_zero) {
            HIP_CALL(hipMemcpy2DAsync(w.rr, sizeof(double2) * ld, B, sizeof(double2) * ldb, sizeof(double2) * M, R,
                                      hipMemcpyDeviceToDevice, s));
        } else if (f32) {
            if (R == 1) HOST_CALL(resid_n_f32a(A32, lda, M, N, X, 0, B, 0, w.rr, 0, 1, w.gws, s));
            else HOST_CALL(mres_f32a(A32, lda, M, N, X, ldx, B, ldb, w.rr, ld, R, w.gws, s));
        } else {
            if (R == 1) HOST_CALL(resid_n_f64(A64, lda, M, N, X, 0, B, 0, w.rr, 0, 1, w.gws, s));
            else HOST_CALL(mres_f64(A64, lda, M, N, X, ldx, B, ldb, w.rr, ld, R, w.gws, s));
        }
        if (f32) {
            if (R == 1) HOST_CALL(gemv_h_f32a(A32, lda, M, N, w.rr, 0, w.s64, 0, 1, s));
            else HOST_CALL(mh_f32a(A32, lda, M, N, w.rr, ld, w.s64, ld, R, w.gws, s));
        } else {
            if (R == 1) HOST_CALL(gemv_h_f64(A64, lda, M, N, w.rr, 0, w.s64, 0, 1, s));
            else HOST_CALL(mh_f64(A64, lda, M, N, w.rr, ld, w.s64, ld, R, s));
        }
        return EXIT_SUCCESS;
    };
    std::vector<double> hg(R);
    auto norms = [&](std::vector<double> &out) -> int {
        kc_norm2<double2><<<R, 1024, 0, s>>>(w.s64, ld, N, w.g0, nullptr, nullptr);
        HIP_CALL(hipGetLastError());
        HIP_CALL(hipMemcpyAsync(hg.data(), w.g0, sizeof(double) * R, hipMemcpyDeviceToHost, s));
        HIP_CALL(hipStreamSynchronize(s));
        for (int q = 0; q < R; q++) out[q] = std::sqrt(std::max(0.0, hg[q]));
        return EXIT_SUCCESS;
    };
    // one operator step of every system: q = A p, s = A^H r
    auto prod_n = [&]() -> int {
        if (f32) {
            if (R == 1) HOST_CALL(gemv_n_f32(A32, lda, M, N, w.p32, 0, w.q32, 0, 1, w.gws, s));
            else HOST_CALL(gemm_n_f32(A32, lda, M, N, w.p32, ld, w.q32, ld, R, w.gws, s, true));
        } else {
            for (int q = 0; q < R; q++)
                HOST_CALL(gemv_n_f64(A64, lda, M, N, w.p + (size_t)q * ld, 0, w.q64 + (size_t)q * ld, 0, 1, w.gws,
                                     s));
        }
        return EXIT_SUCCESS;
    };
    // w = A^H q
    auto prod_h = [&]() -> int {
        if (f32) {
            if (R == 1) HOST_CALL(gemv_h_f32(A32, lda, M, N, w.q32, 0, w.w32, 0, 1, s));
            else if (w.AH) HOST_CALL(gemm_hc_f32(w.AH, M, N, w.q32, ld, w.w32, ld, R, w.gws, s));
            else HOST_CALL(gemm_h_f32(A32, lda, M, N, w.q32, ld, w.w32, ld, R, w.gws, s, true));
        } else {
            for (int q = 0; q < R; q++)
                HOST_CALL(gemv_h_f64(A64, lda, M, N, w.q64 + (size_t)q * ld, 0, w.w64 + (size_t)q * ld, 0, 1, s));
        }
        return EXIT_SUCCESS;
    };

    // n CG steps (n even, so every chunk starts on gamma slot 0)
    auto chunk = [&](int n) -> int {
        for (int it = 0; it < n; it++) {
            const int cur = it & 1;
            double *gc = w.gam + (size_t)cur * R, *gn = w.gam + (size_t)(1 - cur) * R;
            HOST_CALL(prod_n());
            if (f32) kc_norm2<float2><<<R, 1024, 0, s>>>(w.q32, ld, M, w.del, nullptr, nullptr);
            else kc_norm2<double2><<<R, 1024, 0, s>>>(w.q64, ld, M, w.del, nullptr, nullptr);
            HIP_CALL(hipGetLastError());
            HOST_CALL(prod_h());
            if (f32) kc_upd_xs<float2><<<gN, bs, 0, s>>>(N, gc, w.del, w.act, w.p, w.d, w.w32, w.s64, ld);
            else kc_upd_xs<double2><<<gN, bs, 0, s>>>(N, gc, w.del, w.act, w.p, w.d, w.w64, w.s64, ld);
            kc_norm2<double2><<<R, 1024, 0, s>>>(w.s64, ld, N, gn, w.thr, w.act);
            kc_upd_p<double2><<<gN, bs, 0, s>>>(N, gn, gc, w.act, w.s64, w.p, f32 ? w.p32 : nullptr, ld);
            HIP_CALL(hipGetLastError());
        }
        return EXIT_SUCCESS;
    };
    // Chunks of 8 steps run as one captured HIP graph (the host reads the
    // activity flags between chunks).  Not while kernel timing is on: the
    // timing records events around each launch.
    const bool use_graph = !timing_on() && !getenv("BEM_NO_GRAPH");
    hipGraphExec_t exec = nullptr;
    struct ExecGuard {
        hipGraphExec_t *e;
        ~ExecGuard() {
            if (*e) (void)hipGraphExecDestroy(*e);
        }
    } exec_guard{&exec};

    std::vector<double> nb(R), beta(R);
    std::vector<int> done(R, 0);
    // Tolerance of a cycle relative to its start: `tol`, or looser when the
    // cycle starts within a factor 1/tol of outer_tol (margin 0.1, so a
    // further cycle is unlikely); the threshold on |s|^2 goes to the device
    constexpr double kCycleMargin = 0.1;
    const bool fixed_ctol = getenv("BEM_FIXED_CYCLE_TOL") != nullptr;
    std::vector<double> hthr(R);
    HIP_CALL(hipMemset2DAsync(X, sizeof(double2) * ldx, 0, sizeof(double2) * N, R, s));
    HOST_CALL(outer_residual(true));  // from X = 0
    HOST_CALL(norms(nb));
    for (int q = 0; q < R; q++) {
        if (nb[q] == 0) done[q] = 1;
        beta[q] = nb[q];
    }
    bool res_current = true;
    int total = 0, cycles = 0;
    std::vector<int> hact(R);
    for (int cyc = 0; cyc < max_cycles; cyc++) {
        if (!res_current) {
            HOST_CALL(outer_residual(false));
            HOST_CALL(norms(beta));
            res_current = true;
        }
        if (getenv("BEM_TRACE_SOLVE"))
            fprintf(stderr, "cgls cycle %d its %d rel[0] %.3e nb[0] %.3e\n", cyc, total, beta[0] / nb[0], nb[0]);
        bool any = false;
        for (int q = 0; q < R; q++) {
            if (nb[q] > 0 && beta[q] <= outer_tol * nb[q]) done[q] = 1;
            any = any || !done[q];
            double ctol = tol;
            if (outer_tol > 0 && nb[q] > 0 && beta[q] > 0 && !fixed_ctol)
                ctol = std::max(tol, std::min(0.5, kCycleMargin * outer_tol * nb[q] / beta[q]));
            hthr[q] = ctol * ctol * hg[q];
        }
        if (!any || total >= maxit) break;
        cycles++;
        HIP_CALL(hipMemcpyAsync(w.thr, hthr.data(), sizeof(double) * R, hipMemcpyHostToDevice, s));
        // inner CG on the correction: p = s = s64, gamma = g0 = |s|^2 (still in w.g0)
        HIP_CALL(hipMemcpyAsync(w.done, done.data(), sizeof(int) * R, hipMemcpyHostToDevice, s));
        HIP_CALL(hipMemcpyAsync(w.gam, w.g0, sizeof(double) * R, hipMemcpyDeviceToDevice, s));
        kc_act_init<<<(R + 63) / 64, 64, 0, s>>>(R, w.g0, w.done, w.act);
        kc_start<<<gN, bs, 0, s>>>(N, w.s64, w.p, f32 ? w.p32 : nullptr, w.d, ld);
        HIP_CALL(hipGetLastError());
        for (;;) {
            const int n = std::min(kChunk, maxit - total);
            if (n <= 0) break;
            if (n == kChunk && use_graph) {
                if (!exec) {
                    // the chunk is the same every time (pointers fixed for the
                    // solve, scalars in device memory): capture it once
                    hipGraph_t graph = nullptr;
                    HIP_CALL(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                    const int rc = chunk(kChunk);
                    hipError_t e = hipStreamEndCapture(s, &graph);
                    HOST_CALL(rc);
                    HIP_CALL(e);
                    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(graph);
                    HIP_CALL(e);
                }
                HIP_CALL(hipGraphLaunch(exec, s));
            } else {
                HOST_CALL(chunk(n));
            }
            total += n;
            HIP_CALL(hipMemcpyAsync(hact.data(), w.act, sizeof(int) * R, hipMemcpyDeviceToHost, s));
            HIP_CALL(hipStreamSynchronize(s));
            bool live = false;
            for (int q = 0; q < R; q++) live = live || hact[q];
            if (!live) break;
        }
        kc_add<<<gN, bs, 0, s>>>(N, w.d, ld, X, ldx);
        HIP_CALL(hipGetLastError());
        res_current = false;
    }
    if (!res_current) {
        HOST_CALL(outer_residual(false));
        HOST_CALL(norms(beta));
    }
    bool conv = true;
    double worst = 0;
    for (int q = 0; q < R; q++) {
        const double rr = nb[q] > 0 ? beta[q] / nb[q] : 0.0;
        worst = std::max(worst, rr);
        conv = conv && rr <= outer_tol;
    }
    HIP_CALL(hipStreamSynchronize(s));
    if (st) {
        st->iterations = total;
        st->refinements = cycles;
        st->converged = conv ? 1 : 0;
        st->residual = worst;
    }
    return EXIT_SUCCESS;
}

}  // namespace bem
