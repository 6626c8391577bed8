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
// Block variant (block = true, several right-hand sides): block conjugate
// gradients (O'Leary 1980) on the same normal equations, with the residual
// block kept orthonormal (Dubrulle 2001, "BCGrQ"):
//     S_0 = Q_0 xi_0 (QR),  P_0 = Q_0;  per step
//     W = A^H (A P),  alpha = (P^H W)^-1,  X += P alpha xi,
//     Q' zeta = Q - W alpha (QR),  xi <- zeta xi,  P <- Q' + P zeta^H,
// so every system searches the span of all systems' Krylov vectors and the
// residual of system c is |xi[:, c]|.  The two big products are the same MFMA
// kernels as above, for the whole block at once; the block inner products
// (R x R Gram matrices, FP64, fixed order) and the tall-skinny products with
// R x R matrices run on the device, the R x R algebra (Cholesky factors,
// triangular inverses) on the host.  QR is CholeskyQR2 in FP64.  Right-hand
// sides whose fields share spectral content converge together: on the s5
// analogue of the 32-point-source large case, 22 block steps reach 1e-6 where
// CG needs 196.  A breakdown (a Gram matrix that is not numerically positive
// definite, e.g. two equal right-hand sides) ends the block cycle; a cycle
// that cannot start, or breaks down at once, hands the solve to the lockstep
// CG above.
//
// Accuracy comes from FP64 residual refinement, as for GMRES: each cycle
// computes r = b - A x and s = A^H r with FP64 accumulation, runs CG on the
// correction until |s| fell by `tol` (or by only as much as reaching
// 0.1 outer_tol takes, when that is less), and adds it to x; the solve stops
// when every system's |A^H (b - A x)| <= outer_tol |A^H b|.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <initializer_list>
#include <mutex>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "bem_internal.hpp"
#include "small_dense.hpp"
#include "xcd.hpp"

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

// out (FP32) = in (FP64), n rows of each column r = blockIdx.y
__global__ void kc_to_f32(int n, const double2 *__restrict__ in, long ldi, float2 *__restrict__ out, long ldo) {
    const int r = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 v = in[(size_t)r * ldi + i];
    out[(size_t)r * ldo + i] = make_float2((float)v.x, (float)v.y);
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

// ------------------------------------------------------------ block CG ---
constexpr int kMaxBlock = 64;    // right-hand sides of one block (Rb <= 64)
constexpr double kBlockMinWork = 4e8;  // M N R below which block CG takes the lockstep iteration
constexpr int kGramRows = 64;    // rows of one Gram partial
constexpr int kGramSub = 32;     // rows staged in LDS at a time
constexpr int kGramSegs = 16;    // first-level groups of partials in the sum
// output columns per thread of the block product: at N = 40962, Rb = 32,
// 4 columns 25.0-25.4 us, 8 columns 30.6 us, 2 or 16 columns 31-43 us
// (64-, 128- and 256-thread workgroups alike)
constexpr int kMulCols = 4;

// Up to four Gram products U_y^H V_y of one launch (grid.y = y)
constexpr int kMaxPairs = 4;
struct GramPairs {
    const double2 *U[kMaxPairs];
    const double2 *V[kMaxPairs];
};

// part[y][blk][a + Rb b] = sum over the block's rows i of conj(U_y[i, a])
// V_y[i, b] (U_y, V_y: n x Rb, column-major, ld); FP64, each entry summed in
// row order.  Thread t owns the entries e = t + 256 u (a = e % Rb, b = e / Rb).
__global__ __launch_bounds__(256) void kc_gram_part(GramPairs gp, long ld, int n, int Rb, double2 *__restrict__ part) {
    __shared__ double2 Us[kGramSub * kMaxBlock], Vs[kGramSub * kMaxBlock];  // [row][column]
    const int RR = Rb * Rb;
    const int i0 = blockIdx.x * kGramRows;
    const double2 *__restrict__ U = gp.U[blockIdx.y];
    const double2 *__restrict__ V = gp.V[blockIdx.y];
    part += (size_t)blockIdx.y * gridDim.x * RR;
    constexpr int kPer = kMaxBlock * kMaxBlock / 256;
    double2 acc[kPer];
    int ia[kPer], ib[kPer];
#pragma unroll
    for (int u = 0; u < kPer; u++) {
        acc[u] = make_double2(0, 0);
        const int e = min((int)threadIdx.x + 256 * u, RR - 1);
        ia[u] = e % Rb;
        ib[u] = e / Rb;
    }
    for (int sub = 0; sub < kGramRows; sub += kGramSub) {
        for (int e = threadIdx.x; e < kGramSub * Rb; e += 256) {
            const int r = e % kGramSub, c = e / kGramSub;
            const int i = i0 + sub + r;
            double2 u = make_double2(0, 0), v = make_double2(0, 0);
            if (i < n) {
                u = U[(size_t)c * ld + i];
                v = V[(size_t)c * ld + i];
            }
            Us[r * Rb + c] = u;
            Vs[r * Rb + c] = v;
        }
        __syncthreads();
        for (int r = 0; r < kGramSub; r++) {
#pragma unroll
            for (int u = 0; u < kPer; u++) {
                if ((int)threadIdx.x + 256 * u < RR) {
                    const double2 x = Us[r * Rb + ia[u]], y = Vs[r * Rb + ib[u]];
                    acc[u].x = fma(x.x, y.x, fma(x.y, y.y, acc[u].x));
                    acc[u].y = fma(x.x, y.y, fma(-x.y, y.x, acc[u].y));
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < kPer; u++) {
        const int e = threadIdx.x + 256 * u;
        if (e < RR) part[(size_t)blockIdx.x * RR + e] = acc[u];
    }
}

// The same partials on MFMA: part[y][blk][a + Rb b] over kGramRowsM rows per
// workgroup, FP64 v_mfma_f64_16x16x4f64 with the operands read straight from
// the block vectors.  The Rb x Rb result is cut into 16 x 16 tiles; wave w
// takes tiles w, w + 4, .. (TPW of them: 1 while Rb <= 32, else 4).  Lane
// (m, kk) supplies A[m][kk] = U[r + kk, 16 ta + m] and B[kk][m] =
// V[r + kk, 16 tb + m] for rows r, r + 4, ..; conj(u) v in the three-product
// form, with u' = conj(u):
//   P1 = Re u' (Re v + Im v), P2 = (Re u' + Im u') Im v, P3 = (Im u' - Re u') Re v,
//   Re = P1 - P2, Im = P1 + P3.
// Result layout: D[l / 16 + 4 v][l % 16] in lane l, register v.  Rows past n
// and columns past Rb read clamped addresses and are zeroed.
constexpr int kGramRowsM = 128;
template <int TPW, int KB>
__global__ __launch_bounds__(256) void kc_gram_mfma(GramPairs gp, long ld, int n, int Rb, double2 *__restrict__ part) {
    typedef double v4d __attribute__((ext_vector_type(4)));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, kk = lane >> 4;
    const int nt = (Rb + 15) / 16, T = nt * nt, RR = Rb * Rb;
    const double2 *__restrict__ U = gp.U[blockIdx.y];
    const double2 *__restrict__ V = gp.V[blockIdx.y];
    part += ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * RR;
    const int i0 = blockIdx.x * kGramRowsM;
    v4d acc[TPW][3];
    const double2 *pu[TPW], *pv[TPW];
    bool ua[TPW], vb[TPW];
#pragma unroll
    for (int q = 0; q < TPW; q++) {
#pragma unroll
        for (int c = 0; c < 3; c++) acc[q][c] = (v4d){0, 0, 0, 0};
        const int t = min(wave + 4 * q, T - 1);
        const int ca = 16 * (t % nt) + m, cb = 16 * (t / nt) + m;
        ua[q] = ca < Rb;
        vb[q] = cb < Rb;
        pu[q] = U + (size_t)min(ca, Rb - 1) * ld;
        pv[q] = V + (size_t)min(cb, Rb - 1) * ld;
    }
    for (int r = i0; r < i0 + kGramRowsM; r += 4 * KB) {
        double2 u[KB][TPW], v[KB][TPW];
#pragma unroll
        for (int b = 0; b < KB; b++) {
            const int i = r + 4 * b + kk, ic = min(i, n - 1);
#pragma unroll
            for (int q = 0; q < TPW; q++) {
                u[b][q] = pu[q][ic];
                v[b][q] = pv[q][ic];
                if (i >= n || !ua[q]) u[b][q] = make_double2(0.0, 0.0);
                if (i >= n || !vb[q]) v[b][q] = make_double2(0.0, 0.0);
            }
        }
#pragma unroll
        for (int b = 0; b < KB; b++)
#pragma unroll
            for (int q = 0; q < TPW; q++) {
                const double ur = u[b][q].x, ui = -u[b][q].y, vr = v[b][q].x, vi = v[b][q].y;
                acc[q][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(ur, vr + vi, acc[q][0], 0, 0, 0);
                acc[q][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(ur + ui, vi, acc[q][1], 0, 0, 0);
                acc[q][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(ui - ur, vr, acc[q][2], 0, 0, 0);
            }
    }
#pragma unroll
    for (int q = 0; q < TPW; q++) {
        const int t = wave + 4 * q;
        if (t >= T) continue;
        const int a0 = 16 * (t % nt), b = 16 * (t / nt) + m;
        if (b >= Rb) continue;
#pragma unroll
        for (int vv = 0; vv < 4; vv++) {
            const int a = a0 + kk + 4 * vv;
            if (a < Rb)
                part[a + (size_t)Rb * b] =
                    make_double2(acc[q][0][vv] - acc[q][1][vv], acc[q][0][vv] + acc[q][2][vv]);
        }
    }
}

// seg[y][g][e] = sum of part[y][blk][e] over the g-th run of blocks, in order
__global__ void kc_gram_sum1(const double2 *__restrict__ part, int nblk, int RR, double2 *__restrict__ seg) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x, g = blockIdx.y;
    if (e >= RR) return;
    part += (size_t)blockIdx.z * nblk * RR;
    seg += (size_t)blockIdx.z * kGramSegs * RR;
    const int per = (nblk + kGramSegs - 1) / kGramSegs;
    const int b1 = min(nblk, (g + 1) * per);
    double2 sacc = make_double2(0, 0);
    for (int b = g * per; b < b1; b++) {
        const double2 v = part[(size_t)b * RR + e];
        sacc.x += v.x;
        sacc.y += v.y;
    }
    seg[(size_t)g * RR + e] = sacc;
}
// G[y][e] = sum_g seg[y][g][e], in order
__global__ void kc_gram_sum2(const double2 *__restrict__ seg, int RR, double2 *__restrict__ G) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= RR) return;
    seg += (size_t)blockIdx.y * kGramSegs * RR;
    G += (size_t)blockIdx.y * RR;
    double2 sacc = make_double2(0, 0);
#pragma unroll
    for (int g = 0; g < kGramSegs; g++) {
        const double2 v = seg[(size_t)g * RR + e];
        sacc.x += v.x;
        sacc.y += v.y;
    }
    G[e] = sacc;
}

// out[:, b] = Z[:, b] (or 0) + sum_a U[:, a] C[a, b] for an Rb x Rb matrix C
// (column-major); out32, when given, gets the FP32 copy.  out may be Z but
// not U.
// One-dimensional grid of (n / 256 row blocks) x (column groups of kMulCols):
// the column groups of a row block are consecutive logical ids, which
// xcd_remap puts on one XCD in dispatch order, so the groups after the first
// read U from that XCD's L2.
__global__ __launch_bounds__(256) void kc_block_mul(const double2 *__restrict__ U, const double2 *__restrict__ C,
                                                    int Rb, const double2 *Z, double2 *out, float2 *__restrict__ out32,
                                                    long ld, int n) {
    __shared__ double2 Cs[kMaxBlock * kMulCols];  // [column - b0][a]
    const int ncg = (Rb + kMulCols - 1) / kMulCols;
    const int L = xcd_remap(blockIdx.x, gridDim.x);
    const int rb = L / ncg, b0 = (L - rb * ncg) * kMulCols;
    const int nb = min(kMulCols, Rb - b0);
    for (int e = threadIdx.x; e < Rb * nb; e += 256) Cs[e] = C[(size_t)(b0 + e / Rb) * Rb + e % Rb];
    __syncthreads();
    const int i = rb * 256 + threadIdx.x;
    if (i >= n) return;
    double2 acc[kMulCols];
#pragma unroll
    for (int bb = 0; bb < kMulCols; bb++)
        acc[bb] = (Z && bb < nb) ? Z[(size_t)(b0 + bb) * ld + i] : make_double2(0, 0);
    // U's columns in batches of 8 loads in flight (22.7 us against 25.4 with
    // one at a time at N = 40962, Rb = 32; the same sums in the same order)
    constexpr int kMulBatch = 8;
    auto madd = [&](const double2 u, int a) {
#pragma unroll
        for (int bb = 0; bb < kMulCols; bb++) {
            if (bb < nb) {
                const double2 c = Cs[bb * Rb + a];
                acc[bb].x = fma(u.x, c.x, fma(-u.y, c.y, acc[bb].x));
                acc[bb].y = fma(u.x, c.y, fma(u.y, c.x, acc[bb].y));
            }
        }
    };
    int a0 = 0;
    for (; a0 + kMulBatch <= Rb; a0 += kMulBatch) {
        double2 uu[kMulBatch];
#pragma unroll
        for (int q = 0; q < kMulBatch; q++) uu[q] = U[(size_t)(a0 + q) * ld + i];
#pragma unroll
        for (int q = 0; q < kMulBatch; q++) madd(uu[q], a0 + q);
    }
    for (int a = a0; a < Rb; a++) madd(U[(size_t)a * ld + i], a);
#pragma unroll
    for (int bb = 0; bb < kMulCols; bb++) {
        if (bb < nb) {
            const size_t o = (size_t)(b0 + bb) * ld + i;
            out[o] = acc[bb];
            if (out32) out32[o] = make_float2((float)acc[bb].x, (float)acc[bb].y);
        }
    }
}

// S[:, c] = s[:, idx[c]];  X[:, idx[c]] += D[:, c]
__global__ void kc_gather(int N, const double2 *__restrict__ s, const int *__restrict__ idx, double2 *__restrict__ S,
                          long ld) {
    const int c = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) S[(size_t)c * ld + i] = s[(size_t)idx[c] * ld + i];
}
__global__ void kc_scatter_add(int N, const double2 *__restrict__ D, long ldd, const int *__restrict__ idx,
                               double2 *__restrict__ X, long ldx) {
    const int c = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double2 v = D[(size_t)c * ldd + i];
    double2 &x = X[(size_t)idx[c] * ldx + i];
    x.x += v.x;
    x.y += v.y;
}

namespace {

constexpr int kChunk = 8;  // CG steps between reads of the activity flags
using small::cd;
using small::Small;
using small::chol_lower;
using small::herm;
using small::lower_inv;
using small::mul;

// Pinned host blocks of the block CG's R x R transfers, reused across solves
// and threads: hipHostFree waits for the whole device, which would stall the
// other solvers of a batch or server at every solve's end.  The blocks live
// until the process ends (one per solve running at once).
class PinnedPool {
public:
    static constexpr size_t kBytes = sizeof(double2) * (4 + 4 + 6) * 64 * 64;
    static int acquire(double2 **p) {
        {
            std::lock_guard<std::mutex> lk(mu());
            if (!free_list().empty()) {
                *p = free_list().back();
                free_list().pop_back();
                return EXIT_SUCCESS;
            }
        }
        HIP_CALL(hipHostMalloc((void **)p, kBytes, hipHostMallocDefault));
        return EXIT_SUCCESS;
    }
    static void release(double2 *p) {
        if (!p) return;
        std::lock_guard<std::mutex> lk(mu());
        free_list().push_back(p);
    }

private:
    static std::mutex &mu() {
        static std::mutex m;
        return m;
    }
    static std::vector<double2 *> &free_list() {
        static std::vector<double2 *> v;
        return v;
    }
};

struct CglsWS {
    long ld = 0;
    hipStream_t st = nullptr;
    double2 *rr = nullptr, *s64 = nullptr, *d = nullptr, *p = nullptr, *q64 = nullptr, *w64 = nullptr;
    float2 *p32 = nullptr, *q32 = nullptr, *w32 = nullptr;
    float2 *AH = nullptr;  // conjugate-transposed copy of an FP32 A (null: adjoint products read A)
    char *xsplit = nullptr;   // FP16-split products: planes of the right-hand sides
    float *amaxws = nullptr;  // and the bound on |A|
    double *gam = nullptr, *del = nullptr, *g0 = nullptr, *thr = nullptr;
    int *act = nullptr, *done = nullptr;
    char *gws = nullptr;
    // block CG: Q, S (QR input / scratch), P, P (next), W = A^H A P, the
    // correction; FP32 P; Gram partials and sums; R x R matrices (pinned
    // staging slots and their device copies); active column indices
    double2 *bQ = nullptr, *bS = nullptr, *bP = nullptr, *bP2 = nullptr, *bW = nullptr, *bX = nullptr;
    float2 *bP32 = nullptr;
    double2 *gpart = nullptr, *gseg = nullptr, *gsum = nullptr, *dcm = nullptr;
    double2 *hgram = nullptr, *hcm = nullptr;  // one PinnedPool block: Gram sums, then R x R uploads
    int *bidx = nullptr;
    ~CglsWS() {
        sfree(rr, st); sfree(s64, st); sfree(d, st); sfree(p, st); sfree(q64, st); sfree(w64, st);
        sfree(p32, st); sfree(q32, st); sfree(w32, st); sfree(AH, st); sfree(xsplit, st); sfree(amaxws, st);
        sfree(gam, st); sfree(del, st); sfree(g0, st); sfree(thr, st); sfree(act, st); sfree(done, st); sfree(gws, st);
        sfree(bQ, st); sfree(bS, st); sfree(bP, st); sfree(bP2, st); sfree(bW, st); sfree(bX, st); sfree(bP32, st);
        sfree(gpart, st); sfree(gseg, st); sfree(gsum, st); sfree(dcm, st); sfree(bidx, st);
        if (hgram) {
            (void)hipStreamSynchronize(st);  // the pinned slots may feed queued copies
            PinnedPool::release(hgram);
        }
    }
};
constexpr int kCSlots = 4;  // staging slots of R x R matrices (at most 2 uploads between two host waits)
// then the matrices of one block step: area 0 (4, after the step's Gram
// products) and area 1 (2, after a second QR pass); each area is rewritten
// only after a host wait that follows its last upload
constexpr int kStepMats = 6;

}  // namespace

int cgls_device(const float2 *A32, const double2 *A64, long lda, int M, int N, const double2 *B, long ldb,
                double2 *X, long ldx, int R, double tol, int maxit, double outer_tol, int max_cycles,
                bem_solve_stats *st, hipStream_t s, bool block) {
    if (M < N) BEM_FAIL("least squares needs M >= N (got %d x %d)", M, N);
    if (R <= 0) return EXIT_SUCCESS;
    if (block && R > kMaxBlock) BEM_FAIL("block CG takes at most %d right-hand sides (got %d)", kMaxBlock, R);
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
    // Block steps pay three dozen small launches, a host wait and the R x R
    // algebra on the host (about 0.1 ms at R = 16, 0.3 ms at R = 32); below
    // about 4e8 of M N R a lockstep CG step is that cheap that the block's
    // 2x fewer steps do not repay it (s4 with 16 right-hand sides: 7.9 ms
    // against 4.8 ms; s5 with 32: 33.5 ms against 51.3 ms).  BEM_BLOCK_CG_ALWAYS
    // takes the block at every size.
    bool block_on = block && R > 1 &&
                    ((double)M * (double)N * (double)R >= kBlockMinWork || getenv("BEM_BLOCK_CG_ALWAYS"));
    // FP16-split block products: A and the block vectors read as FP16 hi/lo
    // pairs on FP16 MFMA; the adjoint product reads the conjugate-transposed
    // copy, which also leaves the bound amax on |Re A|, |Im A| (per-tile
    // maxima).  At M = N = 40962 with 32 right-hand sides: 2.6 ms a product
    // against 3.37 ms on the FP32 kernels, the same error against the FP64
    // solve.  BEM_GEMM_NO_F16 keeps the FP32 products.
    const bool want_f16s = block_on && f32 && R > 1 && getenv("BEM_GEMM_NO_F16") == nullptr;
    if (want_f16s) HOST_CALL(salloc(&w.amaxws, ws_adjoint_absmax(M, N) / sizeof(float), s));
    if (f32 && R > 1) HOST_CALL(adjoint_copy_make(A32, lda, M, N, R, &w.AH, s, w.amaxws, want_f16s));
    const int gblk = (N + kGramRows - 1) / kGramRows;
    if (block_on) {
        for (double2 **b : {&w.bQ, &w.bS, &w.bP, &w.bP2, &w.bW, &w.bX}) HOST_CALL(salloc(b, nv, s));
        if (f32) HOST_CALL(salloc(&w.bP32, nv, s));
        const size_t mm = (size_t)kMaxBlock * kMaxBlock;
        HOST_CALL(salloc(&w.gpart, (size_t)kMaxPairs * gblk * mm, s));
        HOST_CALL(salloc(&w.gseg, (size_t)kMaxPairs * kGramSegs * mm, s));
        HOST_CALL(salloc(&w.gsum, (size_t)kMaxPairs * mm, s));
        HOST_CALL(salloc(&w.dcm, (size_t)(kCSlots + kStepMats) * mm, s));
        HOST_CALL(salloc(&w.bidx, R, s));
        static_assert(PinnedPool::kBytes >= sizeof(double2) * (kMaxPairs + kCSlots + kStepMats) * kMaxBlock * kMaxBlock,
                      "pinned block too small");
        HOST_CALL(PinnedPool::acquire(&w.hgram));
        w.hcm = w.hgram + kMaxPairs * mm;
    }
    const bool f16s = want_f16s && w.AH;  // (the copy is skipped where memory is short)
    float amax = 0.f;
    if (f16s) {
        HOST_CALL(salloc(&w.xsplit, ws_split_x(std::max(M, N), R), s));
        float *dmax = nullptr;
        HOST_CALL(adjoint_absmax_final(w.amaxws, M, N, &dmax, s));
        HIP_CALL(hipMemcpyAsync(&amax, dmax, sizeof(float), hipMemcpyDeviceToHost, s));
        HIP_CALL(hipStreamSynchronize(s));
    }
    const int bs = 256;
    const dim3 gN((N + bs - 1) / bs, R);

    // FP64-accumulated residual rr = b - A X and s64 = A^H rr
    auto outer_residual = [&](bool x_zero) -> int {
        if (x_zero) {
            HIP_CALL(hipMemcpy2DAsync(w.rr, sizeof(double2) * ld, B, sizeof(double2) * ldb, sizeof(double2) * M, R,
                                      hipMemcpyDeviceToDevice, s));
            if (f16s) {
                // A^H b from X = 0 only sets the first cycle's target and the
                // reference norms (each later residual is FP64), so it runs on
                // the FP16-split product: 2.6 ms instead of 5.7 at M = N = 40962
                kc_to_f32<<<dim3((unsigned)((M + 255) / 256), R), 256, 0, s>>>((int)M, w.rr, ld, w.q32, ld);
                HIP_CALL(hipGetLastError());
                return gemm_hc_f16s_z(w.AH, M, N, amax, w.q32, ld, w.s64, ld, R, w.gws, w.xsplit, s);
            }
        } else if (f32) {
            if (R == 1) HOST_CALL(resid_n_f32a(A32, lda, M, N, X, 0, B, 0, w.rr, 0, 1, w.gws, s));
            else HOST_CALL(mres_f32a(A32, lda, M, N, X, ldx, B, ldb, w.rr, ld, R, w.gws, s));
        } else {
            if (R == 1) HOST_CALL(resid_n_f64(A64, lda, M, N, X, 0, B, 0, w.rr, 0, 1, w.gws, s));
            else HOST_CALL(mres_f64(A64, lda, M, N, X, ldx, B, ldb, w.rr, ld, R, w.gws, s));
        }
        if (f32) {
            if (R == 1) HOST_CALL(gemv_h_f32a(A32, lda, M, N, w.rr, 0, w.s64, 0, 1, s));
            else if (w.AH) HOST_CALL(mh_f32a_hc(w.AH, M, N, w.rr, ld, w.s64, ld, R, w.gws, s));
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

    // ---- block CG pieces
    const bool trace = getenv("BEM_TRACE_SOLVE") != nullptr;
    // the second CholeskyQR pass at every step / S formed on the device with an
    // explicit CholeskyQR2 at every step (both taken by default only when needed)
    const bool force_two_pass = getenv("BEM_BLOCK_CG_TWO_PASS") != nullptr;
    const bool force_explicit = getenv("BEM_BLOCK_CG_EXPLICIT") != nullptr;
    // G[y] (Rb x Rb, host) = U_y^H V_y over the N rows for npair pairs: one
    // launch of the partials, one wait
    const bool gram_valu = getenv("BEM_GRAM_VALU") != nullptr;
    const int gblk_m = (N + kGramRowsM - 1) / kGramRowsM;
    auto grams = [&](const GramPairs &gp, int npair, int Rb, Small *G) -> int {
        const int RR = Rb * Rb;
        int nb = gblk;
        if (gram_valu) {
            kc_gram_part<<<dim3(gblk, npair), 256, 0, s>>>(gp, ld, N, Rb, w.gpart);
        } else if (Rb <= 32) {
            nb = gblk_m;
            kc_gram_mfma<1, 4><<<dim3(gblk_m, npair), 256, 0, s>>>(gp, ld, N, Rb, w.gpart);
        } else {
            nb = gblk_m;
            kc_gram_mfma<4, 2><<<dim3(gblk_m, npair), 256, 0, s>>>(gp, ld, N, Rb, w.gpart);
        }
        kc_gram_sum1<<<dim3((RR + 63) / 64, kGramSegs, npair), 64, 0, s>>>(w.gpart, nb, RR, w.gseg);
        kc_gram_sum2<<<dim3((RR + 255) / 256, npair), 256, 0, s>>>(w.gseg, RR, w.gsum);
        HIP_CALL(hipGetLastError());
        HIP_CALL(hipMemcpyAsync(w.hgram, w.gsum, sizeof(double2) * RR * npair, hipMemcpyDeviceToHost, s));
        HIP_CALL(hipStreamSynchronize(s));
        for (int y = 0; y < npair; y++) {
            G[y].resize((size_t)RR);
            for (int e = 0; e < RR; e++) G[y][e] = cd(w.hgram[(size_t)y * RR + e].x, w.hgram[(size_t)y * RR + e].y);
        }
        return EXIT_SUCCESS;
    };
    auto gram = [&](const double2 *U, const double2 *V, int Rb, Small &G) -> int {
        GramPairs gp{};
        gp.U[0] = U;
        gp.V[0] = V;
        return grams(gp, 1, Rb, &G);
    };
    // out = Z (or 0) + U C for an Rb x Rb host matrix C
    auto block_mul_grid = [&](int Rb) {
        return (unsigned)(((N + 255) / 256) * ((Rb + kMulCols - 1) / kMulCols));
    };
    int cslot = 0;
    auto bmul = [&](const double2 *U, const Small &Cm, int Rb, const double2 *Z, double2 *out,
                    float2 *out32) -> int {
        double2 *hc = w.hcm + (size_t)cslot * kMaxBlock * kMaxBlock;
        double2 *dc = w.dcm + (size_t)cslot * kMaxBlock * kMaxBlock;
        cslot = (cslot + 1) % kCSlots;
        for (int e = 0; e < Rb * Rb; e++) hc[e] = make_double2(Cm[e].real(), Cm[e].imag());
        HIP_CALL(hipMemcpyAsync(dc, hc, sizeof(double2) * Rb * Rb, hipMemcpyHostToDevice, s));
        kc_block_mul<<<block_mul_grid(Rb), 256, 0, s>>>(U, dc, Rb, Z, out, out32, ld, N);
        HIP_CALL(hipGetLastError());
        return EXIT_SUCCESS;
    };
    // the R x R matrices of a block step, uploaded together into a step area;
    // *dc receives the device copy (matrix q at dc + q Rb^2)
    auto upload = [&](int area, std::initializer_list<const Small *> mats, int Rb, const double2 **dc) -> int {
        const size_t mm = (size_t)kMaxBlock * kMaxBlock, RR = (size_t)Rb * Rb;
        const size_t off = (kCSlots + (size_t)area * 4) * mm;
        double2 *h = w.hcm + off;
        size_t q = 0;
        for (const Small *m : mats) {
            for (size_t e = 0; e < RR; e++) h[q * RR + e] = make_double2((*m)[e].real(), (*m)[e].imag());
            q++;
        }
        HIP_CALL(hipMemcpyAsync(w.dcm + off, h, sizeof(double2) * RR * q, hipMemcpyHostToDevice, s));
        *dc = w.dcm + off;
        return EXIT_SUCCESS;
    };
    auto bmul_dev = [&](const double2 *U, const double2 *dc, int Rb, const double2 *Z, double2 *out,
                        float2 *out32) -> int {
        kc_block_mul<<<block_mul_grid(Rb), 256, 0, s>>>(U, dc, Rb, Z, out, out32, ld, N);
        HIP_CALL(hipGetLastError());
        return EXIT_SUCCESS;
    };
    // W = A^H (A P) for the block
    auto bprod = [&](int Rb) -> int {
        if (f32 && f16s) {
            HOST_CALL(gemm_n_f16s(A32, lda, M, N, amax, w.bP32, ld, w.q32, ld, Rb, w.gws, w.xsplit, s));
            HOST_CALL(gemm_hc_f16s_z(w.AH, M, N, amax, w.q32, ld, w.bW, ld, Rb, w.gws, w.xsplit, s));
        } else if (f32) {
            HOST_CALL(gemm_n_f32(A32, lda, M, N, w.bP32, ld, w.q32, ld, Rb, w.gws, s, true));
            if (w.AH) HOST_CALL(gemm_hc_f32z(w.AH, M, N, w.q32, ld, w.bW, ld, Rb, w.gws, s));
            else HOST_CALL(gemm_h_f32z(A32, lda, M, N, w.q32, ld, w.bW, ld, Rb, w.gws, s));
        } else {
            for (int c = 0; c < Rb; c++) {
                HOST_CALL(gemv_n_f64(A64, lda, M, N, w.bP + (size_t)c * ld, 0, w.q64 + (size_t)c * ld, 0, 1, w.gws,
                                     s));
                HOST_CALL(gemv_h_f64(A64, lda, M, N, w.q64 + (size_t)c * ld, 0, w.bW + (size_t)c * ld, 0, 1, s));
            }
        }
        return EXIT_SUCCESS;
    };
    // CholeskyQR2 of the block in bS: bQ zeta = bS (bS is overwritten); ok
    // false when a Gram matrix is not numerically positive definite
    auto cholqr2 = [&](int Rb, Small &zeta, bool &ok) -> int {
        ok = false;
        Small G, L1, L2;
        HOST_CALL(gram(w.bS, w.bS, Rb, G));
        if (!chol_lower(G, Rb, L1)) return EXIT_SUCCESS;
        HOST_CALL(bmul(w.bS, herm(lower_inv(L1, Rb), Rb), Rb, nullptr, w.bQ, nullptr));
        HOST_CALL(gram(w.bQ, w.bQ, Rb, G));
        if (!chol_lower(G, Rb, L2)) return EXIT_SUCCESS;
        HOST_CALL(bmul(w.bQ, herm(lower_inv(L2, Rb), Rb), Rb, nullptr, w.bS, nullptr));
        std::swap(w.bQ, w.bS);
        zeta = mul(herm(L2, Rb), herm(L1, Rb), Rb);
        ok = true;
        return EXIT_SUCCESS;
    };
    int total = 0;
    // one block cycle on the active systems idx (thr: |residual| to reach per
    // system); its = block steps taken, broke = ended by a breakdown
    auto block_cycle = [&](const std::vector<int> &idx, const std::vector<double> &thr, int &its,
                           bool &broke) -> int {
        const int Rb = (int)idx.size();
        its = 0;
        broke = false;
        HIP_CALL(hipMemcpyAsync(w.bidx, idx.data(), sizeof(int) * Rb, hipMemcpyHostToDevice, s));
        const dim3 gb((N + 255) / 256, Rb);
        kc_gather<<<gb, 256, 0, s>>>(N, w.s64, w.bidx, w.bS, ld);
        HIP_CALL(hipGetLastError());
        HIP_CALL(hipMemsetAsync(w.bX, 0, sizeof(double2) * ld * Rb, s));
        Small xi, zeta;
        bool ok = false;
        HOST_CALL(cholqr2(Rb, xi, ok));
        if (!ok) {
            broke = true;
            return EXIT_SUCCESS;
        }
        Small eye((size_t)Rb * Rb, cd(0));
        for (int c = 0; c < Rb; c++) eye[(size_t)c * Rb + c] = 1.0;
        HOST_CALL(bmul(w.bQ, eye, Rb, nullptr, w.bP, f32 ? w.bP32 : nullptr));  // P = Q
        const int RR = Rb * Rb;
        // residual of system c is |xi[:, c]|
        auto converged = [&](const Small &x, double &worst) {
            bool all = true;
            worst = 0;
            for (int c = 0; c < Rb; c++) {
                double nrm = 0;
                for (int r = 0; r < Rb; r++) nrm += std::norm(x[(size_t)c * Rb + r]);
                nrm = std::sqrt(nrm);
                worst = std::max(worst, thr[c] > 0 ? nrm / thr[c] : 0.0);
                all = all && nrm <= thr[c];
            }
            return all;
        };
        auto fro2 = [&](const Small &x) {
            double t = 0;
            for (const cd &v : x) t += std::norm(v);
            return t;
        };
        while (total < maxit) {
            HOST_CALL(bprod(Rb));
            // one launch, one wait: D = P^H W, E = Q^H W, F = W^H W, Q^H Q
            GramPairs gp{};
            gp.U[0] = w.bP; gp.V[0] = w.bW;
            gp.U[1] = w.bQ; gp.V[1] = w.bW;
            gp.U[2] = w.bW; gp.V[2] = w.bW;
            gp.U[3] = w.bQ; gp.V[3] = w.bQ;
            Small Gs[4];
            HOST_CALL(grams(gp, 4, Rb, Gs));
            Small Ld;
            if (!chol_lower(Gs[0], Rb, Ld)) {  // D = P^H A^H A P, Hermitian positive definite
                broke = true;
                break;
            }
            const Small Li = lower_inv(Ld, Rb);
            const Small alpha = mul(herm(Li, Rb), Li, Rb);  // D^-1
            const Small axi = mul(alpha, xi, Rb);
            Small malpha = alpha;
            for (cd &v : malpha) v = -v;
            // S^H S for S = Q - W alpha from the Gram products:
            //   Q^H Q - E alpha - (E alpha)^H + alpha^H F alpha
            const Small Ea = mul(Gs[1], alpha, Rb);
            const Small aFa = mul(herm(alpha, Rb), mul(Gs[2], alpha, Rb), Rb);
            Small G1((size_t)RR);
            for (int j = 0; j < Rb; j++)
                for (int i = 0; i < Rb; i++)
                    G1[(size_t)j * Rb + i] = Gs[3][(size_t)j * Rb + i] - Ea[(size_t)j * Rb + i] -
                                             std::conj(Ea[(size_t)i * Rb + j]) + aFa[(size_t)j * Rb + i];
            const double gterms = std::sqrt(fro2(Gs[3])) + 2 * std::sqrt(fro2(Ea)) + std::sqrt(fro2(aFa));
            Small L1;
            const bool alg = !force_explicit && chol_lower(G1, Rb, L1);
            total++;
            its++;
            double worst = 0;
            bool all = false;
            bool one_pass_taken = false;
            if (alg) {
                const Small L1i = lower_inv(L1, Rb), L1ih = herm(L1i, Rb);
                // Q' = S L1^-H is orthonormal up to the rounding of the Gram
                // terms times |L1^-1|^2 (FP64 sums of N products); one pass
                // when that stays below 1e-9, else a second CholeskyQR pass on
                // an explicit Gram product
                const bool one_pass = !force_two_pass && 1e-14 * gterms * fro2(L1i) <= 1e-9;
                one_pass_taken = one_pass;
                const double2 *dc = nullptr;
                if (one_pass) {
                    zeta = herm(L1, Rb);
                    const Small zh = herm(zeta, Rb);
                    const Small xin = mul(zeta, xi, Rb);
                    all = converged(xin, worst);
                    HOST_CALL(upload(0, {&axi, &malpha, &L1ih, &zh}, Rb, &dc));
                    HOST_CALL(bmul_dev(w.bP, dc, Rb, w.bX, w.bX, nullptr));  // X += P alpha xi
                    if (!all) {
                        HOST_CALL(bmul_dev(w.bW, dc + RR, Rb, w.bQ, w.bS, nullptr));         // S = Q - W alpha
                        HOST_CALL(bmul_dev(w.bS, dc + 2 * RR, Rb, nullptr, w.bQ, nullptr));  // Q = S L1^-H
                        HOST_CALL(bmul_dev(w.bP, dc + 3 * RR, Rb, w.bQ, w.bP2, f32 ? w.bP32 : nullptr));
                        std::swap(w.bP, w.bP2);  // P = Q + P zeta^H
                    }
                    xi = xin;
                } else {
                    HOST_CALL(upload(0, {&axi, &malpha, &L1ih}, Rb, &dc));
                    HOST_CALL(bmul_dev(w.bP, dc, Rb, w.bX, w.bX, nullptr));
                    HOST_CALL(bmul_dev(w.bW, dc + RR, Rb, w.bQ, w.bS, nullptr));
                    HOST_CALL(bmul_dev(w.bS, dc + 2 * RR, Rb, nullptr, w.bQ, nullptr));  // Q1
                    Small G2, L2;
                    HOST_CALL(gram(w.bQ, w.bQ, Rb, G2));
                    if (!chol_lower(G2, Rb, L2)) {
                        broke = true;
                        break;
                    }
                    zeta = mul(herm(L2, Rb), herm(L1, Rb), Rb);
                    const Small zh = herm(zeta, Rb), L2ih = herm(lower_inv(L2, Rb), Rb);
                    const Small xin = mul(zeta, xi, Rb);
                    all = converged(xin, worst);
                    if (!all) {
                        HOST_CALL(upload(1, {&L2ih, &zh}, Rb, &dc));
                        HOST_CALL(bmul_dev(w.bQ, dc, Rb, nullptr, w.bS, nullptr));
                        std::swap(w.bQ, w.bS);  // Q = Q1 L2^-H
                        HOST_CALL(bmul_dev(w.bP, dc + RR, Rb, w.bQ, w.bP2, f32 ? w.bP32 : nullptr));
                        std::swap(w.bP, w.bP2);
                    }
                    xi = xin;
                }
            } else {
                // S^H S from the terms is not positive definite beyond rounding:
                // form S on the device and take CholeskyQR2 of it
                const double2 *dc = nullptr;
                HOST_CALL(upload(0, {&axi, &malpha}, Rb, &dc));
                HOST_CALL(bmul_dev(w.bP, dc, Rb, w.bX, w.bX, nullptr));
                HOST_CALL(bmul_dev(w.bW, dc + RR, Rb, w.bQ, w.bS, nullptr));
                bool ok = false;
                HOST_CALL(cholqr2(Rb, zeta, ok));
                if (!ok) {
                    broke = true;
                    break;
                }
                const Small xin = mul(zeta, xi, Rb);
                all = converged(xin, worst);
                if (!all) {
                    HOST_CALL(bmul(w.bP, herm(zeta, Rb), Rb, w.bQ, w.bP2, f32 ? w.bP32 : nullptr));
                    std::swap(w.bP, w.bP2);
                }
                xi = xin;
            }
            if (trace)
                fprintf(stderr, "block cg step %d (%s): worst |res| / threshold %.3e\n", total,
                        !alg ? "explicit QR" : one_pass_taken ? "one QR pass" : "two QR passes", worst);
            if (all) break;
        }
        kc_scatter_add<<<gb, 256, 0, s>>>(N, w.bX, ld, w.bidx, X, ldx);
        HIP_CALL(hipGetLastError());
        return EXIT_SUCCESS;
    };

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
    int cycles = 0;
    std::vector<int> hact(R);
    for (int cyc = 0; cyc < max_cycles; cyc++) {
        if (!res_current) {
            HOST_CALL(outer_residual(false));
            HOST_CALL(norms(beta));
            res_current = true;
        }
        if (trace) {
            double rmax = 0;
            int above = 0;
            for (int q = 0; q < R; q++) {
                const double rq = nb[q] > 0 ? beta[q] / nb[q] : 0.0;
                rmax = std::max(rmax, rq);
                above += rq > outer_tol;
            }
            fprintf(stderr, "cgls cycle %d its %d rel[0] %.3e max rel %.3e (%d above the tolerance)\n", cyc, total,
                    beta[0] / nb[0], rmax, above);
        }
        bool any = false;
        std::vector<int> idx;
        std::vector<double> bthr;
        for (int q = 0; q < R; q++) {
            if (nb[q] > 0 && beta[q] <= outer_tol * nb[q]) done[q] = 1;
            any = any || !done[q];
            double ctol = tol;
            if (outer_tol > 0 && nb[q] > 0 && beta[q] > 0 && !fixed_ctol)
                ctol = std::max(tol, std::min(0.5, kCycleMargin * outer_tol * nb[q] / beta[q]));
            hthr[q] = ctol * ctol * hg[q];
            if (!done[q]) {
                idx.push_back(q);
                bthr.push_back(ctol * std::sqrt(hg[q]));
            }
        }
        if (!any || total >= maxit) break;
        cycles++;
        if (block_on && idx.size() > 1) {
            int its = 0;
            bool broke = false;
            HOST_CALL(block_cycle(idx, bthr, its, broke));
            if (broke && trace) fprintf(stderr, "block cg: breakdown after %d steps of the cycle\n", its);
            // a cycle that could not start or broke down at once: the lockstep CG takes over
            if (broke && its < 2) block_on = false;
            if (its > 0) {
                res_current = false;
                continue;
            }
        }
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
