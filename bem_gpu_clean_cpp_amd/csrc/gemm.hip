// Multi-right-hand-side complex matrix products on MFMA: Y = A X (M x R) and
// Z = A^H T (N x R) for a complex64 column-major A (M x N, lda).  The
// products run on v_mfma_f32_16x16x4f32 with FP32 operands (the Krylov
// operator), or on v_mfma_f64_16x16x4f64 with A widened to FP64 and FP64
// vectors (the refinement residuals, FP64 accumulation).
//
// The interleaved memory of A is read as a real (2M x N) matrix whose even
// rows hold Re A and odd rows Im A, so one 16-byte load per lane gives 4 real
// rows of one column and feeds 4 MFMAs (one per element).  For Y = A X the
// products P = A~ Re(X) and Q = A~ Im(X) land in the same lane and register
// for the real and imaginary row of each complex row, so Y = (P_re - Q_im,
// P_im + Q_re) combines lane-locally.  For Z = A^H T the second operand is T's
// own interleaved memory (and a swapped, negated copy for Im Z).  Columns (or
// rows, for A^H) are split into chunks over the grid; chunk partials are summed
// in a fixed order by gemm_reduce (deterministic, no atomics).  Each wave
// holds the operands of the next 16 columns (or 32 rows) in registers while
// it multiplies the current ones, so the HBM latency of the A stream is
// covered by MFMA work.
//
// MFMA lane layout (16x16x4, one block): A operand A[i][k] from lane i + 16 k,
// B operand B[k][j] from lane j + 16 k (both precisions); the result
// D[4 (l / 16) + v][l % 16] (f32) or D[l / 16 + 4 v][l % 16] (f64) in lane l,
// register v.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bem_internal.hpp"

namespace bem {

template <typename T>
struct MF;
template <>
struct MF<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    typedef float2 c2;
    static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // result row held by lane group q = lane / 16 in register v
    static __device__ __forceinline__ int drow(int q, int v) { return 4 * q + v; }
};
template <>
struct MF<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    typedef double2 c2;
    static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int drow(int q, int v) { return q + 4 * v; }
};

constexpr int kU = 4;  // MFMA steps (of 4 columns / 8 rows) per prefetched block

// keep v where k holds, else zero, without a branch
template <typename S>
__device__ __forceinline__ S keep(bool k, S v) {
    return k ? v : S(0);
}
__device__ __forceinline__ float4 keep4(bool k01, bool k23, float4 v) {
    return make_float4(keep(k01, v.x), keep(k01, v.y), keep(k23, v.z), keep(k23, v.w));
}
template <typename C2>
__device__ __forceinline__ C2 keep2(bool k, C2 v) {
    C2 r;
    r.x = keep(k, (decltype(r.x + 0))v.x);
    r.y = keep(k, (decltype(r.y + 0))v.y);
    return r;
}

// part[(chunk * R + n) * ldp + i] = sum_{j in chunk} A[i, j] X[j, n]
// wave: 32 complex rows x (16 G) right-hand sides x one column chunk;
// requires lda even and lda >= roundup(M, 32) (rows row0..row0+31 allocated).
// Loads are unconditional at clamped (in-bounds) addresses and masked to zero
// afterwards, so the loop body is straight-line and the next block's loads stay
// in flight while the current block is multiplied (ping-pong buffers).
template <typename T, int G>
__global__ __launch_bounds__(256) void gemm_n_mfma(const float2 *__restrict__ A, long lda, int M, int N,
                                                   const typename MF<T>::c2 *__restrict__ X, long ldx, int R, int rg0,
                                                   int kc, typename MF<T>::c2 *__restrict__ part, long ldp) {
    typedef typename MF<T>::v4 v4;
    typedef typename MF<T>::c2 c2;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row0 = (blockIdx.x * 4 + wave) * 32;
    if (row0 >= M) return;
    const int chunk = blockIdx.y;
    const int j0 = chunk * kc, j1 = min(N, j0 + kc);
    const int m = lane & 15, kk = lane >> 4;
    v4 acc[G][2][4];
#pragma unroll
    for (int g = 0; g < G; g++)
#pragma unroll
        for (int p = 0; p < 2; p++)
#pragma unroll
            for (int t = 0; t < 4; t++) acc[g][p][t] = (v4){0, 0, 0, 0};
    const float4 *Ab = reinterpret_cast<const float4 *>(A + row0 + 2 * m);
    const long lda4 = lda / 2;
    // right-hand sides past R read column R - 1 (their results are not written)
    const c2 *Xg[G];
#pragma unroll
    for (int g = 0; g < G; g++) Xg[g] = X + (size_t)min(rg0 + g * 16 + m, R - 1) * ldx;
    struct Buf {
        float4 a[kU];
        c2 x[kU][G];
    };
    // full blocks: every column inside the chunk, no masking; the address is
    // clamped only for the prefetch past the last pair of blocks
    auto load = [&](int jb, Buf &B) {
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const int Jc = min(jb + 4 * u + kk, j1 - 1);
            B.a[u] = Ab[(size_t)Jc * lda4];
#pragma unroll
            for (int g = 0; g < G; g++) B.x[u][g] = Xg[g][Jc];
        }
    };
    auto load_masked = [&](int jb, Buf &B) {
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const int J = jb + 4 * u + kk;
            const bool in = J < j1;
            const int Jc = min(J, j1 - 1);
            B.a[u] = keep4(in, in, Ab[(size_t)Jc * lda4]);
#pragma unroll
            for (int g = 0; g < G; g++) B.x[u][g] = keep2(in, Xg[g][Jc]);
        }
    };
    auto mul = [&](const Buf &B) {
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const T av[4] = {(T)B.a[u].x, (T)B.a[u].y, (T)B.a[u].z, (T)B.a[u].w};
#pragma unroll
            for (int g = 0; g < G; g++)
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    acc[g][0][t] = MF<T>::mma(av[t], B.x[u][g].x, acc[g][0][t]);
                    acc[g][1][t] = MF<T>::mma(av[t], B.x[u][g].y, acc[g][1][t]);
                }
        }
    };
    Buf b0, b1;
    const int jpair = j0 + (j1 - j0) / (8 * kU) * (8 * kU);  // end of the paired full blocks
    load(j0, b0);
    for (int jb = j0; jb < jpair; jb += 8 * kU) {
        load(jb + 4 * kU, b1);
        __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the products
        mul(b0);
        load(jb + 8 * kU, b0);
        __builtin_amdgcn_sched_barrier(0);
        mul(b1);
    }
    for (int jb = jpair; jb < j1; jb += 4 * kU) {
        load_masked(jb, b1);
        mul(b1);
    }
    // D_t[m'][n'] (m' = drow(lane / 16, v)) is real row 4 m' + t of the tile:
    // complex row 2 m' + t / 2, real part for even t
#pragma unroll
    for (int g = 0; g < G; g++) {
        const int n = rg0 + g * 16 + m;
        if (n >= R) continue;
        c2 *pc = part + ((size_t)chunk * R + n) * ldp;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int mp = MF<T>::drow(kk, v);
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int i = row0 + 2 * mp + h;
                if (i < M) pc[i] = c2{acc[g][0][2 * h][v] - acc[g][1][2 * h + 1][v],
                                      acc[g][0][2 * h + 1][v] + acc[g][1][2 * h][v]};
            }
        }
    }
}

// part[(chunk * R + n) * ldp + j] = sum_{i in chunk} conj(A[i, j]) T[i, n]
// wave: 16 columns of A x (16 G) right-hand sides x one row chunk (i0 even).
// Rows read in pairs: row + 1 may be M, which needs lda, ldt > M for odd M
// (the callers' leading dimensions are M rounded up to 64).
template <typename T, int G>
__global__ __launch_bounds__(256) void gemm_h_mfma(const float2 *__restrict__ A, long lda, int M, int N,
                                                   const typename MF<T>::c2 *__restrict__ Tv, long ldt, int R, int rg0,
                                                   int kr, typename MF<T>::c2 *__restrict__ part, long ldp) {
    typedef typename MF<T>::v4 v4;
    typedef typename MF<T>::c2 c2;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c0 = (blockIdx.x * 4 + wave) * 16;
    if (c0 >= N) return;
    const int chunk = blockIdx.y;
    const int i0 = chunk * kr, i1 = min(M, i0 + kr);
    const int m = lane & 15, kk = lane >> 4;
    const int J = c0 + m;
    v4 accR[G], accI[G];
#pragma unroll
    for (int g = 0; g < G; g++) {
        accR[g] = (v4){0, 0, 0, 0};
        accI[g] = (v4){0, 0, 0, 0};
    }
    // columns past N and right-hand sides past R read the last valid one
    // (their results are not written)
    const float2 *col = A + (size_t)min(J, N - 1) * lda;
    const c2 *Tg[G];
#pragma unroll
    for (int g = 0; g < G; g++) Tg[g] = Tv + (size_t)min(rg0 + g * 16 + m, R - 1) * ldt;
    const int rlast = (i1 - 1) & ~1;  // last even row of the chunk
    struct Buf {
        float4 a[kU];
        c2 t0[kU][G], t1[kU][G];
    };
    auto load = [&](int ib, Buf &B) {
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const int rc = min(ib + 8 * u + 2 * kk, rlast);  // complex rows rc, rc + 1 of this lane
            B.a[u] = *reinterpret_cast<const float4 *>(col + rc);
#pragma unroll
            for (int g = 0; g < G; g++) {
                B.t0[u][g] = Tg[g][rc];
                B.t1[u][g] = Tg[g][rc + 1];
            }
        }
    };
    auto load_masked = [&](int ib, Buf &B) {
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const int row = ib + 8 * u + 2 * kk;
            const bool in0 = row < i1, in1 = row + 1 < i1;
            const int rc = min(row, rlast);
            B.a[u] = keep4(in0, in1, *reinterpret_cast<const float4 *>(col + rc));
#pragma unroll
            for (int g = 0; g < G; g++) {
                B.t0[u][g] = keep2(in0, Tg[g][rc]);
                B.t1[u][g] = keep2(in1, Tg[g][rc + 1]);
            }
        }
    };
    auto mul = [&](const Buf &B) {
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const T av[4] = {(T)B.a[u].x, (T)B.a[u].y, (T)B.a[u].z, (T)B.a[u].w};
#pragma unroll
            for (int g = 0; g < G; g++) {
                // Re Z: sum Ar Tr + Ai Ti ; Im Z: sum Ar Ti - Ai Tr
                const T br[4] = {B.t0[u][g].x, B.t0[u][g].y, B.t1[u][g].x, B.t1[u][g].y};
                const T bi[4] = {B.t0[u][g].y, -B.t0[u][g].x, B.t1[u][g].y, -B.t1[u][g].x};
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    accR[g] = MF<T>::mma(av[t], br[t], accR[g]);
                    accI[g] = MF<T>::mma(av[t], bi[t], accI[g]);
                }
            }
        }
    };
    Buf b0, b1;
    const int ipair = i0 + (i1 - i0) / (16 * kU) * (16 * kU);  // end of the paired full blocks
    load(i0, b0);
    for (int ib = i0; ib < ipair; ib += 16 * kU) {
        load(ib + 8 * kU, b1);
        __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the products
        mul(b0);
        load(ib + 16 * kU, b0);
        __builtin_amdgcn_sched_barrier(0);
        mul(b1);
    }
    for (int ib = ipair; ib < i1; ib += 8 * kU) {
        load_masked(ib, b1);
        mul(b1);
    }
#pragma unroll
    for (int g = 0; g < G; g++) {
        const int n = rg0 + g * 16 + m;
        if (n >= R) continue;
        c2 *pc = part + ((size_t)chunk * R + n) * ldp;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int j = c0 + MF<T>::drow(kk, v);
            if (j < N) pc[j] = c2{accR[g][v], accI[g][v]};
        }
    }
}

// Y[n][i] = (b[n][i] -) sum_chunks part[(chunk * R + n) * ldp + i]  (fixed order)
// (partials of type MF<T>::c2, output OC2 accumulated in S)
template <typename T, typename S, typename OC2>
__global__ __launch_bounds__(256) void gemm_reduce(const typename MF<T>::c2 *__restrict__ part, long ldp, int nchunks,
                                                   int R, int M, const OC2 *__restrict__ b, long ldb,
                                                   OC2 *__restrict__ Y, long ldy) {
    typedef typename MF<T>::c2 C2;
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int n = blockIdx.y;
    if (i >= M) return;
    S sr = 0, si = 0;
    // 8 partials in flight per step, added in chunk order
    int c = 0;
    for (; c + 8 <= nchunks; c += 8) {
        C2 p[8];
#pragma unroll
        for (int u = 0; u < 8; u++) p[u] = part[((size_t)(c + u) * R + n) * ldp + i];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            sr += (S)p[u].x;
            si += (S)p[u].y;
        }
    }
    for (; c < nchunks; c++) {
        const C2 p = part[((size_t)c * R + n) * ldp + i];
        sr += (S)p.x;
        si += (S)p.y;
    }
    if (b) {
        const OC2 bb = b[(size_t)n * ldb + i];
        Y[(size_t)n * ldy + i] = OC2{bb.x - sr, bb.y - si};
    } else {
        Y[(size_t)n * ldy + i] = OC2{sr, si};
    }
}

// column chunks: max(3, 1024 / row tiles).  Fewer, longer chunks than for
// 4096 workgroups: measured Y = A X at R = 16, M = 2562: 25.3 us with 41 chunks,
// 17.3 with 12; at R = 32, M = 10258: 364 us with 12, 252 with 3; at
// M = 40994 three chunks (5.4 ms) beat one (8.2 ms)
// (long_chunks: for products whose rounding a refinement absorbs; the short
// rule, chunks for 4096 workgroups, keeps FP32 sums short where the product
// is final, as in the QR updates and the field operator)
static void gemm_n_shape(long M, long N, int &nch, int &kc, bool long_chunks) {
    const long tiles = (M + 31) / 32;
    long target = long_chunks ? std::max(3L, 1024 / std::max(1L, tiles)) : std::max(1L, 4096 / std::max(1L, tiles));
    kc = (int)std::max(32L, ((N + target - 1) / target + 31) / 32 * 32);
    nch = (int)((N + kc - 1) / kc);
}

// row chunks: max(3, 1024 / column tiles), as for gemm_n (measured Z = A^H T at
// R = 32: M = 10258, 334 -> 321 us; M = 40994, 5.81 -> 5.29 ms against chunks
// for 4096 workgroups)
static void gemm_h_shape(long M, long N, int &nch, int &kr, bool long_chunks) {
    const long tiles = (N + 15) / 16;
    long target = long_chunks ? std::max(3L, 1024 / std::max(1L, tiles)) : std::max(1L, 4096 / std::max(1L, tiles));
    kr = (int)std::max(64L, ((M + target - 1) / target + 63) / 64 * 64);
    nch = (int)((M + kr - 1) / kr);
}

// partials of the FP32 adjoint product alone (Z = A^H T, N x R)
size_t ws_gemm_h_f32(long M, long N, int R) {
    int nch, kr;
    gemm_h_shape(M, N, nch, kr, false);
    return (size_t)nch * (size_t)((N + 63) / 64 * 64) * (size_t)std::max(1, R) * sizeof(float2);
}

// partials of either product at FP64 for R right-hand sides
size_t ws_gemm(long M, long N, int R) {
    int nch = 0;
    for (bool lc : {false, true}) {
        int a, b, c, d;
        gemm_n_shape(M, N, a, b, lc);
        gemm_h_shape(M, N, c, d, lc);
        nch = std::max({nch, a, c});
    }
    const long ldp = (std::max(M, N) + 63) / 64 * 64;
    return (size_t)nch * (size_t)ldp * (size_t)std::max(1, R) * sizeof(double2);
}

template <typename T, typename S = T, typename OC2 = typename MF<T>::c2>
static int gemm_n_impl(const float2 *A, long lda, long M, long N, const typename MF<T>::c2 *X, long ldx,
                       const OC2 *b, long ldb, OC2 *Y, long ldy, int R, void *ws, const char *name, hipStream_t s,
                       bool long_chunks) {
    if (M <= 0 || N <= 0 || R <= 0) return EXIT_SUCCESS;
    if (lda % 2 || lda < (M + 31) / 32 * 32) BEM_FAIL("gemm_n needs an even lda >= roundup(M, 32)");
    int nch, kc;
    gemm_n_shape(M, N, nch, kc, long_chunks);
    const long ldp = (M + 63) / 64 * 64;
    auto *part = (typename MF<T>::c2 *)ws;
    TimedRegion tr;
    timing_begin(tr, name, s);
    const unsigned gx = (unsigned)((M + 127) / 128);
    for (int rg = 0; rg < R; rg += 32) {
        dim3 grid(gx, nch);
        if (R - rg > 16)
            gemm_n_mfma<T, 2><<<grid, 256, 0, s>>>(A, lda, (int)M, (int)N, X, ldx, R, rg, kc, part, ldp);
        else
            gemm_n_mfma<T, 1><<<grid, 256, 0, s>>>(A, lda, (int)M, (int)N, X, ldx, R, rg, kc, part, ldp);
        HIP_CALL(hipGetLastError());
    }
    timing_end(tr, s);
    dim3 g2((unsigned)((M + 255) / 256), R);
    gemm_reduce<T, S, OC2><<<g2, 256, 0, s>>>(part, ldp, nch, R, (int)M, b, ldb, Y, ldy);
    HIP_CALL(hipGetLastError());
    return EXIT_SUCCESS;
}

template <typename T, typename S = T, typename OC2 = typename MF<T>::c2>
static int gemm_h_impl(const float2 *A, long lda, long M, long N, const typename MF<T>::c2 *Tv, long ldt, OC2 *Z,
                       long ldz, int R, void *ws, const char *name, hipStream_t s, bool long_chunks) {
    if (M <= 0 || N <= 0 || R <= 0) return EXIT_SUCCESS;
    if (lda % 2) BEM_FAIL("gemm_h needs an even lda");
    int nch, kr;
    gemm_h_shape(M, N, nch, kr, long_chunks);
    const long ldp = (N + 63) / 64 * 64;
    auto *part = (typename MF<T>::c2 *)ws;
    TimedRegion tr;
    timing_begin(tr, name, s);
    const unsigned gx = (unsigned)((N + 63) / 64);
    for (int rg = 0; rg < R; rg += 32) {
        dim3 grid(gx, nch);
        if (R - rg > 16)
            gemm_h_mfma<T, 2><<<grid, 256, 0, s>>>(A, lda, (int)M, (int)N, Tv, ldt, R, rg, kr, part, ldp);
        else
            gemm_h_mfma<T, 1><<<grid, 256, 0, s>>>(A, lda, (int)M, (int)N, Tv, ldt, R, rg, kr, part, ldp);
        HIP_CALL(hipGetLastError());
    }
    timing_end(tr, s);
    dim3 g2((unsigned)((N + 255) / 256), R);
    gemm_reduce<T, S, OC2><<<g2, 256, 0, s>>>(part, ldp, nch, R, (int)N, (const OC2 *)nullptr, 0, Z, ldz);
    HIP_CALL(hipGetLastError());
    return EXIT_SUCCESS;
}

// Y (M x R, ldy) = A X ; ws >= ws_gemm(M, N, R)
int gemm_n_f32(const float2 *A, long lda, long M, long N, const float2 *X, long ldx, float2 *Y, long ldy, int R,
               void *ws, hipStream_t s, bool long_chunks) {
    return gemm_n_impl<float>(A, lda, M, N, X, ldx, (const float2 *)nullptr, 0, Y, ldy, R, ws, "gemm_n_f32", s,
                              long_chunks);
}

// Z (N x R, ldz) = A^H T, T: M x R (ldt); ws >= ws_gemm(M, N, R)
int gemm_h_f32(const float2 *A, long lda, long M, long N, const float2 *T, long ldt, float2 *Z, long ldz, int R,
               void *ws, hipStream_t s, bool long_chunks) {
    return gemm_h_impl<float>(A, lda, M, N, T, ldt, Z, ldz, R, ws, "gemm_h_f32", s, long_chunks);
}

// FP32 products whose fixed-order reduce writes FP64 vectors (the Krylov basis
// input; refined, so long chunks)
int gemm_n_f32z(const float2 *A, long lda, long M, long N, const float2 *X, long ldx, double2 *Y, long ldy, int R,
                void *ws, hipStream_t s) {
    return gemm_n_impl<float, double, double2>(A, lda, M, N, X, ldx, (const double2 *)nullptr, 0, Y, ldy, R, ws,
                                               "gemm_n_f32", s, true);
}
int gemm_h_f32z(const float2 *A, long lda, long M, long N, const float2 *T, long ldt, double2 *Z, long ldz, int R,
                void *ws, hipStream_t s) {
    return gemm_h_impl<float, double, double2>(A, lda, M, N, T, ldt, Z, ldz, R, ws, "gemm_h_f32", s, true);
}

// out (M x R) = b - A X (b == nullptr: out = A X): FP32 A widened to FP64,
// FP64 vectors, FP64 MFMA accumulation; ws >= ws_gemm(M, N, R)
int mres_f32a(const float2 *A, long lda, long M, long N, const double2 *X, long ldx, const double2 *b, long ldb,
              double2 *out, long ldo, int R, void *ws, hipStream_t s) {
    return gemm_n_impl<double>(A, lda, M, N, X, ldx, b, ldb, out, ldo, R, ws, "mres_f64acc", s, true);
}

// Z (N x R) = A^H T with FP64 accumulation; ws >= ws_gemm(M, N, R)
int mh_f32a(const float2 *A, long lda, long M, long N, const double2 *T, long ldt, double2 *Z, long ldz, int R,
            void *ws, hipStream_t s) {
    return gemm_h_impl<double>(A, lda, M, N, T, ldt, Z, ldz, R, ws, "mh_f64acc", s, true);
}

// ------------------------------------------------ QR rank-nb update ---
// C (M x nc, ldc) -= V (M x 32, ldv) W, W stored per column of C (W[c * ldw + l],
// l < 32; columns l >= nb of V and rows l >= nb of W are zero).  The
// trailing update of the blocked Householder QR: wave tile 32 complex rows x
// (16 G) columns of C, K = 32; no split-K (K is one panel), C updated in place.
template <int G>
__global__ __launch_bounds__(256) void rank_update_mfma(const float2 *__restrict__ V, long ldv, int M,
                                                        const float2 *__restrict__ W, int ldw, int nc,
                                                        float2 *__restrict__ Cm, long ldc) {
    typedef MF<float>::v4 v4;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row0 = (blockIdx.x * 4 + wave) * 32;
    if (row0 >= M) return;
    const int c0 = blockIdx.y * 16 * G;
    const int m = lane & 15, kk = lane >> 4;
    v4 acc[G][2][4];
#pragma unroll
    for (int g = 0; g < G; g++)
#pragma unroll
        for (int p = 0; p < 2; p++)
#pragma unroll
            for (int t = 0; t < 4; t++) acc[g][p][t] = (v4){0, 0, 0, 0};
    const float4 *Vb = reinterpret_cast<const float4 *>(V + row0 + 2 * m);
    const long ldv4 = ldv / 2;
    const float2 *Wg[G];
#pragma unroll
    for (int g = 0; g < G; g++) Wg[g] = W + (size_t)min(c0 + g * 16 + m, nc - 1) * ldw;
    float4 a[8];
    float2 w[8][G];
#pragma unroll
    for (int u = 0; u < 8; u++) {
        const int L = 4 * u + kk;
        a[u] = Vb[(size_t)L * ldv4];
#pragma unroll
        for (int g = 0; g < G; g++) w[u][g] = Wg[g][L];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
        const float av[4] = {a[u].x, a[u].y, a[u].z, a[u].w};
#pragma unroll
        for (int g = 0; g < G; g++)
#pragma unroll
            for (int t = 0; t < 4; t++) {
                acc[g][0][t] = MF<float>::mma(av[t], w[u][g].x, acc[g][0][t]);
                acc[g][1][t] = MF<float>::mma(av[t], w[u][g].y, acc[g][1][t]);
            }
    }
#pragma unroll
    for (int g = 0; g < G; g++) {
        const int n = c0 + g * 16 + m;
        if (n >= nc) continue;
        float2 *cc = Cm + (size_t)n * ldc;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int mp = MF<float>::drow(kk, v);
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int i = row0 + 2 * mp + h;
                if (i < M) {
                    float2 o = cc[i];
                    o.x -= acc[g][0][2 * h][v] - acc[g][1][2 * h + 1][v];
                    o.y -= acc[g][0][2 * h + 1][v] + acc[g][1][2 * h][v];
                    cc[i] = o;
                }
            }
        }
    }
}

// C -= V W as above; V needs 32 columns and ldv >= roundup(M, 32), even
int rank_update_f32(const float2 *V, long ldv, int M, const float2 *W, int ldw, int nc, float2 *C, long ldc,
                    hipStream_t s) {
    if (M <= 0 || nc <= 0) return EXIT_SUCCESS;
    if (ldv % 2 || ldv < (M + 31) / 32 * 32) BEM_FAIL("rank_update needs an even ldv >= roundup(M, 32)");
    dim3 grid((unsigned)((M + 127) / 128), (unsigned)((nc + 31) / 32));
    rank_update_mfma<2><<<grid, 256, 0, s>>>(V, ldv, M, W, ldw, nc, C, ldc);
    HIP_CALL(hipGetLastError());
    return EXIT_SUCCESS;
}

// ------------------------------------------------- FP64-matrix residuals ---
// Residuals and adjoint products for an FP64 matrix (the all-FP64 path) and
// several right-hand sides: VALU FP64 kernels, 8 right-hand sides per pass.
template <typename TA>
struct A2T;
template <>
struct A2T<float> {
    using c2 = float2;
};
template <>
struct A2T<double> {
    using c2 = double2;
};

constexpr int kRG = 8;

// part[(chunk * R + r) * ldp + i] = sum_{j in chunk} A[i, j] X[j, r] for r in [rg0, rg0 + 8)
template <typename TA>
__global__ __launch_bounds__(256) void mres_part(const typename A2T<TA>::c2 *__restrict__ A, long lda, int M, int N,
                                                 const double2 *__restrict__ X, long ldx, int R, int rg0, int kc,
                                                 double2 *__restrict__ part, long ldp) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int chunk = blockIdx.y;
    const int j0 = chunk * kc, j1 = min(N, j0 + kc);
    const int nr = min(kRG, R - rg0);
    double ar[kRG], ai[kRG];
#pragma unroll
    for (int q = 0; q < kRG; q++) {
        ar[q] = 0;
        ai[q] = 0;
    }
    if (i < M) {
        for (int j = j0; j < j1; j++) {
            const typename A2T<TA>::c2 a = A[(size_t)j * lda + i];
            const double xr0 = (double)a.x, xi0 = (double)a.y;
#pragma unroll
            for (int q = 0; q < kRG; q++) {
                if (q < nr) {
                    const double2 x = X[(size_t)(rg0 + q) * ldx + j];
                    ar[q] = fma(xr0, x.x, fma(-xi0, x.y, ar[q]));
                    ai[q] = fma(xr0, x.y, fma(xi0, x.x, ai[q]));
                }
            }
        }
#pragma unroll
        for (int q = 0; q < kRG; q++)
            if (q < nr) part[((size_t)chunk * R + rg0 + q) * ldp + i] = make_double2(ar[q], ai[q]);
    }
}

// out[r][i] = b[r][i] - sum_chunks part (b == nullptr: out = +sum)
__global__ __launch_bounds__(256) void mres_reduce(const double2 *__restrict__ part, long ldp, int nchunks, int R, int M,
                                                   const double2 *__restrict__ b, long ldb, double2 *__restrict__ out,
                                                   long ldo) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int r = blockIdx.y;
    if (i >= M) return;
    double sr = 0, si = 0;
    // 8 partials in flight per step, added in chunk order
    int c = 0;
    for (; c + 8 <= nchunks; c += 8) {
        double2 p[8];
#pragma unroll
        for (int u = 0; u < 8; u++) p[u] = part[((size_t)(c + u) * R + r) * ldp + i];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            sr += p[u].x;
            si += p[u].y;
        }
    }
    for (; c < nchunks; c++) {
        const double2 p = part[((size_t)c * R + r) * ldp + i];
        sr += p.x;
        si += p.y;
    }
    if (b) {
        const double2 bb = b[(size_t)r * ldb + i];
        out[(size_t)r * ldo + i] = make_double2(bb.x - sr, bb.y - si);
    } else {
        out[(size_t)r * ldo + i] = make_double2(sr, si);
    }
}

// Z[j][r] = sum_i conj(A[i, j]) T[i][r], one wave per column, 8 right-hand sides per pass
template <typename TA>
__global__ __launch_bounds__(256) void mh_kernel(const typename A2T<TA>::c2 *__restrict__ A, long lda, int M, int N,
                                                 const double2 *__restrict__ T, long ldt, int R, int rg0,
                                                 double2 *__restrict__ Z, long ldz) {
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= N) return;
    const int nr = min(kRG, R - rg0);
    double zr[kRG], zi[kRG];
#pragma unroll
    for (int q = 0; q < kRG; q++) {
        zr[q] = 0;
        zi[q] = 0;
    }
    const typename A2T<TA>::c2 *col = A + (size_t)j * lda;
    for (int i = lane; i < M; i += 64) {
        const double ar = (double)col[i].x, ai = (double)col[i].y;
#pragma unroll
        for (int q = 0; q < kRG; q++) {
            if (q < nr) {
                const double2 t = T[(size_t)(rg0 + q) * ldt + i];
                zr[q] = fma(ar, t.x, fma(ai, t.y, zr[q]));
                zi[q] = fma(ar, t.y, fma(-ai, t.x, zi[q]));
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kRG; q++) {
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            zr[q] += __shfl_xor(zr[q], m, 64);
            zi[q] += __shfl_xor(zi[q], m, 64);
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < kRG; q++)
            if (q < nr) Z[(size_t)(rg0 + q) * ldz + j] = make_double2(zr[q], zi[q]);
    }
}

static void mres_shape(long M, long N, int &nch, int &kc) {
    const long tiles = (M + 255) / 256;
    long target = std::max(1L, 2048 / std::max(1L, tiles));
    kc = (int)std::max(16L, (N + target - 1) / target);
    nch = (int)((N + kc - 1) / kc);
}

size_t ws_mres(long M, long N, int R) {
    int nch, kc;
    mres_shape(M, N, nch, kc);
    return (size_t)nch * (size_t)((M + 63) / 64 * 64) * (size_t)std::max(1, R) * sizeof(double2);
}

// out (M x R) = b - A X  (b == nullptr: out = A X), FP64 accumulation
template <typename TA>
static int mres_impl(const typename A2T<TA>::c2 *A, long lda, long M, long N, const double2 *X, long ldx,
                     const double2 *b, long ldb, double2 *out, long ldo, int R, void *ws, hipStream_t s) {
    if (M <= 0 || N <= 0 || R <= 0) return EXIT_SUCCESS;
    int nch, kc;
    mres_shape(M, N, nch, kc);
    const long ldp = (M + 63) / 64 * 64;
    double2 *part = (double2 *)ws;
    TimedRegion tr;
    timing_begin(tr, "mres_f64acc", s);
    for (int rg = 0; rg < R; rg += kRG) {
        dim3 grid((unsigned)((M + 255) / 256), nch);
        mres_part<TA><<<grid, 256, 0, s>>>(A, lda, (int)M, (int)N, X, ldx, R, rg, kc, part, ldp);
        HIP_CALL(hipGetLastError());
    }
    timing_end(tr, s);
    dim3 g2((unsigned)((M + 255) / 256), R);
    mres_reduce<<<g2, 256, 0, s>>>(part, ldp, nch, R, (int)M, b, ldb, out, ldo);
    HIP_CALL(hipGetLastError());
    return EXIT_SUCCESS;
}

// Z (N x R) = A^H T, FP64 accumulation
template <typename TA>
static int mh_impl(const typename A2T<TA>::c2 *A, long lda, long M, long N, const double2 *T, long ldt, double2 *Z,
                   long ldz, int R, hipStream_t s) {
    if (M <= 0 || N <= 0 || R <= 0) return EXIT_SUCCESS;
    TimedRegion tr;
    timing_begin(tr, "mh_f64acc", s);
    for (int rg = 0; rg < R; rg += kRG) {
        mh_kernel<TA><<<(unsigned)((N + 3) / 4), 256, 0, s>>>(A, lda, (int)M, (int)N, T, ldt, R, rg, Z, ldz);
        HIP_CALL(hipGetLastError());
    }
    timing_end(tr, s);
    return EXIT_SUCCESS;
}

int mres_f64(const double2 *A, long lda, long M, long N, const double2 *X, long ldx, const double2 *b, long ldb,
             double2 *out, long ldo, int R, void *ws, hipStream_t s) {
    return mres_impl<double>(A, lda, M, N, X, ldx, b, ldb, out, ldo, R, ws, s);
}
int mh_f64(const double2 *A, long lda, long M, long N, const double2 *T, long ldt, double2 *Z, long ldz, int R,
           hipStream_t s) {
    return mh_impl<double>(A, lda, M, N, T, ldt, Z, ldz, R, s);
}

}  // namespace bem
