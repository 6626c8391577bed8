// Complex GEMV kernels for column-major A (HBM-bound, one pass over A).
//
//   gemv_n: y = A x     lanes own pairs of consecutive rows (one 16-byte load
//                       per lane per column, 1 KiB per wave per column);
//                       columns are split into chunks over the grid and the
//                       chunk partials are summed in a fixed order (no atomics).
//   gemv_h: y = A^H x   one wave per column, 16-byte loads down the column,
//                       wave64 butterfly reduction.
// Accumulation type is a template parameter: FP32 A with FP64 accumulation is
// the FP64 refinement residual; FP64 A is the plain FP64 path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include "bem_internal.hpp"

namespace bem {

template <typename T>
struct CT;
template <>
struct CT<float> {
    using c2 = float2;
    using v4 = float4;
};
template <>
struct CT<double> {
    using c2 = double2;
    using v4 = double4;
};

// ----------------------------------------------------------------- gemv_n ---
// part[chunk * ldp + i] = sum_{j in chunk} A[i, j] x[j]
template <typename TA, typename TX, typename TC>
__global__ __launch_bounds__(256) void gemv_n_part(const typename CT<TA>::c2 *__restrict__ A, long lda, int M,
                                                   int N, const typename CT<TX>::c2 *__restrict__ x, int kc,
                                                   typename CT<TC>::c2 *__restrict__ part, long ldp) {
    using A2 = typename CT<TA>::c2;
    using C2 = typename CT<TC>::c2;
    const int i0 = (blockIdx.x * blockDim.x + threadIdx.x) * 2;
    const int chunk = blockIdx.y;
    const int j0 = chunk * kc, j1 = min(N, j0 + kc);
    TC ar0 = 0, ai0 = 0, ar1 = 0, ai1 = 0;
    if (i0 + 1 < M) {
        const A2 *col = A + (size_t)j0 * lda + i0;
        int j = j0;
        for (; j + 8 <= j1; j += 8) {
            A2 a[8][2];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                a[u][0] = col[(size_t)u * lda];
                a[u][1] = col[(size_t)u * lda + 1];
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const typename CT<TX>::c2 xv = x[j + u];
                const TC xr = (TC)xv.x, xi = (TC)xv.y;
                ar0 = fma((TC)a[u][0].x, xr, ar0); ar0 = fma(-(TC)a[u][0].y, xi, ar0);
                ai0 = fma((TC)a[u][0].x, xi, ai0); ai0 = fma((TC)a[u][0].y, xr, ai0);
                ar1 = fma((TC)a[u][1].x, xr, ar1); ar1 = fma(-(TC)a[u][1].y, xi, ar1);
                ai1 = fma((TC)a[u][1].x, xi, ai1); ai1 = fma((TC)a[u][1].y, xr, ai1);
            }
            col += 8 * (size_t)lda;
        }
        for (; j < j1; j++) {
            const A2 a0 = col[0], a1 = col[1];
            const typename CT<TX>::c2 xv = x[j];
            const TC xr = (TC)xv.x, xi = (TC)xv.y;
            ar0 = fma((TC)a0.x, xr, ar0); ar0 = fma(-(TC)a0.y, xi, ar0);
            ai0 = fma((TC)a0.x, xi, ai0); ai0 = fma((TC)a0.y, xr, ai0);
            ar1 = fma((TC)a1.x, xr, ar1); ar1 = fma(-(TC)a1.y, xi, ar1);
            ai1 = fma((TC)a1.x, xi, ai1); ai1 = fma((TC)a1.y, xr, ai1);
            col += lda;
        }
        C2 r0, r1;
        r0.x = ar0; r0.y = ai0; r1.x = ar1; r1.y = ai1;
        part[(size_t)chunk * ldp + i0] = r0;
        part[(size_t)chunk * ldp + i0 + 1] = r1;
    } else if (i0 < M) {
        const A2 *col = A + (size_t)j0 * lda + i0;
        for (int j = j0; j < j1; j++) {
            const A2 a0 = col[0];
            const typename CT<TX>::c2 xv = x[j];
            const TC xr = (TC)xv.x, xi = (TC)xv.y;
            ar0 = fma((TC)a0.x, xr, ar0); ar0 = fma(-(TC)a0.y, xi, ar0);
            ai0 = fma((TC)a0.x, xi, ai0); ai0 = fma((TC)a0.y, xr, ai0);
            col += lda;
        }
        C2 r0;
        r0.x = ar0; r0.y = ai0;
        part[(size_t)chunk * ldp + i0] = r0;
    }
}

// y = alpha * sum_chunks part (+ b): alpha is +-1 here, used for y = A x
// (alpha 1, no b) and r = b - A x (alpha -1, b).  A workgroup handles 64 rows;
// its 4 waves sum interleaved quarters of the chunks, then combine in LDS in a
// fixed order (deterministic).
template <typename TC, typename TY>
__global__ __launch_bounds__(256) void gemv_n_reduce(const typename CT<TC>::c2 *__restrict__ part, long ldp,
                                                     int nchunks, int M, const typename CT<TY>::c2 *__restrict__ b,
                                                     typename CT<TY>::c2 *__restrict__ y, TC alpha) {
    __shared__ TC red[2][4][64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    TC sr = 0, si = 0;
    if (i < M) {
        int c = wid;
        for (; c + 12 < nchunks; c += 16) {
            const typename CT<TC>::c2 p0 = part[(size_t)c * ldp + i], p1 = part[(size_t)(c + 4) * ldp + i],
                                      p2 = part[(size_t)(c + 8) * ldp + i], p3 = part[(size_t)(c + 12) * ldp + i];
            sr += (p0.x + p1.x) + (p2.x + p3.x);
            si += (p0.y + p1.y) + (p2.y + p3.y);
        }
        for (; c < nchunks; c += 4) {
            const typename CT<TC>::c2 p = part[(size_t)c * ldp + i];
            sr += p.x;
            si += p.y;
        }
    }
    red[0][wid][lane] = sr;
    red[1][wid][lane] = si;
    __syncthreads();
    if (wid == 0 && i < M) {
        sr = (red[0][0][lane] + red[0][1][lane]) + (red[0][2][lane] + red[0][3][lane]);
        si = (red[1][0][lane] + red[1][1][lane]) + (red[1][2][lane] + red[1][3][lane]);
        typename CT<TY>::c2 o;
        if (b) {
            o.x = (TY)((TC)b[i].x + alpha * sr);
            o.y = (TY)((TC)b[i].y + alpha * si);
        } else {
            o.x = (TY)(alpha * sr);
            o.y = (TY)(alpha * si);
        }
        y[i] = o;
    }
}

static void gemv_n_shape(long M, long N, int &nchunks, int &kc) {
    // aim for ~2048 workgroups of 512 rows
    const long rowblocks = (M + 511) / 512;
    long target = std::max(1L, 2048 / std::max(1L, rowblocks));
    kc = (int)std::max(8L, (N + target - 1) / target);
    nchunks = (int)((N + kc - 1) / kc);
}

size_t ws_gemv_n(long M, long N, int R) {
    int nch, kc;
    gemv_n_shape(M, N, nch, kc);
    return (size_t)nch * (size_t)((M + 63) / 64 * 64) * sizeof(double2) * (size_t)std::max(1, R);
}

template <typename TA, typename TX, typename TC, typename TY>
static int gemv_n_impl(const typename CT<TA>::c2 *A, long lda, long M, long N, const typename CT<TX>::c2 *x,
                       long ldx, const typename CT<TY>::c2 *b, long ldb, typename CT<TY>::c2 *y, long ldy, int R,
                       void *ws, hipStream_t s, const char *tname) {
    if (M <= 0 || N <= 0) return EXIT_SUCCESS;
    int nch, kc;
    gemv_n_shape(M, N, nch, kc);
    const long ldp = (M + 63) / 64 * 64;
    auto *part = (typename CT<TC>::c2 *)ws;
    for (int r = 0; r < R; r++) {
        TimedRegion tr;
        timing_begin(tr, tname, s);
        dim3 grid((unsigned)((M + 511) / 512), (unsigned)nch);
        gemv_n_part<TA, TX, TC><<<grid, 256, 0, s>>>(A, lda, (int)M, (int)N, x + (size_t)r * ldx, kc, part, ldp);
        HIP_CALL(hipGetLastError());
        timing_end(tr, s);
        gemv_n_reduce<TC, TY><<<(unsigned)((M + 63) / 64), 256, 0, s>>>(
            part, ldp, nch, (int)M, b ? b + (size_t)r * ldb : nullptr, y + (size_t)r * ldy, b ? (TC)-1 : (TC)1);
        HIP_CALL(hipGetLastError());
    }
    return EXIT_SUCCESS;
}

int gemv_n_f32(const float2 *A, long lda, long M, long N, const float2 *x, long ldx, float2 *y, long ldy, int R,
               void *ws, hipStream_t s) {
    return gemv_n_impl<float, float, float, float>(A, lda, M, N, x, ldx, nullptr, 0, y, ldy, R, ws, s, "gemv_n_f32");
}
int resid_n_f32a(const float2 *A, long lda, long M, long N, const double2 *x, long ldx, const double2 *b, long ldb,
                 double2 *r, long ldr, int R, void *ws, hipStream_t s) {
    return gemv_n_impl<float, double, double, double>(A, lda, M, N, x, ldx, b, ldb, r, ldr, R, ws, s, "resid_n_f32a");
}
int gemv_n_f32a(const float2 *A, long lda, long M, long N, const double2 *x, long ldx, double2 *y, long ldy, int R,
                void *ws, hipStream_t s) {
    return gemv_n_impl<float, double, double, double>(A, lda, M, N, x, ldx, nullptr, 0, y, ldy, R, ws, s,
                                                      "resid_n_f32a");
}
int gemv_n_f64(const double2 *A, long lda, long M, long N, const double2 *x, long ldx, double2 *y, long ldy, int R,
               void *ws, hipStream_t s) {
    return gemv_n_impl<double, double, double, double>(A, lda, M, N, x, ldx, nullptr, 0, y, ldy, R, ws, s,
                                                       "gemv_n_f64");
}
int resid_n_f64(const double2 *A, long lda, long M, long N, const double2 *x, long ldx, const double2 *b, long ldb,
                double2 *r, long ldr, int R, void *ws, hipStream_t s) {
    return gemv_n_impl<double, double, double, double>(A, lda, M, N, x, ldx, b, ldb, r, ldr, R, ws, s,
                                                       "gemv_n_f64");
}

// ----------------------------------------------------------------- gemv_h ---
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// y[j] = sum_i conj(A[i, j]) x[i] (optionally y = b - that, for FP64 residuals)
template <typename TA, typename TX, typename TC, typename TY>
__global__ __launch_bounds__(256) void gemv_h_kernel(const typename CT<TA>::c2 *__restrict__ A, long lda, int M,
                                                     int N, const typename CT<TX>::c2 *__restrict__ x,
                                                     typename CT<TY>::c2 *__restrict__ y) {
    using A2 = typename CT<TA>::c2;
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= N) return;
    const A2 *col = A + (size_t)j * lda;
    TC sr = 0, si = 0;
    int i = lane * 2;
    for (; i + 384 + 1 < M; i += 512) {
        A2 a[4][2];
        typename CT<TX>::c2 xv[4][2];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            a[u][0] = col[i + 128 * u];
            a[u][1] = col[i + 128 * u + 1];
            xv[u][0] = x[i + 128 * u];
            xv[u][1] = x[i + 128 * u + 1];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
#pragma unroll
            for (int q = 0; q < 2; q++) {
                sr = fma((TC)a[u][q].x, (TC)xv[u][q].x, sr); sr = fma((TC)a[u][q].y, (TC)xv[u][q].y, sr);
                si = fma((TC)a[u][q].x, (TC)xv[u][q].y, si); si = fma(-(TC)a[u][q].y, (TC)xv[u][q].x, si);
            }
        }
    }
    for (; i + 1 < M; i += 128) {
        const A2 a0 = col[i], a1 = col[i + 1];
        const typename CT<TX>::c2 x0 = x[i], x1 = x[i + 1];
        // conj(a) * x
        sr = fma((TC)a0.x, (TC)x0.x, sr); sr = fma((TC)a0.y, (TC)x0.y, sr);
        si = fma((TC)a0.x, (TC)x0.y, si); si = fma(-(TC)a0.y, (TC)x0.x, si);
        sr = fma((TC)a1.x, (TC)x1.x, sr); sr = fma((TC)a1.y, (TC)x1.y, sr);
        si = fma((TC)a1.x, (TC)x1.y, si); si = fma(-(TC)a1.y, (TC)x1.x, si);
    }
    if (i < M) {
        const A2 a0 = col[i];
        const typename CT<TX>::c2 x0 = x[i];
        sr = fma((TC)a0.x, (TC)x0.x, sr); sr = fma((TC)a0.y, (TC)x0.y, sr);
        si = fma((TC)a0.x, (TC)x0.y, si); si = fma(-(TC)a0.y, (TC)x0.x, si);
    }
    sr = wave_sum(sr);
    si = wave_sum(si);
    if (lane == 0) {
        typename CT<TY>::c2 o;
        o.x = (TY)sr;
        o.y = (TY)si;
        y[j] = o;
    }
}

template <typename TA, typename TX, typename TC, typename TY>
static int gemv_h_impl(const typename CT<TA>::c2 *A, long lda, long M, long N, const typename CT<TX>::c2 *x, long ldx,
                       typename CT<TY>::c2 *y, long ldy, int R, hipStream_t s, const char *tname) {
    if (M <= 0 || N <= 0) return EXIT_SUCCESS;
    for (int r = 0; r < R; r++) {
        TimedRegion tr;
        timing_begin(tr, tname, s);
        gemv_h_kernel<TA, TX, TC, TY><<<(unsigned)((N + 3) / 4), 256, 0, s>>>(A, lda, (int)M, (int)N,
                                                                              x + (size_t)r * ldx, y + (size_t)r * ldy);
        HIP_CALL(hipGetLastError());
        timing_end(tr, s);
    }
    return EXIT_SUCCESS;
}

int gemv_h_f32(const float2 *A, long lda, long M, long N, const float2 *x, long ldx, float2 *y, long ldy, int R,
               hipStream_t s) {
    return gemv_h_impl<float, float, float, float>(A, lda, M, N, x, ldx, y, ldy, R, s, "gemv_h_f32");
}
int gemv_h_f32a(const float2 *A, long lda, long M, long N, const double2 *x, long ldx, double2 *y, long ldy, int R,
                hipStream_t s) {
    return gemv_h_impl<float, double, double, double>(A, lda, M, N, x, ldx, y, ldy, R, s, "gemv_h_f32a");
}
int gemv_h_f64(const double2 *A, long lda, long M, long N, const double2 *x, long ldx, double2 *y, long ldy, int R,
               hipStream_t s) {
    return gemv_h_impl<double, double, double, double>(A, lda, M, N, x, ldx, y, ldy, R, s, "gemv_h_f64");
}

// ------------------------------------------------------ normal equations ---
// w = A^H (A v) in one pass over A, for the CHIEF normal equations with one
// right-hand side.  (A^H A v) = sum_i conj(a_i) (a_i . v) over the rows a_i,
// so with A stored row-major a workgroup can hold a row in registers, form
// t_i = a_i . v with one workgroup reduction and add conj(a_i) t_i to its own
// partial of w before the row leaves the registers.  Column-major A would need
// all of A v before the first dot of A^H, i.e. two passes over A (the pair of
// GEMVs above).  The row-major copy costs one transpose per matrix.
//
//   transpose_rm_ahb_f32  Ar[i * ldr + j] = A[j * lda + i] for j < N, zero
//                     for N <= j < ldr (the padding is read by the pair
//                     loads), and A^H b from the same pass;
//   ne_apply_part     kNeBlocks workgroups of NT = 1024 threads (4 waves per
//                     SIMD), each over a contiguous range of rows.  Thread
//                     t owns the column pairs t + NT k (16-byte loads, 1 KiB
//                     per wave per row) and its partial of w in registers; v
//                     sits in LDS.  Rows go in batches of RB (2 while a
//                     thread has at most 3 pairs, else 1): the batch's loads
//                     are all in flight together, then one reduction and one
//                     barrier (on LDS only) for all RB dots; the reduction
//                     slots alternate by batch parity;
//   ne_reduce         w = sum of the block partials in a fixed order, in FP64
//                     (16 groups of 64 lanes per 64 columns, then a tree).
constexpr int kNeBlocks = 256;
// column pairs a row may have: 1024 threads x 6 pairs for the product,
// 512 x 12 for the residual (N <= 12288)
constexpr long kNeMaxRowPairs = 6144;

// The row-major copy in 64 x 64 LDS tiles (Ar[i * ldr + j] = A[j * lda + i],
// zero for N <= j < ldr, the padding the pair loads read), also forming the
// partial of A^H b over the tile's 64 rows for its 64 columns (FP64): P[blockIdx.x * ldr + j] = sum over the
// tile's rows i of conj(A[i, j]) b[i].  Thread (tx, ty) takes column j0 + tx
// and the rows ty, ty + 4, ... of the tile; the four row groups are added in
// a fixed order.  The pass over A that the first residual of the normal
// equations (A^H b) needs is the one the copy makes anyway.
__global__ __launch_bounds__(256) void k_transpose_rm_ahb(const float2 *__restrict__ A, long lda, int M, int N,
                                                          float2 *__restrict__ Ar, long ldr,
                                                          const double2 *__restrict__ b,
                                                          double2 *__restrict__ P) {
    __shared__ float2 tile[64][65];
    __shared__ double2 s_b[64];
    __shared__ double2 s_p[4][64];
    const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    // all 16 loads of a thread in flight at once: clamped addresses, zero
    // selected outside the matrix (a conditional load per element serialises
    // them, one wait each)
    double2 bv = make_double2(0, 0);
    if (ty == 0 && i0 + tx < M) bv = b[i0 + tx];
    {
        const int i = i0 + tx, ic = min(i, M - 1);
        float2 v[16];
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int j = j0 + ty + 4 * q;
            const float2 x = A[(size_t)min(j, N - 1) * lda + ic];
            v[q] = (i < M && j < N) ? x : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int q = 0; q < 16; q++) tile[ty + 4 * q][tx] = v[q];
    }
    if (ty == 0) s_b[tx] = bv;
    __syncthreads();
    double pr = 0, pi = 0;
    for (int ii = ty; ii < 64; ii += 4) {
        const int i = i0 + ii, j = j0 + tx;
        const float2 a = tile[tx][ii];
        if (i < M && j < ldr) Ar[(size_t)i * ldr + j] = a;
        const double2 bb = s_b[ii];
        pr = fma((double)a.x, bb.x, fma((double)a.y, bb.y, pr));
        pi = fma((double)a.x, bb.y, fma(-(double)a.y, bb.x, pi));
    }
    s_p[ty][tx] = make_double2(pr, pi);
    __syncthreads();
    if (ty == 0) {
        const double2 p0 = s_p[0][tx], p1 = s_p[1][tx], p2 = s_p[2][tx], p3 = s_p[3][tx];
        P[(size_t)blockIdx.x * ldr + j0 + tx] = make_double2((p0.x + p1.x) + (p2.x + p3.x), (p0.y + p1.y) + (p2.y + p3.y));
    }
}

__device__ __forceinline__ float ne_wsum(float v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// Barrier that publishes this workgroup's LDS writes and nothing else: the
// fences are restricted to LDS, so loads from global memory that are still in
// flight (the next row) stay in flight across it.  (__syncthreads() waits for
// every outstanding global load as well.)
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <int NT, int PPT, int RB>
__global__ __launch_bounds__(NT) void ne_apply_part(const float4 *__restrict__ Ar4, long ldr4, int M,
                                                    int rows_per_blk, const float4 *__restrict__ v4, int N,
                                                    float4 *__restrict__ part4, long ldp4) {
    constexpr int NW = NT / 64;
    // v is read by every row: it lives in LDS (NT * PPT pairs, zero past N),
    // which leaves the registers to RB rows and the partial of w
    __shared__ float4 s_v[NT * PPT];
    __shared__ float s_red[2][NW][2 * RB];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int r0 = blockIdx.x * rows_per_blk, r1 = min(M, r0 + rows_per_blk);
    const int npairs = (N + 1) / 2;
    for (int p = tid; p < NT * PPT; p += NT) {
        float4 vv = p < npairs ? v4[p] : make_float4(0.f, 0.f, 0.f, 0.f);
        if (2 * p + 1 >= N) {
            vv.z = 0.f;
            vv.w = 0.f;
        }
        s_v[p] = vv;
    }
    float4 w[PPT];
    // byte offsets of the thread's pairs inside a row, the same for every row
    unsigned off[PPT];
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        off[k] = (unsigned)min(tid + NT * k, (int)ldr4 - 1) * 16u;
        w[k] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    lds_barrier();  // s_v
    int par = 0;
    for (int i0 = r0; i0 < r1; i0 += RB, par ^= 1) {
        // RB rows in flight at once; rows past the block's range are clamped
        // to a valid row and get t = 0
        float4 a[RB][PPT];
#pragma unroll
        for (int u = 0; u < RB; u++) {
            const char *row = (const char *)(Ar4 + (size_t)min(i0 + u, M - 1) * ldr4);
#pragma unroll
            for (int k = 0; k < PPT; k++) a[u][k] = *(const float4 *)(row + off[k]);
        }
        float pr[RB], pi[RB];
#pragma unroll
        for (int u = 0; u < RB; u++) {
            pr[u] = 0.f;
            pi[u] = 0.f;
        }
#pragma unroll
        for (int k = 0; k < PPT; k++) {
            const float4 vr = s_v[tid + NT * k];
#pragma unroll
            for (int u = 0; u < RB; u++) {
                pr[u] += a[u][k].x * vr.x - a[u][k].y * vr.y + a[u][k].z * vr.z - a[u][k].w * vr.w;
                pi[u] += a[u][k].x * vr.y + a[u][k].y * vr.x + a[u][k].z * vr.w + a[u][k].w * vr.z;
            }
        }
#pragma unroll
        for (int u = 0; u < RB; u++) {
            pr[u] = ne_wsum(pr[u]);
            pi[u] = ne_wsum(pi[u]);
        }
        if (lane == 0) {
#pragma unroll
            for (int u = 0; u < RB; u++) {
                s_red[par][wid][2 * u] = pr[u];
                s_red[par][wid][2 * u + 1] = pi[u];
            }
        }
        lds_barrier();
#pragma unroll
        for (int u = 0; u < RB; u++) {
            float tr = 0.f, ti = 0.f;
#pragma unroll
            for (int q = 0; q < NW; q++) {
                tr += s_red[par][q][2 * u];
                ti += s_red[par][q][2 * u + 1];
            }
            if (i0 + u >= r1) {
                tr = 0.f;
                ti = 0.f;
            }
            // w += conj(a) t
#pragma unroll
            for (int k = 0; k < PPT; k++) {
                w[k].x += a[u][k].x * tr + a[u][k].y * ti;
                w[k].y += a[u][k].x * ti - a[u][k].y * tr;
                w[k].z += a[u][k].z * tr + a[u][k].w * ti;
                w[k].w += a[u][k].z * ti - a[u][k].w * tr;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        const int p = tid + NT * k;
        if (p < npairs) part4[(size_t)blockIdx.x * ldp4 + p] = w[k];
    }
}

// w[c] = sum over the nblk block partials of column c, for 64 columns per
// workgroup: 16 groups of 64 lanes each add every 16th partial (8 loads in
// flight), then the 16 group sums are added in a fixed pairwise tree, so the
// result does not depend on timing.
constexpr int kNeRedGroups = 16;
template <typename P>
__device__ __forceinline__ void ne_reduce_cols(const P *__restrict__ part, long ldp, int nblk, int N,
                                               double2 *__restrict__ w) {
    constexpr int NG = kNeRedGroups;
    __shared__ double2 s_sum[NG][64];
    const int cl = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int c = blockIdx.x * 64 + cl;
    const int cc = min(c, N - 1);
    double sr = 0, si = 0;
    int b = g;
    for (; b + NG * 7 < nblk; b += NG * 8) {
        P q[8];
#pragma unroll
        for (int u = 0; u < 8; u++) q[u] = part[(size_t)(b + NG * u) * ldp + cc];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            sr += q[u].x;
            si += q[u].y;
        }
    }
    for (; b < nblk; b += NG) {
        const P q = part[(size_t)b * ldp + cc];
        sr += q.x;
        si += q.y;
    }
    s_sum[g][cl] = make_double2(sr, si);
    __syncthreads();
    if (g == 0 && c < N) {
        double2 a[NG];
#pragma unroll
        for (int j = 0; j < NG; j++) a[j] = s_sum[j][cl];
#pragma unroll
        for (int h = NG / 2; h >= 1; h >>= 1)
#pragma unroll
            for (int j = 0; j < h; j++) {
                a[j].x += a[j + h].x;
                a[j].y += a[j + h].y;
            }
        w[c] = a[0];
    }
}

__global__ __launch_bounds__(64 * kNeRedGroups) void ne_reduce(const float2 *__restrict__ part, long ldp, int nblk,
                                                                int N, double2 *__restrict__ w) {
    ne_reduce_cols(part, ldp, nblk, N, w);
}

bool ne_fused_fits(long M, long N) { return M > 0 && N > 0 && (N + 1) / 2 <= kNeMaxRowPairs; }
long ne_ldr(long N) { return (N + 63) / 64 * 64; }
static void ne_shape(long M, int &rows_per_blk, int &nblk) {
    rows_per_blk = (int)((M + kNeBlocks - 1) / kNeBlocks);
    nblk = (int)((M + rows_per_blk - 1) / rows_per_blk);
}
size_t ws_ne_apply(long M, long N) {
    int rpb, nblk;
    ne_shape(M, rpb, nblk);
    return sizeof(float2) * (size_t)nblk * (size_t)ne_ldr(N);
}


int ne_apply_f32(const float2 *Ar, long ldr, long M, long N, const float2 *v, double2 *w, void *ws, hipStream_t s) {
    if (!ne_fused_fits(M, N)) BEM_FAIL("fused normal-equation product: N = %ld is over its limit", N);
    if (ldr != ne_ldr(N)) BEM_FAIL("fused normal-equation product: ldr %ld, expected %ld", ldr, ne_ldr(N));
    int rpb, nblk;
    ne_shape(M, rpb, nblk);
    const long npairs = (N + 1) / 2, ldp = ldr;
    TimedRegion tr, tq;
    timing_begin(tr, "ne_apply_f32", s);
    timing_begin(tq, "ne_apply_part", s);
    const float4 *A4 = (const float4 *)Ar, *v4 = (const float4 *)v;
    float4 *p4 = (float4 *)ws;
    const long ldr4 = ldr / 2, ldp4 = ldp / 2;
    // 1024 threads (4 waves per SIMD) with ceil(pairs / 1024) pairs each; two
    // rows per batch while the pairs take few registers, then one (measured at
    // N = 10242: 1024 threads x 6 pairs, one row, 155 us; 768 x 7, two rows,
    // 165 us; two rows at 768 x 7 against three or four: 162, 184, 261 us)
#define NE_LAUNCH(PPT_, RB_) \
    ne_apply_part<1024, PPT_, RB_><<<nblk, 1024, 0, s>>>(A4, ldr4, (int)M, rpb, v4, (int)N, p4, ldp4)
    if (npairs <= 1024) {
        NE_LAUNCH(1, 2);
    } else if (npairs <= 2048) {
        NE_LAUNCH(2, 2);
    } else if (npairs <= 3072) {
        NE_LAUNCH(3, 2);
    } else if (npairs <= 4096) {
        NE_LAUNCH(4, 1);
    } else if (npairs <= 5120) {
        NE_LAUNCH(5, 1);
    } else {
        NE_LAUNCH(6, 1);
    }
#undef NE_LAUNCH
    HIP_CALL(hipGetLastError());
    timing_end(tq, s);
    ne_reduce<<<(unsigned)((N + 63) / 64), 64 * kNeRedGroups, 0, s>>>((const float2 *)ws, ldp, nblk, (int)N, w);
    HIP_CALL(hipGetLastError());
    timing_end(tr, s);
    return EXIT_SUCCESS;
}

// r = A^H (b - A x) with FP64 accumulation in one pass over the row-major copy
// (the FP64 refinement residual of the normal equations; x = null gives
// A^H b).  Same layout as ne_apply_part, one row at a time.  x is read in
// FP64 for every row: at 16 bytes a column it stays in L2, and reading it
// leaves the registers to the row and the FP64 partial of r (x4 is a copy of x
// padded with zeros to the row length, so every pair load is valid).  a_i . x is
// formed in FP64 (FP32 entries widened), t_i = b_i - a_i . x, and the
// partial of r += conj(a_i) t_i in FP64.  The partials are reduced in a fixed
// order in FP64.
__device__ __forceinline__ double ne_wsum_d(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

template <int NT, int PPT>
__global__ __launch_bounds__(NT) void ne_resid_part(const float4 *__restrict__ Ar4, long ldr4, int M, int rows_per_blk,
                                                    const double4 *__restrict__ x4, const double2 *__restrict__ b,
                                                    int N, double4 *__restrict__ part4, long ldp4) {
    constexpr int NW = NT / 64;
    __shared__ double s_red[2][NW][2];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int r0 = blockIdx.x * rows_per_blk, r1 = min(M, r0 + rows_per_blk);
    const int npairs = (N + 1) / 2;
    double4 w[PPT];
    unsigned off[PPT];
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        off[k] = (unsigned)min(tid + NT * k, npairs - 1);
        w[k] = make_double4(0, 0, 0, 0);
    }
    for (int i = r0; i < r1; i++) {
        const float4 *row = Ar4 + (size_t)i * ldr4;
        float4 a[PPT];
#pragma unroll
        for (int k = 0; k < PPT; k++) a[k] = row[off[k]];
        double dr = 0, di = 0;
        if (x4) {
#pragma unroll
            for (int k = 0; k < PPT; k++) {
                const int p = tid + NT * k;
                if (p < npairs) {
                    const double4 xv = x4[p];
                    dr = fma((double)a[k].x, xv.x, dr);
                    dr = fma(-(double)a[k].y, xv.y, dr);
                    dr = fma((double)a[k].z, xv.z, dr);
                    dr = fma(-(double)a[k].w, xv.w, dr);
                    di = fma((double)a[k].x, xv.y, di);
                    di = fma((double)a[k].y, xv.x, di);
                    di = fma((double)a[k].z, xv.w, di);
                    di = fma((double)a[k].w, xv.z, di);
                }
            }
        }
        dr = ne_wsum_d(dr);
        di = ne_wsum_d(di);
        const int par = i & 1;
        if (lane == 0) {
            s_red[par][wid][0] = dr;
            s_red[par][wid][1] = di;
        }
        lds_barrier();
        double tr = 0, ti = 0;
#pragma unroll
        for (int q = 0; q < NW; q++) {
            tr += s_red[par][q][0];
            ti += s_red[par][q][1];
        }
        const double2 bi = b[i];
        tr = bi.x - tr;
        ti = bi.y - ti;
        // r += conj(a) t
#pragma unroll
        for (int k = 0; k < PPT; k++) {
            w[k].x = fma((double)a[k].x, tr, fma((double)a[k].y, ti, w[k].x));
            w[k].y = fma((double)a[k].x, ti, fma(-(double)a[k].y, tr, w[k].y));
            w[k].z = fma((double)a[k].z, tr, fma((double)a[k].w, ti, w[k].z));
            w[k].w = fma((double)a[k].z, ti, fma(-(double)a[k].w, tr, w[k].w));
        }
    }
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        const int p = tid + NT * k;
        if (p < npairs) part4[(size_t)blockIdx.x * ldp4 + p] = w[k];
    }
}

// the same reduction of FP64 block partials
__global__ __launch_bounds__(64 * kNeRedGroups) void ne_reduce_d(const double2 *__restrict__ part, long ldp, int nblk,
                                                                  int N, double2 *__restrict__ w) {
    ne_reduce_cols(part, ldp, nblk, N, w);
}

size_t ws_ne_resid(long M, long N) { return 2 * ws_ne_apply(M, N) + sizeof(double2) * (size_t)ne_ldr(N); }
size_t ws_ne_rhs(long M, long N) { return sizeof(double2) * (size_t)((M + 63) / 64) * (size_t)ne_ldr(N); }

// Ar = the row-major copy of A and r = A^H b (FP64) in one pass over A
int transpose_rm_ahb_f32(const float2 *A, long lda, long M, long N, float2 *Ar, long ldr, const double2 *b,
                         double2 *r, void *ws, hipStream_t s) {
    if (ldr != ne_ldr(N)) BEM_FAIL("row-major copy: ldr %ld, expected %ld", ldr, ne_ldr(N));
    const int nbx = (int)((M + 63) / 64);
    dim3 g((unsigned)nbx, (unsigned)(ldr / 64));
    double2 *P = (double2 *)ws;
    k_transpose_rm_ahb<<<g, 256, 0, s>>>(A, lda, (int)M, (int)N, Ar, ldr, b, P);
    HIP_CALL(hipGetLastError());
    ne_reduce_d<<<(unsigned)((N + 63) / 64), 64 * kNeRedGroups, 0, s>>>(P, ldr, nbx, (int)N, r);
    HIP_CALL(hipGetLastError());
    return EXIT_SUCCESS;
}

// xpad[i] = x[i] for i < n, 0 up to ldp (one launch instead of a copy and a fill)
__global__ void k_pad_copy(const double2 *__restrict__ x, int n, double2 *__restrict__ xpad, int ldp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ldp) xpad[i] = i < n ? x[i] : make_double2(0.0, 0.0);
}

int ne_resid_f32a(const float2 *Ar, long ldr, long M, long N, const double2 *x, const double2 *b, double2 *r, void *ws,
                  hipStream_t s) {
    if (!ne_fused_fits(M, N)) BEM_FAIL("fused normal-equation residual: N = %ld is over its limit", N);
    if (ldr != ne_ldr(N)) BEM_FAIL("fused normal-equation residual: ldr %ld, expected %ld", ldr, ne_ldr(N));
    int rpb, nblk;
    ne_shape(M, rpb, nblk);
    const long npairs = (N + 1) / 2, ldp = ldr;  // part: nblk rows of ldp double2
    TimedRegion tr;
    timing_begin(tr, "ne_resid_f32a", s);
    const float4 *A4 = (const float4 *)Ar;
    double4 *p4 = (double4 *)ws;
    const long ldr4 = ldr / 2, ldp4 = ldp / 2;
    // x padded with zeros to ldr after the partials
    double2 *xpad = nullptr;
    if (x) {
        xpad = (double2 *)((char *)ws + 2 * ws_ne_apply(M, N));
        k_pad_copy<<<(unsigned)((ldr + 255) / 256), 256, 0, s>>>(x, (int)N, xpad, (int)ldr);
        HIP_CALL(hipGetLastError());
    }
#define NE_LAUNCH(NT_, PPT_) \
    ne_resid_part<NT_, PPT_><<<nblk, NT_, 0, s>>>(A4, ldr4, (int)M, rpb, (const double4 *)xpad, b, (int)N, p4, ldp4)
    // FP64 partials of r take 8 registers a pair: past 2 pairs per thread at
    // 1024 threads the rows go to 512-thread workgroups (256 registers a lane)
    if (npairs <= 1024) {
        NE_LAUNCH(1024, 1);
    } else if (npairs <= 2048) {
        NE_LAUNCH(1024, 2);
    } else if (npairs <= 512 * 8) {
        NE_LAUNCH(512, 8);
    } else {
        NE_LAUNCH(512, 12);
    }
#undef NE_LAUNCH
    HIP_CALL(hipGetLastError());
    ne_reduce_d<<<(unsigned)((N + 63) / 64), 64 * kNeRedGroups, 0, s>>>((const double2 *)ws, ldp, nblk, (int)N, r);
    HIP_CALL(hipGetLastError());
    timing_end(tr, s);
    return EXIT_SUCCESS;
}

}  // namespace bem
