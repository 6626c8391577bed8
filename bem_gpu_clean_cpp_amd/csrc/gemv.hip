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

}  // namespace bem
