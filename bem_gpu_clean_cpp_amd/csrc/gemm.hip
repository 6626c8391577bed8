// Multi-right-hand-side complex matrix products on MFMA (v_mfma_f32_16x16x4f32):
// Y = A X (M x R) and Z = A^H T (N x R), A complex64 column-major (M x N, lda).
//
// The interleaved memory of A is read as a real (2M x N) matrix Ar~ whose even
// rows hold Re A and odd rows Im A, so one 16-byte load per lane gives 4 real
// rows of one column and feeds 4 MFMAs (one per element).  For Y = A X the
// products P = Ar~ Re(X) and Q = Ar~ Im(X) land in the same lane and register
// for the real and imaginary row of each complex row, so Y = (P_re - Q_im,
// P_im + Q_re) combines lane-locally.  For Z = A^H T the second operand is T's
// own interleaved memory (and a swapped, negated copy for Im Z).  Columns (or
// rows, for A^H) are split into chunks over the grid; chunk partials are summed
// in a fixed order by gemm_reduce (deterministic, no atomics).
//
// MFMA lane layout (16x16x4, one block): A operand A[i][k] from lane i + 16 k,
// B operand B[k][j] from lane j + 16 k, result D[4 (l / 16) + v][l % 16] in
// lane l, register v.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bem_internal.hpp"

namespace bem {

typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4f mfma4(float a, float b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// part[(chunk * R + n) * ldp + i] = sum_{j in chunk} A[i, j] X[j, n]
// wave: 32 complex rows x (16 G) right-hand sides x one column chunk;
// requires lda even and rows row0..row0+31 inside the allocation (lda >= roundup(M, 32)).
template <int G>
__global__ __launch_bounds__(256) void gemm_n_mfma(const float2 *__restrict__ A, long lda, int M, int N,
                                                   const float2 *__restrict__ X, long ldx, int R, int rg0, int kc,
                                                   float2 *__restrict__ part, long ldp) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row0 = (blockIdx.x * 4 + wave) * 32;
    if (row0 >= M) return;
    const int chunk = blockIdx.y;
    const int j0 = chunk * kc, j1 = min(N, j0 + kc);
    const int m = lane & 15, kk = lane >> 4;
    v4f acc[G][2][4];
#pragma unroll
    for (int g = 0; g < G; g++)
#pragma unroll
        for (int p = 0; p < 2; p++)
#pragma unroll
            for (int t = 0; t < 4; t++) acc[g][p][t] = (v4f){0.f, 0.f, 0.f, 0.f};
    const float4 *Ab = reinterpret_cast<const float4 *>(A + row0 + 2 * m);
    for (int j = j0; j < j1; j += 4) {
        const int J = j + kk;
        float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
        if (J < j1) a = Ab[(size_t)J * (lda / 2)];
#pragma unroll
        for (int g = 0; g < G; g++) {
            const int n = rg0 + g * 16 + m;
            float2 xv = make_float2(0.f, 0.f);
            if (J < j1 && n < R) xv = X[(size_t)n * ldx + J];
            const float av[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
            for (int t = 0; t < 4; t++) {
                acc[g][0][t] = mfma4(av[t], xv.x, acc[g][0][t]);
                acc[g][1][t] = mfma4(av[t], xv.y, acc[g][1][t]);
            }
        }
    }
    // D_t[m'][n'] (m' = 4 (lane / 16) + v) is real row 4 m' + t of the tile:
    // complex row 2 m' + t / 2, real part for even t
#pragma unroll
    for (int g = 0; g < G; g++) {
        const int n = rg0 + g * 16 + m;
        if (n >= R) continue;
        float2 *pc = part + ((size_t)chunk * R + n) * ldp;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int mp = 4 * kk + v;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int i = row0 + 2 * mp + h;
                if (i < M) {
                    const float yr = acc[g][0][2 * h][v] - acc[g][1][2 * h + 1][v];
                    const float yi = acc[g][0][2 * h + 1][v] + acc[g][1][2 * h][v];
                    pc[i] = make_float2(yr, yi);
                }
            }
        }
    }
}

// part[(chunk * R + n) * ldp + j] = sum_{i in chunk} conj(A[i, j]) T[i, n]
// wave: 16 columns of A x (16 G) right-hand sides x one row chunk (multiple of 8).
template <int G>
__global__ __launch_bounds__(256) void gemm_h_mfma(const float2 *__restrict__ A, long lda, int M, int N,
                                                   const float2 *__restrict__ T, long ldt, int R, int rg0, int kr,
                                                   float2 *__restrict__ part, long ldp) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c0 = (blockIdx.x * 4 + wave) * 16;
    if (c0 >= N) return;
    const int chunk = blockIdx.y;
    const int i0 = chunk * kr, i1 = min(M, i0 + kr);
    const int m = lane & 15, kk = lane >> 4;
    const int J = c0 + m;
    v4f accR[G], accI[G];
#pragma unroll
    for (int g = 0; g < G; g++) {
        accR[g] = (v4f){0.f, 0.f, 0.f, 0.f};
        accI[g] = (v4f){0.f, 0.f, 0.f, 0.f};
    }
    for (int i = i0; i < i1; i += 8) {
        const int row = i + 2 * kk;  // complex rows row, row + 1 of this lane
        float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
        if (J < N && row < i1) {
            a = *reinterpret_cast<const float4 *>(A + (size_t)J * lda + row);
            if (row + 1 >= i1) {
                a.z = 0.f;
                a.w = 0.f;
            }
        }
        const float av[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
        for (int g = 0; g < G; g++) {
            const int n = rg0 + g * 16 + m;
            float4 tt = make_float4(0.f, 0.f, 0.f, 0.f);
            if (n < R && row < i1) {
                tt = *reinterpret_cast<const float4 *>(T + (size_t)n * ldt + row);
                if (row + 1 >= i1) {
                    tt.z = 0.f;
                    tt.w = 0.f;
                }
            }
            // Re Z: sum Ar Tr + Ai Ti ; Im Z: sum Ar Ti - Ai Tr
            const float br[4] = {tt.x, tt.y, tt.z, tt.w};
            const float bi[4] = {tt.y, -tt.x, tt.w, -tt.z};
#pragma unroll
            for (int t = 0; t < 4; t++) {
                accR[g] = mfma4(av[t], br[t], accR[g]);
                accI[g] = mfma4(av[t], bi[t], accI[g]);
            }
        }
    }
#pragma unroll
    for (int g = 0; g < G; g++) {
        const int n = rg0 + g * 16 + m;
        if (n >= R) continue;
        float2 *pc = part + ((size_t)chunk * R + n) * ldp;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int j = c0 + 4 * kk + v;
            if (j < N) pc[j] = make_float2(accR[g][v], accI[g][v]);
        }
    }
}

// Y[n][i] = sum_chunks part[(chunk * R + n) * ldp + i]  (fixed order)
__global__ __launch_bounds__(256) void gemm_reduce(const float2 *__restrict__ part, long ldp, int nchunks, int R,
                                                   int M, float2 *__restrict__ Y, long ldy) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int n = blockIdx.y;
    if (i >= M) return;
    float sr = 0.f, si = 0.f;
    for (int c = 0; c < nchunks; c++) {
        const float2 p = part[((size_t)c * R + n) * ldp + i];
        sr += p.x;
        si += p.y;
    }
    Y[(size_t)n * ldy + i] = make_float2(sr, si);
}

static void gemm_n_shape(long M, long N, int &nch, int &kc) {
    const long tiles = (M + 31) / 32;
    long target = std::max(1L, 4096 / std::max(1L, tiles));
    kc = (int)std::max(16L, ((N + target - 1) / target + 3) / 4 * 4);
    nch = (int)((N + kc - 1) / kc);
}

static void gemm_h_shape(long M, long N, int &nch, int &kr) {
    const long tiles = (N + 15) / 16;
    long target = std::max(1L, 4096 / std::max(1L, tiles));
    kr = (int)std::max(64L, ((M + target - 1) / target + 7) / 8 * 8);
    nch = (int)((M + kr - 1) / kr);
}

size_t ws_gemm(long M, long N, int R) {
    int a, b, c, d;
    gemm_n_shape(M, N, a, b);
    gemm_h_shape(M, N, c, d);
    const long ldp = (std::max(M, N) + 63) / 64 * 64;
    return (size_t)std::max(a, c) * (size_t)ldp * (size_t)std::max(1, R) * sizeof(float2);
}

// Y (M x R, ldy) = A X ; ws >= ws_gemm(M, N, R)
int gemm_n_f32(const float2 *A, long lda, long M, long N, const float2 *X, long ldx, float2 *Y, long ldy, int R,
               void *ws, hipStream_t s) {
    if (M <= 0 || N <= 0 || R <= 0) return EXIT_SUCCESS;
    if (lda % 2 || lda < (M + 31) / 32 * 32) BEM_FAIL("gemm_n_f32 needs an even lda >= roundup(M, 32)");
    int nch, kc;
    gemm_n_shape(M, N, nch, kc);
    const long ldp = (M + 63) / 64 * 64;
    float2 *part = (float2 *)ws;
    TimedRegion tr;
    timing_begin(tr, "gemm_n_f32", s);
    const unsigned gx = (unsigned)((M + 127) / 128);
    for (int rg = 0; rg < R; rg += 32) {
        dim3 grid(gx, nch);
        if (R - rg > 16) gemm_n_mfma<2><<<grid, 256, 0, s>>>(A, lda, (int)M, (int)N, X, ldx, R, rg, kc, part, ldp);
        else gemm_n_mfma<1><<<grid, 256, 0, s>>>(A, lda, (int)M, (int)N, X, ldx, R, rg, kc, part, ldp);
        HIP_CALL(hipGetLastError());
    }
    timing_end(tr, s);
    dim3 g2((unsigned)((M + 255) / 256), R);
    gemm_reduce<<<g2, 256, 0, s>>>(part, ldp, nch, R, (int)M, Y, ldy);
    HIP_CALL(hipGetLastError());
    return EXIT_SUCCESS;
}

// Z (N x R, ldz) = A^H T, T: M x R (ldt even); ws >= ws_gemm(M, N, R)
int gemm_h_f32(const float2 *A, long lda, long M, long N, const float2 *T, long ldt, float2 *Z, long ldz, int R,
               void *ws, hipStream_t s) {
    if (M <= 0 || N <= 0 || R <= 0) return EXIT_SUCCESS;
    if (lda % 2 || ldt % 2) BEM_FAIL("gemm_h_f32 needs even lda and ldt");
    int nch, kr;
    gemm_h_shape(M, N, nch, kr);
    const long ldp = (N + 63) / 64 * 64;
    float2 *part = (float2 *)ws;
    TimedRegion tr;
    timing_begin(tr, "gemm_h_f32", s);
    const unsigned gx = (unsigned)((N + 63) / 64);
    for (int rg = 0; rg < R; rg += 32) {
        dim3 grid(gx, nch);
        if (R - rg > 16) gemm_h_mfma<2><<<grid, 256, 0, s>>>(A, lda, (int)M, (int)N, T, ldt, R, rg, kr, part, ldp);
        else gemm_h_mfma<1><<<grid, 256, 0, s>>>(A, lda, (int)M, (int)N, T, ldt, R, rg, kr, part, ldp);
        HIP_CALL(hipGetLastError());
    }
    timing_end(tr, s);
    dim3 g2((unsigned)((N + 255) / 256), R);
    gemm_reduce<<<g2, 256, 0, s>>>(part, ldp, nch, R, (int)N, Z, ldz);
    HIP_CALL(hipGetLastError());
    return EXIT_SUCCESS;
}

// ------------------------------------------------- FP64-accumulated residuals ---
// The refinement residuals for many right-hand sides: FP32 (or FP64) A, FP64
// vectors, FP64 accumulation.  VALU kernels, 8 right-hand sides per pass.
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
    for (int c = 0; c < nchunks; c++) {
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

int mres_f32a(const float2 *A, long lda, long M, long N, const double2 *X, long ldx, const double2 *b, long ldb,
              double2 *out, long ldo, int R, void *ws, hipStream_t s) {
    return mres_impl<float>(A, lda, M, N, X, ldx, b, ldb, out, ldo, R, ws, s);
}
int mres_f64(const double2 *A, long lda, long M, long N, const double2 *X, long ldx, const double2 *b, long ldb,
             double2 *out, long ldo, int R, void *ws, hipStream_t s) {
    return mres_impl<double>(A, lda, M, N, X, ldx, b, ldb, out, ldo, R, ws, s);
}
int mh_f32a(const float2 *A, long lda, long M, long N, const double2 *T, long ldt, double2 *Z, long ldz, int R,
            hipStream_t s) {
    return mh_impl<float>(A, lda, M, N, T, ldt, Z, ldz, R, s);
}
int mh_f64(const double2 *A, long lda, long M, long N, const double2 *T, long ldt, double2 *Z, long ldz, int R,
           hipStream_t s) {
    return mh_impl<double>(A, lda, M, N, T, ldt, Z, ldz, R, s);
}

}  // namespace bem
