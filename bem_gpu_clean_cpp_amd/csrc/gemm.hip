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
#include <array>
#include <climits>
#include <cmath>
#include <map>
#include <mutex>
#include <type_traits>

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
//
// Several matrices at once (grid.z = matrix, astride apart): matrix z takes the
// right-hand sides z rper .. z rper + rper - 1 (rg0 counts within them); a
// single matrix has astride 0 and rper = R.
// (KU: MFMA steps per prefetched block; WV: waves per SIMD it is built for)
template <typename T, int G, bool G3, int KU = kU, int WV = 1>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WV))) void gemm_n_mfma(
    const float2 *__restrict__ A, long lda, int M, int N, const typename MF<T>::c2 *__restrict__ X, long ldx, int R,
    int rg0, int kc, typename MF<T>::c2 *__restrict__ part, long ldp, long astride, int rper) {
    typedef typename MF<T>::v4 v4;
    typedef typename MF<T>::c2 c2;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row0 = (blockIdx.x * 4 + wave) * 32;
    if (row0 >= M) return;
    A += (size_t)blockIdx.z * astride;
    rg0 += blockIdx.z * rper;
    const int rend = min(R, (int)(blockIdx.z + 1) * rper);  // end of this matrix's right-hand sides
    const int chunk = blockIdx.y;
    const int j0 = chunk * kc, j1 = min(N, j0 + kc);
    const int m = lane & 15, kk = lane >> 4;
    constexpr int NA = G3 ? 6 : 8;  // accumulators per group of 16 right-hand sides
    v4 acc[G][NA];
#pragma unroll
    for (int g = 0; g < G; g++)
#pragma unroll
        for (int t = 0; t < NA; t++) acc[g][t] = (v4){0, 0, 0, 0};
    const float4 *Ab = reinterpret_cast<const float4 *>(A + row0 + 2 * m);
    const long lda4 = lda / 2;
    // right-hand sides past this matrix's last read that one (their results
    // are not written)
    const c2 *Xg[G];
#pragma unroll
    for (int g = 0; g < G; g++) Xg[g] = X + (size_t)min(rg0 + g * 16 + m, rend - 1) * ldx;
    struct Buf {
        float4 a[KU];
        c2 x[KU][G];
    };
    // full blocks: every column inside the chunk, no masking; the address is
    // clamped only for the prefetch past the last pair of blocks
    auto load = [&](int jb, Buf &B) {
#pragma unroll
        for (int u = 0; u < KU; u++) {
            const int Jc = min(jb + 4 * u + kk, j1 - 1);
            B.a[u] = Ab[(size_t)Jc * lda4];
#pragma unroll
            for (int g = 0; g < G; g++) B.x[u][g] = Xg[g][Jc];
        }
    };
    auto load_masked = [&](int jb, Buf &B) {
#pragma unroll
        for (int u = 0; u < KU; u++) {
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
        for (int u = 0; u < KU; u++) {
            const T av[4] = {(T)B.a[u].x, (T)B.a[u].y, (T)B.a[u].z, (T)B.a[u].w};
#pragma unroll
            for (int g = 0; g < G; g++) {
                const T xr = B.x[u][g].x, xi = B.x[u][g].y;
                if constexpr (G3) {
                    // acc[g][3 h + 0..2] = P1, P2, P3 of complex row 2 m' + h
                    const T xs = xr + xi;
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        const T ar = av[2 * h], ai = av[2 * h + 1];
                        acc[g][3 * h] = MF<T>::mma(ar, xs, acc[g][3 * h]);
                        acc[g][3 * h + 1] = MF<T>::mma(ar + ai, xi, acc[g][3 * h + 1]);
                        acc[g][3 * h + 2] = MF<T>::mma(ai - ar, xr, acc[g][3 * h + 2]);
                    }
                } else {
                    // acc[g][2 t + 0 / 1]: real row t of the pair against Re X / Im X
#pragma unroll
                    for (int t = 0; t < 4; t++) {
                        acc[g][2 * t] = MF<T>::mma(av[t], xr, acc[g][2 * t]);
                        acc[g][2 * t + 1] = MF<T>::mma(av[t], xi, acc[g][2 * t + 1]);
                    }
                }
            }
        }
    };
    Buf b0, b1;
    const int jpair = j0 + (j1 - j0) / (8 * KU) * (8 * KU);  // end of the paired full blocks
    load(j0, b0);
    for (int jb = j0; jb < jpair; jb += 8 * KU) {
        load(jb + 4 * KU, b1);
        __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the products
        mul(b0);
        load(jb + 8 * KU, b0);
        __builtin_amdgcn_sched_barrier(0);
        mul(b1);
    }
    for (int jb = jpair; jb < j1; jb += 4 * KU) {
        load_masked(jb, b1);
        mul(b1);
    }
    // result row m' = drow(lane / 16, v) of the MFMAs belongs to complex rows
    // 2 m' and 2 m' + 1 of the tile (h = 0, 1)
#pragma unroll
    for (int g = 0; g < G; g++) {
        const int n = rg0 + g * 16 + m;
        if (n >= rend) continue;
        c2 *pc = part + ((size_t)chunk * R + n) * ldp;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int mp = MF<T>::drow(kk, v);
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int i = row0 + 2 * mp + h;
                if (i >= M) continue;
                if constexpr (G3)
                    pc[i] = c2{acc[g][3 * h][v] - acc[g][3 * h + 1][v], acc[g][3 * h][v] + acc[g][3 * h + 2][v]};
                else
                    pc[i] = c2{acc[g][4 * h][v] - acc[g][4 * h + 3][v], acc[g][4 * h + 1][v] + acc[g][4 * h + 2][v]};
            }
        }
    }
}

// part[(chunk * R + n) * ldp + j] = sum_{i in chunk} conj(A[i, j]) T[i, n]
// wave: 16 columns of A x (16 G) right-hand sides x one row chunk (i0 even).
// Rows read in pairs: row + 1 may be M, which needs lda, ldt > M for odd M
// (the callers' leading dimensions are M rounded up to 64).
template <typename T, int G, bool G3>
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
    constexpr int NA = G3 ? 3 : 2;  // accumulators per group of 16 right-hand sides
    v4 acc[G][NA];
#pragma unroll
    for (int g = 0; g < G; g++)
#pragma unroll
        for (int t = 0; t < NA; t++) acc[g][t] = (v4){0, 0, 0, 0};
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
                if constexpr (G3) {
                    // conj(a) t in the three-product form, with a' = conj(a):
                    // P1 = Re a' (Re t + Im t), P2 = (Re a' + Im a') Im t,
                    // P3 = (Im a' - Re a') Re t; Z = (P1 - P2, P1 + P3)
                    const T tr[2] = {B.t0[u][g].x, B.t1[u][g].x}, ti[2] = {B.t0[u][g].y, B.t1[u][g].y};
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        const T ar = av[2 * h], ai = av[2 * h + 1];
                        acc[g][0] = MF<T>::mma(ar, tr[h] + ti[h], acc[g][0]);
                        acc[g][1] = MF<T>::mma(ar - ai, ti[h], acc[g][1]);
                        acc[g][2] = MF<T>::mma(-(ai + ar), tr[h], acc[g][2]);
                    }
                } else {
                    // Re Z: sum Ar Tr + Ai Ti ; Im Z: sum Ar Ti - Ai Tr
                    const T br[4] = {B.t0[u][g].x, B.t0[u][g].y, B.t1[u][g].x, B.t1[u][g].y};
                    const T bi[4] = {B.t0[u][g].y, -B.t0[u][g].x, B.t1[u][g].y, -B.t1[u][g].x};
#pragma unroll
                    for (int t = 0; t < 4; t++) {
                        acc[g][0] = MF<T>::mma(av[t], br[t], acc[g][0]);
                        acc[g][1] = MF<T>::mma(av[t], bi[t], acc[g][1]);
                    }
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
            if (j >= N) continue;
            if constexpr (G3)
                pc[j] = c2{acc[g][0][v] - acc[g][1][v], acc[g][0][v] + acc[g][2][v]};
            else
                pc[j] = c2{acc[g][0][v], acc[g][1][v]};
        }
    }
}

// ------------------------------------------- FP32, 32 right-hand sides ---
// The same products for a group of up to 32 right-hand sides on
// v_mfma_f32_32x32x2f32 (32 x 32 output, 2 steps of k per instruction).
// Against the 16x16x4 kernels a wave covers twice the rows (or columns) for
// the same operand loads, so the right-hand-side block is read half as often
// per flop, and each lane holds its share of the 32 x 32 output in 16
// registers per accumulator.  The operands are prefetched into a ring of
// buffers (each buffer: one or two k steps, one 16-byte load of A per step and
// one load of the right-hand-side block per lane), so several steps of loads
// are in flight while the oldest is multiplied.
//
// The right-hand-side block X is column-major, so one X load touches 32 cache
// lines (one per right-hand side) against 8 for the 1 KB of A.  At
// M = N = 40962, R = 32 the texture addresser was busy 69 % of the cycles and
// the MFMA pipes 62 % (three-product form, one k step per buffer); pairing
// the steps so that each X load serves two of them took the product from
// 3.80 to 3.49 ms.  A row-major X, whose loads are 512 contiguous bytes, had
// measured slower (4.43 against 4.27 ms, four-product form): every wave then
// reads the same X row at the same time.
//
// Output layout of v_mfma_f32_32x32x2f32 (checked on the device with an
// in-bounds one-hot probe): lane l, register v holds D[8 (v / 4) + 4 (l / 32)
// + v % 4][l % 32]; operands A[i][k] from lane i + 32 k, B[k][j] from lane
// j + 32 k.
typedef float v16f __attribute__((ext_vector_type(16)));
constexpr int kRingN = 4;   // Y = A X: buffers in the ring
constexpr int kRingPX = 2;  // (buffers of two k steps, paired X loads)
constexpr int kF32LdsPad = 65;  // float2 per right-hand side in the LDS partials (bank spread)
constexpr int kWavesN = 3;  // waves per SIMD it is built for
constexpr int kRingH = 2;   // Z = A^H T
constexpr int kStepsH = 4;  // k steps per buffer of the adjoint product
constexpr int kWavesH = 2;

__device__ __forceinline__ int d32_row(int lane, int v) { return 8 * (v / 4) + 4 * (lane >> 5) + (v & 3); }

// part[(chunk * R + n) * ldp + i] = sum_{j in chunk} A[i, j] X[j, n] for
// n in [rg0, rg0 + 32): wave = RT tiles of 64 complex rows.  Lane
// (i = l % 32, k = l / 32) loads complex rows 2i, 2i + 1 of each tile in column
// j + k (four real rows per tile) and X[j + k][rg0 + i] (the B operand, one
// load for all RT tiles).  Two forms of the complex product:
//
// G3 = false, four real products per complex multiply-add (8 MFMAs per k
// step): MFMA t accumulates real row 4 i' + t of the tile against Re X (P_t)
// and Im X (Q_t), and
//   Y[2i'] = (P0 - Q1, Q0 + P1), Y[2i' + 1] = (P2 - Q3, Q2 + P3).
//
// G3 = true, three real products (6 MFMAs per k step; the 3M method, Higham,
// "Stability of a method for multiplying complex matrices with three real
// matrix multiplications", SIAM J. Matrix Anal. Appl. 13 (1992)), in the form
// where each part of the result is the sum of two products:
//   P1 = Re A (Re X + Im X), P2 = (Re A + Im A) Im X, P3 = (Im A - Re A) Re X,
//   Re Y = P1 - P2, Im Y = P1 + P3.
// The sums are formed per lane from the loaded operands (3 adds per lane per k
// step); the rounding error stays within a small multiple of the four-product
// form's, normwise (|A| |X| scale).
//
// Needs lda >= roundup(M, 64).
//
// LDSR: a workgroup is one block of 64 RT rows over four consecutive column
// sub-chunks of kc columns, one per wave; the four partials of each row tile
// are summed in LDS in wave order and written as contiguous 64-row columns
// (partial chunk blockIdx.y).  Without it, a workgroup is four row blocks
// over one chunk, each wave writing its own partial.
template <bool G3, bool PX, int RING = kRingN, int WAVES = kWavesN, int RT = 1, bool LDSR = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WAVES, WAVES))) void gemm_n_mfma32(
    const float2 *__restrict__ A, long lda, int M, int N, const float2 *__restrict__ X, long ldx, int R, int rg0,
    int kc, float2 *__restrict__ part, long ldp) {
    // RT row tiles of 64 complex rows per wave share every X load
    const int lane = threadIdx.x & 63;
    const int wave = LDSR ? __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) : threadIdx.x >> 6;
    const int row0 = LDSR ? blockIdx.x * 64 * RT : (blockIdx.x * 4 + wave) * 64 * RT;
    if (row0 >= M) return;  // (LDSR: never, the grid covers M exactly)
    const int chunk = blockIdx.y;
    const int j0 = LDSR ? min(N, (chunk * 4 + wave) * kc) : chunk * kc, j1 = min(N, j0 + kc);
    const int i = lane & 31, k = lane >> 5;
    constexpr int NA = G3 ? 6 : 8;  // accumulators per row tile
    v16f acc[RT][NA];
#pragma unroll
    for (int r = 0; r < RT; r++)
#pragma unroll
        for (int t = 0; t < NA; t++) acc[r][t] = (v16f){0};
    // row tiles past M read the last tile's rows (their results are not written)
    const float4 *Ab[RT];
#pragma unroll
    for (int r = 0; r < RT; r++)
        Ab[r] = reinterpret_cast<const float4 *>(A + min(row0 + 64 * r, (M - 1) / 64 * 64) + 2 * i);
    const long lda4 = lda / 2;
    // right-hand sides past R read column R - 1 (their results are not written)
    const float2 *Xr = X + (size_t)min(rg0 + i, R - 1) * ldx;
    // PX: a buffer holds two k steps, lane k taking columns jb + 2k and
    // jb + 2k + 1, so its X values for both are one 16-byte load (half the X
    // load instructions, each touching the same 32 cache lines)
    constexpr int S = PX ? 2 : 1;  // k steps per buffer
    struct Buf {
        float4 a[S][RT];
        typename std::conditional<PX, float4, float2>::type x;
    };
    const int jx_last = (j1 - 1) & ~1;  // last even column of the chunk
    auto load = [&](int jb, Buf &B) {
        if constexpr (PX) {
#pragma unroll
            for (int t = 0; t < 2; t++) {
                const size_t o = (size_t)min(jb + 2 * k + t, j1 - 1) * lda4;
#pragma unroll
                for (int r = 0; r < RT; r++) B.a[t][r] = Ab[r][o];
            }
            B.x = *reinterpret_cast<const float4 *>(Xr + min(jb + 2 * k, jx_last));
        } else {
            const int Jc = min(jb + k, j1 - 1);
#pragma unroll
            for (int r = 0; r < RT; r++) B.a[0][r] = Ab[r][(size_t)Jc * lda4];
            B.x = Xr[Jc];
        }
    };
    auto load_masked = [&](int jb, Buf &B) {
        if constexpr (PX) {
            bool in[2];
#pragma unroll
            for (int t = 0; t < 2; t++) {
                const int J = jb + 2 * k + t;
                in[t] = J < j1;
                const size_t o = (size_t)min(J, j1 - 1) * lda4;
#pragma unroll
                for (int r = 0; r < RT; r++) B.a[t][r] = keep4(in[t], in[t], Ab[r][o]);
            }
            B.x = keep4(in[0], in[1], *reinterpret_cast<const float4 *>(Xr + min(jb + 2 * k, jx_last)));
        } else {
            const int J = jb + k;
            const bool in = J < j1;
            const int Jc = min(J, j1 - 1);
#pragma unroll
            for (int r = 0; r < RT; r++) B.a[0][r] = keep4(in, in, Ab[r][(size_t)Jc * lda4]);
            B.x = keep2(in, Xr[Jc]);
        }
    };
    auto mul1 = [&](v16f *ac, const float4 a, const float2 x) {
        if constexpr (G3) {
            // ac[3 h + 0..2] = P1, P2, P3 of complex row 2 i' + h
            const float xs = x.x + x.y;
            const float ar[2] = {a.x, a.z}, ai[2] = {a.y, a.w};
#pragma unroll
            for (int h = 0; h < 2; h++) {
                ac[3 * h] = __builtin_amdgcn_mfma_f32_32x32x2f32(ar[h], xs, ac[3 * h], 0, 0, 0);
                ac[3 * h + 1] = __builtin_amdgcn_mfma_f32_32x32x2f32(ar[h] + ai[h], x.y, ac[3 * h + 1], 0, 0, 0);
                ac[3 * h + 2] = __builtin_amdgcn_mfma_f32_32x32x2f32(ai[h] - ar[h], x.x, ac[3 * h + 2], 0, 0, 0);
            }
        } else {
            // ac[2 t + 0 / 1] = P_t / Q_t
            const float av[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
            for (int t = 0; t < 4; t++) {
                ac[2 * t] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[t], x.x, ac[2 * t], 0, 0, 0);
                ac[2 * t + 1] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[t], x.y, ac[2 * t + 1], 0, 0, 0);
            }
        }
    };
    auto mul = [&](const Buf &B) {
#pragma unroll
        for (int r = 0; r < RT; r++) {
            if constexpr (PX) {
                mul1(acc[r], B.a[0][r], make_float2(B.x.x, B.x.y));
                mul1(acc[r], B.a[1][r], make_float2(B.x.z, B.x.w));
            } else {
                mul1(acc[r], B.a[0][r], B.x);
            }
        }
    };
    constexpr int W = 2 * S;  // columns per buffer
    Buf b[RING];
    const int jend = j0 + (j1 - j0) / (W * RING) * (W * RING);  // end of the full rounds of the ring
#pragma unroll
    for (int w = 0; w < RING - 1; w++) load(j0 + W * w, b[w]);
    for (int jb = j0; jb < jend; jb += W * RING) {
#pragma unroll
        for (int w = 0; w < RING; w++) {
            load(jb + W * (w + RING - 1), b[(w + RING - 1) % RING]);
            __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the products
            mul(b[w]);
        }
    }
    for (int jb = jend; jb < j1; jb += W) {
        load_masked(jb, b[0]);
        mul(b[0]);
    }
    if constexpr (LDSR) {
        // 16 right-hand sides at a time through 33 KB of LDS (three workgroups
        // per CU still fit at one row tile a wave)
        __shared__ float2 red[4 * 16 * kF32LdsPad];
#pragma unroll
        for (int rt = 0; rt < RT; rt++) {
#pragma unroll
            for (int hf = 0; hf < 2; hf++) {
                if (rt + hf > 0) __syncthreads();  // the previous sums have read red
                if ((i >> 4) == hf) {
                    float2 *rw = red + (size_t)(wave * 16 + (i & 15)) * kF32LdsPad;
#pragma unroll
                    for (int v = 0; v < 16; v++) {
                        const int d = 2 * d32_row(lane, v);
#pragma unroll
                        for (int h = 0; h < 2; h++) {
                            if constexpr (G3)
                                rw[d + h] = make_float2(acc[rt][3 * h][v] - acc[rt][3 * h + 1][v],
                                                        acc[rt][3 * h][v] + acc[rt][3 * h + 2][v]);
                            else
                                rw[d + h] = make_float2(acc[rt][4 * h][v] - acc[rt][4 * h + 3][v],
                                                        acc[rt][4 * h + 1][v] + acc[rt][4 * h + 2][v]);
                        }
                    }
                }
                __syncthreads();
                const int r = threadIdx.x & 63, rowr = row0 + 64 * rt + r;
                if (rowr < M) {
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const int nl = (threadIdx.x >> 6) + 4 * q, n = rg0 + 16 * hf + nl;
                        if (n >= R) break;
                        float2 sum = red[(size_t)nl * kF32LdsPad + r];
#pragma unroll
                        for (int w = 1; w < 4; w++) {
                            const float2 o = red[(size_t)(w * 16 + nl) * kF32LdsPad + r];
                            sum.x += o.x;
                            sum.y += o.y;
                        }
                        part[((size_t)chunk * R + n) * ldp + rowr] = sum;
                    }
                }
            }
        }
        return;
    }
    const int n = rg0 + i;
    if (n >= R) return;
    float2 *pc = part + ((size_t)chunk * R + n) * ldp;
#pragma unroll
    for (int rt = 0; rt < RT; rt++) {
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const int r = row0 + 64 * rt + 2 * d32_row(lane, v);
#pragma unroll
            for (int h = 0; h < 2; h++) {
                if (r + h >= M) continue;
                if constexpr (G3)
                    pc[r + h] = make_float2(acc[rt][3 * h][v] - acc[rt][3 * h + 1][v],
                                            acc[rt][3 * h][v] + acc[rt][3 * h + 2][v]);
                else
                    pc[r + h] = make_float2(acc[rt][4 * h][v] - acc[rt][4 * h + 3][v],
                                            acc[rt][4 * h + 1][v] + acc[rt][4 * h + 2][v]);
            }
        }
    }
}

// part[(chunk * R + n) * ldp + j] = sum_{i in chunk} conj(A[i, j]) T[i, n]
// for n in [rg0, rg0 + 32): wave = 32 columns.  Lane (j = l % 32, k = l / 32)
// loads complex rows r, r + 1 (r = ib + 4u + 2k) of column c0 + j and of
// right-hand side rg0 + j: the k index runs over real rows, with
//   Re Z = sum Ar Tr + Ai Ti,  Im Z = sum Ar Ti - Ai Tr.
// Rows read in pairs: row r + 1 may be M, which needs lda, ldt > M for odd M.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kWavesH, kWavesH))) void gemm_h_mfma32(
    const float2 *__restrict__ A, long lda, int M, int N, const float2 *__restrict__ Tv, long ldt, int R, int rg0,
    int kr, float2 *__restrict__ part, long ldp) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c0 = (blockIdx.x * 4 + wave) * 32;
    if (c0 >= N) return;
    const int chunk = blockIdx.y;
    const int i0 = chunk * kr, i1 = min(M, i0 + kr);
    const int j = lane & 31, k = lane >> 5;
    v16f zr = (v16f){0}, zi = (v16f){0};
    // columns past N and right-hand sides past R read the last valid one
    const float2 *col = A + (size_t)min(c0 + j, N - 1) * lda;
    const float2 *Tn = Tv + (size_t)min(rg0 + j, R - 1) * ldt;
    const int rlast = (i1 - 1) & ~1;  // last even row of the chunk
    struct Buf {
        float4 a[kStepsH], t[kStepsH];
    };
    auto load = [&](int ib, Buf &B) {
#pragma unroll
        for (int u = 0; u < kStepsH; u++) {
            const int rc = min(ib + 4 * u + 2 * k, rlast);
            B.a[u] = *reinterpret_cast<const float4 *>(col + rc);
            B.t[u] = *reinterpret_cast<const float4 *>(Tn + rc);
        }
    };
    auto load_masked = [&](int ib, Buf &B) {
#pragma unroll
        for (int u = 0; u < kStepsH; u++) {
            const int row = ib + 4 * u + 2 * k;
            const bool in0 = row < i1, in1 = row + 1 < i1;
            const int rc = min(row, rlast);
            B.a[u] = keep4(in0, in1, *reinterpret_cast<const float4 *>(col + rc));
            B.t[u] = keep4(in0, in1, *reinterpret_cast<const float4 *>(Tn + rc));
        }
    };
    auto mul = [&](const Buf &B) {
#pragma unroll
        for (int u = 0; u < kStepsH; u++) {
            const float4 a = B.a[u], t = B.t[u];
            zr = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, t.x, zr, 0, 0, 0);
            zi = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, t.y, zi, 0, 0, 0);
            zr = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, t.y, zr, 0, 0, 0);
            zi = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, -t.x, zi, 0, 0, 0);
            zr = __builtin_amdgcn_mfma_f32_32x32x2f32(a.z, t.z, zr, 0, 0, 0);
            zi = __builtin_amdgcn_mfma_f32_32x32x2f32(a.z, t.w, zi, 0, 0, 0);
            zr = __builtin_amdgcn_mfma_f32_32x32x2f32(a.w, t.w, zr, 0, 0, 0);
            zi = __builtin_amdgcn_mfma_f32_32x32x2f32(a.w, -t.z, zi, 0, 0, 0);
        }
    };
    Buf b[kRingH];
    const int step = 4 * kStepsH;  // complex rows per buffer
    const int iend = i0 + (i1 - i0) / (step * kRingH) * (step * kRingH);
#pragma unroll
    for (int w = 0; w < kRingH - 1; w++) load(i0 + w * step, b[w]);
    for (int ib = i0; ib < iend; ib += step * kRingH) {
#pragma unroll
        for (int w = 0; w < kRingH; w++) {
            load(ib + (w + kRingH - 1) * step, b[(w + kRingH - 1) % kRingH]);
            __builtin_amdgcn_sched_barrier(0);
            mul(b[w]);
        }
    }
    for (int ib = iend; ib < i1; ib += step) {
        load_masked(ib, b[0]);
        mul(b[0]);
    }
    const int n = rg0 + j;
    if (n >= R) return;
    float2 *pc = part + ((size_t)chunk * R + n) * ldp;
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int jc = c0 + d32_row(lane, v);
        if (jc < N) pc[jc] = make_float2(zr[v], zi[v]);
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

// Chunk length for a product over n columns (rows, for A^H T) split into
// `target` chunks, a multiple of `align`.  FP32 products also cap it at
// kMaxAccum: the MFMA accumulators sum a chunk in FP32, and their rounding
// grows with its length, while the chunk partials are summed in the
// reduce's own precision (FP64 where a refinement follows).  At
// M = N = 40962, R = 32 (block CG on the normal equations, refined to
// 1e-10), three chunks of 13.7k columns left the second refinement cycle at
// 1.3e-10 whatever its inner tolerance, so a third cycle was needed; 32
// chunks reached 3.7e-11, 64 (640 columns) 1.2e-11 and 128 7.9e-12, and the
// products ran as fast with 32 or 64 chunks as with 3.
constexpr long kMaxAccum = 640;
static int chunk_len(long n, long target, long align, bool fp32) {
    long len = std::max(align, ((n + target - 1) / target + align - 1) / align * align);
    if (fp32) len = std::min(len, kMaxAccum);  // a multiple of 32 and of 64
    return (int)len;
}
static long chunk_target(long tiles, bool long_chunks) {
    return long_chunks ? std::max(3L, 1024 / std::max(1L, tiles)) : std::max(1L, 4096 / std::max(1L, tiles));
}

// column chunks: max(3, 1024 / row tiles).  Fewer, longer chunks than for
// 4096 workgroups: measured Y = A X at R = 16, M = 2562: 25.3 us with 41 chunks,
// 17.3 with 12; at R = 32, M = 10258: 364 us with 12, 252 with 3; at
// M = 40994 three chunks (5.4 ms) beat one (8.2 ms)
// (long_chunks: for products whose rounding a refinement absorbs; the short
// rule, chunks for 4096 workgroups, keeps FP32 sums short where the product
// is final, as in the QR updates and the field operator)
static void gemm_n_shape(long M, long N, int &nch, int &kc, bool long_chunks, bool fp32) {
    kc = chunk_len(N, chunk_target((M + 31) / 32, long_chunks), 32, fp32);
    nch = (int)((N + kc - 1) / kc);
}

// row chunks: max(3, 1024 / column tiles), as for gemm_n (measured Z = A^H T at
// R = 32: M = 10258, 334 -> 321 us; M = 40994, 5.81 -> 5.29 ms against chunks
// for 4096 workgroups)
static void gemm_h_shape(long M, long N, int &nch, int &kr, bool long_chunks, bool fp32) {
    kr = chunk_len(M, chunk_target((N + 15) / 16, long_chunks), 64, fp32);
    nch = (int)((M + kr - 1) / kr);
}

// the same rules for the 32-right-hand-side kernels, whose waves cover 64
// rows (Y = A X) or 32 columns (Z = A^H T); they are FP32 only
static void gemm_n_shape32(long M, long N, int &nch, int &kc, bool long_chunks) {
    kc = chunk_len(N, chunk_target((M + 63) / 64, long_chunks), 32, true);
    nch = (int)((N + kc - 1) / kc);
}
static void gemm_h_shape32(long M, long N, int &nch, int &kr, bool long_chunks) {
    kr = chunk_len(M, chunk_target((N + 31) / 32, long_chunks), 64, true);
    nch = (int)((M + kr - 1) / kr);
}

// the 32-right-hand-side kernels take every FP32 product of more than 16
static bool use32(int R) { return R > 16 && !getenv("BEM_GEMM_16X16"); }

// ---------------------------------------------------------- tile tuner ---
// The FP32 products a refinement follows (the Krylov and CG operators) can
// take a tuned shape per size class instead of the rules above: the kernel
// family (16 or 32 right-hand sides per wave) and the number of split-K
// chunks, chosen by timing every candidate on the device
// (bem_autotune_products).  Classes: the product, M and N in half-octave
// buckets, R in steps of 16.  Chunk counts stay <= kMaxTunedChunks, which the
// workspace sizes allow for; a tuned shape still keeps chunks within
// kMaxAccum columns (more chunks than tuned where N is large).
constexpr int kMaxTunedChunks = 16;
namespace {
struct PShape {
    int wide = 0, chunks = 0;
};
std::mutex g_pt_mu;
std::map<std::array<int, 4>, PShape> g_ptune;
thread_local const PShape *t_pforce = nullptr;  // candidate under test
int p_half_octave(long v) { return (int)std::lround(2.0 * std::log2((double)std::max(1L, v))); }
std::array<int, 4> pkey(int which, long M, long N, int R) {
    return {which, p_half_octave(M), p_half_octave(N), (R + 15) / 16};
}
bool tuned_product(int which, long M, long N, int R, PShape &out) {
    if (t_pforce) {
        out = *t_pforce;
        return true;
    }
    std::lock_guard<std::mutex> lk(g_pt_mu);
    auto it = g_ptune.find(pkey(which, M, N, R));
    if (it == g_ptune.end()) return false;
    out = it->second;
    return true;
}
}  // namespace

// partials of the FP32 adjoint product alone (Z = A^H T, N x R)
size_t ws_gemm_h_f32(long M, long N, int R) {
    int nch, kr, n32, k32;
    gemm_h_shape(M, N, nch, kr, false, true);
    gemm_h_shape32(M, N, n32, k32, false);
    return (size_t)std::max({nch, n32, kMaxTunedChunks}) * (size_t)((N + 63) / 64 * 64) * (size_t)std::max(1, R) *
           sizeof(float2);
}

// partials of either product at FP64 for R right-hand sides (and of the
// adjoint product through the conjugate-transposed copy)
size_t ws_gemm(long M, long N, int R) {
    int nch = 0;
    for (bool lc : {false, true}) {
        int a, b, c, d, e, f, g, h, u, v, x, y;
        gemm_n_shape(M, N, a, b, lc, true);
        gemm_h_shape(M, N, c, d, lc, true);
        gemm_n_shape32(M, N, e, f, lc);
        gemm_h_shape32(M, N, g, h, lc);
        gemm_n_shape(N, M, u, v, lc, true);
        gemm_n_shape32(N, M, x, y, lc);
        nch = std::max({nch, a, c, e, g, u, x});
    }
    nch = std::max(nch, kMaxTunedChunks);
    const long ldp = (std::max(M, N) + 63) / 64 * 64;
    return (size_t)nch * (size_t)ldp * (size_t)std::max(1, R) * sizeof(double2);
}

// complex products in the three-product form (BEM_GEMM_4M=1: four products;
// read at every launch, so a process can compare both)
static bool use_3m() { return !getenv("BEM_GEMM_4M"); }

// The 32-column Y = A X kernel with two 64-row tiles per wave: every X load
// then feeds 12 MFMAs instead of 6 (246 VGPRs, 2 waves per SIMD).  At
// M = N = 40962, R = 32, 64 chunks: 3.34 ms against 3.44 ms, the same bits
// (each element's sum runs in the same order).  It halves the waves, so it
// is taken only while the launch keeps at least 1024 workgroups.
// BEM_GEMM_ROW_TILES = 1 or 2 forces either form (read at every launch).
static bool two_row_tiles(long M, int nch) {
    if (const char *e = getenv("BEM_GEMM_ROW_TILES")) return atoi(e) == 2;
    return (M + 511) / 512 * (long)nch >= 1024;
}

// nmat > 1: nmat matrices astride elements apart, matrix z applied to the
// right-hand sides z R / nmat .. (z + 1) R / nmat - 1 (16 x 16 x 4 kernels)
template <typename T, typename S = T, typename OC2 = typename MF<T>::c2>
static int gemm_n_impl(const float2 *A, long lda, long M, long N, const typename MF<T>::c2 *X, long ldx,
                       const OC2 *b, long ldb, OC2 *Y, long ldy, int R, void *ws, const char *name, hipStream_t s,
                       bool long_chunks, long astride = 0, int nmat = 1) {
    if (M <= 0 || N <= 0 || R <= 0) return EXIT_SUCCESS;
    if (lda % 2 || lda < (M + 31) / 32 * 32) BEM_FAIL("gemm_n needs an even lda >= roundup(M, 32)");
    if (nmat < 1 || R % nmat || (nmat > 1 && astride < lda * N) || nmat > 65535)
        BEM_FAIL("gemm_n: %d matrices %ld apart for %d right-hand sides", nmat, astride, R);
    const int rper = R / nmat;
    PShape ps;
    const bool tuned = nmat == 1 && sizeof(T) == 4 && long_chunks && !getenv("BEM_GEMM_16X16") &&
                       tuned_product(0, M, N, R, ps);
    const bool w32 = nmat == 1 && sizeof(T) == 4 && (tuned ? ps.wide != 0 : use32(R));
    if (w32 && lda < (M + 63) / 64 * 64) BEM_FAIL("gemm_n needs lda >= roundup(M, 64) for more than 16 right-hand sides");
    int nch, kc;
    if (tuned) {
        kc = chunk_len(N, ps.chunks, 32, true);
        nch = (int)((N + kc - 1) / kc);
    } else if (w32) {
        gemm_n_shape32(M, N, nch, kc, long_chunks);
    } else {
        gemm_n_shape(M, N, nch, kc, long_chunks, sizeof(T) == 4);
    }
    const long ldp = (M + 63) / 64 * 64;
    auto *part = (typename MF<T>::c2 *)ws;
    // paired X loads read X[(n) ldx + j .. j + 1] as one 16-byte load (an X
    // that cannot be read in pairs takes the four-product form)
    const bool px = w32 && ldx % 2 == 0 && ((uintptr_t)X & 15) == 0 && (N % 2 == 0 || ldx > N);
    // The three-product kernel with paired X loads sums the partials of four
    // column sub-chunks per workgroup in LDS, so it writes a quarter of the
    // partials: chunks of four sub-chunks of about kc columns (an even count,
    // up to 3 % longer, so that no near-empty last chunk is left).  At M =
    // N = 40962, R = 32: 16 chunks of 4 x 672 columns instead of 65 of 640.
    const bool lds = px && use_3m();
    int nred = nch, kc4 = kc;
    if (lds) {
        const long n4 = std::max(1L, (long)std::ceil((double)N / (4.0 * kc * 1.03)));
        kc4 = (int)(((N + 4 * n4 - 1) / (4 * n4) + 1) / 2 * 2);
        nred = (int)((N + 4L * kc4 - 1) / (4L * kc4));
    }
    const bool rt2 = lds && two_row_tiles(M, nch);
    TimedRegion tr;
    timing_begin(tr, name, s);
    const unsigned gx = (unsigned)((M + 127) / 128);
    for (int rg = 0; rg < rper; rg += 32) {
        if (w32) {
            if (rt2) {
                const dim3 g2((unsigned)((M + 127) / 128), nred);
                gemm_n_mfma32<true, true, kRingPX, 2, 2, true><<<g2, 256, 0, s>>>(
                    A, lda, (int)M, (int)N, (const float2 *)X, ldx, R, rg, kc4, (float2 *)part, ldp);
            } else if (lds) {
                const dim3 g1((unsigned)((M + 63) / 64), nred);
                gemm_n_mfma32<true, true, kRingPX, kWavesN, 1, true><<<g1, 256, 0, s>>>(
                    A, lda, (int)M, (int)N, (const float2 *)X, ldx, R, rg, kc4, (float2 *)part, ldp);
            } else {
                const dim3 grid((unsigned)((M + 255) / 256), nch);
                gemm_n_mfma32<false, false><<<grid, 256, 0, s>>>(A, lda, (int)M, (int)N, (const float2 *)X, ldx, R,
                                                                 rg, kc, (float2 *)part, ldp);
            }
        } else {
            dim3 grid(gx, nch, nmat);
            const bool two = rper - rg > 16;
            // FP64, 32 right-hand sides: 2 MFMA steps per prefetched block
            // leave 3 waves per SIMD (168 VGPRs) where 4 steps held one (210
            // VGPRs + 96 AGPRs): M = N = 40962, 6.46-6.50 -> 5.70-5.74 ms
            constexpr int ku2 = sizeof(T) == 8 ? 2 : kU, wv2 = sizeof(T) == 8 ? 2 : 1;
            if (use_3m() && two)
                gemm_n_mfma<T, 2, true, ku2, wv2><<<grid, 256, 0, s>>>(A, lda, (int)M, (int)N, X, ldx, R, rg, kc,
                                                                       part, ldp, astride, rper);
            else if (use_3m())
                gemm_n_mfma<T, 1, true><<<grid, 256, 0, s>>>(A, lda, (int)M, (int)N, X, ldx, R, rg, kc, part, ldp,
                                                             astride, rper);
            else if (two)
                gemm_n_mfma<T, 2, false><<<grid, 256, 0, s>>>(A, lda, (int)M, (int)N, X, ldx, R, rg, kc, part, ldp,
                                                              astride, rper);
            else
                gemm_n_mfma<T, 1, false><<<grid, 256, 0, s>>>(A, lda, (int)M, (int)N, X, ldx, R, rg, kc, part, ldp,
                                                              astride, rper);
        }
        HIP_CALL(hipGetLastError());
    }
    timing_end(tr, s);
    dim3 g2((unsigned)((M + 255) / 256), R);
    gemm_reduce<T, S, OC2><<<g2, 256, 0, s>>>(part, ldp, w32 ? nred : nch, R, (int)M, b, ldb, Y, ldy);
    HIP_CALL(hipGetLastError());
    return EXIT_SUCCESS;
}

template <typename T, typename S = T, typename OC2 = typename MF<T>::c2>
static int gemm_h_impl(const float2 *A, long lda, long M, long N, const typename MF<T>::c2 *Tv, long ldt, OC2 *Z,
                       long ldz, int R, void *ws, const char *name, hipStream_t s, bool long_chunks) {
    if (M <= 0 || N <= 0 || R <= 0) return EXIT_SUCCESS;
    if (lda % 2) BEM_FAIL("gemm_h needs an even lda");
    PShape ps;
    const bool tuned = sizeof(T) == 4 && long_chunks && !getenv("BEM_GEMM_16X16") && tuned_product(1, M, N, R, ps);
    const bool w32 = sizeof(T) == 4 && (tuned ? ps.wide != 0 : use32(R));
    int nch, kr;
    if (tuned) {
        kr = chunk_len(M, ps.chunks, 64, true);
        nch = (int)((M + kr - 1) / kr);
    } else if (w32) {
        gemm_h_shape32(M, N, nch, kr, long_chunks);
    } else {
        gemm_h_shape(M, N, nch, kr, long_chunks, sizeof(T) == 4);
    }
    const long ldp = (N + 63) / 64 * 64;
    auto *part = (typename MF<T>::c2 *)ws;
    TimedRegion tr;
    timing_begin(tr, name, s);
    const unsigned gx = (unsigned)((N + 63) / 64);
    for (int rg = 0; rg < R; rg += 32) {
        if (w32) {
            gemm_h_mfma32<<<dim3((unsigned)((N + 127) / 128), nch), 256, 0, s>>>(
                A, lda, (int)M, (int)N, (const float2 *)Tv, ldt, R, rg, kr, (float2 *)part, ldp);
        } else {
            dim3 grid(gx, nch);
            const bool two = R - rg > 16;
            if (use_3m() && two)
                gemm_h_mfma<T, 2, true><<<grid, 256, 0, s>>>(A, lda, (int)M, (int)N, Tv, ldt, R, rg, kr, part, ldp);
            else if (use_3m())
                gemm_h_mfma<T, 1, true><<<grid, 256, 0, s>>>(A, lda, (int)M, (int)N, Tv, ldt, R, rg, kr, part, ldp);
            else if (two)
                gemm_h_mfma<T, 2, false><<<grid, 256, 0, s>>>(A, lda, (int)M, (int)N, Tv, ldt, R, rg, kr, part, ldp);
            else
                gemm_h_mfma<T, 1, false><<<grid, 256, 0, s>>>(A, lda, (int)M, (int)N, Tv, ldt, R, rg, kr, part, ldp);
        }
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

// ------------------------------------------ conjugate-transposed copy ---
// AH[j + i ldah] = conj(A[i + j lda]) for i < M, j < N; rows N .. ldah - 1 of
// AH are zero.  64 x 64 tiles through LDS, so both the reads of A and the
// writes of AH are 512-byte column segments (two complex values per lane).
// With bmax, the tile's max(|Re|, |Im|) goes to bmax[tile] as well (the
// bound of the FP16-split products, taken while A passes anyway).
__global__ __launch_bounds__(256) void k_adjoint_copy(const float2 *__restrict__ A, long lda, int M, int N,
                                                      float2 *__restrict__ AH, long ldah, float *__restrict__ bmax) {
    __shared__ float2 tile[64][65];  // [column of A][row of A], one pad
    __shared__ float red[4];
    const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 64;
    const int rr = (threadIdx.x & 31) * 2, cc = threadIdx.x >> 5;
    float m = 0.f;
    for (int c = cc; c < 64; c += 8) {
        const int j = j0 + c, i = i0 + rr;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (j < N) {
            if (i + 1 < M) {
                v = *reinterpret_cast<const float4 *>(A + (size_t)j * lda + i);
            } else if (i < M) {
                const float2 a = A[(size_t)j * lda + i];
                v.x = a.x;
                v.y = a.y;
            }
        }
        tile[c][rr] = make_float2(v.x, v.y);
        tile[c][rr + 1] = make_float2(v.z, v.w);
        m = fmaxf(m, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
    }
    if (bmax) {
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    }
    __syncthreads();
    if (bmax && threadIdx.x == 0)
        bmax[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    for (int r = cc; r < 64; r += 8) {
        const int i = i0 + r;
        if (i >= M) break;
        const float2 a = tile[rr][r], b = tile[rr + 1][r];
        *reinterpret_cast<float4 *>(AH + (size_t)i * ldah + j0 + rr) = make_float4(a.x, -a.y, b.x, -b.y);
    }
}

// The copy costs a read and a write of A once per solve; each adjoint product
// then runs at the Y = A X kernel's rate (4.16 against 4.77 ms at M = 40994,
// N = 40962, R = 32).  Only the 32-right-hand-side kernels are measured.
bool adjoint_copy_pays(long M, long N, int R) {
    return M > 0 && N > 0 && use32(R) && !getenv("BEM_NO_ADJOINT_COPY");
}
long adjoint_copy_ld(long N) { return (N + 63) / 64 * 64; }
size_t adjoint_copy_bytes(long M, long N) { return sizeof(float2) * (size_t)adjoint_copy_ld(N) * (size_t)M; }

int adjoint_copy_fill(const float2 *A, long lda, long M, long N, float2 *AH, hipStream_t s, float *bmax) {
    if (M <= 0 || N <= 0) return EXIT_SUCCESS;
    if (lda % 2 || lda < M) BEM_FAIL("adjoint copy needs an even lda >= M");
    if ((N + 63) / 64 > 65535) BEM_FAIL("adjoint copy: N = %ld is too large", N);
    TimedRegion tr;
    timing_begin(tr, "adjoint_copy", s);
    k_adjoint_copy<<<dim3((unsigned)((M + 63) / 64), (unsigned)((N + 63) / 64)), 256, 0, s>>>(
        A, lda, (int)M, (int)N, AH, adjoint_copy_ld(N), bmax);
    HIP_CALL(hipGetLastError());
    timing_end(tr, s);
    return EXIT_SUCCESS;
}

int adjoint_copy_make(const float2 *A, long lda, long M, long N, int R, float2 **AH, hipStream_t s, float *bmax,
                      bool force) {
    *AH = nullptr;
    if (M <= 0 || N <= 0 || getenv("BEM_NO_ADJOINT_COPY") || (!force && !adjoint_copy_pays(M, N, R)))
        return EXIT_SUCCESS;
    const size_t bytes = adjoint_copy_bytes(M, N);
    size_t free_b = 0, total_b = 0;
    HIP_CALL(hipMemGetInfo(&free_b, &total_b));
    if (free_b < bytes + ((size_t)2 << 30)) return EXIT_SUCCESS;  // keep 2 GiB for the rest of the solve
    HIP_CALL(hipMallocAsync((void **)AH, bytes, s));
    const int rc = adjoint_copy_fill(A, lda, M, N, *AH, s, bmax);
    if (rc != EXIT_SUCCESS) {
        (void)hipFreeAsync(*AH, s);
        *AH = nullptr;
    }
    return rc;
}

// Z (N x R) = A^H T through the copy: the Y = A X product of the N x M matrix AH
int gemm_hc_f32(const float2 *AH, long M, long N, const float2 *T, long ldt, float2 *Z, long ldz, int R, void *ws,
                hipStream_t s) {
    return gemm_n_impl<float>(AH, adjoint_copy_ld(N), N, M, T, ldt, (const float2 *)nullptr, 0, Z, ldz, R, ws,
                              "gemm_hc_f32", s, true);
}
int gemm_hc_f32z(const float2 *AH, long M, long N, const float2 *T, long ldt, double2 *Z, long ldz, int R, void *ws,
                 hipStream_t s) {
    return gemm_n_impl<float, double, double2>(AH, adjoint_copy_ld(N), N, M, T, ldt, (const double2 *)nullptr, 0, Z,
                                               ldz, R, ws, "gemm_hc_f32", s, true);
}

// several matrices at once (astride apart, R / nmat right-hand sides each):
// the Krylov product with FP64 output and the FP64 residual b - A X
int gemm_n_f32z_batched(const float2 *A, long lda, long astride, int nmat, long M, long N, const float2 *X, long ldx,
                        double2 *Y, long ldy, int R, void *ws, hipStream_t s) {
    return gemm_n_impl<float, double, double2>(A, lda, M, N, X, ldx, (const double2 *)nullptr, 0, Y, ldy, R, ws,
                                               "gemm_n_f32", s, true, astride, nmat);
}
int mres_f32a_batched(const float2 *A, long lda, long astride, int nmat, long M, long N, const double2 *X, long ldx,
                      const double2 *b, long ldb, double2 *out, long ldo, int R, void *ws, hipStream_t s) {
    return gemm_n_impl<double>(A, lda, M, N, X, ldx, b, ldb, out, ldo, R, ws, "mres_f64acc", s, true, astride, nmat);
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
// the same through the conjugate-transposed copy, on the Y = A X kernel
// (M = N = 40962, R = 32: 6.1 ms against 7.6 ms for the A^H kernel, which
// holds one wave per SIMD at FP64)
int mh_f32a_hc(const float2 *AH, long M, long N, const double2 *T, long ldt, double2 *Z, long ldz, int R, void *ws,
               hipStream_t s) {
    return gemm_n_impl<double>(AH, adjoint_copy_ld(N), N, M, T, ldt, (const double2 *)nullptr, 0, Z, ldz, R, ws,
                               "mh_f64acc", s, true);
}

// ------------------------------------------------- FP16-split products ---
// Y = A X for up to 32 right-hand sides per launch group on
// v_mfma_f32_32x32x16_f16, with A and X each read as a pair of FP16 values
// (hi = fp16(v), lo = fp16(v - hi): about 22 bits) and
//   A X ~ A_hi X_hi + A_hi X_lo + A_lo X_hi
// accumulated in FP32 (the FP16 products are exact in FP32; the dropped
// A_lo X_lo is 2^-22 of a term).  A stays the FP32 column-major matrix: a
// lane splits its operands in registers.  X is split once per product into
// FP16 planes.  Power-of-two scales keep the values in FP16's normal range:
// |A| 2^sa <= 2^13 from amax (the caller's bound on |Re A|, |Im A|), and per
// right-hand side |X| 2^sx <= 2^10 from the column's maximum; the result is
// multiplied by 2^-(sa + sx).  Complex products in the four-product form
// (its operands need no sums before the split).
//
// MFMA lane layout of v_mfma_f32_32x32x16_f16 (checked on the device with an
// in-bounds one-hot probe): A[i][8 (l / 32) + h] from lane l = i + 32 (l / 32),
// half h; B[8 (l / 32) + h][j] from lane j + 32 (l / 32); the result as for
// v_mfma_f32_32x32x2f32.  Lane (i, hh = l / 32) loads complex rows 2i, 2i + 1
// of columns jb + 8 hh .. jb + 8 hh + 7 (eight 16-byte loads of A per 16
// columns, as many as the FP32 kernel) and the four planes of X for its 8
// columns (16 bytes each).
//
// Measured at M = N = 40962, R = 32: 2.55-2.61 ms a product in the block CG
// solve (5.2-5.3 TB/s of A; a read-only kernel with this access pattern
// reaches 6.2 TB/s), against 3.30-3.44 ms for the FP32 kernel (three-product
// form, paired X, two row tiles).  The waves wait on the A stream: MFMA busy
// 24 % of the cycles at 1.97 GHz.
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
constexpr int kF16Cols = 16;     // columns per MFMA step (K)
constexpr int kF16Chunk = 640;   // columns a wave sums in FP32 (about kMaxAccum)

// cpart[n kColmaxBlocks + b] = max of max(|Re X[j, n]|, |Im X[j, n]|) over
// the rows j = b 256 + t (mod kColmaxBlocks 256): workgroup (b, n); the split
// takes the max of a column's partials.  (One workgroup per right-hand side
// took 43 us at N = 40962, R = 32: 32 workgroups on 256 CUs.)
constexpr int kColmaxBlocks = 16;
__global__ __launch_bounds__(256) void k_colmax(const float2 *__restrict__ X, long ldx, int N, float *__restrict__ cpart) {
    __shared__ float red[4];
    const int n = blockIdx.y;
    float m = 0.f;
    for (int j = blockIdx.x * 256 + threadIdx.x; j < N; j += kColmaxBlocks * 256) {
        const float2 x = X[(size_t)n * ldx + j];
        m = fmaxf(m, fmaxf(fabsf(x.x), fabsf(x.y)));
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) cpart[n * kColmaxBlocks + blockIdx.x] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

// power of two p with v p <= 2^e (v > 0), else 1
__device__ __forceinline__ float pow2_scale(float v, int e) {
    if (!(v > 0.f) || !isfinite(v)) return 1.f;
    return ldexpf(1.f, e - ilogbf(v) - 1);
}

// Xp[((n kb-major) * 4 + p) * 16 + t] for right-hand side n, 16-column block kb,
// planes p = Re hi, Re lo, Im hi, Im lo; xinv[n] = 2^-(sa + sx_n); cpart holds
// the kColmaxBlocks partial maxima of each column (k_colmax)
__global__ __launch_bounds__(256) void k_split_x(const float2 *__restrict__ X, long ldx, int N, int R,
                                                 const float *__restrict__ cpart, float sa, _Float16 *__restrict__ Xp,
                                                 int nkb, float *__restrict__ xinv) {
    __shared__ float cmax;
    const int n = blockIdx.y, e = blockIdx.x * 256 + threadIdx.x;  // e = kb * 16 + t
    if (threadIdx.x < 64) {
        float m = threadIdx.x < kColmaxBlocks ? cpart[n * kColmaxBlocks + threadIdx.x] : 0.f;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
        if (threadIdx.x == 0) cmax = m;
    }
    __syncthreads();
    if (n >= R || e >= nkb * kF16Cols) return;
    const float sx = pow2_scale(cmax, 10);
    if (e == 0) xinv[n] = 1.f / (sa * sx);
    float2 x = e < N ? X[(size_t)n * ldx + e] : make_float2(0.f, 0.f);
    x.x *= sx;
    x.y *= sx;
    const _Float16 rh = (_Float16)x.x, ih = (_Float16)x.y;
    _Float16 *o = Xp + ((size_t)n * nkb + (e >> 4)) * 64 + (e & 15);
    o[0] = rh;
    o[16] = (_Float16)(x.x - (float)rh);
    o[32] = ih;
    o[48] = (_Float16)(x.y - (float)ih);
}

// part[(chunk * R + n) * ldp + i] = 2^-(sa + sx_n) sum_{j in chunk} A~[i, j] X~[j, n]
// for n in [rg0, rg0 + 32).  A workgroup takes one 64-row tile over a chunk of
// 4 kc columns: wave w the w-th kc columns (RING buffers of 16 columns), then
// the four partials are summed in LDS in wave order and written as contiguous
// 64-row columns.  Measured at M = N = 40962, R = 32: 2.58-2.65 ms against
// 2.91-2.95 ms with one partial per wave written from registers (four times
// the partial bytes: 672 MB, which cost about 0.5 ms of the A stream whatever
// their write pattern) and 2.41 ms with no partials written at all.  Longer
// per-wave sums (2 kc columns) were no faster and less accurate.
constexpr int kF16Waves = 4;    // waves (column sub-chunks) per workgroup
constexpr int kF16LdsPad = 65;  // float2 per right-hand side in the LDS partials (bank spread)
template <int RING>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void gemm_n_f16s_mfma(
    const float2 *__restrict__ A, long lda, int M, int N, const _Float16 *__restrict__ Xp, int nkb, int R, int rg0,
    float sa, const float *__restrict__ xinv, int kc, float2 *__restrict__ part, long ldp) {
    __shared__ float2 red[kF16Waves * 32 * kF16LdsPad];
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int row0 = blockIdx.x * 64;
    const int chunk = blockIdx.y;
    const int j0 = min(N, (chunk * kF16Waves + wave) * kc), j1 = min(N, j0 + kc);
    const int i = lane & 31, hh = lane >> 5;
    v16f acc[8];
#pragma unroll
    for (int t = 0; t < 8; t++) acc[t] = (v16f){0};
    // right-hand sides past R read the planes of R - 1 (their results are not written)
    const size_t xlane = (size_t)min(rg0 + i, R - 1) * nkb * 8 + hh;  // half8 index of the lane's planes
    // a full 16-column step: wave-uniform base plus the lane's byte offset (< 2^32)
    const unsigned aoff = (unsigned)(((size_t)(8 * hh) * lda + 2 * i) * sizeof(float2));
    const unsigned xoff = (unsigned)(xlane * sizeof(half8));
    const char *Aw = reinterpret_cast<const char *>(A + row0 + (size_t)j0 * lda);
    const size_t cstep = (size_t)lda * sizeof(float2);  // bytes per column
    const char *Xw = reinterpret_cast<const char *>(Xp) + (size_t)(j0 / kF16Cols) * 8 * sizeof(half8);
    struct Buf {
        float4 a[8];
        half8 x[4];
    };
    const int nfull = (j1 - j0) / kF16Cols;  // kc is a multiple of 16: only the last chunk has a partial step
    auto load_full = [&](int st, Buf &B) {
        const char *base = Aw + (size_t)(kF16Cols * st) * cstep;
#pragma unroll
        for (int t = 0; t < 8; t++) B.a[t] = *reinterpret_cast<const float4 *>(base + t * cstep + aoff);
        const char *xb = Xw + (size_t)st * 8 * sizeof(half8);
#pragma unroll
        for (int p = 0; p < 4; p++) B.x[p] = *reinterpret_cast<const half8 *>(xb + 2 * p * sizeof(half8) + xoff);
    };
    auto load_tail = [&](int jb, Buf &B) {
        const float4 *Ab = reinterpret_cast<const float4 *>(A + row0 + 2 * i);
        const long lda4 = lda / 2;
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const int c = jb + 8 * hh + t;
            const float4 v = Ab[(size_t)min(c, j1 - 1) * lda4];
            B.a[t] = c < j1 ? v : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        const half8 *Xb = reinterpret_cast<const half8 *>(Xp) + xlane;
        const int kb = min(jb / kF16Cols, nkb - 1);
#pragma unroll
        for (int p = 0; p < 4; p++) B.x[p] = Xb[(size_t)kb * 8 + 2 * p];
    };
    auto mul = [&](const Buf &B) {
        // the lane's four real rows (Re, Im of rows 2i and 2i + 1) over its 8 columns
        half8 hi[4], lo[4];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const float v[4] = {B.a[t].x * sa, B.a[t].y * sa, B.a[t].z * sa, B.a[t].w * sa};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const _Float16 h = (_Float16)v[q];
                hi[q][t] = h;
                lo[q][t] = (_Float16)(v[q] - (float)h);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            // real row q against Re X (planes 0, 1) and Im X (planes 2, 3)
            acc[2 * q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(hi[q], B.x[0], acc[2 * q], 0, 0, 0);
            acc[2 * q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(hi[q], B.x[1], acc[2 * q], 0, 0, 0);
            acc[2 * q] = __builtin_amdgcn_mfma_f32_32x32x16_f16(lo[q], B.x[0], acc[2 * q], 0, 0, 0);
            acc[2 * q + 1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(hi[q], B.x[2], acc[2 * q + 1], 0, 0, 0);
            acc[2 * q + 1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(hi[q], B.x[3], acc[2 * q + 1], 0, 0, 0);
            acc[2 * q + 1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(lo[q], B.x[2], acc[2 * q + 1], 0, 0, 0);
        }
    };
    Buf b[RING];
#pragma unroll
    for (int w = 0; w < RING - 1; w++)
        if (w < nfull) load_full(w, b[w]);
    for (int s0 = 0; s0 < nfull; s0 += RING) {
#pragma unroll
        for (int w = 0; w < RING; w++) {
            const int st = s0 + w;
            if (st >= nfull) break;
            if (st + RING - 1 < nfull) load_full(st + RING - 1, b[(w + RING - 1) % RING]);
            __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the products
            mul(b[w]);
        }
    }
    if (j0 + kF16Cols * nfull < j1) {
        load_tail(j0 + kF16Cols * nfull, b[0]);
        mul(b[0]);
    }
    // the wave's complex partial of right-hand side rg0 + i, rows 2 d and 2 d + 1
    float2 *rw = red + (size_t)(wave * 32 + i) * kF16LdsPad;
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int d = d32_row(lane, v);
        rw[2 * d] = make_float2(acc[0][v] - acc[3][v], acc[1][v] + acc[2][v]);
        rw[2 * d + 1] = make_float2(acc[4][v] - acc[7][v], acc[5][v] + acc[6][v]);
    }
    __syncthreads();
    const int r = threadIdx.x & 63;
    if (row0 + r >= M) return;
#pragma unroll
    for (int k = 0; k < 32 / kF16Waves; k++) {
        const int nl = (threadIdx.x >> 6) + kF16Waves * k, n = rg0 + nl;
        if (n >= R) break;
        float2 sum = red[(size_t)nl * kF16LdsPad + r];
#pragma unroll
        for (int w = 1; w < kF16Waves; w++) {
            const float2 o = red[(size_t)(w * 32 + nl) * kF16LdsPad + r];
            sum.x += o.x;
            sum.y += o.y;
        }
        const float u = xinv[n];
        part[((size_t)chunk * R + n) * ldp + row0 + r] = make_float2(sum.x * u, sum.y * u);
    }
}

// out[g] = max of bm[b] over b = g 256 + t (mod gridDim 256)
__global__ __launch_bounds__(256) void k_absmax_final(const float *__restrict__ bm, int nb, float *__restrict__ out) {
    __shared__ float red[4];
    float m = 0.f;
    for (int b = blockIdx.x * 256 + threadIdx.x; b < nb; b += gridDim.x * 256) m = fmaxf(m, bm[b]);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

// split workspace: planes of up to 32 right-hand sides, their maxima and scales
size_t ws_split_x(long N, int R) {
    const size_t nkb = (size_t)(N + kF16Cols - 1) / kF16Cols;
    return sizeof(_Float16) * (size_t)std::max(1, R) * nkb * 64 +
           (kColmaxBlocks + 1) * sizeof(float) * (size_t)std::max(1, R) + 256;
}

// the bound on |A| from the per-tile maxima k_adjoint_copy wrote: ws >=
// ws_adjoint_absmax(M, N) floats' bytes, filled by adjoint_copy_make(.., ws);
// *out (device) = the max
// (two levels: 256 workgroups, then one; one workgroup over the 410k tiles at
// M = N = 40962 took 0.64 ms)
constexpr int kAbsmaxGroups = 256;
size_t ws_adjoint_absmax(long M, long N) {
    return sizeof(float) * ((size_t)((M + 63) / 64) * (size_t)((N + 63) / 64) + kAbsmaxGroups + 1);
}
int adjoint_absmax_final(float *bmax, long M, long N, float **out, hipStream_t s) {
    const long nb = ((M + 63) / 64) * ((N + 63) / 64);
    if (nb <= 0 || nb > INT_MAX) BEM_FAIL("adjoint absmax: %ld tiles", nb);
    k_absmax_final<<<kAbsmaxGroups, 256, 0, s>>>(bmax, (int)nb, bmax + nb);
    k_absmax_final<<<1, 256, 0, s>>>(bmax + nb, kAbsmaxGroups, bmax + nb + kAbsmaxGroups);
    HIP_CALL(hipGetLastError());
    *out = bmax + nb + kAbsmaxGroups;
    return EXIT_SUCCESS;
}

template <typename S, typename OC2>
static int gemm_f16s_impl(const float2 *A, long lda, long M, long N, float amax, const float2 *X, long ldx, OC2 *Y,
                          long ldy, int R, void *ws, void *xs, const char *name, hipStream_t s) {
    if (M <= 0 || N <= 0 || R <= 0) return EXIT_SUCCESS;
    if (lda % 2 || lda < (M + 63) / 64 * 64) BEM_FAIL("FP16-split product needs an even lda >= roundup(M, 64)");
    if (!(amax >= 0.f) || !std::isfinite(amax)) BEM_FAIL("FP16-split product: bad bound %g on |A|", (double)amax);
    // 2^sa with amax 2^sa <= 2^13
    const float sa = amax > 0.f ? std::ldexp(1.0f, 13 - std::ilogb(amax) - 1) : 1.0f;
    const int nkb = (int)((N + kF16Cols - 1) / kF16Cols);
    // chunks of four wave sub-chunks of about kF16Chunk columns (a multiple of
    // 16, up to 3 % longer): N = 40962 = 16 x 2560 + 2 takes 16 chunks of
    // 4 x 656 columns rather than a 17th of 2 columns (641 workgroups that
    // write a whole partial, and 21.3 rounds of 512 workgroups instead of 20.03)
    const long nch0 = std::max(1L, (long)std::ceil((double)N / (kF16Waves * kF16Chunk * 1.03)));
    const int kc = (int)(((N + kF16Waves * nch0 - 1) / (kF16Waves * nch0) + kF16Cols - 1) / kF16Cols * kF16Cols);
    const int nch = (int)((N + (long)kF16Waves * kc - 1) / ((long)kF16Waves * kc));
    // the kernel's lane offsets are 32-bit: 15 columns of A, and the planes of all R
    if ((double)lda * 16 * sizeof(float2) >= 4294967296.0 || (double)R * nkb * 8 * sizeof(half8) >= 4294967296.0)
        BEM_FAIL("FP16-split product: lda = %ld or %d x %ld right-hand-side planes too large", lda, R, N);
    const long ldp = (M + 63) / 64 * 64;
    auto *part = (float2 *)ws;
    auto *Xp = (_Float16 *)xs;
    float *cpart = (float *)((char *)xs + sizeof(_Float16) * (size_t)R * nkb * 64);
    float *xinv = cpart + (size_t)R * kColmaxBlocks;
    TimedRegion tr;
    timing_begin(tr, name, s);
    k_colmax<<<dim3(kColmaxBlocks, R), 256, 0, s>>>(X, ldx, (int)N, cpart);
    k_split_x<<<dim3((unsigned)((nkb * kF16Cols + 255) / 256), R), 256, 0, s>>>(X, ldx, (int)N, R, cpart, sa, Xp,
                                                                                 nkb, xinv);
    for (int rg = 0; rg < R; rg += 32)
        gemm_n_f16s_mfma<2><<<dim3((unsigned)((M + 63) / 64), nch), 256, 0, s>>>(
            A, lda, (int)M, (int)N, Xp, nkb, R, rg, sa, xinv, kc, part, ldp);
    HIP_CALL(hipGetLastError());
    timing_end(tr, s);
    dim3 g2((unsigned)((M + 255) / 256), R);
    gemm_reduce<float, S, OC2><<<g2, 256, 0, s>>>(part, ldp, nch, R, (int)M, (const OC2 *)nullptr, 0, Y, ldy);
    HIP_CALL(hipGetLastError());
    return EXIT_SUCCESS;
}

// Y (M x R) = A X; amax >= max |Re A|, |Im A|; ws >= ws_gemm(M, N, R), xs >= ws_split_x(N, R)
int gemm_n_f16s(const float2 *A, long lda, long M, long N, float amax, const float2 *X, long ldx, float2 *Y, long ldy,
                int R, void *ws, void *xs, hipStream_t s) {
    return gemm_f16s_impl<float, float2>(A, lda, M, N, amax, X, ldx, Y, ldy, R, ws, xs, "gemm_n_f16s", s);
}
// Z (N x R) = A^H T through the conjugate-transposed copy, FP64 out
int gemm_hc_f16s_z(const float2 *AH, long M, long N, float amax, const float2 *T, long ldt, double2 *Z, long ldz,
                   int R, void *ws, void *xs, hipStream_t s) {
    return gemm_f16s_impl<double, double2>(AH, adjoint_copy_ld(N), N, M, amax, T, ldt, Z, ldz, R, ws, xs,
                                           "gemm_hc_f16s", s);
}

// ------------------------------------------------------ tuner driver ---
// values in [-0.5, 0.5) from a counter hash: the tuner's operands
__global__ void k_fill_tune(float2 *a, size_t n, unsigned long long seed) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        unsigned long long x = 2 * i + seed;
        float v[2];
        for (int h = 0; h < 2; h++, x++) {
            unsigned long long z = x + 0x9E3779B97F4A7C15ull;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            v[h] = (float)((z >> 40) * (1.0 / 16777216.0)) - 0.5f;
        }
        a[i] = make_float2(v[0], v[1]);
    }
}

// Times both FP32 products of an M x N matrix with R right-hand sides for
// every candidate shape (1 warm-up + 3 timed launches with the reduce, HIP
// events on s) and keeps the fastest per product for the size class.
// shape[0..1]: Y = A X (rhs per wave, chunks), shape[2..3]: Z = A^H T;
// ms[0..3]: best and default time of Y = A X, then of Z = A^H T.
int autotune_products(long M, long N, int R, int *shape, double *ms, hipStream_t s) {
    if (M < N || N <= 0 || R <= 0 || R > kMaxRhs) BEM_FAIL("bad product sizes %ld x %ld, %d right-hand sides", M, N, R);
    const long lda = (M + 63) / 64 * 64;
    float2 *A = nullptr, *X = nullptr, *Y = nullptr;
    char *ws = nullptr;
    HOST_CALL(salloc(&A, (size_t)lda * N, s));
    HOST_CALL(salloc(&X, (size_t)lda * R, s));
    HOST_CALL(salloc(&Y, (size_t)lda * R, s));
    HOST_CALL(salloc(&ws, ws_gemm(M, N, R), s));
    k_fill_tune<<<2048, 256, 0, s>>>(A, (size_t)lda * N, 11);
    k_fill_tune<<<256, 256, 0, s>>>(X, (size_t)lda * R, 12);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = EXIT_SUCCESS;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = EXIT_FAILURE;
    std::vector<PShape> cand;
    for (int wide : {0, 1})
        for (int c : {1, 2, 3, 4, 6, 8, 12, 16}) cand.push_back(PShape{wide, c});
    for (int which = 0; which < 2 && rc == EXIT_SUCCESS; which++) {
        // the rules' own shape (no tuned entry) first: the default time
        {
            std::lock_guard<std::mutex> lk(g_pt_mu);
            g_ptune.erase(pkey(which, M, N, R));
        }
        PShape best{};
        double bms = 1e300, dms = -1;
        for (int ci = -1; ci < (int)cand.size() && rc == EXIT_SUCCESS; ci++) {
            t_pforce = ci < 0 ? nullptr : &cand[ci];
            float t = 0;
            for (int rep = 0; rep < 4 && rc == EXIT_SUCCESS; rep++) {
                if (rep == 1) (void)hipEventRecord(e0, s);
                rc = which == 0 ? gemm_n_f32(A, lda, M, N, X, lda, Y, lda, R, ws, s, true)
                                : gemm_h_f32(A, lda, M, N, X, lda, Y, lda, R, ws, s, true);
            }
            t_pforce = nullptr;
            (void)hipEventRecord(e1, s);
            (void)hipEventSynchronize(e1);
            (void)hipEventElapsedTime(&t, e0, e1);
            const double avg = t / 3.0;
            if (ci < 0) {
                dms = avg;
            } else if (avg < bms) {
                bms = avg;
                best = cand[ci];
            }
        }
        if (rc != EXIT_SUCCESS) break;
        {
            std::lock_guard<std::mutex> lk(g_pt_mu);
            g_ptune[pkey(which, M, N, R)] = best;
        }
        if (shape) {
            shape[2 * which] = best.wide ? 32 : 16;
            shape[2 * which + 1] = best.chunks;
        }
        if (ms) {
            ms[2 * which] = bms;
            ms[2 * which + 1] = dms;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    sfree(A, s);
    sfree(X, s);
    sfree(Y, s);
    sfree(ws, s);
    (void)hipStreamSynchronize(s);
    HOST_CALL(rc);
    return EXIT_SUCCESS;
}

// the tuned product shapes as text lines "p which Mb Nb Rb wide chunks"
void product_tune_save(FILE *f) {
    std::lock_guard<std::mutex> lk(g_pt_mu);
    for (const auto &kv : g_ptune)
        fprintf(f, "p %d %d %d %d %d %d\n", kv.first[0], kv.first[1], kv.first[2], kv.first[3], kv.second.wide,
                kv.second.chunks);
}
bool product_tune_parse(const char *line) {
    int w, a, b, r, wide, c;
    if (sscanf(line, " p %d %d %d %d %d %d", &w, &a, &b, &r, &wide, &c) != 6) return false;
    if ((w != 0 && w != 1) || c < 1 || c > kMaxTunedChunks || (wide != 0 && wide != 1)) return false;
    std::lock_guard<std::mutex> lk(g_pt_mu);
    g_ptune[{w, a, b, r}] = PShape{wide, c};
    return true;
}
void product_tune_clear() {
    std::lock_guard<std::mutex> lk(g_pt_mu);
    g_ptune.clear();
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
