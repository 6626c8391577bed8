// Blocked Householder least squares on the device: the job of qrSolver
// (numerical.cu:580-644: geqrf, unmqr with Q^H, trsm) and of the reference's
// own blocked WY QR (QR_thin, numerical.cu:246-504), without host round trips
// inside a panel.
//
//   qr_panel    one workgroup factors a panel of nb columns in place: for each
//               column a Householder reflector H = I - tau v v^H (v_0 = 1,
//               tau real, beta = -phase(alpha) ||x||), applied to the rest of
//               the panel; then the triangular factor T of the compact WY form
//               H_1 ... H_nb = I - V T V^H (Schreiber & Van Loan 1989).
//   qr_vhc      W = T^H (V^H C) for the trailing columns C (one workgroup per
//               column, the nb dot products in one pass over the column);
//   qr_update   C -= V W (the rank-nb trailing update, GEMM-shaped);
//   FP32 trailing updates run on MFMA instead: V^H C as the adjoint product
//               (C^H V, gemm.hip) against an explicit copy of the panel's V,
//               W = T^H (V^H C), then C -= V W with the rank-32 MFMA update;
//   trsv_upper  R x = (Q^H b)[0:N] by column-oriented back substitution.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <vector>

#include "bem_internal.hpp"

namespace bem {

template <typename T>
struct QT;
template <>
struct QT<float> {
    using c2 = float2;
};
template <>
struct QT<double> {
    using c2 = double2;
};

constexpr int kNB = 32;  // panel width
constexpr int kCG = 8;   // panel columns updated together
constexpr int kColThreads = 256;  // rows per workgroup of the tall-panel column launches

template <typename T>
__device__ __forceinline__ T wsum(T v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// V[i, l] of panel starting at column j0: unit at row j0+l, stored below it.
template <typename C2>
__device__ __forceinline__ C2 vget(const C2 *A, long lda, int j0, int l, int i) {
    C2 z;
    if (i == j0 + l) {
        z.x = 1;
        z.y = 0;
    } else if (i > j0 + l) {
        z = A[(size_t)(j0 + l) * lda + i];
    } else {
        z.x = 0;
        z.y = 0;
    }
    return z;
}

template <typename T>
__global__ __launch_bounds__(1024) void qr_panel(typename QT<T>::c2 *A, long lda, int M, int j0, int nb,
                                                 typename QT<T>::c2 *Tout /* nb x nb, ld kNB */, int make_t) {
    using C2 = typename QT<T>::c2;
    __shared__ T red[2][16];
    __shared__ T s_tau[kNB];
    __shared__ C2 s_Y[kNB][kNB];  // v_l^H v_m for l < m
    __shared__ C2 s_T[kNB][kNB];
    __shared__ T s_red[16][2 * kCG];
    __shared__ T s_d[2 * kCG];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, nw = blockDim.x >> 6;
    for (int jj = 0; jj < nb; jj++) {
        const int j = j0 + jj;
        C2 *col = A + (size_t)j * lda;
        // s = sum_{i>j} |A[i,j]|^2
        T s = 0;
        for (int i = j + 1 + tid; i < M; i += blockDim.x) s += col[i].x * col[i].x + col[i].y * col[i].y;
        s = wsum(s);
        if (lane == 0) red[0][wid] = s;
        __syncthreads();
        T tot = 0;
        for (int q = 0; q < nw; q++) tot += red[0][q];
        const C2 alpha = col[j];
        const T aa = sqrt(alpha.x * alpha.x + alpha.y * alpha.y);
        const T xnorm = sqrt(aa * aa + tot);
        T tau = 0;
        C2 scale;
        scale.x = 0;
        scale.y = 0;
        C2 beta;
        beta = alpha;
        if (tot > 0 || (xnorm > 0 && alpha.y != 0)) {
            const T phr = aa > 0 ? alpha.x / aa : (T)1, phi = aa > 0 ? alpha.y / aa : (T)0;
            beta.x = -phr * xnorm;
            beta.y = -phi * xnorm;
            // scale = 1 / (alpha - beta)
            const T dr = alpha.x - beta.x, di = alpha.y - beta.y;
            const T dd = dr * dr + di * di;
            scale.x = dr / dd;
            scale.y = -di / dd;
            tau = (T)2 / ((T)1 + tot / dd);
        }
        __syncthreads();  // everyone read alpha / red
        for (int i = j + 1 + tid; i < M; i += blockDim.x) {
            const C2 v = col[i];
            C2 o;
            o.x = v.x * scale.x - v.y * scale.y;
            o.y = v.x * scale.y + v.y * scale.x;
            col[i] = o;
        }
        if (tid == 0) {
            col[j] = beta;
            s_tau[jj] = tau;
        }
        __syncthreads();
        // apply H^H = H (tau real) to the remaining panel columns, 8 at a time:
        // every thread takes rows i > j for all 8 columns, one block reduction
        // of the 8 dot products, then the rank-1 update of those rows
        for (int c0 = j + 1; c0 < j0 + nb; c0 += kCG) {
            const int ncg = min(kCG, j0 + nb - c0);
            T dr[kCG], di[kCG];
#pragma unroll
            for (int c = 0; c < kCG; c++) {
                dr[c] = 0;
                di[c] = 0;
            }
            for (int i = j + 1 + tid; i < M; i += blockDim.x) {
                const C2 v = col[i];
#pragma unroll
                for (int c = 0; c < kCG; c++) {
                    if (c < ncg) {
                        const C2 a = A[(size_t)(c0 + c) * lda + i];
                        dr[c] += v.x * a.x + v.y * a.y;
                        di[c] += v.x * a.y - v.y * a.x;
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < kCG; c++) {
                const T r = wsum(dr[c]), im = wsum(di[c]);
                if (lane == 0) {
                    s_red[wid][2 * c] = r;
                    s_red[wid][2 * c + 1] = im;
                }
            }
            __syncthreads();
            if (tid < 2 * ncg) {
                T t = 0;
                for (int q = 0; q < nw; q++) t += s_red[q][tid];
                const int c = tid >> 1;
                const C2 aj = A[(size_t)(c0 + c) * lda + j];  // row j: v_j = 1
                s_d[tid] = (t + ((tid & 1) ? aj.y : aj.x)) * tau;
            }
            __syncthreads();
            for (int i = j + 1 + tid; i < M; i += blockDim.x) {
                const C2 v = col[i];
#pragma unroll
                for (int c = 0; c < kCG; c++) {
                    if (c < ncg) {
                        const T er = s_d[2 * c], ei = s_d[2 * c + 1];
                        C2 *pa = A + (size_t)(c0 + c) * lda + i;
                        C2 a = *pa;
                        a.x -= er * v.x - ei * v.y;
                        a.y -= er * v.y + ei * v.x;
                        *pa = a;
                    }
                }
            }
            if (tid < ncg) {
                C2 *pa = A + (size_t)(c0 + tid) * lda + j;
                C2 a = *pa;
                a.x -= s_d[2 * tid];
                a.y -= s_d[2 * tid + 1];
                *pa = a;
            }
            __syncthreads();
        }
    }
    if (!make_t) {  // T is formed by qr_make_t from an MFMA V^H V: pass the taus on its diagonal
        for (int p = tid; p < kNB * kNB; p += blockDim.x) {
            const int l = p / kNB, m = p % kNB;
            C2 z;
            z.x = (l == m && l < nb) ? s_tau[l] : (T)0;
            z.y = 0;
            Tout[p] = z;
        }
        return;
    }
    // Y[l][m] = v_l^H v_m for l < m (dot over rows >= j0+m)
    for (int p = wid; p < nb * nb; p += nw) {
        const int l = p / nb, m = p % nb;
        if (l >= m) continue;
        T dr = 0, di = 0;
        for (int i = j0 + m + lane; i < M; i += 64) {
            const C2 vl = vget(A, lda, j0, l, i), vm = vget(A, lda, j0, m, i);
            dr += vl.x * vm.x + vl.y * vm.y;
            di += vl.x * vm.y - vl.y * vm.x;
        }
        dr = wsum(dr);
        di = wsum(di);
        if (lane == 0) {
            s_Y[l][m].x = dr;
            s_Y[l][m].y = di;
        }
    }
    __syncthreads();
    // T upper triangular (LAPACK larft, forward columnwise): T_mm = tau_m,
    // T[0:m, m] = -tau_m T[0:m, 0:m] Y[0:m, m]; column m in parallel over rows
    for (int p = tid; p < kNB * kNB; p += blockDim.x) {
        s_T[p / kNB][p % kNB].x = 0;
        s_T[p / kNB][p % kNB].y = 0;
    }
    __syncthreads();
    for (int m = 0; m < nb; m++) {
        if (tid < m) {
            T zr = 0, zi = 0;
            for (int q = tid; q < m; q++) {
                const C2 t = s_T[tid][q], y = s_Y[q][m];
                zr += t.x * y.x - t.y * y.y;
                zi += t.x * y.y + t.y * y.x;
            }
            s_T[tid][m].x = -s_tau[m] * zr;
            s_T[tid][m].y = -s_tau[m] * zi;
        } else if (tid == m) {
            s_T[m][m].x = s_tau[m];
        }
        __syncthreads();
    }
    for (int p = tid; p < kNB * kNB; p += blockDim.x) Tout[p] = s_T[p / kNB][p % kNB];
}

// Sum of 2 kNB = 64 per-lane values over the wave, scattered: on return lane
// l holds the wave total of slot l in x[0].  Each halving step sends one half
// of the live slots to the partner lane and keeps the other half (the half
// picked by the lane's bit H), so 32 + 16 + ... + 1 = 63 shuffles replace the
// 64 x 6 of one butterfly sum per slot.
template <int H, typename T>
__device__ __forceinline__ void reduce_scatter_step(T (&x)[2 * kNB], int lane) {
    const bool up = (lane & H) != 0;
#pragma unroll
    for (int i = 0; i < H; i++) {
        const T send = up ? x[i] : x[i + H];
        const T keep = up ? x[i + H] : x[i];
        x[i] = keep + __shfl_xor(send, H, 64);
    }
}
template <typename T>
__device__ __forceinline__ void wave_reduce_scatter64(T (&x)[2 * kNB], int lane) {
    reduce_scatter_step<32>(x, lane);
    reduce_scatter_step<16>(x, lane);
    reduce_scatter_step<8>(x, lane);
    reduce_scatter_step<4>(x, lane);
    reduce_scatter_step<2>(x, lane);
    reduce_scatter_step<1>(x, lane);
}

// FP32: the two wide steps as lane-half swaps (v_permlane32_swap exchanges
// lanes 32-63 of its first operand with lanes 0-31 of its second,
// v_permlane16_swap the odd 16-lane rows of the first with the even rows of
// the second), so the lane whose bit H is clear ends with its own and its
// partner's x[i] and the other lane with both x[i + H]: one add, no select.
// The narrow steps read the partner through DPP: row_ror:8 pairs l with l ^ 8
// inside a row, row_half_mirror pairs l with 7 - l inside 8 lanes (opposite
// bit 2, same higher bits, which is all the scatter needs), quad_perm
// [2,3,0,1] and [1,0,3,2] pair across bits 1 and 0.
template <int H>
__device__ __forceinline__ void reduce_scatter_swap(float (&x)[2 * kNB]) {
#pragma unroll
    for (int i = 0; i < H; i++) {
        const unsigned a = __float_as_uint(x[i]), b = __float_as_uint(x[i + H]);
        if constexpr (H == 32) {
            const auto r = __builtin_amdgcn_permlane32_swap(a, b, false, false);
            x[i] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
        } else {
            const auto r = __builtin_amdgcn_permlane16_swap(a, b, false, false);
            x[i] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
        }
    }
}
template <int H, int CTRL>
__device__ __forceinline__ void reduce_scatter_dpp(float (&x)[2 * kNB], int lane) {
    const bool up = (lane & H) != 0;
#pragma unroll
    for (int i = 0; i < H; i++) {
        const float send = up ? x[i] : x[i + H];
        const float keep = up ? x[i + H] : x[i];
        x[i] = keep + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(send), CTRL, 0xF, 0xF, false));
    }
}
template <>
__device__ __forceinline__ void wave_reduce_scatter64<float>(float (&x)[2 * kNB], int lane) {
    reduce_scatter_swap<32>(x);
    reduce_scatter_swap<16>(x);
    reduce_scatter_dpp<8, 0x128>(x, lane);  // row_ror:8
    reduce_scatter_dpp<4, 0x141>(x, lane);  // row_half_mirror
    reduce_scatter_dpp<2, 0x4E>(x, lane);   // quad_perm [2,3,0,1]
    reduce_scatter_dpp<1, 0xB1>(x, lane);   // quad_perm [1,0,3,2]
}

// Panel of exactly kNB columns and at most NT * RPT rows (rows j0..M-1) with
// the panel in registers: thread t holds rows j0 + t + NT q (q < RPT) of all
// kNB columns, so a column step costs three barriers and no memory traffic.
// The column loop is unrolled so every register index is static.
//
// Column step jj (j = j0 + jj): the norm below the diagonal (wave sums ->
// s_nrm) gives the reflector; then one scattered wave reduction forms
// sum_r conj(v(r)) a_c(r) for every other column c (-> s_red, summed over the
// waves by 2 kNB threads -> s_fin).  For c > jj that is v^H a_c, which with
// tau updates a_c.  For c < jj the column holds v_c, exact on every row where
// v is nonzero (r >= j > j0 + c), so the sum is conj(v_c^H v); those give
// column jj of U = diag(1/tau) + the strictly upper part of V^H V, and after
// the last column T = U^-1 (H_1 ... H_kNB = I - V T V^H; Schreiber & Van Loan
// 1989, Joffrain et al. 2006).  Each shared array is written only after a
// barrier that follows its last read, so three barriers per column suffice.
//
// Writes the Householder vectors and R back in place, T to Tout (kNB x kNB,
// row-major, zero below the diagonal) and the explicit V (unit diagonal, zero
// above it) to Vd[l * ldv + (i - j0)] for rows i = j0..M-1.
template <int NT, int RPT>
__global__ __launch_bounds__(NT) void qr_panel_reg(float2 *A, long lda, int M, int j0, float2 *Tout, float2 *Vd,
                                                   long ldv) {
    constexpr int NW = NT / 64;
    __shared__ float s_nrm[NW];
    __shared__ float s_red[NW][2 * kNB];
    __shared__ float2 s_fin[kNB];
    __shared__ float2 s_U[kNB][kNB];  // strictly upper part of V^H V (row q, column m)
    __shared__ float s_tau[kNB];
    __shared__ float s_bc[2];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    float2 a[RPT][kNB];
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int r = j0 + tid + NT * q;
#pragma unroll
        for (int c = 0; c < kNB; c++) a[q][c] = r < M ? A[(size_t)(j0 + c) * lda + r] : make_float2(0.f, 0.f);
    }
#pragma clang loop unroll(full)
    for (int jj = 0; jj < kNB; jj++) {
        const int j = j0 + jj;
        // |x|^2 over rows > j, and alpha = A[j, j] to everyone
        float sq = 0.f;
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int r = j0 + tid + NT * q;
            if (r > j && r < M) sq += a[q][jj].x * a[q][jj].x + a[q][jj].y * a[q][jj].y;
            if (r == j) {
                s_bc[0] = a[q][jj].x;
                s_bc[1] = a[q][jj].y;
            }
        }
        sq = wsum(sq);
        if (lane == 0) s_nrm[wid] = sq;
        __syncthreads();
        float tot = 0.f;
#pragma unroll
        for (int w = 0; w < NW; w++) tot += s_nrm[w];
        const float2 alpha = make_float2(s_bc[0], s_bc[1]);
        const float aa = sqrtf(alpha.x * alpha.x + alpha.y * alpha.y);
        const float xnorm = sqrtf(aa * aa + tot);
        float tau = 0.f;
        float2 scale = make_float2(0.f, 0.f), beta = alpha;
        if (tot > 0 || (xnorm > 0 && alpha.y != 0)) {
            const float phr = aa > 0 ? alpha.x / aa : 1.f, phi = aa > 0 ? alpha.y / aa : 0.f;
            beta = make_float2(-phr * xnorm, -phi * xnorm);
            const float dr = alpha.x - beta.x, di = alpha.y - beta.y;
            const float dd = dr * dr + di * di;
            scale = make_float2(dr / dd, -di / dd);
            tau = 2.f / (1.f + tot / dd);
        }
        // v = x / (alpha - beta) below the diagonal (v_j = 1, zero above);
        // R[j, j] = beta
        float2 v[RPT];
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int r = j0 + tid + NT * q;
            if (r > j && r < M) {
                const float2 x = a[q][jj];
                a[q][jj] = make_float2(x.x * scale.x - x.y * scale.y, x.x * scale.y + x.y * scale.x);
                v[q] = a[q][jj];
            } else {
                if (r == j) a[q][jj] = beta;
                v[q] = make_float2(r == j ? 1.f : 0.f, 0.f);
            }
        }
        {
            // slots 2c, 2c + 1: sum_r conj(v(r)) a_c(r) for c != jj (slot pair
            // jj zero), one scattered wave reduction for all of them
            float x[2 * kNB];
#pragma unroll
            for (int c = 0; c < kNB; c++) {
                float pr = 0.f, pi = 0.f;
                if (c != jj) {
#pragma unroll
                    for (int q = 0; q < RPT; q++) {
                        pr += v[q].x * a[q][c].x + v[q].y * a[q][c].y;
                        pi += v[q].x * a[q][c].y - v[q].y * a[q][c].x;
                    }
                }
                x[2 * c] = pr;
                x[2 * c + 1] = pi;
            }
            wave_reduce_scatter64(x, lane);
            s_red[wid][lane] = x[0];
            __syncthreads();
            if (tid < 2 * kNB) {
                float t = 0.f;
#pragma unroll
                for (int w = 0; w < NW; w++) t += s_red[w][tid];
                ((float *)s_fin)[tid] = (tid >> 1) > jj ? t * tau : t;
            }
            __syncthreads();
#pragma unroll
            for (int c = jj + 1; c < kNB; c++) {
                const float2 e = s_fin[c];
#pragma unroll
                for (int q = 0; q < RPT; q++) {
                    a[q][c].x -= e.x * v[q].x - e.y * v[q].y;
                    a[q][c].y -= e.x * v[q].y + e.y * v[q].x;
                }
                if ((c & 3) == 3) __builtin_amdgcn_sched_barrier(0);
            }
            // column jj of U: v_q^H v = conj(s_fin[q]) for q < jj
            if (tid < jj) {
                s_U[tid][jj] = make_float2(s_fin[tid].x, -s_fin[tid].y);
            } else if (tid == jj) {
                s_tau[jj] = tau;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int r = j0 + tid + NT * q;
        if (r < M) {
#pragma unroll
            for (int c = 0; c < kNB; c++) {
                A[(size_t)(j0 + c) * lda + r] = a[q][c];
                const int d = j0 + c;
                Vd[(size_t)c * ldv + (r - j0)] = r > d ? a[q][c] : make_float2(r == d ? 1.f : 0.f, 0.f);
            }
        }
    }
    __syncthreads();
    // T = U^-1 with U = diag(1/tau) + the strictly upper part of V^H V (the
    // inverse of larft's T; Joffrain et al. 2006).  Lane j of wave 0 back-
    // substitutes column j: t_j = tau_j, t_i = -tau_i sum_{k=i+1..j} U[i][k] t_k,
    // written with tau_i rather than 1 / U[i][i], so a tau of zero gives the zero
    // row and column larft gives.  The per-column recurrence of larft inside the
    // column loop was a serial chain of up to 31 dependent shared-memory reads
    // per column; this is one triangular solve per panel.
    // t_k is zero above row j until written, so the sums need no bound on k,
    // the loads of a row of U need no condition, and only t_j is kept by a
    // select
    if (tid < kNB) {
        const int j = tid;
        float2 t[kNB];
#pragma unroll
        for (int i = 0; i < kNB; i++) t[i] = make_float2(i == j ? s_tau[i] : 0.f, 0.f);
#pragma unroll
        for (int i = kNB - 2; i >= 0; i--) {
            float2 u[kNB];
#pragma unroll
            for (int k = i + 1; k < kNB; k++) u[k] = s_U[i][k];
            float sr = 0.f, si = 0.f;
#pragma unroll
            for (int k = i + 1; k < kNB; k++) {
                sr += u[k].x * t[k].x - u[k].y * t[k].y;
                si += u[k].x * t[k].y + u[k].y * t[k].x;
            }
            const float ti = s_tau[i];
            if (i != j) t[i] = make_float2(-ti * sr, -ti * si);
        }
#pragma unroll
        for (int i = 0; i < kNB; i++) Tout[i * kNB + j] = t[i];
    }
}

// Tall FP32 panels (more rows than the register panel holds): one launch per
// column, the rows split over workgroups (one row per thread).  Launch jj
// applies the reflector of column jj to the workgroup's rows and then forms
// the partials of column jj + 1 from the updated rows: |x|^2 below its
// diagonal and sum conj(x_r) a_c(r) for the later columns c, x being the
// column below the diagonal before scaling.  With v = x / (alpha - beta) below
// the diagonal and v_j = 1,
//   v^H a_c = a_c(j) + conj(scale) sum_{r > j} conj(x_r) a_c(r),
// so the reflector's products need no pass of their own after its norm.  The
// partials are summed in a fixed order by every workgroup of the next launch.
// Row j of the panel (alpha and the a_c(j)) travels in its own buffer,
// written by the workgroup holding the row, because that workgroup overwrites
// the row during the launch that reads it.  jj = -1 only forms the partials of
// column 0.  Column indices are runtime values here, so the row is walked with
// compile-time column loops and compares, never by a runtime register index.
// Leaves the reflectors and R in place and tau_jj on Tout's diagonal.
template <int NT>
__global__ __launch_bounds__(NT) void qr_panel_col(float2 *A, long lda, int M, int j0, int jj, const float *zin,
                                                   const float2 *rin, float *zout, float2 *rout, int nblk,
                                                   float2 *Tout) {
    __shared__ float s_red[NT / 64][2 * kNB];
    __shared__ float s_z[2 * kNB];
    __shared__ float2 s_d[kNB];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int i = j0 + blockIdx.x * NT + tid;
    const bool live = i < M;
    const int ic = min(i, M - 1);
    float2 a[kNB];
#pragma unroll
    for (int c = 0; c < kNB; c++) a[c] = A[(size_t)(j0 + c) * lda + ic];
    if (jj >= 0) {
        const int j = j0 + jj;
        if (tid < 2 * kNB) {
            float t = 0.f;
            for (int q = 0; q < nblk; q++) t += zin[(size_t)q * 2 * kNB + tid];
            s_z[tid] = t;
        }
        __syncthreads();
        const float tot = s_z[2 * jj];
        const float2 alpha = rin[jj];
        const float aa = sqrtf(alpha.x * alpha.x + alpha.y * alpha.y);
        const float xnorm = sqrtf(aa * aa + tot);
        float tau = 0.f;
        float2 scale = make_float2(0.f, 0.f), beta = alpha;
        if (tot > 0 || (xnorm > 0 && alpha.y != 0)) {
            const float phr = aa > 0 ? alpha.x / aa : 1.f, phi = aa > 0 ? alpha.y / aa : 0.f;
            beta = make_float2(-phr * xnorm, -phi * xnorm);
            const float dr = alpha.x - beta.x, di = alpha.y - beta.y;
            const float dd = dr * dr + di * di;
            scale = make_float2(dr / dd, -di / dd);
            tau = 2.f / (1.f + tot / dd);
        }
        if (tid < kNB) {
            const int c = tid;
            float2 d = make_float2(0.f, 0.f);
            if (c > jj) {
                // tau (a_c(j) + conj(scale) S_c)
                const float sx = s_z[2 * c], sy = s_z[2 * c + 1];
                const float2 ac = rin[c];
                d.x = tau * (ac.x + scale.x * sx + scale.y * sy);
                d.y = tau * (ac.y + scale.x * sy - scale.y * sx);
            }
            s_d[c] = d;
        }
        if (blockIdx.x == 0 && tid == 0) Tout[jj * kNB + jj] = make_float2(tau, 0.f);
        __syncthreads();
        if (live && i >= j) {
            float2 x = make_float2(0.f, 0.f);
#pragma unroll
            for (int c = 0; c < kNB; c++)
                if (c == jj) x = a[c];
            const bool below = i > j;
            const float2 v = below ? make_float2(x.x * scale.x - x.y * scale.y, x.x * scale.y + x.y * scale.x)
                                   : make_float2(1.f, 0.f);
#pragma unroll
            for (int c = 0; c < kNB; c++) {
                if (c == jj) {
                    a[c] = below ? v : beta;
                } else if (c > jj) {
                    const float2 d = s_d[c];
                    a[c].x -= d.x * v.x - d.y * v.y;
                    a[c].y -= d.x * v.y + d.y * v.x;
                }
                if (c >= jj) A[(size_t)(j0 + c) * lda + i] = a[c];
            }
        }
    }
    if (jj + 1 < kNB) {
        const int k = jj + 1, j1 = j0 + k;
        float2 x = make_float2(0.f, 0.f);
#pragma unroll
        for (int c = 0; c < kNB; c++)
            if (c == k) x = a[c];
        const bool below = live && i > j1;
        float z[2 * kNB];
#pragma unroll
        for (int c = 0; c < kNB; c++) {
            float pr = 0.f, pi = 0.f;
            if (below) {
                if (c == k) {
                    pr = x.x * x.x + x.y * x.y;
                } else if (c > k) {
                    pr = x.x * a[c].x + x.y * a[c].y;
                    pi = x.x * a[c].y - x.y * a[c].x;
                }
            }
            z[2 * c] = pr;
            z[2 * c + 1] = pi;
        }
        if (live && i == j1) {
#pragma unroll
            for (int c = 0; c < kNB; c++)
                if (c >= k) rout[c] = a[c];
        }
        wave_reduce_scatter64(z, lane);
        s_red[wid][lane] = z[0];
        __syncthreads();
        if (tid < 2 * kNB) {
            float t = 0.f;
#pragma unroll
            for (int w = 0; w < NT / 64; w++) t += s_red[w][tid];
            zout[(size_t)blockIdx.x * 2 * kNB + tid] = t;
        }
    }
}

// T of the compact WY form from the taus (on Tout's diagonal) and
// Y[m * ldy + l] = (V^H V)[l][m]: LAPACK larft, forward columnwise, as in qr_panel
template <typename T>
__global__ void qr_make_t(typename QT<T>::c2 *Tout, int nb, const typename QT<T>::c2 *Y, int ldy) {
    using C2 = typename QT<T>::c2;
    __shared__ C2 s_T[kNB][kNB];
    __shared__ C2 s_Y[kNB][kNB + 1];
    __shared__ T s_tau[kNB];
    const int tid = threadIdx.x;
    if (tid < kNB) s_tau[tid] = Tout[tid * kNB + tid].x;
    for (int p = tid; p < kNB * kNB; p += blockDim.x) {
        s_T[p / kNB][p % kNB].x = 0;
        s_T[p / kNB][p % kNB].y = 0;
        // the serial column loop below reads Y from shared memory only
        s_Y[p / kNB][p % kNB] = Y[(size_t)(p / kNB) * ldy + p % kNB];
    }
    __syncthreads();
    for (int m = 0; m < nb; m++) {
        if (tid < m) {
            T zr = 0, zi = 0;
            for (int q = tid; q < m; q++) {
                const C2 t = s_T[tid][q], y = s_Y[m][q];
                zr += t.x * y.x - t.y * y.y;
                zi += t.x * y.y + t.y * y.x;
            }
            s_T[tid][m].x = -s_tau[m] * zr;
            s_T[tid][m].y = -s_tau[m] * zi;
        } else if (tid == m) {
            s_T[m][m].x = s_tau[m];
        }
        __syncthreads();
    }
    for (int p = tid; p < kNB * kNB; p += blockDim.x) Tout[p] = s_T[p / kNB][p % kNB];
}

// W[:, c] = op(T) (V^H C[:, c]); op = conj-transpose (apply Q^H) or identity (apply Q)
template <typename T>
__global__ __launch_bounds__(256) void qr_vhc(const typename QT<T>::c2 *A, long lda, int M, int j0, int nb,
                                              const typename QT<T>::c2 *Tm, const typename QT<T>::c2 *C, long ldc,
                                              typename QT<T>::c2 *W, int conjT) {
    using C2 = typename QT<T>::c2;
    __shared__ T red[2][4][kNB];
    __shared__ C2 z[kNB];
    const int c = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const C2 *cc = C + (size_t)c * ldc;
    T zr[kNB], zi[kNB];
#pragma unroll
    for (int l = 0; l < kNB; l++) {
        zr[l] = 0;
        zi[l] = 0;
    }
    for (int i = j0 + tid; i < M; i += 256) {
        const C2 a = cc[i];
#pragma unroll
        for (int l = 0; l < kNB; l++) {
            if (l < nb) {
                const C2 v = vget(A, lda, j0, l, i);
                zr[l] += v.x * a.x + v.y * a.y;
                zi[l] += v.x * a.y - v.y * a.x;
            }
        }
    }
#pragma unroll
    for (int l = 0; l < kNB; l++) {
        const T r = wsum(zr[l]), im = wsum(zi[l]);
        if (lane == 0) {
            red[0][wid][l] = r;
            red[1][wid][l] = im;
        }
    }
    __syncthreads();
    if (tid < nb) {
        C2 s;
        s.x = red[0][0][tid] + red[0][1][tid] + red[0][2][tid] + red[0][3][tid];
        s.y = red[1][0][tid] + red[1][1][tid] + red[1][2][tid] + red[1][3][tid];
        z[tid] = s;
    }
    __syncthreads();
    if (tid < nb) {
        // w = T^H z (conjT) or T z
        T wr = 0, wi = 0;
        for (int q = 0; q < nb; q++) {
            C2 t = conjT ? Tm[q * kNB + tid] : Tm[tid * kNB + q];
            if (conjT) t.y = -t.y;
            wr += t.x * z[q].x - t.y * z[q].y;
            wi += t.x * z[q].y + t.y * z[q].x;
        }
        C2 o;
        o.x = wr;
        o.y = wi;
        W[(size_t)c * kNB + tid] = o;
    }
}

// C[i, c] -= sum_l V[i, l] W[l, c]
template <typename T>
__global__ __launch_bounds__(256) void qr_update(const typename QT<T>::c2 *A, long lda, int M, int j0, int nb,
                                                 const typename QT<T>::c2 *W, typename QT<T>::c2 *C, long ldc) {
    using C2 = typename QT<T>::c2;
    __shared__ C2 w[kNB];
    const int c = blockIdx.y;
    const int i = j0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (threadIdx.x < nb) w[threadIdx.x] = W[(size_t)c * kNB + threadIdx.x];
    __syncthreads();
    if (i >= M) return;
    T sr = 0, si = 0;
    for (int l = 0; l < nb; l++) {
        const C2 v = vget(A, lda, j0, l, i);
        sr += v.x * w[l].x - v.y * w[l].y;
        si += v.x * w[l].y + v.y * w[l].x;
    }
    C2 *p = C + (size_t)c * ldc + i;
    C2 o = *p;
    o.x -= sr;
    o.y -= si;
    *p = o;
}

// Vd[l][i - j0] = V[i, l] of the panel (explicit unit diagonal and zeros above
// it; columns l >= nb zero), rows j0..M-1, 32 columns, ld ldvd
template <typename T>
__global__ void k_build_v(const typename QT<T>::c2 *A, long lda, int M, int j0, int nb, typename QT<T>::c2 *Vd,
                          long ldvd) {
    using C2 = typename QT<T>::c2;
    const int i = j0 + blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y;
    if (i >= M) return;
    C2 z;
    if (l < nb) {
        z = vget(A, lda, j0, l, i);
    } else {
        z.x = 0;
        z.y = 0;
    }
    Vd[(size_t)l * ldvd + (i - j0)] = z;
}

// W[c][l] = (T^H V^H C)[l][c] = conj(sum_q T[q][l] Z[q][c]) with Z[q][c] =
// (C^H V)[c][q] (the adjoint product's output); rows l >= nb zero
template <typename T>
__global__ void k_qr_w(const typename QT<T>::c2 *Tm, int nb, const typename QT<T>::c2 *Z, long ldz, int nc,
                       typename QT<T>::c2 *W) {
    using C2 = typename QT<T>::c2;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y;
    if (c >= nc) return;
    T sr = 0, si = 0;
    if (l < nb) {
        for (int q = 0; q <= l; q++) {  // T upper triangular
            const C2 t = Tm[q * kNB + l], z = Z[(size_t)q * ldz + c];
            sr += t.x * z.x - t.y * z.y;
            si += t.x * z.y + t.y * z.x;
        }
    }
    C2 o;
    o.x = sr;
    o.y = -si;
    W[(size_t)c * kNB + l] = o;
}

// R x = b (upper, non-unit), one workgroup per right-hand side, in place in B.
template <typename T>
__global__ __launch_bounds__(1024) void trsv_upper(const typename QT<T>::c2 *Rm, long ldr, int N,
                                                   typename QT<T>::c2 *B, long ldb) {
    using C2 = typename QT<T>::c2;
    __shared__ C2 xj;
    C2 *b = B + (size_t)blockIdx.x * ldb;
    for (int j = N - 1; j >= 0; j--) {
        if (threadIdx.x == 0) {
            const C2 r = Rm[(size_t)j * ldr + j], v = b[j];
            const T d = r.x * r.x + r.y * r.y;
            C2 x;
            x.x = (v.x * r.x + v.y * r.y) / d;
            x.y = (v.y * r.x - v.x * r.y) / d;
            b[j] = x;
            xj = x;
        }
        __syncthreads();
        const C2 x = xj;
        const C2 *rc = Rm + (size_t)j * ldr;
        for (int i = threadIdx.x; i < j; i += blockDim.x) {
            const C2 r = rc[i];
            C2 o = b[i];
            o.x -= r.x * x.x - r.y * x.y;
            o.y -= r.x * x.y + r.y * x.x;
            b[i] = o;
        }
        __syncthreads();
    }
}

// Whole least-squares solve of a small factored system in one workgroup per
// right-hand side, b held in shared memory (M entries, dynamic): Q^H b panel
// by panel (z = V^H b as one scattered reduction of the kNB products, w = T^H z,
// b -= V w), then R x = b[0:N] by blocks of kNB columns from the bottom: one
// wave solves the diagonal block with the block's rows of R in registers, then
// the whole workgroup removes the block's unknowns from the rows above.  The
// same arithmetic as qr_vhc + qr_update + trsv_upper without their 3 np + N
// launches and barrier steps.
template <typename T, int NT>
__global__ __launch_bounds__(NT) void qr_solve_small(const typename QT<T>::c2 *A, long lda, int M, int N,
                                                     const typename QT<T>::c2 *Tb, typename QT<T>::c2 *B, long ldb) {
    using C2 = typename QT<T>::c2;
    extern __shared__ unsigned char qss_raw[];
    C2 *sb = (C2 *)qss_raw;
    __shared__ T s_red[NT / 64][2 * kNB];
    __shared__ C2 s_z[kNB], s_w[kNB];
    __shared__ C2 s_T[kNB * kNB];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    C2 *b = B + (size_t)blockIdx.x * ldb;
    for (int i = tid; i < M; i += NT) sb[i] = b[i];
    // Every load below is unconditional from a clamped, valid address, with
    // the unit diagonal, the zeros above it and columns past N selected
    // afterwards, so the loads of a thread are all in flight together.
    const int np = (N + kNB - 1) / kNB;
    for (int p = 0; p < np; p++) {
        const int j0 = p * kNB, nb = min(kNB, N - j0);
        for (int e = tid; e < kNB * kNB; e += NT) s_T[e] = Tb[(size_t)p * kNB * kNB + e];
        __syncthreads();  // sb (first panel) and s_T
        T x[2 * kNB];
#pragma unroll
        for (int l = 0; l < 2 * kNB; l++) x[l] = 0;
        for (int i = j0 + tid; i < M; i += NT) {
            const C2 bi = sb[i];
#pragma unroll
            for (int l0 = 0; l0 < kNB; l0 += kCG) {
                C2 raw[kCG];
#pragma unroll
                for (int u = 0; u < kCG; u++) raw[u] = A[(size_t)min(j0 + l0 + u, N - 1) * lda + i];
#pragma unroll
                for (int u = 0; u < kCG; u++) {
                    const int l = l0 + u;
                    const bool below = i > j0 + l && l < nb, diag = i == j0 + l && l < nb;
                    const T vx = below ? raw[u].x : (diag ? (T)1 : (T)0), vy = below ? raw[u].y : (T)0;
                    x[2 * l] += vx * bi.x + vy * bi.y;
                    x[2 * l + 1] += vx * bi.y - vy * bi.x;
                }
                __builtin_amdgcn_sched_barrier(0);  // one chunk's loads live at a time
            }
        }
        wave_reduce_scatter64(x, lane);
        s_red[wid][lane] = x[0];
        __syncthreads();
        if (tid < 2 * kNB) {
            T t = 0;
#pragma unroll
            for (int w = 0; w < NT / 64; w++) t += s_red[w][tid];
            ((T *)s_z)[tid] = t;
        }
        __syncthreads();
        if (tid < nb) {
            // w = T^H z; T upper triangular
            T wr = 0, wi = 0;
            for (int q = 0; q <= tid; q++) {
                const C2 t = s_T[q * kNB + tid];
                wr += t.x * s_z[q].x + t.y * s_z[q].y;
                wi += t.x * s_z[q].y - t.y * s_z[q].x;
            }
            s_w[tid].x = wr;
            s_w[tid].y = wi;
        }
        __syncthreads();
        for (int i = j0 + tid; i < M; i += NT) {
            T sr = 0, si = 0;
#pragma unroll
            for (int l0 = 0; l0 < kNB; l0 += kCG) {
                C2 raw[kCG];
#pragma unroll
                for (int u = 0; u < kCG; u++) raw[u] = A[(size_t)min(j0 + l0 + u, N - 1) * lda + i];
#pragma unroll
                for (int u = 0; u < kCG; u++) {
                    const int l = l0 + u;
                    const bool below = i > j0 + l && l < nb, diag = i == j0 + l && l < nb;
                    const T vx = below ? raw[u].x : (diag ? (T)1 : (T)0), vy = below ? raw[u].y : (T)0;
                    sr += vx * s_w[l].x - vy * s_w[l].y;
                    si += vx * s_w[l].y + vy * s_w[l].x;
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            sb[i].x -= sr;
            sb[i].y -= si;
        }
        __syncthreads();  // sb, and s_T / s_w before the next panel
    }
    for (int p = np - 1; p >= 0; p--) {
        const int c0 = p * kNB, nb = min(kNB, N - c0);
        if (wid == 0) {
            // lane l owns row c0 + l of the diagonal block
            const int r = c0 + lane, rc = min(r, N - 1);
            C2 rr[kNB];
#pragma unroll
            for (int m = 0; m < kNB; m++) rr[m] = A[(size_t)min(c0 + m, N - 1) * lda + rc];
#pragma unroll
            for (int m = 0; m < kNB; m++) {
                if (!(lane < nb && m < nb && m >= lane)) {
                    rr[m].x = 0;
                    rr[m].y = 0;
                }
            }
            C2 bl = sb[rc];
            if (lane >= nb) {
                bl.x = 0;
                bl.y = 0;
            }
#pragma unroll
            for (int m = kNB - 1; m >= 0; m--) {
                if (m < nb) {
                    // lane m divides by the diagonal; everyone takes x_m from it
                    C2 xm;
                    const T d = rr[m].x * rr[m].x + rr[m].y * rr[m].y;
                    xm.x = lane == m ? (bl.x * rr[m].x + bl.y * rr[m].y) / d : 0;
                    xm.y = lane == m ? (bl.y * rr[m].x - bl.x * rr[m].y) / d : 0;
                    xm.x = __shfl(xm.x, m, 64);
                    xm.y = __shfl(xm.y, m, 64);
                    if (lane < m) {
                        bl.x -= rr[m].x * xm.x - rr[m].y * xm.y;
                        bl.y -= rr[m].x * xm.y + rr[m].y * xm.x;
                    } else if (lane == m) {
                        bl = xm;
                    }
                }
            }
            if (lane < nb) sb[r] = bl;
        }
        __syncthreads();
        for (int i = tid; i < c0; i += NT) {
            T sr = 0, si = 0;
#pragma unroll
            for (int m0 = 0; m0 < kNB; m0 += kCG) {
                C2 raw[kCG];
#pragma unroll
                for (int u = 0; u < kCG; u++) raw[u] = A[(size_t)min(c0 + m0 + u, N - 1) * lda + i];
#pragma unroll
                for (int u = 0; u < kCG; u++) {
                    if (m0 + u < nb) {
                        const C2 xv = sb[c0 + m0 + u];
                        sr += raw[u].x * xv.x - raw[u].y * xv.y;
                        si += raw[u].x * xv.y + raw[u].y * xv.x;
                    }
                }
            }
            sb[i].x -= sr;
            sb[i].y -= si;
        }
        __syncthreads();
    }
    for (int i = tid; i < N; i += NT) b[i] = sb[i];
}

// Q^H B for systems past the one-workgroup solve: one launch per panel, the
// rows split over workgroups of NT threads (one row per thread, one
// right-hand side per blockIdx.y).  Launch p finishes panel p on the block's
// rows (z = sum of the partials of V_p^H b in a fixed order, w = T_p^H z,
// b -= V_p w) and then forms the block's partial of V_{p+1}^H b over the same
// rows, whose values are final for panel p+1.  p = -1 only forms the partials
// of panel 0.  Partials alternate between two buffers by panel parity.
template <typename T, int NT>
__global__ __launch_bounds__(NT) void qr_apply_step(const typename QT<T>::c2 *A, long lda, int M, int N,
                                                    const typename QT<T>::c2 *Tb, typename QT<T>::c2 *B, long ldb,
                                                    int p, const T *zin, T *zout, int nblk) {
    using C2 = typename QT<T>::c2;
    __shared__ T s_red[NT / 64][2 * kNB];
    __shared__ C2 s_z[kNB], s_w[kNB];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, col = blockIdx.y;
    const int i = blockIdx.x * NT + tid;
    const bool live = i < M;
    C2 *b = B + (size_t)col * ldb;
    C2 bi;
    bi.x = 0;
    bi.y = 0;
    if (live) bi = b[i];
    const int ic = min(i, M - 1);
    const int np = (N + kNB - 1) / kNB;
    if (p >= 0) {
        const int j0 = p * kNB, nb = min(kNB, N - j0);
        if (tid < 2 * kNB) {
            T t = 0;
            for (int q = 0; q < nblk; q++) t += zin[((size_t)col * nblk + q) * 2 * kNB + tid];
            ((T *)s_z)[tid] = t;
        }
        __syncthreads();
        if (tid < nb) {
            // w = T^H z; T upper triangular
            const C2 *Tm = Tb + (size_t)p * kNB * kNB;
            T wr = 0, wi = 0;
            for (int q = 0; q <= tid; q++) {
                const C2 t = Tm[q * kNB + tid];
                wr += t.x * s_z[q].x + t.y * s_z[q].y;
                wi += t.x * s_z[q].y - t.y * s_z[q].x;
            }
            s_w[tid].x = wr;
            s_w[tid].y = wi;
        }
        __syncthreads();
        T sr = 0, si = 0;
#pragma unroll
        for (int l0 = 0; l0 < kNB; l0 += kCG) {
            C2 raw[kCG];
#pragma unroll
            for (int u = 0; u < kCG; u++) raw[u] = A[(size_t)min(j0 + l0 + u, N - 1) * lda + ic];
#pragma unroll
            for (int u = 0; u < kCG; u++) {
                const int l = l0 + u;
                const bool below = i > j0 + l && l < nb, diag = i == j0 + l && l < nb;
                const T vx = below ? raw[u].x : (diag ? (T)1 : (T)0), vy = below ? raw[u].y : (T)0;
                sr += vx * s_w[l].x - vy * s_w[l].y;
                si += vx * s_w[l].y + vy * s_w[l].x;
            }
        }
        if (live && i >= j0) {
            bi.x -= sr;
            bi.y -= si;
            b[i] = bi;
        }
    }
    if (p + 1 < np) {
        const int j0 = (p + 1) * kNB, nb = min(kNB, N - j0);
        T x[2 * kNB];
#pragma unroll
        for (int l = 0; l < 2 * kNB; l++) x[l] = 0;
#pragma unroll
        for (int l0 = 0; l0 < kNB; l0 += kCG) {
            C2 raw[kCG];
#pragma unroll
            for (int u = 0; u < kCG; u++) raw[u] = A[(size_t)min(j0 + l0 + u, N - 1) * lda + ic];
#pragma unroll
            for (int u = 0; u < kCG; u++) {
                const int l = l0 + u;
                const bool below = live && i > j0 + l && l < nb, diag = i == j0 + l && l < nb;
                const T vx = below ? raw[u].x : (diag ? (T)1 : (T)0), vy = below ? raw[u].y : (T)0;
                x[2 * l] = vx * bi.x + vy * bi.y;
                x[2 * l + 1] = vx * bi.y - vy * bi.x;
            }
        }
        wave_reduce_scatter64(x, lane);
        s_red[wid][lane] = x[0];
        __syncthreads();
        if (tid < 2 * kNB) {
            T t = 0;
#pragma unroll
            for (int w = 0; w < NT / 64; w++) t += s_red[w][tid];
            zout[((size_t)col * nblk + blockIdx.x) * 2 * kNB + tid] = t;
        }
    }
}

// One block column of R x = b for systems past the one-workgroup solve: every
// workgroup solves the nb x nb diagonal block at c0 redundantly (one wave,
// the block's rows of R in registers; b[c0..c0+nb) is not written in this
// step), workgroup 0 stores the unknowns in X, and every workgroup removes
// them from its rows below c0 of the working right-hand side B.
template <typename T, int NT>
__global__ __launch_bounds__(NT) void trsv_block_step(const typename QT<T>::c2 *A, long lda, int N,
                                                      typename QT<T>::c2 *B, long ldb, typename QT<T>::c2 *X,
                                                      long ldx, int c0, int nb) {
    using C2 = typename QT<T>::c2;
    __shared__ C2 s_x[kNB];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, col = blockIdx.y;
    C2 *b = B + (size_t)col * ldb;
    if (wid == 0) {
        const int r = c0 + lane, rc = min(r, N - 1);
        C2 rr[kNB];
#pragma unroll
        for (int m = 0; m < kNB; m++) rr[m] = A[(size_t)min(c0 + m, N - 1) * lda + rc];
#pragma unroll
        for (int m = 0; m < kNB; m++) {
            if (!(lane < nb && m < nb && m >= lane)) {
                rr[m].x = 0;
                rr[m].y = 0;
            }
        }
        C2 bl = b[rc];
        if (lane >= nb) {
            bl.x = 0;
            bl.y = 0;
        }
#pragma unroll
        for (int m = kNB - 1; m >= 0; m--) {
            if (m < nb) {
                C2 xm;
                const T d = rr[m].x * rr[m].x + rr[m].y * rr[m].y;
                xm.x = lane == m ? (bl.x * rr[m].x + bl.y * rr[m].y) / d : 0;
                xm.y = lane == m ? (bl.y * rr[m].x - bl.x * rr[m].y) / d : 0;
                xm.x = __shfl(xm.x, m, 64);
                xm.y = __shfl(xm.y, m, 64);
                if (lane < m) {
                    bl.x -= rr[m].x * xm.x - rr[m].y * xm.y;
                    bl.y -= rr[m].x * xm.y + rr[m].y * xm.x;
                } else if (lane == m) {
                    bl = xm;
                }
            }
        }
        if (lane < nb) {
            s_x[lane] = bl;
            if (blockIdx.x == 0) X[(size_t)col * ldx + r] = bl;
        }
    }
    __syncthreads();
    const int i = blockIdx.x * NT + tid;
    if (i < c0) {
        T sr = 0, si = 0;
#pragma unroll
        for (int m0 = 0; m0 < kNB; m0 += kCG) {
            C2 raw[kCG];
#pragma unroll
            for (int u = 0; u < kCG; u++) raw[u] = A[(size_t)min(c0 + m0 + u, N - 1) * lda + i];
#pragma unroll
            for (int u = 0; u < kCG; u++) {
                if (m0 + u < nb) {
                    sr += raw[u].x * s_x[m0 + u].x - raw[u].y * s_x[m0 + u].y;
                    si += raw[u].x * s_x[m0 + u].y + raw[u].y * s_x[m0 + u].x;
                }
            }
        }
        b[i].x -= sr;
        b[i].y -= si;
    }
}

// ------------------------------------------------------------------ host ---
template <typename T>
struct QRFactor {
    using C2 = typename QT<T>::c2;
    C2 *A = nullptr;  // factored copy (owned by caller)
    long lda = 0;
    int M = 0, N = 0;
    C2 *Tb = nullptr;  // per-panel T factors
    C2 *W = nullptr;   // workspace nb x max(ncols)
    int wcols = 0;
    C2 *Vd = nullptr;  // explicit panel V (FP32 MFMA trailing update)
    C2 *Z = nullptr;   // (C^H V) of the trailing columns
    C2 *Yz = nullptr;  // V^H V of the panel (kNB x kNB)
    char *gws = nullptr;
    long ldv = 0;
    // partials and panel rows of the one-launch-per-column tall panel
    float *zcol = nullptr;
    long zcol_stride = 0;
    float2 *rcol = nullptr;
    hipStream_t st = nullptr;
    ~QRFactor() {
        sfree(Tb, st);
        sfree(W, st);
        sfree(Vd, st);
        sfree(Z, st);
        sfree(Yz, st);
        sfree(gws, st);
        sfree(zcol, st);
        sfree(rcol, st);
    }
};

template <typename T>
static int qr_apply(const QRFactor<T> &f, typename QT<T>::c2 *C, long ldc, int nc, int conjT, hipStream_t s) {
    if (nc <= 0) return EXIT_SUCCESS;
    if (nc > f.wcols) BEM_FAIL("QR workspace too small");
    const int np = (f.N + kNB - 1) / kNB;
    for (int pi = 0; pi < np; pi++) {
        const int p = conjT ? pi : np - 1 - pi;  // Q^H: H_nb..H_1 order of panels 0..; Q: reverse
        const int j0 = p * kNB, nb = std::min(kNB, f.N - j0);
        qr_vhc<T><<<nc, 256, 0, s>>>(f.A, f.lda, f.M, j0, nb, f.Tb + (size_t)p * kNB * kNB, C, ldc, f.W, conjT);
        HIP_CALL(hipGetLastError());
        dim3 g((unsigned)((f.M - j0 + 255) / 256), (unsigned)nc);
        qr_update<T><<<g, 256, 0, s>>>(f.A, f.lda, f.M, j0, nb, f.W, C, ldc);
        HIP_CALL(hipGetLastError());
    }
    return EXIT_SUCCESS;
}

template <typename T>
static int qr_factor(QRFactor<T> &f, int extra_cols, hipStream_t s) {
    using C2 = typename QT<T>::c2;
    const int np = (f.N + kNB - 1) / kNB;
    f.st = s;
    HOST_CALL(salloc(&f.Tb, (size_t)np * kNB * kNB, s));
    f.wcols = std::max(f.N, extra_cols);
    HOST_CALL(salloc(&f.W, (size_t)kNB * f.wcols, s));
    // FP32: trailing updates on MFMA (Z = C^H V with the adjoint product, then
    // C -= V W with the rank-32 update); FP64 keeps the VALU kernels
    const bool mfma = sizeof(T) == 4 && !getenv("BEM_QR_VALU");
    const bool global_panel = getenv("BEM_QR_GPANEL") != nullptr;
    const bool one_wg_panel = getenv("BEM_QR_ONE_WG_PANEL") != nullptr;
    if (mfma) {
        f.ldv = (f.lda + 63) / 64 * 64;
        HOST_CALL(salloc(&f.Vd, (size_t)f.ldv * kNB, s));
        HOST_CALL(salloc(&f.Z, (size_t)kNB * f.N, s));
        HOST_CALL(salloc(&f.Yz, (size_t)kNB * kNB, s));
        size_t wsz = 0;
        for (int p = 0; p < np; p++) {
            const int j0 = p * kNB, nb = std::min(kNB, f.N - j0);
            wsz = std::max(wsz, ws_gemm_h_f32(f.M - j0, kNB, kNB));
            if (f.N - (j0 + nb) > 0) wsz = std::max(wsz, ws_gemm_h_f32(f.M - j0, f.N - (j0 + nb), kNB));
        }
        HOST_CALL(salloc(&f.gws, std::max<size_t>(wsz, 1), s));
        f.zcol_stride = (long)((f.M + kColThreads - 1) / kColThreads) * 2 * kNB;
        HOST_CALL(salloc(&f.zcol, (size_t)2 * f.zcol_stride, s));
        HOST_CALL(salloc(&f.rcol, (size_t)2 * kNB, s));
    }
    for (int p = 0; p < np; p++) {
        const int j0 = p * kNB, nb = std::min(kNB, f.N - j0);
        TimedRegion tr;
        timing_begin(tr, sizeof(T) == 4 ? "qr_panel_f32" : "qr_panel_f64", s);
        C2 *Tp = f.Tb + (size_t)p * kNB * kNB;
        const int Mp = f.M - j0;
        // panels of up to 1024 rows: the register kernel also forms T and the
        // explicit V; taller ones (and a narrow last panel) run the global
        // panel kernel, then V, V^H V on MFMA and T from it
        const bool fit = mfma && nb == kNB && !global_panel;
        const int reg = !fit ? 0 : Mp <= 256 ? 1 : Mp <= 512 ? 2 : Mp <= 1024 ? 3 : 0;
        float2 *Tf = (float2 *)Tp, *Vf = (float2 *)f.Vd;
        if (reg == 1) {
            qr_panel_reg<256, 1><<<1, 256, 0, s>>>((float2 *)f.A, f.lda, f.M, j0, Tf, Vf, f.ldv);
        } else if (reg == 2) {
            qr_panel_reg<512, 1><<<1, 512, 0, s>>>((float2 *)f.A, f.lda, f.M, j0, Tf, Vf, f.ldv);
        } else if (reg == 3) {
            qr_panel_reg<512, 2><<<1, 512, 0, s>>>((float2 *)f.A, f.lda, f.M, j0, Tf, Vf, f.ldv);
        } else if (fit && !one_wg_panel) {
            // tall panel: one launch per column over kColThreads-row workgroups
            const int nblk = (Mp + kColThreads - 1) / kColThreads;
            for (int jj = -1; jj < kNB; jj++) {
                const int bi = jj & 1, bo = (jj + 1) & 1;
                qr_panel_col<kColThreads><<<nblk, kColThreads, 0, s>>>(
                    (float2 *)f.A, f.lda, f.M, j0, jj, f.zcol + (size_t)bi * f.zcol_stride,
                    (const float2 *)f.rcol + (size_t)bi * kNB, f.zcol + (size_t)bo * f.zcol_stride,
                    (float2 *)f.rcol + (size_t)bo * kNB, nblk, Tf);
                HIP_CALL(hipGetLastError());
            }
        } else {
            qr_panel<T><<<1, 1024, 0, s>>>(f.A, f.lda, f.M, j0, nb, Tp, mfma ? 0 : 1);
        }
        HIP_CALL(hipGetLastError());
        if (mfma && !reg) {
            dim3 gv((unsigned)((Mp + 255) / 256), kNB);
            k_build_v<T><<<gv, 256, 0, s>>>(f.A, f.lda, f.M, j0, nb, f.Vd, f.ldv);
            HIP_CALL(hipGetLastError());
            HOST_CALL(gemm_h_f32((const float2 *)f.Vd, f.ldv, Mp, kNB, (const float2 *)f.Vd, f.ldv, (float2 *)f.Yz,
                                 kNB, kNB, f.gws, s));
            qr_make_t<T><<<1, 64, 0, s>>>(Tp, nb, f.Yz, kNB);
            HIP_CALL(hipGetLastError());
        }
        timing_end(tr, s);
        const int nc = f.N - (j0 + nb);
        if (nc > 0) {
            C2 *Ct = f.A + (size_t)(j0 + nb) * f.lda;
            TimedRegion tu;
            timing_begin(tu, sizeof(T) == 4 ? "qr_trailing_f32" : "qr_trailing_f64", s);
            if (mfma) {
                float2 *Cb = (float2 *)(Ct + j0);  // rows j0.., columns j0+nb..
                HOST_CALL(gemm_h_f32(Cb, f.lda, Mp, nc, (const float2 *)f.Vd, f.ldv, (float2 *)f.Z, f.N, kNB, f.gws,
                                     s));
                dim3 gw((unsigned)((nc + 255) / 256), kNB);
                k_qr_w<T><<<gw, 256, 0, s>>>(f.Tb + (size_t)p * kNB * kNB, nb, f.Z, f.N, nc, f.W);
                HIP_CALL(hipGetLastError());
                HOST_CALL(rank_update_f32((const float2 *)f.Vd, f.ldv, Mp, (const float2 *)f.W, kNB, nc, Cb, f.lda,
                                          s));
            } else {
                qr_vhc<T><<<nc, 256, 0, s>>>(f.A, f.lda, f.M, j0, nb, f.Tb + (size_t)p * kNB * kNB, Ct, f.lda,
                                             f.W, 1);
                HIP_CALL(hipGetLastError());
                dim3 g((unsigned)((f.M - j0 + 255) / 256), (unsigned)nc);
                qr_update<T><<<g, 256, 0, s>>>(f.A, f.lda, f.M, j0, nb, f.W, Ct, f.lda);
                HIP_CALL(hipGetLastError());
            }
            timing_end(tu, s);
        }
    }
    return EXIT_SUCCESS;
}

// Least squares solve with a factored A: B (M x R, ldb) <- solution in B[0:N].
constexpr size_t kSmallSolveLds = 48 * 1024;  // b of up to 6144 FP32 / 3072 FP64 rows
constexpr int kSmallSolveRhs = 64;

template <typename T>
static int qr_solve_factored(const QRFactor<T> &f, typename QT<T>::c2 *B, long ldb, int R, hipStream_t s) {
    using C2 = typename QT<T>::c2;
    if (R <= 0) return EXIT_SUCCESS;
    // the whole solve in one workgroup per right-hand side, b in shared memory
    // (on request; the row-split launches below measured faster at every size
    // from 129 to 4096 rows)
    const size_t lds = sizeof(C2) * (size_t)f.M;
    if (lds <= kSmallSolveLds && R <= kSmallSolveRhs && getenv("BEM_QR_ONE_WG_SOLVE")) {
        // 2 kNB partial sums per thread plus a chunk of 8 loads: the register
        // room of 512 (FP32) or 256 (FP64) threads
        constexpr int nt = sizeof(T) == 4 ? 512 : 256;
        qr_solve_small<T, nt><<<R, nt, lds, s>>>(f.A, f.lda, f.M, f.N, f.Tb, B, ldb);
        HIP_CALL(hipGetLastError());
        return EXIT_SUCCESS;
    }
    if (getenv("BEM_QR_PER_COLUMN_SOLVE")) {
        HOST_CALL(qr_apply<T>(f, B, ldb, R, 1, s));
        trsv_upper<T><<<R, 1024, 0, s>>>(f.A, f.lda, f.N, B, ldb);
        HIP_CALL(hipGetLastError());
        return EXIT_SUCCESS;
    }
    // one launch per panel for Q^H B and one per block column for the back
    // substitution, the rows spread over workgroups
    constexpr int nt = 256;
    const int np = (f.N + kNB - 1) / kNB;
    const int nblk = (f.M + nt - 1) / nt;
    T *zb = nullptr;
    C2 *X = nullptr;
    HOST_CALL(salloc(&zb, (size_t)2 * R * nblk * 2 * kNB, s));
    HOST_CALL(salloc(&X, (size_t)f.N * R, s));
    int rc = EXIT_SUCCESS;
    for (int p = -1; p < np && rc == EXIT_SUCCESS; p++) {
        const T *zin = zb + (size_t)((p + 2) & 1) * R * nblk * 2 * kNB;
        T *zout = zb + (size_t)((p + 1) & 1) * R * nblk * 2 * kNB;
        qr_apply_step<T, nt><<<dim3(nblk, R), nt, 0, s>>>(f.A, f.lda, f.M, f.N, f.Tb, B, ldb, p, zin, zout, nblk);
        if (hipGetLastError() != hipSuccess) rc = EXIT_FAILURE;
    }
    for (int p = np - 1; p >= 0 && rc == EXIT_SUCCESS; p--) {
        const int c0 = p * kNB, nb = std::min(kNB, f.N - c0);
        const int g = std::max(1, (c0 + nt - 1) / nt);
        trsv_block_step<T, nt><<<dim3(g, R), nt, 0, s>>>(f.A, f.lda, f.N, B, ldb, X, f.N, c0, nb);
        if (hipGetLastError() != hipSuccess) rc = EXIT_FAILURE;
    }
    if (rc == EXIT_SUCCESS &&
        hipMemcpy2DAsync(B, sizeof(C2) * ldb, X, sizeof(C2) * f.N, sizeof(C2) * f.N, R, hipMemcpyDeviceToDevice,
                         s) != hipSuccess)
        rc = EXIT_FAILURE;
    sfree(zb, s);
    sfree(X, s);
    if (rc != EXIT_SUCCESS) BEM_FAIL("QR solve launches failed");
    return EXIT_SUCCESS;
}

// Plain least squares: A (overwritten with its factorisation), B -> solution.
template <typename T>
static int qr_lsq_impl(typename QT<T>::c2 *A, long lda, int M, int N, typename QT<T>::c2 *B, long ldb, int R,
                       hipStream_t s) {
    if (M < N || N <= 0) BEM_FAIL("least squares needs M >= N > 0 (got %d x %d)", M, N);
    QRFactor<T> f;
    f.A = A;
    f.lda = lda;
    f.M = M;
    f.N = N;
    HOST_CALL(qr_factor<T>(f, R, s));
    HOST_CALL(qr_solve_factored<T>(f, B, ldb, R, s));
    HIP_CALL(hipStreamSynchronize(s));
    return EXIT_SUCCESS;
}

int qr_lsq_f32(float2 *A, long lda, int M, int N, float2 *B, long ldb, int R, hipStream_t s) {
    return qr_lsq_impl<float>(A, lda, M, N, B, ldb, R, s);
}
int qr_lsq_f64(double2 *A, long lda, int M, int N, double2 *B, long ldb, int R, hipStream_t s) {
    return qr_lsq_impl<double>(A, lda, M, N, B, ldb, R, s);
}

__global__ void k_identity(float2 *Q, long ldq, int M) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i < M) Q[(size_t)j * ldq + i] = make_float2(i == j ? 1.f : 0.f, 0.f);
}

// Least squares plus explicit Q: A <- R (upper triangle, zero below), Q (M x M).
int qr_lsq_with_q_f32(float2 *A, long lda, int M, int N, float2 *B, long ldb, int R, float2 *Q, long ldq,
                      hipStream_t s) {
    if (M < N || N <= 0) BEM_FAIL("least squares needs M >= N > 0 (got %d x %d)", M, N);
    QRFactor<float> f;
    f.A = A;
    f.lda = lda;
    f.M = M;
    f.N = N;
    HOST_CALL(qr_factor<float>(f, std::max(R, M), s));
    if (Q) {
        dim3 g((unsigned)((M + 255) / 256), (unsigned)M);
        k_identity<<<g, 256, 0, s>>>(Q, ldq, M);
        HIP_CALL(hipGetLastError());
        HOST_CALL(qr_apply<float>(f, Q, ldq, M, 0, s));
    }
    HOST_CALL(qr_solve_factored<float>(f, B, ldb, R, s));
    HIP_CALL(hipStreamSynchronize(s));
    return EXIT_SUCCESS;
}

// ------------------------------------------------ refined LS (FP32 + FP64) ---
int resid_n_f32a(const float2 *A, long lda, long M, long N, const double2 *x, long ldx, const double2 *b, long ldb,
                 double2 *r, long ldr, int R, void *ws, hipStream_t s);

__global__ void k_z2c(const double2 *__restrict__ in, float2 *__restrict__ out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_float2((float)in[i].x, (float)in[i].y);
}
__global__ void k_addc2z(const float2 *__restrict__ d, double2 *__restrict__ x, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        x[i].x += d[i].x;
        x[i].y += d[i].y;
    }
}
__global__ void k_sumsq(const double2 *__restrict__ v, int n, double *out);
__global__ void k_sumsq_c(const float2 *__restrict__ v, int n, double *out);
__global__ void k_sumsq(const double2 *__restrict__ v, int n, double *out) {
    __shared__ double red[16];
    double s = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += v[i].x * v[i].x + v[i].y * v[i].y;
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int q = 0; q < (int)(blockDim.x >> 6); q++) t += red[q];
        *out = t;
    }
}
__global__ void k_sumsq_c(const float2 *__restrict__ v, int n, double *out) {
    __shared__ double red[16];
    double s = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += (double)v[i].x * v[i].x + (double)v[i].y * v[i].y;
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int q = 0; q < (int)(blockDim.x >> 6); q++) t += red[q];
        *out = t;
    }
}

// FP32 QR least squares wrapped in FP64 iterative refinement: the fixed-point
// iteration x <- x + QRsolve(b - A x) with the residual accumulated in FP64
// from the FP32 matrix.  A32 (kept), B64: M x R FP64 rhs, X64: N x R FP64
// solution (out).  Stops when ||correction|| <= refine_tol ||x|| or when the
// correction no longer shrinks (an inconsistent least-squares system keeps a
// nonzero residual, so the residual itself is not the test).
int qr_refined_f32(const float2 *A32, long lda, int M, int N, const double2 *B64, long ldb, double2 *X64, long ldx,
                   int R, double refine_tol, int max_refine, bem_solve_stats *st, hipStream_t s) {
    if (M < N || N <= 0) BEM_FAIL("least squares needs M >= N > 0 (got %d x %d)", M, N);
    QRFactor<float> f;
    HOST_CALL(salloc(&f.A, (size_t)lda * N, s));
    HIP_CALL(hipMemcpyAsync(f.A, A32, sizeof(float2) * lda * N, hipMemcpyDeviceToDevice, s));
    f.lda = lda;
    f.M = M;
    f.N = N;
    float2 *rb = nullptr;
    double2 *r64 = nullptr;
    double *dn = nullptr;
    char *gws = nullptr;
    int refinements = 0;
    double worst = 0;
    bool conv = true;
    int rc = EXIT_FAILURE;
    const int bs = 256;
    if (qr_factor<float>(f, 1, s) != EXIT_SUCCESS) goto out;
    if (salloc(&rb, (size_t)lda, s) || salloc(&r64, (size_t)lda, s) || salloc(&dn, 2, s)) goto out;
    if (salloc(&gws, ws_gemv_n(M, N, 1), s)) goto out;
    for (int r = 0; r < R; r++) {
        const double2 *b = B64 + (size_t)r * ldb;
        double2 *x = X64 + (size_t)r * ldx;
        if (hipMemsetAsync(x, 0, sizeof(double2) * N, s) != hipSuccess) goto out;
        double prev = 1e300, rel = 0;
        for (int it = 0; it <= max_refine; it++) {
            if (resid_n_f32a(A32, lda, M, N, x, 0, b, 0, r64, 0, 1, gws, s) != EXIT_SUCCESS) goto out;
            k_z2c<<<(M + bs - 1) / bs, bs, 0, s>>>(r64, rb, M);
            if (qr_solve_factored<float>(f, rb, lda, 1, s) != EXIT_SUCCESS) goto out;
            k_addc2z<<<(N + bs - 1) / bs, bs, 0, s>>>(rb, x, N);
            k_sumsq_c<<<1, 1024, 0, s>>>(rb, N, dn);
            k_sumsq<<<1, 1024, 0, s>>>(x, N, dn + 1);
            double h[2];
            if (hipMemcpyAsync(h, dn, 2 * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess) goto out;
            if (hipStreamSynchronize(s) != hipSuccess) goto out;
            rel = std::sqrt(h[0] / std::max(h[1], 1e-300));
            refinements++;
            if (rel <= refine_tol || rel >= 0.5 * prev) break;
            prev = rel;
        }
        worst = std::max(worst, rel);
        conv = conv && rel <= std::max(refine_tol, 1e-5);
    }
    rc = EXIT_SUCCESS;
out:
    if (st) {
        st->refinements = refinements;
        st->iterations = 0;
        st->converged = conv ? 1 : 0;
        st->residual = worst;
    }
    sfree(f.A, s);
    sfree(rb, s);
    sfree(r64, s);
    sfree(dn, s);
    sfree(gws, s);
    (void)hipStreamSynchronize(s);
    if (rc != EXIT_SUCCESS) BEM_FAIL("refined QR least squares failed");
    return rc;
}

}  // namespace bem
