// Small dense complex algebra on the host for the block Krylov solvers: the
// R x R (R <= 64) Cholesky factor, triangular inverse, conjugate transpose and
// product, column-major, FP64.  Host-only code: it is compiled into libbem
// and, on its own, into the CPU tests (tests/native).
//
// The loops spell the complex arithmetic out on (re, im) pairs: std::complex
// products carry an inf/NaN fallback that costs more than the arithmetic, and
// a block step of 32 systems does about 10^5 of them.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstddef>
#include <vector>

namespace bem {
namespace small {

using cd = std::complex<double>;
using Small = std::vector<cd>;
inline double *dp(Small &v) { return reinterpret_cast<double *>(v.data()); }
inline const double *dp(const Small &v) { return reinterpret_cast<const double *>(v.data()); }

// G = L L^H (lower L) from G's lower triangle; false when a pivot is not
// positive beyond rounding (at most 1e-14 of the largest diagonal entry)
inline bool chol_lower(const Small &G, int n, Small &L) {
    L.assign((size_t)n * n, cd(0));
    const double *g = dp(G);
    double *l = dp(L);
    double gmax = 0;
    for (int j = 0; j < n; j++) gmax = std::max(gmax, g[2 * ((size_t)j * n + j)]);
    if (!(gmax > 0) || !std::isfinite(gmax)) return false;
    for (int j = 0; j < n; j++) {
        double d = g[2 * ((size_t)j * n + j)];
        for (int k = 0; k < j; k++) {
            const double *lk = l + 2 * ((size_t)k * n + j);
            d -= lk[0] * lk[0] + lk[1] * lk[1];
        }
        if (!(d > 1e-14 * gmax)) return false;
        const double ljj = std::sqrt(d), rl = 1.0 / ljj;
        l[2 * ((size_t)j * n + j)] = ljj;
        for (int i = j + 1; i < n; i++) {
            double vr = g[2 * ((size_t)j * n + i)], vi = g[2 * ((size_t)j * n + i) + 1];
            for (int k = 0; k < j; k++) {
                // v -= L[i, k] conj(L[j, k])
                const double *a = l + 2 * ((size_t)k * n + i), *c = l + 2 * ((size_t)k * n + j);
                vr -= a[0] * c[0] + a[1] * c[1];
                vi -= a[1] * c[0] - a[0] * c[1];
            }
            l[2 * ((size_t)j * n + i)] = vr * rl;
            l[2 * ((size_t)j * n + i) + 1] = vi * rl;
        }
    }
    return true;
}

// inverse of a lower-triangular L with a real diagonal (forward substitution
// per column)
inline Small lower_inv(const Small &L, int n) {
    Small X((size_t)n * n, cd(0));
    const double *l = dp(L);
    double *x = dp(X);
    for (int j = 0; j < n; j++) {
        double *xj = x + 2 * (size_t)j * n;
        xj[2 * j] = 1.0 / l[2 * ((size_t)j * n + j)];
        for (int i = j + 1; i < n; i++) {
            double vr = 0, vi = 0;
            for (int k = j; k < i; k++) {
                const double *a = l + 2 * ((size_t)k * n + i), *b = xj + 2 * k;
                vr += a[0] * b[0] - a[1] * b[1];
                vi += a[0] * b[1] + a[1] * b[0];
            }
            const double r = -1.0 / l[2 * ((size_t)i * n + i)];
            xj[2 * i] = vr * r;
            xj[2 * i + 1] = vi * r;
        }
    }
    return X;
}

// A^H
inline Small herm(const Small &A, int n) {
    Small B((size_t)n * n);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) B[(size_t)j * n + i] = std::conj(A[(size_t)i * n + j]);
    return B;
}

// A B.  The real and imaginary parts of a column of the product accumulate
// in separate arrays (no aliasing, contiguous lanes), then interleave.
inline Small mul(const Small &A, const Small &B, int n) {
    Small C((size_t)n * n);
    const double *__restrict__ a = dp(A);
    const double *__restrict__ b = dp(B);
    double *__restrict__ c = dp(C);
    std::vector<double> ar((size_t)n * n), ai((size_t)n * n);
    for (size_t e = 0; e < (size_t)n * n; e++) {
        ar[e] = a[2 * e];
        ai[e] = a[2 * e + 1];
    }
    std::vector<double> cr(n), ci(n);
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < n; i++) cr[i] = ci[i] = 0.0;
        for (int k = 0; k < n; k++) {
            const double br = b[2 * ((size_t)j * n + k)], bi = b[2 * ((size_t)j * n + k) + 1];
            if (br == 0.0 && bi == 0.0) continue;
            const double *__restrict__ xr = ar.data() + (size_t)k * n;
            const double *__restrict__ xi = ai.data() + (size_t)k * n;
            double *__restrict__ yr = cr.data();
            double *__restrict__ yi = ci.data();
            for (int i = 0; i < n; i++) {
                yr[i] += xr[i] * br - xi[i] * bi;
                yi[i] += xr[i] * bi + xi[i] * br;
            }
        }
        for (int i = 0; i < n; i++) {
            c[2 * ((size_t)j * n + i)] = cr[i];
            c[2 * ((size_t)j * n + i) + 1] = ci[i];
        }
    }
    return C;
}

}  // namespace small
}  // namespace bem
