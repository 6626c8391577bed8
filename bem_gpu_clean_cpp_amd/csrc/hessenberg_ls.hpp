// The small least-squares problem of one GMRES cycle, min |beta e1 - H y|,
// kept as a QR factorization by Givens rotations that is updated one
// Hessenberg column at a time (the reference's givens_coeffs, apply_givens
// and Rsolver, GMRES.cu:26-88).  Host-only code: it is compiled into libbem
// and, on its own, into the CPU tests.
//
// Columns are stored packed: column j holds rows 0..j+1 at offset
// j (j + 3) / 2, so a cycle of K steps keeps about K^2 / 2 entries, not
// (K + 1) K.
#pragma once

#include <cmath>
#include <complex>
#include <cstddef>
#include <vector>

namespace bem {

class HessenbergLS {
public:
    typedef std::complex<double> zc;

    // a new cycle of at most K columns whose residual starts at beta
    void reset(int K, double beta) {
        K_ = K;
        len_ = 0;
        H_.resize(off(K));
        cs_.resize(K);
        sn_.resize(K);
        g_.assign((size_t)K + 1, zc(0.0));
        g_[0] = beta;
    }

    // Append column j (j == len(), j < K): h[0..j] are the Gram-Schmidt
    // coefficients, hn >= 0 the norm of the new basis vector.  The earlier
    // rotations are applied, then a new one zeroes hn.  Returns false, and
    // leaves len() at j, when the rotated column is zero (H is singular and
    // the column cannot be used); the residual estimate is then unchanged.
    bool add_column(int j, const zc *h, double hn) {
        zc *c = &H_[off(j)];
        for (int l = 0; l <= j; l++) c[l] = h[l];
        c[j + 1] = hn;
        for (int l = 0; l < j; l++) {
            const zc t0 = cs_[l] * c[l] + sn_[l] * c[l + 1];
            const zc t1 = -std::conj(sn_[l]) * c[l] + std::conj(cs_[l]) * c[l + 1];
            c[l] = t0;
            c[l + 1] = t1;
        }
        const double mag = std::sqrt(std::norm(c[j]) + std::norm(c[j + 1]));
        if (!(mag > 0.0)) return false;
        cs_[j] = std::conj(c[j]) / mag;
        sn_[j] = std::conj(c[j + 1]) / mag;
        c[j] = cs_[j] * c[j] + sn_[j] * c[j + 1];
        c[j + 1] = 0.0;
        g_[j + 1] = -std::conj(sn_[j]) * g_[j];
        g_[j] = cs_[j] * g_[j];
        len_ = j + 1;
        return true;
    }

    // columns of the triangular factor in use
    int len() const { return len_; }
    // |beta e1 - H y| for the best y over the first len() columns
    double residual() const { return std::abs(g_[len_]); }

    // y[0..len()) by back substitution on the triangular factor
    void solve(zc *y) const {
        for (int i = len_ - 1; i >= 0; i--) {
            zc acc = g_[i];
            for (int l = i + 1; l < len_; l++) acc -= H_[off(l) + i] * y[l];
            y[i] = acc / H_[off(i) + i];
        }
    }

    static size_t off(int j) { return (size_t)j * (size_t)(j + 3) / 2; }

private:
    int K_ = 0, len_ = 0;
    std::vector<zc> H_, cs_, sn_, g_;
};

}  // namespace bem
