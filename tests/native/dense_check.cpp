// Driver for the CPU tests of the host-side dense algebra (small_dense.hpp)
// and the GMRES least-squares update (hessenberg_ls.hpp).  Reads one command
// and its data from stdin, writes the results to stdout as text; the test
// compares them with numpy.
//   chol n   G (n x n, column-major, re im pairs)   -> ok, L
//   linv n   L                                       -> inverse of L
//   mul n    A B                                     -> A B
//   herm n   A                                       -> A^H
//   lsq K beta ncols, then per column: hn and h[0..j] -> per column: ok and
//            the residual estimate; then len and y
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

#include "hessenberg_ls.hpp"
#include "small_dense.hpp"

using bem::small::cd;
using bem::small::Small;

static Small read_mat(int n) {
    Small a((size_t)n * n);
    for (auto &v : a) {
        double r, i;
        std::cin >> r >> i;
        v = cd(r, i);
    }
    return a;
}
static void write_mat(const Small &a) {
    for (const auto &v : a) std::printf("%.17g %.17g\n", v.real(), v.imag());
}

int main() {
    std::string cmd;
    std::cin >> cmd;
    if (cmd == "chol") {
        int n;
        std::cin >> n;
        Small G = read_mat(n), L;
        std::printf("%d\n", bem::small::chol_lower(G, n, L) ? 1 : 0);
        write_mat(L);
    } else if (cmd == "linv") {
        int n;
        std::cin >> n;
        write_mat(bem::small::lower_inv(read_mat(n), n));
    } else if (cmd == "mul") {
        int n;
        std::cin >> n;
        Small A = read_mat(n), B = read_mat(n);
        write_mat(bem::small::mul(A, B, n));
    } else if (cmd == "herm") {
        int n;
        std::cin >> n;
        write_mat(bem::small::herm(read_mat(n), n));
    } else if (cmd == "lsq") {
        int K, ncols;
        double beta;
        std::cin >> K >> beta >> ncols;
        bem::HessenbergLS ls;
        ls.reset(K, beta);
        std::vector<cd> h((size_t)K + 1);
        for (int j = 0; j < ncols; j++) {
            double hn;
            std::cin >> hn;
            for (int l = 0; l <= j; l++) {
                double r, i;
                std::cin >> r >> i;
                h[l] = cd(r, i);
            }
            const bool ok = ls.add_column(ls.len(), h.data(), hn);
            std::printf("%d %.17g\n", ok ? 1 : 0, ls.residual());
        }
        std::vector<cd> y((size_t)K);
        ls.solve(y.data());
        std::printf("%d\n", ls.len());
        for (int i = 0; i < ls.len(); i++) std::printf("%.17g %.17g\n", y[i].real(), y[i].imag());
    } else {
        std::fprintf(stderr, "unknown command %s\n", cmd.c_str());
        return 2;
    }
    return 0;
}
