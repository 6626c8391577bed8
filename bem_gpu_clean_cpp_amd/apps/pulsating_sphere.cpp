// The reference's driver (main.cpp:24-107) on this library: pulsating sphere,
// f = 171.5 Hz, sphere.obj, CHIEF points, QR least squares, then the exterior
// pressure at r = 1.2 .. 3.0 on the -x axis (the commented block main.cpp:74-99
// that produced sphere171Hz_CHIEF / sphere171Hz_NoCHIEF), written with
// wrtCplxMat.  Only the header changes for a reference caller.
//
// usage: pulsating_sphere <mesh.obj> <out file> [num CHIEF = 1] [seed = 1]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/bem.hpp"

using namespace bemgpu;

int main(int argc, char **argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s <mesh.obj> <out file> [numCHIEF] [seed]\n", argv[0]);
        return EXIT_FAILURE;
    }
    const int numChief = argc > 3 ? atoi(argv[3]) : 1;
    const unsigned long long seed = argc > 4 ? strtoull(argv[4], nullptr, 10) : 1;
    const double PI = 3.1415926535897932;
    float f = 171.5;
    float k = 2 * PI * f / 343.21;
    mesh m;
    if (m.readObj(argv[1]) != EXIT_SUCCESS) return EXIT_FAILURE;
    if (m.findBB(0.0001) != EXIT_SUCCESS) return EXIT_FAILURE;
    if (numChief > 0 && m.genCHIEF(numChief, 0.1, seed) != EXIT_SUCCESS) return EXIT_FAILURE;
    gaussQuad gss(6);
    if (gss.sendToDevice() != EXIT_SUCCESS) return EXIT_FAILURE;
    cartCoord src(1000000, 1000000, 1000000);
    const int numSrcs = 1;
    const int M = m.getNumPnts() + m.getNumChief(), N = m.getNumPnts();
    std::vector<cuFloatComplex> A((size_t)M * N), B((size_t)M * numSrcs);
    if (bemSystem(m, k, &src, numSrcs, A.data(), M, B.data(), M) != EXIT_SUCCESS) return EXIT_FAILURE;
    if (qrSolver(A.data(), M, N, M, B.data(), numSrcs, M) != EXIT_SUCCESS) return EXIT_FAILURE;
    const float radius = 3, step = 0.2;
    const int numLocs = (int)std::lround((radius - 1) / step) + 1;
    std::vector<cuFloatComplex> pressure(numLocs);
    pressure[0] = B[0];
    for (int i = 1; i < numLocs; i++) pressure[i] = genExtPressure(k, m, src, cartCoord(-(1 + step * i), 0, 0), B.data());
    if (wrtCplxMat(pressure.data(), 1, numLocs, 1, argv[2]) != EXIT_SUCCESS) return EXIT_FAILURE;
    printf("N=%d M=%d surface pressure at node 0: (%f,%f)\n", N, M, B[0].x, B[0].y);
    return EXIT_SUCCESS;
}
