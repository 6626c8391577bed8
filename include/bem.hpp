// bem.hpp -- the reference's C++ entry points, over the C-ABI of bem.h.
//
// A caller of ZiqiFrancisFan/BEM_GPU_Clean_CPP switches by including this
// header instead of mesh.h / numerical.h / GMRES.h and linking libbem.so.
// Names, argument order and meaning, and the EXIT_SUCCESS / EXIT_FAILURE
// return convention are those of the reference; every function here is a thin
// inline forwarder to the extern "C" function noted beside it.
//
//   reference (file:line)                         here
//   cartCoord (mesh.h:26-99)                      bemgpu::cartCoord
//   triElem (mesh.h:140-209)                      bemgpu::triElem
//   mesh (mesh.h:216-259)                         bemgpu::mesh
//   gaussQuad (numerical.h:166-187)               bemgpu::gaussQuad
//   bemSystem (mesh.h:222-223)                    bemgpu::bemSystem
//   atomicGenSystem (atomicFuncs.h:39-42)         bemgpu::atomicGenSystem
//   qrSolver (numerical.h:196-197)                bemgpu::qrSolver
//   lsqSolver (numerical.h:189-190)               bemgpu::lsqSolver
//   GMRES (GMRES.h:104-105)                       bemgpu::GMRES
//   genExtPressure (mesh.h:425-426)               bemgpu::genExtPressure
//   wrtCplxMat (numerical.h:193-194)              bemgpu::wrtCplxMat
//   printComplexMatrix (numerical.h:143)          bemgpu::printComplexMatrix
//   cmptSurfArea (mesh.h:265)                     bemgpu::cmptSurfArea
#ifndef BEM_HPP
#define BEM_HPP

#include <cstdlib>

#include "bem.h"

namespace bemgpu {

typedef bem_complex cuFloatComplex;  // interleaved {x, y}, 8 bytes

inline cuFloatComplex make_cuFloatComplex(float x, float y) { return bem_complex{x, y}; }

// 12-byte point, layout of cartCoord
struct cartCoord : bem_cart_coord {
    cartCoord() : bem_cart_coord{{0.f, 0.f, 0.f}} {}
    cartCoord(float x, float y, float z) : bem_cart_coord{{x, y, z}} {}
    void set(float x, float y, float z) {
        coords[0] = x;
        coords[1] = y;
        coords[2] = z;
    }
};
static_assert(sizeof(cartCoord) == 12, "cartCoord layout");

typedef bem_tri_elem triElem;
static_assert(sizeof(triElem) == 40, "triElem layout");

// mesh (mesh.h:216-259); the device copy is kept by the library and refreshed
// when the mesh changes.
class mesh {
   public:
    mesh() : h_(bem_mesh_create()) {}
    ~mesh() { bem_mesh_destroy(h_); }
    mesh(const mesh &) = delete;
    mesh &operator=(const mesh &) = delete;
    int readObj(const char *file) { return bem_mesh_read_obj(h_, file); }
    int findBB(const float threshold) { return bem_mesh_find_bb(h_, threshold); }
    // the reference reseeds cuRAND from 0; the seed here selects a SplitMix64 stream
    int genCHIEF(const int num, const float threshold, unsigned long long seed = 0) {
        return bem_mesh_gen_chief(h_, num, threshold, seed);
    }
    int setCHIEF(const cartCoord *pnts, int num) { return bem_mesh_set_chief(h_, pnts, num); }
    int getNumChief() const { return bem_mesh_num_chief(h_); }
    int getNumPnts() const { return bem_mesh_num_pnts(h_); }
    int getNumElems() const { return bem_mesh_num_elems(h_); }
    const bem_mesh *handle() const { return h_; }
    bem_mesh *handle() { return h_; }

   private:
    bem_mesh *h_;
};

// gaussQuad(order) computes the Gauss-Legendre table; sendToDevice publishes it
// to the kernels (numerical.cu:152-156, 227-231).
class gaussQuad {
   public:
    explicit gaussQuad(int order = 6) : n_(order) {}
    int sendToDevice() { return bem_quad_init(n_); }

   private:
    int n_;
};

inline int bemSystem(const mesh &m, const float k, const cartCoord *srcs, const int numSrcs, cuFloatComplex *A,
                     const int lda, cuFloatComplex *B, const int ldb) {
    return bem_system(m.handle(), k, srcs, numSrcs, A, lda, B, ldb);
}

inline int atomicGenSystem(const float k, const triElem *elems, const int numElems, const cartCoord *pnts,
                           const int numNods, const int numCHIEF, const cartCoord *srcs, const int numSrcs,
                           cuFloatComplex *A, const int lda, cuFloatComplex *B, const int ldb) {
    return bem_gen_system(k, elems, numElems, pnts, numNods, numCHIEF, srcs, numSrcs, A, lda, B, ldb);
}

inline int qrSolver(const cuFloatComplex *A, const int mA, const int nA, const int ldA, cuFloatComplex *B,
                    const int nB, const int ldB) {
    return bem_qr_solve(A, mA, nA, ldA, B, nB, ldB);
}

inline int lsqSolver(cuFloatComplex *A_h, const int m, const int n, const int lda, cuFloatComplex *B_h,
                     const int nrhs, const int ldb, cuFloatComplex *Q_h) {
    return bem_lsq_solve(A_h, m, n, lda, B_h, nrhs, ldb, Q_h);
}

inline int GMRES(const cuFloatComplex *A_h, const cuFloatComplex *b_h, const int m, const int maxIter,
                 const float threshold, cuFloatComplex *x_h) {
    return bem_gmres(A_h, b_h, m, maxIter, threshold, x_h);
}

inline cuFloatComplex genExtPressure(const float k, const mesh &m, const cartCoord src, const cartCoord x,
                                     const cuFloatComplex *surfPressure) {
    return bem_ext_pressure(k, m.handle(), src, x, surfPressure);
}

inline int wrtCplxMat(const cuFloatComplex *mat, const int numRows, const int numCols, const int ldm,
                      const char *file) {
    return bem_wrt_cplx_mat(mat, numRows, numCols, ldm, file);
}

inline void printComplexMatrix(cuFloatComplex *A, const int row, const int col, const int lda) {
    bem_print_complex_matrix(A, row, col, lda);
}

inline float cmptSurfArea(const mesh &m) { return bem_mesh_surf_area(m.handle()); }

}  // namespace bemgpu

#endif  // BEM_HPP
