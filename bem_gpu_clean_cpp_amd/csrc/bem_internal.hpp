// Internal declarations shared by the host and HIP sources of libbem.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/bem.h"

namespace bem {

// ---------------------------------------------------------------- errors ---
void set_error(const char *file, int line, const char *fmt, ...);
// a callee failed and has set its message: prefix it with this call site
void chain_error(const char *file, int line, const char *what);

#define BEM_FAIL(...)                                       \
    do {                                                    \
        ::bem::set_error(__FILE__, __LINE__, __VA_ARGS__);  \
        return EXIT_FAILURE;                                \
    } while (0)

#define HIP_CALL(x)                                                                        \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            ::bem::set_error(__FILE__, __LINE__, "HIP error '%s' in %s", hipGetErrorString(e_), #x); \
            return EXIT_FAILURE;                                                           \
        }                                                                                  \
    } while (0)

#define HOST_CALL(x)                                                     \
    do {                                                                 \
        if ((x) != EXIT_SUCCESS) {                                       \
            ::bem::chain_error(__FILE__, __LINE__, #x);                  \
            return EXIT_FAILURE;                                         \
        }                                                                \
    } while (0)

// ------------------------------------------------------------ constants ---
constexpr int kQOrder = 6;            // INTORDER (numerical.h:40-42)
constexpr int kNQ = kQOrder * kQOrder; // 36 points of the collapsed rule
constexpr int kMaxRhs = 64;
constexpr double kPi = 3.1415926535897932;

struct Config {
    int qorder = 0;
    double qp[32] = {0};
    double qw[32] = {0};
    float density = 1.2041f;  // numerical.cu:12
    float speed = 343.21f;    // numerical.cu:14
    float strength = 0.0f;    // STRENGTH, numerical.h:128-130
    // near-singular FP64 integration: (row, element) pairs closer than
    // near_factor times the element's longest edge are integrated by
    // adaptive subdivision (0: off, the reference's rule everywhere)
    double near_factor = 0.0;
    int near_levels = 12;
};
Config &config();
int ensure_quad();  // default order-6 tables if bem_quad_init was not called

// Per-point tables of the 6x6 collapsed rule (element independent):
// xi1, xi2 of the regular map and W = 0.25 w_n w_m rho_m (mesh.cu:623-636).
struct QuadTables {
    float xi1[kNQ], xi2[kNQ], w[kNQ];
    double dxi1[kNQ], dxi2[kNQ], dw[kNQ];
};
void quad_tables(QuadTables &t);

// ------------------------------------------------------------- context ---
struct Context {
    hipStream_t stream = nullptr;
    int device = -1;
};
int ctx_init(Context **out);

// Kernel timing with HIP events on the library stream.
bool timing_on();
struct TimedRegion {
    hipEvent_t a = nullptr, b = nullptr;
    const char *name = nullptr;
};
void timing_begin(TimedRegion &r, const char *name, hipStream_t s);
void timing_end(TimedRegion &r, hipStream_t s);

// ---------------------------------------------------------------- mesh ---
struct Mesh {
    std::vector<float> pnts;   // 3 * N
    std::vector<float> chief;  // 3 * C
    std::vector<bem_tri_elem> elems;
    float bb[6] = {0, 0, 0, 0, 0, 0};
    bool bb_set = false;
    uint64_t version = 0;  // bumped on every change
    int N() const { return (int)(pnts.size() / 3); }
    int C() const { return (int)(chief.size() / 3); }
    int E() const { return (int)elems.size(); }
};
int validate_mesh(const Mesh &m);

// Column tiles for row-owned assembly: a partition of the nodes into P
// spatially compact parts (recursive coordinate bisection); for each part, the
// elements with at least one vertex in it, with per-entry flags:
//   bits 0-2  vertex j of the element lies in this part
//   bits 3-5  this entry is the first, in list order, to touch vertex j's column
//   bit  6    this part owns the element (counts its c and B terms once)
struct TilePlan {
    int P = 0;
    std::vector<int> ptr;      // P + 1
    std::vector<int> entries;  // element index << 8 | flags
    std::vector<int> isolated; // nodes referenced by no element
};
void build_tile_plan(const Mesh &m, int P, TilePlan &plan);
// The same column tiles with every element integrated once, by the tile that
// owns most of its vertices (else the tile of vertex 0).  A vertex owned by
// another tile is a ghost: its contributions go to a side column of its own
// per (tile, node), and ghost_reduce adds the side columns of each node to A
// in a fixed order.
struct GhostPlan {
    int P = 0;
    std::vector<int> ptr;      // P + 1
    std::vector<int> entries;  // element index << 8 | flags (bits 0-2 owned vertex, 3-5 first touch)
    std::vector<int> slots;    // 3 per entry: side column of a ghost vertex, -1 for an owned one
    int nslots = 0;
    std::vector<int> gptr;     // per ghost node: its side columns gslot[gptr[g] .. gptr[g + 1])
    std::vector<int> gnode;    // the node
    std::vector<int> gown;     // 1 when the node's own tile also writes its column
    std::vector<int> gslot;
    double dup = 0;            // entries of the halo plan per element (for the choice of plan)
};
void build_ghost_plan(const Mesh &m, int P, GhostPlan &g);

// Element record read by the assembly kernels with scalar loads (128 B):
//   p3, e1 = p1-p3, e2 = p2-p3, n (unit normal of e1 x e2), J = |e1 x e2|,
//   beta = ca/cb, gamma = cc/cb, nodes, and for the separable FP32 rule
//   f = e2 - e1 and vv[n] = |e1 + theta_n f|^2 at the six Gauss-Legendre
//   theta_n (element-only terms of the squared distance).
template <typename R>
struct alignas(16) ElemRec {
    R p3[3];
    R e1[3];
    R e2[3];
    R n[3];
    R J;
    R beta[2];
    R gamma[2];
    int nodes[3];
    R f[3];
    R vv[6];
    int pad[(sizeof(R) == 4) ? 3 : 1];
};
static_assert(sizeof(ElemRec<float>) == 128, "record size");

// ------------------------------------------------------------ assembly ---
// Device-resident problem: mesh on the device and assembly scratch.
struct DeviceProblem {
    int N = 0, C = 0, M = 0, E = 0;
    long ldA = 0;                 // leading dimension of the device A (M rounded up to 64)
    float4 *pts = nullptr;        // M points (nodes then CHIEF), FP32
    double *pts64 = nullptr;      // 3 * M, FP64 copy
    int *elems = nullptr;         // 3 * E node indices
    float2 *bc = nullptr;         // 3 * E (ca, cb, cc)
    ElemRec<float> *rec32 = nullptr;
    ElemRec<double> *rec64 = nullptr;
    float *sing32 = nullptr;      // E * 3 * 16: singular contributions
    double *sing64 = nullptr;
    int P = 0;                    // column tiles
    int qg = 1;                   // lanes per row in the quadrature kernel
    int impedance = 0;            // some element has ca != 0
    double diam = 0;              // bounding-box diagonal of the M points: bounds every |x - y|
    int *tile_ptr = nullptr;
    int *tile_entries = nullptr;
    int n_isolated = 0;
    int *isolated = nullptr;
    // element-once plan of the FP32 fast kernel (GhostPlan), when it pays
    int use_ghost = 0;
    int gP = 0;                   // its column tiles
    int *gtile_ptr = nullptr, *gtile_entries = nullptr, *gtile_slots = nullptr;
    int nslots = 0, ngnodes = 0;
    int *gred_ptr = nullptr, *gred_node = nullptr, *gred_own = nullptr, *gred_slot = nullptr;
    float2 *side = nullptr;       // nslots * Mp side columns
    float *cpart = nullptr;       // P * Mp  c partial sums
    float2 *bpart = nullptr;      // P * Mp  boundary-condition partial sums
    double *cpart64 = nullptr;
    double2 *bpart64 = nullptr;
    double2 *bterm = nullptr;     // M: sum_e (cc/cb) sum_j g_j(x_i)
    long Mp = 0;
    // near-singular FP64 correction (bem_set_near_field): host copies of the
    // points and elements for the pair search, the (row, element) pairs in
    // row order, and each pair's correction (13 doubles: g, h complex x 3, c)
    std::vector<double> hpts;     // 3 * M
    std::vector<int> helems;      // 3 * E
    double near_factor = 0;       // factor the pairs were found with (0: none)
    int near_levels = 0;
    int near_npairs = 0;
    std::vector<int> hnear_ptr;   // M + 1: pairs of row i at [ptr[i], ptr[i + 1])
    int *near_ptr = nullptr;
    int *near_row = nullptr, *near_elem = nullptr;
    double *near_out = nullptr;
    uint64_t mesh_version = 0;
    hipStream_t st = nullptr;     // buffers are stream-ordered on this stream
};
int device_problem_build(const Mesh &m, DeviceProblem &dp, hipStream_t s);
void device_problem_free(DeviceProblem &dp);
size_t device_problem_bytes(int M, int N, int E);
// assembly launch-shape tuning (assembly.hip)
int autotune_assembly(const Mesh &m, double k, int precision, int *best_qg, int *best_waves, double *best_ms,
                      double *default_ms, hipStream_t s);
int tune_save(const char *file);
int tune_load(const char *file);
void tune_clear();
// multi-right-hand-side product tuning (gemm.hip)
int autotune_products(long M, long N, int R, int *shape, double *ms, hipStream_t s);
void product_tune_save(FILE *f);
bool product_tune_parse(const char *line);
void product_tune_clear();

struct RhsSpec {
    int kind = BEM_RHS_POINT;
    int nrhs = 0;
    float amp = 0.f;
    std::vector<float> params;  // 3 * nrhs
};

// Assemble A (ld dp.ldA, M x N) and dp.bterm on the device.
int assemble_fp32(DeviceProblem &dp, float k, float2 *A, hipStream_t s);
int assemble_fp64(DeviceProblem &dp, double k, double2 *A, hipStream_t s);
// B[:, s] = incident_s - bterm (M x nrhs, ldB)
int rhs_fp32(DeviceProblem &dp, double k, float2 *B, long ldB, const RhsSpec &rhs, hipStream_t s);
int rhs_fp64(DeviceProblem &dp, double k, double2 *B, long ldB, const RhsSpec &rhs, hipStream_t s);
// Exterior field: out[P x R] = inc - sum_e (...) for nodal solutions X (N x R).
int field_fp32(DeviceProblem &dp, float k, const float4 *xs, int P, const float2 *X, long ldx, int R,
               const RhsSpec &rhs, float2 *out, long ldo, hipStream_t s);

// The same field for one right-hand side (an incident point source of the
// given strength at src) with every temporary in the library's fixed-size
// static device scratch: no allocation.  xs: P points of 3 floats, X: N nodal
// values, out: P values, all in device memory; queued on s, no wait.
int field_fp32_noalloc(const DeviceProblem &dp, float k, const float *xs, int P, const float2 *X, float strength,
                       const float *src, float2 *out, hipStream_t s);
size_t field_noalloc_bytes(int E, int P);
size_t field_noalloc_capacity();

// Exterior field as a GEMM: fdp = mesh with the field points as extra rows.
int field_matrix_fp32(DeviceProblem &fdp, float k, float2 *F, long ldF, double2 *fbterm, hipStream_t s);
int field_apply_fp32(DeviceProblem &fdp, double k, const float2 *F, long ldF, const double2 *fbterm, const float2 *X,
                     long ldx, const RhsSpec &rhs, float2 *Y, float2 *out, long ldo, void *ws, hipStream_t s);

// ---------------------------------------------------------- linear algebra ---
size_t ws_gemv_n(long M, long N, int R);
// w = A^H (A v) in one pass over a row-major copy Ar of A (ldr = ne_ldr(N));
// FP32 A and v, FP64 w; for N up to the limit of ne_fused_fits
bool ne_fused_fits(long M, long N);
long ne_ldr(long N);
size_t ws_ne_apply(long M, long N);
int ne_apply_f32(const float2 *Ar, long ldr, long M, long N, const float2 *v, double2 *w, void *ws, hipStream_t s);
// r = A^H (b - A x) (x = null: A^H b) with FP64 accumulation, one pass over the row-major copy
size_t ws_ne_resid(long M, long N);
// the row-major copy and r = A^H b (FP64) in the same pass over A
size_t ws_ne_rhs(long M, long N);
int transpose_rm_ahb_f32(const float2 *A, long lda, long M, long N, float2 *Ar, long ldr, const double2 *b,
                         double2 *r, void *ws, hipStream_t s);
int ne_resid_f32a(const float2 *Ar, long ldr, long M, long N, const double2 *x, const double2 *b, double2 *r, void *ws,
                  hipStream_t s);
int gemv_n_f32(const float2 *A, long lda, long M, long N, const float2 *x, long ldx, float2 *y, long ldy, int R,
               void *ws, hipStream_t s);
int gemv_h_f32(const float2 *A, long lda, long M, long N, const float2 *x, long ldx, float2 *y, long ldy, int R,
               hipStream_t s);
int resid_n_f32a(const float2 *A, long lda, long M, long N, const double2 *x, long ldx, const double2 *b, long ldb,
                 double2 *r, long ldr, int R, void *ws, hipStream_t s);
int gemv_h_f32a(const float2 *A, long lda, long M, long N, const double2 *x, long ldx, double2 *y, long ldy, int R,
                hipStream_t s);
int gemv_n_f64(const double2 *A, long lda, long M, long N, const double2 *x, long ldx, double2 *y, long ldy, int R,
               void *ws, hipStream_t s);
int gemv_h_f64(const double2 *A, long lda, long M, long N, const double2 *x, long ldx, double2 *y, long ldy, int R,
               hipStream_t s);

// multi-right-hand-side products (gemm.hip)
size_t ws_gemm(long M, long N, int R);
size_t ws_gemm_h_f32(long M, long N, int R);
// long_chunks: fewer, longer FP32 column (row) chunks, faster for products a
// refinement corrects; the default keeps the FP32 sums short (QR, field)
int gemm_n_f32(const float2 *A, long lda, long M, long N, const float2 *X, long ldx, float2 *Y, long ldy, int R,
               void *ws, hipStream_t s, bool long_chunks = false);
int gemm_h_f32(const float2 *A, long lda, long M, long N, const float2 *T, long ldt, float2 *Z, long ldz, int R,
               void *ws, hipStream_t s, bool long_chunks = false);
int gemm_n_f32z(const float2 *A, long lda, long M, long N, const float2 *X, long ldx, double2 *Y, long ldy, int R,
                void *ws, hipStream_t s);
int gemm_h_f32z(const float2 *A, long lda, long M, long N, const float2 *T, long ldt, double2 *Z, long ldz, int R,
                void *ws, hipStream_t s);
// Adjoint products through a conjugate-transposed copy AH = A^H (N x M,
// column-major, ldah = adjoint_copy_ld(N)): Z = A^H T then runs on the Y = A X
// kernel, whose wave reads whole column segments of its operand.
bool adjoint_copy_pays(long M, long N, int R);
long adjoint_copy_ld(long N);
size_t adjoint_copy_bytes(long M, long N);
// allocates *AH and fills it when the copy pays and the HBM is free (else *AH
// stays null); adjoint_copy_fill refreshes an allocated copy
// (with bmax, >= ws_adjoint_absmax(M, N) bytes, the copy also leaves per-tile
// maxima of |Re A|, |Im A| there for adjoint_absmax_final; force makes it for
// 16 or fewer right-hand sides too, where the FP16-split products need it)
int adjoint_copy_make(const float2 *A, long lda, long M, long N, int R, float2 **AH, hipStream_t s,
                      float *bmax = nullptr, bool force = false);
int adjoint_copy_fill(const float2 *A, long lda, long M, long N, float2 *AH, hipStream_t s, float *bmax = nullptr);
int gemm_hc_f32(const float2 *AH, long M, long N, const float2 *T, long ldt, float2 *Z, long ldz, int R, void *ws,
                hipStream_t s);
int gemm_hc_f32z(const float2 *AH, long M, long N, const float2 *T, long ldt, double2 *Z, long ldz, int R, void *ws,
                 hipStream_t s);
// FP16-split products (A and X read as FP16 hi/lo pairs on FP16 MFMA, FP32
// accumulation): amax bounds |Re A|, |Im A| (adjoint_absmax_final after a
// copy made with bmax); xs >= ws_split_x
size_t ws_adjoint_absmax(long M, long N);
int adjoint_absmax_final(float *bmax, long M, long N, float **out, hipStream_t s);
size_t ws_split_x(long N, int R);
int gemm_n_f16s(const float2 *A, long lda, long M, long N, float amax, const float2 *X, long ldx, float2 *Y, long ldy,
                int R, void *ws, void *xs, hipStream_t s);
int gemm_hc_f16s_z(const float2 *AH, long M, long N, float amax, const float2 *T, long ldt, double2 *Z, long ldz,
                   int R, void *ws, void *xs, hipStream_t s);
size_t ws_mres(long M, long N, int R);
// nmat matrices astride elements apart, matrix z on right-hand sides
// z R / nmat .. (z + 1) R / nmat - 1: the FP32 product with FP64 output and
// the FP64-accumulated residual b - A X (ws >= ws_gemm(M, N, R))
int gemm_n_f32z_batched(const float2 *A, long lda, long astride, int nmat, long M, long N, const float2 *X, long ldx,
                        double2 *Y, long ldy, int R, void *ws, hipStream_t s);
int mres_f32a_batched(const float2 *A, long lda, long astride, int nmat, long M, long N, const double2 *X, long ldx,
                      const double2 *b, long ldb, double2 *out, long ldo, int R, void *ws, hipStream_t s);
// C -= V W (QR trailing update, V with 32 columns)
int rank_update_f32(const float2 *V, long ldv, int M, const float2 *W, int ldw, int nc, float2 *C, long ldc,
                    hipStream_t s);
int mres_f32a(const float2 *A, long lda, long M, long N, const double2 *X, long ldx, const double2 *b, long ldb,
              double2 *out, long ldo, int R, void *ws, hipStream_t s);
int mres_f64(const double2 *A, long lda, long M, long N, const double2 *X, long ldx, const double2 *b, long ldb,
             double2 *out, long ldo, int R, void *ws, hipStream_t s);
int mh_f32a(const float2 *A, long lda, long M, long N, const double2 *T, long ldt, double2 *Z, long ldz, int R,
            void *ws, hipStream_t s);
// mh_f32a through the conjugate-transposed copy AH (adjoint_copy_make)
int mh_f32a_hc(const float2 *AH, long M, long N, const double2 *T, long ldt, double2 *Z, long ldz, int R, void *ws,
               hipStream_t s);
int mh_f64(const double2 *A, long lda, long M, long N, const double2 *T, long ldt, double2 *Z, long ldz, int R,
           hipStream_t s);

// GMRES workspace (Krylov basis, row-major copy of A, partials) kept by a
// caller between solves of the same shape: allocating and freeing the
// row-major copy costs about 1 ms of host time per solve at N = 10242
struct GmresCache;
void gmres_cache_free(GmresCache *&c);
// X holds the initial guess (x_zero: known to be zero, which saves one residual).
// cache: null to allocate the workspace for this call only; else *cache is
// reused when its shape matches and replaced when it does not (free it with
// gmres_cache_free after the stream's work on it has been ordered)
// astride, nmat > 1: nmat square FP32 matrices astride elements apart, matrix
// z for systems z R / nmat .. (z + 1) R / nmat - 1 (frequencies in lockstep)
int gmres_device(const float2 *A32, const double2 *A64, long lda, int M, int N, const double2 *B, long ldb,
                 double2 *X, long ldx, bool x_zero, int R, int restart, double tol, int maxit, double outer_tol,
                 int max_cycles, bem_solve_stats *st, hipStream_t s, GmresCache **cache = nullptr, long astride = 0,
                 int nmat = 1);
// CGLS (conjugate gradients on A^H A x = A^H b) for R systems in lockstep
// (block = false) or as one block CG over all R (block = true, R <= 64),
// FP64 residual refinement; X is overwritten (the solve starts from zero)
int cgls_device(const float2 *A32, const double2 *A64, long lda, int M, int N, const double2 *B, long ldb,
                double2 *X, long ldx, int R, double tol, int maxit, double outer_tol, int max_cycles,
                bem_solve_stats *st, hipStream_t s, bool block = false);
int qr_lsq_f32(float2 *A, long lda, int M, int N, float2 *B, long ldb, int R, hipStream_t s);
int qr_lsq_f64(double2 *A, long lda, int M, int N, double2 *B, long ldb, int R, hipStream_t s);
int qr_lsq_with_q_f32(float2 *A, long lda, int M, int N, float2 *B, long ldb, int R, float2 *Q, long ldq,
                      hipStream_t s);
int qr_refined_f32(const float2 *A32, long lda, int M, int N, const double2 *B64, long ldb, double2 *X64, long ldx,
                   int R, double refine_tol, int max_refine, bem_solve_stats *st, hipStream_t s);

// host-side mesh operations
int read_obj(Mesh &m, const char *file);
int find_bb(Mesh &m, float threshold);
int gen_chief(Mesh &m, int num, float threshold, uint64_t seed);
int gen_icosphere(Mesh &m, int s, float radius);
int gen_ellipsoid(Mesh &m, int s, float ax, float ay, float az, int nb, int lmax, float amp, uint64_t seed);
double signed_volume(const Mesh &m);
int fix_orientation(Mesh &m, int *flipped);
int points_inside(const Mesh &m, const float *pts, int n, uint64_t seed, int *out);
double surf_area(const Mesh &m);
int gauss_legendre(int n, double *x, double *w);
int quad_init(int order);
int timing_enable(int on);
int timing_reset();
int timing_get(const char *name, double *total_ms, long *count);
const char *last_error();
void clear_error();

// Memory helpers
template <typename T>
int dalloc(T **p, size_t n) {
    if (n == 0) n = 1;
    HIP_CALL(hipMalloc((void **)p, n * sizeof(T)));
    return EXIT_SUCCESS;
}
template <typename T>
void dfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}
// Stream-ordered temporaries: hipFree synchronises the whole device, which would
// serialise solvers running on different streams.
template <typename T>
int salloc(T **p, size_t n, hipStream_t s) {
    if (n == 0) n = 1;
    HIP_CALL(hipMallocAsync((void **)p, n * sizeof(T), s));
    return EXIT_SUCCESS;
}
template <typename T>
void sfree(T *&p, hipStream_t s) {
    if (p) (void)hipFreeAsync((void *)p, s);
    p = nullptr;
}

}  // namespace bem
